// ref_driver.cpp — C wrappers around reference functions compiled from
// /root/reference/src/{Random,Algorithm,XSec}.cpp (oracle/_ref only).
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle against the
// reference's own code where that code builds without external libraries.
#include <vector>
#include "Point.h"
#include "Algorithm.h"
#include "XSec.h"

// Random.cpp:121-204 (declared there with C++ linkage)
double Urand(void);
unsigned long long RN_skip_ahead(unsigned long long* s, long long* n);
void RN_init_problem(unsigned long long* new_seed, int* print_info);
void RN_init_particle(unsigned long long* nps);

extern "C" {

void ref_rn_init_problem(unsigned long long seed)
{
    int quiet = 0;
    RN_init_problem(&seed, &quiet);
}
void ref_rn_init_particle(unsigned long long nps) { RN_init_particle(&nps); }
double ref_urand(void) { return Urand(); }
unsigned long long ref_skip_ahead(unsigned long long seed, long long n)
{
    return RN_skip_ahead(&seed, &n);
}
int ref_binary_search(double x, const double* v, int n)
{
    std::vector<double> vec(v, v + n);
    return binary_search(x, vec);
}
double ref_interpolate(double x, double x1, double x2, double y1, double y2)
{
    return interpolate(x, x1, x2, y1, y2);
}
double ref_geometry_quad(double a, double b, double c) { return geometry_quad(a, b, c); }
void ref_scatter_direction(const double* d, double mu, double* o)
{
    Point p = scatter_direction(Point(d[0], d[1], d[2]), mu);
    o[0] = p.x; o[1] = p.y; o[2] = p.z;
}
double ref_exponential_sample(double param) { return exponential_sample(param); }
// XSTable::xs with the bin index of Nuclide::checkE (Nuclide.cpp:18-24)
double ref_xstable(const double* grid, const double* y, int n, double E)
{
    std::vector<double> g(grid, grid + n), v(y, y + n);
    XSTable t(v);
    const int idx = binary_search(E, g);
    return t.xs((unsigned long long)(long long)idx, E, g);
}

}  // extern "C"
