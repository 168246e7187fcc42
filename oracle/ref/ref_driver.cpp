// ref_driver.cpp — C wrappers around reference functions compiled from the
// reference's own sources (oracle/_ref only): src/{Random,Algorithm,XSec}.cpp
// always, and, where the HDF5 headers that Geometry.h pulls in (through
// Estimator.h) are installed, src/{Distribution,Reaction,Particle,Geometry}.cpp
// as well (REF_FULL).  No HDF5 routine is called or linked: the headers only
// declare the Estimator's report types.
// TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle against the
// reference's own code.
#include <memory>
#include <string>
#include <vector>
#include "Point.h"
#include "Algorithm.h"
#include "XSec.h"
#ifdef REF_FULL
#include "Distribution.h"
#include "Geometry.h"
#include "Particle.h"
#include "Reaction.h"
#endif

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

#ifdef REF_FULL
int ref_full(void) { return 1; }

// DistributionIsotropicDirection::sample (Distribution.cpp:78-92) on the
// global stream
void ref_isotropic_direction(double* o)
{
    DistributionIsotropicDirection d;
    const Point p = d.sample();
    o[0] = p.x; o[1] = p.y; o[2] = p.z;
}

// ReactionScatter::sample (Reaction.cpp:27-118): free-gas elastic scatter of
// a particle of energy E and direction u off a nuclide of mass ratio A, on
// the global stream; *E and u receive the particle's new energy and direction
void ref_free_gas_scatter(double A, double* E, double* u)
{
    std::shared_ptr<Distribution<double>> mu(new DistributionIsotropicScatter());
    ReactionScatter r(nullptr, mu, A);
    Particle P(Point(0.0, 0.0, 0.0), Point(u[0], u[1], u[2]), *E, 0.0, 1.0, 0, nullptr);
    r.sample(P);
    *E = P.energy();
    u[0] = P.dir().x; u[1] = P.dir().y; u[2] = P.dir().z;
}

// Surface::eval / distance / reflect (Geometry.cpp:29-222) of a surface given
// by the flat description of include/mc_api.h (type, c[]): out[0] = eval(p),
// out[1] = distance of a particle at p going along u, out[2..4] = its
// direction after reflect (quadrics leave it unchanged)
void ref_surface(int type, const double* c, const double* p, const double* u, double* out)
{
    std::shared_ptr<Surface> S;
    switch (type) {
    case 1: S.reset(new SurfacePlaneX("s", 0, 0, c[0])); break;
    case 2: S.reset(new SurfacePlaneY("s", 0, 0, c[0])); break;
    case 3: S.reset(new SurfacePlaneZ("s", 0, 0, c[0])); break;
    case 4: S.reset(new SurfacePlane("s", 0, 0, c[0], c[1], c[2], c[3])); break;
    case 5: S.reset(new SurfaceSphere("s", 0, 0, c[0], c[1], c[2], c[3])); break;
    case 6: S.reset(new SurfaceCylinderX("s", 0, 0, c[0], c[1], c[2])); break;
    case 8: S.reset(new SurfaceCylinderZ("s", 0, 0, c[0], c[1], c[2])); break;
    default: out[0] = out[1] = out[2] = out[3] = out[4] = 0.0; return;
    }
    Particle P(Point(p[0], p[1], p[2]), Point(u[0], u[1], u[2]), 1.0, 0.0, 1.0, 0, nullptr);
    out[0] = S->eval(P.pos());
    out[1] = S->distance(P);
    S->reflect(P);
    out[2] = P.dir().x; out[3] = P.dir().y; out[4] = P.dir().z;
}
#else
int ref_full(void) { return 0; }
#endif

}  // extern "C"
