// simulator.hpp — drop-in C++ host interface of the reference's fixed-source
// Simulator (include/simulator.h:21-150, Estimator.h:257-269, 384-423),
// implemented on the C ABI of libmcold.so (include/mc_api.h).
//
//   mcold::Simulator S(io_dir);         // Simulator::Simulator(io_dir), setup.cpp:30
//   S.start();                          // Simulator::start(), handler.cpp:11-48
//   S.Estimators[0]->tally(0).mean;     // Estimator::tally(i), Estimator.cpp:424-427
//   S.report(io_dir);                   // Simulator::report(io_dir), report.cpp:9-50
//   S.Surfaces / Cells / Materials / Nuclides   (simulator.h:118-122), read-only
//
// Errors that the reference reports by printing and exiting are thrown here
// as mcold::Error carrying the mc_status code.  The GPU engine is created by
// the first call that needs the device (start, Material::SigmaT), so a deck
// can be loaded and inspected on a machine without one.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mc_api.h"

namespace mcold {

struct Error : std::runtime_error {
    mc_status status;
    Error(mc_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

struct Point {                       // Point.h
    double x = 0.0, y = 0.0, z = 0.0;
    Point() = default;
    Point(double a, double b, double c) : x(a), y(b), z(c) {}
};

// The particle state the geometry queries take (Particle.h:12-64: position,
// direction, energy, weight).
class Particle {
public:
    Particle(const Point& p, const Point& d, double E = 2.0e6, double w = 1.0)
        : p_pos(p), p_dir(d), p_energy(E), p_weight(w) {}
    Point pos() const { return p_pos; }
    Point dir() const { return p_dir; }
    double energy() const { return p_energy; }
    double weight() const { return p_weight; }
    void set_direction(const Point& d) { p_dir = d; }
    void move(double d) { p_pos = Point(p_pos.x + p_dir.x * d, p_pos.y + p_dir.y * d, p_pos.z + p_dir.z * d); }

private:
    Point p_pos, p_dir;
    double p_energy, p_weight;
};

class Simulator;

// Surface (Geometry.h:41-64): the deck's surface, evaluated on the host with
// the same routines the device kernels use (for callers' setup-time queries;
// the transport evaluates them on the device).
class Surface {
public:
    Surface(std::string n, int id, const mc_surface_desc& d) : s_name(std::move(n)), s_id(id), s_desc(d) {}
    std::string name() const { return s_name; }
    int ID() const { return s_id; }
    int bc() const { return s_desc.bc; }            // 0 transmission, 1 reflective, -1 vacuum
    int type() const { return s_desc.type; }        // MC_SURF_*
    double eval(const Point& p) const;              // Geometry.cpp:29-69
    double distance(const Particle& P) const;       // Geometry.cpp:76-188
    void reflect(Particle& P) const;                // Geometry.cpp:195-222

private:
    std::string s_name;
    int s_id;
    mc_surface_desc s_desc;
};

// Nuclide (Nuclide.h:15-63), read-only: name, mass ratio, table size and
// scattering model.
class Nuclide {
public:
    Nuclide(std::string n, const mc_nuclide_desc& d) : n_name(std::move(n)), n_desc(d) {}
    std::string name() const { return n_name; }
    double A() const { return n_desc.A; }
    bool tabulated() const { return n_desc.kind == MC_NUC_TABLE; }
    int grid_points() const { return tabulated() ? n_desc.grid_count : 0; }
    bool free_gas() const { return n_desc.kinematics == MC_KIN_FREE_GAS && n_desc.A < MC_MASS_INFINITE; }

private:
    std::string n_name;
    mc_nuclide_desc n_desc;
};

// Material (Material.h:15-46): its nuclides and densities; SigmaT(E) is the
// engine's cross-section lookup on the device (Material.cpp:50-57).
class Material {
public:
    Material(std::string n, int index, Simulator* owner,
             std::vector<std::pair<std::shared_ptr<Nuclide>, double>> nd)
        : m_name(std::move(n)), m_index(index), m_owner(owner), m_nuclides(std::move(nd)) {}
    std::string name() const { return m_name; }
    const std::vector<std::pair<std::shared_ptr<Nuclide>, double>>& nuclides() const { return m_nuclides; }
    double SigmaT(double E) const;
    std::vector<double> SigmaT(const std::vector<double>& E) const;

private:
    std::string m_name;
    int m_index;
    Simulator* m_owner;
    std::vector<std::pair<std::shared_ptr<Nuclide>, double>> m_nuclides;
};

// Cell (Geometry.h:195-220)
class Cell {
public:
    Cell(std::string n, int id, double imp, std::shared_ptr<Material> m,
         std::vector<std::pair<std::shared_ptr<Surface>, int>> s)
        : c_name(std::move(n)), c_id(id), c_importance(imp), c_material(std::move(m)), c_surfaces(std::move(s)) {}
    std::string name() const { return c_name; }
    int ID() const { return c_id; }
    double importance() const { return c_importance; }
    const std::shared_ptr<Material>& material() const { return c_material; }
    const std::vector<std::pair<std::shared_ptr<Surface>, int>>& surfaces() const { return c_surfaces; }
    // Simulator::test_point (general.cpp:13-20): p lies in the cell
    bool test_point(const Point& p) const;

private:
    std::string c_name;
    int c_id;
    double c_importance;
    std::shared_ptr<Material> c_material;
    std::vector<std::pair<std::shared_ptr<Surface>, int>> c_surfaces;
};

// Estimator.h:257-269 (after end_simulation only mean and uncer are set)
struct Tally {
    double hist = 0.0, sum = 0.0, squared = 0.0;
    double mean = 0.0, uncer = 0.0;
};

class Estimator {
public:
    Estimator(std::string name, std::vector<Tally> t) : e_name(std::move(name)), e_tally(std::move(t)) {}
    std::string name() const { return e_name; }
    Tally tally(const int i) const { return e_tally.at(i); }
    unsigned long long tally_size() const { return e_tally.size(); }

private:
    std::string e_name;
    std::vector<Tally> e_tally;
};

class Simulator {
public:
    // Reads <io_dir>/input.xml; ZAID nuclides from xs_dir (the reference
    // reads ./xs_library relative to the working directory, setup.cpp:326).
    explicit Simulator(const std::string io_dir, const std::string xs_dir = "./xs_library",
                       const mc_engine_cfg* cfg = nullptr);
    ~Simulator();
    Simulator(const Simulator&) = delete;
    Simulator& operator=(const Simulator&) = delete;

    void start();
    void report(const std::string io_dir);

    std::string simulation_name;
    unsigned long long Nsample = 0;
    unsigned long long Ncycle = 1, Npassive = 0;
    unsigned long long Ntrack = 0;
    std::string mode = "fixed source";
    // THE OBJECTS (simulator.h:118-122), read-only views of the deck
    std::vector<std::shared_ptr<Surface>> Surfaces;
    std::vector<std::shared_ptr<Cell>> Cells;
    std::vector<std::shared_ptr<Nuclide>> Nuclides;
    std::vector<std::shared_ptr<Material>> Materials;
    std::vector<std::shared_ptr<Estimator>> Estimators;
    double device_seconds = 0.0;
    std::string report_path;          // the file the last report() wrote

    // Simulator::search_cell (general.cpp:26-34): the first cell holding p, or null
    std::shared_ptr<Cell> search_cell(const Point& p) const;

private:
    friend class Material;
    mc_engine* device_engine();       // created on first use
    const mc_engine_cfg* engine_cfg = nullptr;
    mc_engine_cfg cfg_copy{};
    mc_problem* problem = nullptr;
    mc_engine* engine = nullptr;
    std::vector<double> sum, sq, mean, uncer;
    bool started = false;
};

}  // namespace mcold
