// mc_problem.hpp — one fixed-source problem as flat arrays (the mc_problem
// handle behind the C ABI), built from a deck in the reference's XML schema.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mc_api.h"

namespace mc {

struct Error : std::runtime_error {
    mc_status status;
    Error(mc_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

struct Problem {
    std::string name;
    std::vector<std::string> surface_names, cell_names, material_names, nuclide_names,
        estimator_names;
    std::vector<std::vector<std::string>> estimator_scores;   // names as in the deck
    std::vector<std::vector<double>> estimator_attach_ids;    // filter grid (IDs)
    std::vector<std::vector<double>> estimator_egrid;

    std::vector<mc_surface_desc> surfaces;
    std::vector<mc_cell_desc> cells;
    std::vector<int32_t> cell_surf, cell_sense;
    std::vector<mc_material_desc> materials;
    std::vector<int32_t> mat_nuclide;
    std::vector<double> mat_density;
    std::vector<mc_nuclide_desc> nuclides;
    std::vector<double> xs_energy, xs_scatter, xs_capture, xs_total;
    std::vector<mc_source_desc> sources;
    std::vector<double> source_cdf;
    std::vector<mc_estimator_desc> estimators;
    std::vector<int32_t> att_estimator, att_index;
    std::vector<double> egrid;
    std::vector<mc_param_desc> params;
    std::vector<int32_t> mat_param;
    double wr = 0.001, ws = 1.0;       // simulator.h:90-91
    uint64_t n_histories = 0;
    uint64_t seed = 1;                  // Random.cpp:105
    int32_t comb_bank_max = 0, comb_teeth = 0;   // particle comb (0: off), setup.cpp:57-67
    int32_t n_bins = 0;

    mc_problem_desc desc{};
    void finalize();                    // fills desc from the vectors

    int search_cell(const double p[3]) const;   // general.cpp:26-34
};

// Builds a problem from deck text (reference schema, setup.cpp:30-1069).
// xs_dir: directory of <ZAID>.txt cross-section files.
void load_problem_xml(Problem& P, const std::string& xml, const std::string& xs_dir);

// Builds a problem from caller-owned flat arrays (mc_problem_create):
// copies them, checks every index and names the records by position.
void load_problem_desc(Problem& P, const mc_problem_desc& d);

// Reads one xs_library file (format: xs_library/README.txt:1-7).  Accepts
// rows of 5 or 6 numbers (the reference reader assumes 6, setup.cpp:366).
void load_xs_file(const std::string& path, double& A, std::vector<double>& E,
                  std::vector<double>& s, std::vector<double>& c, std::vector<double>& f);

// output.h5 in the reference's layout (report.cpp:9-50, Estimator.cpp:368-422)
// through the HDF5 C library loaded at run time (mc_report.cpp); throws
// MC_ERR_UNSUPPORTED when none can be loaded.
bool hdf5_available();
void report_h5(const Problem& P, const mc_tally_out& t, uint64_t n_histories, const std::string& path);

// Reference helpers reproduced for setup-time arithmetic.
int ref_binary_search(double x, const double* v, int n);   // Algorithm.cpp:46-64

}  // namespace mc
