// simulator.hpp — drop-in C++ host interface of the reference's fixed-source
// Simulator (include/simulator.h:21-150, Estimator.h:257-269, 384-423),
// implemented on the C ABI of libmcold.so (include/mc_api.h).
//
//   mcold::Simulator S(io_dir);         // Simulator::Simulator(io_dir), setup.cpp:30
//   S.start();                          // Simulator::start(), handler.cpp:11-48
//   S.Estimators[0]->tally(0).mean;     // Estimator::tally(i), Estimator.cpp:424-427
//   S.report(io_dir);                   // Simulator::report(io_dir), report.cpp:9-50
//
// Errors that the reference reports by printing and exiting are thrown here
// as mcold::Error carrying the mc_status code.
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
    std::vector<std::shared_ptr<Estimator>> Estimators;
    double device_seconds = 0.0;

private:
    mc_problem* problem = nullptr;
    mc_engine* engine = nullptr;
    std::vector<double> sum, sq, mean, uncer;
    bool started = false;
};

}  // namespace mcold
