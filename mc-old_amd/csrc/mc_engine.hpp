// mc_engine.hpp — the GPU engine behind mc_engine handles.
#pragma once

#include <memory>
#include <vector>

#include "../../include/mc_api.h"

namespace mc {

struct Problem;

class Engine {
public:
    Engine(const std::vector<const Problem*>& problems, const mc_engine_cfg* cfg);
    ~Engine();
    void info(mc_engine_info* out) const;
    void offsets(int p, int32_t* bin_off, int32_t* n_bins, int32_t* par_off, int32_t* n_par) const;
    void forward(const mc_run_cfg* run, mc_tally_out* out);
    void backward(const mc_run_cfg* run, const double* grad_out, double* grad, double* sigma, mc_tally_out* out);
    void xs_lookup(const double* E, const int32_t* mat, double* out, uint64_t n);
    void xs_lookup_noalloc(const double* E, const int32_t* mat, double* out, uint64_t n);
    void xs_bench(uint64_t n, int iters, uint64_t seed, double* seconds, double* checksum);
    void gather_bench(uint64_t n, int iters, uint64_t seed, uint64_t table_bytes, int dep, double* seconds);
    uint64_t lookup_table_bytes() const;
    void set_cfg(const mc_engine_cfg& cfg);
    void kernel_stats(mc_kernel_stats* out, bool reset);
    void copy_ring(uint64_t* capacity, uint64_t* growths, uint64_t* waits) const;
    mc_engine_cfg cfg() const;
    int n_problems() const;
    uint64_t deck_histories(int p) const;

private:
    struct Impl;
    std::unique_ptr<Impl> d;
};

void selftest_math(const double* in, double* out, uint64_t n);

}  // namespace mc
