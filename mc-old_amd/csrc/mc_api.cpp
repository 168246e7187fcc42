// mc_api.cpp — the C ABI (include/mc_api.h).  Every entry point catches
// mc::Error and std::exception, records the message for mc_last_error() and
// returns a status: nothing escapes the boundary.
#include <cstdio>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mc_api.h"
#include "mc_engine.hpp"
#include "mc_problem.hpp"

struct mc_problem {
    mc::Problem P;
};
struct mc_engine {
    std::unique_ptr<mc::Engine> E;
};

namespace {

thread_local std::string g_error;

template <class F>
mc_status guarded(F&& f)
{
    try {
        f();
        g_error.clear();
        return MC_OK;
    } catch (const mc::Error& e) {
        g_error = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        g_error = "out of host memory";
        return MC_ERR_CAPACITY;
    } catch (const std::exception& e) {
        g_error = e.what();
        return MC_ERR_INPUT;
    }
}

std::string read_file(const std::string& path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) throw mc::Error(MC_ERR_IO, "cannot read " + path);
    std::ostringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

}  // namespace

extern "C" {

const char* mc_last_error(void) { return g_error.c_str(); }
const char* mc_version(void) { return "mc-old_amd 0.1 (gfx950)"; }

mc_status mc_problem_from_xml(const char* io_dir, const char* xs_dir, mc_problem** out)
{
    return guarded([&] {
        if (!io_dir || !out) throw mc::Error(MC_ERR_ARGUMENT, "null argument");
        std::string dir(io_dir);
        if (!dir.empty() && dir.back() != '/') dir += '/';
        const std::string xml = read_file(dir + "input.xml");     // setup.cpp:33
        auto p = std::make_unique<mc_problem>();
        mc::load_problem_xml(p->P, xml, xs_dir ? xs_dir : "./xs_library");
        *out = p.release();
    });
}

mc_status mc_problem_from_xml_string(const char* xml, const char* xs_dir, mc_problem** out)
{
    return guarded([&] {
        if (!xml || !out) throw mc::Error(MC_ERR_ARGUMENT, "null argument");
        auto p = std::make_unique<mc_problem>();
        mc::load_problem_xml(p->P, xml, xs_dir ? xs_dir : "./xs_library");
        *out = p.release();
    });
}

mc_status mc_problem_create(const mc_problem_desc* desc, mc_problem** out)
{
    return guarded([&] {
        if (!desc || !out) throw mc::Error(MC_ERR_ARGUMENT, "null argument");
        auto p = std::make_unique<mc_problem>();
        mc::load_problem_desc(p->P, *desc);
        *out = p.release();
    });
}

mc_status mc_problem_get_desc(const mc_problem* p, const mc_problem_desc** out)
{
    return guarded([&] {
        if (!p || !out) throw mc::Error(MC_ERR_ARGUMENT, "null argument");
        *out = &p->P.desc;
    });
}

const char* mc_problem_estimator_name(const mc_problem* p, int32_t e)
{
    if (!p || e < 0 || e >= (int32_t)p->P.estimator_names.size()) return nullptr;
    return p->P.estimator_names[e].c_str();
}

const char* mc_problem_object_name(const mc_problem* p, int32_t kind, int32_t i)
{
    if (!p || i < 0) return nullptr;
    const mc::Problem& P = p->P;
    const std::vector<std::string>* v = kind == MC_OBJ_SURFACE     ? &P.surface_names
                                        : kind == MC_OBJ_CELL      ? &P.cell_names
                                        : kind == MC_OBJ_MATERIAL  ? &P.material_names
                                        : kind == MC_OBJ_NUCLIDE   ? &P.nuclide_names
                                        : kind == MC_OBJ_ESTIMATOR ? &P.estimator_names
                                                                   : nullptr;
    if (!v || i >= (int32_t)v->size()) return nullptr;
    return (*v)[i].c_str();
}

void mc_problem_destroy(mc_problem* p) { delete p; }

mc_status mc_engine_create(mc_problem* const* problems, int32_t n, const mc_engine_cfg* cfg, mc_engine** out)
{
    return guarded([&] {
        if (!problems || n <= 0 || !out) throw mc::Error(MC_ERR_ARGUMENT, "null argument");
        std::vector<const mc::Problem*> ps;
        for (int i = 0; i < n; ++i) {
            if (!problems[i]) throw mc::Error(MC_ERR_ARGUMENT, "null problem");
            ps.push_back(&problems[i]->P);
        }
        auto e = std::make_unique<mc_engine>();
        e->E = std::make_unique<mc::Engine>(ps, cfg);
        *out = e.release();
    });
}

mc_status mc_engine_info_get(const mc_engine* e, mc_engine_info* out)
{
    return guarded([&] {
        if (!e || !out) throw mc::Error(MC_ERR_ARGUMENT, "null argument");
        e->E->info(out);
    });
}

mc_status mc_engine_bin_offset(const mc_engine* e, int32_t p, int32_t* offset, int32_t* n_bins)
{
    return guarded([&] {
        if (!e) throw mc::Error(MC_ERR_ARGUMENT, "null argument");
        e->E->offsets(p, offset, n_bins, nullptr, nullptr);
    });
}

mc_status mc_engine_param_offset(const mc_engine* e, int32_t p, int32_t* offset, int32_t* n_params)
{
    return guarded([&] {
        if (!e) throw mc::Error(MC_ERR_ARGUMENT, "null argument");
        e->E->offsets(p, nullptr, nullptr, offset, n_params);
    });
}

mc_status mc_engine_kernel_stats(mc_engine* e, mc_kernel_stats* out, int32_t reset)
{
    return guarded([&] {
        if (!e || !out) throw mc::Error(MC_ERR_ARGUMENT, "null argument");
        e->E->kernel_stats(out, reset != 0);
    });
}

mc_status mc_engine_copy_ring(const mc_engine* e, uint64_t* capacity, uint64_t* growths, uint64_t* waits)
{
    return guarded([&] {
        if (!e) throw mc::Error(MC_ERR_ARGUMENT, "null engine");
        e->E->copy_ring(capacity, growths, waits);
    });
}

mc_status mc_forward(mc_engine* e, const mc_run_cfg* run, mc_tally_out* out)
{
    return guarded([&] {
        if (!e) throw mc::Error(MC_ERR_ARGUMENT, "null engine");
        e->E->forward(run, out);
    });
}

mc_status mc_backward(mc_engine* e, const mc_run_cfg* run, const double* grad_out, double* grad,
                      double* grad_sigma, mc_tally_out* out)
{
    return guarded([&] {
        if (!e) throw mc::Error(MC_ERR_ARGUMENT, "null engine");
        e->E->backward(run, grad_out, grad, grad_sigma, out);
    });
}

mc_status mc_xs_lookup(mc_engine* e, const double* energy, const int32_t* mat, double* sigt, uint64_t n)
{
    return guarded([&] {
        if (!e || (n && (!energy || !mat || !sigt))) throw mc::Error(MC_ERR_ARGUMENT, "null argument");
        if (n) e->E->xs_lookup(energy, mat, sigt, n);
    });
}

mc_status mc_xs_lookup_noalloc(mc_engine* e, const double* energy, const int32_t* mat, double* sigt, uint64_t n)
{
    return guarded([&] {
        if (!e || (n && (!energy || !mat || !sigt))) throw mc::Error(MC_ERR_ARGUMENT, "null argument");
        if (n) e->E->xs_lookup_noalloc(energy, mat, sigt, n);
    });
}

mc_status mc_xs_lookup_bench(mc_engine* e, uint64_t n, int32_t iters, uint64_t seed, double* seconds,
                             double* checksum)
{
    return guarded([&] {
        if (!e || !seconds || n == 0) throw mc::Error(MC_ERR_ARGUMENT, "bad argument");
        e->E->xs_bench(n, iters, seed, seconds, checksum);
    });
}

mc_status mc_gather_bench(mc_engine* e, uint64_t n, int32_t iters, uint64_t seed, uint64_t table_bytes,
                          int32_t dependent_reads, double* seconds, uint64_t* lookup_table_bytes)
{
    return guarded([&] {
        if (!e || !seconds || n == 0) throw mc::Error(MC_ERR_ARGUMENT, "bad argument");
        if (lookup_table_bytes) *lookup_table_bytes = e->E->lookup_table_bytes();
        const uint64_t tb = table_bytes ? table_bytes : e->E->lookup_table_bytes();
        e->E->gather_bench(n, iters, seed, tb, dependent_reads, seconds);
    });
}

mc_status mc_selftest_math(const double* in, double* out, uint64_t n)
{
    return guarded([&] {
        if (!in || !out) throw mc::Error(MC_ERR_ARGUMENT, "null argument");
        if (n) mc::selftest_math(in, out, n);
    });
}

void mc_engine_destroy(mc_engine* e) { delete e; }

mc_status mc_report_text(const mc_problem* p, const mc_tally_out* t, uint64_t n_histories, const char* path)
{
    return guarded([&] {
        if (!p || !t || !path || !t->mean || !t->uncer) throw mc::Error(MC_ERR_ARGUMENT, "null argument");
        const mc::Problem& P = p->P;
        std::ofstream o(path);
        if (!o) throw mc::Error(MC_ERR_IO, std::string("cannot write ") + path);
        o.precision(17);
        // report.cpp:23-46
        o << "/summary/Ncycle 1\n/summary/Nsample " << n_histories << "\n/summary/Npassive 0\n"
          << "/summary/Ntrack " << t->flights << "\n/summary/mode fixed source\n"
          << "/summary/survival_roulette/wr " << P.wr << "\n/summary/survival_roulette/ws " << P.ws << "\n";
        // Estimator::report, Estimator.cpp:368-422
        for (size_t e = 0; e < P.estimators.size(); ++e) {
            const auto& E = P.estimators[e];
            const std::string& name = P.estimator_names[e];
            std::string indexing = E.attach_kind == MC_ATTACH_SURFACE ? "[surface]" : "[cell]";
            if (E.egrid_count) indexing += "[energy]";
            o << "/" << name << " indexing " << indexing << "\n";
            o << "/" << name << "/" << (E.attach_kind == MC_ATTACH_SURFACE ? "surface" : "cell") << " id#";
            for (double v : P.estimator_attach_ids[e]) o << " " << v;
            o << "\n";
            if (E.egrid_count) {
                o << "/" << name << "/energy eV";
                for (double v : P.estimator_egrid[e]) o << " " << v;
                o << "\n";
            }
            const int per = E.n_bins / E.n_scores;
            for (int s = 0; s < E.n_scores; ++s) {
                const std::string sn = P.estimator_scores[e][s];
                o << "/" << name << "/" << sn << "/mean";
                for (int k = 0; k < per; ++k) o << " " << t->mean[E.bin_offset + s * per + k];
                o << "\n/" << name << "/" << sn << "/uncertainty";
                for (int k = 0; k < per; ++k) o << " " << t->uncer[E.bin_offset + s * per + k];
                o << "\n";
            }
        }
    });
}

mc_status mc_report_h5(const mc_problem* p, const mc_tally_out* t, uint64_t n_histories, const char* path)
{
    return guarded([&] {
        if (!p || !t || !path || !t->mean || !t->uncer) throw mc::Error(MC_ERR_ARGUMENT, "null argument");
        mc::report_h5(p->P, *t, n_histories, path);
    });
}

mc_status mc_autotune(mc_engine* e, uint64_t sample_histories, mc_tune_result* out)
{
    return guarded([&] {
        if (!e || !out || sample_histories < 2) throw mc::Error(MC_ERR_ARGUMENT, "bad argument");
        mc::Engine& E = *e->E;
        const int np = E.n_problems();
        // sample: every problem keeps its share of the deck histories
        std::vector<uint64_t> nh(np);
        uint64_t total = 0;
        for (int p = 0; p < np; ++p) total += E.deck_histories(p);
        for (int p = 0; p < np; ++p) {
            const double f = (double)E.deck_histories(p) / (double)total;
            nh[p] = std::max<uint64_t>(2, (uint64_t)(f * (double)sample_histories));
        }
        uint64_t sampled = 0;
        for (auto v : nh) sampled += v;
        mc_run_cfg run{};
        run.mode = MC_RUN_BATCHED;
        run.n_histories = nh.data();
        const mc_engine_cfg base = E.cfg();
        mc_engine_cfg best = base;
        double best_rate = -1.0;
        const int blocks[] = {256, 512};
        const int grids[] = {2, 4};
        const uint64_t banks[] = {1ull << 19, 1ull << 20, 1ull << 21};
        const int drains[] = {1, -1};
        // flights per particle per iteration (forward runs)
        const std::vector<int> events = {1, 2, 4, 6, 8, 16};
        for (uint64_t B : banks) {
            if (B > 4 * sampled && B != banks[0]) continue;
            for (int blk : blocks)
                for (int gpc : grids)
                    for (int dd : drains)
                    for (int ev : events) {
                        mc_engine_cfg c = base;
                        c.block = blk;
                        c.grid_per_cu = gpc;
                        c.bank_capacity = B;
                        c.iters_per_graph = 16;
                        c.drain_divisor = dd;
                        c.events_per_launch = ev;
                        c.use_graph = base.use_graph;   // tune the launch mode the engine uses
                        c.profile = 0;
                        E.set_cfg(c);
                        mc_tally_out t{};
                        E.forward(&run, &t);   // warm (graph capture)
                        E.forward(&run, &t);
                        const double rate = (double)sampled / t.seconds;
                        if (rate > best_rate) { best_rate = rate; best = c; }
                    }
        }
        E.set_cfg(best);
        out->block = best.block;
        out->grid_per_cu = best.grid_per_cu;
        out->iters_per_graph = best.iters_per_graph;
        out->drain_divisor = best.drain_divisor;
        out->bank_capacity = best.bank_capacity;
        out->histories_per_second = best_rate;
        out->events_per_launch = best.events_per_launch;
    });
}

}  // extern "C"
