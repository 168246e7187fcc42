// mc_engine.hip — host side of the GPU engine: device upload of one or many
// problems, the event loop (hipGraph replay of captured iterations), tallies,
// gradients, the standalone XS lookup and the launch-shape autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "mc_device.hpp"
#include "mc_engine.hpp"
#include "mc_problem.hpp"

namespace mc {

// kernels (mc_kernels.hip)
__global__ void k_refill_count(KArgs A);
template <bool kGrad> __global__ void k_refill_scatter(KArgs A);
template <bool kGrad> __global__ void k_resume(KArgs A);
__global__ void k_count_events(const uint32_t* wg_cnt, uint32_t n_wg, Counters* c);
__global__ void k_xs_lookup(KArgs A, int use_lds);
template <bool kTables, bool kGrad, bool kFuse, int kGeo> __global__ void k_flight(KArgs A);
template <bool kTables, bool kGrad> __global__ void k_collision(KArgs A);
template <bool kTables> __global__ void k_cross(KArgs A);
template <bool kTables, bool kGrad> __global__ void k_drain(KArgs A, uint32_t max_events);
__global__ void k_drain_close(Counters* c);
template <int kU> __global__ void k_xs_lookup_array(KArgs A, const double* E, const int32_t* mat, double* out, uint64_t n);
constexpr int kXsUnroll = 4;      // lookups in flight per thread in the standalone kernel (measured best of 1, 2, 4)
template <int kDep>
__global__ void k_gather_lines(const double* E, const int32_t* mat, double* out, uint64_t n, const double* table,
                               uint64_t nlines);
__global__ void k_xs_inputs(double* E, int32_t* mat, uint64_t n, double lnmin, double lnspan,
                            const int32_t* mats, int32_t nmat, uint64_t seed);
__global__ void k_fill_f64(double* p, uint64_t n, double v);
__global__ void k_reduce_rows(const double* part, uint64_t rows, uint64_t n, double* out);
__global__ void k_moments(const double* sum, const double* sq, const double* N, uint64_t n, double* osum,
                          double* osq, double* omean, double* ouncer);
__global__ void k_fill_u8(uint8_t* p, uint64_t n, uint8_t v);
__global__ void k_fill_i32(int32_t* p, uint64_t n, int32_t v);
__global__ void k_selftest_math(const double* in, double* out, uint64_t n);

#define HIP_CHECK(x)                                                                          \
    do {                                                                                      \
        hipError_t _e = (x);                                                                  \
        if (_e != hipSuccess)                                                                 \
            throw Error(MC_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(_e));      \
    } while (0)

// ------------------------------------------------- allocation-free lookups
// mc_xs_lookup_noalloc stages its inputs and outputs in this array instead of
// allocating device buffers: energies at byte 0, Sigma_t at 8 kScratchLookups,
// materials at 16 kScratchLookups.  Each device holds its own copy from the
// module's load until the process ends; every call on a device reuses that
// device's copy, so calls that overlap on one device race.  Nothing in it
// needs to start at zero: a call writes every entry it reads.
constexpr uint64_t kScratchLookups = 1ull << 19;
constexpr size_t kScratchBytes = (size_t)kScratchLookups * 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The device address of `scratch` on device dev (the array's name in host
// code is only a stand-in), looked up once per device with dev current.
static unsigned char* device_scratch(int dev)
{
    static std::mutex mu;
    static std::map<int, unsigned char*> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(dev);
    if (it != addr.end()) return it->second;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw Error(MC_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    addr[dev] = (unsigned char*)p;
    return (unsigned char*)p;
}

// ----------------------------------------------------------- device buffers
// Blocks of the per-run key-to-problem index.
constexpr uint32_t kHbIdxMax = 65536;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    void alloc(size_t n)
    {
        if (n <= bytes && p) return;
        if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
        if (n == 0) n = 8;
        HIP_CHECK(hipMalloc(&p, n));
        bytes = n;
    }
    template <class T> T* as() const { return (T*)p; }
};

template <class T>
static void upload(DevBuf& b, const std::vector<T>& v, hipStream_t st)
{
    b.alloc(sizeof(T) * v.size());
    if (!v.empty()) HIP_CHECK(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
}

static uint64_t lcg_pow(uint64_t g, uint64_t e)
{
    uint64_t r = 1;
    for (; e; e >>= 1) {
        if (e & 1) r = (r * g) & kRnMask;
        g = (g * g) & kRnMask;
    }
    return r;
}

struct Engine::Impl {
    int device = 0;
    unsigned char* scratch_ptr = nullptr;   // this device's `scratch` (allocation-free lookups)
    hipStream_t stream = nullptr;      // the engine's own stream (graphs are captured on it)
    hipStream_t cur = nullptr;         // the stream of the current run
    int n_cu = 0;
    mc_engine_cfg cfg{};
    // per problem, copied at creation (the engine does not keep the problems)
    std::vector<uint64_t> deck_hist;
    int n_mat0 = 0;                    // materials of problem 0 (standalone lookups)

    // concatenated host data
    std::vector<mc_surface_desc> surf;
    std::vector<mc_cell_desc> cell;
    std::vector<int32_t> cell_surf, cell_sense, nbr_begin, nbr;
    std::vector<double> crec;          // per-cell flight records (as doubles: 8-byte aligned)
    int32_t crec_stride = 0;
    std::vector<mc_material_desc> mat;
    std::vector<int32_t> mat_nuc, mat_param, mat_grid, me_tab;
    std::vector<double> me_cst;
    std::vector<DMatRec> mrec;
    std::vector<DMatLook> mlook;
    std::vector<double> rec_te;
    std::vector<uint16_t> bucket16;
    int lds_grid = -1, lds_nb = 0;
    std::vector<double> rec_t, rec_sc;
    std::vector<double> mat_dens;
    std::vector<mc_param_desc> param;
    std::vector<DNuc> nuc;
    std::vector<DGrid> grid;
    std::vector<double> grid_E, tab_s, tab_c, tab_t;
    std::vector<int32_t> bucket;
    std::vector<mc_estimator_desc> est;
    std::vector<int32_t> att_est, att_idx;
    std::vector<double> egrid;
    std::vector<DProb> prob;
    std::vector<mc_source_desc> src;
    std::vector<double> src_cdf;
    std::vector<int32_t> cdf_off;
    std::vector<int32_t> mat_off;      // per problem: first material index
    std::vector<uint64_t> jump;
    int n_bins = 0, n_params = 0, bin_stride = 1, par_stride = 1;
    bool has_tables = false;
    bool split_any = false;            // some problem's cells differ in importance
    bool comb_any = false;             // some problem that can split has a particle comb
    uint64_t comb_cap = 0;             // entries of a slot's comb bank
    std::vector<int32_t> comb_par;     // per problem (bank_max, teeth)
    double emin = 0, emax = 0;

    // device data
    DevBuf d_crec;
    DevBuf d_surf, d_cell, d_cell_surf, d_cell_sense, d_nbr_begin, d_nbr, d_mat, d_mat_nuc, d_mat_dens,
        d_mat_ns, d_mat_nc, d_mat_tconst, d_mat_param, d_mat_grid, d_me_tab, d_me_cst, d_mrec, d_mlook, d_rec_te, d_bucket16, d_rec_t, d_rec_sc, d_param, d_theta, d_mpar, d_nfold, d_mfold, d_nuc, d_grid,
        d_grid_E, d_bucket, d_tab_s,
        d_tab_c, d_tab_t, d_est, d_att_est, d_att_idx, d_egrid, d_prob, d_src, d_src_cdf, d_cdf_off, d_jump,
        d_run_hb, d_run_jd, d_hb_idx, d_binN, d_parN;
    std::vector<int32_t> h_hb_idx;    // host copy of the run's key-to-problem index
    // bank
    DevBuf b_x, b_y, b_z, b_u, b_v, b_w, b_E, b_wgt, b_sigt, b_rng, b_hid, b_cell, b_prob, b_surf, b_nfl,
        b_status, b_tcount, b_tbins, b_hist, b_gidx, b_S, b_H, b_Qw, b_cs, b_cc, b_lp, b_vmat, q_S, q_A, q_C, d_ctr, d_wgcnt, b_mat,
        d_wgxfl;
    // split families and the copy ring (problems that can split only)
    DevBuf c_x, c_y, c_z, c_u, c_v, c_w, c_E, c_wgt, c_rng, c_cell, c_size, d_comb;   // comb banks
    DevBuf b_fam, b_fcount, b_ppend, b_pfirst, d_wgheld, r_S, r_vmat, b_famG, r_x, r_y, r_z, r_u, r_v, r_w, r_E, r_wgt, r_rng, r_hid, r_cell, r_prob, r_fam,
        r_nfl;
    uint64_t ring_cap = 0;
    uint64_t ring_growths = 0;         // times a run doubled the copy ring
    uint64_t last_waits = 0;           // the last run's waits for ring room
    uint32_t h_active_cap = 0;
    uint64_t bank_cap = 0;
    bool bank_grad = false;
    // tallies
    DevBuf t_sum, t_sq, t_gout, t_gsum, t_gsq, t_trace;
    DevBuf t_psum, t_psq, t_pgsum, t_pgsq;   // per-workgroup partial moments
    uint64_t part_rows = 0;
    Counters* h_ctr = nullptr;   // pinned
    // graphs
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_iters[2] = {0, 0};
    uint64_t generation = 0, graph_gen[2] = {~0ull, ~0ull};
    const void* graph_trace[2] = {nullptr, nullptr};   // trace pointer captured in each graph
    uint64_t graph_tcount[2] = {0, 0};                 // and the trace count (record stride)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> kev;     // per-kernel timing events (profile mode)
    mc_kernel_stats kstats{};

    ~Impl()
    {
        for (auto& g : graph)
            if (g) (void)hipGraphExecDestroy(g);
        if (h_ctr) (void)hipHostFree(h_ctr);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        for (auto e : kev) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void build(const std::vector<const Problem*>& ps);
    void upload_static();
    void ensure_bank(bool grad);
    void ensure_partials();
    KArgs args(bool grad, mc_trace_record* dtrace, uint64_t trace_count) const;
    void launch_iteration(const KArgs& A, bool grad, hipStream_t st, hipEvent_t* ev);
    void run_loop(bool grad, KArgs& A);
    bool grow_ring(uint64_t head, uint64_t tail);
    void run(const mc_run_cfg* run, bool grad, const double* gout, mc_tally_out* out, double* grad_mean,
             double* grad_sigma);
    dim3 grid_dim() const { return dim3((unsigned)std::max(1, n_cu * cfg.grid_per_cu)); }
    // Launch shape of the standalone lookup kernel: with a bucket table in LDS
    // it takes most of a CU's LDS, so one workgroup of 1024 per CU.
    size_t lookup_lds() const
    {
        // the bucket table, then problem 0's materials
        return lds_grid >= 0 ? ((size_t)(lds_nb + 1) * 2 + 15) / 16 * 16 + sizeof(DLdsMat) * kLdsMats : 0;
    }
    dim3 lookup_grid() const { return lds_grid >= 0 ? dim3((unsigned)n_cu) : grid_dim(); }
    dim3 lookup_block() const { return lds_grid >= 0 ? dim3(1024) : dim3((unsigned)cfg.block); }
};

// ---------------------------------------------------------------- building
void Engine::Impl::build(const std::vector<const Problem*>& ps)
{
    for (const Problem* P : ps) deck_hist.push_back(P->n_histories);
    n_mat0 = (int)ps[0]->materials.size();
    std::map<std::vector<double>, int> grid_ids;
    std::map<std::string, int> nuc_ids;
    for (size_t p = 0; p < ps.size(); ++p) {
        const Problem& P = *ps[p];
        const mc_problem_desc& d = P.desc;
        const int surf_off = (int)surf.size(), cell_off = (int)cell.size(), cs_off = (int)cell_surf.size();
        const int mat_o = (int)mat.size(), mn_off = (int)mat_nuc.size(), est_off = (int)est.size();
        const int att_off = (int)att_est.size(), eg_off = (int)egrid.size(), src_off = (int)src.size();
        const int par_off = (int)param.size(), mp_off = (int)mat_param.size();
        // nuclides: deduplicated by content
        std::vector<int> nmap(d.n_nuclides);
        for (int n = 0; n < d.n_nuclides; ++n) {
            const mc_nuclide_desc& N = d.nuclides[n];
            std::string key((const char*)&N.A, sizeof(double));
            key.append((const char*)&N.kind, sizeof(int32_t));
            key.append((const char*)&N.kinematics, sizeof(int32_t));
            key.append((const char*)&N.xs_scatter, 3 * sizeof(double));
            if (N.kind == MC_NUC_TABLE) {
                key.append((const char*)(d.xs_energy + N.grid_begin), sizeof(double) * N.grid_count);
                key.append((const char*)(d.xs_scatter + N.grid_begin), sizeof(double) * N.grid_count);
                key.append((const char*)(d.xs_capture + N.grid_begin), sizeof(double) * N.grid_count);
                key.append((const char*)(d.xs_total + N.grid_begin), sizeof(double) * N.grid_count);
            }
            auto it = nuc_ids.find(key);
            if (it != nuc_ids.end()) { nmap[n] = it->second; continue; }
            DNuc D{};
            D.A = N.A;
            D.kind = N.kind;
            D.kin = N.kinematics;
            D.xs_s = N.xs_scatter;
            D.xs_c = N.xs_capture;
            D.xs_t = N.xs_total;
            D.grid = -1;
            if (N.kind == MC_NUC_TABLE) {
                has_tables = true;
                std::vector<double> E(d.xs_energy + N.grid_begin, d.xs_energy + N.grid_begin + N.grid_count);
                auto g = grid_ids.find(E);
                if (g == grid_ids.end()) {
                    DGrid G{};
                    G.e_off = (int)grid_E.size();
                    G.n = (int)E.size();
                    grid_E.insert(grid_E.end(), E.begin(), E.end());
                    if (std::is_sorted(E.begin(), E.end()) && E.front() > 0.0) {
                        // bucket table over the bit pattern: lo[k] = binary_search(low(k));
                        // exact for non-decreasing grids (repeated energies included:
                        // "largest i with E_i < x" is then one boundary)
                        const int64_t k0 = (int64_t)(double_to_bits(E.front()) >> kBucketShift);
                        const int64_t k1 = (int64_t)(double_to_bits(E.back()) >> kBucketShift);
                        G.key_base = k0;
                        G.nb = (int)(k1 - k0 + 1);
                        G.b_off = (int)bucket.size();
                        for (int64_t k = 0; k <= G.nb; ++k) {
                            const double low = bits_to_double((uint64_t)(k0 + k) << kBucketShift);
                            bucket.push_back(binary_search(low, E.data(), (int)E.size()));
                        }
                    } else {
                        G.key_base = 0;
                        G.nb = 0;
                        G.b_off = -1;                 // full search (grid_index)
                    }
                    grid.push_back(G);
                    g = grid_ids.emplace(E, (int)grid.size() - 1).first;
                    const double lo = *std::min_element(E.begin(), E.end());
                    const double hi = *std::max_element(E.begin(), E.end());
                    if (lo > 0.0 && (emin == 0 || lo < emin)) emin = lo;
                    emax = std::max(emax, hi);
                }
                D.grid = g->second;
                D.tab = (int)tab_s.size();
                D.n = N.grid_count;
                tab_s.insert(tab_s.end(), d.xs_scatter + N.grid_begin, d.xs_scatter + N.grid_begin + N.grid_count);
                tab_c.insert(tab_c.end(), d.xs_capture + N.grid_begin, d.xs_capture + N.grid_begin + N.grid_count);
                tab_t.insert(tab_t.end(), d.xs_total + N.grid_begin, d.xs_total + N.grid_begin + N.grid_count);
            }
            nuc.push_back(D);
            nmap[n] = (int)nuc.size() - 1;
            nuc_ids.emplace(key, nmap[n]);
        }
        for (int s = 0; s < d.n_surfaces; ++s) {
            mc_surface_desc S = d.surfaces[s];
            S.est_begin += att_off;
            surf.push_back(S);
        }
        // a crossing from a cell of importance I > 0 into one of higher
        // importance splits the particle (importance 0 only kills)
        double imin = 0.0, imax = 0.0;
        for (int c = 0; c < d.n_cells; ++c) {
            const double I = d.cells[c].importance;
            if (I > 0.0 && (imin == 0.0 || I < imin)) imin = I;
            imax = std::max(imax, I);
        }
        if (imin > 0.0 && imax > imin) {
            split_any = true;
            if (d.comb_teeth > 0) {
                // a bank holds up to bank_max - 1 copies plus what one walk adds
                comb_any = true;
                comb_cap = std::max<uint64_t>(comb_cap, std::max<uint64_t>((uint64_t)d.comb_bank_max - 1 + kCombSlack,
                                                                           (uint64_t)d.comb_teeth));
            }
        }
        for (int c = 0; c < d.n_cells; ++c) {
            mc_cell_desc C = d.cells[c];
            if (C.material >= 0) C.material += mat_o;
            C.surf_begin += cs_off;
            C.tl_begin += att_off;
            C.col_begin += att_off;
            cell.push_back(C);
        }
        int ncs = 0;
        for (int c = 0; c < d.n_cells; ++c) ncs = std::max(ncs, d.cells[c].surf_begin + d.cells[c].surf_count);
        for (int k = 0; k < ncs; ++k) {
            cell_surf.push_back(d.cell_surf[k] + surf_off);
            cell_sense.push_back(d.cell_sense[k]);
        }
        int nmn = 0;
        for (int m = 0; m < d.n_materials; ++m) {
            mc_material_desc M = d.materials[m];
            nmn = std::max(nmn, M.nuc_begin + M.nuc_count);
            M.nuc_begin += mn_off;
            M.param_begin += mp_off;
            mat.push_back(M);
        }
        for (int k = 0; k < nmn; ++k) {
            mat_nuc.push_back(nmap[d.mat_nuclide[k]]);
            mat_dens.push_back(d.mat_density[k]);
        }
        for (int k = 0; k < d.n_params; ++k) {
            mc_param_desc D = d.params[k];
            D.material += mat_o;
            D.cell += cell_off;
            param.push_back(D);
        }
        int nmp = 0;
        for (int m = 0; m < d.n_materials; ++m)
            nmp = std::max(nmp, d.materials[m].param_begin + d.materials[m].param_count);
        for (int k = 0; k < nmp; ++k) mat_param.push_back(d.mat_param[k] + par_off);
        for (int e = 0; e < d.n_estimators; ++e) {
            mc_estimator_desc E = d.estimators[e];
            E.egrid_begin += eg_off;
            E.bin_offset += n_bins;
            est.push_back(E);
        }
        int natt = 0;
        for (int s = 0; s < d.n_surfaces; ++s) natt = std::max(natt, d.surfaces[s].est_begin + d.surfaces[s].est_count);
        for (int c = 0; c < d.n_cells; ++c) {
            natt = std::max(natt, d.cells[c].tl_begin + d.cells[c].tl_count);
            natt = std::max(natt, d.cells[c].col_begin + d.cells[c].col_count);
        }
        for (int k = 0; k < natt; ++k) {
            att_est.push_back(d.att_estimator[k] + est_off);
            att_idx.push_back(d.att_index[k]);
        }
        int neg = 0;
        for (int e = 0; e < d.n_estimators; ++e)
            neg = std::max(neg, d.estimators[e].egrid_begin + d.estimators[e].egrid_count);
        egrid.insert(egrid.end(), d.egrid, d.egrid + neg);
        for (int s = 0; s < d.n_sources; ++s) {
            mc_source_desc S = d.sources[s];
            S.cell += cell_off;
            src.push_back(S);
        }
        cdf_off.push_back((int)src_cdf.size());
        src_cdf.insert(src_cdf.end(), d.source_cdf, d.source_cdf + d.n_sources + 1);
        DProb R{};
        R.cell_begin = cell_off;
        R.cell_count = d.n_cells;
        R.src_begin = src_off;
        R.src_count = d.n_sources;
        R.bin_begin = n_bins;
        R.bin_count = d.n_bins;
        R.param_begin = par_off;
        R.param_count = d.n_params;
        R.mat_begin = mat_o;
        R.mat_count = d.n_materials;
        R.seed = d.seed;
        comb_par.push_back(d.comb_bank_max);
        comb_par.push_back(d.comb_teeth);
        R.wr = d.wr;
        R.ws = d.ws;
        prob.push_back(R);
        mat_off.push_back(mat_o);
        n_bins += d.n_bins;
        n_params += d.n_params;
        bin_stride = std::max(bin_stride, d.n_bins);
        par_stride = std::max(par_stride, d.n_params);
    }
    // per material: the one grid shared by its tabulated nuclides (-1 none,
    // -2 several); per entry: table offset or the constant values
    for (size_t m = 0; m < mat.size(); ++m) {
        int g = -1;
        for (int k = 0; k < mat[m].nuc_count; ++k) {
            const DNuc& N = nuc[mat_nuc[mat[m].nuc_begin + k]];
            if (N.kind != MC_NUC_TABLE) continue;
            g = (g == -1 || g == N.grid) ? N.grid : -2;
        }
        mat_grid.push_back(g);
    }
    for (size_t j = 0; j < mat_nuc.size(); ++j) {
        const DNuc& N = nuc[mat_nuc[j]];
        me_tab.push_back(N.kind == MC_NUC_TABLE ? N.tab : -1);
        me_cst.push_back(N.xs_t);     // order t, s, c
        me_cst.push_back(N.xs_s);
        me_cst.push_back(N.xs_c);
    }
    // interleaved records for materials whose nuclides all sit on one grid,
    // shared by materials with the same nuclide list (cloned cells)
    std::map<std::vector<int32_t>, DMatRec> rec_ids;
    for (size_t m = 0; m < mat.size(); ++m) {
        DMatRec R{-1, -1, 0, 0, 0, 0};
        const int n = mat[m].nuc_count;
        bool all_tab = mat_grid[m] >= 0 && n > 0;
        std::vector<int32_t> key{mat_grid[m]};
        for (int k = 0; k < n && all_tab; ++k) {
            const int32_t tb = me_tab[mat[m].nuc_begin + k];
            all_tab = tb >= 0;
            key.push_back(tb);
        }
        if (all_tab) {
            auto it = rec_ids.find(key);
            if (it != rec_ids.end()) {
                R = it->second;
            } else {
                const DGrid& G = grid[mat_grid[m]];
                R.t_stride = (2 * n + 7) / 8 * 8;           // whole 64-byte units
                R.c_at = 2 * std::max(n, kFastNuc);
                R.sc_stride = (R.c_at + 2 * n + 7) / 8 * 8;
                R.t_off = (int64_t)rec_t.size();
                R.sc_off = (int64_t)rec_sc.size();
                rec_t.resize(rec_t.size() + (size_t)G.n * R.t_stride, 0.0);
                rec_sc.resize(rec_sc.size() + (size_t)G.n * R.sc_stride, 0.0);
                for (int i = 0; i < G.n; ++i) {
                    const int i2 = i + 1 < G.n ? i + 1 : i;
                    double* rt = rec_t.data() + R.t_off + (int64_t)i * R.t_stride;
                    double* rs = rec_sc.data() + R.sc_off + (int64_t)i * R.sc_stride;
                    for (int k = 0; k < n; ++k) {
                        const int tb = me_tab[mat[m].nuc_begin + k];
                        rt[2 * k] = tab_t[tb + i];
                        rt[2 * k + 1] = tab_t[tb + i2];
                        rs[2 * k] = tab_s[tb + i];
                        rs[2 * k + 1] = tab_s[tb + i2];
                        rs[R.c_at + 2 * k] = tab_c[tb + i];
                        rs[R.c_at + 2 * k + 1] = tab_c[tb + i2];
                    }
                }
                rec_ids.emplace(key, R);
            }
        }
        mrec.push_back(R);
        // fast Sigma_t path: records, at most kFastNuc nuclides, one point per bucket
        DMatLook L{};
        L.rec_off = -1;
        L.te_off = -1;
        L.grid = mat_grid[m];
        if (R.t_off >= 0 && n <= kFastNuc && grid[mat_grid[m]].b_off >= 0) {
            const DGrid& G = grid[mat_grid[m]];
            int crowd = 0;
            for (int k = 0; k < G.nb; ++k) crowd = std::max(crowd, bucket[G.b_off + k + 1] - bucket[G.b_off + k]);
            if (crowd <= 1) {
                L.key_base = G.key_base;
                L.scaled = mat[m].param_count > 0;
                L.rec_off = L.scaled ? R.sc_off : R.t_off;
                L.rec_stride = L.scaled ? R.sc_stride : R.t_stride;
                L.e_off = G.e_off;
                L.n = G.n;
                L.b_off = G.b_off;
                L.nb = G.nb;
                L.nuc_count = n;
                L.nuc_begin = mat[m].nuc_begin;
                if (!L.scaled) {
                    // records (E_i, sigma_t[i] per nuclide) for the LDS path
                    L.te_stride = (1 + n + 3) / 4 * 4;          // 32 or 64 bytes
                    L.te_off = (int64_t)rec_te.size();
                    rec_te.resize(rec_te.size() + (size_t)G.n * L.te_stride, 0.0);
                    for (int i = 0; i < G.n; ++i) {
                        double* r = rec_te.data() + L.te_off + (int64_t)i * L.te_stride;
                        r[0] = grid_E[G.e_off + i];
                        for (int k = 0; k < n; ++k) r[1 + k] = tab_t[me_tab[mat[m].nuc_begin + k] + i];
                    }
                }
            }
        }
        mlook.push_back(L);
    }
    if (rec_t.empty()) { rec_t.push_back(0.0); rec_sc.push_back(0.0); }
    rec_te.resize(rec_te.size() + kTeVals, 0.0);   // a record read past the last one stays in bounds
    // the grid whose bucket table the lookup kernels stage in LDS: the one of
    // the first material with an LDS path, if its table fits and indices fit u16
    for (const DMatLook& L : mlook) {
        if (L.te_off < 0) continue;
        const DGrid& G = grid[L.grid];
        if (G.n < 65535 && (size_t)(G.nb + 1) * 2 <= kLdsBucketBytes) {
            lds_grid = L.grid;
            lds_nb = G.nb;
            for (int k = 0; k <= G.nb; ++k) bucket16.push_back((uint16_t)(bucket[G.b_off + k] + 1));
        }
        break;
    }
    if (bucket16.empty()) bucket16.push_back(0);
    // per surface: cells that list it, in index order
    std::vector<std::vector<int>> lists(surf.size());
    for (size_t c = 0; c < cell.size(); ++c)
        for (int k = 0; k < cell[c].surf_count; ++k) {
            auto& L = lists[cell_surf[cell[c].surf_begin + k]];
            if (L.empty() || L.back() != (int)c) L.push_back((int)c);
        }
    nbr_begin.push_back(0);
    for (auto& L : lists) {
        nbr.insert(nbr.end(), L.begin(), L.end());
        nbr_begin.push_back((int)nbr.size());
    }
    // per-cell flight records when every cell has at most kRecSurfMax surfaces
    int smax = 0;
    for (const mc_cell_desc& C : cell) smax = std::max(smax, C.surf_count);
    if (!cell.empty() && smax <= kRecSurfMax) {
        crec_stride = (int32_t)(sizeof(DCellHead) + sizeof(DCellSurf) * std::max(smax, 1));
        crec.assign(cell.size() * (size_t)crec_stride / sizeof(double), 0.0);
        for (size_t c = 0; c < cell.size(); ++c) {
            char* r = (char*)crec.data() + c * (size_t)crec_stride;
            const mc_cell_desc& C = cell[c];
            const DCellHead H{C.material, C.surf_count, C.tl_begin, C.tl_count, C.col_begin, C.col_count, 0, 0};
            std::memcpy(r, &H, sizeof H);
            for (int k = 0; k < C.surf_count; ++k) {
                const int sf = cell_surf[C.surf_begin + k];
                DCellSurf S{sf, surf[sf].type, {}};
                for (int q = 0; q < 5; ++q) S.c[q] = surf[sf].c[q];
                std::memcpy(r + sizeof(DCellHead) + k * sizeof(DCellSurf), &S, sizeof S);
            }
        }
    }
    // LCG jump multipliers: jump[256 j + v] = g^(stride * v * 256^j), one per
    // byte value of a history key
    jump.assign(8 * 256, 1);
    uint64_t base = lcg_pow(kRnMult, kRnStride);
    for (int j = 0; j < 8; ++j) {
        for (int v = 1; v < 256; ++v) jump[256 * j + v] = (jump[256 * j + v - 1] * base) & kRnMask;
        base = lcg_pow(base, 256);
    }
}

void Engine::Impl::upload_static()
{
    upload(d_surf, surf, stream);
    upload(d_cell, cell, stream);
    upload(d_cell_surf, cell_surf, stream);
    upload(d_cell_sense, cell_sense, stream);
    upload(d_crec, crec, stream);
    upload(d_nbr_begin, nbr_begin, stream);
    upload(d_nbr, nbr, stream);
    upload(d_mat, mat, stream);
    upload(d_mat_nuc, mat_nuc, stream);
    upload(d_mat_dens, mat_dens, stream);
    upload(d_mat_ns, mat_dens, stream);
    upload(d_mat_nc, mat_dens, stream);
    upload(d_mat_param, mat_param, stream);
    upload(d_mat_grid, mat_grid, stream);
    upload(d_me_tab, me_tab, stream);
    upload(d_me_cst, me_cst, stream);
    upload(d_mrec, mrec, stream);
    upload(d_mlook, mlook, stream);
    upload(d_rec_te, rec_te, stream);
    upload(d_bucket16, bucket16, stream);
    upload(d_rec_t, rec_t, stream);
    upload(d_rec_sc, rec_sc, stream);
    upload(d_param, param, stream);
    upload(d_theta, std::vector<double>(std::max(1, n_params), 1.0), stream);
    upload(d_nuc, nuc, stream);
    upload(d_grid, grid, stream);
    upload(d_grid_E, grid_E, stream);
    upload(d_bucket, bucket, stream);
    upload(d_tab_s, tab_s, stream);
    upload(d_tab_c, tab_c, stream);
    upload(d_tab_t, tab_t, stream);
    upload(d_est, est, stream);
    upload(d_att_est, att_est, stream);
    upload(d_att_idx, att_idx, stream);
    upload(d_egrid, egrid, stream);
    upload(d_prob, prob, stream);
    upload(d_src, src, stream);
    upload(d_src_cdf, src_cdf, stream);
    upload(d_cdf_off, cdf_off, stream);
    upload(d_jump, jump, stream);
    if (comb_any) upload(d_comb, comb_par, stream);
    if (split_any) d_wgheld.alloc(sizeof(uint32_t) * 8192);
    d_run_hb.alloc(sizeof(uint64_t) * (prob.size() + 1));
    d_run_jd.alloc(sizeof(uint64_t) * prob.size());
    d_binN.alloc(sizeof(double) * std::max(1, n_bins));
    d_parN.alloc(sizeof(double) * std::max(1, n_params));
    d_hb_idx.alloc(sizeof(int32_t) * (kHbIdxMax + 3));
    t_sum.alloc(sizeof(double) * std::max(1, n_bins));
    t_sq.alloc(sizeof(double) * std::max(1, n_bins));
    t_gout.alloc(sizeof(double) * std::max(1, n_bins));
    t_gsum.alloc(sizeof(double) * std::max(1, n_params));
    t_gsq.alloc(sizeof(double) * std::max(1, n_params));
    d_ctr.alloc(sizeof(Counters));
    d_wgcnt.alloc(sizeof(uint32_t) * 2 * 8192);   // per-workgroup counts (grid <= 8192 blocks)
    d_wgxfl.alloc(sizeof(uint64_t) * 8192);
    HIP_CHECK(hipStreamSynchronize(stream));
}

constexpr uint64_t kRingPerSlot = 8;

// Copy-ring growth: entry k of a ring of capacity R sits at k mod R.
template <class T>
__global__ void k_ring_move(const T* __restrict__ src, T* __restrict__ dst, uint64_t head, uint64_t tail,
                            uint64_t cap0, uint64_t cap1)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = head + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < tail; k += stride)
        dst[k % cap1] = src[k % cap0];
}

// Doubles the copy ring, moving the copies [head, tail) to their places in
// the larger one; false (nothing changed) when the device has no memory for
// it or the ring holds 2^31 entries.
bool Engine::Impl::grow_ring(uint64_t head, uint64_t tail)
{
    const uint64_t R0 = ring_cap, R = 2 * R0;
    if (R > (1ull << 31)) return false;
    const uint64_t per = 8 * 10 + 4 * 4 + (bank_grad ? 8 * (uint64_t)par_stride + 4 : 0);
    size_t avail = 0, total = 0;
    HIP_CHECK(hipMemGetInfo(&avail, &total));
    if ((double)R * per > 0.8 * (double)avail) return false;   // old and new rings coexist while moving
    const dim3 g(1024), b(256);
    auto move = [&](DevBuf& buf, size_t esz, uint64_t rows) {
        DevBuf nb;
        nb.alloc(esz * R * rows);
        for (uint64_t k = 0; k < rows; ++k) {
            if (esz == 8)
                hipLaunchKernelGGL(k_ring_move<uint64_t>, g, b, 0, cur, buf.as<uint64_t>() + k * R0,
                                   nb.as<uint64_t>() + k * R, head, tail, R0, R);
            else
                hipLaunchKernelGGL(k_ring_move<uint32_t>, g, b, 0, cur, buf.as<uint32_t>() + k * R0,
                                   nb.as<uint32_t>() + k * R, head, tail, R0, R);
        }
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(cur));       // before the old buffer is freed
        std::swap(buf.p, nb.p);
        std::swap(buf.bytes, nb.bytes);
    };
    for (DevBuf* r : {&r_x, &r_y, &r_z, &r_u, &r_v, &r_w, &r_E, &r_wgt, &r_rng, &r_hid}) move(*r, 8, 1);
    for (DevBuf* r : {&r_cell, &r_prob, &r_fam, &r_nfl}) move(*r, 4, 1);
    if (bank_grad) {
        // rows of S are ring-major (stride = the capacity)
        move(r_S, 8, (uint64_t)par_stride);
        move(r_vmat, 4, 1);
    }
    ring_cap = R;
    ++ring_growths;
    ++generation;                                   // graphs hold the ring's addresses
    return true;
}

// Row stride (elements) of the parameter-major S and H arrays for a bank of
// cap slots: cap plus 32 KiB + 256 B, so that the rows of one slot do not sit a
// power of two apart (a fold updates up to eight of them at once).
inline uint64_t lr_stride(uint64_t cap) { return cap + 4128; }

// Device memory the comb banks of one engine may take.
constexpr uint64_t kCombBankBytes = 16ull << 30;

void Engine::Impl::ensure_bank(bool grad)
{
    uint64_t B = cfg.bank_capacity;
    if (comb_any) {
        // every slot carries a bank of comb_cap copies (76 B each)
        const uint64_t fit = kCombBankBytes / (comb_cap * 76ull) / 1024 * 1024;
        B = std::max<uint64_t>(1024, std::min(B, fit));
    }
    if (bank_cap == B && (!grad || bank_grad)) return;
    const bool realloc_all = bank_cap != B;
    if (realloc_all) {
        for (DevBuf* b : {&b_x, &b_y, &b_z, &b_u, &b_v, &b_w, &b_E, &b_wgt, &b_sigt}) b->alloc(8 * B);
        b_rng.alloc(8 * B);
        b_hid.alloc(8 * B);
        b_cell.alloc(4 * B);
        b_prob.alloc(4 * B);
        b_mat.alloc(4 * B);
        b_surf.alloc(4 * B);
        b_gidx.alloc(4 * B);
        b_nfl.alloc(4 * B);
        b_status.alloc(B);
        b_tcount.alloc(B);
        b_tbins.alloc(4 * (uint64_t)kTouched * B);
        b_hist.alloc(8 * B * (uint64_t)bin_stride);
        q_A.alloc(4 * B);
        q_C.alloc(4 * B);
        if (split_any) {
            // copies wait in the ring while no slot is free: 8 per bank slot
            const uint64_t R = kRingPerSlot * B;
            b_fam.alloc(4 * B);
            b_fcount.alloc(4 * B);
            b_ppend.alloc(4 * B);
            b_pfirst.alloc(4 * B);
            if (comb_any) {
                const uint64_t R = B * comb_cap;
                for (DevBuf* c : {&c_x, &c_y, &c_z, &c_u, &c_v, &c_w, &c_E, &c_wgt, &c_rng}) c->alloc(8 * R);
                c_cell.alloc(4 * R);
                c_size.alloc(4 * B);
            }
            for (DevBuf* r : {&r_x, &r_y, &r_z, &r_u, &r_v, &r_w, &r_E, &r_wgt, &r_rng, &r_hid}) r->alloc(8 * R);
            for (DevBuf* r : {&r_cell, &r_prob, &r_fam, &r_nfl}) r->alloc(4 * R);
            ring_cap = R;
        }
        bank_cap = B;
        bank_grad = false;
    }
    if (grad && (!bank_grad || realloc_all)) {
        b_S.alloc(8 * lr_stride(B) * (uint64_t)par_stride);
        b_H.alloc(8 * lr_stride(B) * (uint64_t)par_stride);
        b_Qw.alloc(8 * B);
        b_cs.alloc(8 * B);
        b_cc.alloc(8 * B);
        b_lp.alloc(8 * B);
        b_vmat.alloc(4 * B);
        q_S.alloc(4 * B);
        if (split_any) {
            // split copies carry S; families sum their members' gradients
            r_S.alloc(8 * ring_cap * (uint64_t)par_stride);
            r_vmat.alloc(4 * ring_cap);
            b_famG.alloc(8 * B * (uint64_t)par_stride);
        }
        bank_grad = true;
    }
    ++generation;
}

// One private row of tally moments per workgroup when that fits in 2^26
// entries per array, else a single shared row (stride 0).
void Engine::Impl::ensure_partials()
{
    const uint64_t wgs = (uint64_t)grid_dim().x;
    const uint64_t width = (uint64_t)std::max(1, std::max(n_bins, n_params));
    const uint64_t rows = wgs * width <= (1ull << 26) ? wgs : 1;
    if (rows == part_rows) return;
    t_psum.alloc(8 * rows * (uint64_t)std::max(1, n_bins));
    t_psq.alloc(8 * rows * (uint64_t)std::max(1, n_bins));
    t_pgsum.alloc(8 * rows * (uint64_t)std::max(1, n_params));
    t_pgsq.alloc(8 * rows * (uint64_t)std::max(1, n_params));
    part_rows = rows;
    ++generation;
}

KArgs Engine::Impl::args(bool grad, mc_trace_record* dtrace, uint64_t trace_count) const
{
    KArgs A{};
    A.surf = d_surf.as<mc_surface_desc>();
    A.cell = d_cell.as<mc_cell_desc>();
    A.cell_surf = d_cell_surf.as<int32_t>();
    A.crec = crec.empty() ? nullptr : d_crec.as<char>();
    A.crec_stride = crec_stride;
    A.cell_sense = d_cell_sense.as<int32_t>();
    A.nbr_begin = d_nbr_begin.as<int32_t>();
    A.nbr = d_nbr.as<int32_t>();
    A.mat = d_mat.as<mc_material_desc>();
    A.mat_nuc = d_mat_nuc.as<int32_t>();
    A.mat_dens = d_mat_dens.as<double>();
    A.mat_ns = d_mat_ns.as<double>();
    A.mat_nc = d_mat_nc.as<double>();
    A.mat_tconst = d_mat_tconst.as<double>();
    A.mat_param = d_mat_param.as<int32_t>();
    A.mat_grid = d_mat_grid.as<int32_t>();
    A.me_tab = d_me_tab.as<int32_t>();
    A.me_cst = d_me_cst.as<double>();
    A.mrec = d_mrec.as<DMatRec>();
    A.mlook = d_mlook.as<DMatLook>();
    A.rec_te = d_rec_te.as<double>();
    A.bucket16 = d_bucket16.as<uint16_t>();
    A.lds_grid = lds_grid;
    A.lds_nb = lds_nb;
    A.rec_t = d_rec_t.as<double>();
    A.rec_sc = d_rec_sc.as<double>();
    A.param = d_param.as<mc_param_desc>();
    A.theta = d_theta.as<double>();
    A.mpar = d_mpar.as<DParFold>();
    A.nfold = d_nfold.as<DNucFold>();
    A.mfold = d_mfold.as<int32_t>();
    A.nuc = d_nuc.as<DNuc>();
    A.grid = d_grid.as<DGrid>();
    A.grid_E = d_grid_E.as<double>();
    A.bucket = d_bucket.as<int32_t>();
    A.tab_s = d_tab_s.as<double>();
    A.tab_c = d_tab_c.as<double>();
    A.tab_t = d_tab_t.as<double>();
    A.est = d_est.as<mc_estimator_desc>();
    A.att_est = d_att_est.as<int32_t>();
    A.att_idx = d_att_idx.as<int32_t>();
    A.egrid = d_egrid.as<double>();
    A.prob = d_prob.as<DProb>();
    A.src = d_src.as<mc_source_desc>();
    A.src_cdf = d_src_cdf.as<double>();
    A.cdf_off = d_cdf_off.as<int32_t>();
    A.jump = d_jump.as<uint64_t>();
    A.jump_copy = lcg_pow(kRnMult, 1ull << kCopyJumpLog2);
    A.split_any = split_any ? 1 : 0;
    if (split_any) {
        A.fam = b_fam.as<int32_t>();
        A.fcount = b_fcount.as<int32_t>();
        A.ppend = b_ppend.as<int32_t>();
        A.pfirst = b_pfirst.as<int32_t>();
        if (comb_any) {
            A.bx = c_x.as<double>(); A.by = c_y.as<double>(); A.bz = c_z.as<double>();
            A.bu = c_u.as<double>(); A.bv = c_v.as<double>(); A.bw = c_w.as<double>();
            A.bE = c_E.as<double>();
            A.bwgt = c_wgt.as<double>();
            A.brng = c_rng.as<uint64_t>();
            A.bcell = c_cell.as<int32_t>();
            A.bsize = c_size.as<int32_t>();
            A.bcap = (uint32_t)comb_cap;
            A.comb = d_comb.as<int32_t>();
        }
        A.wg_held = d_wgheld.as<uint32_t>();
        A.cx = r_x.as<double>(); A.cy = r_y.as<double>(); A.cz = r_z.as<double>();
        A.cu = r_u.as<double>(); A.cv = r_v.as<double>(); A.cw = r_w.as<double>();
        A.cE = r_E.as<double>();
        A.cwgt = r_wgt.as<double>();
        A.crng = r_rng.as<uint64_t>();
        A.chid = r_hid.as<uint64_t>();
        A.ccell = r_cell.as<int32_t>();
        A.cprob = r_prob.as<int32_t>();
        A.cfam = r_fam.as<int32_t>();
        A.cnfl = r_nfl.as<uint32_t>();
        A.ccap = ring_cap;
    }
    A.run_hb = d_run_hb.as<uint64_t>();
    A.run_jd = d_run_jd.as<uint64_t>();
    A.hb_idx = d_hb_idx.as<int32_t>();
    A.n_prob = (int32_t)prob.size();
    A.n_mat = (int32_t)mat.size();
    A.n_cell = (int32_t)cell.size();
    A.n_surf = (int32_t)surf.size();
    A.n_cell_surf = (int32_t)cell_surf.size();
    A.bin_stride = bin_stride;
    A.par_stride = par_stride;
    A.sstride = lr_stride(bank_cap);
    A.cap = bank_cap;
    A.x = b_x.as<double>(); A.y = b_y.as<double>(); A.z = b_z.as<double>();
    A.u = b_u.as<double>(); A.v = b_v.as<double>(); A.w = b_w.as<double>();
    A.E = b_E.as<double>();
    A.wgt = b_wgt.as<double>();
    A.sigt = b_sigt.as<double>();
    A.rng = b_rng.as<uint64_t>();
    A.hid = b_hid.as<uint64_t>();
    A.pcell = b_cell.as<int32_t>();
    A.pprob = b_prob.as<int32_t>();
    A.pmat = b_mat.as<int32_t>();
    A.psurf = b_surf.as<int32_t>();
    A.gidx = b_gidx.as<int32_t>();
    A.nflight = b_nfl.as<uint32_t>();
    A.status = b_status.as<uint8_t>();
    A.tcount = b_tcount.as<uint8_t>();
    A.tbins = b_tbins.as<uint32_t>();
    A.wg_cnt = d_wgcnt.as<uint32_t>();
    A.wg_xfl = d_wgxfl.as<uint64_t>();
    // a run that can split takes one flight per particle per iteration: its
    // copies wait in the ring for free slots, which appear once per iteration
    A.events_per_launch = split_any ? 1u : (uint32_t)(grad ? cfg.grad_events_per_launch : cfg.events_per_launch);
    A.hist = b_hist.as<double>();
    A.S = grad ? b_S.as<double>() : nullptr;
    A.H = grad ? b_H.as<double>() : nullptr;
    A.Qw = grad ? b_Qw.as<double>() : nullptr;
    A.cs = grad ? b_cs.as<double>() : nullptr;
    A.cc = grad ? b_cc.as<double>() : nullptr;
    A.lpend = grad ? b_lp.as<double>() : nullptr;
    A.vmat = grad ? b_vmat.as<uint32_t>() : nullptr;
    A.rS = grad && split_any ? r_S.as<double>() : nullptr;
    A.rvmat = grad && split_any ? r_vmat.as<uint32_t>() : nullptr;
    A.famG = grad && split_any ? b_famG.as<double>() : nullptr;
    A.qS = grad ? q_S.as<uint32_t>() : nullptr;
    A.qA = q_A.as<uint32_t>();
    A.qC = q_C.as<uint32_t>();
    A.ctr = d_ctr.as<Counters>();
    A.psum = t_psum.as<double>();
    A.psq = t_psq.as<double>();
    A.bin_row = part_rows > 1 ? (uint64_t)n_bins : 0;
    A.gout = t_gout.as<double>();
    A.pgsum = t_pgsum.as<double>();
    A.pgsq = t_pgsq.as<double>();
    A.par_row = part_rows > 1 ? (uint64_t)n_params : 0;
    A.trace = dtrace;
    A.trace_count = trace_count;
    return A;
}

// Kernels of one event iteration (the first entries of mc_kernel_stats).
constexpr int kIterKernels = 5;
constexpr int kDrainSlot = 5;

// One event iteration.  With ev (2 * MC_N_KERNELS events) each kernel is
// bracketed by a start/stop event pair for per-kernel timing.
void Engine::Impl::launch_iteration(const KArgs& A, bool grad, hipStream_t st, hipEvent_t* ev)
{
    const dim3 g = grid_dim(), b((unsigned)cfg.block);
    auto mark = [&](int k, int edge) {
        if (ev) HIP_CHECK(hipEventRecord(ev[2 * k + edge], st));
    };
    mark(0, 0);
    if (split_any) {
        if (grad) hipLaunchKernelGGL(k_resume<true>, g, b, 0, st, A);
        else hipLaunchKernelGGL(k_resume<false>, g, b, 0, st, A);
    }
    hipLaunchKernelGGL(k_refill_count, g, b, 0, st, A);
    mark(0, 1);
    mark(1, 0);
    if (grad) hipLaunchKernelGGL(k_refill_scatter<true>, g, b, 0, st, A);
    else hipLaunchKernelGGL(k_refill_scatter<false>, g, b, 0, st, A);
    mark(1, 1);
    mark(2, 0);
    if (has_tables) hipLaunchKernelGGL(k_xs_lookup, g, b, 0, st, A, 0);
    mark(2, 1);
    mark(3, 0);
    // every forward run, and gradient runs of several flights per iteration,
    // collide in the flight kernel
    const bool fused = !grad || A.events_per_launch > 1;
    const size_t acc_lds = sizeof(double) * kAccEntries * b.x;   // the fused gradient kernel's S sums
    // fused runs read the geometry from an LDS copy when it fits; forward
    // runs otherwise from the per-cell flight records when there are (in the
    // gradient kernel the record's pointer costs more registers than the
    // reads it saves)
    const bool lds_geo = A.n_cell <= kGeoCells && A.n_surf <= kGeoSurfs && A.n_cell_surf <= kGeoCellSurf;
    const bool rec_geo = !lds_geo && !grad && A.crec;
    if (has_tables) {
        if (!grad && lds_geo) hipLaunchKernelGGL((k_flight<true, false, true, kGeoLds>), g, b, 0, st, A);
        else if (rec_geo) hipLaunchKernelGGL((k_flight<true, false, true, kGeoRec>), g, b, 0, st, A);
        else if (!grad) hipLaunchKernelGGL((k_flight<true, false, true, kGeoArrays>), g, b, 0, st, A);
        else if (fused && lds_geo) hipLaunchKernelGGL((k_flight<true, true, true, kGeoLds>), g, b, acc_lds, st, A);
        else if (fused) hipLaunchKernelGGL((k_flight<true, true, true, kGeoArrays>), g, b, acc_lds, st, A);
        else hipLaunchKernelGGL((k_flight<true, true, false, kGeoArrays>), g, b, 0, st, A);
    } else {
        if (!grad && lds_geo) hipLaunchKernelGGL((k_flight<false, false, true, kGeoLds>), g, b, 0, st, A);
        else if (rec_geo) hipLaunchKernelGGL((k_flight<false, false, true, kGeoRec>), g, b, 0, st, A);
        else if (!grad) hipLaunchKernelGGL((k_flight<false, false, true, kGeoArrays>), g, b, 0, st, A);
        else if (fused && lds_geo) hipLaunchKernelGGL((k_flight<false, true, true, kGeoLds>), g, b, acc_lds, st, A);
        else if (fused) hipLaunchKernelGGL((k_flight<false, true, true, kGeoArrays>), g, b, acc_lds, st, A);
        else hipLaunchKernelGGL((k_flight<false, true, false, kGeoArrays>), g, b, 0, st, A);
    }
    // gradient runs cross surfaces in a kernel of their own (timed with the flight)
    if (grad) {
        if (has_tables) hipLaunchKernelGGL(k_cross<true>, g, b, 0, st, A);
        else hipLaunchKernelGGL(k_cross<false>, g, b, 0, st, A);
    }
    mark(3, 1);
    mark(4, 0);
    if (!fused) {
        if (has_tables) hipLaunchKernelGGL((k_collision<true, true>), g, b, 0, st, A);
        else hipLaunchKernelGGL((k_collision<false, true>), g, b, 0, st, A);
    }
    mark(4, 1);
    // event counts for the statistics of profiled runs (outside the timed pairs)
    if (ev) hipLaunchKernelGGL(k_count_events, dim3(1), dim3(256), 0, st, A.wg_cnt, g.x, A.ctr);
}

void Engine::Impl::run_loop(bool grad, KArgs& A)
{
    const int gi = grad ? 1 : 0;
    const int K = std::max(1, cfg.iters_per_graph);
    const bool profile = cfg.profile != 0;
    const bool use_graph = cfg.use_graph && !profile;
    auto ensure_graph = [&]() {
        if (!graph[gi] || graph_gen[gi] != generation || graph_iters[gi] != K ||
            graph_trace[gi] != (const void*)A.trace || graph_tcount[gi] != A.trace_count) {
            if (graph[gi]) { HIP_CHECK(hipGraphExecDestroy(graph[gi])); graph[gi] = nullptr; }
            hipGraph_t gr;
            HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < K; ++k) launch_iteration(A, grad, stream, nullptr);
            HIP_CHECK(hipStreamEndCapture(stream, &gr));
            HIP_CHECK(hipGraphInstantiate(&graph[gi], gr, nullptr, nullptr, 0));
            HIP_CHECK(hipGraphDestroy(gr));
            graph_gen[gi] = generation;
            graph_iters[gi] = K;
            graph_trace[gi] = (const void*)A.trace;
            graph_tcount[gi] = A.trace_count;
        }
    };
    if (profile && kev.empty()) {
        kev.resize(2 * MC_N_KERNELS);
        for (auto& e : kev) HIP_CHECK(hipEventCreate(&e));
    }
    const auto t0 = std::chrono::steady_clock::now();
    const double max_s = getenv("MC_MAX_SECONDS") ? atof(getenv("MC_MAX_SECONDS")) : 3600.0;
    const dim3 g = grid_dim(), b((unsigned)cfg.block);
    uint64_t drain_flights = 0, drain_coll = 0;
    auto refill = [&]() {
        if (split_any) {
            if (grad) hipLaunchKernelGGL(k_resume<true>, g, b, 0, cur, A);
            else hipLaunchKernelGGL(k_resume<false>, g, b, 0, cur, A);
        }
        hipLaunchKernelGGL(k_refill_count, g, b, 0, cur, A);
        if (grad) hipLaunchKernelGGL(k_refill_scatter<true>, g, b, 0, cur, A);
        else hipLaunchKernelGGL(k_refill_scatter<false>, g, b, 0, cur, A);
    };
    auto read_ctr = [&]() {
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(h_ctr, A.ctr, sizeof(Counters), hipMemcpyDeviceToHost, cur));
        HIP_CHECK(hipStreamSynchronize(cur));
    };
    // Back-pressure of the copy ring: a split whose copies do not all fit
    // leaves its particle waiting at the surface until they do.  Once a chunk
    // has seen a particle wait, the ring doubles (its copies move to their
    // places in the larger ring).  A refill that queued no particle while the
    // ring holds copies found every slot waiting: the ring must grow, and a
    // device without the memory for it ends the run.
    uint64_t waits_seen = 0;                     // the run's counters start at zero
    // returns true when the run is over
    auto settle = [&]() -> bool {
        const uint64_t held = split_any ? h_ctr->ctail - (h_ctr->chead + h_ctr->cplaced) : 0;
        const bool stuck = h_ctr->nA == 0 && held > 0;
        if (h_ctr->nA == 0 && !stuck) return true;
        if (split_any && (stuck || h_ctr->waits != waits_seen)) {
            waits_seen = h_ctr->waits;
            if (grow_ring(h_ctr->chead + h_ctr->cplaced, h_ctr->ctail)) {
                A = args(grad, A.trace, A.trace_count);
            } else if (stuck) {
                throw Error(MC_ERR_CAPACITY, "the copy ring (" + std::to_string(ring_cap) +
                                                 " entries) is full, every particle waits for room in it, "
                                                 "and the device has no memory to grow it");
            }
        }
        return false;
    };
    for (uint64_t chunk = 0;; ++chunk) {
        // The tail of a run: every history has started and few particles are
        // left, so an iteration would mostly pay launch latencies.  One drain
        // launch carries them to the end of their histories instead.
        const bool drain = chunk > 0 && cfg.drain_divisor > 0 &&
                           h_ctr->next_hist + h_ctr->claimed >= h_ctr->end_hist &&
                           (uint64_t)h_ctr->nA * (uint64_t)cfg.drain_divisor <= (uint64_t)h_active_cap;
        if (drain) {
            // the roll-over in the refill below adds the last iteration's counts
            const uint64_t tA0 = h_ctr->tA + h_ctr->nA, tC0 = h_ctr->tC + h_ctr->nC, x0 = h_ctr->xfl;
            refill();
            if (profile) HIP_CHECK(hipEventRecord(kev[2 * kDrainSlot], cur));
            const uint32_t kMaxEvents = 1u << 14;
            if (has_tables) {
                if (grad) hipLaunchKernelGGL((k_drain<true, true>), g, b, 0, cur, A, kMaxEvents);
                else hipLaunchKernelGGL((k_drain<true, false>), g, b, 0, cur, A, kMaxEvents);
            } else {
                if (grad) hipLaunchKernelGGL((k_drain<false, true>), g, b, 0, cur, A, kMaxEvents);
                else hipLaunchKernelGGL((k_drain<false, false>), g, b, 0, cur, A, kMaxEvents);
            }
            if (profile) HIP_CHECK(hipEventRecord(kev[2 * kDrainSlot + 1], cur));
            // the drain counted its own events; the refill after it must not
            // count the drained particles again
            hipLaunchKernelGGL(k_drain_close, dim3(1), dim3(64), 0, cur, A.ctr);
            refill();
            read_ctr();
            drain_flights += h_ctr->tA - tA0 - (h_ctr->xfl - x0);
            drain_coll += h_ctr->tC - tC0;
            if (profile) {
                float ms = 0.f;
                HIP_CHECK(hipEventSynchronize(kev[2 * kDrainSlot + 1]));
                HIP_CHECK(hipEventElapsedTime(&ms, kev[2 * kDrainSlot], kev[2 * kDrainSlot + 1]));
                kstats.seconds[kDrainSlot] += 1e-3 * ms;
                kstats.launches[kDrainSlot] += 1;
            }
            if (settle()) break;
            // histories longer than the drain's event limit: back to iterations,
            // whose first roll-over must not count this refill's particles
            hipLaunchKernelGGL(k_drain_close, dim3(1), dim3(64), 0, cur, A.ctr);
            continue;
        }
        if (use_graph) {
            ensure_graph();
            HIP_CHECK(hipGraphLaunch(graph[gi], cur));
        } else if (profile) {
            for (int k = 0; k < K; ++k) {
                launch_iteration(A, grad, cur, kev.data());
                HIP_CHECK(hipEventSynchronize(kev[2 * kIterKernels - 1]));
                for (int j = 0; j < kIterKernels; ++j) {
                    // no lookup kernel without tables, no collision kernel when the
                    // flight kernel collides
                    if ((j == 2 && !has_tables) || (j == 4 && (!grad || A.events_per_launch > 1))) continue;
                    float ms = 0.f;
                    HIP_CHECK(hipEventElapsedTime(&ms, kev[2 * j], kev[2 * j + 1]));
                    kstats.seconds[j] += 1e-3 * ms;
                    kstats.launches[j] += 1;
                }
                kstats.items[0] += h_active_cap;
                kstats.items[1] += h_active_cap;
            }
        } else {
            for (int k = 0; k < K; ++k) launch_iteration(A, grad, cur, nullptr);
        }
        read_ctr();
        if (settle()) break;
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (el > max_s)
            throw Error(MC_ERR_CAPACITY, "event loop exceeded " + std::to_string(max_s) +
                                         " s (MC_MAX_SECONDS); live particles: " + std::to_string(h_ctr->nA));
    }
    if (profile) {
        // event counts: the lookup kernel saw every live particle of the
        // iterations, the flight kernel took their flights (several per
        // particle when it collides and flies on), then collisions; the
        // drain's flights are its own
        const uint64_t it_flights = h_ctr->tA - drain_flights;
        kstats.items[2] += has_tables ? it_flights - h_ctr->xfl : 0;
        kstats.items[3] += it_flights;
        kstats.items[4] += h_ctr->tC - drain_coll;
        kstats.items[kDrainSlot] += drain_flights;
    }
}

void Engine::Impl::run(const mc_run_cfg* run, bool grad, const double* gout, mc_tally_out* out,
                       double* grad_mean, double* grad_sigma)
{
    HIP_CHECK(hipSetDevice(device));
    if (grad && comb_any)
        throw Error(MC_ERR_UNSUPPORTED, "gradients of problems with a particle comb: the comb's selection "
                                        "of copies has no likelihood-ratio term here");
    cur = (run && run->stream) ? (hipStream_t)run->stream : stream;
    const bool dev = run && run->device_ptrs;
    const int np = (int)prob.size();
    std::vector<uint64_t> nh(np), hb(np + 1, 0);
    for (int p = 0; p < np; ++p) {
        nh[p] = (run && run->n_histories) ? run->n_histories[p] : deck_hist[p];
        if (nh[p] < 2) throw Error(MC_ERR_ARGUMENT, "each problem needs at least 2 histories");
        hb[p + 1] = hb[p] + nh[p];
    }
    ensure_bank(grad);
    // theta -> per-reaction densities (as the oracle: N_s = theta * N)
    std::vector<double> th(std::max(1, n_params), 1.0);
    if (run && run->theta)
        for (int k = 0; k < n_params; ++k) th[k] = run->theta[k];
    std::vector<double> ns(mat_dens), nc(mat_dens);
    for (int k = 0; k < n_params; ++k) {
        const mc_param_desc& D = param[k];
        const int j = mat[D.material].nuc_begin + D.nuc_slot;
        if (D.reactions & MC_REACTION_SCATTER) ns[j] = th[k] * mat_dens[j];
        if (D.reactions & MC_REACTION_CAPTURE) nc[j] = th[k] * mat_dens[j];
    }
    upload(d_theta, th, cur);
    upload(d_mat_ns, ns, cur);
    upload(d_mat_nc, nc, cur);
    // Sigma_t of materials whose nuclides all have constant cross sections,
    // summed as material_sigt sums it (same terms, same order)
    std::vector<double> tconst(std::max<size_t>(1, mat.size()), std::nan(""));
    for (size_t m = 0; m < mat.size(); ++m) {
        const mc_material_desc& M = mat[m];
        bool cst = true;
        for (int k = 0; k < M.nuc_count; ++k) cst = cst && me_tab[M.nuc_begin + k] < 0;
        if (!cst) continue;
        if (M.param_count == 0) {
            double t = 0.0;
            for (int k = 0; k < M.nuc_count; ++k) {
                const int j = M.nuc_begin + k;
                t += me_cst[3 * j + XS_T] * mat_dens[j];
            }
            tconst[m] = t;
        } else {
            double sv = 0.0, cv = 0.0;
            for (int k = 0; k < M.nuc_count; ++k) {
                const int j = M.nuc_begin + k;
                sv += me_cst[3 * j + XS_S] * ns[j];
                cv += me_cst[3 * j + XS_C] * nc[j];
            }
            tconst[m] = sv + cv;
        }
    }
    upload(d_mat_tconst, tconst, cur);
    // the collision fold's packed parameter table (1/theta as the kernels would
    // divide it: IEEE division on both sides)
    std::vector<int> owner(std::max(1, n_params), 0);
    for (int p = 0; p < np; ++p)
        for (int k = prob[p].param_begin; k < prob[p].param_begin + prob[p].param_count; ++k) owner[k] = p;
    std::vector<DParFold> mpar(std::max<size_t>(1, mat_param.size()));
    for (size_t m = 0; m < mat.size(); ++m)
        for (int q = 0; q < mat[m].param_count; ++q) {
            const int kg = mat_param[mat[m].param_begin + q];
            const mc_param_desc& D = param[kg];
            DParFold& F = mpar[mat[m].param_begin + q];
            F.kl = kg - prob[owner[kg]].param_begin;
            F.slot = D.nuc_slot;
            F.reactions = D.reactions;
            F.pad = 0;
            F.N = mat_dens[mat[m].nuc_begin + D.nuc_slot];
            F.inv_theta = 1.0 / th[kg];
        }
    upload(d_mpar, mpar, cur);
    std::vector<DNucFold> nfold(std::max<size_t>(1, mat_nuc.size()), DNucFold{-1, -1, 0.0});
    std::vector<int32_t> mfold(std::max<size_t>(1, mat.size()), 0);
    for (size_t m = 0; m < mat.size(); ++m) {
        bool fast = mat[m].param_count > 0 && mat[m].nuc_count <= 8;
        for (int q = 0; q < mat[m].param_count && fast; ++q) {
            const DParFold& F = mpar[mat[m].param_begin + q];
            DNucFold& N = nfold[mat[m].nuc_begin + F.slot];
            if (F.reactions == MC_REACTION_SCATTER && N.kl_s < 0) {
                N.kl_s = F.kl;
                N.inv_theta_s = F.inv_theta;
            } else if (F.reactions == MC_REACTION_CAPTURE && N.kl_c < 0) {
                N.kl_c = F.kl;
            } else {
                fast = false;
            }
        }
        mfold[m] = fast ? 1 : 0;
    }
    upload(d_nfold, nfold, cur);
    upload(d_mfold, mfold, cur);
    HIP_CHECK(hipMemcpyAsync(d_run_hb.p, hb.data(), sizeof(uint64_t) * (np + 1), hipMemcpyHostToDevice, cur));
    // history keys: problem p's j-th history of the run has run key hb[p] + j
    // and key key_offset[p] + j; its stream start times g^(stride (key_offset
    // - hb)) moves it from the one to the other (exponents wrap mod 2^64, a
    // multiple of the generator's period 2^61)
    std::vector<uint64_t> jd(np, 1);
    if (run && run->key_offset)
        for (int p = 0; p < np; ++p) jd[p] = lcg_pow(kRnMult, kRnStride * (run->key_offset[p] - hb[p]));
    HIP_CHECK(hipMemcpyAsync(d_run_jd.p, jd.data(), sizeof(uint64_t) * np, hipMemcpyHostToDevice, cur));
    {
        // key -> problem index: blocks of 2^sh keys, at most kHbIdxMax of them;
        // a key's problem lies between those of its block's two ends
        const uint64_t total = hb[np];
        uint32_t sh = 0;
        while ((total >> sh) >= kHbIdxMax) ++sh;
        const uint32_t nj = (uint32_t)(total >> sh) + 1;
        std::vector<int32_t>& idx = h_hb_idx;
        idx.assign(nj + 3, 0);
        idx[0] = (int32_t)sh;
        idx[1] = (int32_t)nj;
        int p = 0;
        for (uint32_t j = 0; j <= nj; ++j) {
            uint64_t key = (uint64_t)j << sh;
            if (total > 0 && key > total - 1) key = total - 1;
            while (p + 1 < (int)np && hb[p + 1] <= key) ++p;
            idx[2 + j] = p;
        }
        HIP_CHECK(hipMemcpyAsync(d_hb_idx.p, idx.data(), sizeof(int32_t) * (nj + 3), hipMemcpyHostToDevice, cur));
    }
    if (grad) {
        if (dev) {
            HIP_CHECK(hipMemcpyAsync(t_gout.p, gout, sizeof(double) * n_bins, hipMemcpyDeviceToDevice, cur));
        } else {
            std::vector<double> g(gout, gout + n_bins);
            HIP_CHECK(hipMemcpyAsync(t_gout.p, g.data(), sizeof(double) * n_bins, hipMemcpyHostToDevice, cur));
        }
    }
    // trace buffer: the caller's device buffer, or the engine's, copied back
    mc_trace_record* dtrace = nullptr;
    const uint64_t tcount = (run && run->trace) ? run->trace_count : 0;
    if (tcount) {
        const size_t bytes = sizeof(mc_trace_record) * tcount * (size_t)np;
        if (dev) {
            dtrace = run->trace;
        } else {
            if (t_trace.bytes < bytes) { t_trace.alloc(bytes); ++generation; }
            dtrace = t_trace.as<mc_trace_record>();
        }
        hipLaunchKernelGGL(k_fill_u8, dim3(1024), dim3(256), 0, cur, (uint8_t*)dtrace, (uint64_t)bytes, (uint8_t)0);
    }
    ensure_partials();
    KArgs A = args(grad, dtrace, tcount);     // run_loop renews it when the copy ring grows
    const dim3 fg(1024), fb(256);
    const uint64_t nbp = part_rows * (uint64_t)std::max(1, n_bins), npp = part_rows * (uint64_t)std::max(1, n_params);
    hipLaunchKernelGGL(k_fill_f64, fg, fb, 0, cur, A.psum, nbp, 0.0);
    hipLaunchKernelGGL(k_fill_f64, fg, fb, 0, cur, A.psq, nbp, 0.0);
    hipLaunchKernelGGL(k_fill_f64, fg, fb, 0, cur, A.hist, bank_cap * (uint64_t)bin_stride, 0.0);
    hipLaunchKernelGGL(k_fill_u8, fg, fb, 0, cur, A.tcount, bank_cap, (uint8_t)0);
    if (grad) {
        hipLaunchKernelGGL(k_fill_f64, fg, fb, 0, cur, A.pgsum, npp, 0.0);
        hipLaunchKernelGGL(k_fill_f64, fg, fb, 0, cur, A.pgsq, npp, 0.0);
        // every history's closeout clears the S and H entries it used, so the
        // bank starts each run cleared
        hipLaunchKernelGGL(k_fill_f64, fg, fb, 0, cur, A.S, lr_stride(bank_cap) * (uint64_t)par_stride, 0.0);
        hipLaunchKernelGGL(k_fill_f64, fg, fb, 0, cur, A.H, lr_stride(bank_cap) * (uint64_t)par_stride, 0.0);
        if (split_any)
            hipLaunchKernelGGL(k_fill_f64, fg, fb, 0, cur, A.famG, bank_cap * (uint64_t)par_stride, 0.0);
    }
    HIP_CHECK(hipEventRecord(ev0, cur));
    const bool plain = run && run->mode == MC_RUN_PLAIN;
    uint64_t flights = 0;
    uint32_t err_lost = 0, err_split = 0, err_bank = 0;
    uint64_t waits = 0;
    for (int p = 0; p < (plain ? np : 1); ++p) {
        Counters c{};
        c.next_hist = plain ? hb[p] : 0;
        c.end_hist = plain ? hb[p + 1] : hb[np];
        // slots worth scanning: no more than the histories of this run (and,
        // in gradient runs, than a share of the bank)
        const uint64_t live = grad ? std::max<uint64_t>(1024, bank_cap / (uint64_t)cfg.grad_live_divisor) : bank_cap;
        c.active_cap = (uint32_t)std::min<uint64_t>(std::min(live, bank_cap), c.end_hist - c.next_hist);
        h_active_cap = c.active_cap;
        HIP_CHECK(hipMemcpyAsync(A.ctr, &c, sizeof(Counters), hipMemcpyHostToDevice, cur));
        hipLaunchKernelGGL(k_fill_u8, fg, fb, 0, cur, A.status, bank_cap, (uint8_t)ST_EMPTY);
        hipLaunchKernelGGL(k_fill_u8, fg, fb, 0, cur, (uint8_t*)A.wg_xfl, sizeof(uint64_t) * 8192, (uint8_t)0);
        if (split_any) {
            hipLaunchKernelGGL(k_fill_i32, fg, fb, 0, cur, A.fam, bank_cap, -1);
            hipLaunchKernelGGL(k_fill_i32, fg, fb, 0, cur, A.fcount, bank_cap, 0);
            hipLaunchKernelGGL(k_fill_i32, fg, fb, 0, cur, A.ppend, bank_cap, 0);
            if (comb_any) hipLaunchKernelGGL(k_fill_i32, fg, fb, 0, cur, A.bsize, bank_cap, 0);
        }
        run_loop(grad, A);
        flights += h_ctr->tA;      // every entry of queue A is one flight
        waits += h_ctr->waits;
        err_bank += h_ctr->err_bank;
        err_lost += h_ctr->err_lost;
        err_split += h_ctr->err_split;
    }
    HIP_CHECK(hipEventRecord(ev1, cur));
    HIP_CHECK(hipEventSynchronize(ev1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
    last_waits = waits;
    if (err_bank)
        throw Error(MC_ERR_CAPACITY, std::to_string(err_bank) + " split copies found a history's comb bank full (" +
                                         std::to_string(comb_cap) + " entries: bank_max - 1 + " +
                                         std::to_string(kCombSlack) + ")");
    if (err_lost) throw Error(MC_ERR_LOST_PARTICLE, std::to_string(err_lost) + " particle(s) lost");
    if (err_split)
        throw Error(MC_ERR_ARGUMENT, "a particle split in a problem without an importance rise");
    // reduce the per-workgroup partial moments (workgroups in order)
    const dim3 rg(256), rb(256);
    hipLaunchKernelGGL(k_reduce_rows, rg, rb, 0, cur, A.psum, part_rows, (uint64_t)n_bins, t_sum.as<double>());
    hipLaunchKernelGGL(k_reduce_rows, rg, rb, 0, cur, A.psq, part_rows, (uint64_t)n_bins, t_sq.as<double>());
    if (grad) {
        hipLaunchKernelGGL(k_reduce_rows, rg, rb, 0, cur, A.pgsum, part_rows, (uint64_t)n_params,
                           t_gsum.as<double>());
        hipLaunchKernelGGL(k_reduce_rows, rg, rb, 0, cur, A.pgsq, part_rows, (uint64_t)n_params,
                           t_gsq.as<double>());
    }
    if (dev) {
        // Estimator::end_cycle / end_simulation (Estimator.cpp:347-367) on the
        // device, into the caller's device arrays: the same operations as the
        // host path below, so the same bits
        std::vector<double> binN(std::max(1, n_bins)), parN(std::max(1, n_params));
        for (int p = 0; p < np; ++p) {
            for (int b = prob[p].bin_begin; b < prob[p].bin_begin + prob[p].bin_count; ++b) binN[b] = (double)nh[p];
            for (int k = prob[p].param_begin; k < prob[p].param_begin + prob[p].param_count; ++k)
                parN[k] = (double)nh[p];
        }
        HIP_CHECK(hipMemcpyAsync(d_binN.p, binN.data(), sizeof(double) * binN.size(), hipMemcpyHostToDevice, cur));
        HIP_CHECK(hipMemcpyAsync(d_parN.p, parN.data(), sizeof(double) * parN.size(), hipMemcpyHostToDevice, cur));
        if (out && n_bins > 0)
            hipLaunchKernelGGL(k_moments, dim3(64), dim3(256), 0, cur, t_sum.as<double>(), t_sq.as<double>(),
                               d_binN.as<double>(), (uint64_t)n_bins, out->sum, out->sq, out->mean, out->uncer);
        if (grad && n_params > 0)
            hipLaunchKernelGGL(k_moments, dim3(64), dim3(256), 0, cur, t_gsum.as<double>(), t_gsq.as<double>(),
                               d_parN.as<double>(), (uint64_t)n_params, (double*)nullptr, (double*)nullptr,
                               grad_mean, grad_sigma);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(cur));
        if (out) {
            out->flights = flights;
            out->seconds = ms * 1e-3;
        }
        return;
    }
    HIP_CHECK(hipStreamSynchronize(cur));
    std::vector<double> sum(n_bins), sq(n_bins);
    HIP_CHECK(hipMemcpy(sum.data(), t_sum.p, sizeof(double) * n_bins, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(sq.data(), t_sq.p, sizeof(double) * n_bins, hipMemcpyDeviceToHost));
    if (out) {
        out->flights = flights;
        out->seconds = ms * 1e-3;
        for (int p = 0; p < np; ++p) {
            const double N = (double)nh[p];
            for (int b = prob[p].bin_begin; b < prob[p].bin_begin + prob[p].bin_count; ++b) {
                if (out->sum) out->sum[b] = sum[b];
                if (out->sq) out->sq[b] = sq[b];
                // Estimator::end_cycle / end_simulation, Estimator.cpp:347-367
                const double mean = sum[b] / N;
                const double u2 = (sq[b] / N - mean * mean) / (N - 1.0);
                if (out->mean) out->mean[b] = mean;
                if (out->uncer) out->uncer[b] = std::sqrt(u2);
            }
        }
    }
    if (grad) {
        std::vector<double> gs(n_params), gq(n_params);
        HIP_CHECK(hipMemcpy(gs.data(), t_gsum.p, sizeof(double) * n_params, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(gq.data(), t_gsq.p, sizeof(double) * n_params, hipMemcpyDeviceToHost));
        for (int p = 0; p < np; ++p) {
            const double N = (double)nh[p];
            for (int k = prob[p].param_begin; k < prob[p].param_begin + prob[p].param_count; ++k) {
                const double m = gs[k] / N;
                if (grad_mean) grad_mean[k] = m;
                if (grad_sigma) grad_sigma[k] = std::sqrt((gq[k] / N - m * m) / (N - 1.0));
            }
        }
    }
    if (tcount) {
        HIP_CHECK(hipMemcpy(run->trace, dtrace, sizeof(mc_trace_record) * tcount * (size_t)np,
                            hipMemcpyDeviceToHost));
    }
}

// ---------------------------------------------------------------- Engine API
Engine::Engine(const std::vector<const Problem*>& ps, const mc_engine_cfg* c) : d(new Impl)
{
    if (ps.empty()) throw Error(MC_ERR_ARGUMENT, "no problems");
    mc_engine_cfg cfg{};
    if (c) cfg = *c;
    if (cfg.block == 0) cfg.block = 256;
    if (cfg.grid_per_cu == 0) cfg.grid_per_cu = 4;
    if (cfg.iters_per_graph == 0) cfg.iters_per_graph = 16;
    if (cfg.bank_capacity == 0) cfg.bank_capacity = 1ull << 21;
    if (cfg.drain_divisor == 0) cfg.drain_divisor = 4;
    if (cfg.events_per_launch == 0) cfg.events_per_launch = 16;
    if (cfg.events_per_launch < 1 || cfg.events_per_launch > 64)
        throw Error(MC_ERR_ARGUMENT, "events_per_launch must be in 1..64");
    // gradient runs: measured best on the 24-parameter shield (1e8 histories)
    if (cfg.grad_events_per_launch == 0) cfg.grad_events_per_launch = 24;
    if (cfg.grad_live_divisor == 0) cfg.grad_live_divisor = 4;
    if (cfg.grad_events_per_launch < 1 || cfg.grad_events_per_launch > 64 || cfg.grad_live_divisor < 1)
        throw Error(MC_ERR_ARGUMENT, "grad_events_per_launch must be in 1..64, grad_live_divisor at least 1");
    if (cfg.block % 64 || cfg.block > 1024) throw Error(MC_ERR_ARGUMENT, "block must be a multiple of 64, <= 1024");
    if (cfg.bank_capacity > (1ull << 31)) throw Error(MC_ERR_CAPACITY, "bank capacity above 2^31");
    d->cfg = cfg;
    d->device = cfg.device;
    int ndev = 0;
    HIP_CHECK(hipGetDeviceCount(&ndev));
    if (ndev == 0) throw Error(MC_ERR_DEVICE, "no HIP device");
    HIP_CHECK(hipSetDevice(d->device));
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, d->device));
    d->n_cu = prop.multiProcessorCount;
    if ((int64_t)d->n_cu * cfg.grid_per_cu > 8192 || cfg.grid_per_cu < 1)
        throw Error(MC_ERR_ARGUMENT, "grid_per_cu gives more than 8192 workgroups");
    HIP_CHECK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    d->scratch_ptr = device_scratch(d->device);
    HIP_CHECK(hipHostMalloc((void**)&d->h_ctr, sizeof(Counters), hipHostMallocDefault));
    HIP_CHECK(hipEventCreate(&d->ev0));
    HIP_CHECK(hipEventCreate(&d->ev1));
    d->build(ps);
    d->upload_static();
    if (d->lookup_lds() > 0) {
        // the standalone lookup stages up to kLdsBucketBytes of bucket table in LDS
        HIP_CHECK(hipFuncSetAttribute((const void*)k_xs_lookup_array<kXsUnroll>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)d->lookup_lds()));
    }
}

Engine::~Engine() = default;

void Engine::info(mc_engine_info* o) const
{
    o->n_problems = (int32_t)d->prob.size();
    o->n_bins = d->n_bins;
    o->n_params = d->n_params;
    o->bank_capacity = d->cfg.bank_capacity;
    o->block = d->cfg.block;
    o->grid_per_cu = d->cfg.grid_per_cu;
    o->iters_per_graph = d->cfg.iters_per_graph;
    o->n_cu = d->n_cu;
}

void Engine::offsets(int p, int32_t* bin_off, int32_t* nb, int32_t* par_off, int32_t* npar) const
{
    if (p < 0 || p >= (int)d->prob.size()) throw Error(MC_ERR_ARGUMENT, "problem index out of range");
    if (bin_off) *bin_off = d->prob[p].bin_begin;
    if (nb) *nb = d->prob[p].bin_count;
    if (par_off) *par_off = d->prob[p].param_begin;
    if (npar) *npar = d->prob[p].param_count;
}

void Engine::forward(const mc_run_cfg* run, mc_tally_out* out) { d->run(run, false, nullptr, out, nullptr, nullptr); }

void Engine::backward(const mc_run_cfg* run, const double* gout, double* grad, double* sig, mc_tally_out* out)
{
    if (!gout) throw Error(MC_ERR_ARGUMENT, "grad_out is required");
    if (d->n_params == 0) throw Error(MC_ERR_ARGUMENT, "the problems declare no <parameters>");
    d->run(run, true, gout, out, grad, sig);
}

void Engine::xs_lookup(const double* E, const int32_t* mat, double* out, uint64_t n)
{
    HIP_CHECK(hipSetDevice(d->device));
    for (uint64_t i = 0; i < n; ++i)
        if (mat[i] < -1 || mat[i] >= (int32_t)d->n_mat0)
            throw Error(MC_ERR_ARGUMENT, "material index out of range");
    DevBuf bE, bM, bO;
    bE.alloc(8 * n);
    bM.alloc(4 * n);
    bO.alloc(8 * n);
    HIP_CHECK(hipMemcpyAsync(bE.p, E, 8 * n, hipMemcpyHostToDevice, d->stream));
    HIP_CHECK(hipMemcpyAsync(bM.p, mat, 4 * n, hipMemcpyHostToDevice, d->stream));
    // problem 0's materials start at index 0
    const KArgs A = d->args(false, nullptr, 0);
    hipLaunchKernelGGL(k_xs_lookup_array<kXsUnroll>, d->lookup_grid(), d->lookup_block(), d->lookup_lds(), d->stream, A,
                       bE.as<double>(),
                       bM.as<int32_t>(), bO.as<double>(), n);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(out, bO.p, 8 * n, hipMemcpyDeviceToHost, d->stream));
    HIP_CHECK(hipStreamSynchronize(d->stream));
}

// mc_xs_lookup with the device staging in `scratch` (at this device's address
// found when the engine was created): no device memory is allocated or freed.
void Engine::xs_lookup_noalloc(const double* E, const int32_t* mat, double* out, uint64_t n)
{
    if (n > kScratchLookups)
        throw Error(MC_ERR_CAPACITY, "mc_xs_lookup_noalloc takes at most " + std::to_string(kScratchLookups) +
                                         " lookups per call (" + std::to_string(n) + " given)");
    HIP_CHECK(hipSetDevice(d->device));
    for (uint64_t i = 0; i < n; ++i)
        if (mat[i] < -1 || mat[i] >= (int32_t)d->n_mat0)
            throw Error(MC_ERR_ARGUMENT, "material index out of range");
    double* sE = (double*)d->scratch_ptr;
    double* sO = (double*)(d->scratch_ptr + 8 * kScratchLookups);
    int32_t* sM = (int32_t*)(d->scratch_ptr + 16 * kScratchLookups);
    HIP_CHECK(hipMemcpyAsync(sE, E, 8 * n, hipMemcpyHostToDevice, d->stream));
    HIP_CHECK(hipMemcpyAsync(sM, mat, 4 * n, hipMemcpyHostToDevice, d->stream));
    const KArgs A = d->args(false, nullptr, 0);
    hipLaunchKernelGGL(k_xs_lookup_array<kXsUnroll>, d->lookup_grid(), d->lookup_block(), d->lookup_lds(), d->stream, A,
                       (const double*)sE, (const int32_t*)sM, sO, n);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(out, sO, 8 * n, hipMemcpyDeviceToHost, d->stream));
    HIP_CHECK(hipStreamSynchronize(d->stream));
}

void Engine::xs_bench(uint64_t n, int iters, uint64_t seed, double* sec, double* checksum)
{
    HIP_CHECK(hipSetDevice(d->device));
    if (!d->has_tables) throw Error(MC_ERR_ARGUMENT, "no tabulated nuclides to look up");
    std::vector<int32_t> mats;
    for (int m = 0; m < d->n_mat0; ++m) mats.push_back(m);
    DevBuf bE, bM, bO, bmats;
    bE.alloc(8 * n);
    bM.alloc(4 * n);
    bO.alloc(8 * n);
    upload(bmats, mats, d->stream);
    const double lnmin = std::log(d->emin), lnspan = std::log(d->emax) - std::log(d->emin);
    hipLaunchKernelGGL(k_xs_inputs, dim3(2048), dim3(256), 0, d->stream, bE.as<double>(), bM.as<int32_t>(), n,
                       lnmin, lnspan, bmats.as<int32_t>(), (int32_t)mats.size(), seed);
    const KArgs A = d->args(false, nullptr, 0);
    const dim3 g = d->lookup_grid(), b = d->lookup_block();
    const size_t lds = d->lookup_lds();
    auto launch = [&]() {
        hipLaunchKernelGGL(k_xs_lookup_array<kXsUnroll>, g, b, lds, d->stream, A, bE.as<double>(), bM.as<int32_t>(),
                           bO.as<double>(), n);
    };
    launch();   // warm-up
    HIP_CHECK(hipEventRecord(d->ev0, d->stream));
    for (int k = 0; k < iters; ++k) launch();
    HIP_CHECK(hipEventRecord(d->ev1, d->stream));
    HIP_CHECK(hipEventSynchronize(d->ev1));
    HIP_CHECK(hipGetLastError());
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, d->ev0, d->ev1));
    *sec = ms * 1e-3 / std::max(1, iters);
    if (checksum) {
        std::vector<double> o(std::min<uint64_t>(n, 1 << 16));
        HIP_CHECK(hipMemcpy(o.data(), bO.p, 8 * o.size(), hipMemcpyDeviceToHost));
        double s = 0;
        for (double v : o) s += v;
        *checksum = s;
    }
}

// Times k_gather_lines on the same synthetic inputs as xs_bench: `dep`
// dependent 128-byte line reads per lookup from a table of table_bytes.
void Engine::gather_bench(uint64_t n, int iters, uint64_t seed, uint64_t table_bytes, int dep, double* sec)
{
    HIP_CHECK(hipSetDevice(d->device));
    if (dep != 1 && dep != 2) throw Error(MC_ERR_ARGUMENT, "dependent reads per lookup must be 1 or 2");
    const uint64_t nlines = std::max<uint64_t>(1, table_bytes / 128);
    std::vector<int32_t> mats{0};
    DevBuf bE, bM, bO, bT, bmats;
    bE.alloc(8 * n);
    bM.alloc(4 * n);
    bO.alloc(8 * n);
    bT.alloc(128 * nlines);
    upload(bmats, mats, d->stream);
    hipLaunchKernelGGL(k_fill_f64, dim3(1024), dim3(256), 0, d->stream, bT.as<double>(), 16 * nlines, 1.0);
    const double lnmin = std::log(d->emin > 0 ? d->emin : 1e-5), lnspan = std::log(std::max(d->emax, 2.0)) - lnmin;
    hipLaunchKernelGGL(k_xs_inputs, dim3(2048), dim3(256), 0, d->stream, bE.as<double>(), bM.as<int32_t>(), n,
                       lnmin, lnspan, bmats.as<int32_t>(), 1, seed);
    const dim3 g = d->lookup_grid(), b = d->lookup_block();
    auto launch = [&]() {
        if (dep == 1)
            hipLaunchKernelGGL(k_gather_lines<1>, g, b, 0, d->stream, bE.as<double>(), bM.as<int32_t>(),
                               bO.as<double>(), n, bT.as<double>(), nlines);
        else
            hipLaunchKernelGGL(k_gather_lines<2>, g, b, 0, d->stream, bE.as<double>(), bM.as<int32_t>(),
                               bO.as<double>(), n, bT.as<double>(), nlines);
    };
    launch();
    HIP_CHECK(hipEventRecord(d->ev0, d->stream));
    for (int k = 0; k < iters; ++k) launch();
    HIP_CHECK(hipEventRecord(d->ev1, d->stream));
    HIP_CHECK(hipEventSynchronize(d->ev1));
    HIP_CHECK(hipGetLastError());
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, d->ev0, d->ev1));
    *sec = ms * 1e-3 / std::max(1, iters);
}

uint64_t Engine::lookup_table_bytes() const
{
    // what the standalone lookup reads besides its streams: the records of
    // the LDS path (or the nuclide tables) and the grids
    return 8 * (uint64_t)(d->rec_te.size() + d->rec_t.size() + d->grid_E.size());
}

void Engine::copy_ring(uint64_t* capacity, uint64_t* growths, uint64_t* waits) const
{
    if (capacity) *capacity = d->ring_cap;
    if (growths) *growths = d->ring_growths;
    if (waits) *waits = d->last_waits;
}

void Engine::kernel_stats(mc_kernel_stats* out, bool reset)
{
    if (out) *out = d->kstats;
    if (reset) d->kstats = mc_kernel_stats{};
}

void Engine::set_cfg(const mc_engine_cfg& c)
{
    if ((int64_t)d->n_cu * c.grid_per_cu > 8192 || c.grid_per_cu < 1 || c.block % 64 || c.block < 64 ||
        c.block > 1024 || c.bank_capacity == 0 || c.bank_capacity > (1ull << 31) || c.events_per_launch < 0 ||
        c.events_per_launch > 64)
        throw Error(MC_ERR_ARGUMENT, "invalid launch configuration");
    d->cfg.profile = c.profile;
    d->cfg.block = c.block;
    d->cfg.grid_per_cu = c.grid_per_cu;
    d->cfg.iters_per_graph = c.iters_per_graph;
    d->cfg.use_graph = c.use_graph;
    d->cfg.drain_divisor = c.drain_divisor == 0 ? 4 : c.drain_divisor;
    d->cfg.events_per_launch = c.events_per_launch == 0 ? 16 : c.events_per_launch;
    if (c.grad_events_per_launch >= 1 && c.grad_events_per_launch <= 64)
        d->cfg.grad_events_per_launch = c.grad_events_per_launch;
    if (c.grad_live_divisor >= 1) d->cfg.grad_live_divisor = c.grad_live_divisor;
    if (c.bank_capacity != d->cfg.bank_capacity) d->cfg.bank_capacity = c.bank_capacity;
    ++d->generation;
}

mc_engine_cfg Engine::cfg() const { return d->cfg; }

int Engine::n_problems() const { return (int)d->prob.size(); }
uint64_t Engine::deck_histories(int p) const { return d->deck_hist[p]; }

void selftest_math(const double* in, double* out, uint64_t n)
{
    DevBuf bi, bo;
    bi.alloc(8 * n);
    bo.alloc(32 * n);
    HIP_CHECK(hipMemcpy(bi.p, in, 8 * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_selftest_math, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, bi.as<double>(),
                       bo.as<double>(), n);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(out, bo.p, 32 * n, hipMemcpyDeviceToHost));
}

}  // namespace mc
