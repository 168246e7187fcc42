// mc_device.hpp — device-side data layout and per-event physics of the
// event-based transport loop (gfx950, wave64).
//
// Layout in HBM:
//   * problem data (surfaces, cells, materials, nuclides, estimators) are the
//     flat mc_*_desc records of include/mc_api.h, concatenated over the
//     engine's problems with indices rebased;
//   * cross-section tables: one energy grid per distinct grid (nuclides with
//     identical grids share it) plus per-nuclide sigma_s, sigma_c, sigma_t;
//     each grid carries a bucket table over the IEEE bit pattern of E
//     (2^kBucketBits buckets per octave) narrowing the bin search;
//   * the particle bank is structure-of-arrays, one array per field, indexed
//     by slot; per-slot history tallies are rows of `hist` ([slot][bin]).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mc_physics.hpp"

namespace mc {

constexpr int kBucketBits = 10;                // buckets per octave = 1024
constexpr int kBucketShift = 52 - kBucketBits;

// slot status; flight marks the next event, surface/collision resolve it
// ST_HELD: a member of a split family ended here; the slot has no particle,
// but its tally row is a family's row while fcount[slot] > 0 (then only split
// copies, which score into their own family's row, may take the slot).
// ST_WAIT: the particle split at a surface and waits there for room in the
// copy ring for the rest of its copies (k_resume retries each iteration).
// ST_POP: the particle of a history with a particle comb ended while copies
// wait on its slot's own bank; k_resume combs the bank and starts the top copy.
enum : uint8_t { ST_EMPTY = 0, ST_ALIVE = 1, ST_DEAD = 2, ST_SURF = 3, ST_COLL = 4, ST_HELD = 5, ST_WAIT = 6,
                 ST_POP = 7 };

// Entries a history's bank may gain from one particle's walk beyond the
// comb's bank_max (its capacity is max(bank_max - 1 + kCombSlack, teeth)).
constexpr int kCombSlack = 512;

constexpr int kTouched = 8;        // per-slot list of tally bins touched by the history

// Split copy i (from 0) of a particle starts (i + 1) 2^kCopyJumpLog2 draws
// ahead of the particle's stream.  A history's stream is the block of
// kRnStride draws at key * kRnStride; 2^50 draws are 7.4e9 such blocks, so a
// copy's stream lies beyond the blocks of every key below 7.4e9, and nested
// copies add further multiples (2048 before the 2^61-draw period wraps).
constexpr int kCopyJumpLog2 = 50;

struct DNuc {            // device nuclide (deduplicated across problems)
    double A;
    double xs_s, xs_c, xs_t;
    int32_t kind, grid, tab, n;   // tab: offset into xs_s/xs_c/xs_t tables
    int32_t kin, pad;             // elastic-scatter kinematics (MC_KIN_*)
};

struct DGrid {
    int32_t e_off, n;             // energies in grid_E[e_off .. e_off+n)
    int32_t b_off, nb;            // buckets in bucket[b_off .. b_off+nb]; b_off < 0:
                                  // no bucket table (a grid whose energies decrease
                                  // somewhere: the search runs over the whole grid)
    int64_t key_base;             // (bits(E_0) >> kBucketShift)
};

struct DProb {
    int32_t cell_begin, cell_count;
    int32_t src_begin, src_count;
    int32_t bin_begin, bin_count;
    int32_t param_begin, param_count;
    int32_t mat_begin, mat_count;
    uint64_t seed;
    double wr, ws;
};

struct Counters {
    uint32_t nA, nS, nC;          // queue lengths (written by one workgroup, no atomics)
    uint32_t active_cap;          // slots in use this run: min(capacity, histories)
    uint64_t next_hist;           // next history key to start
    uint64_t end_hist;            // one past the last key of this run
    uint64_t claimed;             // histories started by the last refill
    uint64_t flights;             // Ntrack
    uint32_t err_lost, err_split;
    uint64_t tA, tS, tC;          // particles processed by flight / surface / collision
    uint32_t qa_ident;            // 1: every slot below active_cap is live and queue A is
                                  // the identity (not written; kernels use the slot index)
    uint32_t fcur;                // next queue-A entry a continuous-energy forward flight
                                  // workgroup takes (reset by each refill)
    uint64_t chead, ctail;        // copy ring: split copies waiting for a slot are [chead, ctail)
    uint32_t cplaced, pad2;       // copies the last refill placed
    uint64_t xfl;                 // flights taken beyond the first of an iteration (included in tA)
    uint64_t waits;               // splits and turns that found the ring without room
    uint32_t err_bank, pad3;      // split copies a history's comb bank had no room for
};

// Entry i of queue A.
__device__ __forceinline__ uint32_t qa_at(const uint32_t* qA, uint32_t ident, uint32_t i)
{
    return ident ? i : qA[i];
}

constexpr int kFastNuc = 4;      // nuclides a record's fast path holds in registers

// Interleaved cross-section records of a material whose nuclides are all
// tabulated on one grid.  Record i of rec_t holds, per nuclide k in material
// order, (sigma_t[i], sigma_t[i+1]) at 2k; record i of rec_sc holds the pairs
// of sigma_s at 2k and of sigma_c at c_at + 2k.  They are copies of the nuclide
// tables, so a lookup reads the same values in one or two cache lines instead
// of one gather per nuclide and reaction.  The last record repeats its value.
struct DMatRec {
    int64_t t_off, sc_off;        // offsets (doubles) of record 0, -1: none
    int32_t t_stride, sc_stride;  // record sizes in doubles
    int32_t c_at, pad;            // sigma_c pairs start at c_at: 2 max(n, kFastNuc)
};

// Per material, everything the fast Sigma_t path reads, in one load: the
// shared grid (its bucket table holds at most one grid point per bucket) and
// the material's records (rec_t, or rec_sc for parameterised materials).
// rec_off < 0: no fast path (more than kFastNuc nuclides, constants, mixed
// grids or crowded buckets).
struct DMatLook {
    int64_t key_base;
    int64_t rec_off;
    int32_t e_off, n, b_off, nb;
    int32_t rec_stride, nuc_count, nuc_begin, scaled;
    int64_t te_off;               // rec_te record 0 (unscaled materials), -1: none
    int32_t te_stride, grid;
};

// rec_te: record i = (E_i, sigma_t[i] of each nuclide in material order), 32 or
// 64 B, for the standalone lookup's LDS path, whose bucket table sits in LDS:
// bucket -> the two records of the interval (adjacent, one or two lines; a
// third when E lies above the bucket's grid point).
constexpr int kTeVals = 1 + kFastNuc;
constexpr size_t kLdsBucketBytes = 96 * 1024;   // largest bucket table staged in LDS

// One parameter of a material as the collision's likelihood-ratio fold reads
// it, packed per run: problem-local index, nuclide slot, reactions, the
// entry's density N and 1/theta.
struct DParFold {
    int32_t kl, slot, reactions, pad;
    double N, inv_theta;
};

// The same per nuclide entry, for materials where each nuclide has at most one
// scatter-only and one capture-only parameter (mfold[m] = 1): problem-local
// indices (-1: none) and 1/theta of the scatter parameter.
// Score-function sums kept in LDS by the fused gradient flight kernel: for
// materials of at most kAccNuc nuclides, kAccEntries doubles per thread of
// dynamic LDS (the launch passes kAccEntries * 8 * blockDim bytes).
constexpr int kAccNuc = 4;
constexpr int kAccEntries = 2 * kAccNuc;

struct DNucFold {
    int32_t kl_s, kl_c;
    double inv_theta_s;
};

// Everything a kernel needs, passed by value.
struct KArgs {
    // problem data
    const mc_surface_desc* surf;
    const mc_cell_desc* cell;
    const int32_t* cell_surf;
    const int32_t* cell_sense;
    const int32_t* nbr_begin;     // per surface: candidate cells after a crossing
    const int32_t* nbr;
    const mc_material_desc* mat;
    const int32_t* mat_nuc;       // device nuclide index per material entry
    const double* mat_dens;       // N
    const double* mat_ns;         // per-run scatter density (parameterised materials)
    const double* mat_nc;         // per-run capture density
    const double* mat_tconst;     // per run: Sigma_t of all-constant materials (NaN otherwise)
    const int32_t* mat_param;     // global parameter index
    const int32_t* mat_grid;      // per material: shared grid of its tables, -1 none, -2 mixed
    const int32_t* me_tab;        // per material entry: table offset, -1 constant nuclide
    const double* me_cst;         // per material entry: constant (t, s, c)
    const DMatRec* mrec;          // per material: interleaved cross-section records
    const DMatLook* mlook;        // per material: fast Sigma_t path
    const double* rec_te;
    const uint16_t* bucket16;     // the LDS grid's bucket table as lb + 1
    int32_t lds_grid, lds_nb;     // grid staged in LDS by the lookup kernels (-1: none)
    const double* rec_t;
    const double* rec_sc;
    const mc_param_desc* param;   // material/nuc_slot rebased
    const double* theta;
    const DParFold* mpar;         // per mat_param entry (gradient runs)
    const DNucFold* nfold;        // per material entry (gradient runs)
    const int32_t* mfold;         // per material: 1 when nfold describes all its parameters
    const DNuc* nuc;
    const DGrid* grid;
    const double* grid_E;
    const int32_t* bucket;
    const double* tab_s;
    const double* tab_c;
    const double* tab_t;
    const mc_estimator_desc* est; // bin_offset is global
    const int32_t* att_est;
    const int32_t* att_idx;
    const double* egrid;
    const DProb* prob;
    const mc_source_desc* src;
    const double* src_cdf;        // per problem: src_count + 1 entries at cdf_off
    const int32_t* cdf_off;
    const uint64_t* jump;         // LCG jump multipliers, 256 per byte of a history key
    uint64_t jump_copy;           // g^(2^kCopyJumpLog2): split copy i starts (i + 1) 2^50 draws on
    int32_t split_any;            // some problem can split (importance rises between cells)
    int32_t pad_split;
    const uint64_t* run_hb;       // per problem first run key of this run (n_prob + 1)
    // per problem g^(stride (key_offset - run_hb)) mod 2^63: moves a history
    // from its run key to the key the caller gave it
    const uint64_t* run_jd;
    // per run: {shift, n, then for j = 0..n the problem holding key
    // min(j << shift, last key)}; n <= 65536 blocks of 2^shift keys
    const int32_t* hb_idx;
    // one-speed forward runs: flights (each with its collision) a particle
    // takes per event iteration; per flight workgroup, the flights beyond the
    // first of each particle, folded into the counters by the next refill
    uint32_t events_per_launch;
    uint64_t* wg_xfl;
    int32_t n_prob;
    int32_t n_mat;                // materials of all problems
    int32_t n_cell, n_surf, n_cell_surf;   // geometry of all problems
    const char* crec;             // per-cell flight records (DCellHead, DCellSurf...), or null
    int32_t crec_stride;          // bytes from one cell's record to the next
    int32_t bin_stride;           // hist row length (max bins per problem)
    int32_t par_stride;           // rows of S and H (max params per problem)
    int32_t pad;
    uint64_t cap;                 // bank capacity
    // bank (SoA)
    double *x, *y, *z, *u, *v, *w, *E, *wgt, *sigt;
    uint64_t *rng, *hid;
    int32_t *pcell, *pprob, *psurf;
    int32_t* pmat;                // material of the particle's cell (-1 void)
    int32_t* gidx;                // interval of E on pmat's shared grid, from the lookup kernel
    uint32_t* nflight;
    uint8_t* status;
    uint8_t* tcount;              // touched-bin count per slot (kTouched + 1: overflow)
    uint32_t* tbins;              // [slot][kTouched] touched bins (problem-local)
    uint32_t* wg_cnt;             // per-workgroup counts for the two-pass compaction
    // split families (when split_any): per slot the family of its particle, as
    // the slot whose tally row the family shares (-1: none), and per row slot the
    // family's members not yet ended (copies in the ring included)
    int32_t *fam, *fcount;
    // a particle whose split copies did not all fit in the ring waits at the
    // surface: ppend copies still to place, the next being copy pfirst
    int32_t *ppend, *pfirst;
    // problems with a particle comb: per slot a LIFO bank of its history's
    // split copies (entry i of slot s at s * bcap + i; bsize entries), run one
    // after another in the slot as the reference runs its bank
    double *bx, *by, *bz, *bu, *bv, *bw, *bE, *bwgt;
    uint64_t* brng;
    int32_t *bcell, *bsize;
    uint32_t bcap;
    const int32_t* comb;          // per problem (bank_max, teeth), (0, 0): no comb; null: no problem has one
    uint32_t* wg_held;            // per workgroup: ended slots whose row a family holds
    // copy ring (capacity ccap): particles made by importance splitting
    double *cx, *cy, *cz, *cu, *cv, *cw, *cE, *cwgt;
    uint64_t *crng, *chid;
    int32_t *ccell, *cprob, *cfam;
    uint32_t* cnfl;
    uint64_t ccap;
    // gradient runs that split: a copy's score function S (parameter-major,
    // row stride ccap) and visited-materials mask, taken from the particle at
    // the split; per family row slot the family's gradient sum (row stride cap)
    double* rS;
    uint32_t* rvmat;
    double* famG;
    double* hist;
    // likelihood-ratio state: score function S and H = sum of dS * Q, parameter-
    // major [param][slot] with row stride sstride (cap plus a pad that keeps
    // the rows of one slot off a power-of-two spacing, so the parameters a
    // fold touches spread over memory channels), and per slot the scores' gradient weight Q, the pending coefficients cs, cc
    // of dSigma_s/dtheta, dSigma_c/dtheta and the unfolded flight length lpend
    double* S;
    double* H;
    uint64_t sstride;
    double *Qw, *cs, *cc, *lpend;
    // per slot: the materials with parameters the history has been in (bit
    // m - mat_begin; bit 31: a material of index 31 or above), whose S and H
    // entries its closeout reads and clears
    uint32_t* vmat;
    uint32_t* qS;                 // gradient runs: surface crossings, per flight workgroup segment
    uint32_t *qA, *qC;
    Counters* ctr;
    // tallies: per-workgroup partial moments [workgroup][bin], reduced at the
    // end of a run; a row stride of 0 makes every workgroup share one row
    double* psum;
    double* psq;
    uint64_t bin_row;             // row stride of psum/psq (engine bins, or 0)
    const double* gout;           // gradient weights per bin
    double* pgsum;
    double* pgsq;
    uint64_t par_row;             // row stride of pgsum/pgsq (engine params, or 0)
    mc_trace_record* trace;       // device buffer [prob][trace_count]
    uint64_t trace_count;
};

// ------------------------------------------------------------------ XS lookup

// Nuclide::checkE's binary_search (Algorithm.cpp:46-64) narrowed by the
// grid's bucket table: the result equals the full search exactly.  A grid
// that is not non-decreasing (the reference reads such files as they are)
// has no table and takes the full search, whose probes alone define the
// result there.
__device__ __forceinline__ int grid_index(const KArgs& A, const DGrid& g, double E)
{
    if (g.b_off < 0) return binary_search(E, A.grid_E + g.e_off, g.n);
    const int64_t k = (int64_t)(double_to_bits(E) >> kBucketShift) - g.key_base;
    if (k < 0) return -1;
    if (k >= g.nb) return g.n - 1;
    const int32_t* b = A.bucket + g.b_off + k;
    int left = b[0] + 1, right = b[1];
    const double* e = A.grid_E + g.e_off;
    while (left <= right) {
        const int mid = (left + right) >> 1;
        if (e[mid] < E) left = mid + 1;
        else right = mid - 1;
    }
    return right;
}

// Interpolation point on one grid, reused by every nuclide sharing the grid.
struct XsPoint {
    int grid = -1;
    int idx = 0, n = 0;
    double a = 0.0, b = 0.0;   // (E-E2)/(E1-E2) and (E-E1)/(E2-E1)
};

__device__ __forceinline__ void xs_point(const KArgs& A, int gid, double E, XsPoint& P)
{
    if (P.grid == gid) return;
    const DGrid g = A.grid[gid];
    P.grid = gid;
    P.n = g.n;
    P.idx = grid_index(A, g, E);
    if (P.idx >= 0 && P.idx < g.n - 1) {
        const double E1 = A.grid_E[g.e_off + P.idx], E2 = A.grid_E[g.e_off + P.idx + 1];
        P.a = (E - E2) / (E1 - E2);
        P.b = (E - E1) / (E2 - E1);
    }
}

// The same point from an interval index found earlier at this E.
__device__ __forceinline__ void xs_point_on(const KArgs& A, int gid, int e_off, int n, int idx, double E,
                                            XsPoint& P)
{
    P.grid = gid;
    P.n = n;
    P.idx = idx;
    if (idx >= 0 && idx < n - 1) {
        const double E1 = A.grid_E[e_off + idx], E2 = A.grid_E[e_off + idx + 1];
        P.a = (E - E2) / (E1 - E2);
        P.b = (E - E1) / (E2 - E1);
    }
}

__device__ __forceinline__ void xs_point_at(const KArgs& A, int gid, int idx, double E, XsPoint& P)
{
    const DGrid g = A.grid[gid];
    xs_point_on(A, gid, g.e_off, g.n, idx, E, P);
}

// XSTable::xs (XSec.cpp:9-40) with the interpolation of Algorithm.cpp:103-105
__device__ __forceinline__ double xs_table(const double* y, const XsPoint& P)
{
    if (P.idx == P.n - 1) return y[P.n - 1];
    if (P.idx == -1) return y[0];
    return P.a * y[P.idx] + P.b * y[P.idx + 1];
}

enum { XS_T = 0, XS_S = 1, XS_C = 2 };

__device__ __forceinline__ double nuclide_xs(const KArgs& A, int n, int which, double E, XsPoint& P)
{
    const DNuc N = A.nuc[n];
    if (N.kind == MC_NUC_CONSTANT) return which == XS_S ? N.xs_s : which == XS_C ? N.xs_c : N.xs_t;
    xs_point(A, N.grid, E, P);
    const double* t = which == XS_S ? A.tab_s : which == XS_C ? A.tab_c : A.tab_t;
    return xs_table(t + N.tab, P);
}

struct Xs3 {
    double t, s, c;
};

// Interval of E on material m's shared grid, when all its tabulated nuclides
// share one (mat_grid[m] >= 0); one search serves every nuclide.
__device__ __forceinline__ int material_point(const KArgs& A, int m, double E, XsPoint& P)
{
    const int g = A.mat_grid[m];
    if (g >= 0) xs_point(A, g, E, P);
    return g;
}

// material_point from the interval the lookup kernel stored for slot s; valid
// while the particle keeps the material and energy it had at the lookup.
__device__ __forceinline__ int material_point_cached(const KArgs& A, int m, uint32_t s, double E, XsPoint& P)
{
    const int g = A.mat_grid[m];
    if (g >= 0) xs_point_at(A, g, A.gidx[s], E, P);
    return g;
}

// Microscopic cross section of material entry j (mat_nuclide[j]).  Constant
// nuclides read me_cst; tabulated ones read their table at the shared
// interval, or fall back to the nuclide's own grid.
__device__ __forceinline__ double entry_xs(const KArgs& A, int j, int which, double E, XsPoint& P, int g)
{
    const int tb = A.me_tab[j];
    if (tb < 0) return A.me_cst[3 * j + which];
    if (g >= 0) {
        const double* t = which == XS_S ? A.tab_s : which == XS_C ? A.tab_c : A.tab_t;
        return xs_table(t + tb, P);
    }
    return nuclide_xs(A, A.mat_nuc[j], which, E, P);
}

// xs_table on a record pair (y[i], y[i+1]) taken at record max(idx, 0).
__device__ __forceinline__ double xs_pair(const double* r, const XsPoint& P)
{
    if (P.idx == P.n - 1 || P.idx == -1) return r[0];
    return P.a * r[0] + P.b * r[1];
}

__device__ __forceinline__ const double* rec_at(const double* base, int64_t off, int32_t stride, const XsPoint& P)
{
    return base + off + (int64_t)(P.idx > 0 ? P.idx : 0) * stride;
}

// material_sigt by the fast path: bucket, then the three grid points and the
// two records the interval can use, all loaded at once, then the choice.  Same
// interval (binary_search, Algorithm.cpp:46-64), same interpolation and the
// same sum as material_sigt, so the same bits.  False: m has no fast path.
__device__ __forceinline__ bool material_sigt_fast_at(const KArgs& A, const DMatLook& L, const double* dens,
                                                     const double* ns, const double* nc, double E, double& sig,
                                                     int& idx);

__device__ __forceinline__ bool material_sigt_fast(const KArgs& A, int m, double E, double& sig, int& idx)
{
    const DMatLook L = A.mlook[m];
    const int j = L.nuc_begin > 0 ? L.nuc_begin : 0;
    return material_sigt_fast_at(A, L, A.mat_dens + j, A.mat_ns + j, A.mat_nc + j, E, sig, idx);
}

// material_sigt_fast for the material of lookup record L; dens, ns and nc are
// its nuclides' densities (unscaled materials) or per-reaction densities
// (scaled ones).
__device__ __forceinline__ bool material_sigt_fast_at(const KArgs& A, const DMatLook& L, const double* dens,
                                                     const double* ns, const double* nc, double E, double& sig,
                                                     int& idx)
{
    if (L.rec_off < 0) return false;
    const int64_t key = (int64_t)(double_to_bits(E) >> kBucketShift) - L.key_base;
    int i, r0, r1;
    double E1 = 0.0, E2 = 0.0;
    bool up = false;
    if (key < 0) {
        i = -1; r0 = r1 = 0;
    } else if (key >= L.nb) {
        i = L.n - 1; r0 = r1 = L.n - 1;
    } else {
        const int lb = A.bucket[L.b_off + key], rb = A.bucket[L.b_off + key + 1];
        const int left = lb + 1;
        const double* e = A.grid_E + L.e_off;
        const double em = e[lb > 0 ? lb : 0];
        const double e0 = e[left < L.n - 1 ? left : L.n - 1];
        const double e1 = e[left + 1 < L.n - 1 ? left + 1 : L.n - 1];
        up = rb >= left && e0 < E;               // the bucket's one point lies below E
        i = up ? left : lb;
        E1 = up ? e0 : em;
        E2 = up ? e1 : e0;
        r0 = lb > 0 ? lb : 0;
        r1 = left < L.n - 1 ? left : L.n - 1;
    }
    const double* base = L.scaled ? A.rec_sc : A.rec_t;
    const double* ra = base + L.rec_off + (int64_t)r0 * L.rec_stride;
    const double* rb2 = base + L.rec_off + (int64_t)r1 * L.rec_stride;
    // the whole of both candidate records (a record has at least kV doubles
    // when it serves the fast path), then the one the interval uses
    constexpr int kV = 4 * kFastNuc;
    const int cnt = L.scaled ? kV : 2 * kFastNuc;
    double v[kV];
#pragma unroll
    for (int q = 0; q < kV; ++q) {
        v[q] = 0.0;
        if (q < cnt) {
            const double x0 = ra[q], x1 = rb2[q];
            v[q] = up ? x1 : x0;
        }
    }
    XsPoint P;
    P.idx = i;
    P.n = L.n;
    if (i >= 0 && i < L.n - 1) {
        P.a = (E - E2) / (E1 - E2);
        P.b = (E - E1) / (E2 - E1);
    }
    idx = i;
    if (!L.scaled) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < kFastNuc; ++k)
            if (k < L.nuc_count) t += xs_pair(v + 2 * k, P) * dens[k];
        sig = t;
        return true;
    }
    double s = 0.0, c = 0.0;
#pragma unroll
    for (int k = 0; k < kFastNuc; ++k) {
        if (k < L.nuc_count) {
            s += xs_pair(v + 2 * k, P) * ns[k];
            c += xs_pair(v + 2 * kFastNuc + 2 * k, P) * nc[k];
        }
    }
    sig = s + c;
    return true;
}

// Material::SigmaT/S/C (Material.cpp:18-57); parameterised materials use the
// per-reaction densities and Sigma_t = Sigma_s + Sigma_c.
__device__ __forceinline__ double material_sigt(const KArgs& A, int m, double E, int* idx = nullptr)
{
    double fs;
    int fi;
    if (material_sigt_fast(A, m, E, fs, fi)) {
        if (idx) *idx = fi;
        return fs;
    }
    const mc_material_desc M = A.mat[m];
    XsPoint P;
    const int g = material_point(A, m, E, P);
    if (idx) *idx = P.idx;
    const DMatRec R = A.mrec[m];
    if (R.t_off >= 0) {
        if (M.param_count == 0) {
            const double* r = rec_at(A.rec_t, R.t_off, R.t_stride, P);
            double t = 0.0;
            for (int k = 0; k < M.nuc_count; ++k) t += xs_pair(r + 2 * k, P) * A.mat_dens[M.nuc_begin + k];
            return t;
        }
        const double* r = rec_at(A.rec_sc, R.sc_off, R.sc_stride, P);
        double s = 0.0, c = 0.0;
        for (int k = 0; k < M.nuc_count; ++k) {
            const int j = M.nuc_begin + k;
            s += xs_pair(r + 2 * k, P) * A.mat_ns[j];
            c += xs_pair(r + R.c_at + 2 * k, P) * A.mat_nc[j];
        }
        return s + c;
    }
    if (M.param_count == 0) {
        double t = 0.0;
        for (int k = 0; k < M.nuc_count; ++k) {
            const int j = M.nuc_begin + k;
            t += entry_xs(A, j, XS_T, E, P, g) * A.mat_dens[j];
        }
        return t;
    }
    double s = 0.0, c = 0.0;
    for (int k = 0; k < M.nuc_count; ++k) {
        const int j = M.nuc_begin + k;
        s += entry_xs(A, j, XS_S, E, P, g) * A.mat_ns[j];
        c += entry_xs(A, j, XS_C, E, P, g) * A.mat_nc[j];
    }
    return s + c;
}

// Sigma_t of an unscaled material on the LDS grid: the bucket from LDS, then
// the record of the interval's lower point, which carries the grid energies
// the choice and the interpolation need.  Same interval, same values, same
// arithmetic as material_sigt.  key_base, te_off, te_stride, n, nb and
// nuc_count are the material's DMatLook fields, dens its nuclide densities.
__device__ __forceinline__ double sigt_lds_core(const KArgs& A, int64_t key_base, int64_t te_off, int te_stride,
                                                int n, int nb, int nuc_count, const double* dens,
                                                const uint16_t* sb, double E, int& idx)
{
    const int64_t key = (int64_t)(double_to_bits(E) >> kBucketShift) - key_base;
    const double* base = A.rec_te + te_off;
    const int64_t st = te_stride;
    const int last = n - 1;
    double a[kTeVals], b[kTeVals];       // records i and i + 1 of the interval
    int i;
    if (key < 0 || key >= nb) {
        i = key < 0 ? -1 : last;
        const double* r = base + (i > 0 ? i : 0) * st;
#pragma unroll
        for (int q = 0; q < kTeVals; ++q) { a[q] = r[q]; b[q] = a[q]; }
    } else {
        const int lb = (int)sb[key] - 1, rb = (int)sb[key + 1] - 1;
        const int r0 = lb > 0 ? lb : 0, r1 = r0 + 1 < last ? r0 + 1 : last;
        const double* ra = base + r0 * st;
        const double* rb1 = base + r1 * st;
#pragma unroll
        for (int q = 0; q < kTeVals; ++q) { a[q] = ra[q]; b[q] = rb1[q]; }
        // the bucket's one grid point (E_{lb+1}) lies below E: the next interval
        const bool up = rb > lb && (lb >= 0 ? b[0] : a[0]) < E;
        i = up ? lb + 1 : lb;
        if (up && lb >= 0) {
            const int r2 = lb + 2 < last ? lb + 2 : last;
            const double* rc = base + r2 * st;
#pragma unroll
            for (int q = 0; q < kTeVals; ++q) { a[q] = b[q]; b[q] = rc[q]; }
        }
    }
    const bool interp = i >= 0 && i < last;
    double pa = 0.0, pb = 0.0;
    if (interp) {
        pa = (E - b[0]) / (a[0] - b[0]);
        pb = (E - a[0]) / (b[0] - a[0]);
    }
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < kFastNuc; ++k)
        if (k < nuc_count) t += (interp ? pa * a[1 + k] + pb * b[1 + k] : a[1 + k]) * dens[k];
    idx = i;
    return t;
}

__device__ __forceinline__ double material_sigt_lds(const KArgs& A, const DMatLook& L, const uint16_t* sb, double E,
                                                   int& idx)
{
    return sigt_lds_core(A, L.key_base, L.te_off, L.te_stride, L.n, L.nb, L.nuc_count, A.mat_dens + L.nuc_begin, sb,
                         E, idx);
}

// The flight kernel's copy in LDS of the first kFlightLooks materials' lookup
// records and nuclide densities (and per-reaction densities), which the
// lookup after each collision reads.
// Per-cell flight records, a fixed stride apart (problem sets whose geometry
// does not fit the LDS copy below): the cell fields a flight and a collision
// read, then its surfaces' ids, types and coefficients in the cell's order,
// so that a flight reads its geometry in one place instead of through the
// cell, its surface list and the surfaces.
struct DCellHead {
    int32_t material, surf_count, tl_begin, tl_count, col_begin, col_count, pad0, pad1;
};
struct DCellSurf {
    int32_t sid, type;
    double c[5];                  // mc_surface_desc c[0..4]: all surface_distance reads
};
constexpr int kRecSurfMax = 8;    // cells of at most this many surfaces

// Where the fused flight kernel reads the geometry from.
enum { kGeoArrays = 0, kGeoLds = 1, kGeoRec = 2 };

// The fused flight kernel's copy in LDS of the whole geometry (cells,
// their surface lists, surfaces) when it is small enough, read by every
// flight and collision instead of the engine's arrays.
constexpr int kGeoCells = 64, kGeoSurfs = 64, kGeoCellSurf = 256;
struct GeoLds {
    mc_cell_desc cell[kGeoCells];
    int32_t cell_surf[kGeoCellSurf];
    mc_surface_desc surf[kGeoSurfs];
};

// The lookup record also carries what the material's collisions read: its
// cross-section records and, per nuclide slot, the target mass and kinematics.
struct DFlightLook {
    DMatLook L;
    DMatRec R;
    double dens[kFastNuc], ns[kFastNuc], nc[kFastNuc];   // per nuclide slot
    double mass[kFastNuc];
    int32_t kin[kFastNuc];
};
constexpr int kFlightLooks = 32;

// Sigma_t for the flight kernel's lookups: the material's record from LDS when
// staged, else material_sigt (the same bits either way).
__device__ __forceinline__ double flight_sigt(const KArgs& A, const DFlightLook* lk, int nlk, int m, double E, int& idx)
{
    if (m < nlk) {
        double s;
        if (material_sigt_fast_at(A, lk[m].L, lk[m].dens, lk[m].ns, lk[m].nc, E, s, idx)) return s;
    }
    return material_sigt(A, m, E, &idx);
}

// The standalone lookup's copy in LDS of what its path reads per material
// (the DMatLook fields and the nuclide densities), so that a lookup reads
// its material from LDS instead of from memory.  te_off < 0: the material
// takes material_sigt.
struct DLdsMat {
    int64_t key_base, te_off;
    int32_t te_stride, n, nb, nuc_count;
    double dens[kFastNuc];
};
constexpr int kLdsMats = 256;                    // materials the standalone lookup stages

// The lookup kernels' Sigma_t: the LDS path when the material has one, else
// material_sigt.
__device__ __forceinline__ double lookup_sigt(const KArgs& A, const uint16_t* sb, int m, double E, int& idx)
{
    if (sb) {
        const DMatLook L = A.mlook[m];
        if (L.te_off >= 0 && L.grid == A.lds_grid) return material_sigt_lds(A, L, sb, E, idx);
    }
    return material_sigt(A, m, E, &idx);
}

// Stages the LDS grid's bucket table (lds_nb + 1 entries) in dynamic LDS;
// returns nullptr when there is none.  Every thread of the block must call it.
__device__ __forceinline__ const uint16_t* stage_buckets(const KArgs& A, uint16_t* sb)
{
    if (A.lds_grid < 0) return nullptr;
    const int n = A.lds_nb + 1;
    for (int k = threadIdx.x; k < n; k += blockDim.x) sb[k] = A.bucket16[k];
    __syncthreads();
    return sb;
}

__device__ __forceinline__ Xs3 material_xs3(const KArgs& A, int m, double E)
{
    const mc_material_desc M = A.mat[m];
    XsPoint P;
    const int g = material_point(A, m, E, P);
    Xs3 r = {0.0, 0.0, 0.0};
    if (M.param_count == 0) {
        for (int k = 0; k < M.nuc_count; ++k) {
            const int j = M.nuc_begin + k;
            const double N = A.mat_dens[j];
            r.t += entry_xs(A, j, XS_T, E, P, g) * N;
            r.s += entry_xs(A, j, XS_S, E, P, g) * N;
            r.c += entry_xs(A, j, XS_C, E, P, g) * N;
        }
    } else {
        for (int k = 0; k < M.nuc_count; ++k) {
            const int j = M.nuc_begin + k;
            r.s += entry_xs(A, j, XS_S, E, P, g) * A.mat_ns[j];
            r.c += entry_xs(A, j, XS_C, E, P, g) * A.mat_nc[j];
        }
        r.t = r.s + r.c;
    }
    return r;
}

// d Sigma_s / d theta_k and d Sigma_c / d theta_k for parameter q of material m
struct DSig {
    int k;
    double ds, dc;
};

__device__ __forceinline__ DSig param_dsig(const KArgs& A, const mc_material_desc& M, int q, double E, XsPoint& P,
                                          int g)
{
    DSig r;
    r.k = A.mat_param[M.param_begin + q];
    const mc_param_desc D = A.param[r.k];
    const int j = M.nuc_begin + D.nuc_slot;
    const double N = A.mat_dens[j];
    r.ds = (D.reactions & MC_REACTION_SCATTER) ? entry_xs(A, j, XS_S, E, P, g) * N : 0.0;
    r.dc = (D.reactions & MC_REACTION_CAPTURE) ? entry_xs(A, j, XS_C, E, P, g) * N : 0.0;
    return r;
}

// ------------------------------------------------------------------ wave ops

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Ballot + popcount compaction: one atomic per wave, ranks in lane order.
__device__ __forceinline__ uint32_t wave_append(bool pred, uint32_t* counter)
{
    const uint64_t mask = __ballot(pred);
    if (mask == 0) return 0;
    const int lane = (int)lane_id();
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

}  // namespace mc
