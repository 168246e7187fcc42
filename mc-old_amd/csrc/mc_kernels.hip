// mc_kernels.hip — event-based transport kernels for gfx950 (wave64).
//
// One event iteration over the structure-of-arrays particle bank:
//   k_resume          (problems that can split) particles waiting at a surface
//                     for room in the copy ring place the copies it now has
//                     room for; those with none left take their roulette
//   k_refill_count    per-workgroup counts of live, free and family-held slots
//                     (and the counter roll-over from the previous iteration)
//   k_refill_scatter  prefix over the counts; split copies waiting in the
//                     copy ring, else new histories, go into free slots; queue
//                     A = live slots plus the new ones, in slot order, not
//                     written while every slot is live (Simulator::start's
//                     sample loop, handler.cpp:19-36)
//   k_xs_lookup       Sigma_t(E) for every particle in A (Material::SigmaT,
//                     Material.cpp:50-57) — continuous-energy problems only
//   k_flight          distance to collision vs nearest surface, move, track-
//                     length tallies (random_walk + move_particle, general.cpp:
//                     40-83, 177-211); flights that end on a surface are crossed
//                     by the workgroup (gradient runs: by k_cross): boundary
//                     conditions, cell search, surface tallies, importance
//                     split or roulette, weight roulette (general.cpp:89-115,
//                     population_control.cpp:9-49); a flight's collision
//                     follows in the same thread and a surviving particle
//                     flies on, up to events_per_launch flights per launch
//                     (gradient runs at one flight per launch instead queue
//                     collisions in the workgroup's own segment of queue C)
//   k_collision       (gradient runs at one flight per launch) per workgroup,
//                     the segment its flight workgroup filled:
//                     collision tallies, implicit capture, nuclide selection,
//                     elastic scatter, weight roulette (general.cpp:121-163)
// k_drain carries the last particles of a run from event to event through the
// same per-particle routines.  Queue A is built without atomics: workgroups
// rank themselves from per-workgroup counts and rank their slots with block
// scans, so it keeps slot order and consecutive lanes touch consecutive bank
// entries.  Queue C is block-local: a wave ballot ranks its lanes and one LDS
// atomic per wave places the wave.  A history that ends is closed out where it
// ends: the bins it touched are added to the FP64 moment arrays
// (Estimator::end_history, Estimator.cpp:339-346).
#include "mc_device.hpp"

namespace mc {

// ------------------------------------------------------------ block helpers

// Sum over the block of two counters.  lds: at least 32 words.
__device__ __forceinline__ void block_sum2(uint32_t& a, uint32_t& b, uint32_t* lds)
{
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0) { lds[2 * wave] = a; lds[2 * wave + 1] = b; }
    __syncthreads();
    a = 0;
    b = 0;
    for (uint32_t k = 0; k < nw; ++k) { a += lds[2 * k]; b += lds[2 * k + 1]; }
    __syncthreads();
}

// Block-contiguous partition of n items over the grid, in whole blocks.
__device__ __forceinline__ void block_range(uint32_t n, uint32_t& lo, uint32_t& hi)
{
    const uint32_t nb = gridDim.x, bd = blockDim.x;
    uint32_t per = (n + nb - 1) / nb;
    per = (per + bd - 1) / bd * bd;
    lo = blockIdx.x * per;
    hi = lo + per < n ? lo + per : n;
    if (lo > n) lo = n;
}

// Prefix over per-workgroup count pairs: sums of workgroups before this one and
// grand totals.
__device__ __forceinline__ void wg_prefix(const uint32_t* cnt, uint32_t& a_before, uint32_t& b_before,
                                          uint32_t& a_tot, uint32_t& b_tot, uint32_t* lds)
{
    uint32_t ab = 0, bb = 0, at = 0, bt = 0;
    for (uint32_t v = threadIdx.x; v < gridDim.x; v += blockDim.x) {
        const uint32_t a = cnt[2 * v], b = cnt[2 * v + 1];
        if (v < blockIdx.x) { ab += a; bb += b; }
        at += a;
        bt += b;
    }
    block_sum2(ab, bb, lds);
    block_sum2(at, bt, lds);
    a_before = ab;
    b_before = bb;
    a_tot = at;
    b_tot = bt;
}

// The same value, hidden from the optimiser: addresses derived from it are
// computed afresh after this point instead of being kept live (in registers,
// or spilled) from an earlier computation.
__device__ __forceinline__ uint32_t opaque(uint32_t v)
{
    asm volatile("" : "+v"(v));
    return v;
}

#define MC_GRID_LOOP(count)                                                                \
    for (uint32_t _i = blockIdx.x * blockDim.x + threadIdx.x; _i < (count); _i += gridDim.x * blockDim.x)

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ int find_problem(const KArgs& A, uint64_t hid)
{
    // the run's index gives the problems of the 2^shift-key block holding hid
    const int32_t* t = A.hb_idx;
    uint64_t j = hid >> t[0];
    if (j >= (uint64_t)t[1]) j = (uint64_t)t[1] - 1;
    int lo = t[2 + j], hi = t[3 + j];   // largest p with run_hb[p] <= hid
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (A.run_hb[mid] <= hid) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ bool test_point(const KArgs& A, int c, Vec3 p)
{
    // Simulator::test_point, general.cpp:13-20
    const mc_cell_desc C = A.cell[c];
    for (int k = 0; k < C.surf_count; ++k) {
        const int j = C.surf_begin + k;
        const mc_surface_desc& S = A.surf[A.cell_surf[j]];
        if (surface_eval(S.type, S.c, p) * (double)A.cell_sense[j] < 0) return false;
    }
    return true;
}

// Cell after crossing surface s: the cells that list s, in index order, then
// the problem's full linear search (general.cpp:26-34) when none contains p.
__device__ __forceinline__ int find_cell(const KArgs& A, int s, const DProb& PR, Vec3 p)
{
    for (int k = A.nbr_begin[s]; k < A.nbr_begin[s + 1]; ++k)
        if (test_point(A, A.nbr[k], p)) return A.nbr[k];
    for (int c = PR.cell_begin; c < PR.cell_begin + PR.cell_count; ++c)
        if (test_point(A, c, p)) return c;
    return -1;
}

// Adds v to the history's tally bin b (problem-local) and records the bin the
// first time it becomes non-zero.
__device__ __forceinline__ void hist_add(const KArgs& A, uint32_t s, int b, double v)
{
    if (A.split_any) {
        // a member of a split family scores into the family's row (shared, so atomic;
        // the family closes out by a full scan of its row)
        const int f = A.fam[s];
        if (f >= 0) {
            if (v != 0.0) atomicAdd(A.hist + (size_t)f * (size_t)A.bin_stride + b, v);
            return;
        }
    }
    double* h = A.hist + (size_t)s * (size_t)A.bin_stride + b;
    const double old = *h;
    *h = old + v;
    if (old == 0.0 && v != 0.0) {
        const uint8_t n = A.tcount[s];
        if (n < kTouched) {
            A.tbins[(size_t)s * kTouched + n] = (uint32_t)b;
            A.tcount[s] = n + 1;
        } else if (n == kTouched) {
            A.tcount[s] = kTouched + 1;      // overflow: close out by full scan
        }
    }
}

// Estimator::score (Estimator.cpp:298-336) for one attachment.
// Gradient runs: lp is the flight length not yet folded into S (the track's own
// length for track-length scores, the pending flight at a collision, 0 after a
// surface crossing), sigt the cell's Sigma_t at E and coll_end whether the
// track ends in a collision.  A score of weight gv adds gv to Q and, in a
// material with parameters, its extra terms to the coefficients cs, cc:
//   d ln(score)/d theta = S - lp dSigma_t + [track, coll_end] dSigma_t/Sigma_t + d ln Sigma_x.
template <bool kGrad>
__device__ __forceinline__ void score_attachment(const KArgs& A, uint32_t s, int att, double l, int cellid, double E, double w,
                                 const DProb& PR, double lp, double sigt, bool coll_end)
{
    const int e = A.att_est[att];
    const mc_estimator_desc& Es = A.est[e];
    int ebin = 0, NE = 1;
    const int egc = Es.egrid_count;
    if (egc > 0) {
        NE = egc - 1;
        ebin = binary_search(E, A.egrid + Es.egrid_begin, egc);
        if (ebin < 0 || ebin >= NE) return;            // Estimator.cpp:174
    }
    const int fi = A.att_idx[att];
    const int m = A.cell[cellid].material;
    const int nsc = Es.n_scores;
    const int natt = Es.n_attach;
    Xs3 X = {0.0, 0.0, 0.0};
    bool have_x = false;
    for (int sidx = 0; sidx < nsc; ++sidx) {
        const int kern = Es.score_kernel[sidx], sc = Es.score[sidx];
        if (!have_x && (kern == MC_KERNEL_COLLISION || sc != MC_SCORE_FLUX)) {
            if (m >= 0) X = material_xs3(A, m, E);
            have_x = true;
        }
        double kv;                                      // Estimator.cpp:17-41
        if (kern == MC_KERNEL_TRACK_LENGTH) kv = w * l;
        else if (kern == MC_KERNEL_COLLISION) kv = w / X.t;
        else kv = w;
        double v, sx = 1.0;                             // Estimator.cpp:48-124
        if (sc == MC_SCORE_ABSORPTION) { sx = X.c; v = X.c * kv; }
        else if (sc == MC_SCORE_SCATTER) { sx = X.s; v = X.s * kv; }
        else if (sc == MC_SCORE_CAPTURE) { sx = X.c; v = X.c * kv; }
        else if (sc == MC_SCORE_TOTAL) { sx = X.t; v = X.t * kv; }
        else v = kv;
        const int bin = Es.bin_offset + sidx * natt * NE + fi * NE + ebin;
        hist_add(A, s, bin - PR.bin_begin, v);
        if (kGrad) {
            const double g = A.gout[bin];
            if (g != 0.0 && v != 0.0) {
                const double gv = g * v;
                A.Qw[s] += gv;
                if (m >= 0 && A.mat[m].param_count > 0) {
                    double a = -lp;
                    if (kern == MC_KERNEL_TRACK_LENGTH && coll_end) a += 1.0 / sigt;
                    double bs = a, bc = a;
                    if (sc == MC_SCORE_SCATTER) bs += 1.0 / sx;
                    else if (sc == MC_SCORE_ABSORPTION || sc == MC_SCORE_CAPTURE) bc += 1.0 / sx;
                    else if (sc == MC_SCORE_TOTAL) { bs += 1.0 / sx; bc += 1.0 / sx; }
                    A.cs[s] += gv * bs;
                    A.cc[s] += gv * bc;
                }
            }
        }
    }
}

// Pending likelihood-ratio state of one slot (see KArgs::S).
struct LrPending {
    double lp, cs, cc, Q;
};

__device__ __forceinline__ LrPending lr_pending(const KArgs& A, uint32_t s)
{
    return LrPending{A.lpend[s], A.cs[s], A.cc[s], A.Qw[s]};
}

// The same without the flight length (lp = 0), for a caller that has it.
__device__ __forceinline__ LrPending lr_pending_cs(const KArgs& A, uint32_t s)
{
    return LrPending{0.0, A.cs[s], A.cc[s], A.Qw[s]};
}

// Fold one parameter's pending terms into S and H: the scores' coefficients
// cs, cc of dSigma_s/dtheta, dSigma_c/dtheta, then the increment
// v = -lp dSigma_t/dtheta + extra of S and its H term v Q.  Terms that are
// exactly zero are skipped (S and H never hold -0, so the values are
// unchanged).  The updates are device atomics whose result is not used: an
// entry belongs to one particle, so they apply in program order and leave
// exactly the sums of load-add-store, without a memory round trip per
// parameter (h - x and h + (-x) are the same IEEE operation).
__device__ __forceinline__ void fold_one(const KArgs& A, uint32_t s, int kl, double ds, double dc, double extra,
                                         const LrPending& L)
{
    // parameter-major rows: lanes of a wave touch neighbouring slots
    const size_t o = (size_t)kl * A.sstride + s;
    if (L.cs != 0.0 || L.cc != 0.0) atomicAdd(&A.H[o], -(L.cs * ds + L.cc * dc));
    double v = -L.lp * (ds + dc);
    v += extra;
    if (v != 0.0) {
        atomicAdd(&A.S[o], v);
        if (L.Q != 0.0) atomicAdd(&A.H[o], v * L.Q);
    }
}

// Increments of S summed in LDS (the fused gradient flight kernel): a lane's
// particle stays in one material from its first event of a launch to its
// last (a surface crossing ends its launch), so the increments of that
// material's parameters can be summed in the workgroup's LDS and added to S
// once, when the particle stops, instead of once per collision.  Materials
// on the folds' fast path (each nuclide with at most one scatter and one
// capture parameter) of at most kAccNuc nuclides: entry 2k (scatter) and
// 2k + 1 (capture) of nuclide slot k; other materials of at most kAccEntries
// parameters: entry q of their q-th parameter.  Entry e of a lane is
// acc[e * blockDim.x] (consecutive lanes in consecutive words).  H updates, rare (the particle has scored), stay
// direct.  S then receives the same increments, summed in a different order
// (rounding only).
// fold_one with the increment of S added to the LDS entry *acc.
__device__ __forceinline__ void fold_acc(const KArgs& A, uint32_t s, int kl, double ds, double dc, double extra,
                                         const LrPending& L, double* acc)
{
    const size_t o = (size_t)kl * A.sstride + s;
    if (L.cs != 0.0 || L.cc != 0.0) atomicAdd(&A.H[o], -(L.cs * ds + L.cc * dc));
    double v = -L.lp * (ds + dc);
    v += extra;
    if (v != 0.0) {
        // the entry is this lane's alone: an LDS add whose result is not used
        __hip_atomic_fetch_add(acc, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (L.Q != 0.0) atomicAdd(&A.H[o], v * L.Q);
    }
}

// Whether a lane in material m sums its S increments in LDS.
__device__ __forceinline__ bool acc_material(const KArgs& A, int m)
{
    if (m < 0) return false;
    const mc_material_desc M = A.mat[m];
    return A.mfold[m] ? M.nuc_count <= kAccNuc : (M.param_count > 0 && M.param_count <= kAccEntries);
}

// Adds the lane's LDS sums for material m to S and clears them.
__device__ __forceinline__ void flush_acc(const KArgs& A, uint32_t s, int m, double* acc)
{
    const mc_material_desc M = A.mat[m];
    const uint32_t bd = blockDim.x;
    if (!A.mfold[m]) {
        for (int q = 0; q < M.param_count; ++q) {
            double* e = acc + (size_t)q * bd;
            const double v = e[0];
            e[0] = 0.0;
            if (v != 0.0) atomicAdd(&A.S[(size_t)A.mpar[M.param_begin + q].kl * A.sstride + s], v);
        }
        return;
    }
    for (int k = 0; k < M.nuc_count; ++k) {
        const DNucFold F = A.nfold[M.nuc_begin + k];
        double* e = acc + (size_t)(2 * k) * bd;
        const double vs = e[0], vc = e[bd];
        e[0] = 0.0;
        e[bd] = 0.0;
        if (F.kl_s >= 0 && vs != 0.0) atomicAdd(&A.S[(size_t)F.kl_s * A.sstride + s], vs);
        if (F.kl_c >= 0 && vc != 0.0) atomicAdd(&A.S[(size_t)F.kl_c * A.sstride + s], vc);
    }
}

// S and H entries are read back (at a history's end) in the kernel whose
// atomics updated them: coherent loads, so no stale first-level cache line
// is read.
__device__ __forceinline__ double lr_load(const double* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bit of material m in a slot's visited-materials mask.
__device__ __forceinline__ uint32_t mat_bit(const DProb& PR, int m)
{
    const int l = m - PR.mat_begin;
    return l < 31 ? 1u << l : 1u << 31;
}

// Surface crossing: fold the flight that ended on the surface, in material m
// of the cell it leaves (about one flight in a hundred in a shield).  A
// crossing runs alone in its lane, so its time is its chain of memory round
// trips: every value a fold needs is loaded before its first update is
// issued (a material whose nuclides each carry at most one scatter and one
// capture parameter and whose record holds their sigma_s and sigma_c pairs
// reads them from that one record).
__device__ void fold_surface(const KArgs& A, uint32_t s, const DProb& PR, int m, const mc_material_desc& M, double E,
                             XsPoint& P, int gm)
{
    const LrPending L = lr_pending(A, s);
    const DMatRec R = A.mrec[m];
    if (gm >= 0 && R.sc_off >= 0 && A.mfold[m]) {
        constexpr int kN = 4;                        // nuclides per chunk
        const double* r = rec_at(A.rec_sc, R.sc_off, R.sc_stride, P);
        for (int k0 = 0; k0 < M.nuc_count; k0 += kN) {
            int ks[kN], kc[kN];
            double ds[kN], dc[kN];
#pragma unroll
            for (int t = 0; t < kN; ++t) {
                if (k0 + t < M.nuc_count) {
                    const int k = k0 + t, j = M.nuc_begin + k;
                    const DNucFold F = A.nfold[j];
                    ks[t] = F.kl_s;
                    kc[t] = F.kl_c;
                    const double N = A.mat_dens[j];
                    ds[t] = xs_pair(r + 2 * k, P) * N;
                    dc[t] = xs_pair(r + R.c_at + 2 * k, P) * N;
                }
            }
#pragma unroll
            for (int t = 0; t < kN; ++t) {
                if (k0 + t < M.nuc_count) {
                    if (ks[t] >= 0) fold_one(A, s, ks[t], ds[t], 0.0, 0.0, L);
                    if (kc[t] >= 0) fold_one(A, s, kc[t], 0.0, dc[t], 0.0, L);
                }
            }
        }
    } else {
        for (int q = 0; q < M.param_count; ++q) {
            const DSig d = param_dsig(A, M, q, E, P, gm);
            fold_one(A, s, d.k - PR.param_begin, d.ds, d.dc, 0.0, L);
        }
    }
    if (L.cs != 0.0 || L.cc != 0.0) {
        A.cs[s] = 0.0;
        A.cc[s] = 0.0;
    }
}

__device__ __forceinline__ void close_bin(const KArgs& A, double* row, const DProb& PR, int b)
{
    const double h = row[b];
    if (h != 0.0) {
        // this workgroup's private row: no contention with other workgroups
        const size_t k = (size_t)blockIdx.x * A.bin_row + (size_t)(PR.bin_begin + b);
        atomicAdd(&A.psum[k], h);
        atomicAdd(&A.psq[k], h * h);
        row[b] = 0.0;
    }
}

// Closes out up to kCloseChunk of a particle's parameters, the t-th being
// kof(t) (t < n): their S and H entries are all read before the first update
// is issued.  The gradient term g = Q S - H goes to the workgroup's moments
// or, for a member of split family fam >= 0, to the family's sum; S and H
// are cleared for the slot's next particle.
constexpr int kCloseChunk = 4;

template <class KOf>
__device__ __forceinline__ void close_chunk(const KArgs& A, const DProb& PR, double* Srow, double* Hrow, size_t ks,
                                            double Q, bool scored, int fam, int n, KOf kof)
{
    int kk[kCloseChunk];
    double sv[kCloseChunk], hv[kCloseChunk];
#pragma unroll
    for (int t = 0; t < kCloseChunk; ++t)
        if (t < n) kk[t] = kof(t);
#pragma unroll
    for (int t = 0; t < kCloseChunk; ++t) {
        if (t < n && scored) {
            sv[t] = lr_load(Srow + (size_t)kk[t] * ks);
            hv[t] = lr_load(Hrow + (size_t)kk[t] * ks);
        }
    }
#pragma unroll
    for (int t = 0; t < kCloseChunk; ++t) {
        if (t < n) {
            const int k = kk[t];
            if (scored) {
                const double g = Q * sv[t] - hv[t];
                if (g != 0.0) {
                    if (fam >= 0) {
                        atomicAdd(&A.famG[(size_t)k * A.cap + fam], g);     // the family's sum
                    } else {
                        const size_t j = (size_t)blockIdx.x * A.par_row + (size_t)(PR.param_begin + k);
                        atomicAdd(&A.pgsum[j], g);
                        atomicAdd(&A.pgsq[j], g * g);
                    }
                }
                __hip_atomic_store(Hrow + (size_t)k * ks, 0.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __hip_atomic_store(Srow + (size_t)k * ks, 0.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Gradient closeout of particle s (Estimator::end_history for the score
// function): S and H are non-zero only for parameters of the materials the
// particle has been in.  Their entries are read (when it scored: a history
// that touched no bin made no score with v != 0, so Q, H and the pending
// coefficients are zero and it adds nothing) and cleared for the slot's next
// particle.  Its sum Q S_end - H goes to the workgroup's moments, or for a
// member of a split family (fam >= 0) to the family's sum, which the family's
// last member closes out.
template <bool kFlush>
__device__ __forceinline__ void grad_closeout(const KArgs& A, uint32_t s, const DProb& PR, double E, bool scored, int fam)
{
    {
        const uint32_t vm = A.vmat[s];
        if (vm != 0) {
            double* Srow = A.S + s;
            double* Hrow = A.H + s;
            const size_t ks = A.sstride;
            const double Q = A.Qw[s];
            if (kFlush && scored) {
                // scores still pending (surface scores of the cell the particle ends in)
                const double cs = A.cs[s], cc = A.cc[s];
                const int m = A.pmat[s];
                if ((cs != 0.0 || cc != 0.0) && m >= 0) {
                    const mc_material_desc M = A.mat[m];
                    XsPoint P;
                    const int gm = material_point(A, m, E, P);
                    for (int q = 0; q < M.param_count; ++q) {
                        const DSig d = param_dsig(A, M, q, E, P, gm);
                        atomicAdd(&Hrow[(size_t)(d.k - PR.param_begin) * ks], -(cs * d.ds + cc * d.dc));
                    }
                }
            }
            // the history's gradient sum over its scores gv * S = Q S_end - H,
            // kCloseChunk parameters at a time with their S and H loaded
            // before any update is issued (one memory round trip per chunk)
            if (vm & (1u << 31)) {
                for (int k0 = 0; k0 < PR.param_count; k0 += kCloseChunk)
                    close_chunk(A, PR, Srow, Hrow, ks, Q, scored, fam, PR.param_count - k0,
                                [&](int t) { return k0 + t; });
            } else {
                for (uint32_t b = vm; b != 0; b &= b - 1) {
                    const mc_material_desc M = A.mat[PR.mat_begin + __ffs(b) - 1];
                    for (int q0 = 0; q0 < M.param_count; q0 += kCloseChunk)
                        close_chunk(A, PR, Srow, Hrow, ks, Q, scored, fam, M.param_count - q0,
                                    [&](int t) { return A.mat_param[M.param_begin + q0 + t] - PR.param_begin; });
                }
            }
        }
    }
}

// A member of a split family ends: the last one closes out the family's row
// (Estimator::end_history for the whole history, parent and copies).  The
// slot's status becomes ST_HELD: its row may still be a family's.
template <bool kGrad>
__device__ uint8_t finish_family(const KArgs& A, uint32_t s, const DProb& PR)
{
    const int f = A.fam[s];
    A.fam[s] = -1;
    __threadfence();                              // this member's scores before the count
    if (atomicSub(&A.fcount[f], 1) == 1) {
        __threadfence();                          // every member's scores after it
        double* row = A.hist + (size_t)f * (size_t)A.bin_stride;
        for (int b = 0; b < PR.bin_count; ++b) close_bin(A, row, PR, b);
        A.tcount[f] = 0;
        if (kGrad) {
            // the history's gradient: the sum over its members
            for (int k = 0; k < PR.param_count; ++k) {
                double* pg = A.famG + (size_t)k * A.cap + f;
                const double G = lr_load(pg);
                if (G != 0.0) {
                    const size_t j = (size_t)blockIdx.x * A.par_row + (size_t)(PR.param_begin + k);
                    atomicAdd(&A.pgsum[j], G);
                    atomicAdd(&A.pgsq[j], G * G);
                    __hip_atomic_store(pg, 0.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
    }
    return ST_HELD;
}

// End of a history: close out its tally row (Estimator::end_history).
// Returns the status the caller stores in slot s.
// kFlush: scores may still be pending (not after a collision, which folds them).
template <bool kGrad, bool kFlush>
__device__ __forceinline__ uint8_t finish_history(const KArgs& A, uint32_t s, int p, const DProb& PR, uint64_t rng, Vec3 pos,
                                  double E, uint32_t nfl)
{
    if constexpr (!kGrad) {
        // particle comb: the history goes on with the next copy of its bank
        // (only a comb problem's particles push copies on their slot's bank)
        if (A.bsize && A.bsize[s] > 0) return ST_POP;
    }
    if (A.split_any && A.fam[s] >= 0) {
        if (kGrad) grad_closeout<kFlush>(A, s, PR, E, true, A.fam[s]);
        return finish_family<kGrad>(A, s, PR);
    }
    double* row = A.hist + (size_t)s * (size_t)A.bin_stride;
    const uint32_t nt = A.tcount[s];
    const uint32_t* tb = A.tbins + (size_t)s * kTouched;
    double tot = 0.0;
    if (A.trace) {
        // sum of the history's bins in ascending bin order (the oracle's order)
        if (nt <= kTouched) {
            int list[kTouched];
            for (uint32_t j = 0; j < nt; ++j) list[j] = tb[j];
            for (uint32_t j = 1; j < nt; ++j)
                for (uint32_t k = j; k > 0 && list[k - 1] > list[k]; --k) {
                    const int t = list[k]; list[k] = list[k - 1]; list[k - 1] = t;
                }
            int prev = -1;
            for (uint32_t j = 0; j < nt; ++j)
                if (list[j] != prev) { tot += row[list[j]]; prev = list[j]; }
        } else {
            for (int b = 0; b < PR.bin_count; ++b) tot += row[b];
        }
    }
    if (nt <= kTouched) {
        for (uint32_t j = 0; j < nt; ++j) close_bin(A, row, PR, tb[j]);
    } else {
        for (int b = 0; b < PR.bin_count; ++b) close_bin(A, row, PR, b);
    }
    A.tcount[s] = 0;
    if (A.trace) {
        const uint64_t local = A.hid[s] - A.run_hb[p];
        if (local < A.trace_count) {
            mc_trace_record& T = A.trace[(size_t)p * A.trace_count + local];
            T.rng = rng;
            T.flights = nfl;
            T.x = pos.x; T.y = pos.y; T.z = pos.z;
            T.E = E;
            T.score = tot;
        }
    }
    if (kGrad) grad_closeout<kFlush>(A, s, PR, E, nt != 0, -1);
    return ST_DEAD;
}

// ---------------------------------------------------------------- kernels

// Pass 1 of the refill: live and free slots per workgroup.  The first thread
// also rolls the counters over from the previous iteration (no other block
// reads them before the next kernel).
__global__ void __launch_bounds__(1024) k_refill_count(KArgs A)
{
    __shared__ uint32_t lds[32];
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        Counters* c = A.ctr;
        c->tA += c->nA;            // event counts of the previous iteration
        c->tS += c->nS;
        c->tC += c->nC;
        c->next_hist += c->claimed;
        c->claimed = 0;
        c->chead += c->cplaced;    // copies the last refill placed left the ring
        c->cplaced = 0;
        c->nS = 0;
        c->nC = 0;
    }
    if (A.events_per_launch > 1 && blockIdx.x == 0) {
        // flights the last flight launch took beyond the first per particle
        // (the flight kernel runs on this grid, one entry per workgroup)
        __shared__ unsigned long long xs;
        if (threadIdx.x == 0) xs = 0;
        unsigned long long x = 0;
        for (uint32_t b = threadIdx.x; b < gridDim.x; b += blockDim.x) {
            x += A.wg_xfl[b];
            A.wg_xfl[b] = 0;
        }
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
        __syncthreads();
        if (lane_id() == 0 && x != 0) atomicAdd(&xs, x);
        __syncthreads();
        if (threadIdx.x == 0 && xs != 0) {
            A.ctr->tA += xs;
            A.ctr->xfl += xs;
        }
    }
    uint32_t lo, hi;
    block_range(A.ctr->active_cap, lo, hi);
    uint32_t live = 0, freec = 0, held = 0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const uint8_t st = A.status[i];
        live += st == ST_ALIVE;
        freec += st == ST_DEAD || st == ST_EMPTY;
        if (st == ST_HELD) {
            if (A.fcount[i] > 0) ++held;
            else ++freec;
        }
    }
    block_sum2(live, freec, lds);
    if (A.split_any) {
        uint32_t zero = 0;
        block_sum2(held, zero, lds);
    }
    if (threadIdx.x == 0) {
        A.wg_cnt[2 * blockIdx.x] = live;
        A.wg_cnt[2 * blockIdx.x + 1] = freec;
        if (A.split_any) A.wg_held[blockIdx.x] = held;
    }
}

// A new history in slot i: SourceBank::get_source + SourcePoint::get_source
// (Source.cpp:20-45).  jstate is the LCG multiplier g^(stride * hid) of its
// run key; times run_jd[p] it is g^(stride * key) of the history's key, so
// the stream starts where RN_init_particle (Random.cpp:196-204) puts it.
template <bool kGrad>
__device__ __forceinline__ void start_history(const KArgs& A, uint32_t i, uint64_t hid, int p, uint64_t jstate)
{
    const DProb PR = A.prob[p];
    uint64_t rng = (((jstate * A.run_jd[p]) & kRnMask) * PR.seed) & kRnMask;
    const double xi = urand(rng);
    int si = binary_search(xi, A.src_cdf + A.cdf_off[p], PR.src_count + 1);
    si = si < 0 ? 0 : (si >= PR.src_count ? PR.src_count - 1 : si);   // cdf[n] may round below 1
    const mc_source_desc& S = A.src[PR.src_begin + si];
    double ux, uy, uz;
    if (S.dir_kind == MC_DIR_DELTA) {
        ux = S.dir[0]; uy = S.dir[1]; uz = S.dir[2];
    } else {                                        // Distribution.cpp:78-92
        const double mu = 2.0 * urand(rng) - 1.0;
        double ca, sa;
        sincos_2pi(urand(rng), ca, sa);
        const double c = sqrt(1.0 - mu * mu);
        uy = ca * c;
        uz = sa * c;
        ux = mu;
    }
    double E0;
    if (S.energy_kind == MC_ENERGY_UNIFORM)          // Distribution.cpp:30-33
        E0 = S.energy[0] + urand(rng) * (S.energy[1] - S.energy[0]);
    else
        E0 = S.energy[0];
    A.x[i] = S.pos[0]; A.y[i] = S.pos[1]; A.z[i] = S.pos[2];
    A.u[i] = ux; A.v[i] = uy; A.w[i] = uz;
    A.E[i] = E0;
    A.wgt[i] = 1.0;
    A.rng[i] = rng;
    A.hid[i] = hid;
    A.pcell[i] = S.cell;
    A.pmat[i] = A.cell[S.cell].material;
    A.pprob[i] = p;
    A.nflight[i] = 0;
    A.tcount[i] = 0;
    A.status[i] = ST_ALIVE;
    if (kGrad) {
        // S and H of the slot are zero (cleared by the previous history's closeout)
        const int m0 = A.cell[S.cell].material;
        A.vmat[i] = m0 >= 0 && A.mat[m0].param_count > 0 ? mat_bit(PR, m0) : 0u;
        A.Qw[i] = 0.0;
        A.cs[i] = 0.0;
        A.cc[i] = 0.0;
        A.lpend[i] = 0.0;
    }
}

// A split copy from ring entry k takes slot i (its family's row stays where
// the family began; the slot's own row is not touched).
template <bool kGrad>
__device__ __forceinline__ void start_copy(const KArgs& A, uint32_t i, uint64_t k)
{
    const int c = A.ccell[k];
    if (kGrad) {
        // S of the particle at the split; the slot's S and H are clear (its
        // last history's closeout cleared them), and so are Q and the rest
        const int p = A.cprob[k];
        const DProb PR = A.prob[p];
        A.vmat[i] = A.rvmat[k];
        for (int q = 0; q < PR.param_count; ++q) A.S[(size_t)q * A.sstride + i] = A.rS[(size_t)q * A.ccap + k];
        A.Qw[i] = 0.0;
        A.cs[i] = 0.0;
        A.cc[i] = 0.0;
        A.lpend[i] = 0.0;
    }
    A.x[i] = A.cx[k]; A.y[i] = A.cy[k]; A.z[i] = A.cz[k];
    A.u[i] = A.cu[k]; A.v[i] = A.cv[k]; A.w[i] = A.cw[k];
    A.E[i] = A.cE[k];
    A.wgt[i] = A.cwgt[k];
    A.rng[i] = A.crng[k];
    A.hid[i] = A.chid[k];
    A.pcell[i] = c;
    A.pmat[i] = A.cell[c].material;
    A.pprob[i] = A.cprob[k];
    A.nflight[i] = A.cnfl[k];
    A.fam[i] = A.cfam[k];
    A.status[i] = ST_ALIVE;
}

// Block-wide exclusive prefix of two counts per thread (wave shuffles, then
// the wave totals through LDS); also returns the block totals.
__device__ __forceinline__ void block_scan2(uint32_t a, uint32_t b, uint32_t& a_ex, uint32_t& b_ex, uint32_t& a_tot,
                                            uint32_t& b_tot, uint32_t* lds)
{
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t ia = a, ib = b;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t ta = __shfl_up(ia, o), tb = __shfl_up(ib, o);
        if (lane >= (uint32_t)o) { ia += ta; ib += tb; }
    }
    if (lane == 63) { lds[2 * wave] = ia; lds[2 * wave + 1] = ib; }
    __syncthreads();
    uint32_t wa = 0, wb = 0, ta = 0, tb = 0;
    for (uint32_t k = 0; k < nw; ++k) {
        const uint32_t x = lds[2 * k], y = lds[2 * k + 1];
        if (k < wave) { wa += x; wb += y; }
        ta += x;
        tb += y;
    }
    __syncthreads();
    a_ex = wa + ia - a;
    b_ex = wb + ib - b;
    a_tot = ta;
    b_tot = tb;
}

// Pass 2 of the refill: slot-ordered queue A and new histories, no atomics.
// The block finds the problem and the RN stream multiplier once, for its
// first new key; every other key of the block is a short step from it.  A
// block that starts many histories walks its range in tiles of consecutive
// slots, one slot per lane (the new histories' stores coalesce), ranking each
// tile with wave ballots and one LDS exchange; one that starts few (the
// steady state of long histories) gives each thread a run of consecutive
// slots and ranks them all with a single block scan.
struct RefillCtx {
    uint64_t next, lim, chead, hb, jb;
    int pb;
    bool cmode, ident;
};

// Slot i, free and of rank fidx among the free slots, gets a particle: a
// split copy from the ring, or the history of key next + fidx.
template <bool kGrad>
__device__ __forceinline__ void refill_slot(const KArgs& A, const RefillCtx& R, uint32_t i, uint64_t fidx, int& p)
{
    if (R.cmode) {
        start_copy<kGrad>(A, i, (R.chead + fidx) % A.ccap);
        return;
    }
    const uint64_t hid = R.next + fidx;
    for (int step = 0; p + 1 < A.n_prob && A.run_hb[p + 1] <= hid; ++step) {
        if (step == 4) { p = find_problem(A, hid); break; }
        ++p;
    }
    start_history<kGrad>(A, i, hid, p, history_seed(R.jb, hid - R.hb, A.jump));
}

__device__ __forceinline__ bool refill_free(const KArgs& A, uint8_t st, uint32_t i, bool cmode)
{
    return st == ST_DEAD || st == ST_EMPTY || (st == ST_HELD && (cmode || A.fcount[i] == 0));
}

// Slots a refill workgroup lists before it starts their particles (at least
// twice the largest workgroup).
constexpr uint32_t kRefillList = 2048;

template <bool kGrad>
__global__ void __launch_bounds__(1024) k_refill_scatter(KArgs A)
{
    __shared__ uint32_t lds[64];
    __shared__ uint32_t s_list[kRefillList];
    __shared__ uint64_t s_jb;
    __shared__ int s_pb;
    const uint32_t cap = A.ctr->active_cap;
    uint32_t lo, hi;
    block_range(cap, lo, hi);
    const uint64_t next = A.ctr->next_hist, end = A.ctr->end_hist;
    const uint64_t remaining = end > next ? end - next : 0;
    // split copies waiting in the ring go first: a refill that has any places
    // copies only, into free slots and slots whose row a family holds
    const uint64_t chead = A.ctr->chead;
    const uint64_t pend = A.split_any ? A.ctr->ctail - chead : 0;
    const bool cmode = pend > 0;
    uint32_t live_before, free_before, live_tot, free_tot;
    wg_prefix(A.wg_cnt, live_before, free_before, live_tot, free_tot, lds);
    if (cmode) {
        uint32_t hb = 0, ht = 0;
        for (uint32_t v = threadIdx.x; v < gridDim.x; v += blockDim.x) {
            const uint32_t h = A.wg_held[v];
            if (v < blockIdx.x) hb += h;
            ht += h;
        }
        block_sum2(hb, ht, lds);
        free_before += hb;
        free_tot += ht;
    }
    RefillCtx R;
    R.next = next;
    R.chead = chead;
    R.cmode = cmode;
    // lim: free slots (in rank order) that get a particle
    R.lim = cmode ? (pend < free_tot ? pend : (uint64_t)free_tot)
                  : ((uint64_t)free_tot < remaining ? (uint64_t)free_tot : remaining);
    // every free slot gets a history: all slots live, queue A is the identity
    R.ident = !cmode && (uint64_t)free_tot <= remaining && live_tot + free_tot == cap;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        A.ctr->claimed = cmode ? 0 : R.lim;
        A.ctr->cplaced = cmode ? (uint32_t)R.lim : 0u;
        A.ctr->nA = live_tot + (uint32_t)R.lim;
        A.ctr->qa_ident = R.ident ? 1u : 0u;
        A.ctr->fcur = 0;
    }
    R.hb = next + (uint64_t)free_before;                  // the block's first new key
    // thread 0 finds the block's seed while the statuses load; the first
    // block scan's barrier publishes it
    const bool seed = !cmode && (uint64_t)free_before < R.lim;
    auto find_seed = [&]() {
        if (threadIdx.x == 0 && seed) {
            s_pb = find_problem(A, R.hb);
            s_jb = history_seed(1, R.hb, A.jump);          // g^(stride * hb)
        }
    };
    int p = 0;
    // particles this block places (its free slots of rank below lim)
    const uint64_t placed = R.lim > free_before ? R.lim - free_before : 0;
    if (placed * 16 >= hi - lo) {
        const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
        const uint64_t below = (1ull << lane) - 1ull;
        uint64_t fbase = free_before, lbase = live_before;
        // the slots that get a particle are listed in LDS in rank order and
        // started by the whole block at once (one dependent chain of loads per
        // thread per flush, not one per tile); entry k has rank list_f0 + k
        uint32_t listed = 0;
        uint64_t list_f0 = free_before;
        // statuses of the next tile are loaded ahead of this tile's scan
        uint8_t st_next = lo + threadIdx.x < hi ? A.status[lo + threadIdx.x] : ST_SURF;
        find_seed();
        bool first = true;
        for (uint32_t base = lo; base < hi; base += blockDim.x) {
            const uint32_t i = base + threadIdx.x;
            const uint8_t st = st_next;                    // ST_SURF: beyond the range
            st_next = i + blockDim.x < hi ? A.status[i + blockDim.x] : ST_SURF;
            const bool live = st == ST_ALIVE, fr = st != ST_SURF && refill_free(A, st, i, cmode);
            const uint64_t ml = __ballot(live), mf = __ballot(fr);
            if (lane == 0) {
                lds[2 * wave] = (uint32_t)__popcll(ml);
                lds[2 * wave + 1] = (uint32_t)__popcll(mf);
            }
            __syncthreads();
            uint32_t wl = 0, wf = 0, tl = 0, tf = 0;
            for (uint32_t k = 0; k < nw; ++k) {
                const uint32_t x = lds[2 * k], y = lds[2 * k + 1];
                if (k < wave) { wl += x; wf += y; }
                tl += x;
                tf += y;
            }
            __syncthreads();
            if (first) {
                R.pb = s_pb;
                R.jb = s_jb;
                p = R.pb;
                first = false;
            }
            const uint64_t fidx = fbase + wf + (uint64_t)__popcll(mf & below);
            const uint64_t act = lbase + wl + (uint64_t)__popcll(ml & below) + (fidx < R.lim ? fidx : R.lim);
            if (live) {
                if (!R.ident) A.qA[act] = i;
            } else if (fr) {
                if (fidx < R.lim) {
                    s_list[fidx - list_f0] = i;
                    if (!R.ident) A.qA[act] = i;
                } else if (st == ST_DEAD) {
                    A.status[i] = ST_EMPTY;
                }
            }
            // slots of this tile that get a particle (the same in every thread)
            const uint64_t got = R.lim > fbase ? (R.lim - fbase < tf ? R.lim - fbase : (uint64_t)tf) : 0;
            listed += (uint32_t)got;
            fbase += tf;
            lbase += tl;
            if (listed + blockDim.x > kRefillList || base + blockDim.x >= hi) {
                __syncthreads();
                for (uint32_t k = threadIdx.x; k < listed; k += blockDim.x) {
                    int pk = R.pb;
                    refill_slot<kGrad>(A, R, s_list[k], list_f0 + k, pk);
                }
                __syncthreads();
                list_f0 += listed;
                listed = 0;
            }
        }
    } else {
        const uint32_t per = (hi - lo + blockDim.x - 1) / blockDim.x;
        const uint32_t a = min(hi, lo + threadIdx.x * per), b = min(hi, a + per);
        uint32_t nlive = 0, nfree = 0;
        for (uint32_t i = a; i < b; ++i) {
            const uint8_t st = A.status[i];
            nlive += st == ST_ALIVE;
            nfree += refill_free(A, st, i, cmode);
        }
        find_seed();
        uint32_t live_ex, free_ex, lt, ft;
        block_scan2(nlive, nfree, live_ex, free_ex, lt, ft, lds);
        R.pb = s_pb;
        R.jb = s_jb;
        p = R.pb;
        uint64_t fidx = (uint64_t)free_before + free_ex;      // rank of my next free slot
        uint64_t act = (uint64_t)live_before + live_ex + (fidx < R.lim ? fidx : R.lim);
        for (uint32_t i = a; i < b; ++i) {
            const uint8_t st = A.status[i];
            if (st == ST_ALIVE) {
                if (!R.ident) A.qA[act] = i;
                ++act;
            } else if (refill_free(A, st, i, cmode)) {
                if (fidx < R.lim) {
                    refill_slot<kGrad>(A, R, i, fidx, p);
                    if (!R.ident) A.qA[act] = i;
                    ++act;
                } else if (st == ST_DEAD) {
                    A.status[i] = ST_EMPTY;
                }
                ++fidx;
            }
        }
    }
}

// Dynamic LDS of the lookup kernels: the bucket table of the LDS grid.
extern __shared__ uint16_t mc_lds_buckets[];

// use_lds: stage the bucket table in LDS (worth it when a launch has many
// lookups spread over the whole grid; in the event loop the energies are
// concentrated and the table lines are in cache, so the loop passes 0).
__global__ void __launch_bounds__(1024) k_xs_lookup(KArgs A, int use_lds)
{
    const uint16_t* sb = use_lds ? stage_buckets(A, mc_lds_buckets) : nullptr;
    // two independent lookups per thread per step keep more gathers in flight
    const uint32_t n = A.ctr->nA, ident = A.ctr->qa_ident;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += 2 * stride) {
        const uint32_t i2 = i + stride;
        const uint32_t s = qa_at(A.qA, ident, i);
        const uint32_t s2 = i2 < n ? qa_at(A.qA, ident, i2) : s;
        const int m = A.pmat[s], m2 = A.pmat[s2];
        const double E = A.E[s], E2 = A.E[s2];
        int g = 0, g2 = 0;
        const double r = m >= 0 ? lookup_sigt(A, sb, m, E, g) : 0.0;
        const double r2 = m2 >= 0 ? lookup_sigt(A, sb, m2, E2, g2) : 0.0;
        // the interval is kept for the collision and surface kernels (same E)
        A.sigt[s] = r;
        A.gidx[s] = g;
        if (i2 < n) {
            A.sigt[s2] = r2;
            A.gidx[s2] = g2;
        }
    }
}

// weight_roulette, population_control.cpp:9-15
__device__ __forceinline__ void weight_roulette(double& w, bool& alive, uint64_t& rng, const DProb& PR)
{
    if (w < PR.wr) {
        if (urand(rng) < w / PR.ws) w = PR.ws;
        else { alive = false; w = 0.0; }
    }
}

// surface_hit, general.cpp:89-115, then cell_importance (population_control.cpp:
// 21-49) and weight_roulette, for a flight of slot s that ended on surface sf at
// p.  Writes the particle back and returns its new status.  Kept out of line:
// about one flight in a hundred crosses a surface in a shield, and the flight
// kernel's registers stay for the common path.
// Importance splitting into n (population_control.cpp:21-49): the particle
// goes on as one member of its history's split family and n - 1 copies of it
// wait in the copy ring for a slot; copy i (from 0) starts (i + 1) 2^50 draws
// ahead of the particle's stream.  The family's members score into the tally
// row of the slot where the family began, and the last member to end closes
// it out.
//
// Back-pressure: copies the ring has no room for stay with the particle,
// which waits at the surface (ppend[s] copies still to place, the next being
// copy pfirst[s]) and places them at its next turns, as the refill moves
// copies into free slots; it flies on once all are placed.  Its state is the
// state at the split until then, so every copy is the one an immediate
// placement would have made.  The host grows the ring when particles wait.

// The particle of slot s becomes a member of a split family of n members.
__device__ __forceinline__ void join_family(const KArgs& A, uint32_t s, int n)
{
    const int f = A.fam[s];
    if (f < 0) {
        A.fam[s] = (int)s;
        atomicAdd(&A.fcount[s], n);          // the particle and its copies
    } else {
        atomicAdd(&A.fcount[f], n - 1);
    }
}

// Places copies first .. first + count - 1 of the split of slot s's particle
// (state p, u, E, w, rng, nfl in cell `cell`) in the ring, as many as there is
// room for, and returns how many.  Only the refill takes copies out of the
// ring, so outside it the room can only shrink: a compare-and-swap on the
// tail reserves a run of entries.
template <bool kGrad>
__device__ uint32_t push_copies(const KArgs& A, uint32_t s, uint32_t first, uint32_t count, Vec3 p, Vec3 u,
                                double E, double w, uint64_t rng, uint32_t nfl, int cell, int p_id)
{
    // copies the last refill placed have left the ring (the head is rolled
    // past them at the next refill)
    const uint64_t head = A.ctr->chead + A.ctr->cplaced;
    unsigned long long* tailp = (unsigned long long*)&A.ctr->ctail;
    unsigned long long t0 = __hip_atomic_load(tailp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t take = 0;
    for (;;) {
        const uint64_t used = (uint64_t)t0 - head;
        const uint64_t room = used < A.ccap ? A.ccap - used : 0;
        take = room < count ? (uint32_t)room : count;
        if (take == 0) return 0;
        const unsigned long long prev = atomicCAS(tailp, t0, t0 + take);
        if (prev == t0) break;
        t0 = prev;
    }
    const int f = A.fam[s];
    const uint64_t hid = A.hid[s];
    if (kGrad) {
        // the copies carry the particle's score function S (its H, Q and
        // pending coefficients stay with it) and its visited materials, the
        // cell it enters included
        const DProb PR = A.prob[p_id];
        uint32_t vm = A.vmat[s];
        const int m1 = A.cell[cell].material;
        if (m1 >= 0 && A.mat[m1].param_count > 0) vm |= mat_bit(PR, m1);
        for (uint32_t i = 0; i < take; ++i) A.rvmat[(t0 + i) % A.ccap] = vm;
        for (int k = 0; k < PR.param_count; ++k) {
            const double v = lr_load(A.S + (size_t)k * A.sstride + s);
            for (uint32_t i = 0; i < take; ++i) A.rS[(size_t)k * A.ccap + (t0 + i) % A.ccap] = v;
        }
    }
    uint64_t r = rng;
    for (uint32_t i = 0; i <= first; ++i) r = (r * A.jump_copy) & kRnMask;    // copy `first`
    for (uint32_t i = 0; i < take; ++i) {
        if (i > 0) r = (r * A.jump_copy) & kRnMask;
        const uint64_t k = (t0 + i) % A.ccap;
        A.cx[k] = p.x; A.cy[k] = p.y; A.cz[k] = p.z;
        A.cu[k] = u.x; A.cv[k] = u.y; A.cw[k] = u.z;
        A.cE[k] = E;
        A.cwgt[k] = w;
        A.crng[k] = r;
        A.chid[k] = hid;
        A.ccell[k] = cell;
        A.cprob[k] = p_id;
        A.cfam[k] = f;
        A.cnfl[k] = nfl;
    }
    return take;
}

// A split into n at a crossing: the family grows by n - 1 and the copies go
// to the ring; returns how many are left to place (0: all placed).
template <bool kGrad>
__device__ __forceinline__ uint32_t spawn_copies(const KArgs& A, uint32_t s, int n, Vec3 p, Vec3 u, double E, double w,
                                                 uint64_t rng, uint32_t nfl, int cell, int p_id)
{
    if (!A.split_any) {                      // (cannot happen: a split needs an importance rise)
        atomicAdd(&A.ctr->err_split, 1u);
        return 0;
    }
    join_family(A, s, n);
    const uint32_t nc = (uint32_t)(n - 1);
    return nc - push_copies<kGrad>(A, s, 0, nc, p, u, E, w, rng, nfl, cell, p_id);
}

// Problems with a particle comb keep a history's split copies on its slot's
// own bank (handler.cpp:19-29: the reference's Pbank, last in first out):
// copies first .. first + count - 1 of a split go on top in order, copy i
// starting (i + 1) 2^50 draws ahead of the particle's stream as ring copies
// do.  A full bank drops the copy and counts it (the run then fails).
__device__ __forceinline__ void push_bank(const KArgs& A, uint32_t s, uint32_t first, uint32_t count, Vec3 p, Vec3 u,
                                          double E, double w, uint64_t rng, int cell)
{
    int b = A.bsize[s];
    uint64_t r = rng;
    for (uint32_t i = 0; i < first; ++i) r = (r * A.jump_copy) & kRnMask;
    for (uint32_t i = 0; i < count; ++i) {
        r = (r * A.jump_copy) & kRnMask;
        if ((uint32_t)b >= A.bcap) {
            atomicAdd(&A.ctr->err_bank, 1u);
            continue;
        }
        const size_t k = (size_t)s * A.bcap + (uint32_t)b;
        A.bx[k] = p.x; A.by[k] = p.y; A.bz[k] = p.z;
        A.bu[k] = u.x; A.bv[k] = u.y; A.bw[k] = u.z;
        A.bE[k] = E;
        A.bwgt[k] = w;
        A.brng[k] = r;
        A.bcell[k] = cell;
        ++b;
    }
    A.bsize[s] = b;
}

// The end of one particle of a comb history (status ST_POP): Simulator::
// particle_comb (population_control.cpp:55-84) on the slot's bank, then the
// top copy starts in the slot (handler.cpp:23).  A bank of bank_max or more
// is combed to teeth copies: with W its total weight and w_avg =
// W / teeth, a tooth at u w_avg, u the next draw of the ended particle's
// stream (the reference draws from its one global stream), then every
// w_avg; walking the bank in order, a copy whose running weight sum passes
// the tooth is kept at weight w_avg and the tooth moves on (once per copy).
// Entries past the last kept one are the bank's first entry as it was, its
// weight included (the reference fills them from Pbank[0]); the k-th of
// them (from 1) starts k 2^50 draws ahead of that entry's stream, so no two
// entries share one.  The popped copy keeps the slot's history, tally row
// and flight count.
__device__ __forceinline__ void comb_pop(const KArgs& A, uint32_t s)
{
    const int p = A.pprob[s];
    const size_t base = (size_t)s * A.bcap;
    int n = A.bsize[s];
    if (n >= A.comb[2 * p]) {
        const int teeth = A.comb[2 * p + 1];
        uint64_t rng = A.rng[s];
        double W = 0.0;
        for (int i = 0; i < n; ++i) W += A.bwgt[base + i];
        const double w_avg = W / (double)teeth;
        double tooth = urand(rng) * w_avg;
        const double x0 = A.bx[base], y0 = A.by[base], z0 = A.bz[base];
        const double u0 = A.bu[base], v0 = A.bv[base], ww0 = A.bw[base];
        const double E0 = A.bE[base], w0 = A.bwgt[base];
        const uint64_t r0 = A.brng[base];
        const int c0 = A.bcell[base];
        double sum = 0.0;
        int j = 0;
        for (int i = 0; i < n && j < teeth; ++i) {     // j <= i: kept copies move down in place
            sum += A.bwgt[base + i];
            if (sum > tooth) {
                const size_t a = base + i, b = base + j;
                if (a != b) {
                    A.bx[b] = A.bx[a]; A.by[b] = A.by[a]; A.bz[b] = A.bz[a];
                    A.bu[b] = A.bu[a]; A.bv[b] = A.bv[a]; A.bw[b] = A.bw[a];
                    A.bE[b] = A.bE[a];
                    A.brng[b] = A.brng[a];
                    A.bcell[b] = A.bcell[a];
                }
                A.bwgt[b] = w_avg;
                tooth += w_avg;
                ++j;
            }
        }
        uint64_t r = r0;
        for (int k = j; k < teeth; ++k) {
            r = (r * A.jump_copy) & kRnMask;
            const size_t b = base + k;
            A.bx[b] = x0; A.by[b] = y0; A.bz[b] = z0;
            A.bu[b] = u0; A.bv[b] = v0; A.bw[b] = ww0;
            A.bE[b] = E0;
            A.bwgt[b] = w0;
            A.brng[b] = r;
            A.bcell[b] = c0;
        }
        n = teeth;
    }
    --n;
    const size_t k = base + n;
    const int c = A.bcell[k];
    A.x[s] = A.bx[k]; A.y[s] = A.by[k]; A.z[s] = A.bz[k];
    A.u[s] = A.bu[k]; A.v[s] = A.bv[k]; A.w[s] = A.bw[k];
    A.E[s] = A.bE[k];
    A.wgt[s] = A.bwgt[k];
    A.rng[s] = A.brng[k];
    A.pcell[s] = c;
    A.pmat[s] = A.cell[c].material;
    A.bsize[s] = n;
    A.status[s] = ST_ALIVE;
}

template <bool kGrad>
__device__ __forceinline__ uint8_t surface_crossing(const KArgs& A, uint32_t s, int sf, Vec3 p, Vec3 u, double E,
                                                 double w, uint64_t rng, uint32_t nfl, int cell_old, int p_id)
{
    const mc_surface_desc& Sd = A.surf[sf];
    const DProb PR = A.prob[p_id];
    int cell = cell_old;
    bool alive = true, waiting = false;
    if (kGrad) {
        // the flight that ended here, in the cell it leaves
        const int m = A.cell[cell_old].material;
        if (m >= 0) {
            const mc_material_desc M = A.mat[m];
            if (M.param_count > 0) {
                // a shared grid implies tables, so the lookup kernel stored the interval
                XsPoint P;
                const int gm = material_point_cached(A, m, s, E, P);
                fold_surface(A, s, PR, m, M, E, P, gm);
            }
        }
    }
    if (Sd.bc == MC_BC_TRANSMISSION) {
        p.x += u.x * kEpsFloat;
        p.y += u.y * kEpsFloat;
        p.z += u.z * kEpsFloat;
        cell = find_cell(A, sf, PR, p);
        if (cell < 0) {
            atomicAdd(&A.ctr->err_lost, 1u);
            cell = cell_old;
            alive = false;
            w = 0.0;
        }
    } else if (Sd.bc == MC_BC_VACUUM) {
        alive = false;
        w = 0.0;
    } else {
        u = surface_reflect(Sd.type, Sd.c, u);
        p.x += u.x * kEpsFloat;
        p.y += u.y * kEpsFloat;
        p.z += u.z * kEpsFloat;
    }
    for (int a = 0; a < Sd.est_count; ++a)
        score_attachment<kGrad>(A, s, Sd.est_begin + a, 0.0, cell, E, w, PR, 0.0, 0.0, false);
    // cell_importance, population_control.cpp:21-49
    const double Iold = A.cell[cell_old].importance;
    const double Inew = A.cell[cell].importance;
    if (Inew != Iold) {
        const double rat = Inew / Iold;
        if (rat < 1.0) {
            if (urand(rng) < rat) w = w / rat;
            else { alive = false; w = 0.0; }
        } else {
            const int nsplit = (int)floor(rat + urand(rng));
            w = w / (double)nsplit;
            if (nsplit > 1) {
                // comb problems: the copies go on the slot's own bank, at the
                // next k_resume, while the particle waits here
                const bool comb = !kGrad && A.comb && A.comb[2 * p_id + 1] > 0;
                const uint32_t left =
                    comb ? (uint32_t)(nsplit - 1) : spawn_copies<kGrad>(A, s, nsplit, p, u, E, w, rng, nfl, cell, p_id);
                if (left > 0) {
                    // the particle waits here, before its roulette, with the
                    // state its remaining copies are made from
                    A.ppend[s] = (int32_t)left;
                    A.pfirst[s] = nsplit - 1 - (int32_t)left;
                    if (!comb) atomicAdd((unsigned long long*)&A.ctr->waits, 1ull);
                    waiting = true;
                }
            }
        }
    }
    if (!waiting) weight_roulette(w, alive, rng, PR);
    A.x[s] = p.x; A.y[s] = p.y; A.z[s] = p.z;
    A.u[s] = u.x; A.v[s] = u.y; A.w[s] = u.z;
    A.wgt[s] = w;
    A.rng[s] = rng;
    A.nflight[s] = nfl;
    A.pcell[s] = cell;
    const int m1 = A.cell[cell].material;
    A.pmat[s] = m1;
    if (kGrad && cell != cell_old && m1 >= 0 && A.mat[m1].param_count > 0) {
        const uint32_t vm = A.vmat[s], bit = mat_bit(PR, m1);
        if (!(vm & bit)) A.vmat[s] = vm | bit;
    }
    if (waiting) return ST_WAIT;
    if (alive) return ST_ALIVE;
    return finish_history<kGrad, true>(A, s, p_id, PR, rng, p, E, nfl);
}

// A turn of a particle waiting at a surface (status ST_WAIT): it places the
// copies the ring has room for (a comb problem's: all of them, on its slot's
// bank) and, once all are placed, takes the roulette it stopped before
// (surface_crossing).  Returns ST_WAIT while copies are
// left, else ST_ALIVE (it flies on) or the status its end gives the slot.
template <bool kGrad>
__device__ __forceinline__ uint8_t resume_parked(const KArgs& A, uint32_t s)
{
    const uint32_t left = (uint32_t)A.ppend[s], first = (uint32_t)A.pfirst[s];
    const Vec3 p = {A.x[s], A.y[s], A.z[s]};
    const Vec3 u = {A.u[s], A.v[s], A.w[s]};
    const double E = A.E[s];
    double w = A.wgt[s];
    uint64_t rng = A.rng[s];
    const uint32_t nfl = A.nflight[s];
    const int cell = A.pcell[s], p_id = A.pprob[s];
    const DProb PR = A.prob[p_id];
    uint32_t put;
    if (!kGrad && A.comb && A.comb[2 * p_id + 1] > 0) {
        push_bank(A, s, first, left, p, u, E, w, rng, cell);     // a comb problem's own bank
        put = left;
    } else {
        put = push_copies<kGrad>(A, s, first, left, p, u, E, w, rng, nfl, cell, p_id);
    }
    if (put < left) {
        A.ppend[s] = (int32_t)(left - put);
        A.pfirst[s] = (int32_t)(first + put);
        atomicAdd((unsigned long long*)&A.ctr->waits, 1ull);
        return ST_WAIT;
    }
    A.ppend[s] = 0;
    bool alive = true;
    weight_roulette(w, alive, rng, PR);
    A.wgt[s] = w;
    A.rng[s] = rng;
    if (alive) return ST_ALIVE;
    return finish_history<kGrad, true>(A, s, p_id, PR, rng, p, E, nfl);
}

// First kernel of an iteration of a run that can split: waiting particles
// retry, and comb histories start their next copy, before the refill counts
// the slots.
template <bool kGrad>
__global__ void __launch_bounds__(1024) k_resume(KArgs A)
{
    uint32_t lo, hi;
    block_range(A.ctr->active_cap, lo, hi);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const uint8_t s0 = A.status[i];
        if (!kGrad && s0 == ST_POP) {
            comb_pop(A, i);
            continue;
        }
        if (s0 != ST_WAIT) continue;
        const uint8_t st = resume_parked<kGrad>(A, i);
        if (!kGrad && st == ST_POP) {
            // its roulette ended it and its bank holds copies: the next one
            // starts now, so that no slot is left between statuses when the
            // refill counts live particles
            comb_pop(A, i);
            continue;
        }
        if (st != ST_WAIT) A.status[i] = st;
    }
}
template __global__ void k_resume<false>(KArgs);
template __global__ void k_resume<true>(KArgs);

// One flight of slot s: distance to collision and to the nearest surface,
// the move and its track-length scores; a flight ending on a surface crosses
// it here.  Returns the new status (ST_COLL: a collision is due).
// kDefer: a flight ending on a surface stores the particle and the surface
// and returns ST_SURF; the crossing is left to cross_deferred.
// State of a particle held in registers across the events of one launch of
// the flight kernel: kReg routines read and update it instead of the bank,
// and the caller stores it once when the particle stops.
struct PState {
    Vec3 pos, dir;
    double E, w, sigt;
    double lp;                    // gradient runs: the last flight's length, until its collision folds it
    uint64_t rng;
    int cell, prob, gidx;
    uint32_t nfl;
    bool pend;                    // gradient runs: the slot's cs, cc or Qw may be non-zero
};

__device__ __forceinline__ PState load_state(const KArgs& A, uint32_t s, bool tables)
{
    PState P;
    P.cell = A.pcell[s];
    P.prob = A.pprob[s];
    P.pos = {A.x[s], A.y[s], A.z[s]};
    P.dir = {A.u[s], A.v[s], A.w[s]};
    P.E = A.E[s];
    P.w = A.wgt[s];
    P.rng = A.rng[s];
    P.nfl = A.nflight[s];
    P.sigt = tables ? A.sigt[s] : 0.0;
    P.lp = 0.0;                   // every launch's first event is a flight
    P.pend = true;
    P.gidx = tables ? A.gidx[s] : 0;
    return P;
}

// The fields an event can change (the cell changes only at a crossing, which
// stores its own).
__device__ __forceinline__ void store_state(const KArgs& A, uint32_t s, const PState& P, bool tables)
{
    A.x[s] = P.pos.x; A.y[s] = P.pos.y; A.z[s] = P.pos.z;
    A.u[s] = P.dir.x; A.v[s] = P.dir.y; A.w[s] = P.dir.z;
    A.E[s] = P.E;
    A.wgt[s] = P.w;
    A.rng[s] = P.rng;
    A.nflight[s] = P.nfl;
    if (tables) {
        A.sigt[s] = P.sigt;
        A.gidx[s] = P.gidx;
    }
}

// kGeo: where the geometry is read from (kGeoLds: the LDS copy *gl;
// kGeoRec: the cell's flight record; kGeoArrays: cell, cell_surf, surf).
template <bool kTables, bool kGrad, bool kDefer, bool kReg = false, int kGeo = kGeoArrays>
__device__ __forceinline__ uint8_t flight_one(const KArgs& A, uint32_t s, bool& crossed, PState* R = nullptr,
                                              const GeoLds* gl = nullptr)
{
    const int c = kReg ? R->cell : A.pcell[s];
    mc_cell_desc C{};
    const DCellSurf* csr = nullptr;
    if constexpr (kGeo == kGeoRec) {
        const char* rec = A.crec + (size_t)c * A.crec_stride;
        const DCellHead H = *(const DCellHead*)rec;
        C.material = H.material;
        C.surf_count = H.surf_count;
        C.tl_begin = H.tl_begin;
        C.tl_count = H.tl_count;
        csr = (const DCellSurf*)(rec + sizeof(DCellHead));
    } else {
        C = kGeo == kGeoLds ? gl->cell[c] : A.cell[c];
    }
    Vec3 p = kReg ? R->pos : Vec3{A.x[s], A.y[s], A.z[s]};
    const Vec3 u = kReg ? R->dir : Vec3{A.u[s], A.v[s], A.w[s]};
    // surface_intersect, general.cpp:54-67
    double dist = kMaxFloat;
    int S = -1;
    for (int k = 0; k < C.surf_count; ++k) {
        if constexpr (kGeo == kGeoRec) {
            const double d = surface_distance(csr[k].type, csr[k].c, p, u);
            if (d < dist) { dist = d; S = csr[k].sid; }
        } else {
            const int sf = kGeo == kGeoLds ? gl->cell_surf[C.surf_begin + k] : A.cell_surf[C.surf_begin + k];
            const mc_surface_desc& Sd = kGeo == kGeoLds ? gl->surf[sf] : A.surf[sf];
            const double d = surface_distance(Sd.type, Sd.c, p, u);
            if (d < dist) { dist = d; S = sf; }
        }
    }
    // collision_distance, general.cpp:40-48
    uint64_t rng = kReg ? R->rng : A.rng[s];
    const double E = kReg ? R->E : A.E[s];
    double dcol, sigt = 0.0;
    if (C.material >= 0) {
        // without tables every material is all-constant: its per-run Sigma_t
        sigt = kTables ? (kReg ? R->sigt : A.sigt[s]) : A.mat_tconst[C.material];
        dcol = -mc_log(urand(rng)) / sigt;
    } else {
        dcol = kMaxFloatLess;
    }
    const bool hit = dcol > dist;
    const double l = hit ? dist : dcol;
    const int p_id = kReg ? R->prob : A.pprob[s];
    const DProb PR = A.prob[p_id];
    // gradient runs: the flight's -l dSigma_t/dtheta is folded into S at the
    // collision or surface crossing that ends it, where the per-nuclide
    // cross sections at E are evaluated anyway
    // (state in registers: a flight that ends in a collision hands its length
    // to the collision in R; only one that ends on a surface stores it)
    if (kGrad && kReg) R->lp = l;
    if (kGrad && (!kReg || hit) && C.material >= 0 && A.mat[C.material].param_count > 0) A.lpend[s] = l;
    // move_particle, general.cpp:73-83
    p.x += u.x * l;
    p.y += u.y * l;
    p.z += u.z * l;
    const uint32_t nfl = (kReg ? R->nfl : A.nflight[s]) + 1;
    const double w = kReg ? R->w : A.wgt[s];
    for (int a = 0; a < C.tl_count; ++a)
        score_attachment<kGrad>(A, s, C.tl_begin + a, l, c, E, w, PR, l, sigt, !hit);
    if (kGrad && kReg && C.tl_count > 0) R->pend = true;
    uint8_t st;
    crossed = false;
    if (kReg) {
        R->pos = p;
        R->rng = rng;
        R->nfl = nfl;
    }
    if (!hit) {
        if (!kReg) {
            A.x[s] = p.x; A.y[s] = p.y; A.z[s] = p.z;
            A.rng[s] = rng;
            A.nflight[s] = nfl;
        }
        st = ST_COLL;
    } else if (S >= 0) {
        crossed = true;
        if (kDefer) {
            if (!kReg) {
                A.x[s] = p.x; A.y[s] = p.y; A.z[s] = p.z;
                A.rng[s] = rng;
                A.nflight[s] = nfl;
            }
            A.psurf[s] = S;
            st = ST_SURF;
        } else {
            st = surface_crossing<kGrad>(A, s, S, p, u, E, w, rng, nfl, c, p_id);
        }
    } else {                        // no surface ahead: lost particle
        atomicAdd(&A.ctr->err_lost, 1u);
        if (kReg) {
            R->w = 0.0;
        } else {
            A.x[s] = p.x; A.y[s] = p.y; A.z[s] = p.z;
            A.rng[s] = rng;
            A.nflight[s] = nfl;
            A.wgt[s] = 0.0;
        }
        st = finish_history<kGrad, false>(A, s, p_id, PR, rng, p, E, nfl);
    }
    if (!kReg) A.status[s] = st;
    return st;
}

// The surface crossing of a flight that flight_one<..., true> deferred.
template <bool kGrad>
__device__ __forceinline__ void cross_deferred(const KArgs& A, uint32_t s)
{
    const Vec3 p = {A.x[s], A.y[s], A.z[s]};
    const Vec3 u = {A.u[s], A.v[s], A.w[s]};
    A.status[s] = surface_crossing<kGrad>(A, s, A.psurf[s], p, u, A.E[s], A.wgt[s], A.rng[s], A.nflight[s],
                                          A.pcell[s], A.pprob[s]);
}

// Flights of queue A over the block's range.  Flights that end on a surface
// are queued in LDS (an LDS atomic gives each a place; the order does not
// matter, every crossing is independent) and crossed by the whole block
// once the queue fills or the range ends, so the rare crossing does not hold
// a wave's other lanes and its registers stay out of the flight loop.
struct PState;
template <bool kTables, bool kGrad, bool kInlineFreeGas = true, bool kReg = false, int kGeo = kGeoArrays>
__device__ __forceinline__ uint8_t collision_one(const KArgs& A, uint32_t s, PState* reg = nullptr,
                                                 double* sacc = nullptr, const GeoLds* gl = nullptr,
                                                 const DFlightLook* lk = nullptr, int nlk = 0);

// kGeo: kGeoLds (fused launches whose geometry fits GeoLds): flights and
// collisions read the cells and surfaces from the workgroup's LDS copy;
// kGeoRec: from the per-cell flight records.
template <bool kTables, bool kGrad, bool kFuse, int kGeo>
__global__ void __launch_bounds__(1024) k_flight(KArgs A)
{
    // kFuse: the flight's collision happens in the same pass (a kernel of its
    // own would re-read the state), and a particle that survives it flies
    // again, up to events_per_launch flights (continuous energy: after a
    // lookup of Sigma_t at its new energy), before the bank is refilled.
    // Every particle still takes its own events in order; only how many of
    // them one iteration holds changes.  Without kFuse (gradient runs at one
    // flight per iteration) collisions go to queue C for k_collision.
    uint32_t extra = 0;                 // flights beyond the first, this thread
    // fused gradient launches: kAccEntries doubles per thread of dynamic LDS,
    // cleared here (each thread its own column)
    extern __shared__ double lds_acc[];
    if constexpr (kGrad && kFuse) {
#pragma unroll
        for (int e = 0; e < kAccEntries; ++e) lds_acc[e * blockDim.x + threadIdx.x] = 0.0;
    }
    constexpr uint32_t kXq = 2048;      // >= the largest workgroup
    __shared__ uint32_t xq[kXq];
    __shared__ uint32_t xn, cn, sn;
    __shared__ unsigned long long xf;
    __shared__ GeoLds geo;
    if constexpr (kGeo == kGeoLds) {
        for (int q = threadIdx.x; q < A.n_cell; q += blockDim.x) geo.cell[q] = A.cell[q];
        for (int q = threadIdx.x; q < A.n_surf; q += blockDim.x) geo.surf[q] = A.surf[q];
        for (int q = threadIdx.x; q < A.n_cell_surf; q += blockDim.x) geo.cell_surf[q] = A.cell_surf[q];
    }
    // the lookup records of the first materials (the fused loop's lookups)
    __shared__ DFlightLook lk[kFlightLooks];
    const int nlk = kFuse && kTables ? (A.n_mat < kFlightLooks ? A.n_mat : kFlightLooks) : 0;
    for (int q = threadIdx.x; q < nlk; q += blockDim.x) {
        const DMatLook L = A.mlook[q];
        lk[q].L = L;
#pragma unroll
        for (int k = 0; k < kFastNuc; ++k) {
            const bool on = L.rec_off >= 0 && k < L.nuc_count;
            lk[q].dens[k] = on ? A.mat_dens[L.nuc_begin + k] : 0.0;
            lk[q].ns[k] = on ? A.mat_ns[L.nuc_begin + k] : 0.0;
            lk[q].nc[k] = on ? A.mat_nc[L.nuc_begin + k] : 0.0;
            lk[q].mass[k] = on ? A.nuc[A.mat_nuc[L.nuc_begin + k]].A : 0.0;
            lk[q].kin[k] = on ? A.nuc[A.mat_nuc[L.nuc_begin + k]].kin : 0;
        }
        lk[q].R = A.mrec[q];
    }
    const uint32_t n = A.ctr->nA;
    uint32_t lo, hi;
    block_range(n, lo, hi);
    // a range of at most kXq particles queues at most kXq crossings (one per
    // particle): one-speed launches then cross them all after the block's last
    // pass, and its waves need not meet after every pass to check the queue's
    // fill (continuous-energy launches measured faster crossing as they go)
    const bool one_flush = !kTables && hi - lo <= kXq;
    // continuous-energy forward launches: workgroups take blockDim-sized
    // chunks of queue A from a cursor instead of fixed ranges, so that none
    // holds the launch's end with a range of slow particles (a chunk's
    // successor is claimed at its start and read between the pass's barriers)
    constexpr bool kDyn = kTables && kFuse && !kGrad;
    __shared__ uint32_t first, nxt;     // (the first chunk apart: nxt is rewritten at once)
    if (threadIdx.x == 0) {
        xn = 0; cn = 0; sn = 0; xf = 0;
        if (kDyn) first = atomicAdd(&A.ctr->fcur, blockDim.x);
    }
    __syncthreads();
    if (kDyn) {
        lo = first;
        hi = n;
    }
    const uint32_t ident = A.ctr->qa_ident;
    uint32_t s_next = lo + threadIdx.x < hi ? qa_at(A.qA, ident, lo + threadIdx.x) : 0u;   // read one item ahead
    for (uint32_t base = lo; base < hi;) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t s = s_next;
        if (kDyn) {
            if (threadIdx.x == 0) nxt = atomicAdd(&A.ctr->fcur, blockDim.x);
        } else if (i + blockDim.x < hi) {
            s_next = qa_at(A.qA, ident, i + blockDim.x);
        }
        uint8_t st = ST_DEAD;
        if (i < hi) {
            bool crossed;
            uint32_t nfl = 0;
            if constexpr (kFuse) {
                // the state stays in registers from event to event
                PState R = load_state(A, s, kTables);
                if constexpr (kGrad) R.pend = A.Qw[s] != 0.0 || A.cs[s] != 0.0 || A.cc[s] != 0.0;
                const int m = kTables ? A.pmat[s] : -1;
                // gradient runs: this lane's S sums in LDS, when its material allows
                double* acc = nullptr;
                if constexpr (kGrad) {
                    if (acc_material(A, kTables ? m : A.cell[R.cell].material)) acc = lds_acc + threadIdx.x;
                }
                for (;;) {
                    // the slot's field addresses are recomputed where they are
                    // used (one add each) rather than held across the events
                    const uint32_t se = opaque(s);
                    st = flight_one<kTables, kGrad, true, true, kGeo>(A, se, crossed, &R, &geo);
                    ++nfl;
                    if (st != ST_COLL) break;
                    st = collision_one<kTables, kGrad, false, true, kGeo>(A, se, &R, acc, &geo, lk, nlk);
                    if (kGrad && st != ST_ALIVE) acc = nullptr;     // its end flushed the sums
                    if (st != ST_ALIVE || nfl >= A.events_per_launch) break;
                    // once fewer than a quarter of the wave's lanes are still
                    // flying, the rest stop too and fly on next iteration (the
                    // wave would otherwise idle most lanes; every particle's
                    // events keep their order either way)
                    if (__popcll(__ballot(1)) < 16) break;
                    if constexpr (kTables) {
                        // Sigma_t and its grid interval at the energy the collision left
                        int g = 0;
                        R.sigt = m >= 0 ? flight_sigt(A, lk, nlk, m, R.E, g) : 0.0;
                        R.gidx = g;
                    }
                }
                const uint32_t so = opaque(s);
                if (kGrad && acc) flush_acc(A, so, A.cell[R.cell].material, acc);
                store_state(A, so, R, kTables);
                A.status[so] = st;
            } else {
                // one flight; its collision goes to queue C
                st = flight_one<kTables, kGrad, true>(A, s, crossed);
                nfl = 1;
            }
            extra += nfl - 1;
        }
        // collisions go to this block's segment of queue C (slots of a wave in
        // lane order, one LDS atomic per wave); crossings to the LDS queue
        const uint32_t ci = wave_append(st == ST_COLL, &cn);
        if (!kFuse && st == ST_COLL) A.qC[lo + ci] = s;
        const uint32_t xi = wave_append(st == ST_SURF, &xn);
        if constexpr (kGrad) {
            // gradient runs: crossings go to this block's segment of queue S
            // and k_cross takes them (the crossing's folds and closeout would
            // otherwise raise this kernel's register demand)
            if (st == ST_SURF) A.qS[lo + xi] = s;
        } else {
            if (st == ST_SURF) xq[xi] = s;
            if (!one_flush) {
                __syncthreads();
                const uint32_t q = xn;      // the same in every thread between the barriers
                const uint32_t next = kDyn ? nxt : base + blockDim.x;
                __syncthreads();
                if (q + blockDim.x > kXq || next >= hi) {
                    for (uint32_t j = threadIdx.x; j < q; j += blockDim.x) cross_deferred<kGrad>(A, xq[j]);
                    __syncthreads();
                    if (threadIdx.x == 0) { sn += q; xn = 0; }
                    __syncthreads();
                }
                if (kDyn) {
                    base = next;
                    s_next = base + threadIdx.x < hi ? qa_at(A.qA, ident, base + threadIdx.x) : 0u;
                    continue;
                }
            }
        }
        base += blockDim.x;
    }
    if constexpr (!kGrad) {
        if (one_flush) {
            // every crossing of the range is queued: the block crosses them once
            __syncthreads();
            const uint32_t q = xn;
            for (uint32_t j = threadIdx.x; j < q; j += blockDim.x) cross_deferred<kGrad>(A, xq[j]);
            if (threadIdx.x == 0) sn += q;
        }
    }
    if constexpr (kGrad) {
        __syncthreads();
        if (threadIdx.x == 0) sn = xn;
        __syncthreads();
    }
    if constexpr (kFuse) {
        unsigned long long w = extra;
        for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o);
        if (lane_id() == 0 && w != 0) atomicAdd(&xf, w);
        __syncthreads();
        if (threadIdx.x == 0 && xf != 0) A.wg_xfl[blockIdx.x] += xf;   // this block's own entry
    }
    // the collision kernel's workgroup b takes this block's segment
    if (threadIdx.x == 0) {
        A.wg_cnt[2 * blockIdx.x] = sn;
        A.wg_cnt[2 * blockIdx.x + 1] = cn;
    }
}

// Gradient runs: the surface crossings k_flight queued, workgroup b taking
// flight workgroup b's segment of queue S.
template <bool kTables>
__global__ void __launch_bounds__(1024) k_cross(KArgs A)
{
    uint32_t lo, hi;
    block_range(A.ctr->nA, lo, hi);
    const uint32_t n = A.wg_cnt[2 * blockIdx.x];
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) cross_deferred<true>(A, A.qS[lo + i]);
}

// Event counts of an iteration (profiled runs only): the flight's per-workgroup
// surface-crossing and collision counts summed into the counters.
__global__ void k_count_events(const uint32_t* wg_cnt, uint32_t n_wg, Counters* c)
{
    __shared__ uint32_t lds[32];
    uint32_t ns = 0, nc = 0;
    for (uint32_t b = threadIdx.x; b < n_wg; b += blockDim.x) {
        ns += wg_cnt[2 * b];
        nc += wg_cnt[2 * b + 1];
    }
    block_sum2(ns, nc, lds);
    if (threadIdx.x == 0) {
        c->nS = ns;
        c->nC = nc;
    }
}

// One collision of slot s (general.cpp:121-163): collision tallies, implicit
// capture, nuclide selection, elastic scatter, weight roulette.  Returns the
// new status.
// Free-gas scatter of a collision's particle state: inline in the forward
// collision kernel, out of line where registers are tighter (the drain and
// the gradient collision kernels).
struct FreeGasState {
    double E;
    Vec3 u;
    uint64_t rng;
};

__device__ __forceinline__ FreeGasState free_gas_inline(double A, FreeGasState f)
{
    free_gas_scatter(A, f.E, f.u, f.rng);
    return f;
}

__device__ __noinline__ FreeGasState free_gas_outline(double A, FreeGasState f)
{
    free_gas_scatter(A, f.E, f.u, f.rng);
    return f;
}

// kReg: the particle's state is *reg, read and updated there; the returned
// status is not stored (the caller stores the state and the status).
// lk, nlk (fused launches): a material with a staged lookup record takes its
// descriptor, grid, records, densities and nuclide masses from LDS.
template <bool kTables, bool kGrad, bool kInlineFreeGas, bool kReg, int kGeo>
__device__ __forceinline__ uint8_t collision_one(const KArgs& A, uint32_t s, PState* reg, double* sacc,
                                                 const GeoLds* gl, const DFlightLook* lk, int nlk)
{
    const int c = kReg ? reg->cell : A.pcell[s];
    mc_cell_desc C{};
    if constexpr (kGeo == kGeoRec) {
        const DCellHead H = *(const DCellHead*)(A.crec + (size_t)c * A.crec_stride);
        C.material = H.material;
        C.col_begin = H.col_begin;
        C.col_count = H.col_count;
    } else {
        C = kGeo == kGeoLds ? gl->cell[c] : A.cell[c];
    }
    const int p_id = kReg ? reg->prob : A.pprob[s];
    const DProb PR = A.prob[p_id];
    double w = kReg ? reg->w : A.wgt[s];
    uint64_t rng = kReg ? reg->rng : A.rng[s];
    double E = kReg ? reg->E : A.E[s];
    bool alive = true;
    Vec3 u = kReg ? reg->dir : Vec3{A.u[s], A.v[s], A.w[s]};
    if (C.material < 0) {                    // void: kill at collision
        alive = false;
        w = 0.0;
    } else {
        if (C.col_count > 0) {
            const double lp = kGrad ? (kReg ? reg->lp : A.lpend[s]) : 0.0;
            for (int a = 0; a < C.col_count; ++a)
                score_attachment<kGrad>(A, s, C.col_begin + a, 0.0, c, E, w, PR, lp, 0.0, false);
            if (kGrad && kReg) reg->pend = true;
        }
        const int m = C.material;
        const DFlightLook* F = kTables && kReg && m < nlk && lk[m].L.rec_off >= 0 ? &lk[m] : nullptr;
        mc_material_desc M;
        if (F && !kGrad) M = {F->L.nuc_begin, F->L.nuc_count, 0, F->L.scaled};
        else M = A.mat[m];
        const bool scaled = M.param_count != 0;
        XsPoint P;
        int gm;
        if constexpr (kTables && kReg) {
            if (F) {
                gm = F->L.grid;
                xs_point_on(A, gm, F->L.e_off, F->L.n, reg->gidx, E, P);
            } else {
                gm = A.mat_grid[m];
                if (gm >= 0) xs_point_at(A, gm, reg->gidx, E, P);
            }
        } else {
            gm = kTables ? material_point_cached(A, m, s, E, P) : material_point(A, m, E, P);
        }
        // Sigma_s, Sigma_c (and Sigma_t when no lookup kernel ran) in one pass.
        // The first kNucReg entries keep in registers what later steps reuse:
        // the products sigma_s * N for the nuclide choice (the same terms
        // Sigma_s sums) or, in gradient runs, the microscopic sigma_s, from
        // which both the choice and dSigma_s/dtheta follow.
        constexpr int kNucReg = 8;
        double ps[kGrad ? 1 : kNucReg];
        double rs[kGrad ? kNucReg : 1];
        Xs3 X = {0.0, 0.0, 0.0};
        // sigma_s and sigma_c of every nuclide from the material's record (one
        // or two cache lines) when it has one, else from the nuclide tables
        const DMatRec R = F ? F->R : A.mrec[m];
        const double* rsc = kTables && R.sc_off >= 0 ? rec_at(A.rec_sc, R.sc_off, R.sc_stride, P) : nullptr;
        const int n2 = R.c_at;
#pragma unroll
        for (int k = 0; k < kNucReg; ++k) {
            if (!kGrad) ps[k] = 0.0;
            if (k < M.nuc_count) {
                const int j = M.nuc_begin + k;
                const bool f = F && k < kFastNuc;
                const double Nd = f ? F->dens[k] : A.mat_dens[j];
                const double xs_s = rsc ? xs_pair(rsc + 2 * k, P) : entry_xs(A, j, XS_S, E, P, gm);
                const double xs_c = rsc ? xs_pair(rsc + n2 + 2 * k, P) : entry_xs(A, j, XS_C, E, P, gm);
                const double prod = xs_s * (scaled ? (f ? F->ns[k] : A.mat_ns[j]) : Nd);
                if constexpr (kGrad) {
                    rs[k] = xs_s;
                } else {
                    ps[k] = prod;
                }
                X.s += prod;
                X.c += xs_c * (scaled ? (f ? F->nc[k] : A.mat_nc[j]) : Nd);
                if (!kTables && !scaled) X.t += entry_xs(A, j, XS_T, E, P, gm) * Nd;
            }
        }
        for (int k = kNucReg; k < M.nuc_count; ++k) {
            const int j = M.nuc_begin + k;
            const double Nd = A.mat_dens[j];
            const double xs_s = rsc ? xs_pair(rsc + 2 * k, P) : entry_xs(A, j, XS_S, E, P, gm);
            const double xs_c = rsc ? xs_pair(rsc + n2 + 2 * k, P) : entry_xs(A, j, XS_C, E, P, gm);
            X.s += xs_s * (scaled ? A.mat_ns[j] : Nd);
            X.c += xs_c * (scaled ? A.mat_nc[j] : Nd);
            if (!kTables && !scaled) X.t += entry_xs(A, j, XS_T, E, P, gm) * Nd;
        }
        // Sigma_t: the lookup kernel's value for this particle is the same sum
        if (scaled) X.t = X.s + X.c;
        else if (kTables) X.t = kReg ? reg->sigt : A.sigt[s];
        // implicit capture, general.cpp:154-156 (Sigma_f = 0)
        const double implicit = X.c + 0.0;
        w = w * (X.t - implicit) / X.t;
        // Material::nuclide_scatter, Material.cpp:106-115
        const double ur = X.s * urand(rng);
        double acc = 0.0;
        int sel = -1;
#pragma unroll
        for (int k = 0; k < kNucReg; ++k) {
            if (sel < 0 && k < M.nuc_count) {
                if constexpr (kGrad) {
                    const int j = M.nuc_begin + k;
                    acc += rs[k] * (scaled ? A.mat_ns[j] : A.mat_dens[j]);
                } else {
                    acc += ps[k];
                }
                if (acc > ur) sel = k;
            }
        }
        for (int k = kNucReg; sel < 0 && k < M.nuc_count; ++k) {
            const int j = M.nuc_begin + k;
            const double xs_s = rsc ? xs_pair(rsc + 2 * k, P) : entry_xs(A, j, XS_S, E, P, gm);
            acc += xs_s * (scaled ? A.mat_ns[j] : A.mat_dens[j]);
            if (acc > ur) sel = k;
        }
        if constexpr (kGrad) {
            // Fold the flight that ended here (its 1/Sigma_t collision density
            // cancels the implicit-capture weight factor's -dSigma_t/Sigma_t)
            // and the nuclide selection, [scatter off sel]/theta, into S and H.
            // dSigma_s/dtheta = sigma_s N and dSigma_c/dtheta = sigma_c N of the
            // parameter's entry: sigma_s from the registers above, sigma_c
            // read again (the record or table lines are in cache).
            if (M.param_count > 0 && A.mfold[m]) {
                // per nuclide: its scatter and capture parameters, sigma_s from registers
                // (a particle that has not scored has nothing pending)
                LrPending L = kReg && !reg->pend ? LrPending{0.0, 0.0, 0.0, 0.0} : lr_pending_cs(A, s);
                L.lp = kReg ? reg->lp : A.lpend[s];
#pragma unroll
                for (int k = 0; k < kNucReg; ++k) {
                    if (k < M.nuc_count) {
                        const int j = M.nuc_begin + k;
                        const DNucFold F = A.nfold[j];
                        const double N = A.mat_dens[j];
                        const double xtra = sel == k ? F.inv_theta_s : 0.0;
                        if (F.kl_s >= 0) {
                            if (kReg && sacc)   // (the material has at most kAccNuc nuclides)
                                fold_acc(A, s, F.kl_s, rs[k] * N, 0.0, xtra, L, sacc + (size_t)(2 * k) * blockDim.x);
                            else
                                fold_one(A, s, F.kl_s, rs[k] * N, 0.0, xtra, L);
                        }
                        if (F.kl_c >= 0) {
                            const double xs_c = rsc ? xs_pair(rsc + n2 + 2 * k, P) : entry_xs(A, j, XS_C, E, P, gm);
                            if (kReg && sacc)
                                fold_acc(A, s, F.kl_c, 0.0, xs_c * N, 0.0, L, sacc + (size_t)(2 * k + 1) * blockDim.x);
                            else
                                fold_one(A, s, F.kl_c, 0.0, xs_c * N, 0.0, L);
                        }
                    }
                }
                if (L.cs != 0.0 || L.cc != 0.0) {
                    A.cs[s] = 0.0;
                    A.cc[s] = 0.0;
                }
            } else if (M.param_count > 0) {
                // (a particle that has not scored has nothing pending)
                LrPending L = kReg && !reg->pend ? LrPending{0.0, 0.0, 0.0, 0.0} : lr_pending_cs(A, s);
                L.lp = kReg ? reg->lp : A.lpend[s];
                for (int q = 0; q < M.param_count; ++q) {
                    const DParFold F = A.mpar[M.param_begin + q];
                    const int slot = F.slot, j = M.nuc_begin + slot;
                    const double N = F.N;
                    double ds = 0.0, dc = 0.0;
                    if (F.reactions & MC_REACTION_SCATTER) {
                        double xs_s = 0.0;
                        if (slot < kNucReg) {
#pragma unroll
                            for (int t = 0; t < kNucReg; ++t)
                                if (t == slot) xs_s = rs[t];
                        } else {
                            xs_s = rsc ? xs_pair(rsc + 2 * slot, P) : entry_xs(A, j, XS_S, E, P, gm);
                        }
                        ds = xs_s * N;
                    }
                    if (F.reactions & MC_REACTION_CAPTURE)
                        dc = (rsc ? xs_pair(rsc + n2 + 2 * slot, P) : entry_xs(A, j, XS_C, E, P, gm)) * N;
                    const double extra =
                        (sel >= 0 && slot == sel && (F.reactions & MC_REACTION_SCATTER)) ? F.inv_theta : 0.0;
                    if (kReg && sacc)   // (the material has at most kAccEntries parameters)
                        fold_acc(A, s, F.kl, ds, dc, extra, L, sacc + (size_t)q * blockDim.x);
                    else
                        fold_one(A, s, F.kl, ds, dc, extra, L);
                }
                if (L.cs != 0.0 || L.cc != 0.0) {
                    A.cs[s] = 0.0;
                    A.cc[s] = 0.0;
                }
            }
        }
        const int ss = sel >= 0 ? sel : 0;
        double n_mass;
        int n_kin;
        if (F) {
            n_mass = F->mass[ss];
            n_kin = F->kin[ss];
        } else {
            const DNuc N = A.nuc[A.mat_nuc[M.nuc_begin + ss]];
            n_mass = N.A;
            n_kin = N.kin;
        }
        if (sel >= 0 && n_kin == MC_KIN_FREE_GAS && n_mass < MC_MASS_INFINITE) {
            const FreeGasState f = kInlineFreeGas ? free_gas_inline(n_mass, FreeGasState{E, u, rng})
                                                  : free_gas_outline(n_mass, FreeGasState{E, u, rng});
            E = f.E;
            u = f.u;
            rng = f.rng;
        } else if (sel >= 0) {
            // elastic scatter, target at rest (T -> 0 limit of Reaction.cpp:27-118)
            const double mu_cm = 2.0 * urand(rng) - 1.0;
            double mu = mu_cm;
            if (n_mass < MC_MASS_INFINITE) {
                const double Am = n_mass;
                const double q = Am * Am + 2.0 * Am * mu_cm + 1.0;
                E = E * q / ((Am + 1.0) * (Am + 1.0));
                mu = (1.0 + Am * mu_cm) / sqrt(q);
            }
            double ca, sa;
            sincos_2pi(urand(rng), ca, sa);
            u = scatter_direction(u, mu, ca, sa);
        }
    }
    weight_roulette(w, alive, rng, PR);
    if constexpr (kReg) {
        reg->dir = u;
        reg->E = E;
        reg->w = w;
        reg->rng = rng;
        if (alive) return ST_ALIVE;
        if (kGrad && sacc) flush_acc(A, s, C.material, sacc);    // the closeout reads S
        return finish_history<kGrad, false>(A, s, p_id, PR, rng, reg->pos, E, reg->nfl);
    } else {
        A.u[s] = u.x; A.v[s] = u.y; A.w[s] = u.z;
        A.E[s] = E;
        A.wgt[s] = w;
        A.rng[s] = rng;
        if (alive) {
            A.status[s] = ST_ALIVE;
            return ST_ALIVE;
        } else {
            const Vec3 p = {A.x[s], A.y[s], A.z[s]};
            A.status[s] = finish_history<kGrad, false>(A, s, p_id, PR, rng, p, E, A.nflight[s]);
        }
        return ST_DEAD;
    }
}

template <bool kTables, bool kGrad>
__global__ void __launch_bounds__(1024) k_collision(KArgs A)
{
    // workgroup b takes the collisions flight workgroup b queued (the same
    // partition of queue A); the next entry is read one item ahead
    uint32_t lo, hi;
    block_range(A.ctr->nA, lo, hi);
    const uint32_t n = A.wg_cnt[2 * blockIdx.x + 1];
    const uint32_t* q = A.qC + lo;
    uint32_t i = threadIdx.x;
    uint32_t s_next = i < n ? q[i] : 0u;
    for (; i < n; i += blockDim.x) {
        const uint32_t s = s_next;
        if (i + blockDim.x < n) s_next = q[i + blockDim.x];
        collision_one<kTables, kGrad, !kGrad>(A, s);
    }
}

// Drain of the last particles of a run.  Once no history is left to start and
// the bank has thinned out, an event iteration costs its launch latencies for
// few particles; here each thread instead carries its particles from event to
// event (lookup, flight, surface crossing or collision) through the same
// per-particle routines, so every history takes exactly the same steps.  At
// most max_events flights per particle per launch; a particle still alive
// after that stays in the bank for the next iteration.
template <bool kTables, bool kGrad>
__global__ void __launch_bounds__(1024) k_drain(KArgs A, uint32_t max_events)
{
    __shared__ uint32_t lds[32];
    const uint32_t n = A.ctr->nA;
    uint32_t nfl = 0, nsurf = 0, ncoll = 0;
    MC_GRID_LOOP(n)
    {
        const uint32_t s0 = qa_at(A.qA, A.ctr->qa_ident, _i);
        for (uint32_t e = 0; e < max_events; ++e) {
            const uint32_t s = opaque(s0);     // field addresses recomputed per event
            if (kTables) {
                const int m = A.pmat[s];
                if (m >= 0) {
                    int g = 0;
                    A.sigt[s] = material_sigt(A, m, A.E[s], &g);
                    A.gidx[s] = g;
                }
            }
            bool crossed;
            uint8_t st = flight_one<kTables, kGrad, false>(A, s, crossed);
            ++nfl;
            nsurf += crossed;
            if (st == ST_COLL) {
                ++ncoll;
                st = collision_one<kTables, kGrad, false>(A, s);
            }
            if (st != ST_ALIVE) break;
        }
    }
    // event counts: nfl and nsurf summed per block, then one atomic each
    block_sum2(nfl, nsurf, lds);
    uint32_t zero = 0;
    block_sum2(ncoll, zero, lds);
    if (threadIdx.x == 0) {
        if (nfl) atomicAdd((unsigned long long*)&A.ctr->tA, (unsigned long long)nfl);
        if (nsurf) atomicAdd((unsigned long long*)&A.ctr->tS, (unsigned long long)nsurf);
        if (ncoll) atomicAdd((unsigned long long*)&A.ctr->tC, (unsigned long long)ncoll);
    }
}

// After a drain: its events are counted, so the next roll-over adds none.
__global__ void k_drain_close(Counters* c)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        c->nA = 0;
        c->nS = 0;
        c->nC = 0;
    }
}

// Standalone lookup over arrays (mc_xs_lookup and its timing entry point).
// kU independent lookups per thread per step; the input and output streams
// bypass the caches' retention (non-temporal) so the tables stay in L2.
template <int kU>
__global__ void __launch_bounds__(1024) k_xs_lookup_array(KArgs A, const double* __restrict__ E,
                                                         const int32_t* __restrict__ mat,
                                                         double* __restrict__ out, uint64_t n)
{
    const uint16_t* sb = stage_buckets(A, mc_lds_buckets);
    // problem 0's materials (the lookup's) follow the bucket table in LDS
    DLdsMat* lm = nullptr;
    int nm = 0;
    if (sb) {
        lm = (DLdsMat*)((char*)mc_lds_buckets + ((size_t)(A.lds_nb + 1) * 2 + 15) / 16 * 16);
        nm = A.prob[0].mat_count < kLdsMats ? A.prob[0].mat_count : kLdsMats;
        for (int q = threadIdx.x; q < nm; q += blockDim.x) {
            const DMatLook L = A.mlook[q];
            DLdsMat R;
            R.key_base = L.key_base;
            R.te_off = L.te_off >= 0 && L.grid == A.lds_grid ? L.te_off : -1;
            R.te_stride = L.te_stride;
            R.n = L.n;
            R.nb = L.nb;
            R.nuc_count = L.nuc_count;
#pragma unroll
            for (int k = 0; k < kFastNuc; ++k) R.dens[k] = k < L.nuc_count ? A.mat_dens[L.nuc_begin + k] : 0.0;
            lm[q] = R;
        }
        __syncthreads();
    }
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += kU * stride) {
        int m[kU];
        double e[kU], r[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint64_t iu = i + u * stride;
            m[u] = iu < n ? __builtin_nontemporal_load(mat + iu) : -1;
            e[u] = iu < n ? __builtin_nontemporal_load(E + iu) : 1.0;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            int g = 0;
            if (m[u] < 0) {
                r[u] = 0.0;
            } else if (m[u] < nm && lm[m[u]].te_off >= 0) {
                const DLdsMat& L = lm[m[u]];
                r[u] = sigt_lds_core(A, L.key_base, L.te_off, L.te_stride, L.n, L.nb, L.nuc_count, L.dens, sb, e[u], g);
            } else {
                r[u] = lookup_sigt(A, sb, m[u], e[u], g);
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint64_t iu = i + u * stride;
            if (iu < n) __builtin_nontemporal_store(r[u], out + iu);
        }
    }
}
template __global__ void k_xs_lookup_array<4>(KArgs, const double*, const int32_t*, double*, uint64_t);

// The random-gather ceiling of the standalone lookup: the same input and
// output streams (E 8 B, material 4 B in; 8 B out, non-temporal) and, per
// lookup, kDep dependent reads of a 128-byte line of a table of nlines lines
// at a position hashed from E (the lookup's record reads, with no
// arithmetic and no search).  Its rate bounds what any lookup reading that
// many lines of a table that size can reach.
template <int kDep>
__global__ void __launch_bounds__(1024) k_gather_lines(const double* __restrict__ E, const int32_t* __restrict__ mat,
                                                       double* __restrict__ out, uint64_t n,
                                                       const double* __restrict__ table, uint64_t nlines)
{
    constexpr int kU = 4;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += kU * stride) {
        uint64_t h[kU];
        int32_t m[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint64_t iu = i + u * stride;
            const double e = iu < n ? __builtin_nontemporal_load(E + iu) : 1.0;
            m[u] = iu < n ? __builtin_nontemporal_load(mat + iu) : 0;
            h[u] = double_to_bits(e) * 0x9E3779B97F4A7C15ull;
        }
        double r[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            uint64_t line = (h[u] >> 20) % nlines;
            double v = 0.0;
#pragma unroll
            for (int d = 0; d < kDep; ++d) {
                const double* l = table + 16 * line;
                const double a = l[0], b = l[15];
                v += a + b;
                line = (line + (uint64_t)double_to_bits(a) + 1) % nlines;   // the next read depends on this one
            }
            r[u] = v * (double)m[u];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint64_t iu = i + u * stride;
            if (iu < n) __builtin_nontemporal_store(r[u], out + iu);
        }
    }
}
template __global__ void k_gather_lines<1>(const double*, const int32_t*, double*, uint64_t, const double*, uint64_t);
template __global__ void k_gather_lines<2>(const double*, const int32_t*, double*, uint64_t, const double*, uint64_t);

// Synthetic lookup inputs: E log-uniform in [emin, emax], material uniform.
__global__ void k_xs_inputs(double* E, int32_t* mat, uint64_t n, double lnmin, double lnspan,
                            const int32_t* mats, int32_t nmat, uint64_t seed)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t r = (seed + 2 * i + 1) & kRnMask;
        urand(r);
        const double a = urand(r), b = urand(r);
        E[i] = exp(lnmin + a * lnspan);
        int k = (int)(b * (double)nmat);
        mat[i] = mats[k < nmat ? k : nmat - 1];
    }
}

// out[b] = sum over rows w of part[w * n + b] (rows in order)
// One block per bin: threads take rows t, t + blockDim, ...; then a fixed
// tree in LDS (deterministic for a given launch shape).  blockDim <= 256.
__global__ void k_reduce_rows(const double* part, uint64_t rows, uint64_t n, double* out)
{
    __shared__ double red[256];
    for (uint64_t b = blockIdx.x; b < n; b += gridDim.x) {
        double s = 0.0;
        for (uint64_t w = threadIdx.x; w < rows; w += blockDim.x) s += part[w * n + b];
        red[threadIdx.x] = s;
        __syncthreads();
        for (unsigned o = blockDim.x / 2; o > 0; o >>= 1) {
            if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[b] = red[0];
        __syncthreads();
    }
}

// Estimator::end_cycle / end_simulation (Estimator.cpp:347-367) per bin or
// parameter with N histories: mean = sum / N, uncertainty = sqrt((sq / N -
// mean^2) / (N - 1)); the sums themselves where osum / osq are given.
__global__ void k_moments(const double* sum, const double* sq, const double* N, uint64_t n, double* osum,
                          double* osq, double* omean, double* ouncer)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double s = sum[i], q = sq[i], Nh = N[i];
        if (osum) osum[i] = s;
        if (osq) osq[i] = q;
        const double mean = s / Nh;
        const double u2 = (q / Nh - mean * mean) / (Nh - 1.0);
        if (omean) omean[i] = mean;
        if (ouncer) ouncer[i] = sqrt(u2);
    }
}

__global__ void k_fill_f64(double* p, uint64_t n, double v)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_u8(uint8_t* p, uint64_t n, uint8_t v)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_i32(int32_t* p, uint64_t n, int32_t v)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Device evaluation of the shared math routines, for CPU/GPU equality tests:
// out[4i..4i+3] = log(x), cos(2 pi x), sin(2 pi x), sqrt(x) for x = in[i].
__global__ void k_selftest_math(const double* in, double* out, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = in[i];
    double c, s;
    sincos_2pi(x, c, s);
    out[4 * i] = mc_log(x);
    out[4 * i + 1] = c;
    out[4 * i + 2] = s;
    out[4 * i + 3] = sqrt(x);
}

// explicit instantiations used by the host
template __global__ void k_refill_scatter<false>(KArgs);
template __global__ void k_refill_scatter<true>(KArgs);
template __global__ void k_drain<false, false>(KArgs, uint32_t);
template __global__ void k_drain<true, false>(KArgs, uint32_t);
template __global__ void k_drain<false, true>(KArgs, uint32_t);
template __global__ void k_drain<true, true>(KArgs, uint32_t);
template __global__ void k_collision<false, true>(KArgs);
template __global__ void k_collision<true, true>(KArgs);
template __global__ void k_flight<false, false, true, kGeoArrays>(KArgs);
template __global__ void k_flight<false, true, true, kGeoArrays>(KArgs);
template __global__ void k_flight<false, false, true, kGeoLds>(KArgs);
template __global__ void k_flight<false, true, true, kGeoLds>(KArgs);
template __global__ void k_flight<false, false, true, kGeoRec>(KArgs);
template __global__ void k_flight<false, true, false, kGeoArrays>(KArgs);
template __global__ void k_flight<true, false, true, kGeoArrays>(KArgs);
template __global__ void k_flight<true, true, true, kGeoArrays>(KArgs);
template __global__ void k_flight<true, false, true, kGeoLds>(KArgs);
template __global__ void k_flight<true, true, true, kGeoLds>(KArgs);
template __global__ void k_flight<true, false, true, kGeoRec>(KArgs);
template __global__ void k_flight<true, true, false, kGeoArrays>(KArgs);
template __global__ void k_cross<false>(KArgs);
template __global__ void k_cross<true>(KArgs);

}  // namespace mc
