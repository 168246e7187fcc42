// mc_kernels.hip — event-based transport kernels for gfx950.
//
// One event iteration over the structure-of-arrays particle bank:
//   k_reset      zero the queue counters (graph-safe: a kernel, not a memset)
//   k_refill     compact live slots into queue A (wave ballot + prefix) and
//                start new histories in free slots (Simulator::start's
//                sample loop, handler.cpp:19-36)
//   k_xs_lookup  Sigma_t(E) for every particle in A (Material::SigmaT,
//                Material.cpp:50-57) — continuous-energy problems only
//   k_flight     distance to collision vs nearest surface, move, track-length
//                tallies; appends to the surface or collision queue
//                (random_walk + move_particle, general.cpp:40-83, 177-211)
//   k_surface    surface crossing: boundary conditions, cell search, surface
//                tallies, importance, weight roulette (general.cpp:89-115,
//                population_control.cpp:9-49)
//   k_collision  collision tallies, implicit capture, nuclide selection,
//                elastic scatter, weight roulette (general.cpp:121-163)
// A history that ends is closed out where it ends: its per-slot tally row is
// added to the FP64 moment arrays (Estimator::end_history, Estimator.cpp:339-346).
#include "mc_device.hpp"

namespace mc {

#define MC_WAVE_LOOP(count)                                                              \
    const uint32_t _lane = lane_id();                                                    \
    const uint32_t _wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;                 \
    const uint32_t _nwaves = (gridDim.x * blockDim.x) >> 6;                              \
    for (uint32_t _base = _wave * 64u; _base < (count); _base += _nwaves * 64u)

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ int find_problem(const KArgs& A, uint64_t hid)
{
    int lo = 0, hi = A.n_prob - 1;   // largest p with run_hb[p] <= hid
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

// Estimator::score (Estimator.cpp:298-336) for one attachment.
template <bool kGrad>
__device__ void score_attachment(const KArgs& A, int att, double l, int cellid, double E, double w,
                                 const DProb& PR, double* hrow, const double* Srow, double* Grow)
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
        hrow[bin - PR.bin_begin] += v;
        if (kGrad) {
            const double g = A.gout[bin];
            if (g != 0.0 && v != 0.0) {
                const double gv = g * v;
                for (int k = 0; k < PR.param_count; ++k) Grow[k] += gv * Srow[k];
                if (m >= 0) {
                    const mc_material_desc M = A.mat[m];
                    XsPoint P;
                    for (int q = 0; q < M.param_count; ++q) {
                        const DSig d = param_dsig(A, M, q, E, P);
                        double dl = 0.0;
                        if (sc == MC_SCORE_SCATTER) dl += d.ds / sx;
                        else if (sc == MC_SCORE_ABSORPTION || sc == MC_SCORE_CAPTURE) dl += d.dc / sx;
                        else if (sc == MC_SCORE_TOTAL) dl += (d.ds + d.dc) / sx;
                        if (kern == MC_KERNEL_COLLISION) dl -= (d.ds + d.dc) / X.t;
                        Grow[d.k - PR.param_begin] += gv * dl;
                    }
                }
            }
        }
    }
}

// End of a history: close out its tally row (Estimator::end_history).
template <bool kGrad>
__device__ void finish_history(const KArgs& A, uint32_t s, int p, const DProb& PR, uint64_t rng, Vec3 pos,
                               double E, uint32_t nfl)
{
    double* row = A.hist + (size_t)s * (size_t)A.bin_stride;
    double tot = 0.0;
    for (int b = 0; b < PR.bin_count; ++b) {
        const double h = row[b];
        tot += h;
        if (h != 0.0) {
            atomicAdd(&A.sum[PR.bin_begin + b], h);
            atomicAdd(&A.sq[PR.bin_begin + b], h * h);
            row[b] = 0.0;
        }
    }
    atomicAdd((unsigned long long*)&A.ctr->flights, (unsigned long long)nfl);
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
    if (kGrad) {
        double* G = A.G + (size_t)s * (size_t)A.par_stride;
        for (int k = 0; k < PR.param_count; ++k) {
            const double g = G[k];
            if (g != 0.0) {
                atomicAdd(&A.gsum[PR.param_begin + k], g);
                atomicAdd(&A.gsq[PR.param_begin + k], g * g);
                G[k] = 0.0;
            }
        }
    }
}

// ---------------------------------------------------------------- kernels

__global__ void k_reset(Counters* c)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) { c->nA = 0; c->nS = 0; c->nC = 0; }
}

template <bool kGrad>
__global__ void __launch_bounds__(1024) k_refill(KArgs A)
{
    const uint32_t cap = (uint32_t)A.cap;
    const uint64_t end_h = A.ctr->end_hist;
    MC_WAVE_LOOP(cap)
    {
        const uint32_t i = _base + _lane;
        const bool valid = i < cap;
        const uint8_t st = valid ? A.status[i] : ST_ALIVE;
        const bool keep = valid && st == ST_ALIVE;
        const bool need = valid && st != ST_ALIVE;
        // claim history keys for free slots: one atomic per wave
        const uint64_t mask = __ballot(need);
        bool got = false;
        uint64_t hid = 0;
        if (mask != 0 && __atomic_load_n(&A.ctr->next_hist, __ATOMIC_RELAXED) < end_h) {
            const int leader = __ffsll((unsigned long long)mask) - 1;
            unsigned long long base = 0;
            if ((int)_lane == leader)
                base = atomicAdd((unsigned long long*)&A.ctr->next_hist, (unsigned long long)__popcll(mask));
            base = __shfl(base, leader);
            hid = base + (uint64_t)__popcll(mask & ((1ull << _lane) - 1ull));
            got = need && hid < end_h;
        }
        if (got) {
            // SourceBank::get_source + SourcePoint::get_source (Source.cpp:20-45)
            const int p = find_problem(A, hid);
            const DProb PR = A.prob[p];
            uint64_t rng = history_seed(PR.seed, hid, A.jump);
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
            A.pprob[i] = p;
            A.nflight[i] = 0;
            A.status[i] = ST_ALIVE;
            if (kGrad) {
                double* Srow = A.S + (size_t)i * (size_t)A.par_stride;
                for (int k = 0; k < PR.param_count; ++k) Srow[k] = 0.0;
            }
        } else if (need && st != ST_EMPTY) {
            A.status[i] = ST_EMPTY;
        }
        const bool active = keep || got;
        const uint32_t pos = wave_append(active, &A.ctr->nA);
        if (active) A.qA[pos] = i;
    }
}

__global__ void __launch_bounds__(1024) k_xs_lookup(KArgs A)
{
    const uint32_t n = A.ctr->nA;
    MC_WAVE_LOOP(n)
    {
        const uint32_t i = _base + _lane;
        if (i < n) {
            const uint32_t s = A.qA[i];
            const int m = A.cell[A.pcell[s]].material;
            A.sigt[s] = m >= 0 ? material_sigt(A, m, A.E[s]) : 0.0;
        }
    }
}

template <bool kTables, bool kGrad>
__global__ void __launch_bounds__(1024) k_flight(KArgs A)
{
    const uint32_t n = A.ctr->nA;
    MC_WAVE_LOOP(n)
    {
        const uint32_t i = _base + _lane;
        const bool valid = i < n;
        bool hit = false;
        uint32_t s = 0;
        if (valid) {
            s = A.qA[i];
            const int c = A.pcell[s];
            const mc_cell_desc C = A.cell[c];
            Vec3 p = {A.x[s], A.y[s], A.z[s]};
            const Vec3 u = {A.u[s], A.v[s], A.w[s]};
            // surface_intersect, general.cpp:54-67
            double dist = kMaxFloat;
            int S = -1;
            for (int k = 0; k < C.surf_count; ++k) {
                const int sf = A.cell_surf[C.surf_begin + k];
                const mc_surface_desc& Sd = A.surf[sf];
                const double d = surface_distance(Sd.type, Sd.c, p, u);
                if (d < dist) { dist = d; S = sf; }
            }
            // collision_distance, general.cpp:40-48
            uint64_t rng = A.rng[s];
            const double E = A.E[s];
            double dcol, sigt = 0.0;
            if (C.material >= 0) {
                sigt = kTables ? A.sigt[s] : material_sigt(A, C.material, E);
                dcol = -mc_log(urand(rng)) / sigt;
            } else {
                dcol = kMaxFloatLess;
            }
            hit = dcol > dist;
            const double l = hit ? dist : dcol;
            const int p_id = A.pprob[s];
            const DProb PR = A.prob[p_id];
            if (kGrad && C.material >= 0) {
                const mc_material_desc M = A.mat[C.material];
                double* Srow = A.S + (size_t)s * (size_t)A.par_stride;
                XsPoint P;
                for (int q = 0; q < M.param_count; ++q) {
                    const DSig d = param_dsig(A, M, q, E, P);
                    const double dt = d.ds + d.dc;
                    Srow[d.k - PR.param_begin] += -l * dt + (hit ? 0.0 : dt / sigt);
                }
            }
            // move_particle, general.cpp:73-83
            p.x += u.x * l;
            p.y += u.y * l;
            p.z += u.z * l;
            const uint32_t nfl = A.nflight[s] + 1;
            const double w = A.wgt[s];
            if (C.tl_count > 0) {
                double* hrow = A.hist + (size_t)s * (size_t)A.bin_stride;
                const double* Srow = kGrad ? A.S + (size_t)s * (size_t)A.par_stride : nullptr;
                double* Grow = kGrad ? A.G + (size_t)s * (size_t)A.par_stride : nullptr;
                for (int a = 0; a < C.tl_count; ++a)
                    score_attachment<kGrad>(A, C.tl_begin + a, l, c, E, w, PR, hrow, Srow, Grow);
            }
            A.x[s] = p.x; A.y[s] = p.y; A.z[s] = p.z;
            A.rng[s] = rng;
            A.nflight[s] = nfl;
            A.psurf[s] = S;
            if (hit && S < 0) {            // no surface ahead: lost particle
                atomicAdd(&A.ctr->err_lost, 1u);
                A.wgt[s] = 0.0;
                finish_history<kGrad>(A, s, p_id, PR, rng, p, E, nfl);
                A.status[s] = ST_DEAD;
                hit = false;
                s = 0xffffffffu;
            }
        }
        const bool to_surf = valid && hit;
        const bool to_coll = valid && !hit && s != 0xffffffffu;
        const uint32_t ps = wave_append(to_surf, &A.ctr->nS);
        if (to_surf) A.qS[ps] = s;
        const uint32_t pc = wave_append(to_coll, &A.ctr->nC);
        if (to_coll) A.qC[pc] = s;
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

template <bool kGrad>
__global__ void __launch_bounds__(1024) k_surface(KArgs A)
{
    const uint32_t n = A.ctr->nS;
    MC_WAVE_LOOP(n)
    {
        const uint32_t i = _base + _lane;
        if (i >= n) continue;
        const uint32_t s = A.qS[i];
        const int sf = A.psurf[s];
        const mc_surface_desc& Sd = A.surf[sf];
        const int p_id = A.pprob[s];
        const DProb PR = A.prob[p_id];
        Vec3 p = {A.x[s], A.y[s], A.z[s]};
        Vec3 u = {A.u[s], A.v[s], A.w[s]};
        double w = A.wgt[s];
        uint64_t rng = A.rng[s];
        const double E = A.E[s];
        const int cell_old = A.pcell[s];
        int cell = cell_old;
        bool alive = true;
        // surface_hit, general.cpp:89-115
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
        if (Sd.est_count > 0) {
            double* hrow = A.hist + (size_t)s * (size_t)A.bin_stride;
            const double* Srow = kGrad ? A.S + (size_t)s * (size_t)A.par_stride : nullptr;
            double* Grow = kGrad ? A.G + (size_t)s * (size_t)A.par_stride : nullptr;
            for (int a = 0; a < Sd.est_count; ++a)
                score_attachment<kGrad>(A, Sd.est_begin + a, 0.0, cell, E, w, PR, hrow, Srow, Grow);
        }
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
                if (nsplit > 1) atomicAdd(&A.ctr->err_split, 1u);   // splitting: unsupported here
            }
        }
        weight_roulette(w, alive, rng, PR);
        A.x[s] = p.x; A.y[s] = p.y; A.z[s] = p.z;
        A.u[s] = u.x; A.v[s] = u.y; A.w[s] = u.z;
        A.wgt[s] = w;
        A.rng[s] = rng;
        A.pcell[s] = cell;
        if (alive) {
            A.status[s] = ST_ALIVE;
        } else {
            finish_history<kGrad>(A, s, p_id, PR, rng, p, E, A.nflight[s]);
            A.status[s] = ST_DEAD;
        }
    }
}

template <bool kGrad>
__global__ void __launch_bounds__(1024) k_collision(KArgs A)
{
    const uint32_t n = A.ctr->nC;
    MC_WAVE_LOOP(n)
    {
        const uint32_t i = _base + _lane;
        if (i >= n) continue;
        const uint32_t s = A.qC[i];
        const int c = A.pcell[s];
        const mc_cell_desc C = A.cell[c];
        const int p_id = A.pprob[s];
        const DProb PR = A.prob[p_id];
        double w = A.wgt[s];
        uint64_t rng = A.rng[s];
        double E = A.E[s];
        bool alive = true;
        Vec3 u = {A.u[s], A.v[s], A.w[s]};
        if (C.material < 0) {                    // void: kill at collision
            alive = false;
            w = 0.0;
        } else {
            if (C.col_count > 0) {
                double* hrow = A.hist + (size_t)s * (size_t)A.bin_stride;
                const double* Srow = kGrad ? A.S + (size_t)s * (size_t)A.par_stride : nullptr;
                double* Grow = kGrad ? A.G + (size_t)s * (size_t)A.par_stride : nullptr;
                for (int a = 0; a < C.col_count; ++a)
                    score_attachment<kGrad>(A, C.col_begin + a, 0.0, c, E, w, PR, hrow, Srow, Grow);
            }
            const int m = C.material;
            const Xs3 X = material_xs3(A, m, E);
            // implicit capture, general.cpp:154-156 (Sigma_f = 0)
            const double implicit = X.c + 0.0;
            w = w * (X.t - implicit) / X.t;
            // Material::nuclide_scatter, Material.cpp:106-115
            const mc_material_desc M = A.mat[m];
            const double ur = X.s * urand(rng);
            double acc = 0.0;
            int sel = -1;
            {
                XsPoint P;
                for (int k = 0; k < M.nuc_count; ++k) {
                    const int j = M.nuc_begin + k;
                    const double N = M.param_count ? A.mat_ns[j] : A.mat_dens[j];
                    acc += nuclide_xs(A, A.mat_nuc[j], XS_S, E, P) * N;
                    if (acc > ur) { sel = k; break; }
                }
            }
            if (kGrad) {
                double* Srow = A.S + (size_t)s * (size_t)A.par_stride;
                XsPoint P;
                for (int q = 0; q < M.param_count; ++q) {
                    const DSig d = param_dsig(A, M, q, E, P);
                    double v = -(d.ds + d.dc) / X.t;
                    const mc_param_desc D = A.param[d.k];
                    if (sel >= 0 && D.nuc_slot == sel && (D.reactions & MC_REACTION_SCATTER))
                        v += 1.0 / A.theta[d.k];
                    Srow[d.k - PR.param_begin] += v;
                }
            }
            if (sel >= 0) {
                // elastic scatter, target at rest (T -> 0 limit of Reaction.cpp:27-118)
                const DNuc N = A.nuc[A.mat_nuc[M.nuc_begin + sel]];
                const double mu_cm = 2.0 * urand(rng) - 1.0;
                double mu = mu_cm;
                if (N.A < MC_MASS_INFINITE) {
                    const double Am = N.A;
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
        A.u[s] = u.x; A.v[s] = u.y; A.w[s] = u.z;
        A.E[s] = E;
        A.wgt[s] = w;
        A.rng[s] = rng;
        if (alive) {
            A.status[s] = ST_ALIVE;
        } else {
            const Vec3 p = {A.x[s], A.y[s], A.z[s]};
            finish_history<kGrad>(A, s, p_id, PR, rng, p, E, A.nflight[s]);
            A.status[s] = ST_DEAD;
        }
    }
}

// Standalone lookup over arrays (mc_xs_lookup and its timing entry point).
__global__ void __launch_bounds__(1024) k_xs_lookup_array(KArgs A, const double* __restrict__ E,
                                                         const int32_t* __restrict__ mat,
                                                         double* __restrict__ out, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int m = mat[i];
        out[i] = m >= 0 ? material_sigt(A, m, E[i]) : 0.0;
    }
}

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
template __global__ void k_refill<false>(KArgs);
template __global__ void k_refill<true>(KArgs);
template __global__ void k_flight<false, false>(KArgs);
template __global__ void k_flight<true, false>(KArgs);
template __global__ void k_flight<false, true>(KArgs);
template __global__ void k_flight<true, true>(KArgs);
template __global__ void k_surface<false>(KArgs);
template __global__ void k_surface<true>(KArgs);
template __global__ void k_collision<false>(KArgs);
template __global__ void k_collision<true>(KArgs);

}  // namespace mc
