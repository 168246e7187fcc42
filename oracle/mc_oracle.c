/*
 * mc_oracle.c — CPU restatement of ilhamv/MC-old's fixed-source random walk.
 *
 * TEST INFRASTRUCTURE ONLY (see mc_oracle.h).  Each function cites the
 * reference file:line it restates.  Deliberate differences from the
 * reference, all shared with the GPU engine:
 *   - the implicit-fission block of collision() (general.cpp:135-151) is
 *     absent: non-multiplying media only; it consumed two random numbers;
 *   - elastic scattering is the reference's free-gas model at 293.6 K
 *     (ReactionScatter::sample, Reaction.cpp:27-118) or, where a deck's
 *     nuclide asks for kinematics="target_at_rest", its zero-temperature
 *     limit; a nuclide with A >= MC_MASS_INFINITE scatters isotropically in
 *     the lab without changing energy (one-speed);
 *   - log() and cos/sin(2 pi xi) are the project's own routines below, so
 *     that CPU and GPU histories are bit-identical;
 *   - every history owns an LCG stream started by RN_init_particle
 *     (Random.cpp:196-204) at key = history index, instead of one global
 *     stream; a split copy's stream starts 2^50 draws per copy ahead of its
 *     parent's (see cell_importance);
 *   - the particle comb draws its tooth from the stream of the particle
 *     whose walk just ended, and its fill entries get jumped streams (see
 *     particle_comb).
 *
 * Build: gcc -O2 -ffp-contract=off (no FMA contraction: bit-exactness).
 */
#include "mc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

/* ------------------------------------------------------------ constants */
/* Constants.h:11-13 */
static const double EPSILON_FLOAT = (double)FLT_EPSILON;
static const double MAX_FLOAT = (double)FLT_MAX;
#define MAX_FLOAT_LESS (0.9 * MAX_FLOAT)

/* Random.cpp:92-101: MCNP5 63-bit generator #5 */
#define RN_MULT   3512401965023503517ULL
#define RN_STRIDE 152917ULL
#define RN_MASK   (~0ULL >> 1)
#define RN_PERIOD (1ULL << 61)
#define RN_NORM   (1.0 / 9223372036854775808.0)   /* 2^-63 */

/* ==================================================================== RNG */
/* Random.cpp:121-126 */
uint64_t mco_lcg_next(uint64_t s) { return (RN_MULT * s) & RN_MASK; }

double mco_urand(uint64_t* s)
{
    *s = mco_lcg_next(*s);
    return (double)(*s) * RN_NORM;
}

/* Random.cpp:130-149 (RN_ADD = 0) */
uint64_t mco_skip_ahead(uint64_t seed, int64_t n)
{
    int64_t nskip = n;
    while (nskip < 0) nskip += (int64_t)RN_PERIOD;
    uint64_t k = (uint64_t)nskip & RN_MASK;
    uint64_t gen = 1, g = RN_MULT, inc = 0, c = 0;
    for (; k; k >>= 1) {
        if (k & 1) {
            gen = (gen * g) & RN_MASK;
            inc = (inc * g + c) & RN_MASK;
        }
        c = (g * c + c) & RN_MASK;
        g = (g * g) & RN_MASK;
    }
    return (gen * seed + inc) & RN_MASK;
}

/* Random.cpp:196-204 */
uint64_t mco_init_particle(uint64_t seed0, uint64_t nps)
{
    return mco_skip_ahead(seed0, (int64_t)(nps * RN_STRIDE));
}

/* ============================================================ math helpers */
/*
 * Natural logarithm for positive normal x.  Method: x = 2^e m with
 * m in (sqrt(1/2), sqrt(2)], s = (m-1)/(m+1), ln m = 2 atanh(s) summed as the
 * odd series 2s(1 + s^2/3 + s^4/5 + ...) to s^22 (|s| <= 0.1716, truncation
 * below 1e-18), and e ln2 added in Cody-Waite hi/lo parts.  Only +,-,*,/ on
 * doubles, so every IEEE-754 machine gets the same bits.
 */
double mco_log(double x)
{
    static const double LN2_HI = 0.6931471804855391;      /* ln2, low 20 bits cleared */
    static const double LN2_LO = 7.440617110012397e-11;   /* ln2 - LN2_HI */
    static const double SQRT2 = 1.4142135623730951;
    uint64_t b;
    memcpy(&b, &x, 8);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    b = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    double m;
    memcpy(&m, &b, 8);
    if (m > SQRT2) { m = m * 0.5; e += 1; }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double r = 1.0 / 23.0;
    r = 1.0 / 21.0 + z * r;
    r = 1.0 / 19.0 + z * r;
    r = 1.0 / 17.0 + z * r;
    r = 1.0 / 15.0 + z * r;
    r = 1.0 / 13.0 + z * r;
    r = 1.0 / 11.0 + z * r;
    r = 1.0 / 9.0 + z * r;
    r = 1.0 / 7.0 + z * r;
    r = 1.0 / 5.0 + z * r;
    r = 1.0 / 3.0 + z * r;
    const double lnm = 2.0 * s + (2.0 * s) * (z * r);
    const double de = (double)e;
    return de * LN2_HI + (lnm + de * LN2_LO);
}

/*
 * cos(2 pi xi), sin(2 pi xi) for xi in [0,1].  The angle is reduced exactly on
 * xi: t = 4 xi = q + r (q quadrant, r in [0,1)); r <= 1/2 gives theta = r pi/2,
 * otherwise the co-function of (1-r) pi/2, so theta <= pi/4, where the Taylor
 * series of sin to theta^17 and cos to theta^18 are below 1e-19.  Coefficients
 * are (-1)^k/(2k+1)! and (-1)^k/(2k)! rounded to double.
 */
void mco_sincos_2pi(double xi, double* c, double* s)
{
    static const double PI_HALF = 1.5707963267948966;
    const double t = 4.0 * xi;
    int q = (int)t;
    const double r = t - (double)q;
    q &= 3;
    const int swap = r > 0.5;
    const double th = (swap ? (1.0 - r) : r) * PI_HALF;
    const double z = th * th;
    double ps = -8.22063524662433e-18;
    ps = 2.8114572543455206e-15 + z * ps;
    ps = -7.647163731819816e-13 + z * ps;
    ps = 1.6059043836821613e-10 + z * ps;
    ps = -2.505210838544172e-08 + z * ps;
    ps = 2.7557319223985893e-06 + z * ps;
    ps = -0.0001984126984126984 + z * ps;
    ps = 0.008333333333333333 + z * ps;
    ps = -0.16666666666666666 + z * ps;
    const double sn = th + th * (z * ps);
    double pc = -1.5619206968586225e-16;
    pc = 4.779477332387385e-14 + z * pc;
    pc = -1.1470745597729725e-11 + z * pc;
    pc = 2.08767569878681e-09 + z * pc;
    pc = -2.755731922398589e-07 + z * pc;
    pc = 2.48015873015873e-05 + z * pc;
    pc = -0.001388888888888889 + z * pc;
    pc = 0.041666666666666664 + z * pc;
    pc = -0.5 + z * pc;
    const double cs = 1.0 + z * pc;
    const double c0 = swap ? sn : cs;
    const double s0 = swap ? cs : sn;
    switch (q) {
    case 0: *c = c0; *s = s0; break;
    case 1: *c = -s0; *s = c0; break;
    case 2: *c = -c0; *s = -s0; break;
    default: *c = s0; *s = -c0; break;
    }
}

/* Algorithm.cpp:46-64 */
int mco_binary_search(double x, const double* v, int n)
{
    int left = 0, right = n - 1, mid;
    while (left <= right) {
        mid = (left + right) / 2;
        if (v[mid] < x) left = mid + 1;
        else right = mid - 1;
    }
    return right;
}

/* Algorithm.cpp:103-105 */
double mco_interpolate(double x, double x1, double x2, double y1, double y2)
{
    return (x - x2) / (x1 - x2) * y1 + (x - x1) / (x2 - x1) * y2;
}

/* Algorithm.cpp:16-38 */
double mco_geometry_quad(double a, double b, double c)
{
    const double D = b * b - 4.0 * a * c;
    if (D <= 0.0) return MAX_FLOAT;
    const double sqrtD = sqrt(D);
    const double ai = 0.5 / a;
    double r1 = ai * (-1.0 * b - sqrtD);
    double r2 = ai * (-1.0 * b + sqrtD);
    if (r1 < 0) r1 = MAX_FLOAT;
    if (r2 < 0) r2 = MAX_FLOAT;
    return fmin(r1, r2);
}

/* Algorithm.cpp:67-101 (azimuth 2 pi xi from one draw) */
void mco_scatter_direction(const double d[3], double mu0, uint64_t* rng, double o[3])
{
    double cos_azi, sin_azi;
    mco_sincos_2pi(mco_urand(rng), &cos_azi, &sin_azi);
    const double Ac = sqrt(1.0 - mu0 * mu0);
    if (d[2] != 1.0) {
        const double B = sqrt(1.0 - d[2] * d[2]);
        const double C = Ac / B;
        o[0] = d[0] * mu0 + (d[0] * d[2] * cos_azi - d[1] * sin_azi) * C;
        o[1] = d[1] * mu0 + (d[1] * d[2] * cos_azi + d[0] * sin_azi) * C;
        o[2] = d[2] * mu0 - cos_azi * Ac * B;
    } else {
        const double B = sqrt(1.0 - d[1] * d[1]);
        const double C = Ac / B;
        o[0] = d[0] * mu0 + (d[0] * d[1] * cos_azi - d[2] * sin_azi) * C;
        o[2] = d[2] * mu0 + (d[2] * d[1] * cos_azi + d[0] * sin_azi) * C;
        o[1] = d[1] * mu0 - cos_azi * Ac * B;
    }
}

/* =============================================================== geometry */
/* Geometry.cpp:29-69 */
double mco_surface_eval(const mc_surface_desc* S, const double p[3])
{
    const double* c = S->c;
    switch (S->type) {
    case MC_SURF_PLANE_X: return p[0] - c[0];
    case MC_SURF_PLANE_Y: return p[1] - c[0];
    case MC_SURF_PLANE_Z: return p[2] - c[0];
    case MC_SURF_PLANE: return c[0] * p[0] + c[1] * p[1] + c[2] * p[2] - c[3];
    case MC_SURF_SPHERE: {
        const double xt = p[0] - c[0], yt = p[1] - c[1], zt = p[2] - c[2];
        return xt * xt + yt * yt + zt * zt - c[4];
    }
    case MC_SURF_CYL_X: {
        const double yt = p[1] - c[0], zt = p[2] - c[1];
        return yt * yt + zt * zt - c[4];
    }
    case MC_SURF_CYL_Z: {
        const double xt = p[0] - c[0], yt = p[1] - c[1];
        return xt * xt + yt * yt - c[4];
    }
    }
    return 0.0;
}

static double plane_axis_distance(double pos, double dir, double loc)
{
    /* Geometry.cpp:76-126 */
    if (fabs(dir) > EPSILON_FLOAT) {
        const double dist = (loc - pos) / dir;
        if (dist > 0.0) return dist;
        return MAX_FLOAT;
    }
    return MAX_FLOAT;
}

/* Geometry.cpp:76-188 */
double mco_surface_distance(const mc_surface_desc* S, const double p[3], const double u[3])
{
    const double* c = S->c;
    switch (S->type) {
    case MC_SURF_PLANE_X: return plane_axis_distance(p[0], u[0], c[0]);
    case MC_SURF_PLANE_Y: return plane_axis_distance(p[1], u[1], c[0]);
    case MC_SURF_PLANE_Z: return plane_axis_distance(p[2], u[2], c[0]);
    case MC_SURF_PLANE: {
        const double denom = c[0] * u[0] + c[1] * u[1] + c[2] * u[2];
        if (fabs(denom) > EPSILON_FLOAT) {
            const double dist = (c[3] - c[0] * p[0] - c[1] * p[1] - c[2] * p[2]) / denom;
            if (dist > 0.0) return dist;
            return MAX_FLOAT;
        }
        return MAX_FLOAT;
    }
    case MC_SURF_SPHERE: {
        const double b = 2.0 * ((p[0] - c[0]) * u[0] + (p[1] - c[1]) * u[1] + (p[2] - c[2]) * u[2]);
        return mco_geometry_quad(1.0, b, mco_surface_eval(S, p));
    }
    case MC_SURF_CYL_X: {
        const double a = 1.0 - u[0] * u[0];
        const double b = 2.0 * ((p[1] - c[0]) * u[1] + (p[2] - c[1]) * u[2]);
        return mco_geometry_quad(a, b, mco_surface_eval(S, p));
    }
    case MC_SURF_CYL_Z: {
        const double a = 1.0 - u[2] * u[2];
        const double b = 2.0 * ((p[0] - c[0]) * u[0] + (p[1] - c[1]) * u[1]);
        return mco_geometry_quad(a, b, mco_surface_eval(S, p));
    }
    }
    return MAX_FLOAT;
}

/* Geometry.cpp:195-222 (quadrics do not reflect) */
void mco_surface_reflect(const mc_surface_desc* S, double u[3])
{
    const double* c = S->c;
    switch (S->type) {
    case MC_SURF_PLANE_X: u[0] = -u[0]; break;
    case MC_SURF_PLANE_Y: u[1] = -u[1]; break;
    case MC_SURF_PLANE_Z: u[2] = -u[2]; break;
    case MC_SURF_PLANE: {
        const double K = (c[0] * u[0] + c[1] * u[1] + c[2] * u[2]);
        const double qx = u[0] - c[4] * K;
        const double qy = u[1] - c[5] * K;
        const double qz = u[2] - c[6] * K;
        u[0] = qx; u[1] = qy; u[2] = qz;
        break;
    }
    default: break;
    }
}

/* ========================================================= cross sections */
/* XSec.cpp:9-40 with Nuclide::checkE's binary search, Nuclide.cpp:18-24 */
static double table_xs(const double* grid, const double* y, int n, double E)
{
    const int idx = mco_binary_search(E, grid, n);
    if (idx == n - 1) return y[n - 1];
    if (idx == -1) return y[0];
    return mco_interpolate(E, grid[idx], grid[idx + 1], y[idx], y[idx + 1]);
}

double mco_nuclide_xs(const mc_problem_desc* d, int nuc, int which, double E)
{
    const mc_nuclide_desc* N = &d->nuclides[nuc];
    if (N->kind == MC_NUC_CONSTANT) {
        if (which == MCO_XS_SCATTER) return N->xs_scatter;
        if (which == MCO_XS_CAPTURE) return N->xs_capture;
        return N->xs_total;
    }
    const double* grid = d->xs_energy + N->grid_begin;
    const double* y = which == MCO_XS_SCATTER ? d->xs_scatter
                    : which == MCO_XS_CAPTURE ? d->xs_capture : d->xs_total;
    return table_xs(grid, y + N->grid_begin, N->grid_count, E);
}

/* Per-run effective densities: theta scales a (material, nuclide, reaction). */
typedef struct {
    double* ns;   /* scatter density per mat_nuclide entry */
    double* nc;   /* capture density */
} scaled_t;

static void build_scaled(const mc_problem_desc* d, const double* theta, scaled_t* sc)
{
    int total = 0;
    for (int m = 0; m < d->n_materials; ++m) {
        const int e = d->materials[m].nuc_begin + d->materials[m].nuc_count;
        if (e > total) total = e;
    }
    sc->ns = (double*)malloc(sizeof(double) * (size_t)(total + 1));
    sc->nc = (double*)malloc(sizeof(double) * (size_t)(total + 1));
    for (int k = 0; k < total; ++k) { sc->ns[k] = d->mat_density[k]; sc->nc[k] = d->mat_density[k]; }
    for (int k = 0; k < d->n_params; ++k) {
        const mc_param_desc* P = &d->params[k];
        const int j = d->materials[P->material].nuc_begin + P->nuc_slot;
        const double t = theta ? theta[k] : 1.0;
        if (P->reactions & MC_REACTION_SCATTER) sc->ns[j] = t * d->mat_density[j];
        if (P->reactions & MC_REACTION_CAPTURE) sc->nc[j] = t * d->mat_density[j];
    }
}

/* Material.cpp:18-57.  Unparameterised materials sum sigma_x * N in nuclide
 * order exactly as the reference; a material carrying gradient parameters
 * uses per-reaction densities and Sigma_t = Sigma_s + Sigma_c.            */
typedef struct { double t, s, c; } xs3;

static xs3 material_xs3(const mc_problem_desc* d, const scaled_t* sc, int mat, double E)
{
    xs3 r = {0.0, 0.0, 0.0};
    const mc_material_desc* M = &d->materials[mat];
    if (M->param_count == 0) {
        for (int k = 0; k < M->nuc_count; ++k) {
            const int j = M->nuc_begin + k, n = d->mat_nuclide[j];
            const double N = d->mat_density[j];
            r.t += mco_nuclide_xs(d, n, MCO_XS_TOTAL, E) * N;
        }
        for (int k = 0; k < M->nuc_count; ++k) {
            const int j = M->nuc_begin + k, n = d->mat_nuclide[j];
            r.s += mco_nuclide_xs(d, n, MCO_XS_SCATTER, E) * d->mat_density[j];
        }
        for (int k = 0; k < M->nuc_count; ++k) {
            const int j = M->nuc_begin + k, n = d->mat_nuclide[j];
            r.c += mco_nuclide_xs(d, n, MCO_XS_CAPTURE, E) * d->mat_density[j];
        }
    } else {
        for (int k = 0; k < M->nuc_count; ++k) {
            const int j = M->nuc_begin + k, n = d->mat_nuclide[j];
            r.s += mco_nuclide_xs(d, n, MCO_XS_SCATTER, E) * sc->ns[j];
        }
        for (int k = 0; k < M->nuc_count; ++k) {
            const int j = M->nuc_begin + k, n = d->mat_nuclide[j];
            r.c += mco_nuclide_xs(d, n, MCO_XS_CAPTURE, E) * sc->nc[j];
        }
        r.t = r.s + r.c;
    }
    return r;
}

double mco_material_xs(const mc_problem_desc* d, const double* theta, int mat, int which, double E)
{
    scaled_t sc;
    build_scaled(d, theta, &sc);
    const xs3 r = material_xs3(d, &sc, mat, E);
    free(sc.ns);
    free(sc.nc);
    return which == MCO_XS_SCATTER ? r.s : which == MCO_XS_CAPTURE ? r.c : r.t;
}

/* ================================================================ particle */
typedef struct {
    double p[3], u[3];
    double E, w;
    int cell, cell_old, surf_old, alive;
    uint64_t rng;
    /* likelihood-ratio state (gradient runs).  A score of gradient weight gv
     * made when the score function is S contributes gv * S to the gradient;
     * summed over a particle that is Q * S_end - H, with Q the sum of gv and
     * H the sum over increments dS of dS * (Q when dS was added).  Terms a
     * score adds on top of S (its own d ln Sigma_x, the pending flight) are
     * kept as coefficients cs, cc of dSigma_s/dtheta and dSigma_c/dtheta of
     * the current material at E, folded into H with the pending flight.    */
    double* S;
    double* H;
    double Q, cs, cc;
    double lpend;     /* length of the last flight, not yet folded into S     */
} opart;

typedef struct {
    const mc_problem_desc* d;
    scaled_t sc;
    double* hist;       /* per-history tally (Tally::hist)        */
    int* touched;       /* bins with hist != 0 this history       */
    unsigned char* is_touched;
    int n_touched;
    uint64_t flights;
    const double* theta;
    const double* gout; /* gradient weights per bin (NULL: off)   */
    double* G;          /* per-history gradient                   */
    int P;              /* number of parameters                   */
    /* LIFO bank of split particles (handler.cpp:21-29) */
    opart* bank;
    int bank_n, bank_cap;
    int error;
} ostate;

static int search_cell(const mc_problem_desc* d, const double p[3])
{
    /* general.cpp:13-34 */
    for (int c = 0; c < d->n_cells; ++c) {
        const mc_cell_desc* C = &d->cells[c];
        int inside = 1;
        for (int k = 0; k < C->surf_count; ++k) {
            const int s = d->cell_surf[C->surf_begin + k];
            if (mco_surface_eval(&d->surfaces[s], p) * (double)d->cell_sense[C->surf_begin + k] < 0) {
                inside = 0;
                break;
            }
        }
        if (inside) return c;
    }
    return -1;
}

static void kill(opart* P) { P->alive = 0; P->w = 0.0; }   /* Particle.cpp:77-81 */

static void move(opart* P, double l)    /* Particle.cpp:66-76 (time omitted) */
{
    P->p[0] += P->u[0] * l;
    P->p[1] += P->u[1] * l;
    P->p[2] += P->u[2] * l;
}

/* derivatives of Sigma_{t,s,c} w.r.t. the parameters acting on material m */
typedef struct { int k; double ds, dc; } dsig;

static int material_dsig(ostate* st, int mat, double E, dsig* out)
{
    const mc_problem_desc* d = st->d;
    if (mat < 0) return 0;
    const mc_material_desc* M = &d->materials[mat];
    for (int q = 0; q < M->param_count; ++q) {
        const int k = d->mat_param[M->param_begin + q];
        const mc_param_desc* Pd = &d->params[k];
        const int j = M->nuc_begin + Pd->nuc_slot, n = d->mat_nuclide[j];
        const double N = d->mat_density[j];
        out[q].k = k;
        out[q].ds = (Pd->reactions & MC_REACTION_SCATTER) ? mco_nuclide_xs(d, n, MCO_XS_SCATTER, E) * N : 0.0;
        out[q].dc = (Pd->reactions & MC_REACTION_CAPTURE) ? mco_nuclide_xs(d, n, MCO_XS_CAPTURE, E) * N : 0.0;
    }
    return M->param_count;
}

static void add_hist(ostate* st, int bin, double v)
{
    if (!st->is_touched[bin]) { st->is_touched[bin] = 1; st->touched[st->n_touched++] = bin; }
    st->hist[bin] += v;
}

/* Estimator::score, Estimator.cpp:298-336, for one attachment.
 * l: track length (0 for surface and collision events); coll_end: the
 * flight being scored ends in a collision (its length density then carries
 * Sigma_t, whose log-derivative joins the score's).
 *
 * Likelihood-ratio bookkeeping is deferred: a flight of length l leaves
 * lpend = l instead of adding -l dSigma_t/dtheta to S, and the next
 * collision or surface crossing folds it in.  A score made before that fold
 * adds the pending term itself, so every score sees the same total
 * d ln p / d theta as if S were updated at the flight.                    */
static void estimator_score(ostate* st, opart* P, int att, double l, int coll_end)
{
    const mc_problem_desc* d = st->d;
    const int e = d->att_estimator[att];
    const mc_estimator_desc* Es = &d->estimators[e];
    int ebin = 0, NE = 1;
    if (Es->egrid_count > 0) {
        NE = Es->egrid_count - 1;
        ebin = mco_binary_search(P->E, d->egrid + Es->egrid_begin, Es->egrid_count);
        if (ebin < 0 || ebin >= NE) return;           /* Estimator.cpp:174 */
    }
    const int fi = d->att_index[att];
    const int mat = d->cells[P->cell].material;
    xs3 X = {0.0, 0.0, 0.0};
    int have_x = 0;
    dsig ds[64];
    int nds = 0;
    for (int s = 0; s < Es->n_scores; ++s) {
        const int kern = Es->score_kernel[s], sc = Es->score[s];
        if (!have_x && (st->gout || kern == MC_KERNEL_COLLISION || sc != MC_SCORE_FLUX)) {
            if (mat >= 0) X = material_xs3(d, &st->sc, mat, P->E);
            if (st->gout) nds = material_dsig(st, mat, P->E, ds);
            have_x = 1;
        }
        double kv;                                       /* Estimator.cpp:17-41 */
        if (kern == MC_KERNEL_TRACK_LENGTH) kv = P->w * l;
        else if (kern == MC_KERNEL_COLLISION) kv = P->w / X.t;
        else kv = P->w;
        double v, sx = 1.0;                               /* Estimator.cpp:48-124 */
        switch (sc) {
        case MC_SCORE_ABSORPTION: sx = X.c; v = X.c * kv; break;
        case MC_SCORE_SCATTER: sx = X.s; v = X.s * kv; break;
        case MC_SCORE_CAPTURE: sx = X.c; v = X.c * kv; break;
        case MC_SCORE_TOTAL: sx = X.t; v = X.t * kv; break;
        default: v = kv; break;
        }
        const int bin = Es->bin_offset + s * Es->n_attach * NE + fi * NE + ebin;
        add_hist(st, bin, v);
        if (st->gout && st->gout[bin] != 0.0 && v != 0.0) {
            /* d ln(score)/d theta = S - lpend dSigma_t + [track ends in a
             * collision] dSigma_t/Sigma_t + d ln Sigma_x.  A collision
             * estimator's 1/Sigma_t cancels the collision density of the
             * flight that brought the particle here.                        */
            const double gv = st->gout[bin] * v;
            P->Q += gv;
            if (nds > 0) {
                double a = -P->lpend;
                if (kern == MC_KERNEL_TRACK_LENGTH && coll_end) a += 1.0 / X.t;
                double bs = a, bc = a;
                if (sc == MC_SCORE_SCATTER) bs += 1.0 / sx;
                else if (sc == MC_SCORE_ABSORPTION || sc == MC_SCORE_CAPTURE) bc += 1.0 / sx;
                else if (sc == MC_SCORE_TOTAL) { bs += 1.0 / sx; bc += 1.0 / sx; }
                P->cs += gv * bs;
                P->cc += gv * bc;
            }
        }
    }
}

static void bank_push(ostate* st, const opart* P)
{
    if (st->bank_n == st->bank_cap) {
        st->bank_cap = st->bank_cap ? 2 * st->bank_cap : 64;
        st->bank = (opart*)realloc(st->bank, sizeof(opart) * (size_t)st->bank_cap);
    }
    opart* Q = &st->bank[st->bank_n++];
    *Q = *P;
    if (P->S) {
        /* the copy carries the score function; the scores made so far stay
         * with the parent                                                  */
        Q->S = (double*)malloc(sizeof(double) * (size_t)st->P);
        memcpy(Q->S, P->S, sizeof(double) * (size_t)st->P);
        Q->H = (double*)calloc((size_t)st->P, sizeof(double));
        Q->Q = 0.0;
        Q->cs = 0.0;
        Q->cc = 0.0;
    }
}

/* population_control.cpp:21-49.  Split copies take fresh streams: copy i
 * (from 0) starts (i + 1) 2^50 draws ahead of the parent's current state.
 * A history's stream is the block of RN_STRIDE draws at key * RN_STRIDE;
 * 2^50 draws are 7.4e9 blocks, so a copy's stream lies beyond the blocks of
 * every key below 7.4e9 (runs whose keys stay there cannot meet a copy's
 * stream), and nested copies add further multiples of 2^50 (2048 of them
 * before the 2^61-draw period wraps).                                     */
static void cell_importance(ostate* st, opart* P)
{
    const mc_problem_desc* d = st->d;
    const double Iold = d->cells[P->cell_old].importance;
    const double Inew = d->cells[P->cell].importance;
    if (Inew == Iold) return;
    const double rat = Inew / Iold;
    if (rat < 1.0) {
        if (mco_urand(&P->rng) < rat) P->w = P->w / rat;
        else kill(P);
    } else {
        const int n = (int)floor(rat + mco_urand(&P->rng));
        P->w = P->w / (double)n;
        for (int i = 0; i < n - 1; ++i) {
            opart C = *P;
            C.rng = mco_skip_ahead(P->rng, (int64_t)(i + 1) << MCO_COPY_JUMP_LOG2);
            bank_push(st, &C);
        }
    }
}

/* Particle comb, population_control.cpp:55-84, after each particle's walk
 * (handler.cpp:24-28): a bank of bank_max or more copies is combed to
 * comb_teeth, each selected copy taking the average weight; the selection
 * walks the bank in order with one tooth per selected copy.  Entries past
 * the last selected one stay copies of the bank's first entry, as the
 * reference fills them (bank_tmp(comb_teeth, Pbank[0])), weight included.
 * The tooth's random number is the next draw of the stream of the particle
 * whose walk just ended (the reference draws it from its one global
 * stream), and the fill copies' streams are jumped like split copies' (k-th
 * fill: (k + 1) 2^50 draws ahead of the first entry's), so that no two
 * entries walk the same stream.  No comb in gradient runs.                 */
int mco_comb_plan(const double* w, int n, int teeth, double u, int* src, double* wout, int* fill)
{
    double W = 0.0;
    for (int i = 0; i < n; ++i) W += w[i];
    const double w_avg = W / (double)teeth;
    double tooth = u * w_avg;
    double sum = 0.0;
    int j = 0;
    /* j < teeth: the last tooth lies below W by u w_avg; a first tooth within
     * rounding of 0 could otherwise select one entry more than the reference's
     * bank_tmp holds */
    for (int i = 0; i < n && j < teeth; ++i) {
        sum += w[i];
        if (sum > tooth) {
            src[j] = i;
            wout[j] = w_avg;
            fill[j] = 0;
            tooth += w_avg;
            ++j;
        }
    }
    const int selected = j;
    for (int k = selected; k < teeth; ++k) {
        src[k] = 0;
        wout[k] = w[0];
        fill[k] = k - selected + 1;
    }
    return selected;
}

static void particle_comb(ostate* st, uint64_t* rng)
{
    const int bank_max = st->d->comb_bank_max, teeth = st->d->comb_teeth;
    if (bank_max <= 0 || st->bank_n < bank_max) return;
    const int n = st->bank_n;
    double* w = (double*)malloc(sizeof(double) * (size_t)n);
    double* wout = (double*)malloc(sizeof(double) * (size_t)teeth);
    int* src = (int*)malloc(sizeof(int) * (size_t)teeth);
    int* fill = (int*)malloc(sizeof(int) * (size_t)teeth);
    for (int i = 0; i < n; ++i) w[i] = st->bank[i].w;
    mco_comb_plan(w, n, teeth, mco_urand(rng), src, wout, fill);
    opart* out = (opart*)malloc(sizeof(opart) * (size_t)teeth);
    for (int k = 0; k < teeth; ++k) {
        out[k] = st->bank[src[k]];
        out[k].w = wout[k];
        if (fill[k]) out[k].rng = mco_skip_ahead(st->bank[0].rng, (int64_t)fill[k] << MCO_COPY_JUMP_LOG2);
    }
    free(st->bank);
    st->bank = out;
    st->bank_n = st->bank_cap = teeth;
    free(w); free(wout); free(src); free(fill);
}

/* Fold the pending terms of the current material at E into S and H: the
 * scores' coefficients cs, cc, then the flight's -lpend dSigma_t/dtheta plus
 * extra[q] (the nuclide-selection term at a collision).                    */
static void fold_pending(ostate* st, opart* P, const dsig* ds, int nds, const double* extra)
{
    for (int q = 0; q < nds; ++q) {
        const int k = ds[q].k;
        P->H[k] -= P->cs * ds[q].ds + P->cc * ds[q].dc;
        double v = -P->lpend * (ds[q].ds + ds[q].dc);
        if (extra) v += extra[q];
        P->S[k] += v;
        P->H[k] += v * P->Q;
    }
    P->lpend = 0.0;
    P->cs = 0.0;
    P->cc = 0.0;
}

/* a flight that ended on a surface, in the cell it leaves */
static void fold_surface_flight(ostate* st, opart* P)
{
    dsig ds[64];
    const int nds = material_dsig(st, st->d->cells[P->cell].material, P->E, ds);
    fold_pending(st, P, ds, nds, NULL);
}

/* end of a particle: scores still pending, then its gradient Q S - H */
static void close_particle(ostate* st, opart* P)
{
    if (P->cs != 0.0 || P->cc != 0.0) {
        dsig ds[64];
        const int nds = material_dsig(st, st->d->cells[P->cell].material, P->E, ds);
        for (int q = 0; q < nds; ++q) P->H[ds[q].k] -= P->cs * ds[q].ds + P->cc * ds[q].dc;
        P->cs = 0.0;
        P->cc = 0.0;
    }
    for (int k = 0; k < st->P; ++k) st->G[k] += P->Q * P->S[k] - P->H[k];
}

/* general.cpp:89-115 */
static void surface_hit(ostate* st, opart* P, int S)
{
    const mc_problem_desc* d = st->d;
    const mc_surface_desc* Sd = &d->surfaces[S];
    if (st->gout) fold_surface_flight(st, P);
    P->surf_old = S;
    if (Sd->bc == MC_BC_TRANSMISSION) {
        move(P, EPSILON_FLOAT);
        const int c = search_cell(d, P->p);
        if (c < 0) { st->error = MC_ERR_LOST_PARTICLE; kill(P); return; }
        P->cell_old = P->cell;
        P->cell = c;
    } else if (Sd->bc == MC_BC_VACUUM) {
        kill(P);
        P->cell_old = P->cell;
    } else {
        mco_surface_reflect(Sd, P->u);
        move(P, EPSILON_FLOAT);
        P->cell_old = P->cell;
    }
    for (int a = 0; a < Sd->est_count; ++a) estimator_score(st, P, Sd->est_begin + a, 0.0, 0);
    cell_importance(st, P);
}

/* ReactionScatter::sample, Reaction.cpp:27-118: elastic scatter off a free
 * gas at 293.6 K with a constant cross section.  mu0 isotropic in the centre
 * of mass (Distribution.cpp:74-77); target speed and cosine sampled by the
 * reference's two-branch rejection (its Eq. 37-41); centre-of-mass
 * kinematics; speed and energy with Particle.cpp:42-56's constants.  The
 * reference leaves the order of the two draws in -log(U) - log(U) cos^2
 * unspecified (C++ operand order): they are taken left to right here and on
 * the GPU.  cos(pi/2 xi) is cos(2 pi (xi/4)).                              */
void mco_free_gas_scatter(double A, double* E, double u[3], uint64_t* rng)
{
    const double mu0 = 2.0 * mco_urand(rng) - 1.0;
    const double beta = sqrt(2.0659834e-11 * A);
    const double v = 13831.5926439 * sqrt(*E) * 100.0;
    const double y = beta * v;
    double V, mut;
    do {
        double x;
        if (mco_urand(rng) < 2.0 / (2.0 + 1.7724538509055159 * y)) {
            const double u1 = mco_urand(rng);
            const double u2 = mco_urand(rng);
            x = sqrt(-mco_log(u1 * u2));
        } else {
            double c, sn;
            mco_sincos_2pi(0.25 * mco_urand(rng), &c, &sn);
            const double u1 = mco_urand(rng);
            const double u2 = mco_urand(rng);
            x = sqrt(-mco_log(u1) - mco_log(u2) * c * c);
        }
        V = x / beta;
        mut = 2.0 * mco_urand(rng) - 1.0;
    } while (mco_urand(rng) > sqrt(v * v + V * V - 2.0 * v * V * mut) / (v + V));
    double nd[3], cm[3], vc[3], dc[3], dn[3], vl[3];
    mco_scatter_direction(u, mut, rng, nd);                   /* target direction */
    for (int k = 0; k < 3; ++k) {
        vl[k] = v * u[k];
        cm[k] = (vl[k] + A * (nd[k] * V)) / (1.0 + A);         /* centre-of-mass velocity */
        vc[k] = vl[k] - cm[k];
    }
    const double sc = sqrt(vc[0] * vc[0] + vc[1] * vc[1] + vc[2] * vc[2]);
    for (int k = 0; k < 3; ++k) dc[k] = vc[k] / sc;
    mco_scatter_direction(dc, mu0, rng, dn);
    for (int k = 0; k < 3; ++k) vl[k] = sc * dn[k] + cm[k];
    const double sp = sqrt(vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2]);
    *E = 5.2270376e-13 * sp * sp;
    for (int k = 0; k < 3; ++k) u[k] = vl[k] / sp;
}

/* Elastic scatter off a target at rest, the T -> 0 limit of
 * ReactionScatter::sample (Reaction.cpp:27-118): the centre-of-mass cosine
 * mu_cm = 2 xi - 1 (DistributionIsotropicScatter, Distribution.cpp:74-77),
 * then two-body kinematics with target mass ratio A,
 *   E' = E (A^2 + 2 A mu_cm + 1) / (A + 1)^2,
 *   mu_lab = (1 + A mu_cm) / sqrt(A^2 + 2 A mu_cm + 1),
 * and the new direction from scatter_direction (Algorithm.cpp:67-101).  An
 * infinitely heavy target (A >= MC_MASS_INFINITE) keeps E and scatters with
 * mu_lab = mu_cm (one-speed media).                                         */
void mco_elastic_target_at_rest(double A, double* E, double u[3], uint64_t* rng)
{
    const double mu_cm = 2.0 * mco_urand(rng) - 1.0;
    double mu = mu_cm;
    if (A < MC_MASS_INFINITE) {
        const double q = A * A + 2.0 * A * mu_cm + 1.0;
        *E = *E * q / ((A + 1.0) * (A + 1.0));
        mu = (1.0 + A * mu_cm) / sqrt(q);
    }
    double o[3];
    mco_scatter_direction(u, mu, rng, o);
    u[0] = o[0]; u[1] = o[1]; u[2] = o[2];
}

/* elastic scatter of a collision: free gas (above) or target at rest */
static void scatter(ostate* st, opart* P, int nuc)
{
    const mc_nuclide_desc* N = &st->d->nuclides[nuc];
    if (N->A < MC_MASS_INFINITE && N->kinematics == MC_KIN_FREE_GAS) {
        mco_free_gas_scatter(N->A, &P->E, P->u, &P->rng);
        return;
    }
    mco_elastic_target_at_rest(N->A, &P->E, P->u, &P->rng);
}

/* general.cpp:121-163 without the fission block */
static void collision(ostate* st, opart* P, double d_flight)
{
    const mc_problem_desc* d = st->d;
    const mc_cell_desc* C = &d->cells[P->cell];
    (void)d_flight;
    if (C->material < 0) { kill(P); return; }
    for (int a = 0; a < C->col_count; ++a) estimator_score(st, P, C->col_begin + a, 0.0, 0);
    const int mat = C->material;
    const xs3 X = material_xs3(d, &st->sc, mat, P->E);
    /* implicit capture, general.cpp:154-156 (Sigma_f = 0) */
    const double implicit = X.c + 0.0;
    P->w = P->w * (X.t - implicit) / X.t;
    /* Material::nuclide_scatter, Material.cpp:106-115 */
    const mc_material_desc* M = &d->materials[mat];
    const double u = X.s * mco_urand(&P->rng);
    double s = 0.0;
    int sel = -1;
    for (int k = 0; k < M->nuc_count; ++k) {
        const int j = M->nuc_begin + k, n = d->mat_nuclide[j];
        const double N = M->param_count ? st->sc.ns[j] : d->mat_density[j];
        s += mco_nuclide_xs(d, n, MCO_XS_SCATTER, P->E) * N;
        if (s > u) { sel = k; break; }
    }
    if (st->gout) {
        /* the flight that ended here contributes -lpend dSigma_t + dSigma_t/Sigma_t,
         * the implicit-capture weight factor -dSigma_t/Sigma_t, the nuclide
         * selection [scatter of sel]/theta; the two 1/Sigma_t terms cancel   */
        dsig ds[64];
        double sel_term[64];
        const int nds = material_dsig(st, mat, P->E, ds);
        for (int q = 0; q < nds; ++q) {
            const mc_param_desc* Pd = &d->params[ds[q].k];
            sel_term[q] = (sel >= 0 && Pd->nuc_slot == sel && (Pd->reactions & MC_REACTION_SCATTER))
                              ? 1.0 / (st->theta ? st->theta[ds[q].k] : 1.0) : 0.0;
        }
        fold_pending(st, P, ds, nds, sel_term);
    }
    if (sel < 0) return;
    scatter(st, P, d->mat_nuclide[M->nuc_begin + sel]);
}

/* general.cpp:177-211 */
static void random_walk(ostate* st, opart* P)
{
    const mc_problem_desc* d = st->d;
    while (P->alive) {
        const mc_cell_desc* C = &d->cells[P->cell];
        /* surface_intersect, general.cpp:54-67 */
        double dist = MAX_FLOAT;
        int S = -1;
        for (int k = 0; k < C->surf_count; ++k) {
            const int s = d->cell_surf[C->surf_begin + k];
            const double dd = mco_surface_distance(&d->surfaces[s], P->p, P->u);
            if (dd < dist) { dist = dd; S = s; }
        }
        /* collision_distance, general.cpp:40-48 */
        double dcol, sigt = 0.0;
        if (C->material >= 0) {
            sigt = material_xs3(d, &st->sc, C->material, P->E).t;
            dcol = -mco_log(mco_urand(&P->rng)) / sigt;
        } else {
            dcol = MAX_FLOAT_LESS;
        }
        const int hit = dcol > dist;
        const double l = hit ? dist : dcol;
        if (hit && S < 0) { st->error = MC_ERR_LOST_PARTICLE; kill(P); return; }
        P->lpend = l;   /* folded into S at the collision or surface that ends the flight */
        /* move_particle, general.cpp:73-83 */
        move(P, l);
        st->flights++;
        for (int a = 0; a < C->tl_count; ++a) estimator_score(st, P, C->tl_begin + a, l, !hit);
        if (hit) surface_hit(st, P, S);
        else collision(st, P, l);
        /* weight_roulette, population_control.cpp:9-15 */
        if (P->w < d->wr) {
            if (mco_urand(&P->rng) < P->w / d->ws) P->w = d->ws;
            else kill(P);
        }
        if (st->error) return;
    }
    if (st->gout) close_particle(st, P);
}

/* DistributionIsotropicDirection::sample (Distribution.cpp:78-92): polar
 * cosine mu = 2 xi1 - 1 on the x axis, azimuth 2 pi xi2 in the y-z plane.  */
void mco_isotropic_direction(uint64_t* rng, double u[3])
{
    const double mu = 2.0 * mco_urand(rng) - 1.0;
    double ca, sa;
    mco_sincos_2pi(mco_urand(rng), &ca, &sa);
    const double c = sqrt(1.0 - mu * mu);
    u[1] = ca * c;
    u[2] = sa * c;
    u[0] = mu;
}

/* handler.cpp:11-48 for one cycle of fixed source */
int mco_run(const mc_problem_desc* d, const double* theta, uint64_t key_offset,
            uint64_t n_hist, double* sum, double* sq, uint64_t* flights,
            mc_trace_record* trace, uint64_t trace_count,
            const double* grad_out, double* gsum, double* gsq)
{
    if (!d || !sum || !sq) return MC_ERR_ARGUMENT;
    if (d->comb_bank_max > 0 && grad_out) return MC_ERR_UNSUPPORTED;   /* no comb in gradient runs */
    ostate st;
    memset(&st, 0, sizeof(st));
    st.d = d;
    st.theta = theta;
    build_scaled(d, theta, &st.sc);
    const int nb = d->n_bins > 0 ? d->n_bins : 1;
    st.hist = (double*)calloc((size_t)nb, sizeof(double));
    st.touched = (int*)malloc(sizeof(int) * (size_t)nb);
    st.is_touched = (unsigned char*)calloc((size_t)nb, 1);
    st.P = d->n_params;
    st.gout = (grad_out && gsum && gsq && d->n_params > 0) ? grad_out : NULL;
    double *Sbuf = NULL, *Hbuf = NULL;
    if (st.gout) {
        st.G = (double*)calloc((size_t)st.P, sizeof(double));
        Sbuf = (double*)calloc((size_t)st.P, sizeof(double));
        Hbuf = (double*)calloc((size_t)st.P, sizeof(double));
    }
    for (uint64_t h = 0; h < n_hist && !st.error; ++h) {
        opart P;
        memset(&P, 0, sizeof(P));
        P.rng = mco_init_particle(d->seed, key_offset + h);
        /* SourceBank::get_source, Source.cpp:42-45, then SourcePoint, :20-24 */
        const double xi = mco_urand(&P.rng);
        int si = mco_binary_search(xi, d->source_cdf, d->n_sources + 1);
        si = si < 0 ? 0 : (si >= d->n_sources ? d->n_sources - 1 : si);   /* cdf[n] may round below 1 */
        const mc_source_desc* src = &d->sources[si];
        P.p[0] = src->pos[0]; P.p[1] = src->pos[1]; P.p[2] = src->pos[2];
        if (src->dir_kind == MC_DIR_DELTA) {
            P.u[0] = src->dir[0]; P.u[1] = src->dir[1]; P.u[2] = src->dir[2];
        } else {
            mco_isotropic_direction(&P.rng, P.u);
        }
        if (src->energy_kind == MC_ENERGY_UNIFORM)   /* Distribution.cpp:30-33 */
            P.E = src->energy[0] + mco_urand(&P.rng) * (src->energy[1] - src->energy[0]);
        else
            P.E = src->energy[0];
        P.w = 1.0;
        P.alive = 1;
        P.cell = src->cell;
        P.cell_old = src->cell;
        P.surf_old = -1;
        if (st.gout) {
            memset(Sbuf, 0, sizeof(double) * (size_t)st.P);
            memset(Hbuf, 0, sizeof(double) * (size_t)st.P);
            P.S = Sbuf;
            P.H = Hbuf;
        }
        const uint64_t f0 = st.flights;
        random_walk(&st, &P);
        opart last = P;
        particle_comb(&st, &P.rng);
        while (st.bank_n > 0 && !st.error) {
            opart Q = st.bank[--st.bank_n];
            random_walk(&st, &Q);
            last = Q;
            particle_comb(&st, &Q.rng);
            if (Q.S) { free(Q.S); free(Q.H); }
        }
        /* Estimator::end_history, Estimator.cpp:339-346 */
        double hscore = 0.0;
        if (trace && h < trace_count) {
            for (int b = 0; b < nb; ++b) hscore += st.hist[b];
        }
        for (int t = 0; t < st.n_touched; ++t) {
            const int b = st.touched[t];
            sum[b] += st.hist[b];
            sq[b] += st.hist[b] * st.hist[b];
            st.hist[b] = 0.0;
            st.is_touched[b] = 0;
        }
        st.n_touched = 0;
        if (st.gout) {
            for (int k = 0; k < st.P; ++k) {
                gsum[k] += st.G[k];
                gsq[k] += st.G[k] * st.G[k];
                st.G[k] = 0.0;
            }
        }
        if (trace && h < trace_count) {
            mc_trace_record* T = &trace[h];
            T->rng = last.rng;
            T->flights = st.flights - f0;
            T->x = last.p[0]; T->y = last.p[1]; T->z = last.p[2];
            T->E = last.E;
            T->score = hscore;
        }
    }
    if (flights) *flights += st.flights;
    free(st.hist); free(st.touched); free(st.is_touched);
    free(st.sc.ns); free(st.sc.nc);
    while (st.bank_n > 0) {          /* copies left by an error */
        opart* Q = &st.bank[--st.bank_n];
        if (Q->S) { free(Q->S); free(Q->H); }
    }
    free(st.G); free(Sbuf); free(Hbuf); free(st.bank);
    return st.error;
}
