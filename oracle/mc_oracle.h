/*
 * mc_oracle.h — CPU restatement of ilhamv/MC-old's fixed-source random walk.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this library: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, as the checker.
 *
 * The oracle consumes the flat problem view of include/mc_api.h (the same
 * boundary the product takes) and replays each history exactly as the
 * reference's loops do (handler.cpp:11-48, general.cpp:13-211,
 * population_control.cpp:9-49, Estimator.cpp:298-367), keyed per history by
 * the reference's own RN_init_particle skip-ahead (Random.cpp:196-204).
 */
#ifndef MC_ORACLE_H
#define MC_ORACLE_H

#include <stdint.h>
#include "../include/mc_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* --- Random.cpp:92-204 ---------------------------------------------------- */
uint64_t mco_lcg_next(uint64_t s);
double   mco_urand(uint64_t* s);
uint64_t mco_skip_ahead(uint64_t seed, int64_t nskip);
uint64_t mco_init_particle(uint64_t seed0, uint64_t nps);

/* --- Algorithm.cpp ---------------------------------------------------------- */
int    mco_binary_search(double x, const double* v, int n);
double mco_interpolate(double x, double x1, double x2, double y1, double y2);
double mco_geometry_quad(double a, double b, double c);
void   mco_scatter_direction(const double dir[3], double mu0, uint64_t* rng, double out[3]);
double mco_log(double x);
void   mco_sincos_2pi(double xi, double* c, double* s);

/* --- Reaction.cpp:27-118 ------------------------------------------------------ */
/* free-gas elastic scatter at 293.6 K: updates *E (eV) and the direction u */
void   mco_free_gas_scatter(double A, double* E, double u[3], uint64_t* rng);
/* elastic scatter off a target at rest (the T -> 0 limit): updates *E and u */
void   mco_elastic_target_at_rest(double A, double* E, double u[3], uint64_t* rng);

/* --- Distribution.cpp:78-92 ------------------------------------------------------ */
/* isotropic direction: u[0] = 2 xi1 - 1, azimuth 2 pi xi2 around x           */
void   mco_isotropic_direction(uint64_t* rng, double u[3]);

/* Split copy i (from 0) of a particle starts (i + 1) << MCO_COPY_JUMP_LOG2
 * draws ahead of the particle's stream (population_control.cpp:21-49 pushes
 * copies; the per-history streams need a place for theirs).                  */
#define MCO_COPY_JUMP_LOG2 50

/* --- Geometry.cpp ------------------------------------------------------------ */
double mco_surface_eval(const mc_surface_desc* S, const double p[3]);
double mco_surface_distance(const mc_surface_desc* S, const double p[3], const double u[3]);
void   mco_surface_reflect(const mc_surface_desc* S, double u[3]);

/* --- Material.cpp / Nuclide.cpp / XSec.cpp ------------------------------------ */
enum { MCO_XS_TOTAL = 0, MCO_XS_SCATTER = 1, MCO_XS_CAPTURE = 2 };
double mco_nuclide_xs(const mc_problem_desc* d, int nuc, int which, double E);
double mco_material_xs(const mc_problem_desc* d, const double* theta, int mat, int which, double E);

/* --- whole histories ----------------------------------------------------------
 * Runs histories [0, n_hist) of problem d with RNG keys key_offset + h.
 * sum/sq (n_bins) accumulate the history-score moments (end_history).
 * trace (optional) receives the first trace_count histories.
 * grad_out (optional, n_bins) switches on the likelihood-ratio gradient:
 * gsum/gsq (n_params) accumulate the per-history vector-Jacobian product.
 * Returns MC_OK or an mc_status error.                                       */
/* The particle comb's selection (population_control.cpp:55-84) over bank
 * weights w[0..n) with first tooth u * W / teeth: output entry k is a copy of
 * entry src[k] with weight wout[k]; fill[k] > 0 marks the k-th entry past the
 * last selected one (a copy of entry 0 at its own weight, whose stream the
 * comb jumps fill[k] 2^50 draws ahead).  Returns the number selected.       */
int mco_comb_plan(const double* w, int n, int teeth, double u, int* src, double* wout, int* fill);

int mco_run(const mc_problem_desc* d, const double* theta, uint64_t key_offset,
            uint64_t n_hist, double* sum, double* sq, uint64_t* flights,
            mc_trace_record* trace, uint64_t trace_count,
            const double* grad_out, double* gsum, double* gsq);

#ifdef __cplusplus
}
#endif
#endif
