/*
 * mc_api.h — C ABI of the MI355X fixed-source Monte Carlo transport engine.
 *
 * Plain C, POD structs and pointers only: no C++ or torch types cross this
 * boundary, no exception escapes it, every call returns an mc_status and
 * mc_last_error() describes the most recent failure on the calling thread.
 *
 * What each entry point replaces in the reference (ilhamv/MC-old):
 *   mc_problem_from_xml        Simulator::Simulator(io_dir)      src/simulator/setup.cpp:30-1069
 *   mc_problem_from_xml_string   (same parser, deck held in memory)
 *   mc_problem_create          the same object graph built by a caller from
 *                              flat arrays instead of a deck
 *   mc_problem_desc            the object graph the constructor builds
 *                              (Surfaces/Cells/Materials/Nuclides/Estimators,
 *                              include/simulator.h:118-122)
 *   mc_engine_create           (no counterpart: device upload of one or many problems)
 *   mc_forward                 Simulator::start()                src/simulator/handler.cpp:11-48
 *                              with random_walk                  src/simulator/general.cpp:177-211
 *                              and the estimator closeouts       src/Estimator.cpp:339-367
 *   mc_backward                (no counterpart: likelihood-ratio gradient of the tallies)
 *   mc_xs_lookup               Material::SigmaT(E)               src/Material.cpp:50-57
 *   mc_autotune                (no counterpart: launch-shape search for the batched loop)
 *   mc_report_h5               Simulator::report(io_dir)         src/simulator/report.cpp:9-50
 *   mc_report_text               (the same content as text, where HDF5 is absent)
 *
 * Ownership: the library owns mc_problem / mc_engine handles; the caller owns
 * every output buffer.  An engine copies what it needs from its problems when
 * it is created: a problem may be destroyed once the engines built from it
 * exist.  Threading: one handle per host thread; a handle is not re-entrant;
 * any number of engines per device.  All floating-point tallies are FP64.
 */
#ifndef MC_API_H
#define MC_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
typedef enum mc_status {
    MC_OK = 0,
    MC_ERR_ARGUMENT = 1,      /* null or inconsistent argument                      */
    MC_ERR_INPUT = 2,         /* malformed deck / unknown element (setup.cpp exits)  */
    MC_ERR_IO = 3,            /* unreadable deck or cross-section file              */
    MC_ERR_UNSUPPORTED = 4,   /* valid reference input outside this engine's scope  */
    MC_ERR_DEVICE = 5,        /* HIP runtime failure                                 */
    MC_ERR_LOST_PARTICLE = 6, /* point outside every cell (general.cpp:31-33)        */
    MC_ERR_CAPACITY = 7       /* a fixed-size limit was exceeded                     */
} mc_status;

const char* mc_last_error(void);
const char* mc_version(void);

/* --------------------------------------------------------- flat problem view
 * The loader turns an XML deck (reference schema) into these flat arrays.  IDs
 * are indices in deck order, as in the reference (setup.cpp:527,619).      */

enum { /* surface types, Geometry.h:298-420 */
    MC_SURF_PLANE_X = 1, MC_SURF_PLANE_Y = 2, MC_SURF_PLANE_Z = 3,
    MC_SURF_PLANE = 4, MC_SURF_SPHERE = 5, MC_SURF_CYL_X = 6, MC_SURF_CYL_Z = 8
};
enum { MC_BC_VACUUM = -1, MC_BC_TRANSMISSION = 0, MC_BC_REFLECTIVE = 1 };
enum { MC_NUC_CONSTANT = 1, MC_NUC_TABLE = 2 };
/* elastic-scatter kinematics of a nuclide of finite mass (deck attribute
 * kinematics="free_gas" (default: the reference's, Reaction.cpp:27-118) or
 * "target_at_rest", its T -> 0 limit)                                       */
enum { MC_KIN_TARGET_AT_REST = 0, MC_KIN_FREE_GAS = 1 };
enum { MC_DIR_ISOTROPIC = 1, MC_DIR_DELTA = 2 };
enum { MC_ENERGY_DELTA = 1, MC_ENERGY_UNIFORM = 2 };
enum { MC_ATTACH_SURFACE = 1, MC_ATTACH_CELL = 2 };
enum { /* score kernels, Estimator.cpp:17-41 */
    MC_KERNEL_TRACK_LENGTH = 1, MC_KERNEL_COLLISION = 2, MC_KERNEL_NEUTRON = 3
};
enum { /* scores, Estimator.cpp:48-124; "cross" = FLUX with the NEUTRON kernel */
    MC_SCORE_FLUX = 1, MC_SCORE_ABSORPTION = 2, MC_SCORE_SCATTER = 3,
    MC_SCORE_CAPTURE = 4, MC_SCORE_TOTAL = 5
};
enum { MC_REACTION_SCATTER = 1, MC_REACTION_CAPTURE = 2 };   /* bit mask */

#define MC_MAX_SCORES 8
#define MC_MASS_INFINITE 1.0e300  /* A >= this: infinitely heavy target (one-speed) */

typedef struct mc_surface_desc {
    int32_t type, bc;
    int32_t est_begin, est_count;   /* attached surface estimators (att_* arrays) */
    /* plane_x/y/z: c[0] = position; plane: a,b,c,d and c[4..6] = 2a/|n|^2 ...
     * (SurfacePlane ctor, Geometry.cpp:14-23); sphere: x0,y0,z0,r, c[4] = r*r;
     * cylinder_x: y0,z0,r, c[4] = r*r; cylinder_z: x0,y0,r, c[4] = r*r        */
    double c[8];
} mc_surface_desc;

typedef struct mc_cell_desc {
    int32_t material;               /* -1: void                                  */
    int32_t surf_begin, surf_count; /* cell_surf / cell_sense                    */
    int32_t tl_begin, tl_count;     /* attached track-length estimators (att_*)  */
    int32_t col_begin, col_count;   /* attached collision estimators (att_*)     */
    int32_t pad;
    double importance;
} mc_cell_desc;

typedef struct mc_material_desc {
    int32_t nuc_begin, nuc_count;   /* mat_nuclide / mat_density                 */
    int32_t param_begin, param_count; /* gradient parameters acting on it        */
} mc_material_desc;

typedef struct mc_nuclide_desc {
    double A;                       /* target mass ratio; >= MC_MASS_INFINITE: one-speed */
    int32_t kind;                   /* MC_NUC_CONSTANT / MC_NUC_TABLE            */
    int32_t grid_begin, grid_count; /* xs_energy / xs_scatter / xs_capture / xs_total */
    int32_t kinematics;             /* elastic scatter: MC_KIN_TARGET_AT_REST or
                                       MC_KIN_FREE_GAS (Reaction.cpp:27-118, 293.6 K)  */
    double xs_scatter, xs_capture, xs_total;   /* MC_NUC_CONSTANT values (barn) */
} mc_nuclide_desc;

typedef struct mc_source_desc {
    double pos[3];
    double dir[3];                  /* MC_DIR_DELTA                              */
    double energy[2];               /* delta: [0]; uniform: a, b                 */
    double probability;             /* parsed, ignored as in Source.cpp:65-72    */
    int32_t cell;                   /* search_cell(pos) at setup                 */
    int32_t dir_kind, energy_kind, pad;
} mc_source_desc;

typedef struct mc_estimator_desc {
    int32_t attach_kind;            /* MC_ATTACH_SURFACE / MC_ATTACH_CELL        */
    int32_t n_attach;               /* size of the surface/cell filter           */
    int32_t n_scores;
    int32_t score[MC_MAX_SCORES];
    int32_t score_kernel[MC_MAX_SCORES];
    int32_t egrid_begin, egrid_count; /* energy filter grid in egrid[] (0: none) */
    int32_t bin_offset, n_bins;     /* layout [score][attach][energy]            */
} mc_estimator_desc;

typedef struct mc_param_desc {      /* d(tally)/d(theta): theta scales reaction(s)  */
    int32_t cell;                   /* deck cell the parameter was declared on   */
    int32_t material;               /* material instance (cloned per cell)       */
    int32_t nuc_slot;               /* position in the material's nuclide list   */
    int32_t reactions;              /* MC_REACTION_* mask                        */
} mc_param_desc;

typedef struct mc_problem_desc {
    int32_t n_surfaces, n_cells, n_materials, n_nuclides;
    int32_t n_sources, n_estimators, n_params, n_bins;
    const mc_surface_desc*   surfaces;
    const mc_cell_desc*      cells;
    const int32_t*           cell_surf;     /* surface index          */
    const int32_t*           cell_sense;    /* +1 / -1                */
    const mc_material_desc*  materials;
    const int32_t*           mat_nuclide;
    const double*            mat_density;   /* atoms / (barn cm)      */
    const mc_nuclide_desc*   nuclides;
    const double*            xs_energy;     /* eV                     */
    const double*            xs_scatter;    /* barn                   */
    const double*            xs_capture;
    const double*            xs_total;      /* (s + c) + f per point, setup.cpp:372 */
    const mc_source_desc*    sources;
    const double*            source_cdf;    /* n_sources+1, Source.cpp:65-72 */
    const mc_estimator_desc* estimators;
    const int32_t*           att_estimator; /* estimator index        */
    const int32_t*           att_index;     /* surface/cell filter bin, Estimator.cpp:140,156 */
    const double*            egrid;
    const mc_param_desc*     params;
    const int32_t*           mat_param;     /* material -> parameter list */
    double   wr, ws;                        /* weight roulette, simulator.h:90-91 */
    uint64_t n_histories;                   /* <description samples=...>          */
    uint64_t seed;                          /* LCG seed (default 1)               */
    /* particle comb, <population_control><particle_comb bank_max teeth>
     * (setup.cpp:57-67, population_control.cpp:55-84): after each particle's
     * walk, a history whose bank of split copies holds bank_max or more is
     * combed to comb_teeth copies.  0, 0: no comb.                         */
    int32_t  comb_bank_max, comb_teeth;
} mc_problem_desc;

typedef struct mc_problem mc_problem;

/* Reads <io_dir>/input.xml; ZAID nuclides read <xs_dir>/<ZAID>.txt (NULL:
 * "./xs_library", setup.cpp:326).                                            */
mc_status mc_problem_from_xml(const char* io_dir, const char* xs_dir, mc_problem** out);
mc_status mc_problem_from_xml_string(const char* xml, const char* xs_dir, mc_problem** out);
/* Builds a problem from caller-owned flat arrays (copied; the caller keeps
 * its arrays).  Array lengths follow from the counts and index ranges of the
 * records: cell_surf/cell_sense up to the largest surf_begin + surf_count,
 * mat_nuclide/mat_density up to the largest nuc_begin + nuc_count, xs_* up to
 * the largest grid_begin + grid_count, att_* up to the largest attachment
 * range, egrid up to the largest egrid_begin + egrid_count, mat_param up to
 * the largest param_begin + param_count.  source_cdf may be NULL (equal
 * probabilities, Source.cpp:65-72); a source's cell < 0 is found by the
 * cell search (general.cpp:26-34).  Every index is checked: MC_ERR_INPUT
 * names the first that is out of range.  Names are "surface <i>", "cell
 * <i>", "material <i>", "nuclide <i>" and "Estimator #<i+1>".
 * Replaces building the object graph by Simulator::Simulator (setup.cpp:30-1069). */
mc_status mc_problem_create(const mc_problem_desc* desc, mc_problem** out);
mc_status mc_problem_get_desc(const mc_problem* p, const mc_problem_desc** out);
/* Estimator name / score name lookups for reports (NULL when out of range). */
const char* mc_problem_estimator_name(const mc_problem* p, int32_t e);
/* Deck name of surface, cell, material (a per-cell copy made for gradient
 * parameters is "<material>@<cell>"), nuclide or estimator i (NULL when out
 * of range): the names the reference's objects carry (Geometry::name,
 * Material::name, Nuclide::name, Estimator::name).                          */
enum { MC_OBJ_SURFACE = 1, MC_OBJ_CELL = 2, MC_OBJ_MATERIAL = 3, MC_OBJ_NUCLIDE = 4, MC_OBJ_ESTIMATOR = 5 };
const char* mc_problem_object_name(const mc_problem* p, int32_t kind, int32_t i);
void mc_problem_destroy(mc_problem* p);

/* ------------------------------------------------------------------- engine */
typedef struct mc_engine mc_engine;

typedef struct mc_engine_cfg {
    int32_t device;          /* HIP device ordinal                              */
    int32_t block;           /* workgroup size (0: 256)                         */
    int32_t grid_per_cu;     /* persistent workgroups per CU (0: 4)             */
    int32_t iters_per_graph; /* event iterations per captured graph (0: 16)     */
    uint64_t bank_capacity;  /* particle slots (0: 1<<21)                       */
    int32_t use_graph;       /* 1: replay hipGraphs, 0: plain launches          */
    int32_t profile;         /* 1: plain launches timed per kernel (mc_engine_kernel_stats) */
    int32_t drain_divisor;   /* once every history has started and at most
                                bank/drain_divisor particles live, finish them
                                in one drain launch (0: 4, < 0: never)        */
    int32_t events_per_launch; /* forward runs: flights, each with its
                                collision, a particle takes per event iteration
                                before the bank is refilled (0: 16; 1: one);
                                runs that can split take one                  */
    int32_t grad_events_per_launch; /* the same for gradient runs (0: 24; 1: one,
                                with a collision kernel of its own)           */
    int32_t grad_live_divisor; /* gradient runs keep at most
                                bank_capacity / grad_live_divisor particles
                                live (0: 4): the score-function rows of fewer
                                slots stay in cache                           */
} mc_engine_cfg;

/* Per-kernel device time of the event loop, recorded when cfg.profile = 1:
 * kernels in order refill_count, refill_scatter, xs_lookup, flight (with the
 * surface crossings; in forward runs also the collisions and the further
 * flights of each particle), collision (gradient runs at one flight per
 * iteration), drain.  Items of the flight kernel are flights.               */
#define MC_N_KERNELS 6
typedef struct mc_kernel_stats {
    uint64_t launches[MC_N_KERNELS];
    uint64_t items[MC_N_KERNELS];   /* particles (slots for refill) processed  */
    double seconds[MC_N_KERNELS];
} mc_kernel_stats;

enum { MC_RUN_BATCHED = 0, MC_RUN_PLAIN = 1 };

typedef struct mc_trace_record {   /* per-history state for CPU/GPU bit-equality */
    uint64_t rng;                  /* LCG state after the history               */
    uint64_t flights;              /* move_particle calls (Ntrack)              */
    double x, y, z, E;             /* state when the history ended              */
    double score;                  /* sum over bins of the history's tally      */
} mc_trace_record;

typedef struct mc_run_cfg {
    int32_t mode;                  /* MC_RUN_BATCHED / MC_RUN_PLAIN             */
    int32_t device_ptrs;           /* 1: the mc_tally_out arrays, trace, and
                                      mc_backward's grad_out / grad / grad_sigma
                                      are device pointers on the engine's device */
    const uint64_t* n_histories;   /* per problem, NULL: each deck's samples    */
    const double* theta;           /* n_params of the whole engine, NULL: 1.0   */
    mc_trace_record* trace;        /* NULL: off                                 */
    uint64_t trace_count;          /* histories traced per problem (the first
                                      trace_count of the run's histories)       */
    /* History keys: problem p runs the n_histories[p] histories of keys
     * key_offset[p], key_offset[p] + 1, ...  A history's random-number stream
     * is fixed by its key alone (RN_init_particle, Random.cpp:196-204), so runs
     * over disjoint key ranges are shards of one run: their traces are the
     * full run's and their moment sums add up to its sums.  NULL: problem p
     * starts at the sum of the histories of the problems before it.          */
    const uint64_t* key_offset;
    void* stream;                  /* hipStream_t the run is issued on (NULL: the
                                      engine's own stream); the call returns once
                                      the run's work on it has finished        */
} mc_run_cfg;

typedef struct mc_tally_out {      /* caller-owned, sized engine n_bins         */
    double* sum;                   /* sum over histories of the history score  */
    double* sq;                    /* sum of squares                            */
    double* mean;                  /* end_simulation(): mean                    */
    double* uncer;                 /* end_simulation(): standard error          */
    uint64_t flights;              /* total flights (Ntrack)                    */
    double seconds;                /* device time of the event loop             */
} mc_tally_out;

typedef struct mc_engine_info {
    int32_t n_problems, n_bins, n_params, pad;
    uint64_t bank_capacity;
    int32_t block, grid_per_cu, iters_per_graph, n_cu;
} mc_engine_info;

mc_status mc_engine_create(mc_problem* const* problems, int32_t n_problems,
                           const mc_engine_cfg* cfg, mc_engine** out);
mc_status mc_engine_info_get(const mc_engine* e, mc_engine_info* out);
/* Bin offset of problem p in the engine-wide tally arrays. */
mc_status mc_engine_bin_offset(const mc_engine* e, int32_t p, int32_t* offset, int32_t* n_bins);
mc_status mc_engine_param_offset(const mc_engine* e, int32_t p, int32_t* offset, int32_t* n_params);
/* Accumulated per-kernel statistics since the last reset (reset != 0 clears
 * them after reading).  No reference counterpart: measurement entry point. */
mc_status mc_engine_kernel_stats(mc_engine* e, mc_kernel_stats* out, int32_t reset);
/* The copy ring of split copies (problems whose importances rise between
 * cells): its capacity in entries, the times it has doubled since the engine
 * was made (the capacity is 0 until a run allocates the particle bank), and
 * the last run's waits (splits, and later turns, that found it
 * without room for a copy; the particle then waits at the surface, before
 * its roulette, and places the rest of its copies as room appears).  The
 * reference's bank grows without bound (population_control.cpp:21-49); no
 * counterpart: measurement entry point. */
mc_status mc_engine_copy_ring(const mc_engine* e, uint64_t* capacity, uint64_t* growths, uint64_t* waits);

mc_status mc_forward(mc_engine* e, const mc_run_cfg* run, mc_tally_out* out);

/* Vector-Jacobian product of the tallies w.r.t. theta by likelihood-ratio
 * scores: grad[k] = d/dtheta_k sum_b grad_out[b] * tally_mean[b].
 * grad_sigma receives the standard error of each component.                */
mc_status mc_backward(mc_engine* e, const mc_run_cfg* run, const double* grad_out,
                      double* grad, double* grad_sigma, mc_tally_out* out);

/* Standalone cross-section lookup over host arrays: sigt[i] = Sigma_t of
 * material mat[i] at energy[i] for problem 0 (bit-identical to the event loop). */
mc_status mc_xs_lookup(mc_engine* e, const double* energy, const int32_t* mat,
                       double* sigt, uint64_t n);

/* mc_xs_lookup without any device allocation, for callers that cannot
 * allocate (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The device staging of
 * the inputs and outputs is a fixed array in the library, one per device, held
 * until the process ends: n <= 524288 (2^19) lookups per call, larger n is
 * rejected with MC_ERR_CAPACITY.  Calls that overlap on one device share that
 * array and race; calls on different devices do not.  Same results as
 * mc_xs_lookup, bit for bit.                                                  */
mc_status mc_xs_lookup_noalloc(mc_engine* e, const double* energy, const int32_t* mat, double* sigt,
                               uint64_t n);

/* Times the lookup kernel alone on n device-resident inputs (E log-uniform
 * over the tabulated range, material uniform over problem 0's materials):
 * average seconds per launch over `iters` launches; checksum of the outputs.
 * No reference counterpart: measurement entry point.                       */
mc_status mc_xs_lookup_bench(mc_engine* e, uint64_t n, int32_t iters, uint64_t seed,
                             double* seconds, double* checksum);

/* The random-gather ceiling of mc_xs_lookup_bench: the same input and
 * output streams on the same n synthetic inputs, and per lookup
 * `dependent_reads` (1 or 2) dependent reads of a 128-byte line at a hashed
 * position in a table of table_bytes (0: the size of what the lookup reads,
 * returned in *lookup_table_bytes), no search and no arithmetic: average
 * seconds per launch.  No reference counterpart: measurement entry point.   */
mc_status mc_gather_bench(mc_engine* e, uint64_t n, int32_t iters, uint64_t seed, uint64_t table_bytes,
                          int32_t dependent_reads, double* seconds, uint64_t* lookup_table_bytes);

/* Evaluates the engine's log, cos/sin(2 pi x) and sqrt on the device:
 * out[4i..4i+3] for in[i].  No reference counterpart: CPU/GPU equality test. */
mc_status mc_selftest_math(const double* in, double* out, uint64_t n);

typedef struct mc_tune_result {
    int32_t block, grid_per_cu, iters_per_graph, drain_divisor;
    uint64_t bank_capacity;
    double histories_per_second;
    int32_t events_per_launch, pad;
} mc_tune_result;
/* Times short batched runs over a grid of launch shapes (workgroup size,
 * workgroups per CU, bank capacity, drain threshold, flights per particle per
 * iteration) on a sample of the
 * engine's histories, every problem keeping its share, and installs the
 * fastest.  No reference counterpart (the reference has no launch shapes). */
mc_status mc_autotune(mc_engine* e, uint64_t sample_histories, mc_tune_result* out);

void mc_engine_destroy(mc_engine* e);

/* output.h5 of Simulator::report (report.cpp:9-50, Estimator.cpp:368-422):
 * /summary/{Ncycle,Nsample,Npassive,Ntrack,mode}, /summary/survival_roulette/
 * {wr,ws}, and per estimator /<name>/ (attribute "indexing", one dataset per
 * filter grid with attribute "unit", /<name>/<score>/{mean,uncertainty} shaped
 * by the filter sizes).  Written with the HDF5 C library loaded at run time
 * ($MC_HDF5_LIBRARY, libhdf5.so, /opt/conda/lib/libhdf5.so): MC_ERR_UNSUPPORTED
 * when none loads (mc_report_text writes the same content as text).          */
mc_status mc_report_h5(const mc_problem* p, const mc_tally_out* t, uint64_t n_histories, const char* path);

/* Text report in the output.h5 layout of report.cpp:9-50 / Estimator.cpp:368-422. */
mc_status mc_report_text(const mc_problem* p, const mc_tally_out* t, uint64_t n_histories,
                         const char* path);

#ifdef __cplusplus
}
#endif
#endif /* MC_API_H */
