/*
 * cuhre.h -- C-ABI of the MI355X adaptive cubature engine.
 *
 * Plain pointers and sizes only.  All functions return 0 on success, a
 * positive "not converged" flag where stated, or a negative error code
 * (message available from cuhre_last_error()).  No function calls exit().
 *
 * Reference interface replaced (paths relative to the reference tree):
 *   quad::GPUcuhre<T>::GPUcuhre(argc, argv, dim, key, verbose, numDevices)
 *       quad/GPUquad/GPUquad.h:46-63           -> cuhre_context_create
 *   quad::GPUcuhre<T>::integrate(epsrel, epsabs, integral&, error&,
 *       nregions&, neval&, integT* f, Volume* vol)
 *       quad/GPUquad/GPUquad.h:611-727         -> cuhre_integrate
 *   IntegralGPU::Cuhre typedef, CuhreInteface.h:9-12 (the facade that
 *       include/gpucuhre.hpp mirrors on top of this ABI)
 *   INTEGRATE_GPU_PHASE1 / SampleRegionBlock (per-region rule values)
 *       quad/GPUquad/GPUQuadPhases.cu:37-79,
 *       quad/GPUquad/GPUQuadSample.cu:99-220   -> cuhre_eval_regions
 *   alignRegions + divideIntervalsGPU (compaction + bisection)
 *       quad/GPUquad/GPUKernelquad.cu:63-149   -> cuhre_split_regions
 *   Volume (integration box), quad/GPUquad/GPUQuadVolume.cu:4-25
 *       -> the lows/highs arguments (FP64 instead of float textures).
 */
#ifndef CUHRE_H
#define CUHRE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUHRE_MAXDIM 16
#define CUHRE_MAXTERMS 8

/* error codes */
#define CUHRE_OK 0
#define CUHRE_NOT_CONVERGED 1
#define CUHRE_ERR_ARG (-1)
#define CUHRE_ERR_NONFINITE (-2)
#define CUHRE_ERR_ALLOC (-3)
#define CUHRE_ERR_DEVICE (-4)
#define CUHRE_ERR_UNSUPPORTED (-5)
#define CUHRE_ERR_CAPACITY (-6)

/* integrand families compiled into the library */
#define CUHRE_GAUSSIAN 1     /* exp(-sum a_i^2 (x_i-w_i)^2)                    */
#define CUHRE_PRODUCT_PEAK 2 /* prod 1/(a_i^-2 + (x_i-w_i)^2)                  */
#define CUHRE_OSCILLATORY 3  /* cos(2 pi u + sum a_i x_i)                      */
#define CUHRE_CORNER_PEAK 4  /* (1 + sum a_i x_i)^-exponent                    */
#define CUHRE_CONTINUOUS 5   /* exp(-sum a_i |x_i - w_i|)                      */
#define CUHRE_POLYNOMIAL 6   /* sum_t coef_t prod_i x_i^powers[t][i]           */
#define CUHRE_COSINE_SUM 7   /* -(1/2) sum_i cos(10 x_i)/0.054402111088937     */
#define CUHRE_TABLE6 8       /* B(x1,x2) x3^-1.5 exp(-(x4-.5)^2/.02)(1+x5 x6)  */
#define CUHRE_LC_LT_T 9      /* the reference's LC_LT_t in x = (lt, zt, lc)     */

/* Integrand description (the C-ABI stand-in for the reference's functor
 * with __device__ double operator()(double x[], int dim)).  The table of
 * CUHRE_TABLE6 is a host pointer to nx*ny doubles, table[i*ny + j] =
 * B(i/(nx-1), j/(ny-1)); the library copies it to the device.
 *
 * CUHRE_LC_LT_T (dim 3, lc_lt_t.h:20-50 with GPUQuadInterp2D.cu:49-100):
 * table points to nx + ny + 5*nx*ny doubles: the lt nodes (nx, the
 * interpolation rows), the zt nodes (ny, the columns), then the fields
 * tau, mu, sigma, fmsk, fprj, each field[j*nx + i] = value at (lt_i, zt_j)
 * (the layout of the reference's tau_arr ...).  table_nx = nx, table_ny =
 * ny; the nodes must increase strictly.  Every value is rounded to float,
 * as the reference's tables are float textures, and the fields are
 * interpolated bilinearly in (lt, zt) in the reference's float/double
 * arithmetic.  Beyond the last node the last interval is extended (the
 * reference reads a texel past the end there); below the first node the
 * first interval is extended, as in the reference. */
typedef struct {
  int32_t family;
  int32_t dim;
  double a[CUHRE_MAXDIM];
  double w[CUHRE_MAXDIM];
  double u;
  int32_t exponent;
  int32_t nterms;
  double coef[CUHRE_MAXTERMS];
  int32_t powers[CUHRE_MAXTERMS][CUHRE_MAXDIM];
  const double* table;
  int32_t table_nx, table_ny;
} cuhre_integrand;

/* Result of one integration (the reference returns integral, error,
 * nregions, neval by reference plus an int flag).  fail: 0 converged,
 * 1 not converged (maxeval, capacity, iteration cap or no active region
 * left while the global test still fails). */
typedef struct {
  double integral;
  double error;          /* Genz-Malik error estimate sum            */
  double fp64_err_bound; /* first-order FP64 rounding bound          */
  int64_t nregions;      /* regions evaluated (Phase-1 meaning)      */
  int64_t neval;         /* nregions * points per region             */
  int32_t fail;
  int32_t iterations;
  int64_t peak_regions;  /* largest list evaluated in one iteration  */
  int64_t nfinished;     /* regions retired as finished              */
  double ms_total;       /* host wall time inside the call           */
  double ms_eval;        /* device time of rule-evaluation launches  */
  double ms_split;       /* device time of classify/split launches   */
  int32_t eval_launches;
  int32_t split_launches;
  double eval_flops;     /* algorithmic FP64 flops of the eval launches */
  double ms_eval_largest;  /* device time of the largest eval launch    */
  int32_t nzeroed;       /* two-phase: regions whose error passed 1e10 and
                            were zeroed (the reference's rule); the call
                            then returns 1 even when the sums pass       */
  int32_t pad1;
  int64_t largest_regions; /* regions of that launch                   */
} cuhre_result;

typedef struct cuhre_context cuhre_context;

/* Creates an engine on HIP device `device`.  capacity_regions bounds the
 * region list (0 = what three quarters of the free HBM holds); the list
 * buffers start small and grow on demand. */
int cuhre_context_create(cuhre_context** ctx, int device, int64_t capacity_regions);
int cuhre_context_destroy(cuhre_context* ctx);

/* Engine knobs.  profile != 0 records HIP events around every launch
 * (fills ms_eval/ms_split).  point_segments: 0 automatic, else 1, 2, 4, 8
 * or 16 wavefronts share the points of each group of 64 regions (one region
 * per lane; more segments give short lists more parallelism), or 32: the
 * 32 lanes of a half wavefront share the points of one region (the
 * automatic choice for the shortest lists). */
int cuhre_context_set(cuhre_context* ctx, int profile, int point_segments);

/* Plain-loop scheduling (no counterpart in the reference, whose host loop
 * reads every iteration's sums before the next launch).  on != 0, the
 * default: each iteration's evaluation and split enter the stream before
 * the host has read the previous iteration's state, their grids sized for
 * the largest list that iteration can have and the real length read on the
 * device, so the GPU does not wait on the host between iterations; 0: one
 * iteration at a time.  Results are identical either way.  Profiling
 * (cuhre_context_set) runs one iteration at a time. */
int cuhre_context_queue_ahead(cuhre_context* ctx, int on);

/* Plain adaptive integration over the box [lows, highs].  key selects the
 * rule: 0 or 7 = Genz-Malik degree 7/5 (err = |I7 - I5|); 9 = the
 * reference's degree-9 rule with four null rules (quad/GPUquad/GPUQuadRule.cu:30-226,
 * error as quad/GPUquad/GPUQuadSample.cu:195-216).  Returns the fail flag
 * (0/1) or a negative error code. */
int cuhre_integrate(cuhre_context* ctx, const cuhre_integrand* f, int dim,
                    double epsrel, double epsabs, int64_t maxeval, int key,
                    const double* lows, const double* highs, cuhre_result* out);

/* cuhre_integrate without any allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as cuhre_integrate (bit for bit), with the
 * region lists in a static 64 MiB device array per GPU (taken once per
 * device, held until the process ends) instead of the context's buffers;
 * the context must exist already (its creation allocates).  A run whose
 * list would outgrow the array -- cuhre_noalloc_capacity(dim) regions,
 * about 233k at dim 8 and 540k at dim 3 -- stops with CUHRE_ERR_CAPACITY
 * (out holds the partial result).  Tables of TABLE6 / LC_LT_T must fit in
 * 1 MiB.  Calls on one GPU share the array: they are serialised by a lock
 * inside one process; calls that overlap on one GPU otherwise race. */
int cuhre_integrate_noalloc(cuhre_context* ctx, const cuhre_integrand* f, int dim,
                            double epsrel, double epsabs, int64_t maxeval, int key,
                            const double* lows, const double* highs, cuhre_result* out);
/* Regions per list the static array holds at dim (-1 outside 2..10). */
int64_t cuhre_noalloc_capacity(int dim);

/* The reference's two-phase method (quad/GPUquad/GPUquad.h:611-727):
 * Phase 1 is the breadth-first loop above, at most 100 iterations, until
 * the global test holds, no region is active, or a split leaves
 * phase1_regions regions or more (0: 2^14, FIRST_PHASE_MAXREGIONS); Phase 2
 * then refines each of those regions on its own
 * (quad/GPUquad/GPUQuadPhases.cu:422-536): a max-error heap, at most
 * pool_max bisections (0: 2048, MAX_GLOBALPOOL_SIZE), until the region's
 * own sum meets max(epsrel |sum|, epsabs), children's errors refined from
 * the parent difference (quad/GPUquad/GPUQuadPhases.cu:498-515).  Returns
 * 0 when the global test holds, else 1 with out->fail = the number of
 * Phase-2 regions that missed their own test (the reference's flag).
 * pool_max is at most 2^24.  The Phase-2 workspace ((2 dim + 2) doubles
 * + 5 bytes per slot and region) is kept by the context between calls. */
int cuhre_integrate_two_phase(cuhre_context* ctx, const cuhre_integrand* f, int dim,
                              double epsrel, double epsabs, int key,
                              int64_t phase1_regions, int32_t pool_max,
                              const double* lows, const double* highs,
                              cuhre_result* out);

/* Rule evaluation of n regions given on the host (SoA: lower[m*n + r],
 * width[m*n + r]).  Outputs per region: est, err, axis, abssum (volume *
 * sum |w f| of the basic rule); i5 (degree-5 estimate; 0 under key 9) and
 * dvec (d*n, SoA) may be NULL. */
int cuhre_eval_regions(cuhre_context* ctx, const cuhre_integrand* f, int dim,
                       int key, int64_t n, const double* lower,
                       const double* width, double* est, double* err,
                       int32_t* axis, double* abssum, double* i5, double* dvec);

/* Phase 2 of the two-phase method (quad/GPUquad/GPUQuadPhases.cu:422-536) on
 * n regions given on the host (SoA as cuhre_eval_regions): per region the
 * refined result, error and pool size (1 + bisections).  key 0/7 or 9;
 * pool_max 0: 2048. */
int cuhre_phase2_regions(cuhre_context* ctx, const cuhre_integrand* f, int dim, int key, int64_t n,
                         const double* lower, const double* width, double epsrel, double epsabs,
                         int32_t pool_max, double* res, double* err, int32_t* count);

/* One classify + compact + bisect step on a host list: regions with
 * err > max(epsrel |est|, epsabs) are split on axis[r]; child 0 of the j-th
 * such region lands at 2j, child 1 at 2j+1 (SoA with stride 2A).  out_*
 * must hold 2n*dim doubles.  *n_active receives A. */
int cuhre_split_regions(cuhre_context* ctx, int dim, int64_t n,
                        const double* lower, const double* width,
                        const double* est, const double* err,
                        const int32_t* axis, double epsrel, double epsabs,
                        double* out_lower, double* out_width,
                        int64_t* n_active);

/* Deterministic FP64 sum of n values (the reduction the engine uses). */
int cuhre_sum(cuhre_context* ctx, int64_t n, const double* x, double* out);

/* ---------------------------------------------------------------- batches
 * Many independent integrals of one dimension in one call (no reference
 * counterpart: the reference integrates one functor per call).  Each
 * instance runs the plain method with its own tolerances and budget; all
 * share one region list on the device and the iterations are replayed from
 * a hipGraph.  Families: GAUSSIAN, PRODUCT_PEAK, OSCILLATORY, CORNER_PEAK,
 * CONTINUOUS. */
typedef struct {
  int32_t point_segments;       /* 0: from cuhre_batch_autotune, else 1..16   */
  int32_t grid_blocks;          /* 0: from cuhre_batch_autotune               */
  int32_t iterations_per_check; /* graph replays between host checks (0: 2)   */
  int32_t use_graph;            /* 1: hipGraph replay, 0: direct launches     */
  int32_t persistent;           /* 1: first one persistent launch in which each
                                   workgroup takes instances from a queue and
                                   runs each one's whole loop in its own region
                                   queue (2 x 16384 regions); an instance that
                                   outgrows its queue is handed over, with its
                                   list and sums, to the replayed iterations.
                                   0: replayed iterations only.  Same results
                                   per instance either way (sums to rounding). */
  int64_t initial_capacity;     /* first list allocation in regions (0: 4 Mi);
                                   the list grows on demand                   */
  double finish_budget;         /* finish policy per instance: 0 the plain
                                   per-region test err <= max(epsrel |est|,
                                   epsabs) (results equal plain calls); b in
                                   (0, 1) the shared allotment err <= max(epsrel
                                   ((1 - b) |est| + b |I| vrel), epsabs), I the
                                   instance's estimate in the iteration, vrel the
                                   region's share of its box.  The global test
                                   sum err <= max(epsrel |I|, epsabs) is the
                                   only convergence test either way.          */
} cuhre_batch_config;

typedef struct {
  int32_t iterations;           /* largest per-instance iteration count       */
  int32_t launched_iterations;  /* iterations replayed (>= iterations)        */
  int64_t total_regions;
  int64_t total_neval;
  int32_t not_converged;
  int32_t graph_used;
  int32_t point_segments;
  int32_t grid_blocks;
  double ms_total;
  int64_t peak_list;            /* longest region list of the run            */
  int64_t capacity;             /* regions per list buffer at the end        */
  int32_t grows;                /* list growths during the run               */
  int32_t persistent_finished;  /* instances the persistent launch finished
                                   (-1: no persistent launch)                 */
} cuhre_batch_info;

/* count instances; fs[count]; epsrel/epsabs/maxeval per instance (NULL:
 * 1e-3 / 0 / 1e9); lows/highs count*dim (NULL: unit cube); out[count].
 * Returns 0 when every instance converged, 1 if some did not, <0 on error.
 * out[i].fail: 0 converged, 1 not converged, CUHRE_ERR_NONFINITE (-2) when
 * instance i's integrand returned a non-finite value; the call then returns
 * -2 with every other instance's result complete (and counted in
 * info->not_converged when it did not converge). */
int cuhre_integrate_batch(cuhre_context* ctx, int32_t count, const cuhre_integrand* fs, int dim,
                          const double* epsrel, const double* epsabs, const int64_t* maxeval,
                          const double* lows, const double* highs,
                          const cuhre_batch_config* cfg, cuhre_result* out,
                          cuhre_batch_info* info);

/* Times the batch evaluation kernel over point_segments {1,2,4} x grid
 * blocks {256..4096} on a uniformly refined sample of the batch's own list
 * and stores the fastest shape (also returned in *best; table_ms, if not
 * NULL, receives the 15 timings in that order). */
int cuhre_batch_autotune(cuhre_context* ctx, int32_t count, const cuhre_integrand* fs, int dim,
                         const double* lows, const double* highs, cuhre_batch_config* best,
                         double* table_ms);

/* ------------------------------------------------------ learned split guide
 * A BF16 MLP (32-64-64-64-16, FP32 accumulation on MFMA) trained on the GPU
 * by the library's own forward/backward kernels.  It proposes split axes
 * and early finishes; the Genz-Malik error sum remains the only
 * convergence test.  No reference counterpart. */
#define CUHRE_GUIDE_PARAMS 11472 /* W0 32x64, W1 64x64, W2 64x64, W3 64x16, b0..b3 */

typedef struct cuhre_guide cuhre_guide;

typedef struct {
  int32_t axis_mode;        /* 0: fourth difference; 1: guide when confident   */
  int32_t finish_mode;      /* 0: none; 1: guide proposal + volume-share guard;
                               2: guard alone (no network);
                               3: shared allotment: the per-region test
                                  becomes err <= max(epsrel ((1 - budget) |est|
                                  + budget |I| vrel), epsabs), I the iteration's
                                  estimate, vrel the region's share of the box;
                               4: normalised allotment: err <= max(epsrel |I|
                                  ((1 - budget) |est| / E + budget vrel),
                                  epsabs), E = sum |est| over the finished
                                  regions and the iteration's: the tolerances
                                  add up to epsrel |I| even when the
                                  integrand's signs cancel; for f >= 0 and
                                  budget 0 it is the plain test                */
  float axis_confidence;    /* softmax probability needed to override the axis */
  float finish_threshold;   /* sigmoid output needed to propose a finish       */
  double budget;            /* modes 1, 2: share of tol retired regions may use;
                               modes 3, 4: the allotment's volume share b     */
} cuhre_guide_policy;

typedef struct {
  int64_t retired;          /* regions retired early by the finish rule        */
  int64_t samples;          /* samples logged (collection runs)                */
  int64_t overridden;       /* split axes the network changed                  */
  int32_t guide_iterations;
  int32_t pad;
} cuhre_guide_stats;

typedef struct {
  double loss_first, loss_last;      /* mean loss of the first / last epoch  */
  double axis_accuracy, finish_accuracy; /* last epoch, training samples     */
  int32_t epochs, steps;
  int64_t samples;
  double ms;
} cuhre_train_stats;

int cuhre_guide_create(cuhre_context* ctx, uint64_t seed, cuhre_guide** guide);
int cuhre_guide_destroy(cuhre_guide* guide);
int cuhre_guide_set_params(cuhre_guide* guide, const float* params);
int cuhre_guide_get_params(cuhre_guide* guide, float* params);
/* logits[n*16] of features[n*32] */
int cuhre_guide_forward(cuhre_guide* guide, int64_t n, const float* features, float* logits);
/* gradient of the mean loss over n samples (no update) */
int cuhre_guide_gradient(cuhre_guide* guide, int64_t n, const float* features, const int32_t* axis,
                         const float* finish, int dim, float* grad, double* loss);
/* run that logs up to max_samples (features, look-ahead axis, finish);
 * pol NULL: plain axes; pol->axis_mode 1: the network's confident axes
 * (only axis_mode and axis_confidence are read) */
int cuhre_guide_collect(cuhre_guide* guide, const cuhre_integrand* f, int dim, double epsrel,
                        double epsabs, int64_t maxeval, const double* lows, const double* highs,
                        int64_t max_samples, const cuhre_guide_policy* pol, cuhre_result* out);
int64_t cuhre_guide_samples(const cuhre_guide* guide);
int cuhre_guide_get_samples(cuhre_guide* guide, int64_t first, int64_t count, float* features,
                            int32_t* axis, float* finish);
/* Adam over the logged samples */
int cuhre_guide_train(cuhre_guide* guide, int dim, int epochs, int batch, float lr, uint64_t seed,
                      cuhre_train_stats* stats);
/* the plain loop with the guide between evaluation and split (pol NULL:
 * axis_mode 1, finish_mode 1, confidence 0.6, threshold 0.5, budget 0.25) */
int cuhre_integrate_guided(cuhre_context* ctx, cuhre_guide* guide, const cuhre_integrand* f, int dim,
                           double epsrel, double epsabs, int64_t maxeval, const double* lows,
                           const double* highs, const cuhre_guide_policy* pol, cuhre_result* out,
                           cuhre_guide_stats* stats);

/* Thread-local message of the last error. */
const char* cuhre_last_error(void);

/* Points per region of the rule selected by key for dimension dim. */
int64_t cuhre_points_per_region(int dim, int key);

#ifdef __cplusplus
}
#endif
#endif
