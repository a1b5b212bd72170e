/*
 * cpab_hip.h -- C ABI of the MI355X (gfx950) CPA-velocity integration library
 * (libcpab_hip.so).
 *
 * Drop-in for the reference's GPU backend object `Calcs`
 * (cpab/gpu/Calcs.py:44-1024) and its expm helper `gpu_expm`
 * (cpab/gpu/expm/expm_affine_by_series_2D.py:246-258): same kernel
 * arguments, precompiled HIP kernels instead of a pycuda JIT.
 *
 * Conventions
 *  - Every array argument is a DEVICE pointer owned by the caller (the
 *    reference writes into the caller's `out` buffers the same way).
 *  - Points are AoS, row-major (nPts, dim) float64: pts[i*dim + d]
 *    (cpab/gpu/calc_2Dto2D.cu:380-381).  Per-cell tables are (nC, dim*(dim+1))
 *    row-major float64, i.e. the top dim rows of each affine matrix
 *    (cpab/cpaNd/_PAT.py:31, _CpaCalcs.py:243).  Cell ids are int32 (nPts,).
 *  - dim is 1, 2 or 3; tess is 1 (type I: triangles / tetrahedra) or 2
 *    (type II: intervals / rectangles / boxes), the reference's TESS_TYPE
 *    (cpab/gpu/Calcs.py:161).  Type I needs dim 2 or 3.
 *  - nC0, nC1, nC2, inc_x, inc_y, inc_z are what Calcs.parse_nCs_and_incs
 *    passes (cpab/gpu/Calcs.py:215-233): unused counts 0, unused increments 1.
 *  - Work is enqueued on the context's stream; no call synchronises unless
 *    it says so.  Functions return CPAB_OK or an error code and never throw;
 *    cpab_last_error() describes the last failure of the context.
 *  - A context is used by one host thread at a time; distinct contexts are
 *    independent.
 */
#ifndef CPAB_HIP_H
#define CPAB_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cpab_ctx cpab_ctx;

enum {
  CPAB_OK = 0,
  CPAB_ERR_ARG = 1,         /* bad shape / value: the reference's ValueError */
  CPAB_ERR_UNSUPPORTED = 2, /* e.g. tess I in 1D: NotImplementedError */
  CPAB_ERR_ALIAS = 3,       /* input and output share memory (_CpaCalcs.py:382-389) */
  CPAB_ERR_HIP = 4,         /* a HIP runtime call failed */
  CPAB_ERR_NOMEM = 5
};

/* Context on `device` using the default stream.  Replaces Calcs.__init__
 * (cpab/gpu/Calcs.py:44-63), which JIT-compiled the kernels. */
int cpab_create(int device, cpab_ctx **out);
int cpab_destroy(cpab_ctx *ctx);
const char *cpab_last_error(const cpab_ctx *ctx);
/* Enqueue subsequent work on `hip_stream` (a hipStream_t; NULL = default). */
int cpab_set_stream(cpab_ctx *ctx, void *hip_stream);
int cpab_synchronize(cpab_ctx *ctx);
/* Library / build identification string. */
const char *cpab_version(void);

/* Fused sign, dt scaling and per-cell matrix exponential:
 *   signedAs = sign * As;  Trels = expm(dt * signedAs)   (closed form in 1D).
 * Replaces CpaCalcs.prepare_signedSqAs_and_Tlocals_for_gpu
 * (cpab/cpaNd/_CpaCalcs.py:227-313) with its gpu_expm call.
 * sign is +1 (calc_T_fwd) or -1 (calc_T_inv).  signedAs_out may be NULL. */
int cpab_prep_trels(cpab_ctx *ctx, int dim, int nC, const double *As, int sign, double dt,
                    double *signedAs_out, double *Trels_out);

/* gpu_expm(As, Ts, p): series / closed-form expm of an already scaled table
 * (cpab/gpu/expm/expm_affine_by_series_2D.py:246, _3D.py:228).  dim 2 or 3. */
int cpab_expm(cpab_ctx *ctx, int dim, int nC, const double *As, double *Ts, int p);

/* calc_T: Calcs.calc_T (cpab/gpu/Calcs.py:454-561) and its kernel
 * (cpab/gpu/calc_2Dto2D.cu:348-408, calc_1Dto1D.cu:485-548,
 * calc_3Dto3D.cu:353-416).  Trels/As are the tables of cpab_prep_trels.
 * n_fallback (optional, int32 (nPts,)) receives, per point, the number of
 * steps whose Trel proposal left its cell and were redone by RK2. */
int cpab_calc_T(cpab_ctx *ctx, int dim, int tess, int nC, const double *pos0, double *pos,
                const double *Trels, const double *As, double dt, int nTimeSteps, int nStepsODE,
                int64_t nPts, int nC0, int nC1, int nC2, double inc_x, double inc_y,
                double inc_z, int32_t *n_fallback);

/* calc_T with the per-cell prep in the same launch: prepare_signedSqAs_and_
 * Tlocals_for_gpu followed by calc_T (cpab/cpaNd/_CpaCalcs.py:227-313,
 * 376-558; the kernels above).  As is the raw Avees table (nC x dim(dim+1));
 * every block forms sign * As and Trels = expm(dt * sign * As) itself (the
 * same operations as cpab_prep_trels) and block 0 writes them to
 * signedAs_out and Trels_out (the space's persistent tables, both
 * required), so results, tables and n_fallback are the bits of
 * cpab_prep_trels + cpab_calc_T.  When the tables do not fit LDS the prep
 * runs as its own launch first. */
int cpab_calc_T_prep(cpab_ctx *ctx, int dim, int tess, int nC, const double *pos0, double *pos,
                     const double *As, int sign, double dt, int nTimeSteps, int nStepsODE,
                     int64_t nPts, int nC0, int nC1, int nC2, double inc_x, double inc_y,
                     double inc_z, double *signedAs_out, double *Trels_out, int32_t *n_fallback);

/* calc_T_simple: Calcs.calc_T_simple (cpab/gpu/Calcs.py:339-450), kernel
 * cpab/gpu/calc_2Dto2D.cu:412-445 (nStepsODE*nTimeSteps RK2 substeps). */
int cpab_calc_T_simple(cpab_ctx *ctx, int dim, int tess, int nC, const double *pos0, double *pos,
                       const double *As, double dt, int nTimeSteps, int nStepsODE, int64_t nPts,
                       int nC0, int nC1, int nC2, double inc_x, double inc_y, double inc_z);

/* calc_trajectory: Calcs.calc_trajectory (cpab/gpu/Calcs.py:735-814), kernel
 * cpab/gpu/calc_2Dto2D.cu:499-556.  traj is (nTimeSteps, nPts, dim): row t is
 * the position after t+1 steps (the host prepends pos0, _CpaCalcs.py:866-868). */
int cpab_calc_trajectory(cpab_ctx *ctx, int dim, int tess, int nC, const double *pos0,
                         double *traj, const double *Trels, const double *As, double dt,
                         int nTimeSteps, int nStepsODE, int64_t nPts, int nC0, int nC1, int nC2,
                         double inc_x, double inc_y, double inc_z);

/* calc_grad_theta: the forward sensitivity of calc_T_simple
 * (cpab/gpu/calc_2Dto2D.cu:202-319, 448-491; bound nowhere in the reference,
 * cpab/gpu/Calcs.py:196, 237-335).  Bs: (d, nC, dim, dim+1) basis tables
 * (CpaSpace.BasMats, cpab/cpaNd/_CpaSpace.py:296-300); grad: (nPts, dim, d),
 * dT(x)/dtheta with the same RK2 (q_j' = B_j [p;1] + A_lin q_j); pos: the
 * calc_T_simple end points. */
int cpab_calc_grad_theta(cpab_ctx *ctx, int dim, int tess, int nC, const double *pos0, double *pos,
                         const double *As, const double *Bs, double *grad, int d, double dt,
                         int nTimeSteps, int nStepsODE, int64_t nPts, int nC0, int nC1, int nC2,
                         double inc_x, double inc_y, double inc_z);

/* Reverse-mode counterpart of calc_grad_theta for 2D (the same derivative of
 * calc_T_simple's RK2 map, cpab/gpu/calc_2Dto2D.cu:202-319, 412-445, taken
 * backwards): for nInst instances of nPts points each, given the
 * cotangents lam (nInst, nPts, 2) of the end points, gradA (nInst, nC, 6) =
 * sum_x lam(x)^T dT(x)/dA, A the integrated (signed) tables As (nInst, nC, 6).
 * dtheta follows as B^T gradA (times the sign).  One checkpoint per time step
 * in a library workspace (16 * nTimeSteps bytes per point); the substeps are
 * recomputed.  nStepsODE <= 16.  The tables and per-wave gradient tables
 * live in LDS: nC up to 1706 here, 853 for the calc_T map below
 * (CPAB_ERR_UNSUPPORTED beyond).  pos_out (optional): the end points. */
int cpab_calc_T_simple_vjp(cpab_ctx *ctx, int tess, int nC, int nInst, const double *pos0,
                           const double *lam, const double *As, double dt, int nTimeSteps,
                           int nStepsODE, int64_t nPts, int nC0, int nC1, double inc_x,
                           double inc_y, double *pos_out, double *gradA);

/* The SSD loss of images warped by calc_T_simple and its adjoint, fused, for
 * nInst image pairs: the points are the H x W pixel grid (x = column,
 * y = row), w(x) = src(T(x)) sampled bilinearly in FP64 (the cpab_remap2d
 * convention), loss[i] = 0.5 sum_x (w(x) - dst(x))^2 and gradA[i] = dloss/dA
 * (the loss of _ScaleDependentLogLikelihoodGaussian.py:71-109 up to the
 * factor -1/sigma^2).  src, dst: (nInst, H, W) float32. */
int cpab_ssd_grad2d(cpab_ctx *ctx, int tess, int nC, int nInst, const double *As, const float *src,
                    const float *dst, int H, int W, double dt, int nTimeSteps, int nStepsODE,
                    int nC0, int nC1, double inc_x, double inc_y, double *pos_out, double *loss,
                    double *gradA);

/* Reverse mode of calc_T (the map of calc_T_fwd / calc_T_inv,
 * cpab/gpu/calc_2Dto2D.cu:348-408; bound nowhere in the reference, which
 * differentiates only calc_T_simple): per time step the Trel proposal, kept
 * when it stays in its cell, else nStepsODE RK2 substeps, with the decisions
 * replayed from the forward sweep.  Trels = expm(dt As) as cpab_prep_trels
 * forms them (cpab/gpu/expm/expm_affine_by_series_2D.py:85-237), As the
 * signed tables (nInst, nC, 6).  gradA = sum_x lam(x)^T dT(x)/dAs, through
 * the Trel steps (the adjoint of the Frechet derivative of expm) and the RK2
 * substeps.  The additions follow a fixed order: gradA is the same bit for bit
 * on every run.  Same checkpoint workspace and limits as
 * cpab_calc_T_simple_vjp. */
int cpab_calc_T_vjp(cpab_ctx *ctx, int tess, int nC, int nInst, const double *pos0,
                    const double *lam, const double *Trels, const double *As, double dt,
                    int nTimeSteps, int nStepsODE, int64_t nPts, int nC0, int nC1, double inc_x,
                    double inc_y, double *pos_out, double *gradA);

/* cpab_ssd_grad2d for the calc_T map (Trels as in cpab_calc_T_vjp): the SSD
 * loss of the images warped by calc_T and its gradient in the signed tables.
 * sig (optional, uint32 (nInst, H*W)): a hash of every cell the forward
 * sweep looked up per pixel (equal hashes under two parameter values: the
 * same accept/redo decisions and substep cells, so the map is smooth
 * between them).  n_redo (optional, (nInst,)): the time steps redone by RK2,
 * summed over the pixels. */
int cpab_ssd_grad2d_calc_T(cpab_ctx *ctx, int tess, int nC, int nInst, const double *Trels,
                           const double *As, const float *src, const float *dst, int H, int W,
                           double dt, int nTimeSteps, int nStepsODE, int nC0, int nC1,
                           double inc_x, double inc_y, double *pos_out, double *loss,
                           double *gradA, uint32_t *sig, double *n_redo);

/* calc_v: Calcs.calc_velocities (cpab/gpu/Calcs.py:816-924), kernel
 * cpab/gpu/calc_2Dto2D.cu:561-602.  As is the unsigned table. */
int cpab_calc_v(cpab_ctx *ctx, int dim, int tess, int nC, const double *pts, double *vel,
                const double *As, int64_t nPts, int nC0, int nC1, int nC2, double inc_x,
                double inc_y, double inc_z);

/* calc_cell_idx: Calcs.calc_cell_idx (cpab/gpu/Calcs.py:968-1010), kernel
 * cpab/gpu/calc_2Dto2D.cu:323-344. */
int cpab_calc_cell_idx(cpab_ctx *ctx, int dim, int tess, const double *pts, int32_t *cell_idx,
                       int64_t nPts, int nC0, int nC1, int nC2, double inc_x, double inc_y,
                       double inc_z);

/* remap_fwd / remap_inv of an H x W float image: out[r*W + c] = img(pts[r*W + c]),
 * bilinear, zero outside (cpab/cpa2d/TransformWrapper.py:261-317; the
 * reference delegates to the un-vendored of.gpu.remap.remap2d). */
int cpab_remap2d(cpab_ctx *ctx, const double *pts, const float *img, float *out, int H, int W);

/* Trilinear remap of a D x H x W float volume (cpab/cpa3d/TransformWrapper.py:254-290). */
int cpab_remap3d(cpab_ctx *ctx, const double *pts, const float *vol, float *out, int D, int H,
                 int W);

/* Gaussian SSD log-likelihood sum_i -0.5*(x_i - y_i)^2/sigma^2, written to the
 * device scalar *ll (cpab/cpa2d/model/transformations/register/gpu/gaussian.py:15-34
 * and the sum of _ScaleDependentLogLikelihoodGeneral.py:118). */
int cpab_ssd_ll(cpab_ctx *ctx, const float *x, const float *y, int64_t n, double sigma,
                double *ll);

/* ------------------------------------------------------------------------
 * Batched calc_T over many independent instances (no reference counterpart:
 * the reference makes one Calcs.calc_T call per instance, cpab/gpu/Calcs.py:454-561;
 * this runs a list of such calls as one persistent launch).  Every instance
 * has its own tessellation (same dim and tess across the batch; increments
 * all powers of two or none), tables and points; the points of all
 * instances are concatenated in instance order, as are their unsigned As
 * (cells in instance order) for cpab_batch_prep.  Per point the result is
 * bit-identical to cpab_calc_T on that instance alone. */
typedef struct {
  int nC0, nC1, nC2;
  double inc_x, inc_y, inc_z;
  int64_t nPts;
} cpab_batch_inst;

/* Schedule of the persistent launch: points per tile, threads per block
 * (a multiple of 64, at most 1024), tiles a block takes per queue grab. */
typedef struct {
  int tile;
  int block;
  int chunk;
} cpab_batch_sched;

typedef struct cpab_batch cpab_batch;

int cpab_batch_create(cpab_ctx *ctx, int dim, int tess, int nInst, const cpab_batch_inst *insts,
                      cpab_batch **out);
/* Total cells, total points, doubles of the table buffer, largest nC. */
int cpab_batch_sizes(cpab_batch *b, int64_t *nCells, int64_t *nPts, int64_t *tableDoubles,
                     int *maxNC);
/* prepare_signedSqAs_and_Tlocals_for_gpu (cpab/cpaNd/_CpaCalcs.py:227-313) for
 * every instance at once: As (nCells, dim(dim+1)) unsigned tables in
 * instance order -> tables (per instance: Trels then sign*As). */
int cpab_batch_prep(cpab_batch *b, const double *As, int sign, double dt, double *tables);
/* calc_T of every instance (cpab/gpu/calc_2Dto2D.cu:348-408 per point).
 * sched NULL: {16384, 1024, 4}.  use_graph != 0: the launch is captured once
 * per argument set and replayed.  nfb (optional, int32 per point): rejected
 * Trel proposals. */
int cpab_batch_calc_T(cpab_batch *b, const double *tables, const double *pos0, double *out,
                      double dt, int nTimeSteps, int nStepsODE, const cpab_batch_sched *sched,
                      int use_graph, int32_t *nfb);
int cpab_batch_destroy(cpab_batch *b);

/* One tile of a served request for cpab_calc_T_items: the request's own
 * tables (Trels and signed As of cpab_prep_trels, (nC, dim(dim+1)) each),
 * points and output, the point range [begin, end) of this tile, and the
 * request's tessellation. */
typedef struct {
  const double *trels;
  const double *As;
  const double *pos0;
  double *out;
  int64_t begin, end;
  int nC0, nC1, nC2;
  double inc_x, inc_y, inc_z;
} cpab_work_item;

/* Avees = B theta for m thetas at once: CpaSpace.theta2Avees
 * (cpab/cpaNd/_CpaSpace.py:366-376).  B (D, d) row-major, theta (m, d),
 * Av (m, D).  Each row is summed in index order, so a theta's Avees do not
 * depend on the rest of the batch. */
int cpab_theta2avees(cpab_ctx *ctx, int D, int d, int m, const double *B, const double *theta,
                     double *Av);

/* calc_T over a list of tiles of any requests of one dim and tess (no
 * reference counterpart: the serving loop's continuous batching), one
 * persistent launch; per point bit-identical to cpab_calc_T.  Increments of
 * all items powers of two or none.  sched NULL: {16384 (unused), 1024, 1}. */
int cpab_calc_T_items(cpab_ctx *ctx, int dim, int tess, int nItems, const cpab_work_item *items,
                      double dt, int nTimeSteps, int nStepsODE, const cpab_batch_sched *sched);

/* cpab_calc_T_items without any allocation: no reference counterpart; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * kernels and results as cpab_calc_T_items (bit for bit); the work list and
 * the launch's queue counter live in a static device array of the library
 * (one per GPU, 640 KiB, held until the process ends), so a call takes
 * at most 4096 items (CPAB_ERR_ARG beyond).  Every call on a GPU shares that
 * array: calls that overlap on one GPU (other streams or contexts) race; calls
 * on one stream are ordered.  No kernel timing (cpab_prof_*) for this entry. */
int cpab_calc_T_items_noalloc(cpab_ctx *ctx, int dim, int tess, int nItems,
                              const cpab_work_item *items, double dt, int nTimeSteps,
                              int nStepsODE, const cpab_batch_sched *sched);

/* ------------------------------------------------------------------------
 * Device-resident Metropolis sampling of 2D image registration, for many
 * independent chains at once: Metropolis.run / .step
 * (cpab/cpaNd/inference/Metropolis.py:41-151) with the random-walk proposal
 * of Proposal.__call__ (cpab/cpaNd/inference/Proposal.py:29-64), theta2Avees
 * (cpab/cpaNd/_CpaSpace.py:366-376), calc_T_inv and the Gaussian SSD
 * likelihood (cpab/cpa2d/model/transformations/register/
 * _ScaleDependentLogLikelihoodGaussian.py:34-109).  The pixel (r, c) of the
 * H x W images is the point (x = c, y = r) (x_dense_img,
 * cpab/cpa2d/calcs/_CpaCalcs.py:46-62).  A step is:
 *   theta' = theta + Lprop z,  z ~ N(0, I) (Philox4x32-10 + Box-Muller,
 *            counter (step, chain), key = seed);
 *   Avees  = B theta';  T = calc_T_inv(grid; Avees);
 *   ll'    = sum -0.5 (src(T(x)) - dst(x))^2 / sigma^2  (/ H*W if use_ave_ll);
 *   log_ratio = ll' - ll (+ wlp (lp' - lp) with a prior, Metropolis.py:64-70);
 *   accept if log_ratio > 0, else if log(u) < log_ratio (log(u) > log_ratio
 *   with inverted_acceptance, the reference's test, Metropolis.py:77-82).
 * All arrays are device pointers owned by the caller; the object owns only
 * its scratch.  B is copied (transposed) when the object is created, on the
 * context's stream; later changes to B are not seen. ----------------------- */
typedef struct {
  int nChains;          /* independent chains sharing the images and basis */
  int d;                /* dimension of theta */
  int tess, nC, nC0, nC1;
  double inc_x, inc_y;
  int H, W;             /* image size; the domain is [0, W] x [0, H] */
  double sigma;         /* SSD noise level */
  double dt;            /* integration: calc_T_inv with (dt, nTimeSteps, nStepsODE) */
  int nTimeSteps, nStepsODE;
  int use_ave_ll;       /* divide the log-likelihood by H*W (Metropolis use_ave_ll) */
  int inverted_acceptance;
  uint64_t seed;
  const double *B;      /* (nC*6, d) row-major: Avees = B theta */
  const double *Lprop;  /* (d, d) row-major lower triangular proposal factor */
  const float *src;     /* (H, W) source image, warped by the transform */
  const float *dst;     /* (H, W) target image */
  double *theta;        /* (nChains, d) current state (in/out) */
  double *ll;           /* (nChains,) current log-likelihood (out) */
  int64_t *n_accepted;  /* (nChains,) accepted proposals (in/out) */
  int64_t *step;        /* (nChains,) step counters (in/out; the RNG counter) */
  double *trace;        /* optional (nChains, trace_len): ll after each step */
  int trace_len;
  uint64_t *n_fallback; /* optional (nChains,): accumulates the integration steps
                           whose Trel proposal left its cell (RK2 redone), over the
                           decided steps */
  int spec_depth;       /* proposals evaluated per chain and launch, at most 64: 1 = one
                           step per launch; K > 1 = speculative prefetch of the next K
                           steps on the all-accepted path (same chain bit for bit; slots
                           above a rejection found during the launch are not started,
                           or stop at a step checkpoint);
                           0 = chosen by the library (about 2^20 points per launch, at
                           most 16) */
  int64_t *stats;       /* optional (nChains, 5), accumulated: proposals evaluated in
                           full, rejected Trel proposals of every pixel integrated,
                           launches, steps decided, pixels' worth integrated (a pixel
                           stopped at step t of nT counts t / nT; with speculation a
                           slot above a known rejection is not started or stops early,
                           so the counts of work done depend on timing; the chain
                           does not) */
  int pts_per_wave;     /* pixels a wave integrates (a multiple of 64, at most 4096);
                           0 = chosen by the library from the launch size (64, 128 or
                           256 pixels per wave; many chains one step per launch also
                           get one-wave blocks).  Pixels go to waves in 8 x 8 tiles
                           when W and H are multiples of 8, else in row order.  The
                           log-likelihood's summation order follows the pixels'
                           assignment to blocks, so chains give the same bits alone
                           and in a batch when they run with the same explicit value
                           (4-wave blocks then). */
  double wlp;           /* weight of the log-prior ratio (Metropolis wlp, :15, 64-70) */
  const double *prior_prec; /* optional (d, d) row-major precision J of the Gaussian prior
                           N(0, J^-1) on theta (ScaleDependentLogPrior.calc_lp,
                           cpab/cpaNd/model/_ScaleDependentLogPrior.py: lp = -0.5 theta^T J
                           theta, mu = 0); when set, log_ratio = ll' - ll + wlp (lp' - lp)
                           (Metropolis.py:64-70; Register.py:224-248 passes one with
                           wlp = 1e-4).  NULL: no prior term.  Read at every launch. */
  int integrate_corner; /* 0: the pixel at the domain corner (0, 0) is left out of the
                           integration when a bound on its path, from the tables of the
                           corner's cells, shows that its FP32 bilinear sample is
                           src(0, 0) bit for bit (the log-likelihood is then the same: the
                           corner's lane takes no other pixel either way; that pixel's
                           rejected Trel proposals are not counted in n_fallback and
                           stats).  Nonzero: always integrate it (no reference
                           counterpart: the reference integrates every pixel). */
} cpab_reg2d_desc;

typedef struct cpab_reg2d cpab_reg2d;

int cpab_reg2d_create(cpab_ctx *ctx, const cpab_reg2d_desc *desc, cpab_reg2d **out);
/* ll of the current theta of every chain (Metropolis.set_theta, :25-40). */
int cpab_reg2d_init(cpab_reg2d *r);
/* nSteps Metropolis steps of every chain.  A launch decides one step per
 * chain (spec_depth 1) or 1..K (speculation); use_graph != 0 replays a
 * captured 16-launch hipGraph instead of launching (the remainder is
 * launched directly).  With speculation the call reads the step counters
 * back between batches of launches, so it returns after the steps are done;
 * with spec_depth 1 it only enqueues. */
int cpab_reg2d_run(cpab_reg2d *r, int nSteps, int use_graph);
/* theta_prop (nChains x d): the proposal the next step will evaluate (armed
 * at the end of every step); ll_prop (nChains): the log-likelihood of the
 * proposal evaluated by the last step.  Either pointer may be NULL. */
int cpab_reg2d_last_proposal(cpab_reg2d *r, double *theta_prop, double *ll_prop);
/* The speculation depth in use (spec_depth as resolved at create). */
int cpab_reg2d_spec_depth(cpab_reg2d *r, int *K);
/* Change the speculation depth between runs (up to four times the library's
 * own choice at create, at most 64, or the depth asked for if larger); the armed proposals
 * are recomputed for the new depth, so the chain is unchanged.  Call after
 * cpab_reg2d_init.  Synchronises the context's stream.  The graph captured at
 * each depth is kept, so returning to a depth replays it without capturing. */
int cpab_reg2d_set_spec_depth(cpab_reg2d *r, int K);
/* Step launches so far: by replay of a captured graph and launched directly
 * (no reference counterpart; the reference launches every kernel of a step
 * from the host, Metropolis.py:41-102).  Either pointer may be NULL. */
int cpab_reg2d_launch_counts(cpab_reg2d *r, int64_t *replayed, int64_t *direct);
int cpab_reg2d_destroy(cpab_reg2d *r);

/* Kernel timing with HIP events on the context's stream (no reference
 * counterpart; the reference wraps its launches in GpuTimer,
 * cpab/gpu/Calcs.py:537-563).  While enabled, every integration launch
 * (calc_T, trajectory, simple) is bracketed by events; cpab_prof_read
 * synchronises, returns the launch count and summed milliseconds, and
 * resets the counters. */
int cpab_prof_enable(cpab_ctx *ctx, int on);
int cpab_prof_read(cpab_ctx *ctx, int64_t *n_launches, double *total_ms);

#ifdef __cplusplus
}
#endif

#endif /* CPAB_HIP_H */
