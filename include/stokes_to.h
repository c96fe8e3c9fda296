/*
 * stokes_to.h -- C ABI of libstokes_hip.so, the MI355X (gfx950) implementation of
 * the Stokes/Brinkman topology-optimization hot path of jyl-pages/Stokes_TO
 * (app `fluid_topo`).
 *
 * Conventions
 *  - Every function returns ST_OK (0) or a negative ST_E* code; nothing exits the
 *    process and no C++ exception crosses the ABI.  st_last_error() gives text.
 *  - Host arrays use the ABI layout of the reference's host fields
 *    (SX/reservoir/Field.h:67-72): C-order (i, j, k) with k fastest.
 *      cells  : n*n*n            (2D: n*n)
 *      faces d: dims n+1 along d, n across (2D: x, y only)
 *    A dof vector / descriptor array is [cells | faces_x | faces_y | faces_z],
 *    the order of the reference's solver vectors (ST/FluidSteadyState3D.cpp:101-105).
 *  - Vectors are FP64 at the ABI; the preconditioner precision is internal.
 *  - A context owns one device, one stream and all device state; it is not
 *    thread-safe (like the reference's LinearMapping objects, which keep
 *    internal scratch: CX/LinearMapping.h:4-14).
 *  - Fixed DOFs (Dirichlet faces) have identity rows ("definite fix",
 *    ST/StokeFlowMapping3D.cpp:301-305); face vol equals the free mask
 *    (ST/StokeFlowDescriptor3D.cu:37-50).
 */
#ifndef STOKES_TO_H
#define STOKES_TO_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct st_ctx st_ctx;

enum {
    ST_OK = 0,
    ST_EINVAL = -1,     /* bad argument or unsupported configuration */
    ST_EHIP = -2,       /* HIP runtime failure (allocation, launch) */
    ST_ENOCONV = -3,    /* Krylov iteration cap reached before rtol */
    ST_ESINGULAR = -4,  /* coarsest-level matrix is singular (closed domain without gauge) */
    ST_ESTATE = -5,     /* call order: geometry / penalty / rhs missing */
    ST_ENAN = -6        /* NaN or Inf in an input, the residual or the coarsest-level matrix */
};

enum { ST_FP64 = 0, ST_FP32 = 1, ST_BF16 = 2 };        /* preconditioner precision */
enum { ST_INIT_ZERO = 0, ST_INIT_KEEP = 1, ST_INIT_HOST = 2, ST_INIT_LEARNED = 3 };
enum { ST_OPT_OC = 0, ST_OPT_MMA = 1 };                 /* design update of st_to_step */

typedef struct {
    int dim;             /* 2 or 3 */
    int n;               /* cells per axis; n = 4 * 2^k */
    /* design -> Brinkman penalty: alpha(r) = alpha_max + (alpha_min - alpha_max) r(1+q)/(r+q)
       on r = rho^power, times alpha_scale (FT/FluidTopoOpt3D.cpp:110-123, :145-159) */
    double alpha_max, alpha_min, q, power, alpha_scale;
    int precision;       /* ST_FP64 / ST_FP32 / ST_BF16 multigrid preconditioner */
    double rtol;         /* stop on the true FP64 residual ||b-Ax||/||b|| <= rtol */
    int max_iter;        /* total FGMRES iterations */
    int restart;         /* FGMRES restart length (<= 32) */
    int nu_pre, nu_post; /* smoothing sweeps (each = both colours) */
    double omega;        /* smoother damping */
    int closed;          /* all-ones pressure gauge on the coarsest level (initial value;
                            st_set_geometry's closed argument sets it) */
    double filter_radius;/* density filter radius in cells (<= 1 disables) */
    int optimizer;       /* ST_OPT_OC (default) or ST_OPT_MMA, the reference's optimizer */
    int fused_sweep;     /* 1: FP32/BF16 smoothing sweeps of the tiled levels in one pass (both
                            colours, bit-identical in FP32); 0 (default): one pass per colour,
                            faster on MI355X (the one-pass kernel holds 126 KB of LDS) */
    int warm_order;      /* ST_INIT_KEEP / ST_INIT_LEARNED starts of a sequence of solves
                            (st_to_step, repeated st_solve): 0 the held solution; 1 its linear,
                            2 its quadratic extrapolation from the previous solutions of the
                            sequence, taken when its true FP64 residual is lower; 3 (default) the
                            held solution plus the combination of the last three differences of
                            solutions with the least residual (FP32 precision, 3D, grid side a
                            multiple of 64; elsewhere 2) */
} st_params;

typedef struct {
    int iters;           /* FGMRES iterations (= V-cycles) */
    int converged;       /* 1 if relres <= rtol */
    double relres;       /* true FP64 relative residual at exit */
    double relres0;      /* true relative residual of the initial guess */
    double ms;           /* wall time of the solve incl. hierarchy setup (hipEvents) */
    double setup_ms;     /* hierarchy setup part (coarsening + coarsest inverse) */
    double init_ms;      /* initial-guess part (extrapolation or minimal-residual start, learned
                            predictor inference and their residual checks; the minimal-residual
                            start sets up the hierarchy itself, so there setup_ms is part of it) */
    int learned_used;    /* ST_INIT_LEARNED: 1 if the network's candidate had the smaller true residual */
} st_stats;

/* defaults: the configs' material model (alpha_max 1e4, q 0.1, power 1, scale 1/n^2),
   FP64, rtol 1e-8, FGMRES(30), V(1,2), omega 1, filter radius 1.5, colour passes */
void st_default_params(st_params *p, int dim, int n);

/* replaces FluidSteadyState3D::init (ST/FluidSteadyState3D.cpp:12-90): builds the
   multigrid hierarchy and all device buffers */
int st_create(st_ctx **out, const st_params *p);
void st_destroy(st_ctx *c);
const char *st_last_error(const st_ctx *c);
int st_get_params(const st_ctx *c, st_params *p);
/* number of dofs of the ABI vector and of the design (cells) */
long st_num_dofs(const st_ctx *c);
long st_num_cells(const st_ctx *c);

/* replaces FluidSteadyState3D::init_boundary(cell_vol, cell_fixed, face_vol, face_fixed,
   bool _closed) (ST/FluidSteadyState3D.h:33-34, ST/FluidSteadyState3D.cpp:92-114):
   fixed[cells|faces], 1 = fixed (Dirichlet) DOF; vol is the free mask.  closed = 1 adds
   the all-ones pressure gauge to the coarsest matrix (a domain whose boundary is fixed
   everywhere, ST/StokeLike.cu:605-608), 0 removes it, -1 keeps params.closed.  The
   coarsest inverse is rebuilt on the next solve. */
int st_set_geometry(st_ctx *c, const uint8_t *fixed, int closed);
/* replaces FluidSteadyState3D::update_penalty (ST/FluidSteadyState3D.cpp:130-150):
   pen[cells|faces]; cell penalties must be zero */
int st_set_penalty(st_ctx *c, const double *pen);
/* replaces the host loop of FluidTopoOpt3D::Update_Fluid (FT/FluidTopoOpt3D.cpp:145-159):
   rho[cells] -> face alpha on the device */
int st_set_design(st_ctx *c, const double *rho);
int st_get_design(st_ctx *c, double *rho);
/* device-pointer variant of st_set_design (rho in device memory, ABI cell order) */
int st_set_design_device(st_ctx *c, const double *d_rho);
/* replaces FluidSteadyState3D::update_b (ST/FluidSteadyState3D.cpp:116-127):
   body[dofs] (nullable) are source terms on free rows; dirichlet[dofs] (nullable)
   are values on fixed faces, lifted into the right-hand side */
int st_set_rhs(st_ctx *c, const double *body, const double *dirichlet);
/* replaces FluidSteadyState3D::solve + GMRES::Solve (ST/FluidSteadyState3D.cpp:152-170,
   CX/GMRES.cpp:20-110): FGMRES right-preconditioned by one V-cycle, stopped on the
   true FP64 residual.  init: ST_INIT_ZERO (Dirichlet values, zero elsewhere),
   ST_INIT_KEEP (the device solution of the previous solve), ST_INIT_HOST (x0),
   ST_INIT_LEARNED (the learned predictor: with a previous solution, that solution
   corrected by the predicted change of the coarse state between its design and this
   one; without one, the Dirichlet values plus the predicted coarse state; taken only
   if its true FP64 residual is below that of the candidate without the network, see
   st_stats.learned_used; after 3 rejections in a row the network's candidate is only
   built on every 4th call).  x_out (nullable) receives the solution. */
int st_solve(st_ctx *c, int init, const double *x0, double *x_out, st_stats *stats);
int st_get_solution(st_ctx *c, double *x);
/* forget the held forward / adjoint solutions (a new run): ST_INIT_KEEP then starts
   from the Dirichlet values and ST_INIT_LEARNED from the network's prediction alone.
   No reference counterpart (the reference's GMRES always starts from zero). */
int st_reset_solution(st_ctx *c);
/* the design gradient dJ/drho (after the filter's chain rule) of the last st_to_step,
   the reference's design_grad (FT/FluidTopoOpt3D.cpp:170-198) */
int st_get_gradient(st_ctx *c, double *dc);
/* replaces StokeFlowMapping3DFixed::applyMapping through LinearMapping
   (ST/StokeFlowMapping3D.cu:176-306, CX/LinearMapping.h:4-14): y = A x, host arrays */
int st_apply(st_ctx *c, const double *x, double *y);
/* device-pointer variant (ABI layout in device memory, on the context's stream) */
int st_apply_device(st_ctx *c, const double *d_x, double *d_y);
/* replaces FluidEnergy3D::compute_energy (FT/FluidEnergy3D.cpp:94-109): J = x^T G x / 2 */
int st_objective(st_ctx *c, const double *x, double *J);
/* replaces FluidEnergy3D::compute_gradient_x (FT/FluidEnergy3D.cpp:71-81): gx = G x */
int st_energy_grad(st_ctx *c, const double *x, double *gx);
/* replaces the sensitivity loop of FluidTopoOpt3D::Update_Grad (FT/FluidTopoOpt3D.cpp:170-198):
   dj[cells] = dJ/drho from forward x and adjoint lam (A lam = G x) */
int st_sensitivity(st_ctx *c, const double *rho, const double *x, const double *lam, double *dj);
/* density filter (no counterpart in the reference, which feeds rho to alpha directly,
   FT/FluidTopoOpt3D.cpp:145-168): out = H in / (H 1), or its transpose */
int st_filter(st_ctx *c, const double *in, double *out, int transpose);
/* optimality-criteria design update with bisection on the volume multiplier, in the
   slot of OptimizerMMA::Optimize_MMA (SX/optimization/OptimizerMma.h:138-170);
   volume measured on the filtered density */
int st_oc_update(st_ctx *c, const double *rho, const double *dc, double volfrac, double move, double eta,
                 double rho_min, double *rho_new, double *change);
/* replaces MMASolver::Update with one constraint (EXT/mma/MMASolver.cpp:74-89, the
   update of SX/optimization/OptimizerMma.h:138-170): one method-of-moving-asymptotes
   step of x (host, ncell values, in place) from the objective gradient dfdx, the
   constraint value g (feasible if g <= 0) and its gradient dgdx, inside the box
   [xmin, xmax]; the dual is solved by the same interior-point Newton iteration.
   The context keeps the asymptotes and the two previous iterates across calls, like
   the reference's MMASolver object; st_mma_reset starts over. */
int st_mma_update(st_ctx *c, double *x, const double *dfdx, double g, const double *dgdx, const double *xmin,
                  const double *xmax, int *newton_steps);
int st_mma_reset(st_ctx *c);
/* one full design iteration on the device, the body of Optimize_MMA's loop
   (FT/FluidTopoOpt3D.cpp:4-17): filter, design->alpha, forward solve, objective,
   adjoint solve, sensitivity, filter transpose, and the design update of the device
   design: OC, or with params.optimizer = ST_OPT_MMA the reference's MMA step with the
   reference's volume constraint 1e4 (mean(rho~) - volfrac) <= 0 and box
   [max(rho_min, rho - move), min(1, rho + move)] (eta unused).
   init: ST_INIT_ZERO (both solves from zero, as the reference's GMRES), ST_INIT_KEEP
   (previous forward / adjoint solutions), ST_INIT_LEARNED (forward solve from the
   learned initial guess of st_solve, adjoint from the previous adjoint; needs a
   trained predictor, 3D only) */
int st_to_step(st_ctx *c, int init, double volfrac, double move, double eta, double rho_min, double *J,
               double *vol, double *change, st_stats *fwd, st_stats *adj);

/* J = x^T G x / 2 of st_objective for callers that cannot allocate device memory
   (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
   No counterpart in the reference.  No context: d_fixed (uint8, [cells|faces], 1 =
   fixed), d_pen (FP64 [cells|faces]: zero cells, face alpha, zero on fixed faces) and
   d_x (FP64 dof vector) are device arrays in the ABI layout; J is written to the
   device scalar d_J, on `stream` (a hipStream_t; NULL = the null stream), without a
   host synchronisation.  Scratch is a 16 KiB device array per GPU inside the library
   (no allocation, no memset); calls that overlap on one GPU race on it.  n <= 1024,
   dim 2 or 3 (G of the reference's 2D or 3D objective).  On error st_last_error(NULL)
   gives the message. */
int st_objective_noalloc(int dim, int n, const uint8_t *d_fixed, const double *d_pen, const double *d_x, double *d_J,
                         void *stream);

/* ---- kernel timing (HIP events around each launch on the context stream) ---- */
enum { ST_PROF_SMOOTH = 0, ST_PROF_APPLY = 1, ST_PROF_RESIDUAL = 2, ST_PROF_SMOOTH0 = 3, ST_PROF_SWEEP = 4,
       ST_PROF_NKIND = 5 };
/* ST_PROF_SMOOTH: smoother colour passes with x given (k_march<., SMOOTH>);
   ST_PROF_SMOOTH0: the first pass of a V-cycle, from x = 0 (k_smooth0_pairs);
   ST_PROF_SWEEP: one-pass two-colour sweeps (k_sweep, params.fused_sweep) */
/* on != 0 starts recording the finest-level launches of each kind; 0 stops */
int st_profile(st_ctx *c, int on);
/* total device ms and launch count recorded for `kind` since st_profile(c, 1);
   bytes = algorithmic HBM bytes of one launch of that kind (FP64 unless the MG
   precision applies) */
int st_profile_read(st_ctx *c, int kind, double *total_ms, long *launches, double *bytes_per_launch);

/* ---- learned initial guess (no counterpart in the reference, whose GMRES starts
   from zero, CX/GMRES.cpp:29-40) ----
   A 3-level 3D conv U-Net (16/32/64 channels, 3x3x3, BF16 weights, FP32 accumulate,
   BF16 MFMA kernels) on the n/4 multigrid level maps (block-mean density, boundary
   channels) to the coarse state (p, u).  3D only. */
/* record the pair (density used for the current alpha, current device solution) */
int st_record_sample(st_ctx *c);
/* the network's training loss on the current (design, solution) pair without
   recording it, and the loss of the zero prediction on the same pair: how much of the
   coarse state the trained network predicts on a design it has not seen.  No
   reference counterpart. */
int st_net_eval(st_ctx *c, double *loss, double *loss_zero);
int st_num_samples(const st_ctx *c);
/* Adam training on the recorded samples (batch 1); loss = mean squared error of the
   scaled coarse state after the last epoch */
int st_net_train(st_ctx *c, int epochs, double lr, double *loss);
/* parameter access and a full forward/backward on host features/targets
   (R^3 x 4 each, R = n/4), used by the gradient parity tests */
long st_net_num_params(st_ctx *c);
int st_net_get_params(st_ctx *c, float *params);
int st_net_set_params(st_ctx *c, const float *params);
int st_net_grad(st_ctx *c, const float *features, const float *targets, double *loss, float *grad, float *out);
/* one convolution layer on host arrays, no context: mode 0 forward (+bias, relu),
   1 input gradient of gout, 2 weight and bias gradients of (in, gout); channels-last,
   R^3 grid, zero padding, ntap 27 (3x3x3) or 1 */
int st_net_conv_test(int R, int ntap, int cin, int cout, int relu, int mode, const float *in, const float *w,
                     const float *bias, const float *gout, float *out, float *gw, float *gb);

/* ---- multigrid internals exposed for component parity tests ---- */
int st_mg_levels(const st_ctx *c);
int st_mg_level_n(const st_ctx *c, int level);
/* one V-cycle z = M b (host arrays, ABI layout of level 0) */
int st_mg_vcycle(st_ctx *c, const double *b, double *z);
/* level descriptor: fixed flags [cells|faces] and face alpha (FP64 hierarchy) */
int st_mg_level_desc(st_ctx *c, int level, uint8_t *fixed, double *alpha_faces);
/* one smoother colour step on `level`: xout = S_color(b, xin) */
int st_mg_smooth(st_ctx *c, int level, int color, const double *b, const double *xin, double *xout);
/* the same step with the context's mixed-precision smoother (FP32 or BF16 storage;
   FP64 contexts: st_mg_smooth) */
int st_mg_smooth_mp(st_ctx *c, int level, int color, const double *b, const double *xin, double *xout);
/* bc = R rf (fine level -> level+1), xf_out = xf + P ec */
int st_mg_restrict(st_ctx *c, int level, const double *rf, double *bc);
int st_mg_prolong(st_ctx *c, int level, const double *ec, const double *xf, double *xf_out);

#ifdef __cplusplus
}
#endif
#endif
