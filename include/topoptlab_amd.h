/*
 * topoptlab_amd — C ABI of the MI355X (gfx950) SIMP finite-element hot path.
 *
 * All vector arguments are DEVICE pointers owned by the caller, contiguous
 * float64, in the reference's own numbering: node id
 * (iz*(nelx+1) + ix)*(nely+1) + iy, DOF = ndpn*node + component, element id
 * (iz*nelx + ix)*nely + iy (2D: nely*ix + iy).  DOF vectors are full length
 * (fixed DOFs included); on output fixed DOFs hold 0, on input they are
 * ignored.  Handles own their internal workspaces.  Every call enqueues on
 * the handle's HIP stream; calls that return host scalars synchronise it.
 *
 * Return codes: TL_OK, TL_EINVAL (bad argument; Python raises ValueError),
 * TL_ENOCONV (iterative solve did not converge; RuntimeError, as
 * topoptlab/solve_linsystem.py:349-351), TL_EHIP (HIP runtime error).
 * tl_last_error() returns the message of the last failing call of the
 * calling thread.
 */
#ifndef TOPOPTLAB_AMD_H
#define TOPOPTLAB_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { TL_OK = 0, TL_EINVAL = 1, TL_ENOCONV = 2, TL_EHIP = 3 };
enum { TL_PREC_FP64 = 0, TL_PREC_FP32 = 1, TL_PREC_BF16 = 2, TL_PREC_FP16 = 3 };
enum { TL_PRECOND_NONE = 0, TL_PRECOND_JACOBI = 1, TL_PRECOND_GMG = 2 };

typedef struct tl_problem tl_problem;
typedef struct tl_filter tl_filter;

/* Structured Q4 (ndim 2) / H8 (ndim 3) grid with unit elements. */
typedef struct {
  int32_t ndim;          /* 2 or 3 */
  int32_t nelx, nely, nelz;  /* nelz ignored for ndim == 2 */
  int32_t ndpn;          /* DOFs per node: 1 (Poisson) or ndim (elasticity) */
  const double* KE;      /* HOST, (2^ndim*ndpn)^2 row-major element matrix */
  double eps, penal;     /* modified SIMP: s = eps + (1-eps) x^penal */
  const int64_t* fixed;  /* HOST, Dirichlet (zero) DOF indices */
  int64_t nfixed;
  void* stream;          /* hipStream_t; NULL = default stream */
} tl_grid_desc;

/* Replaces the set-up half of topoptlab/topology_optimization.py:264-371
 * (KE, edofMat, iK/jK, BC arrays): the structured grid is never assembled. */
int tl_create(const tl_grid_desc* desc, tl_problem** out);
void tl_destroy(tl_problem* p);
int tl_set_stream(tl_problem* p, void* stream);
const char* tl_last_error(void);
int64_t tl_ndof(const tl_problem* p);
int64_t tl_nel(const tl_problem* p);

/* Replaces topology_optimization.py:419-420 (scale = simp(xPhys); Kes).
 * xPhys: device, nel float64. */
int tl_set_density(tl_problem* p, const double* xPhys);

/* y = K_ff u (fixed rows/cols removed), matrix-free.  Replaces the CSR
 * product of topology_optimization.py:429-451 + linear_solvers.py:114,130.
 * prec: TL_PREC_FP64 or TL_PREC_FP32; TL_PREC_BF16 / TL_PREC_FP16 run the
 * V-cycle's matrix-core operator (3D H8 elasticity only: u rounded to fp32,
 * element vectors and KE to bf16/fp16 after max-normalisation, fp32
 * accumulation).  Inputs/outputs stay float64. */
int tl_apply_K(tl_problem* p, const double* u, double* y, int32_t prec);

/* Diagonal of K_ff (fixed DOFs: 0). */
int tl_diag_K(tl_problem* p, double* d);

typedef struct {
  int32_t nlevels;       /* grid levels incl. the fine one (>= 2) */
  double omega;          /* damped-Jacobi weight */
  int32_t nu1, nu2;      /* pre-/post-smoothing sweeps */
  int32_t precision;     /* V-cycle arithmetic: TL_PREC_FP64, TL_PREC_FP32, or
                            TL_PREC_BF16 / TL_PREC_FP16 (3D H8 elasticity: fine-
                            level operator on MFMA, fp32 vectors and coarse levels) */
  int32_t gamma;         /* cycle index on levels >= gamma_from: 1 (or 0) = V-cycle, 2 = W */
  int32_t gamma_from;    /* first level (>= 1; 0 means 1) cycled gamma times per visit */
} tl_mg_opts;

/* Galerkin GMG hierarchy for the current density (multigrid.py:89-223,
 * gmg.py:34-224).  Must be called after each tl_set_density that should
 * update the preconditioner. */
int tl_setup_mg(tl_problem* p, const tl_mg_opts* opts);

typedef struct {
  double rtol, atol;     /* stop: ||r||_2 < max(atol, rtol*||b||_2) (linear_solvers.py:107-117) */
  int32_t maxiter;
  int32_t x0_given;      /* 1: u holds the initial guess (warm start) */
  int32_t precond;       /* TL_PRECOND_* */
  int32_t check_every;   /* iterations between host convergence checks */
} tl_solve_opts;

typedef struct {
  int32_t iterations;
  int32_t status;        /* 0 converged, TL_ENOCONV otherwise */
  double relres_recursive;  /* ||r_k|| / ||b|| of the recursive residual */
  double relres_true;       /* ||b - K_ff u|| / ||b||, recomputed in FP64 */
} tl_solve_info;

/* PCG on K_ff u = f (linear_solvers.py:50-136 via
 * solve_linsystem.py:341-352).  f, u: device, ndof float64. */
int tl_solve(tl_problem* p, const double* f, double* u,
             const tl_solve_opts* opts, tl_solve_info* info);

/* ---- Batched solves of many independent instances.  topoptlab solves one
 * problem at a time and loops over load cases (solve_linsystem.py:341-352,
 * topology_optimization.py:463-505); these calls take a whole batch. ---- */
typedef struct {
  int32_t nstreams;    /* concurrent HIP streams (<= 0: 4) */
  int32_t chunk;       /* PCG iterations per host check (rounded up to even; <= 0: 8) */
  int32_t use_graphs;  /* 1: the chunk of PCG iterations between host checks is captured as a
                          hipGraph and replayed (per instance, or the whole lockstep chunk) */
  int32_t lockstep;    /* 1: all systems iterate together, one launch per kernel and place
                          in the cycle covering every system (2D handles with GMG V(1,1);
                          otherwise the per-instance streams are used) */
} tl_batch_opts;

/* Solve K_ff u = f for nrhs[i] right-hand sides on each of n handles
 * (density and, for GMG, tl_setup_mg already done).  f[i], u[i]: device
 * float64 [nrhs[i]][ndof_i]; infos: sum(nrhs) entries, instance-major.
 * Instances run concurrently on nstreams streams; the right-hand sides of
 * one instance run one after another; every solve is the tl_solve PCG with
 * the same arithmetic.  A handle may appear once.  TL_ENOCONV when any
 * right-hand side did not converge (see its info.status). */
int tl_batch_solve(int32_t n, tl_problem* const* handles, const double* const* f,
                   double* const* u, const int32_t* nrhs, const tl_solve_opts* opts,
                   const tl_batch_opts* bopts, tl_solve_info* infos);

/* FP64 fine-grid operator of 3D H8 elasticity: kind 0 (the default) runs
 * the element contraction on vector FMAs, kind 1 on v_mfma_f64_16x16x4f64
 * (column walk); same operator, different summation order (no counterpart in
 * topoptlab: the choice of kernel for K.u of fem.py:141-178). */
int tl_set_fp64_operator(tl_problem* p, int32_t kind);

/* 2D operator workgroup shape: ty = 0 (a line of 256 consecutive nodes and
 * a halo of 2*ny + 2 nodes) or a (256/ty) x ty node tile, ty in {16, 32,
 * 64}; per precision TL_PREC_FP64 / TL_PREC_FP32.  Every shape gives the same
 * arithmetic (same sums in the same order). */
int tl_set_tile2d(tl_problem* p, int32_t prec, int32_t ty);
/* Time every shape (reps launches each, FP64 and FP32 products) on this
 * grid, keep the fastest in the handle and in a process-wide table keyed by
 * (nx, ny, dofs per node); later calls for the same key reuse the table.
 * tile_out[2] (may be NULL): chosen ty for FP64, FP32; us_out[8] (may be
 * NULL): microseconds per launch, [fp64 ty=0,16,32,64, fp32 ...].  Grids
 * with one shape (3D, or a non-pattern element matrix) return zeros. */
int tl_autotune(tl_problem* p, int32_t reps, int32_t* tile_out, double* us_out);
void tl_autotune_clear(void);

/* Compliance and sensitivities summed over ncases load cases
 * (objectives.py:7-55, topology_optimization.py:463-505).
 * u: device [ncases][ndof]; xPhys, dc: device nel; obj: HOST scalar. */
int tl_compliance_sens(tl_problem* p, const double* u, int32_t ncases,
                       const double* xPhys, double* obj, double* dc);

/* Distance filter of filter/matrix_filter.py:101-232 applied as a stencil. */
int tl_filter_create(int32_t ndim, int32_t nelx, int32_t nely, int32_t nelz,
                     double rmin, void* stream, tl_filter** out);
void tl_filter_destroy(tl_filter* f);
int tl_filter_set_stream(tl_filter* f, void* stream);
/* Hs (row sums of H): device nel. */
int tl_filter_hs(tl_filter* f, double* Hs);
/* mode 0: out = H in / Hs            (topology_optimization.py:640)
 * mode 1: out = H (in / Hs)          (topology_optimization.py:557-558)
 * mode 2: out = H(x*in)/Hs/max(1e-3,x) (topology_optimization.py:546-547)
 * mode 3: out = H in                 (the product H @ v itself) */
int tl_filter_apply(tl_filter* f, const double* in, double* out, int32_t mode,
                    const double* x);

/* Node/element averaging TF of the Helmholtz PDE filter
 * (filter/helmholtz_filter.py:223-227, 2D): mode 0: out (nodes) = TF in
 * (elements), each node gets 1/4 of every element around it; mode 1: out
 * (elements) = TF^T in (nodes), 1/4 of each of the element's nodes.  The
 * filter is TF^T KF^-1 TF with KF the screened-Poisson operator, solved by
 * tl_solve on a tl_problem made from that element matrix. */
int tl_helmholtz_transfer(int32_t nelx, int32_t nely, const double* in, double* out,
                          int32_t mode, void* stream);

/* Optimality-criteria bisection of optimizer/optimality_criterion.py:9-79.
 * x, dc, dv, xnew: device n; g, gt: host scalars.  el_flags (device n int32,
 * may be NULL): 1 = passive element (xnew 0), 2 = active element (xnew 1),
 * applied inside every bisection step as at optimality_criterion.py:68-71.
 * The reduction scratch is allocated on the first call per device and
 * shared; concurrent host threads are serialised by a lock. */
int tl_oc_update(int64_t n, const double* x, const double* dc,
                 const double* dv, const int32_t* el_flags, double g, double move,
                 double l1, double l2, double* xnew, double* gt, void* stream);

/* tl_oc_update without any device allocation.  No counterpart in topoptlab:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as tl_oc_update; the reduction partials and
 * the bisection state live in a statically allocated device array of
 * 41,216 bytes per device (held until the process ends), whose size does not
 * depend on n (the grid is capped at 1024 blocks), so no n is too large for
 * it.  Its address is looked up once per device on the first call.  Every
 * call on a device shares that device's array: calls that overlap on one
 * device (several host threads or streams) race and must be serialised by
 * the caller. */
int tl_oc_update_noalloc(int64_t n, const double* x, const double* dc,
                         const double* dv, const int32_t* el_flags, double g, double move,
                         double l1, double l2, double* xnew, double* gt, void* stream);

/* Instrumentation (no counterpart in the reference): while enabled, every
 * launch of the fine-grid operator kernel and of the first coarse level's
 * stencil kernel is bracketed by HIP events on the handle's stream.  kind 0:
 * PCG products (K p, true residual), kind 1: V-cycle residual, kind 2:
 * V-cycle Jacobi sweeps, kind 3: level-1 residual, kind 4: level-1 Jacobi
 * sweeps.  tl_profile_read synchronises, returns the summed kernel time and
 * launch count of a kind and clears that kind. */
int tl_profile(tl_problem* p, int32_t enable);
int tl_profile_read(tl_problem* p, int32_t kind, double* total_ms, int64_t* launches);

/* ---- Learned coarse-grid correction (no counterpart in topoptlab). ----
 * Layers of a small 3D convolutional network on node grids.  Activations:
 * fp32, channel-major [c][z][x][y]; weights: fp16 [cout][cin][27], tap
 * t = (dz+1)*9 + (dx+1)*3 + (dy+1); fp32 accumulation; zero padding.
 * Stride 2 maps 2m+1 nodes to m+1 per axis.  nz, nx, ny are the input
 * grid's node counts; all pointers are device pointers on `stream`.
 * out[co] = relu?(bias[co] + sum_ci w[co][ci] * in[ci]), 1 <= cin <= 64. */
int tl_nn_conv3d(int32_t cin, int32_t cout, int32_t nz, int32_t nx, int32_t ny, int32_t stride,
                 int32_t relu, const float* in, const uint16_t* w16, const float* bias,
                 float* out, void* stream);
/* gin (+)= d(sum gout * conv)/d in, for gout already masked by the ReLU. */
int tl_nn_conv3d_bwd_data(int32_t cin, int32_t cout, int32_t nz, int32_t nx, int32_t ny,
                          int32_t stride, const float* gout, const uint16_t* w16, float* gin,
                          int32_t accumulate, void* stream);
/* gw[co][ci][t], gb[co]: weight and bias gradients (fixed-order sums). */
int tl_nn_conv3d_bwd_weight(int32_t cin, int32_t cout, int32_t nz, int32_t nx, int32_t ny,
                            int32_t stride, const float* in, const float* gout, float* gw,
                            float* gb, void* stream);
/* g[i] = 0 where y[i] <= 0 (ReLU backward on the layer output). */
int tl_nn_relu_mask(int64_t n, const float* y, float* g, void* stream);
/* Trilinear node interpolation from a (cz, cx, cy) grid to (2cz-1, 2cx-1,
 * 2cy-1), channel-wise, and its transpose; accumulate: add into the output. */
int tl_nn_upsample(int32_t c, int32_t cz, int32_t cx, int32_t cy, const float* coarse,
                   float* fine, int32_t accumulate, void* stream);
int tl_nn_upsample_t(int32_t c, int32_t cz, int32_t cx, int32_t cy, const float* gfine,
                     float* gcoarse, int32_t accumulate, void* stream);
/* Mean squared error: grad = 2 (pred - target)/n, *loss (device double) =
 * mean (pred - target)^2; partials: 1024 doubles of scratch. */
int tl_nn_mse(int64_t n, const float* pred, const float* target, float* grad,
              double* partials, double* loss, void* stream);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on fp32 master weights w with
 * moments m, v; w16 (may be NULL) receives the fp16 copy; step >= 1. */
int tl_nn_adam(int64_t n, float* w, const float* g, float* m, float* v, uint16_t* w16,
               int32_t step, float lr, void* stream);

/* Inputs of the network on the level-1 node grid of a 3D H8 problem (even
 * element counts), channel-major [6][nz1][nx1][ny1]: mean element density
 * at the node, the load restricted by full weighting times fscale (3),
 * fixed-node indicator, ones.  xPhys: nel, f: ndof (device, float64). */
int tl_nn_features(tl_problem* p, const double* xPhys, const double* f, double fscale,
                   float* out);
/* out[c][node1] = scale * u at the coinciding fine node (training targets). */
int tl_nn_inject(tl_problem* p, const double* u, double scale, float* out);
/* x (ndof, float64) = trilinear prolongation of a level-1 field y[3][...]
 * (channel-major), zero on fixed DOFs. */
int tl_nn_prolong(tl_problem* p, const float* y, double* x);
/* Energy-optimal scaling of an initial guess: x *= alpha with
 * alpha = (f.x)/(x.K x), which minimises the energy norm of the error over
 * the line through x; the result is never worse than the zero guess.
 * work: ndof float64; *alpha returned on the host (synchronises). */
int tl_energy_scale(tl_problem* p, const double* f, double* x, double* work, double* alpha);

/* Method of moving asymptotes for one constraint: the mmapy.mmasub step of
 * topology_optimization.py:591-607 with the defaults of
 * optimizer/mma_utils.py:8-60 (Svanberg's published algorithm; mmapy is not
 * vendored).  m = 1: a, c, d are scalars. */
typedef struct {
  double move, asyinit, asydecr, asyincr, asymin, asymax, raa0, albefa;
  double a0, a, c, d;
} tl_mma_opts;
/* x (device n): xval on entry, xmma on exit.  low, upp (device n): the
 * asymptotes, updated.  xold1, xold2 (device n): the previous two iterates,
 * read when iter > 2 (the asymptotes restart at x -/+ asyinit*(xmax-xmin)
 * for iter <= 2).  xmin, xmax (device n) may be NULL: 0 and 1.  df0dx, dfdx
 * (device n): objective and constraint gradients; fval: constraint value.
 * *newton_its (may be NULL): Newton steps of the subproblem.  Synchronises
 * the stream (the subproblem's scalars live on the host). */
int tl_mma_update(int64_t n, double* x, const double* xold1, const double* xold2,
                  const double* xmin, const double* xmax, const double* df0dx, double fval,
                  const double* dfdx, double* low, double* upp, int32_t iter,
                  const tl_mma_opts* opts, void* stream, int32_t* newton_its);

/* mean(a) and max|a-b| of device vectors (topology_optimization.py:516,654). */
int tl_mean_maxdiff(int64_t n, const double* a, const double* b, double* mean,
                    double* maxdiff, void* stream);

#ifdef __cplusplus
}
#endif
#endif
