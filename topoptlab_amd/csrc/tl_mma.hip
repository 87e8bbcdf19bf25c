// Method of moving asymptotes on the GPU: one MMA design update for the
// single (volume) constraint of topoptlab's compliance problem.  topoptlab
// calls mmapy.mmasub (topology_optimization.py:591-607) with the defaults
// of optimizer/mma_utils.py:8-60; mmapy implements Svanberg's published
// algorithm (IJNME 24, 1987; "MMA and GCMMA", 2007 note), restated here:
//
//  * asymptotes: x -/+ asyinit*(xmax-xmin) for iter <= 2, else widened or
//    narrowed by asyincr / asydecr from the sign of (x-xold1)(xold1-xold2)
//    and clamped to [asymin, asymax]*(xmax-xmin) from x;
//  * bounds alfa/beta from the asymptotes (albefa) and the move limit;
//  * the convex separable approximation p0, q0, P, Q, b;
//  * the subproblem with artificial variables y, z by the primal-dual
//    interior-point method: relaxed KKT residual, epsi from 1 down to 1e-7
//    (x0.1), Newton steps on the reduced (lam, z) system, step length that
//    keeps every slack positive (factor 1.01), halved until the residual
//    norm decreases (at most 50 halvings, 200 Newton steps per epsi).
//
// Elementwise work and the global sums / maxima run on the device; the few
// scalars of the m = 1 system (y, z, lam, mu, zet, s) live on the host,
// which reads 3-4 reduced values per Newton step.
#include "tl_internal.hpp"

#include <algorithm>
#include <cmath>

namespace tl {
namespace {

constexpr int BS = 256;
constexpr int NR = 4;  // reduced values per pass

struct MmaWork {
  DevBuf vec, partials;
  // the folded scalars, written by the fold kernel straight into pinned,
  // coherent host memory: the host reads them after the stream synchronises,
  // with no copy in between (one per fold, ~90 folds per update)
  double* hout = nullptr;
  int64_t n = 0;
  int ensure(int64_t nn) {
    if (nn > n) {
      TL_TRY(vec.alloc(sizeof(double) * size_t(nn) * 15));
      n = nn;
    }
    if (!partials.ptr) TL_TRY(partials.alloc(sizeof(double) * NR * kMaxRedBlocks));
    if (!hout) TL_HIP(hipHostMalloc(reinterpret_cast<void**>(&hout), sizeof(double) * NR,
                                    hipHostMallocCoherent));
    return TL_OK;
  }
};

MmaWork& work_for_device() {
  static MmaWork w[64];
  int dev = 0;
  (void)hipGetDevice(&dev);
  return w[dev & 63];
}

// Block reduction of NR values; bit k of maxmask: value k is a maximum.
__device__ inline void block_reduce(double (&v)[NR], unsigned maxmask, double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NR; ++k) v[k] = ((maxmask >> k) & 1u) ? wave_max(v[k]) : wave_sum(v[k]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NR; ++k) sh[k * nw + wid] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    double r = sh[k * nw];
    for (int w = 1; w < nw; ++w)
      r = ((maxmask >> k) & 1u) ? fmax(r, sh[k * nw + w]) : r + sh[k * nw + w];
    v[k] = r;
  }
  __syncthreads();
}

__device__ inline void store_red(const double (&v)[NR], double* partials) {
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NR; ++k) partials[k * kMaxRedBlocks + blockIdx.x] = v[k];
}

__global__ __launch_bounds__(BS) void mma_fold_kernel(const double* __restrict__ partials,
                                                      unsigned nb, unsigned maxmask,
                                                      double* __restrict__ out) {
  __shared__ double sh[NR * (BS / 64)];
  double v[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const bool mx = (maxmask >> k) & 1u;
    double r = mx ? -INFINITY : 0.0;
    for (unsigned b = threadIdx.x; b < nb; b += BS) {
      const double p = partials[k * kMaxRedBlocks + b];
      r = mx ? fmax(r, p) : r + p;
    }
    v[k] = r;
  }
  block_reduce(v, maxmask, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NR; ++k) out[k] = v[k];
}

struct MmaVec {
  double *alfa, *beta, *p0, *q0, *P, *Q;
  double *x[2], *xsi[2], *eta[2];
  double *dx, *dxsi, *deta;
};

// asymptotes, bounds, approximation; sum of P/(upp-x) + Q/(x-low)
__global__ __launch_bounds__(BS) void mma_prep_kernel(
    int64_t n, const double* __restrict__ xval, const double* __restrict__ xold1,
    const double* __restrict__ xold2, const double* __restrict__ xmin_a,
    const double* __restrict__ xmax_a, const double* __restrict__ df0dx,
    const double* __restrict__ dfdx, double* __restrict__ low, double* __restrict__ upp, int iter,
    tl_mma_opts o, MmaVec w, double* __restrict__ partials) {
  __shared__ double sh[NR * (BS / 64)];
  double red[NR] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BS) {
    const double xv = xval[i];
    const double xmn = xmin_a ? xmin_a[i] : 0.0, xmx = xmax_a ? xmax_a[i] : 1.0;
    const double rng = xmx - xmn;
    double lo, up;
    if (iter <= 2) {
      lo = xv - o.asyinit * rng;
      up = xv + o.asyinit * rng;
    } else {
      const double zzz = (xv - xold1[i]) * (xold1[i] - xold2[i]);
      const double factor = zzz > 0 ? o.asyincr : (zzz < 0 ? o.asydecr : 1.0);
      lo = xv - factor * (xold1[i] - low[i]);
      up = xv + factor * (upp[i] - xold1[i]);
      lo = fmin(fmax(lo, xv - o.asymax * rng), xv - o.asymin * rng);
      up = fmax(fmin(up, xv + o.asymax * rng), xv + o.asymin * rng);
    }
    low[i] = lo;
    upp[i] = up;
    w.alfa[i] = fmax(fmax(lo + o.albefa * (xv - lo), xv - o.move * rng), xmn);
    w.beta[i] = fmin(fmin(up - o.albefa * (up - xv), xv + o.move * rng), xmx);
    const double xmamiinv = 1.0 / fmax(rng, 1e-5);
    const double ux1 = up - xv, xl1 = xv - lo;
    const double ux2 = ux1 * ux1, xl2 = xl1 * xl1;
    const double g0 = df0dx[i], g1 = dfdx[i];
    double p0 = fmax(g0, 0.0), q0 = fmax(-g0, 0.0);
    const double pq0 = 0.001 * (p0 + q0) + o.raa0 * xmamiinv;
    p0 = (p0 + pq0) * ux2;
    q0 = (q0 + pq0) * xl2;
    double P = fmax(g1, 0.0), Q = fmax(-g1, 0.0);
    const double PQ = 0.001 * (P + Q) + o.raa0 * xmamiinv;
    P = (P + PQ) * ux2;
    Q = (Q + PQ) * xl2;
    w.p0[i] = p0; w.q0[i] = q0; w.P[i] = P; w.Q[i] = Q;
    red[0] += P / ux1 + Q / xl1;
    // subproblem start: x = (alfa+beta)/2, xsi = max(1/(x-alfa), 1), eta likewise
    const double xs = 0.5 * (w.alfa[i] + w.beta[i]);
    w.x[0][i] = xs;
    w.xsi[0][i] = fmax(1.0 / (xs - w.alfa[i]), 1.0);
    w.eta[0][i] = fmax(1.0 / (w.beta[i] - xs), 1.0);
  }
  block_reduce(red, 0u, sh);
  store_red(red, partials);
}

// Residual of the relaxed KKT system, vector part, at x = x0 + steg*dx (and
// xsi, eta likewise) written to set `to` (steg = 0: no update, set `from`).
// Reduced: [0] sum of squares, [1] max |.|, [2] gvec = sum P/ux + Q/xl.
__global__ __launch_bounds__(BS) void mma_resid_kernel(int64_t n, MmaVec w, int from, int to,
                                                       double steg, double lam, double epsi,
                                                       const double* __restrict__ low,
                                                       const double* __restrict__ upp,
                                                       double* __restrict__ partials) {
  __shared__ double sh[NR * (BS / 64)];
  double red[NR] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BS) {
    double x = w.x[from][i], xsi = w.xsi[from][i], eta = w.eta[from][i];
    if (to != from) {
      x += steg * w.dx[i];
      xsi += steg * w.dxsi[i];
      eta += steg * w.deta[i];
      w.x[to][i] = x; w.xsi[to][i] = xsi; w.eta[to][i] = eta;
    }
    const double ux1 = upp[i] - x, xl1 = x - low[i];
    const double plam = w.p0[i] + w.P[i] * lam, qlam = w.q0[i] + w.Q[i] * lam;
    const double rex = plam / (ux1 * ux1) - qlam / (xl1 * xl1) - xsi + eta;
    const double rexsi = xsi * (x - w.alfa[i]) - epsi;
    const double reeta = eta * (w.beta[i] - x) - epsi;
    red[0] += rex * rex + rexsi * rexsi + reeta * reeta;
    red[1] = fmax(red[1], fmax(fabs(rex), fmax(fabs(rexsi), fabs(reeta))));
    red[2] += w.P[i] / ux1 + w.Q[i] / xl1;
  }
  block_reduce(red, 2u, sh);
  store_red(red, partials);
}

// Newton system sums at set `cur`: [0] gvec, [1] sum GG delx/diagx, [2] sum GG^2/diagx
__global__ __launch_bounds__(BS) void mma_newton_kernel(int64_t n, MmaVec w, int cur, double lam,
                                                        double epsi, const double* __restrict__ low,
                                                        const double* __restrict__ upp,
                                                        double* __restrict__ partials) {
  __shared__ double sh[NR * (BS / 64)];
  double red[NR] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BS) {
    const double x = w.x[cur][i], xsi = w.xsi[cur][i], eta = w.eta[cur][i];
    const double ux1 = upp[i] - x, xl1 = x - low[i];
    const double ux2 = ux1 * ux1, xl2 = xl1 * xl1;
    const double plam = w.p0[i] + w.P[i] * lam, qlam = w.q0[i] + w.Q[i] * lam;
    const double GG = w.P[i] / ux2 - w.Q[i] / xl2;
    const double xa = x - w.alfa[i], bx = w.beta[i] - x;
    const double delx = plam / ux2 - qlam / xl2 - epsi / xa + epsi / bx;
    const double diagx = 2.0 * (plam / (ux1 * ux2) + qlam / (xl1 * xl2)) + xsi / xa + eta / bx;
    red[0] += w.P[i] / ux1 + w.Q[i] / xl1;
    red[1] += GG * (delx / diagx);
    red[2] += GG * GG / diagx;
  }
  block_reduce(red, 0u, sh);
  store_red(red, partials);
}

// Newton direction of x, xsi, eta for given dlam; maxima of the step ratios:
// [0] -1.01 dx/(x-alfa), [1] 1.01 dx/(beta-x), [2] -1.01 dxsi/xsi, [3] -1.01 deta/eta
__global__ __launch_bounds__(BS) void mma_dir_kernel(int64_t n, MmaVec w, int cur, double lam,
                                                     double dlam, double epsi,
                                                     const double* __restrict__ low,
                                                     const double* __restrict__ upp,
                                                     double* __restrict__ partials) {
  __shared__ double sh[NR * (BS / 64)];
  double red[NR] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BS) {
    const double x = w.x[cur][i], xsi = w.xsi[cur][i], eta = w.eta[cur][i];
    const double ux1 = upp[i] - x, xl1 = x - low[i];
    const double ux2 = ux1 * ux1, xl2 = xl1 * xl1;
    const double plam = w.p0[i] + w.P[i] * lam, qlam = w.q0[i] + w.Q[i] * lam;
    const double GG = w.P[i] / ux2 - w.Q[i] / xl2;
    const double xa = x - w.alfa[i], bx = w.beta[i] - x;
    const double delx = plam / ux2 - qlam / xl2 - epsi / xa + epsi / bx;
    const double diagx = 2.0 * (plam / (ux1 * ux2) + qlam / (xl1 * xl2)) + xsi / xa + eta / bx;
    const double dx = -delx / diagx - (GG * dlam) / diagx;
    const double dxsi = -xsi + epsi / xa - (xsi * dx) / xa;
    const double deta = -eta + epsi / bx + (eta * dx) / bx;
    w.dx[i] = dx; w.dxsi[i] = dxsi; w.deta[i] = deta;
    red[0] = fmax(red[0], -1.01 * dx / xa);
    red[1] = fmax(red[1], 1.01 * dx / bx);
    red[2] = fmax(red[2], -1.01 * dxsi / xsi);
    red[3] = fmax(red[3], -1.01 * deta / eta);
  }
  block_reduce(red, 15u, sh);
  store_red(red, partials);
}

__global__ void mma_copy_kernel(int64_t n, const double* __restrict__ a, double* __restrict__ b) {
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BS)
    b[i] = a[i];
}

}  // namespace
}  // namespace tl

using namespace tl;

extern "C" int tl_mma_update(int64_t n, double* x, const double* xold1, const double* xold2,
                             const double* xmin, const double* xmax, const double* df0dx,
                             double fval, const double* dfdx, double* low, double* upp,
                             int32_t iter, const tl_mma_opts* o, void* stream,
                             int32_t* newton_its) {
  TL_ARG(n >= 1 && x && df0dx && dfdx && low && upp && o, "null argument or empty design");
  TL_ARG(iter <= 2 || (xold1 && xold2), "xold1 and xold2 are needed for iter > 2");
  TL_ARG(o->move > 0.0 && o->asyinit > 0.0 && o->albefa > 0.0 && o->albefa < 1.0,
         "MMA parameters out of range");
  MmaWork& W = work_for_device();
  TL_TRY(W.ensure(n));
  hipStream_t s = static_cast<hipStream_t>(stream);
  double* base = W.vec.as<double>();
  MmaVec w;
  double** slots[15] = {&w.alfa, &w.beta, &w.p0, &w.q0, &w.P, &w.Q, &w.x[0], &w.x[1],
                        &w.xsi[0], &w.xsi[1], &w.eta[0], &w.eta[1], &w.dx, &w.dxsi, &w.deta};
  for (int k = 0; k < 15; ++k) *slots[k] = base + size_t(k) * size_t(n);
  const unsigned nb = unsigned(std::min<int64_t>((n + BS - 1) / BS, 1024));
  double* partials = W.partials.as<double>();
  double* out = W.hout;
  double h[NR];
  auto fold = [&](unsigned maxmask) -> int {
    hipLaunchKernelGGL(mma_fold_kernel, dim3(1), dim3(BS), 0, s, partials, nb, maxmask, out);
    TL_CHECK_LAUNCH();
    TL_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < NR; ++k) h[k] = static_cast<volatile double*>(out)[k];
    return TL_OK;
  };
  // approximation and the subproblem's start point
  hipLaunchKernelGGL(mma_prep_kernel, dim3(nb), dim3(BS), 0, s, n, x, xold1, xold2, xmin, xmax,
                     df0dx, dfdx, low, upp, int(iter), *o, w, partials);
  TL_CHECK_LAUNCH();
  TL_TRY(fold(0u));
  const double b = h[0] - fval;
  const double a0 = o->a0, a = o->a, c = o->c, d = o->d;
  double y = 1.0, z = 1.0, lam = 1.0, mu = std::max(1.0, 0.5 * c), zet = 1.0, sl = 1.0;
  int cur = 0;
  int its = 0;
  // scalar part of the residual (m = 1) for gvec
  auto scal_res = [&](double gvec, double epsi, double yy, double zz, double ll, double mm,
                      double ze, double ss, double& sq, double& mx) {
    const double r[6] = {c + d * yy - mm - ll, a0 - ze - a * ll, gvec - a * zz - yy + ss - b,
                         mm * yy - epsi, ze * zz - epsi, ll * ss - epsi};
    sq = 0.0;
    mx = 0.0;
    for (double v : r) { sq += v * v; mx = std::max(mx, std::fabs(v)); }
  };
  for (double epsi = 1.0; epsi > 1e-7; epsi *= 0.1) {
    hipLaunchKernelGGL(mma_resid_kernel, dim3(nb), dim3(BS), 0, s, n, w, cur, cur, 0.0, lam, epsi,
                       low, upp, partials);
    TL_CHECK_LAUNCH();
    TL_TRY(fold(2u));
    double sq, mx;
    scal_res(h[2], epsi, y, z, lam, mu, zet, sl, sq, mx);
    double residunorm = std::sqrt(h[0] + sq);
    double residumax = std::max(h[1], mx);
    for (int ittt = 0; residumax > 0.9 * epsi && ittt < 200; ++ittt) {
      ++its;
      hipLaunchKernelGGL(mma_newton_kernel, dim3(nb), dim3(BS), 0, s, n, w, cur, lam, epsi, low,
                         upp, partials);
      TL_CHECK_LAUNCH();
      TL_TRY(fold(0u));
      const double gvec = h[0], S1 = h[1], S2 = h[2];
      const double dely = c + d * y - lam - epsi / y;
      const double delz = a0 - a * lam - epsi / z;
      const double dellam = gvec - a * z - y - b + epsi / lam;
      const double diagy = d + mu / y;
      const double diaglamyi = sl / lam + 1.0 / diagy;
      // reduced system [[Alam, a], [a, -zet/z]] [dlam, dz] = [blam, delz]
      const double blam = dellam + dely / diagy - S1;
      const double Alam = diaglamyi + S2;
      const double A22 = -zet / z;
      const double det = Alam * A22 - a * a;
      const double dlam = (blam * A22 - a * delz) / det;
      const double dz = (Alam * delz - a * blam) / det;
      const double dy = -dely / diagy + dlam / diagy;
      const double dmu = -mu + epsi / y - (mu * dy) / y;
      const double dzet = -zet + epsi / z - zet * dz / z;
      const double ds = -sl + epsi / lam - (sl * dlam) / lam;
      hipLaunchKernelGGL(mma_dir_kernel, dim3(nb), dim3(BS), 0, s, n, w, cur, lam, dlam, epsi,
                         low, upp, partials);
      TL_CHECK_LAUNCH();
      TL_TRY(fold(15u));
      double stm = std::max({h[0], h[1], h[2], h[3], 1.0});
      const double sv[6] = {y, z, lam, mu, zet, sl}, dv[6] = {dy, dz, dlam, dmu, dzet, ds};
      for (int k = 0; k < 6; ++k) stm = std::max(stm, -1.01 * dv[k] / sv[k]);
      double steg = 1.0 / stm;
      const double y0 = y, z0 = z, l0 = lam, m0 = mu, ze0 = zet, s0 = sl;
      double resinew = 2.0 * residunorm;
      const int trial = 1 - cur;
      for (int itto = 0; resinew > residunorm && itto < 50; ++itto) {
        y = y0 + steg * dy; z = z0 + steg * dz; lam = l0 + steg * dlam;
        mu = m0 + steg * dmu; zet = ze0 + steg * dzet; sl = s0 + steg * ds;
        hipLaunchKernelGGL(mma_resid_kernel, dim3(nb), dim3(BS), 0, s, n, w, cur, trial, steg, lam,
                           epsi, low, upp, partials);
        TL_CHECK_LAUNCH();
        TL_TRY(fold(2u));
        scal_res(h[2], epsi, y, z, lam, mu, zet, sl, sq, mx);
        resinew = std::sqrt(h[0] + sq);
        residumax = std::max(h[1], mx);
        steg *= 0.5;
      }
      cur = trial;
      residunorm = resinew;
    }
  }
  hipLaunchKernelGGL(mma_copy_kernel, dim3(nb), dim3(BS), 0, s, n, w.x[cur], x);
  TL_CHECK_LAUNCH();
  TL_HIP(hipStreamSynchronize(s));
  if (newton_its) *newton_its = its;
  return TL_OK;
}
