// Density filter, compliance + sensitivities, OC update and loop reductions.
//
// Filter: filter/matrix_filter.py:101-232 builds H_ij = max(0, rr - |i-j|)
// over the (2*ceil(rmin)-1)^ndim box (rr = ||rmin||_2/sqrt(ndim)) and applies
// it as a CSC product (topology_optimization.py:546-558, 640).  Here H is
// never stored: each element sums its box in increasing column order
// (z, x, y lexicographic == increasing linear index nely*x + y + nelx*nely*z),
// the order in which scipy's CSC product accumulates a row, with explicit
// round-to-nearest multiply/add (no FMA contraction) so results can match the
// CPU bit for bit.
#include "tl_internal.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

namespace tl {
namespace {

constexpr int BS = 256;
constexpr int kMaxBox = 13 * 13 * 13;  // rmin up to 7 in 3D

__device__ inline double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ inline double add_rn(double a, double b) { return __dadd_rn(a, b); }

}  // namespace
}  // namespace tl

struct tl_filter {
  int dim = 2, nelx = 0, nely = 0, nelz = 0, R = 0;
  int64_t nel = 0;
  double rr = 0.0;
  hipStream_t stream = nullptr;
  tl::DevBuf w;    // (2R+1)^dim weights, z,x,y order
  tl::DevBuf hs;   // row sums
  tl::DevBuf tmp;  // nel scratch
  ~tl_filter() { w.release(); hs.release(); tmp.release(); }
};

namespace tl {
namespace {

struct FiltArgs {
  const double* w;
  int dim, nelx, nely, nelz, R;
  int64_t nel;
};

// out_e = sum_{j in box(e), inside} w_ej * v_j   (v == nullptr: v_j = 1)
__device__ inline double box_sum(const FiltArgs& a, int64_t e, const double* __restrict__ v) {
  // 32-bit divisions (fewer than 2^31 elements)
  const uint32_t eu = uint32_t(e), ly = uint32_t(a.nely), lxy = uint32_t(a.nelx) * ly;
  int x, y, z;
  if (a.dim == 2) {
    z = 0;
  } else {
    z = int(eu / lxy);
  }
  const uint32_t rest = eu - uint32_t(z) * lxy;
  x = int(rest / ly); y = int(rest - uint32_t(x) * ly);
  const int W = 2 * a.R + 1;
  const int zr = a.dim == 3 ? a.R : 0;
  double acc = 0.0;
  for (int dz = -zr; dz <= zr; ++dz) {
    const int zz = z + dz;
    if (zz < 0 || zz >= (a.dim == 3 ? a.nelz : 1)) continue;
    for (int dx = -a.R; dx <= a.R; ++dx) {
      const int xx = x + dx;
      if (xx < 0 || xx >= a.nelx) continue;
      const int wbase = ((dz + zr) * W + (dx + a.R)) * W;
      const int64_t jbase = (int64_t(zz) * a.nelx + xx) * a.nely;
      for (int dy = -a.R; dy <= a.R; ++dy) {
        const int yy = y + dy;
        if (yy < 0 || yy >= a.nely) continue;
        const double wv = a.w[wbase + dy + a.R];
        acc = add_rn(acc, v ? mul_rn(wv, v[jbase + yy]) : wv);
      }
    }
  }
  return acc;
}

__global__ void filter_hs_kernel(FiltArgs a, double* __restrict__ hs) {
  const int64_t e = blockIdx.x * int64_t(BS) + threadIdx.x;
  if (e < a.nel) hs[e] = box_sum(a, e, nullptr);
}

// mode 0: out = (H in)/Hs ; mode 1: in pre-divided by Hs ; mode 2: see below
__global__ void filter_apply_kernel(FiltArgs a, const double* __restrict__ v,
                                    const double* __restrict__ hs, const double* __restrict__ x,
                                    double* __restrict__ out, int mode) {
  const int64_t e = blockIdx.x * int64_t(BS) + threadIdx.x;
  if (e >= a.nel) return;
  const double s = box_sum(a, e, v);
  if (mode == 0) out[e] = s / hs[e];
  else if (mode == 1 || mode == 3) out[e] = s;
  else out[e] = (s / hs[e]) / fmax(0.001, x[e]);
}

__global__ void div_kernel(const double* __restrict__ a, const double* __restrict__ b,
                           double* __restrict__ o, int64_t n) {
  const int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x;
  if (i < n) o[i] = a[i] / b[i];
}

__global__ void mul_kernel(const double* __restrict__ a, const double* __restrict__ b,
                           double* __restrict__ o, int64_t n) {
  const int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x;
  if (i < n) o[i] = a[i] * b[i];
}

// ---------------------------------------------------------------------------
// compliance and sensitivities (element-centric gather)
// ---------------------------------------------------------------------------
template <int NDPN, bool D3>
__global__ __launch_bounds__(BS) void compliance_kernel(
    const double* __restrict__ u, int ncases, int64_t ndof, const double* __restrict__ xphys,
    const double* __restrict__ KE, int nelx, int nely, int nelz, double eps, double penal,
    double* __restrict__ dc, double* partials) {
  constexpr int NN = D3 ? 8 : 4;
  constexpr int ND = NN * NDPN;
  const int64_t nel = int64_t(nelx) * nely * (D3 ? nelz : 1);
  double objp = 0.0;
  for (int64_t e = blockIdx.x * int64_t(BS) + threadIdx.x; e < nel;
       e += int64_t(gridDim.x) * BS) {
    // 32-bit divisions (fewer than 2^31 elements)
    const uint32_t eu = uint32_t(e), ly = uint32_t(nely), lxy = uint32_t(nelx) * ly;
    const int ez = D3 ? int(eu / lxy) : 0;
    const uint32_t rest = eu - uint32_t(ez) * lxy;
    const int ex = int(rest / ly), ey = int(rest - uint32_t(ex) * ly);
    const int ny = nely + 1, nx = nelx + 1;
    const double xe = xphys[e];
    const double se = eps + (1.0 - eps) * pow(xe, penal);
    const double dse = penal * (1.0 - eps) * pow(xe, penal - 1.0);
    double dce = 0.0;
    for (int lc = 0; lc < ncases; ++lc) {
      const double* uc = u + int64_t(lc) * ndof;
      double ue[ND];
#pragma unroll
      for (int k = 0; k < NN; ++k) {
        const int64_t node =
            (int64_t(ez + loc_dz(k)) * nx + ex + loc_dx(k)) * ny + ey + loc_dy(k);
#pragma unroll
        for (int c = 0; c < NDPN; ++c) ue[k * NDPN + c] = uc[node * NDPN + c];
      }
      double ce = 0.0;
#pragma unroll
      for (int i = 0; i < ND; ++i) {
        double ki = 0.0;
#pragma unroll
        for (int j = 0; j < ND; ++j) ki += KE[i * ND + j] * ue[j];
        ce += ue[i] * ki;
      }
      objp += se * ce;
      dce -= dse * ce;
    }
    dc[e] = dce;
  }
  __shared__ double sh[BS / 64 + 2];
  double v[1] = {objp};
  block_sum<1>(v, sh);
  store_partials<1>(v, partials);
}

// one block: total of the preceding kernel's nb partials -> out[0]
__global__ __launch_bounds__(BS) void fold_sum_kernel(const double* __restrict__ partials,
                                                      unsigned nb, double* out) {
  __shared__ double sh[BS / 64 + 2];
  double tot[1];
  fold_partials<1>(partials, nb, tot, sh);
  if (threadIdx.x == 0) out[0] = tot[0];
}

// ---------------------------------------------------------------------------
// OC bisection (optimizer/optimality_criterion.py:61-77), device resident
// ---------------------------------------------------------------------------
struct OcState {
  double l1, l2, g, gt, move;
  int done, steps;
};

__global__ __launch_bounds__(BS) void oc_step_kernel(const double* __restrict__ x,
                                                     const double* __restrict__ dc,
                                                     const double* __restrict__ dv,
                                                     const int32_t* __restrict__ flags, int64_t n,
                                                     double* __restrict__ xnew, const OcState* st,
                                                     double* partials) {
  if (st->done) return;
  const double lmid = 0.5 * (st->l2 + st->l1);
  const double move = st->move;
  // compensated (two-sum) accumulation so that the sign decision below is
  // the sign of the exactly rounded total
  double s = 0.0, comp = 0.0;
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BS) {
    const double xi = x[i];
    const double cand = xi * sqrt(-dc[i] / dv[i] / lmid);
    double xn = fmax(0.0, fmax(xi - move, fmin(1.0, fmin(xi + move, cand))));
    if (flags) {  // passive (1) elements are 0, active (2) elements are 1 (:68-71)
      const int32_t fl = flags[i];
      xn = fl == 1 ? 0.0 : (fl == 2 ? 1.0 : xn);
    }
    xnew[i] = xn;
    // explicit rounding ops: a contracted FMA would break the two-sum
    const double term = __dmul_rn(dv[i], __dsub_rn(xn, xi));
    const double t = __dadd_rn(s, term);
    const double bp = __dsub_rn(t, s);
    comp = __dadd_rn(comp, __dadd_rn(__dsub_rn(s, __dsub_rn(t, bp)), __dsub_rn(term, bp)));
    s = t;
  }
  __shared__ double sh[2 * (BS / 64) + 4];
  double v[2] = {s, comp};
  block_sum<2>(v, sh);
  store_partials<2>(v, partials);
}

// one block: gt = g + sum dv*(xnew - x); bisect (optimality_criterion.py:72-77)
__global__ __launch_bounds__(BS) void oc_fold_kernel(const double* __restrict__ partials,
                                                     unsigned nb, OcState* st) {
  if (st->done) return;
  __shared__ double sh[2 * (BS / 64) + 4];
  double tot[2];
  fold_partials<2>(partials, nb, tot, sh);
  if (threadIdx.x == 0) {
    const double lmid = 0.5 * (st->l2 + st->l1);
    const double gt = st->g + (tot[0] + tot[1]);
    st->gt = gt;
    if (gt > 0) st->l1 = lmid;
    else st->l2 = lmid;
    st->steps += 1;
    if (!((st->l2 - st->l1) / (st->l1 + st->l2) > 1e-3)) st->done = 1;
  }
}

__global__ __launch_bounds__(BS) void mean_maxdiff_kernel(const double* __restrict__ a,
                                                          const double* __restrict__ b, int64_t n,
                                                          double* partials) {
  double s = 0.0, m = 0.0;
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BS) {
    s += a[i];
    if (b) m = fmax(m, fabs(a[i] - b[i]));
  }
  m = wave_max(m);
  __shared__ double shm[BS / 64];
  __shared__ double sh[BS / 64 + 2];
  if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = m;
  double v[1] = {s};
  block_sum<1>(v, sh);  // its barriers also publish shm
  double mb = 0.0;
  for (int w = 0; w < BS / 64; ++w) mb = fmax(mb, shm[w]);
  double pv[2] = {v[0], mb};
  store_partials<2>(pv, partials);
}

// one block: out[0] = mean, out[1] = max of the block maxima
__global__ __launch_bounds__(BS) void mean_maxdiff_fold_kernel(const double* __restrict__ partials,
                                                               unsigned nb, int64_t n,
                                                               double* out) {
  __shared__ double sh[BS / 64 + 2];
  double tot[1];
  fold_partials<1>(partials, nb, tot, sh);
  double mx = 0.0;
  for (unsigned b = threadIdx.x; b < nb; b += BS) mx = fmax(mx, partials[kMaxRedBlocks + b]);
  mx = wave_max(mx);
  __shared__ double shm[BS / 64];
  if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m2 = 0.0;
    for (int w = 0; w < BS / 64; ++w) m2 = fmax(m2, shm[w]);
    out[0] = tot[0] / double(n);
    out[1] = m2;
  }
}

unsigned nblk(int64_t n) { return unsigned(std::max<int64_t>(1, (n + BS - 1) / BS)); }

// Process-wide reduction scratch of the handle-free calls, one per device.
// The mutex is held for the whole call (allocation and use), so host threads
// that call tl_oc_update / tl_mean_maxdiff concurrently are serialised.
struct Scratch {
  DevBuf partials, st, out;
  std::mutex mu;
  int ensure() {
    if (!partials.ptr) {
      TL_TRY(partials.alloc(sizeof(double) * 2 * kMaxRedBlocks));
      TL_TRY(st.alloc(sizeof(OcState)));
      TL_TRY(out.alloc(sizeof(double) * 4));
    }
    return TL_OK;
  }
};

Scratch& scratch_for_device() {
  static Scratch s[64];
  int dev = 0;
  (void)hipGetDevice(&dev);
  return s[dev & 63];
}

// The bisection of optimality_criterion.py:61-77 on `stream`, with the
// block partials (two values, kMaxRedBlocks stride, at most kOcBlocks
// blocks) and the state in caller-provided device memory.
constexpr unsigned kOcBlocks = 1024;

int oc_bisect(int64_t n, const double* x, const double* dc, const double* dv,
              const int32_t* el_flags, double g, double move, double l1, double l2,
              double* xnew, double* gt, hipStream_t s, double* partials, OcState* st) {
  OcState h{l1, l2, g, g, move, 0, 0};
  h.done = ((l2 - l1) / (l1 + l2) > 1e-3) ? 0 : 1;
  TL_HIP(hipMemcpyAsync(st, &h, sizeof(OcState), hipMemcpyHostToDevice, s));
  const unsigned nb = unsigned(std::min<int64_t>(nblk(n), kOcBlocks));
  for (int launched = 0; launched < 400;) {
    for (int k = 0; k < 16; ++k) {
      hipLaunchKernelGGL(oc_step_kernel, dim3(nb), dim3(BS), 0, s, x, dc, dv, el_flags, n, xnew,
                         st, partials);
      hipLaunchKernelGGL(oc_fold_kernel, dim3(1), dim3(BS), 0, s, partials, nb, st);
    }
    TL_CHECK_LAUNCH();
    launched += 16;
    TL_HIP(hipMemcpyAsync(&h, st, sizeof(OcState), hipMemcpyDeviceToHost, s));
    TL_HIP(hipStreamSynchronize(s));
    if (h.done) break;
  }
  TL_ARG(h.done, "OC bisection did not terminate within 400 steps");
  *gt = h.gt;
  return TL_OK;
}

// Scratch of tl_oc_update_noalloc, statically allocated on every device the
// library runs on: the block partials of oc_step_kernel (two values per
// block at stride kMaxRedBlocks, at most kOcBlocks blocks) followed by the
// bisection state.  41,216 bytes per device, held until the process ends.
constexpr size_t kNoAllocPartialDoubles = size_t(kMaxRedBlocks) + kOcBlocks;
constexpr size_t kNoAllocStateOffset = sizeof(double) * kNoAllocPartialDoubles;
constexpr size_t kNoAllocScratchBytes = kNoAllocStateOffset + 256;
static_assert(sizeof(OcState) <= 256, "bisection state must fit its scratch slot");
static_assert(kNoAllocStateOffset % 256 == 0, "state slot must stay 256-byte aligned");
static_assert(kOcBlocks <= kMaxRedBlocks, "oc_bisect's grid must fit the partial stride");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Device address of `scratch` per device, looked up once with that device
// current (the symbol names a different copy on every device).
struct NoAllocScratch {
  std::mutex mu;
  void* ptr[64] = {};
};

NoAllocScratch& noalloc_scratch() {
  static NoAllocScratch s;
  return s;
}

int noalloc_scratch_ptr(void** out) {
  int dev = 0;
  TL_HIP(hipGetDevice(&dev));
  TL_ARG(dev >= 0 && dev < 64, "device ordinal out of range for the no-alloc scratch table");
  NoAllocScratch& S = noalloc_scratch();
  std::lock_guard<std::mutex> lock(S.mu);
  if (!S.ptr[dev]) {
    void* p = nullptr;
    TL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TL_ARG(p, "hipGetSymbolAddress returned a null scratch address");
    S.ptr[dev] = p;
  }
  *out = S.ptr[dev];
  return TL_OK;
}

}  // namespace
}  // namespace tl

using namespace tl;

__global__ void helmholtz_transfer_kernel(int nelx, int nely, const double* __restrict__ in,
                                          double* __restrict__ out, int mode) {
  const int ny = nely + 1;
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (mode == 0) {
    if (i >= int64_t(nelx + 1) * ny) return;
    const int ix = int(i / ny), iy = int(i % ny);
    double acc = 0.0;
    for (int ex = ix - 1; ex <= ix; ++ex)
      for (int ey = iy - 1; ey <= iy; ++ey)
        if (ex >= 0 && ex < nelx && ey >= 0 && ey < nely) acc += 0.25 * in[int64_t(ex) * nely + ey];
    out[i] = acc;
  } else {
    if (i >= int64_t(nelx) * nely) return;
    const int ex = int(i / nely), ey = int(i % nely);
    // local node order (BL, BR, TR, TL) = nodes (ex, ey+1), (ex+1, ey+1), (ex+1, ey), (ex, ey)
    const int64_t n1 = int64_t(ex) * ny + ey, n2 = int64_t(ex + 1) * ny + ey;
    double acc = 0.0;
    acc += 0.25 * in[n1 + 1];
    acc += 0.25 * in[n2 + 1];
    acc += 0.25 * in[n2];
    acc += 0.25 * in[n1];
    out[i] = acc;
  }
}

extern "C" {

int tl_filter_create(int32_t ndim, int32_t nelx, int32_t nely, int32_t nelz, double rmin,
                     void* stream, tl_filter** out) {
  TL_ARG(out, "null argument");
  TL_ARG(ndim == 2 || ndim == 3, "ndim must be 2 or 3");
  TL_ARG(nelx > 0 && nely > 0 && (ndim == 2 || nelz > 0), "element counts must be positive");
  TL_ARG(rmin > 0.0 && std::isfinite(rmin), "rmin must be positive");
  TL_ARG(int64_t(nelx) * nely * (ndim == 3 ? nelz : 1) < (int64_t(1) << 31),
         "grid too large for 32-bit element indexing");
  const int R = std::max(0, int(std::ceil(rmin)) - 1);
  const int W = 2 * R + 1;
  TL_ARG((ndim == 3 ? W * W * W : W * W) <= kMaxBox, "rmin too large for the filter stencil");
  auto* f = new tl_filter();
  f->dim = ndim; f->nelx = nelx; f->nely = nely; f->nelz = ndim == 3 ? nelz : 0; f->R = R;
  f->nel = int64_t(nelx) * nely * (ndim == 3 ? nelz : 1);
  f->stream = static_cast<hipStream_t>(stream);
  // rr = ||(rmin,...,rmin)||_2 / sqrt(ndim) evaluated the way numpy's norm
  // does it (a dot product accumulated with fused multiply-adds, then sqrt and
  // a division), so that the filter weights match the reference bit for bit
  double ss = rmin * rmin;
  for (int d = 1; d < ndim; ++d) ss = std::fma(rmin, rmin, ss);
  f->rr = std::sqrt(ss) / std::sqrt(double(ndim));
  const int zr = ndim == 3 ? R : 0;
  std::vector<double> w;
  for (int dz = -zr; dz <= zr; ++dz)
    for (int dx = -R; dx <= R; ++dx)
      for (int dy = -R; dy <= R; ++dy) {
        const double r = std::sqrt(double(dx * dx + dy * dy + dz * dz));
        w.push_back(std::max(f->rr - r, 0.0));
      }
  int rc;
  if ((rc = f->w.alloc(sizeof(double) * w.size())) || (rc = f->hs.alloc(sizeof(double) * f->nel)) ||
      (rc = f->tmp.alloc(sizeof(double) * f->nel))) {
    delete f;
    return rc;
  }
  hipError_t e = hipMemcpy(f->w.ptr, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    delete f;
    set_error(std::string("tl_filter_create: ") + hipGetErrorString(e));
    return TL_EHIP;
  }
  FiltArgs a{f->w.as<double>(), f->dim, f->nelx, f->nely, f->nelz, f->R, f->nel};
  hipLaunchKernelGGL(filter_hs_kernel, dim3(nblk(f->nel)), dim3(BS), 0, f->stream, a,
                     f->hs.as<double>());
  e = hipGetLastError();
  if (e != hipSuccess) {
    delete f;
    set_error(std::string("tl_filter_create: ") + hipGetErrorString(e));
    return TL_EHIP;
  }
  *out = f;
  return TL_OK;
}

void tl_filter_destroy(tl_filter* f) { delete f; }

int tl_filter_set_stream(tl_filter* f, void* stream) {
  TL_ARG(f, "null handle");
  f->stream = static_cast<hipStream_t>(stream);
  return TL_OK;
}

int tl_filter_hs(tl_filter* f, double* Hs) {
  TL_ARG(f && Hs, "null argument");
  TL_HIP(hipMemcpyAsync(Hs, f->hs.ptr, sizeof(double) * f->nel, hipMemcpyDeviceToDevice, f->stream));
  return TL_OK;
}

int tl_filter_apply(tl_filter* f, const double* in, double* out, int32_t mode, const double* x) {
  TL_ARG(f && in && out, "null argument");
  TL_ARG(mode >= 0 && mode <= 3, "filter mode must be 0, 1, 2 or 3");
  TL_ARG(mode != 2 || x, "mode 2 needs the design variables x");
  TL_ARG(in != out, "in and out must not alias");
  FiltArgs a{f->w.as<double>(), f->dim, f->nelx, f->nely, f->nelz, f->R, f->nel};
  const unsigned nb = nblk(f->nel);
  const double* v = in;
  double* t = f->tmp.as<double>();
  if (mode == 1) {
    hipLaunchKernelGGL(div_kernel, dim3(nb), dim3(BS), 0, f->stream, in, f->hs.as<double>(), t, f->nel);
    v = t;
  } else if (mode == 2) {
    hipLaunchKernelGGL(mul_kernel, dim3(nb), dim3(BS), 0, f->stream, x, in, t, f->nel);
    v = t;
  }
  hipLaunchKernelGGL(filter_apply_kernel, dim3(nb), dim3(BS), 0, f->stream, a, v,
                     f->hs.as<double>(), x, out, mode);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

int tl_compliance_sens(tl_problem* P, const double* u, int32_t ncases, const double* xPhys,
                       double* obj, double* dc) {
  TL_ARG(P && u && xPhys && dc && obj, "null argument");
  TL_ARG(ncases >= 1, "ncases must be >= 1");
  const Grid& g = P->g;
  const unsigned nb = unsigned(std::min<int64_t>(nblk(g.nel), kMaxRedBlocks));
  double* out = P->partials.as<double>() + 3 * kMaxRedBlocks;
  const bool d3 = g.dim == 3;
#define TL_COMP(N, D)                                                                          \
  hipLaunchKernelGGL((compliance_kernel<N, D>), dim3(nb), dim3(BS), 0, P->stream, u, ncases,  \
                     P->ndof, xPhys, P->op.KE64.as<double>(), g.nelx, g.nely, g.nelz, P->eps, \
                     P->penal, dc, P->partials.as<double>())
  if (d3) {
    switch (P->ndpn) {
      case 1: TL_COMP(1, true); break;
      case 3: TL_COMP(3, true); break;
      default: set_error("unsupported DOFs per node"); return TL_EINVAL;
    }
  } else {
    switch (P->ndpn) {
      case 1: TL_COMP(1, false); break;
      case 2: TL_COMP(2, false); break;
      default: set_error("unsupported DOFs per node"); return TL_EINVAL;
    }
  }
#undef TL_COMP
  TL_CHECK_LAUNCH();
  hipLaunchKernelGGL(fold_sum_kernel, dim3(1), dim3(BS), 0, P->stream, P->partials.as<double>(),
                     nb, out);
  TL_CHECK_LAUNCH();
  TL_HIP(hipMemcpyAsync(obj, out, sizeof(double), hipMemcpyDeviceToHost, P->stream));
  TL_HIP(hipStreamSynchronize(P->stream));
  return TL_OK;
}

int tl_oc_update(int64_t n, const double* x, const double* dc, const double* dv,
                 const int32_t* el_flags, double g, double move, double l1, double l2,
                 double* xnew, double* gt, void* stream) {
  TL_ARG(x && dc && dv && xnew && gt, "null argument");
  TL_ARG(n > 0, "n must be positive");
  Scratch& S = scratch_for_device();
  std::lock_guard<std::mutex> lock(S.mu);
  TL_TRY(S.ensure());
  return oc_bisect(n, x, dc, dv, el_flags, g, move, l1, l2, xnew, gt,
                   static_cast<hipStream_t>(stream), S.partials.as<double>(), S.st.as<OcState>());
}

int tl_oc_update_noalloc(int64_t n, const double* x, const double* dc, const double* dv,
                         const int32_t* el_flags, double g, double move, double l1, double l2,
                         double* xnew, double* gt, void* stream) {
  TL_ARG(x && dc && dv && xnew && gt, "null argument");
  TL_ARG(n > 0, "n must be positive");
  void* base = nullptr;
  TL_TRY(noalloc_scratch_ptr(&base));
  auto* bytes = static_cast<unsigned char*>(base);
  return oc_bisect(n, x, dc, dv, el_flags, g, move, l1, l2, xnew, gt,
                   static_cast<hipStream_t>(stream), reinterpret_cast<double*>(bytes),
                   reinterpret_cast<OcState*>(bytes + kNoAllocStateOffset));
}

int tl_mean_maxdiff(int64_t n, const double* a, const double* b, double* mean, double* maxdiff,
                    void* stream) {
  TL_ARG(a && mean && n > 0, "null argument");
  Scratch& S = scratch_for_device();
  std::lock_guard<std::mutex> lock(S.mu);
  TL_TRY(S.ensure());
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned nb = unsigned(std::min<int64_t>(nblk(n), 1024));
  hipLaunchKernelGGL(mean_maxdiff_kernel, dim3(nb), dim3(BS), 0, s, a, b, n,
                     S.partials.as<double>());
  hipLaunchKernelGGL(mean_maxdiff_fold_kernel, dim3(1), dim3(BS), 0, s, S.partials.as<double>(),
                     nb, n, S.out.as<double>());
  TL_CHECK_LAUNCH();
  double h[2];
  TL_HIP(hipMemcpyAsync(h, S.out.ptr, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
  TL_HIP(hipStreamSynchronize(s));
  *mean = h[0];
  if (maxdiff) *maxdiff = h[1];
  return TL_OK;
}


// Node <-> element averaging of the Helmholtz filter (TF of
// filter/helmholtz_filter.py:223-227, weight 1/4 per element node).
// mode 0: node[n] = 1/4 * sum of the (up to 4) elements around node n, in
// ascending element order (TF x); mode 1: el[e] = 1/4 * sum of its 4 nodes in
// local order (TF^T y).
int tl_helmholtz_transfer(int32_t nelx, int32_t nely, const double* in, double* out, int32_t mode,
                          void* stream) {
  TL_ARG(in && out && nelx >= 1 && nely >= 1, "null argument or empty grid");
  TL_ARG(mode == 0 || mode == 1, "mode must be 0 (TF x) or 1 (TF^T y)");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t n = mode == 0 ? int64_t(nelx + 1) * (nely + 1) : int64_t(nelx) * nely;
  hipLaunchKernelGGL(helmholtz_transfer_kernel, dim3(unsigned(nblk(n))), dim3(BS), 0, s, nelx,
                     nely, in, out, mode);
  TL_CHECK_LAUNCH();
  return TL_OK;
}
}  // extern "C"
