// Kernels of the learned coarse-grid correction: a small 3D convolutional
// network on node grids (no counterpart in topoptlab).  Activations are fp32,
// channel-major a[c][z][x][y]; convolution weights are fp16 w[cout][cin][27]
// (tap t = (dz+1)*9 + (dx+1)*3 + (dy+1)) with fp32 accumulation; the fp32
// master weights live with the optimiser.  Stride-2 convolutions map 2m+1
// nodes to m+1 (node-aligned, like the multigrid levels); upsampling is the
// trilinear node interpolation (the multigrid prolongation) and its
// transpose.  All reductions are fixed-order (no atomics).
#include "tl_internal.hpp"

#include <algorithm>
#include <cmath>

namespace tl {
namespace {

constexpr int BS = 256;
constexpr int CO = 8;  // output channels per thread (forward)
constexpr int CI = 8;  // input channels per thread (data gradient)

__device__ inline int tap_dz(int t) { return t / 9 - 1; }
__device__ inline int tap_dx(int t) { return (t / 3) % 3 - 1; }
__device__ inline int tap_dy(int t) { return t % 3 - 1; }

// y[co][o] = relu?(b[co] + sum_ci sum_t w[co][ci][t] * a[ci][stride*o + off_t])
__global__ __launch_bounds__(BS) void conv_fwd_kernel(int cin, int cout, int nz, int nx, int ny,
                                                      int oz, int ox, int oy, int stride, int relu,
                                                      const float* __restrict__ a,
                                                      const _Float16* __restrict__ w,
                                                      const float* __restrict__ b,
                                                      float* __restrict__ y) {
  extern __shared__ float wsh[];  // [CO][cin][27] as fp32
  const int co0 = blockIdx.y * CO;
  const int nco = min(CO, cout - co0);
  for (int i = threadIdx.x; i < CO * cin * 27; i += BS) {
    const int c = i / (cin * 27);
    wsh[i] = c < nco ? float(w[int64_t(co0) * cin * 27 + i]) : 0.f;
  }
  __syncthreads();
  const int64_t nout = int64_t(oz) * ox * oy;
  const int64_t o = int64_t(blockIdx.x) * BS + threadIdx.x;
  if (o >= nout) return;
  const int iy = int(o % oy), ix = int((o / oy) % ox), iz = int(o / (int64_t(oy) * ox));
  const int64_t plane = int64_t(nz) * nx * ny;
  float acc[CO];
#pragma unroll
  for (int k = 0; k < CO; ++k) acc[k] = 0.f;
  for (int t = 0; t < 27; ++t) {
    const int Z = stride * iz + tap_dz(t), X = stride * ix + tap_dx(t), Y = stride * iy + tap_dy(t);
    if (Z < 0 || Z >= nz || X < 0 || X >= nx || Y < 0 || Y >= ny) continue;
    const int64_t src = (int64_t(Z) * nx + X) * ny + Y;
    for (int c = 0; c < cin; ++c) {
      const float v = a[int64_t(c) * plane + src];
#pragma unroll
      for (int k = 0; k < CO; ++k) acc[k] = fmaf(wsh[(k * cin + c) * 27 + t], v, acc[k]);
    }
  }
  for (int k = 0; k < nco; ++k) {
    float v = acc[k] + b[co0 + k];
    if (relu) v = fmaxf(v, 0.f);
    y[int64_t(co0 + k) * nout + o] = v;
  }
}

// ga[ci][n] (+)= sum_co sum_t w[co][ci][t] * gy[co][o] over outputs o with
// stride*o + off_t = n
__global__ __launch_bounds__(BS) void conv_bwd_data_kernel(int cin, int cout, int nz, int nx,
                                                           int ny, int oz, int ox, int oy,
                                                           int stride, int accumulate,
                                                           const float* __restrict__ gy,
                                                           const _Float16* __restrict__ w,
                                                           float* __restrict__ ga) {
  extern __shared__ float wsh[];  // [cout][CI][27] as fp32
  const int ci0 = blockIdx.y * CI;
  const int nci = min(CI, cin - ci0);
  for (int i = threadIdx.x; i < cout * CI * 27; i += BS) {
    const int co = i / (CI * 27), k = (i / 27) % CI, t = i % 27;
    wsh[i] = k < nci ? float(w[(int64_t(co) * cin + ci0 + k) * 27 + t]) : 0.f;
  }
  __syncthreads();
  const int64_t nin = int64_t(nz) * nx * ny;
  const int64_t n = int64_t(blockIdx.x) * BS + threadIdx.x;
  if (n >= nin) return;
  const int Y = int(n % ny), X = int((n / ny) % nx), Z = int(n / (int64_t(ny) * nx));
  const int64_t nout = int64_t(oz) * ox * oy;
  float acc[CI];
#pragma unroll
  for (int k = 0; k < CI; ++k) acc[k] = 0.f;
  for (int t = 0; t < 27; ++t) {
    const int pz = Z - tap_dz(t), px = X - tap_dx(t), py = Y - tap_dy(t);
    if (stride == 2 && ((pz | px | py) & 1)) continue;
    const int iz = pz / stride, ix = px / stride, iy = py / stride;
    if (pz < 0 || px < 0 || py < 0 || iz >= oz || ix >= ox || iy >= oy) continue;
    const int64_t src = (int64_t(iz) * ox + ix) * oy + iy;
    for (int co = 0; co < cout; ++co) {
      const float g = gy[int64_t(co) * nout + src];
#pragma unroll
      for (int k = 0; k < CI; ++k) acc[k] = fmaf(wsh[(co * CI + k) * 27 + t], g, acc[k]);
    }
  }
  const int64_t plane = nin;
  for (int k = 0; k < nci; ++k) {
    float* p = ga + int64_t(ci0 + k) * plane + n;
    *p = accumulate ? *p + acc[k] : acc[k];
  }
}

// gw[co][ci][t] = sum_o gy[co][o] * a[ci][stride*o + off_t]; gb[co] = sum_o gy[co][o]
// One workgroup per (co, ci): fixed-order strided partial sums, then the
// fixed block tree.
__global__ __launch_bounds__(BS) void conv_bwd_weight_kernel(int cin, int cout, int nz, int nx,
                                                             int ny, int oz, int ox, int oy,
                                                             int stride,
                                                             const float* __restrict__ a,
                                                             const float* __restrict__ gy,
                                                             float* __restrict__ gw,
                                                             float* __restrict__ gb) {
  __shared__ double red[28][BS / 64];
  const int co = blockIdx.x / cin, ci = blockIdx.x % cin;
  const int64_t nout = int64_t(oz) * ox * oy;
  const int64_t plane = int64_t(nz) * nx * ny;
  double acc[28];
#pragma unroll
  for (int t = 0; t < 28; ++t) acc[t] = 0.0;
  for (int64_t o = threadIdx.x; o < nout; o += BS) {
    const int iy = int(o % oy), ix = int((o / oy) % ox), iz = int(o / (int64_t(oy) * ox));
    const float g = gy[int64_t(co) * nout + o];
#pragma unroll
    for (int t = 0; t < 27; ++t) {
      const int Z = stride * iz + tap_dz(t), X = stride * ix + tap_dx(t), Y = stride * iy + tap_dy(t);
      if (Z < 0 || Z >= nz || X < 0 || X >= nx || Y < 0 || Y >= ny) continue;
      acc[t] += double(g) * double(a[int64_t(ci) * plane + (int64_t(Z) * nx + X) * ny + Y]);
    }
    acc[27] += double(g);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < 28; ++t) {
    const double s = wave_sum(acc[t]);
    if (lane == 0) red[t][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x < 28) {
    double s = 0.0;
    for (int k = 0; k < BS / 64; ++k) s += red[threadIdx.x][k];
    if (threadIdx.x < 27) gw[(int64_t(co) * cin + ci) * 27 + threadIdx.x] = float(s);
    else if (ci == 0) gb[co] = float(s);
  }
}

// g *= (y > 0)
__global__ void relu_mask_kernel(int64_t n, const float* __restrict__ y, float* __restrict__ g) {
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BS)
    if (!(y[i] > 0.f)) g[i] = 0.f;
}

__device__ inline float lin_w(int d) { return d == 0 ? 1.f : 0.5f; }

// fine[c][Z][X][Y] (+)= sum of the coarse neighbours with weights 1 or 1/2 per axis
__global__ void upsample_kernel(int c, int cz, int cx, int cy, const float* __restrict__ in,
                                float* __restrict__ out, int accumulate) {
  const int fz = 2 * cz - 1, fx = 2 * cx - 1, fy = 2 * cy - 1;
  const int64_t nf = int64_t(fz) * fx * fy, nc = int64_t(cz) * cx * cy;
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < nf * c; i += int64_t(gridDim.x) * BS) {
    const int ch = int(i / nf);
    const int64_t f = i % nf;
    const int Y = int(f % fy), X = int((f / fy) % fx), Z = int(f / (int64_t(fy) * fx));
    float s = 0.f;
    for (int az = Z >> 1; az <= (Z + 1) >> 1; ++az)
      for (int ax = X >> 1; ax <= (X + 1) >> 1; ++ax)
        for (int ay = Y >> 1; ay <= (Y + 1) >> 1; ++ay)
          s += lin_w(Z - 2 * az) * lin_w(X - 2 * ax) * lin_w(Y - 2 * ay) *
               in[int64_t(ch) * nc + (int64_t(az) * cx + ax) * cy + ay];
    out[i] = accumulate ? out[i] + s : s;
  }
}

// coarse[c][z][x][y] (+)= transpose of upsample applied to the fine gradient
__global__ void upsample_t_kernel(int c, int cz, int cx, int cy, const float* __restrict__ gf,
                                  float* __restrict__ gc, int accumulate) {
  const int fz = 2 * cz - 1, fx = 2 * cx - 1, fy = 2 * cy - 1;
  const int64_t nf = int64_t(fz) * fx * fy, nc = int64_t(cz) * cx * cy;
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < nc * c; i += int64_t(gridDim.x) * BS) {
    const int ch = int(i / nc);
    const int64_t k = i % nc;
    const int y = int(k % cy), x = int((k / cy) % cx), z = int(k / (int64_t(cy) * cx));
    float s = 0.f;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
          const int Z = 2 * z + dz, X = 2 * x + dx, Y = 2 * y + dy;
          if (Z < 0 || Z >= fz || X < 0 || X >= fx || Y < 0 || Y >= fy) continue;
          s += lin_w(dz) * lin_w(dx) * lin_w(dy) * gf[int64_t(ch) * nf + (int64_t(Z) * fx + X) * fy + Y];
        }
    gc[i] = accumulate ? gc[i] + s : s;
  }
}

// grad = 2 (p - t) / n; per-block partial sums of (p - t)^2
__global__ __launch_bounds__(BS) void mse_kernel(int64_t n, const float* __restrict__ p,
                                                 const float* __restrict__ t,
                                                 float* __restrict__ grad,
                                                 double* __restrict__ partials) {
  __shared__ double sh[BS / 64 + 2];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BS) {
    const float d = p[i] - t[i];
    grad[i] = 2.f * d / float(n);
    acc += double(d) * double(d);
  }
  double v[1] = {acc};
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

__global__ __launch_bounds__(BS) void sum_kernel(const double* __restrict__ partials, int nb,
                                                 double scale, double* __restrict__ out) {
  __shared__ double sh[BS / 64 + 2];
  double v[1] = {0.0};
  for (int i = threadIdx.x; i < nb; i += BS) v[0] += partials[i];
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) *out = v[0] * scale;
}

// Adam (Kingma & Ba 2015, Algorithm 1) on fp32 master weights; fp16 copy for the kernels
__global__ void adam_kernel(int64_t n, float* __restrict__ w, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v,
                            _Float16* __restrict__ w16, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BS) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (w16) w16[i] = _Float16(wi);
  }
}

// Network inputs on the level-1 node grid (fine node (2i, 2j, 2k)) of a 3D
// grid: [0] mean density of the fine elements at the node, [1..3] the load
// restricted by full weighting (the transpose of the trilinear
// prolongation) times fscale, [4] 1 where a DOF of the node is fixed, [5] 1.
__global__ void features_kernel(int nelx, int nely, int nelz, const double* __restrict__ xe,
                                const double* __restrict__ f, const uint8_t* __restrict__ mask,
                                double fscale, float* __restrict__ out) {
  const int nx = nelx + 1, ny = nely + 1, nz = nelz + 1;
  const int cx = nelx / 2 + 1, cy = nely / 2 + 1, cz = nelz / 2 + 1;
  const int64_t nc = int64_t(cx) * cy * cz;
  for (int64_t k = blockIdx.x * int64_t(BS) + threadIdx.x; k < nc; k += int64_t(gridDim.x) * BS) {
    const int y = int(k % cy), x = int((k / cy) % cx), z = int(k / (int64_t(cy) * cx));
    const int X = 2 * x, Y = 2 * y, Z = 2 * z;
    double xs = 0.0;
    int ne = 0;
    for (int ez = Z - 1; ez <= Z; ++ez)
      for (int ex = X - 1; ex <= X; ++ex)
        for (int ey = Y - 1; ey <= Y; ++ey) {
          if (ex < 0 || ex >= nelx || ey < 0 || ey >= nely || ez < 0 || ez >= nelz) continue;
          xs += xe[(int64_t(ez) * nelx + ex) * nely + ey];
          ++ne;
        }
    double fr[3] = {0.0, 0.0, 0.0};
    for (int dz = -1; dz <= 1; ++dz)
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
          const int FZ = Z + dz, FX = X + dx, FY = Y + dy;
          if (FZ < 0 || FZ >= nz || FX < 0 || FX >= nx || FY < 0 || FY >= ny) continue;
          const double wgt = (dz ? 0.5 : 1.0) * (dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0);
          const int64_t fn = (int64_t(FZ) * nx + FX) * ny + FY;
          for (int c = 0; c < 3; ++c) fr[c] += wgt * f[fn * 3 + c];
        }
    const int64_t fn0 = (int64_t(Z) * nx + X) * ny + Y;
    out[k] = float(ne ? xs / ne : 0.0);
    for (int c = 0; c < 3; ++c) out[(1 + c) * nc + k] = float(fr[c] * fscale);
    out[4 * nc + k] = mask[fn0] ? 1.f : 0.f;
    out[5 * nc + k] = 1.f;
  }
}

// out[c][k] = scale * u[fine node (2i, 2j, 2k)][c]  (injection)
__global__ void inject_kernel(int nelx, int nely, int nelz, const double* __restrict__ u,
                              double scale, float* __restrict__ out) {
  const int nx = nelx + 1, ny = nely + 1;
  const int cx = nelx / 2 + 1, cy = nely / 2 + 1, cz = nelz / 2 + 1;
  const int64_t nc = int64_t(cx) * cy * cz;
  for (int64_t k = blockIdx.x * int64_t(BS) + threadIdx.x; k < nc; k += int64_t(gridDim.x) * BS) {
    const int y = int(k % cy), x = int((k / cy) % cx), z = int(k / (int64_t(cy) * cx));
    const int64_t fn = (int64_t(2 * z) * nx + 2 * x) * ny + 2 * y;
    for (int c = 0; c < 3; ++c) out[c * nc + k] = float(u[fn * 3 + c] * scale);
  }
}

// x[fine DOF] = trilinear interpolation of the channel-major level-1 field y
// (zero on fixed DOFs)
__global__ void prolong3_kernel(int nelx, int nely, int nelz, const float* __restrict__ yc,
                                const uint8_t* __restrict__ mask, double* __restrict__ x) {
  const int nx = nelx + 1, ny = nely + 1, nz = nelz + 1;
  const int cx = nelx / 2 + 1, cy = nely / 2 + 1, cz = nelz / 2 + 1;
  const int64_t nc = int64_t(cx) * cy * cz, nf = int64_t(nx) * ny * nz;
  for (int64_t n = blockIdx.x * int64_t(BS) + threadIdx.x; n < nf; n += int64_t(gridDim.x) * BS) {
    const int Y = int(n % ny), X = int((n / ny) % nx), Z = int(n / (int64_t(ny) * nx));
    double s[3] = {0.0, 0.0, 0.0};
    for (int az = Z >> 1; az <= (Z + 1) >> 1; ++az)
      for (int ax = X >> 1; ax <= (X + 1) >> 1; ++ax)
        for (int ay = Y >> 1; ay <= (Y + 1) >> 1; ++ay) {
          const double wgt = double(lin_w(Z - 2 * az)) * lin_w(X - 2 * ax) * lin_w(Y - 2 * ay);
          const int64_t k = (int64_t(az) * cx + ax) * cy + ay;
          for (int c = 0; c < 3; ++c) s[c] += wgt * double(yc[c * nc + k]);
        }
    const uint8_t mk = mask[n];
    for (int c = 0; c < 3; ++c) x[n * 3 + c] = ((mk >> c) & 1) ? 0.0 : s[c];
  }
}

// partial sums of (a.b, c.d)
__global__ __launch_bounds__(BS) void dot2_kernel(int64_t n, const double* __restrict__ a,
                                                  const double* __restrict__ b,
                                                  const double* __restrict__ c,
                                                  const double* __restrict__ d,
                                                  double* __restrict__ partials) {
  __shared__ double sh[2 * (BS / 64) + 2];
  double v[2] = {0.0, 0.0};
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BS) {
    v[0] += a[i] * b[i];
    v[1] += c[i] * d[i];
  }
  block_sum<2>(v, sh);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = v[0];
    partials[1024 + blockIdx.x] = v[1];
  }
}

__global__ __launch_bounds__(BS) void scale_by_ratio_kernel(int64_t n, const double* __restrict__ partials,
                                                            int nb, double* __restrict__ x,
                                                            double* __restrict__ alpha) {
  // every block folds the partials in the same fixed order (identical result)
  __shared__ double sh[2 * (BS / 64) + 2];
  double v[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < nb; i += BS) {
    v[0] += partials[i];
    v[1] += partials[1024 + i];
  }
  block_sum<2>(v, sh);
  const double a = v[1] > 0.0 ? v[0] / v[1] : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) *alpha = a;
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BS)
    x[i] *= a;
}

unsigned grid1(int64_t n) { return unsigned(std::min<int64_t>((n + BS - 1) / BS, 4096)); }

int out_dim(int n, int stride) { return stride == 1 ? n : (n - 1) / 2 + 1; }

}  // namespace
}  // namespace tl

using namespace tl;

extern "C" {

int tl_nn_conv3d(int32_t cin, int32_t cout, int32_t nz, int32_t nx, int32_t ny, int32_t stride,
                 int32_t relu, const float* in, const uint16_t* w16, const float* bias, float* out,
                 void* stream) {
  TL_ARG(in && w16 && bias && out, "null argument");
  TL_ARG(cin >= 1 && cin <= 64 && cout >= 1 && nz >= 1 && nx >= 1 && ny >= 1,
         "bad shape (1 <= cin <= 64)");
  TL_ARG(stride == 1 || (stride == 2 && nz % 2 == 1 && nx % 2 == 1 && ny % 2 == 1),
         "stride must be 1, or 2 on odd node counts");
  const int oz = out_dim(nz, stride), ox = out_dim(nx, stride), oy = out_dim(ny, stride);
  const int64_t nout = int64_t(oz) * ox * oy;
  const dim3 grid{unsigned((nout + BS - 1) / BS), unsigned((cout + CO - 1) / CO), 1u};
  hipLaunchKernelGGL(conv_fwd_kernel, grid, dim3(BS), size_t(CO) * cin * 27 * sizeof(float),
                     static_cast<hipStream_t>(stream), cin, cout, nz, nx, ny, oz, ox, oy, stride,
                     relu, in, reinterpret_cast<const _Float16*>(w16), bias, out);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

int tl_nn_conv3d_bwd_data(int32_t cin, int32_t cout, int32_t nz, int32_t nx, int32_t ny,
                          int32_t stride, const float* gout, const uint16_t* w16, float* gin,
                          int32_t accumulate, void* stream) {
  TL_ARG(gout && w16 && gin, "null argument");
  TL_ARG(cin >= 1 && cout >= 1 && cout <= 64 && nz >= 1 && nx >= 1 && ny >= 1,
         "bad shape (1 <= cout <= 64)");
  TL_ARG(stride == 1 || (stride == 2 && nz % 2 == 1 && nx % 2 == 1 && ny % 2 == 1),
         "stride must be 1, or 2 on odd node counts");
  const int oz = out_dim(nz, stride), ox = out_dim(nx, stride), oy = out_dim(ny, stride);
  const int64_t nin = int64_t(nz) * nx * ny;
  const dim3 grid{unsigned((nin + BS - 1) / BS), unsigned((cin + CI - 1) / CI), 1u};
  hipLaunchKernelGGL(conv_bwd_data_kernel, grid, dim3(BS), size_t(cout) * CI * 27 * sizeof(float),
                     static_cast<hipStream_t>(stream), cin, cout, nz, nx, ny, oz, ox, oy, stride,
                     accumulate, gout, reinterpret_cast<const _Float16*>(w16), gin);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

int tl_nn_conv3d_bwd_weight(int32_t cin, int32_t cout, int32_t nz, int32_t nx, int32_t ny,
                            int32_t stride, const float* in, const float* gout, float* gw,
                            float* gb, void* stream) {
  TL_ARG(in && gout && gw && gb, "null argument");
  TL_ARG(cin >= 1 && cout >= 1 && nz >= 1 && nx >= 1 && ny >= 1, "bad shape");
  TL_ARG(stride == 1 || (stride == 2 && nz % 2 == 1 && nx % 2 == 1 && ny % 2 == 1),
         "stride must be 1, or 2 on odd node counts");
  const int oz = out_dim(nz, stride), ox = out_dim(nx, stride), oy = out_dim(ny, stride);
  hipLaunchKernelGGL(conv_bwd_weight_kernel, dim3(unsigned(cout * cin)), dim3(BS), 0,
                     static_cast<hipStream_t>(stream), cin, cout, nz, nx, ny, oz, ox, oy, stride,
                     in, gout, gw, gb);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

int tl_nn_relu_mask(int64_t n, const float* y, float* g, void* stream) {
  TL_ARG(y && g && n >= 0, "bad argument");
  if (n == 0) return TL_OK;
  hipLaunchKernelGGL(relu_mask_kernel, dim3(grid1(n)), dim3(BS), 0,
                     static_cast<hipStream_t>(stream), n, y, g);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

int tl_nn_upsample(int32_t c, int32_t cz, int32_t cx, int32_t cy, const float* coarse,
                   float* fine, int32_t accumulate, void* stream) {
  TL_ARG(coarse && fine && c >= 1 && cz >= 2 && cx >= 2 && cy >= 2, "bad argument");
  const int64_t n = int64_t(c) * (2 * cz - 1) * (2 * cx - 1) * (2 * cy - 1);
  hipLaunchKernelGGL(upsample_kernel, dim3(grid1(n)), dim3(BS), 0,
                     static_cast<hipStream_t>(stream), c, cz, cx, cy, coarse, fine, accumulate);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

int tl_nn_upsample_t(int32_t c, int32_t cz, int32_t cx, int32_t cy, const float* gfine,
                     float* gcoarse, int32_t accumulate, void* stream) {
  TL_ARG(gfine && gcoarse && c >= 1 && cz >= 2 && cx >= 2 && cy >= 2, "bad argument");
  const int64_t n = int64_t(c) * cz * cx * cy;
  hipLaunchKernelGGL(upsample_t_kernel, dim3(grid1(n)), dim3(BS), 0,
                     static_cast<hipStream_t>(stream), c, cz, cx, cy, gfine, gcoarse, accumulate);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

int tl_nn_mse(int64_t n, const float* pred, const float* target, float* grad, double* partials,
              double* loss, void* stream) {
  TL_ARG(pred && target && grad && partials && loss && n > 0, "bad argument");
  const unsigned nb = std::min<unsigned>(grid1(n), 1024u);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(mse_kernel, dim3(nb), dim3(BS), 0, s, n, pred, target, grad, partials);
  TL_CHECK_LAUNCH();
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(BS), 0, s, partials, int(nb), 1.0 / double(n), loss);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

int tl_nn_adam(int64_t n, float* w, const float* g, float* m, float* v, uint16_t* w16,
               int32_t step, float lr, void* stream) {
  TL_ARG(w && g && m && v && n > 0 && step >= 1, "bad argument");
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = float(1.0 - std::pow(0.9, step)), c2 = float(1.0 - std::pow(0.999, step));
  hipLaunchKernelGGL(adam_kernel, dim3(grid1(n)), dim3(BS), 0, static_cast<hipStream_t>(stream),
                     n, w, g, m, v, reinterpret_cast<_Float16*>(w16), lr, b1, b2, 1e-8f, c1, c2);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

int tl_nn_features(tl_problem* p, const double* xPhys, const double* f, double fscale,
                   float* out) {
  TL_ARG(p && xPhys && f && out, "null argument");
  const Grid& g = p->g;
  TL_ARG(g.dim == 3 && p->ndpn == 3 && g.nelx % 2 == 0 && g.nely % 2 == 0 && g.nelz % 2 == 0,
         "network features: 3D H8 elasticity with even element counts");
  const int64_t nc = int64_t(g.nelx / 2 + 1) * (g.nely / 2 + 1) * (g.nelz / 2 + 1);
  hipLaunchKernelGGL(features_kernel, dim3(grid1(nc)), dim3(BS), 0, p->stream, g.nelx, g.nely,
                     g.nelz, xPhys, f, p->op.mask.as<uint8_t>(), fscale, out);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

int tl_nn_inject(tl_problem* p, const double* u, double scale, float* out) {
  TL_ARG(p && u && out, "null argument");
  const Grid& g = p->g;
  TL_ARG(g.dim == 3 && p->ndpn == 3 && g.nelx % 2 == 0 && g.nely % 2 == 0 && g.nelz % 2 == 0,
         "network targets: 3D H8 elasticity with even element counts");
  const int64_t nc = int64_t(g.nelx / 2 + 1) * (g.nely / 2 + 1) * (g.nelz / 2 + 1);
  hipLaunchKernelGGL(inject_kernel, dim3(grid1(nc)), dim3(BS), 0, p->stream, g.nelx, g.nely,
                     g.nelz, u, scale, out);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

int tl_nn_prolong(tl_problem* p, const float* y, double* x) {
  TL_ARG(p && y && x, "null argument");
  const Grid& g = p->g;
  TL_ARG(g.dim == 3 && p->ndpn == 3 && g.nelx % 2 == 0 && g.nely % 2 == 0 && g.nelz % 2 == 0,
         "network output: 3D H8 elasticity with even element counts");
  hipLaunchKernelGGL(prolong3_kernel, dim3(grid1(g.nnodes)), dim3(BS), 0, p->stream, g.nelx,
                     g.nely, g.nelz, y, p->op.mask.as<uint8_t>(), x);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

int tl_energy_scale(tl_problem* p, const double* f, double* x, double* work, double* alpha) {
  TL_ARG(p && f && x && work && alpha, "null argument");
  TL_TRY(tl_apply_K(p, x, work, TL_PREC_FP64));
  TL_TRY(p->partials.alloc(sizeof(double) * 2048));
  const unsigned nb = std::min<unsigned>(grid1(p->ndof), 1024u);
  double* part = p->partials.as<double>();
  hipLaunchKernelGGL(dot2_kernel, dim3(nb), dim3(BS), 0, p->stream, p->ndof, f, x, x, work, part);
  TL_CHECK_LAUNCH();
  TL_TRY(p->t.alloc(sizeof(double) * 2));
  hipLaunchKernelGGL(scale_by_ratio_kernel, dim3(grid1(p->ndof)), dim3(BS), 0, p->stream,
                     p->ndof, part, int(nb), x, p->t.as<double>());
  TL_CHECK_LAUNCH();
  TL_HIP(hipMemcpyAsync(alpha, p->t.ptr, sizeof(double), hipMemcpyDeviceToHost, p->stream));
  TL_HIP(hipStreamSynchronize(p->stream));
  return TL_OK;
}

}  // extern "C"
