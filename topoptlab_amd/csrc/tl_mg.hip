// Geometric multigrid on structured grids: Galerkin coarse operators stored
// as nodal stencils, stride-2 (bi/tri)linear transfers, damped-Jacobi V-cycle
// and a dense inverse on the coarsest grid.
//
// Reference: topoptlab/multigrid.py:89-223 (V-cycle with Galerkin P^T A P and
// a direct coarsest solve), topoptlab/gmg.py:34-224 (shape-function
// prolongation), linear_solvers.py:306-372 (damped Jacobi).  Deviations kept
// on purpose: zero initial guess on coarse levels (symmetric V-cycle), coarse
// operators built once per density instead of once per cycle, no absolute
// early exit inside the smoother.
//
// Coarse operators: with P_l the nodal interpolation (x) I_ndpn and Dirichlet
// sets that are grid aligned (every fixed DOF interpolates only from fixed
// coarse DOFs -- checked on the host), the Galerkin operator of K_ff equals
// M_{l} P^T A_{l-1} P M_{l} with the fixed rows/cols of level l removed.  Level 1
// is assembled straight from the fine densities: coarse element matrix =
// sum_children s_child * G_c with G_c = Q_c^T KE Q_c (8 constant matrices).
// Deeper levels are Galerkin products of the previous level's stencil.
#include "tl_internal.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace tl {

namespace {

constexpr int BS = 256;
constexpr int GJB = 32;  // Gauss-Jordan block size

__host__ __device__ inline int nstencil(int dim) { return dim == 3 ? 27 : 9; }
__host__ __device__ inline int center_k(int dim) { return dim == 3 ? 13 : 4; }

struct GridD {  // device-side copy of a level grid
  int dim, nelx, nely, nelz, nx, ny, nz;
  int64_t nnodes;
};

GridD to_d(const Grid& g) {
  return GridD{g.dim, g.nelx, g.nely, g.nelz, g.nx, g.ny, g.nz, g.nnodes};
}

// 32-bit divisions (tl_create keeps nnodes * ndpn below 2^31): a 64-bit
// division by a runtime divisor costs about a hundred instructions.
__device__ inline void node_xyz(const GridD& g, int64_t n, int& ix, int& iy, int& iz) {
  const uint32_t u = uint32_t(n), ny = uint32_t(g.ny), nx = uint32_t(g.nx);
  const uint32_t t = u / ny;
  iy = int(u - t * ny);
  ix = int(t % nx);
  iz = int(t / nx);
}

__device__ inline int64_t node_id(const GridD& g, int ix, int iy, int iz) {
  return (int64_t(iz) * g.nx + ix) * g.ny + iy;
}

__device__ inline bool node_in(const GridD& g, int ix, int iy, int iz) {
  return ix >= 0 && ix < g.nx && iy >= 0 && iy < g.ny && iz >= 0 && iz < g.nz;
}

// stencil index of a relative offset (rx, ry, rz)
__device__ inline int kidx(int dim, int rx, int ry, int rz) {
  return dim == 3 ? ((rz + 1) * 3 + (rx + 1)) * 3 + (ry + 1) : (rx + 1) * 3 + (ry + 1);
}

// ---------------------------------------------------------------------------
// level-1 stencil straight from the fine densities
// ---------------------------------------------------------------------------
// One thread per (node, d, c), (d, c) = blockIdx.y: the node's (up to) 8 level-1 elements and
// their 8 children's scales are loaded once and feed all 27 offsets; each
// entry sums elements in (oz, ox, oy) order and children in order.
template <bool D3, int NDPN>
__global__ __launch_bounds__(BS) void st_from_fine_kernel(GridD g1, const double* __restrict__ spf,
                                                          int fnelx, int fnely,
                                                          const double* __restrict__ G1,
                                                          const uint8_t* __restrict__ mask1,
                                                          double* __restrict__ st) {
  constexpr int NEL = D3 ? 8 : 4, NCH = D3 ? 8 : 4, ND = NEL * NDPN, NK = D3 ? 27 : 9;
  // (d, c) per grid row: uniform over the workgroup, so the element-matrix
  // entries are read once per wave (scalar loads), not once per lane
  const int dc = int(blockIdx.y);
  const int d = dc / NDPN, c = dc % NDPN;
  for (int64_t n = blockIdx.x * int64_t(BS) + threadIdx.x; n < g1.nnodes;
       n += int64_t(gridDim.x) * BS) {
    int ix, iy, iz;
    node_xyz(g1, n, ix, iy, iz);
    // child scales of the elements around the node, e = ((oz+1)*2 + (ox+1))*2 + (oy+1)
    double sc[NEL][NCH];
    bool eok[NEL];
#pragma unroll
    for (int e = 0; e < NEL; ++e) {
      const int oz = D3 ? (e >> 2) - 1 : 0, ox = ((e >> 1) & 1) - 1, oy = (e & 1) - 1;
      const int ex = ix + ox, ey = iy + oy, ez = iz + oz;
      eok[e] = ex >= 0 && ex < g1.nelx && ey >= 0 && ey < g1.nely &&
               (!D3 || (ez >= 0 && ez < g1.nelz));
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        const int cy = ch & 1, cx = (ch >> 1) & 1, cz = ch >> 2;
        const int fx = 2 * ex + cx, fy = 2 * ey + cy, fz = 2 * ez + cz;
        const int64_t pidx =
            (D3 ? (int64_t(fz + 1) * (fnelx + 2) + fx + 1) * (fnely + 2)
                : int64_t(fx + 1) * (fnely + 2)) + fy + 1;
        sc[e][ch] = eok[e] ? spf[pidx] : 0.0;
      }
    }
    const bool fn = (mask1[n] >> d) & 1;
    const double* Gdc = G1 + d * ND + c;  // entry (la*NDPN + d)*ND + lb*NDPN + c
#pragma unroll
    for (int k = 0; k < NK; ++k) {
      const int rz = D3 ? k / 9 - 1 : 0, rx = (k / 3) % 3 - 1, ry = k % 3 - 1;
      double acc = 0.0;
      const bool mexists = node_in(g1, ix + rx, iy + ry, iz + rz);
#pragma unroll
      for (int e = 0; e < NEL; ++e) {
        const int oz = D3 ? (e >> 2) - 1 : 0, ox = ((e >> 1) & 1) - 1, oy = (e & 1) - 1;
        const int bx = rx - ox, by = ry - oy, bz = rz - oz;
        if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
        if (!(mexists && eok[e])) continue;
        const int la = loc_of(-ox, -oy, -oz), lb = loc_of(bx, by, bz);
        const int off = la * NDPN * ND + lb * NDPN;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) acc += sc[e][ch] * Gdc[ch * ND * ND + off];
      }
      const bool fm = mexists ? ((mask1[node_id(g1, ix + rx, iy + ry, iz + rz)] >> c) & 1) : true;
      if (fn || fm) acc = (fn && rx == 0 && ry == 0 && rz == 0 && c == d) ? 1.0 : 0.0;
      st[int64_t((k * NDPN + d) * NDPN + c) * g1.nnodes + n] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// stencil Galerkin: level l (fine) -> level l+1 (coarse)
// ---------------------------------------------------------------------------
// One thread per (coarse node, d, c).  A_c(n, n+r) = sum over the fine nodes
// i = 2n + dv around n and their stencil offsets sv of w(dv) A_f(i, i+sv)
// w(dv + sv - 2r), with the 1D transfer weight w(0) = 1, w(+-1) = 1/2 (else 0)
// per axis.  The weights factor per axis, so each fine node's block of 27 (2D:
// 9) values is contracted along z, y and x with 3x3 weight tables (243 instead
// of 729 multiply-adds, no branches on the offsets).
__device__ __forceinline__ double tw1(int e) { return e == 0 ? 1.0 : ((e == 1 || e == -1) ? 0.5 : 0.0); }

template <bool D3>
__global__ __launch_bounds__(BS) void st_coarsen_kernel(GridD gf, GridD gc, int ndpn,
                                                        const double* __restrict__ stf,
                                                        const uint8_t* __restrict__ maskc,
                                                        double* __restrict__ stc) {
  constexpr int NK = D3 ? 27 : 9;
  constexpr int NZ = D3 ? 3 : 1;
  const int64_t tot = gc.nnodes * ndpn * ndpn;
  for (int64_t t = blockIdx.x * int64_t(BS) + threadIdx.x; t < tot;
       t += int64_t(gridDim.x) * BS) {
    const int64_t n = t % gc.nnodes;
    const int dc = int(t / gc.nnodes);
    const int d = dc / ndpn, c = dc % ndpn;
    int ix, iy, iz;
    node_xyz(gc, n, ix, iy, iz);
    double acc[NK];
#pragma unroll
    for (int k = 0; k < NK; ++k) acc[k] = 0.0;
#pragma unroll 1
    for (int dp = 0; dp < NK; ++dp) {
      const int dz = D3 ? dp / 9 - 1 : 0, dx = (dp / 3) % 3 - 1, dy = dp % 3 - 1;
      const int fx = 2 * ix + dx, fy = 2 * iy + dy, fz = D3 ? 2 * iz + dz : 0;
      if (!node_in(gf, fx, fy, fz)) continue;
      const double wi = (dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0) * (dz ? 0.5 : 1.0);
      const int64_t fi = node_id(gf, fx, fy, fz);
      // v[sz][sx][sy]: the fine node's block (d, c) for offset (sx, sy, sz)
      double v[NZ][3][3];
#pragma unroll
      for (int ks = 0; ks < NK; ++ks)
        v[D3 ? ks / 9 : 0][(ks / 3) % 3][ks % 3] =
            stf[int64_t((ks * ndpn + d) * ndpn + c) * gf.nnodes + fi];
      // contract z: zc[rz][sx][sy] = sum_sz w(dz + sz - 2 rz) v[sz][sx][sy]
      double zc[NZ][3][3];
#pragma unroll
      for (int rz = 0; rz < NZ; ++rz)
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int b = 0; b < 3; ++b) {
            if constexpr (D3) {
              double sum = 0.0;
#pragma unroll
              for (int sz = 0; sz < 3; ++sz) sum += tw1(dz + (sz - 1) - 2 * (rz - 1)) * v[sz][a][b];
              zc[rz][a][b] = sum;
            } else {
              zc[rz][a][b] = v[0][a][b];
            }
          }
      // contract y: yc[rz][sx][ry]
      double yc[NZ][3][3];
#pragma unroll
      for (int rz = 0; rz < NZ; ++rz)
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int ry = 0; ry < 3; ++ry) {
            double sum = 0.0;
#pragma unroll
            for (int sy = 0; sy < 3; ++sy) sum += tw1(dy + (sy - 1) - 2 * (ry - 1)) * zc[rz][a][sy];
            yc[rz][a][ry] = sum;
          }
      // contract x and accumulate: acc[rz][rx][ry] += wi * sum_sx w(dx + sx - 2 rx) yc
#pragma unroll
      for (int rz = 0; rz < NZ; ++rz)
#pragma unroll
        for (int rx = 0; rx < 3; ++rx)
#pragma unroll
          for (int ry = 0; ry < 3; ++ry) {
            double sum = 0.0;
#pragma unroll
            for (int sx = 0; sx < 3; ++sx) sum += tw1(dx + (sx - 1) - 2 * (rx - 1)) * yc[rz][sx][ry];
            acc[(rz * 3 + rx) * 3 + ry] += wi * sum;
          }
    }
    const bool fn = (maskc[n] >> d) & 1;
#pragma unroll
    for (int k = 0; k < NK; ++k) {
      const int rz = D3 ? k / 9 - 1 : 0, rx = (k / 3) % 3 - 1, ry = k % 3 - 1;
      const bool mexists = node_in(gc, ix + rx, iy + ry, iz + rz);
      double v = mexists ? acc[k] : 0.0;
      const bool fm = mexists ? ((maskc[node_id(gc, ix + rx, iy + ry, iz + rz)] >> c) & 1) : true;
      if (fn || fm) v = (fn && rx == 0 && ry == 0 && rz == 0 && c == d) ? 1.0 : 0.0;
      stc[int64_t((k * ndpn + d) * ndpn + c) * gc.nnodes + n] = v;
    }
  }
}

// Forward half of the stencil (centre block and the blocks of neighbours
// with a positive offset in (z, x, y) order), fp64 and fp32, and the inverse
// diagonal.  The operator is symmetric, A(i, i-r) = A(i-r, i)^T, so the V-cycle
// reads a backward block from the neighbour's forward half, transposed: half
// the stencil bytes, and the applied operator is exactly symmetric.
__global__ void st_finish_kernel(GridD g, int ndpn, const double* __restrict__ st64,
                                 double* __restrict__ sth64, float* __restrict__ st32,
                                 const uint8_t* __restrict__ mask, double* __restrict__ dinv64,
                                 float* __restrict__ dinv32) {
  const int kc = center_k(g.dim);
  const int nqh = (kc + 1) * ndpn * ndpn;
  const int64_t tot = g.nnodes * nqh;
  for (int64_t t = blockIdx.x * int64_t(BS) + threadIdx.x; t < tot;
       t += int64_t(gridDim.x) * BS) {
    const double v = st64[t + int64_t(kc) * ndpn * ndpn * g.nnodes];
    sth64[t] = v;
    st32[t] = float(v);
  }
  for (int64_t t = blockIdx.x * int64_t(BS) + threadIdx.x; t < g.nnodes * ndpn;
       t += int64_t(gridDim.x) * BS) {
    const int64_t n = t / ndpn;
    const int d = int(t % ndpn);
    const double dg = st64[int64_t((kc * ndpn + d) * ndpn + d) * g.nnodes + n];
    const bool fixed = (mask[n] >> d) & 1;
    const double inv = (fixed || dg == 0.0) ? 0.0 : 1.0 / dg;
    dinv64[t] = inv;
    dinv32[t] = float(inv);
  }
}

// ---------------------------------------------------------------------------
// V-cycle kernels (T = float or double)
// ---------------------------------------------------------------------------
// y = b - A x  (MODE 0)   or   y = x + omega*dinv*(b - A x)  (MODE 1), with A
// from its forward half sth ([kk][d][c][node], kk = k - centre): forward
// blocks of node i, backward blocks as the transposed forward blocks of the
// neighbours.  Workgroups take contiguous node ranges per XCD so that the
// neighbours' blocks (up to one node plane back) are re-read from the XCD's
// own L2.
// (P xc) at fine node (ix, iy, iz): the up to 8 (2D: 4) coarse nodes are read
// from clamped positions, all loads before the first sum, the unused ones with
// weight 0.  Shared by the prolongation and the fused post-smoothing sweep so
// that both add the same value.
template <int NDPN, typename T, bool D3>
__device__ __forceinline__ void prolong_at(const GridD& gc, const T* __restrict__ xc, int ix,
                                           int iy, int iz, T (&acc)[NDPN]) {
#pragma unroll
  for (int d = 0; d < NDPN; ++d) acc[d] = T(0);
  const int x0 = ix >> 1, x1 = (ix + 1) >> 1, y0 = iy >> 1, y1 = (iy + 1) >> 1;
  const int z0 = iz >> 1, z1 = (iz + 1) >> 1;
  const T wx = (ix & 1) ? T(0.5) : T(1), wy = (iy & 1) ? T(0.5) : T(1), wz = (iz & 1) ? T(0.5) : T(1);
#pragma unroll
  for (int a = 0; a < (D3 ? 2 : 1); ++a)
#pragma unroll
    for (int bx = 0; bx < 2; ++bx)
#pragma unroll
      for (int by = 0; by < 2; ++by) {
        const int cz = z0 + a, cx = x0 + bx, cy = y0 + by;
        const bool ok = cz <= z1 && cx <= x1 && cy <= y1;
        const int64_t ci = node_id(gc, min(cx, gc.nx - 1), min(cy, gc.ny - 1), min(cz, gc.nz - 1));
        const T w = ok ? wx * wy * wz : T(0);
#pragma unroll
        for (int d = 0; d < NDPN; ++d) acc[d] += w * xc[ci * NDPN + d];
      }
}

// Coarse correction applied on the fly by the post-smoothing sweep that
// follows a prolongation: x'(m) = x(m) + (P xc)(m) on the free DOFs of m.
template <typename T>
struct CoarseCorr {
  GridD gc;
  const T* xc = nullptr;
  const uint8_t* mask = nullptr;  // the fine level's
};

// One 27-point (2D: 9-point) term A(n, m) x_m, m = n + (rx, ry, rz).  DIM is
// the grid dimension when known at compile time (0: read it from g).  PRO: x_m
// is x'(m) of the coarse correction cc (DIM 2 or 3 only).
template <int NDPN, typename T, int DIM, bool PRO = false>
__device__ __forceinline__ void st_term(const GridD& g, const T* __restrict__ sth,
                                        const T* __restrict__ x, int64_t n, int ix, int iy,
                                        int iz, int rx, int ry, int rz, T (&acc)[NDPN],
                                        const CoarseCorr<T>& cc = CoarseCorr<T>{}) {
  const int dim = DIM ? DIM : g.dim;
  const int kc = center_k(dim);
  const int k = kidx(dim, rx, ry, rz);
  // neighbours outside the grid carry zero weight: clamp the reads
  const int mx = min(max(ix + rx, 0), g.nx - 1), my = min(max(iy + ry, 0), g.ny - 1),
            mz = min(max(iz + rz, 0), g.nz - 1);
  const int64_t m = node_id(g, mx, my, mz);
  T xm[NDPN];
#pragma unroll
  for (int c = 0; c < NDPN; ++c) xm[c] = x[m * NDPN + c];
  if constexpr (PRO) {
    static_assert(DIM == 2 || DIM == 3, "the fused correction needs a compile-time dimension");
    T pc[NDPN];
    prolong_at<NDPN, T, DIM == 3>(cc.gc, cc.xc, mx, my, mz, pc);
    const uint8_t mkm = cc.mask[m];
#pragma unroll
    for (int c = 0; c < NDPN; ++c)
      if (!((mkm >> c) & 1)) xm[c] += pc[c];
  }
  if (k >= kc) {  // forward block of node n (zero towards missing neighbours)
    const T* h = sth + int64_t(k - kc) * NDPN * NDPN * g.nnodes + n;
#pragma unroll
    for (int d = 0; d < NDPN; ++d)
#pragma unroll
      for (int c = 0; c < NDPN; ++c) acc[d] += h[int64_t(d * NDPN + c) * g.nnodes] * xm[c];
  } else {  // A(n, m) = A(m, n)^T: forward block (offset -r) of neighbour m
    const bool ok = node_in(g, ix + rx, iy + ry, iz + rz);
    const T* h = sth + int64_t(2 * kc - k - kc) * NDPN * NDPN * g.nnodes + m;
#pragma unroll
    for (int d = 0; d < NDPN; ++d)
#pragma unroll
      for (int c = 0; c < NDPN; ++c) {
        const T w = h[int64_t(c * NDPN + d) * g.nnodes];
        acc[d] += (ok ? w : T(0)) * xm[c];
      }
  }
}

// Raw buffer loads (gfx9 descriptor, no stride, range = nbytes): one uniform
// base per array and a 32-bit per-lane byte offset, so a row's dozens of
// loads in flight cost one address register per neighbour.
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t brsrc(const T* base, uint32_t nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(base), short(0), int(nbytes), 0x00020000);
}
template <typename T>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  static_assert(sizeof(T) == 8 || sizeof(T) == 4, "8- or 4-byte values");
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// Loads of 3D stencil term K (offset (K/9 - 1, K/3 % 3 - 1, K % 3 - 1) in
// (z, x, y)) at node n: x of the neighbour and the block A(n, m), the forward
// block of n for K >= 13, else the transposed forward block 26 - K of m.
// Each forward-half slab kk (NDPN^2 arrays of nnodes values) is one buffer,
// which the host keeps below 4 GiB (st_dimu).
template <int NDPN, typename T, int K>
__device__ __forceinline__ void st_load_term(const GridD& g, const T* __restrict__ sth,
                                             __amdgpu_buffer_rsrc_t rxb, uint32_t n, int ix,
                                             int iy, int iz, T (&xm)[NDPN],
                                             T (&w)[NDPN * NDPN]) {
  constexpr int NN = NDPN * NDPN;
  constexpr int rz = K / 9 - 1, rx = K / 3 % 3 - 1, ry = K % 3 - 1;
  const uint32_t arr = uint32_t(g.nnodes) * uint32_t(sizeof(T));
  const uint32_t m = (uint32_t(min(max(iz + rz, 0), g.nz - 1)) * uint32_t(g.nx) +
                      uint32_t(min(max(ix + rx, 0), g.nx - 1))) * uint32_t(g.ny) +
                     uint32_t(min(max(iy + ry, 0), g.ny - 1));
#pragma unroll
  for (int c = 0; c < NDPN; ++c)
    xm[c] = bload<T>(rxb, m * uint32_t(NDPN * sizeof(T)), uint32_t(c * sizeof(T)));
  constexpr int kk = K >= 13 ? K - 13 : 13 - K;
  const __amdgpu_buffer_rsrc_t r = brsrc(sth + int64_t(kk) * NN * g.nnodes, uint32_t(NN) * arr);
  if constexpr (K >= 13) {
#pragma unroll
    for (int e = 0; e < NN; ++e) w[e] = bload<T>(r, n * uint32_t(sizeof(T)), uint32_t(e) * arr);
  } else {
    const bool ok = node_in(g, ix + rx, iy + ry, iz + rz);
#pragma unroll
    for (int d = 0; d < NDPN; ++d)
#pragma unroll
      for (int c = 0; c < NDPN; ++c) {
        const T v = bload<T>(r, m * uint32_t(sizeof(T)), uint32_t(c * NDPN + d) * arr);
        w[d * NDPN + c] = ok ? v : T(0);
      }
  }
}

// A group of 3D stencil terms K...: the registers of their loads and their
// sum into acc in the order given.
template <int NDPN, typename T, int... K>
struct StGroup {
  static constexpr int NT = sizeof...(K), NN = NDPN * NDPN;
  T xm[NT][NDPN], w[NT][NN];
  __device__ __forceinline__ void load(const GridD& g, const T* __restrict__ sth,
                                       __amdgpu_buffer_rsrc_t rxb, uint32_t n, int ix, int iy,
                                       int iz) {
    int j = 0;
    ((st_load_term<NDPN, T, K>(g, sth, rxb, n, ix, iy, iz, xm[j], w[j]), ++j), ...);
  }
  __device__ __forceinline__ void sum(T (&acc)[NDPN]) const {
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int d = 0; d < NDPN; ++d)
#pragma unroll
        for (int c = 0; c < NDPN; ++c) acc[d] += w[j][d * NDPN + c] * xm[j][c];
  }
};

// Loads of a group, then its sum.  The scheduling barrier keeps the FMAs
// behind the group's loads (the scheduler would otherwise interleave them one
// term at a time, leaving few loads in flight); the empty asm takes the sums
// (the FMAs stay before it) and may read and write through the kernel's
// pointers (the next group's loads stay after it), so that one group of
// registers is live at a time.
template <int NDPN, typename T, int... K>
__device__ __forceinline__ void st_group(const GridD& g, const T* __restrict__ sth,
                                         const T* __restrict__ x, __amdgpu_buffer_rsrc_t rxb,
                                         uint32_t n, int ix, int iy, int iz, T (&acc)[NDPN]) {
  StGroup<NDPN, T, K...> grp;
  grp.load(g, sth, rxb, n, ix, iy, iz);
  __builtin_amdgcn_sched_barrier(0);
  grp.sum(acc);
#pragma unroll
  for (int d = 0; d < NDPN; ++d) asm volatile("" : "+v"(acc[d]));
  asm volatile("" : : "s"(sth), "s"(x) : "memory");
}

// The 27 terms of node n's row of the 3D stencil, in pairs (K, 26 - K) and
// then the centre: the backward block of a pair is the forward block that
// its owner (a node of the plane below, or of the line or node before) reads
// in the same group, which keeps the two reads close in time and in the L2
// (1.37x the algorithmic bytes fetched instead of 1.62x in plane order).
// GP pairs per group: 1 on large levels (few registers: every workgroup
// resident at once), 3 on small ones (fewer dependent round trips per
// thread: 5 groups instead of 14, for levels whose few waves leave the
// memory latency exposed).
template <int NDPN, typename T, int GP>
__device__ __forceinline__ void st_row_3d(const GridD& g, const T* __restrict__ sth,
                                          const T* __restrict__ x, uint32_t n, int ix, int iy,
                                          int iz, T (&acc)[NDPN]) {
  static_assert(GP == 1 || GP == 3, "1 or 3 pairs per group");
  const __amdgpu_buffer_rsrc_t rxb = brsrc(x, uint32_t(g.nnodes) * NDPN * uint32_t(sizeof(T)));
  if constexpr (GP == 1) {
    st_group<NDPN, T, 0, 26>(g, sth, x, rxb, n, ix, iy, iz, acc);
    st_group<NDPN, T, 1, 25>(g, sth, x, rxb, n, ix, iy, iz, acc);
    st_group<NDPN, T, 2, 24>(g, sth, x, rxb, n, ix, iy, iz, acc);
    st_group<NDPN, T, 3, 23>(g, sth, x, rxb, n, ix, iy, iz, acc);
    st_group<NDPN, T, 4, 22>(g, sth, x, rxb, n, ix, iy, iz, acc);
    st_group<NDPN, T, 5, 21>(g, sth, x, rxb, n, ix, iy, iz, acc);
    st_group<NDPN, T, 6, 20>(g, sth, x, rxb, n, ix, iy, iz, acc);
    st_group<NDPN, T, 7, 19>(g, sth, x, rxb, n, ix, iy, iz, acc);
    st_group<NDPN, T, 8, 18>(g, sth, x, rxb, n, ix, iy, iz, acc);
    st_group<NDPN, T, 9, 17>(g, sth, x, rxb, n, ix, iy, iz, acc);
    st_group<NDPN, T, 10, 16>(g, sth, x, rxb, n, ix, iy, iz, acc);
    st_group<NDPN, T, 11, 15>(g, sth, x, rxb, n, ix, iy, iz, acc);
    st_group<NDPN, T, 12, 14>(g, sth, x, rxb, n, ix, iy, iz, acc);
    st_group<NDPN, T, 13>(g, sth, x, rxb, n, ix, iy, iz, acc);
  } else {
    st_group<NDPN, T, 0, 26, 1, 25, 2, 24>(g, sth, x, rxb, n, ix, iy, iz, acc);
    st_group<NDPN, T, 3, 23, 4, 22, 5, 21>(g, sth, x, rxb, n, ix, iy, iz, acc);
    st_group<NDPN, T, 6, 20, 7, 19, 8, 18>(g, sth, x, rxb, n, ix, iy, iz, acc);
    st_group<NDPN, T, 9, 17, 10, 16, 11, 15>(g, sth, x, rxb, n, ix, iy, iz, acc);
    st_group<NDPN, T, 12, 14, 13>(g, sth, x, rxb, n, ix, iy, iz, acc);
  }
}

// MODE 0: y = b - A x; 1: y = x + omega D^-1 (b - A x); fixed DOFs of y are 0.
// DIMU 0: runtime dimension, a loop over the planes (large 2D levels: fewer
// registers, two waves per SIMD); 2: 2D with every term unrolled, so the
// forward/backward choice is resolved at compile time and the loads of all
// terms are in flight together (small 2D levels, whose few waves cannot hide
// a chain of dependent loads); 4 and 5: 3D, groups of terms whose loads are
// issued together (st_row_3d with 1 or 3 pairs per group), summed in the
// groups' order.  Forms 0 and 2 sum the same terms in the same order.
// PRO (MODE 1, DIMU 2): x is the iterate before the coarse correction cc,
// which is added on the fly (prolongation and first post-smoothing sweep in
// one pass, with the values the two separate passes give).
// The row of node n (n < g.nnodes): one thread, all NDPN DOFs.
template <int NDPN, typename T, int MODE, int DIMU, bool PRO = false>
__device__ __forceinline__ void st_apply_node(const GridD& g, const T* __restrict__ sth,
                                              const T* __restrict__ x, const T* __restrict__ b,
                                              const T* __restrict__ dinv, T omega,
                                              const uint8_t* __restrict__ mask,
                                              T* __restrict__ y, int64_t n,
                                              const CoarseCorr<T>& cc = CoarseCorr<T>{}) {
  int ix, iy, iz;
  node_xyz(g, n, ix, iy, iz);
  T acc[NDPN];
#pragma unroll
  for (int d = 0; d < NDPN; ++d) acc[d] = T(0);
  if constexpr (DIMU == 4 || DIMU == 5) {
    static_assert(!PRO, "the fused correction is a 2D form");
    st_row_3d<NDPN, T, DIMU == 4 ? 1 : 3>(g, sth, x, uint32_t(n), ix, iy, iz, acc);
  } else if constexpr (DIMU == 0) {
    const bool d3 = g.dim == 3;
    for (int rz = (d3 ? -1 : 0); rz <= (d3 ? 1 : 0); ++rz)
#pragma unroll
      for (int rx = -1; rx <= 1; ++rx)
#pragma unroll
        for (int ry = -1; ry <= 1; ++ry)
          st_term<NDPN, T, 0>(g, sth, x, n, ix, iy, iz, rx, ry, rz, acc);
  } else {
#pragma unroll
    for (int rz = (DIMU == 3 ? -1 : 0); rz <= (DIMU == 3 ? 1 : 0); ++rz)
#pragma unroll
      for (int rx = -1; rx <= 1; ++rx)
#pragma unroll
        for (int ry = -1; ry <= 1; ++ry)
          st_term<NDPN, T, DIMU, PRO>(g, sth, x, n, ix, iy, iz, rx, ry, rz, acc, cc);
  }
  const uint8_t mk = mask[n];
  T pn[NDPN];
  if constexpr (PRO) prolong_at<NDPN, T, DIMU == 3>(cc.gc, cc.xc, ix, iy, iz, pn);
#pragma unroll
  for (int d = 0; d < NDPN; ++d) {
    const int64_t i = n * NDPN + d;
    T xi = x[i];
    if constexpr (PRO) xi += pn[d];  // only read on free DOFs (fixed ones give 0)
    T v;
    if (MODE == 0) v = b[i] - acc[d];
    else v = xi + omega * dinv[i] * (b[i] - acc[d]);
    y[i] = ((mk >> d) & 1) ? T(0) : v;
  }
}

template <int NDPN, typename T, int MODE, int DIMU, bool PRO = false>
__global__ __launch_bounds__(BS) void st_apply_kernel(GridD g, const T* __restrict__ sth,
                                                      const T* __restrict__ x,
                                                      const T* __restrict__ b,
                                                      const T* __restrict__ dinv, T omega,
                                                      const uint8_t* __restrict__ mask,
                                                      T* __restrict__ y, const int* done,
                                                      CoarseCorr<T> cc) {
  static_assert(!PRO || (MODE == 1 && DIMU == 2), "fused correction: unrolled 2D Jacobi sweep");
  if (*done) return;
  const unsigned nb = gridDim.x, hw = blockIdx.x;
  const unsigned xcd = hw & 7u, q = nb >> 3, r = nb & 7u;
  const unsigned logical = xcd * q + min(xcd, r) + (hw >> 3);
  const int64_t n = int64_t(logical) * blockDim.x + threadIdx.x;
  if (n >= g.nnodes) return;
  st_apply_node<NDPN, T, MODE, DIMU, PRO>(g, sth, x, b, dinv, omega, mask, y, n, cc);
}

// Small 3D levels (up to 16,384 nodes): four lanes per node.  Lane q of a
// node's quad loads and sums the terms K = q, q + 4,
// ... of its 27-term row (runtime offsets, so the quad's lanes run one
// instruction stream), all of its loads issued before the first sum; the
// quad's partial sums are then combined by two butterfly exchanges (every
// lane ends with the same total) and lanes 0..NDPN-1 write one DOF each.
// Four times the threads of the one-lane form and a quarter of its loads per
// lane: these levels have too few waves to hide a chain of dependent loads.
// On larger levels (30,625 nodes) the one-lane form with three pairs per
// load group is faster: it already has waves enough, and the quad's extra
// threads and exchanges cost more than the shorter chains save.
template <int NDPN, typename T, int MODE>
__global__ __launch_bounds__(64) void st_apply_quad_kernel(GridD g, const T* __restrict__ sth,
                                                           const T* __restrict__ x,
                                                           const T* __restrict__ b,
                                                           const T* __restrict__ dinv, T omega,
                                                           const uint8_t* __restrict__ mask,
                                                           T* __restrict__ y, const int* done) {
  constexpr int NN = NDPN * NDPN, NT = 7;  // terms per lane (lane 3: 6)
  if (*done) return;
  const unsigned nb = gridDim.x, hw = blockIdx.x;
  const unsigned xcd = hw & 7u, qq = nb >> 3, rr = nb & 7u;
  const unsigned logical = xcd * qq + min(xcd, rr) + (hw >> 3);
  const int64_t n = (int64_t(logical) * 64 + threadIdx.x) >> 2;
  const int q = int(threadIdx.x & 3u);
  if (n >= g.nnodes) return;  // a whole quad: its four lanes share n
  int ix, iy, iz;
  node_xyz(g, n, ix, iy, iz);
  T xm[NT][NDPN], w[NT][NN];
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int K = min(q + 4 * j, 26);  // lane 3's seventh slot repeats term 26 with weight 0
    const bool live = q + 4 * j <= 26;
    const int rz = K / 9 - 1, rx = K / 3 % 3 - 1, ry = K % 3 - 1;
    const int64_t m = node_id(g, min(max(ix + rx, 0), g.nx - 1), min(max(iy + ry, 0), g.ny - 1),
                              min(max(iz + rz, 0), g.nz - 1));
#pragma unroll
    for (int c = 0; c < NDPN; ++c) xm[j][c] = x[m * NDPN + c];
    // forward block of n (K >= 13), else the transposed forward block of m
    const bool fwd = K >= 13;
    const bool ok = live && (fwd || node_in(g, ix + rx, iy + ry, iz + rz));
    const T* h = sth + int64_t(fwd ? K - 13 : 13 - K) * NN * g.nnodes + (fwd ? n : m);
#pragma unroll
    for (int d = 0; d < NDPN; ++d)
#pragma unroll
      for (int c = 0; c < NDPN; ++c) {
        const T v = h[int64_t(fwd ? d * NDPN + c : c * NDPN + d) * g.nnodes];
        w[j][d * NDPN + c] = ok ? v : T(0);
      }
  }
  __builtin_amdgcn_sched_barrier(0);
  T acc[NDPN];
#pragma unroll
  for (int d = 0; d < NDPN; ++d) acc[d] = T(0);
#pragma unroll
  for (int j = 0; j < NT; ++j)
#pragma unroll
    for (int d = 0; d < NDPN; ++d)
#pragma unroll
      for (int c = 0; c < NDPN; ++c) acc[d] += w[j][d * NDPN + c] * xm[j][c];
#pragma unroll
  for (int d = 0; d < NDPN; ++d) {
    acc[d] += __shfl_xor(acc[d], 1, 64);
    acc[d] += __shfl_xor(acc[d], 2, 64);
  }
  if (q >= NDPN) return;
  T a = acc[0];
#pragma unroll
  for (int d = 1; d < NDPN; ++d) a = q == d ? acc[d] : a;
  const int64_t i = n * NDPN + q;
  T v;
  if (MODE == 0) v = b[i] - a;
  else v = x[i] + omega * dinv[i] * (b[i] - a);
  y[i] = ((mask[n] >> q) & 1) ? T(0) : v;
}

// x = omega*dinv*b  (first smoothing sweep from a zero guess); b may be fp64
template <typename T, typename TB>
__global__ void jacobi_zero_kernel(const TB* __restrict__ b, const T* __restrict__ dinv, T omega,
                                   T* __restrict__ x, int64_t n, const int* done) {
  if (*done) return;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    x[i] = zero_guess_sweep<T, TB>(omega, dinv[i], b[i]);
}

// coarse b = P^T r (masked).  The 27 (2D: 9) fine neighbours are read from
// clamped positions, those outside the grid with weight 0, so every load is
// issued before the first sum (a loop that stepped over them waited for each
// load in turn); the summation order is unchanged.  With xc set, the coarse
// level's first smoothing sweep from the zero guess, x = omega D^-1 b, is
// written in the same pass (the coarse cycle then starts from it).
template <int NDPN, typename T, bool D3, bool SEED>
__device__ __forceinline__ void restrict_node(const GridD& gf, const GridD& gc,
                                              const T* __restrict__ rf,
                                              const uint8_t* __restrict__ maskc,
                                              T* __restrict__ bc, const T* __restrict__ dinvc,
                                              T omega, T* __restrict__ xc, int64_t n) {
  int ix, iy, iz;
  node_xyz(gc, n, ix, iy, iz);
  T acc[NDPN];
#pragma unroll
  for (int d = 0; d < NDPN; ++d) acc[d] = T(0);
#pragma unroll
  for (int dz = (D3 ? -1 : 0); dz <= (D3 ? 1 : 0); ++dz)
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy) {
        const int fx = 2 * ix + dx, fy = 2 * iy + dy, fz = D3 ? 2 * iz + dz : 0;
        const bool ok = node_in(gf, fx, fy, fz);
        const int64_t fi = node_id(gf, min(max(fx, 0), gf.nx - 1), min(max(fy, 0), gf.ny - 1),
                                   min(max(fz, 0), gf.nz - 1));
        // a zero weight (not a branch) for neighbours outside the grid, so
        // that the loads stay unconditional; their clamped values are finite
        const T w = ok ? T((dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0) * (dz ? 0.5 : 1.0)) : T(0);
#pragma unroll
        for (int d = 0; d < NDPN; ++d) acc[d] += w * rf[fi * NDPN + d];
      }
  const uint8_t mk = maskc[n];
#pragma unroll
  for (int d = 0; d < NDPN; ++d) {
    const T v = ((mk >> d) & 1) ? T(0) : acc[d];
    bc[n * NDPN + d] = v;
    if constexpr (SEED) xc[n * NDPN + d] = zero_guess_sweep<T, T>(omega, dinvc[n * NDPN + d], v);
  }
}

template <int NDPN, typename T, bool D3, bool SEED>
__global__ void restrict_kernel(GridD gf, GridD gc, const T* __restrict__ rf,
                                const uint8_t* __restrict__ maskc, T* __restrict__ bc,
                                const T* __restrict__ dinvc, T omega, T* __restrict__ xc,
                                const int* done) {
  if (*done) return;
  const int64_t n = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (n >= gc.nnodes) return;
  restrict_node<NDPN, T, D3, SEED>(gf, gc, rf, maskc, bc, dinvc, omega, xc, n);
}

// fine x += P xc (masked), node n of the fine grid
template <int NDPN, typename T, bool D3>
__device__ __forceinline__ void prolong_node(const GridD& gf, const GridD& gc,
                                             const T* __restrict__ xc,
                                             const uint8_t* __restrict__ maskf,
                                             T* __restrict__ xf, int64_t n) {
  int ix, iy, iz;
  node_xyz(gf, n, ix, iy, iz);
  T acc[NDPN];
  prolong_at<NDPN, T, D3>(gc, xc, ix, iy, iz, acc);
  const uint8_t mk = maskf[n];
#pragma unroll
  for (int d = 0; d < NDPN; ++d)
    if (!((mk >> d) & 1)) xf[n * NDPN + d] += acc[d];
}

template <int NDPN, typename T, bool D3>
__global__ void prolong_add_kernel(GridD gf, GridD gc, const T* __restrict__ xc,
                                   const uint8_t* __restrict__ maskf, T* __restrict__ xf,
                                   const int* done) {
  if (*done) return;
  const int64_t n = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (n >= gf.nnodes) return;
  prolong_node<NDPN, T, D3>(gf, gc, xc, maskf, xf, n);
}

// ---------------------------------------------------------------------------
// Multi-instance forms for lockstep batched solves: one launch covers the
// blocks of many systems; block b computes local block tiles[b].y of the
// system whose argument struct is descs[tiles[b].x], with the same per-node
// arithmetic as the single-system kernel.
// ---------------------------------------------------------------------------
template <typename T>
struct RestrictArgs {
  GridD gf, gc;
  const T* rf;
  const uint8_t* maskc;
  T* bc;
  const T* dinvc;
  T omega;
  T* xc;
  const int* done;
};

template <int NDPN, typename T, bool D3, bool SEED>
__global__ __launch_bounds__(BS) void restrict_multi_kernel(const RestrictArgs<T>* __restrict__ descs,
                                                            const uint2* __restrict__ tiles) {
  const uint2 t = tiles[blockIdx.x];
  const RestrictArgs<T> a = descs[t.x];
  if (*a.done) return;
  const int64_t n = int64_t(t.y) * BS + threadIdx.x;
  if (n >= a.gc.nnodes) return;
  restrict_node<NDPN, T, D3, SEED>(a.gf, a.gc, a.rf, a.maskc, a.bc, a.dinvc, a.omega, a.xc, n);
}

template <int NDPN, typename T, bool D3, bool SEED>
int restrict_multi_launch(hipStream_t s, const void* descs, const uint2* tiles, unsigned nb,
                          size_t lds) {
  hipLaunchKernelGGL((restrict_multi_kernel<NDPN, T, D3, SEED>), dim3(nb), dim3(BS), lds, s,
                     static_cast<const RestrictArgs<T>*>(descs), tiles);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

template <typename T>
struct ProlongArgs {
  GridD gf, gc;
  const T* xc;
  const uint8_t* maskf;
  T* xf;
  const int* done;
};

template <int NDPN, typename T, bool D3>
__global__ __launch_bounds__(BS) void prolong_multi_kernel(const ProlongArgs<T>* __restrict__ descs,
                                                           const uint2* __restrict__ tiles) {
  const uint2 t = tiles[blockIdx.x];
  const ProlongArgs<T> a = descs[t.x];
  if (*a.done) return;
  const int64_t n = int64_t(t.y) * BS + threadIdx.x;
  if (n >= a.gf.nnodes) return;
  prolong_node<NDPN, T, D3>(a.gf, a.gc, a.xc, a.maskf, a.xf, n);
}

template <int NDPN, typename T, bool D3>
int prolong_multi_launch(hipStream_t s, const void* descs, const uint2* tiles, unsigned nb,
                         size_t lds) {
  hipLaunchKernelGGL((prolong_multi_kernel<NDPN, T, D3>), dim3(nb), dim3(BS), lds, s,
                     static_cast<const ProlongArgs<T>*>(descs), tiles);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

template <typename T>
struct StArgs {
  GridD g;
  const T* sth;
  const T* x;
  const T* b;
  const T* dinv;
  T omega;
  const uint8_t* mask;
  T* y;
  const int* done;
  CoarseCorr<T> cc;
};

template <int NDPN, typename T, int MODE, int DIMU, bool PRO>
__global__ __launch_bounds__(BS) void st_apply_multi_kernel(const StArgs<T>* __restrict__ descs,
                                                            const uint2* __restrict__ tiles) {
  const uint2 t = tiles[blockIdx.x];
  const StArgs<T> a = descs[t.x];
  if (*a.done) return;
  const int64_t n = int64_t(t.y) * BS + threadIdx.x;
  if (n >= a.g.nnodes) return;
  st_apply_node<NDPN, T, MODE, DIMU, PRO>(a.g, a.sth, a.x, a.b, a.dinv, a.omega, a.mask, a.y, n,
                                          a.cc);
}

template <int NDPN, typename T, int MODE, int DIMU, bool PRO>
int st_apply_multi_launch(hipStream_t s, const void* descs, const uint2* tiles, unsigned nb,
                          size_t lds) {
  hipLaunchKernelGGL((st_apply_multi_kernel<NDPN, T, MODE, DIMU, PRO>), dim3(nb), dim3(BS), lds,
                     s, static_cast<const StArgs<T>*>(descs), tiles);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

template <int NDPN, typename T>
int launch_restrict(tl_problem* P, const Grid& gf, const Grid& gc, const T* rf,
                    const uint8_t* maskc, T* bc, const T* dinvc, T omega, T* xc,
                    const int* done) {
  const dim3 grid(unsigned((gc.nnodes + BS - 1) / BS));
  auto go = [&](auto d3_c, auto seed_c) {
    constexpr bool D3_ = decltype(d3_c)::value, SEED_ = decltype(seed_c)::value;
    if (P->rec) {
      P->rec->add(P->rec_phase, P->rec_level, &restrict_multi_launch<NDPN, T, D3_, SEED_>,
                  RestrictArgs<T>{to_d(gf), to_d(gc), rf, maskc, bc, dinvc, omega, xc, done},
                  grid.x, 0);
      return;
    }
    hipLaunchKernelGGL((restrict_kernel<NDPN, T, D3_, SEED_>), grid, dim3(BS), 0, P->stream,
                       to_d(gf), to_d(gc), rf, maskc, bc, dinvc, omega, xc, done);
  };
  using F = std::false_type;
  using Tr = std::true_type;
  if (gf.dim == 3) { if (xc) go(Tr{}, Tr{}); else go(Tr{}, F{}); }
  else { if (xc) go(F{}, Tr{}); else go(F{}, F{}); }
  TL_CHECK_LAUNCH();
  return TL_OK;
}

template <int NDPN, typename T>
int launch_prolong(tl_problem* P, const Grid& gf, const Grid& gc, const T* xc,
                   const uint8_t* maskf, T* xf, const int* done) {
  const dim3 grid(unsigned((gf.nnodes + BS - 1) / BS));
  if (P->rec) {
    const ProlongArgs<T> a{to_d(gf), to_d(gc), xc, maskf, xf, done};
    if (gf.dim == 3)
      P->rec->add(P->rec_phase, P->rec_level, &prolong_multi_launch<NDPN, T, true>, a, grid.x, 0);
    else
      P->rec->add(P->rec_phase, P->rec_level, &prolong_multi_launch<NDPN, T, false>, a, grid.x, 0);
    return TL_OK;
  }
  if (gf.dim == 3)
    hipLaunchKernelGGL((prolong_add_kernel<NDPN, T, true>), grid, dim3(BS), 0, P->stream, to_d(gf),
                       to_d(gc), xc, maskf, xf, done);
  else
    hipLaunchKernelGGL((prolong_add_kernel<NDPN, T, false>), grid, dim3(BS), 0, P->stream,
                       to_d(gf), to_d(gc), xc, maskf, xf, done);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

// x_v[row] = (Cinv b_v)[row] for NV vectors sharing one pass over the row
// (one wavefront per row; inactive vectors skipped); each vector's sum takes
// the terms in the same order as gemv_row below.
template <typename T, int NV>
__device__ __forceinline__ void gemv_row_nv(const double* __restrict__ A, int64_t np, int64_t n,
                                            const T* const (&b)[NV], T* const (&x)[NV],
                                            const bool (&on)[NV], int64_t row, int lane) {
  const double* __restrict__ Ar = A + row * np;
  double s[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) s[v] = 0.0;
  constexpr int U = 8;
  int64_t j = lane;
  for (; j + 64 * (U - 1) < n; j += 64 * U) {
    double a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] = Ar[j + 64 * u];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      if (!on[v]) continue;
      double w[U];
#pragma unroll
      for (int u = 0; u < U; ++u) w[u] = double(b[v][j + 64 * u]);
#pragma unroll
      for (int u = 0; u < U; ++u) s[v] += a[u] * w[u];
    }
  }
  for (; j < n; j += 64) {
    const double a = Ar[j];
#pragma unroll
    for (int v = 0; v < NV; ++v)
      if (on[v]) s[v] += a * double(b[v][j]);
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    if (!on[v]) continue;
    const double t = wave_sum(s[v]);
    if (lane == 0) x[v][row] = T(t);
  }
}

// x[row] = (Cinv b)[row], one wavefront per row (Cinv fp64, padded leading
// dimension np)
template <typename T>
__device__ __forceinline__ void gemv_row(const double* __restrict__ A, int64_t np, int64_t n,
                                         const T* __restrict__ b, T* __restrict__ x, int64_t row,
                                         int lane) {
  const double* __restrict__ Ar = A + row * np;
  double s = 0.0;
  // U loads per lane in flight before the first FMA (a plain loop waited for
  // each load: ~30 dependent round trips per row); the FMAs keep the column
  // order j = lane, lane + 64, ...
  constexpr int U = 8;
  int64_t j = lane;
  for (; j + 64 * (U - 1) < n; j += 64 * U) {
    double a[U], v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      a[u] = Ar[j + 64 * u];
      v[u] = double(b[j + 64 * u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) s += a[u] * v[u];
  }
  for (; j < n; j += 64) s += Ar[j] * double(b[j]);
  s = wave_sum(s);
  if (lane == 0) x[row] = T(s);
}

// x = Cinv b on the coarsest grid
template <typename T>
__global__ __launch_bounds__(BS) void dense_gemv_kernel(const double* __restrict__ A, int64_t np,
                                                        int64_t n, const T* __restrict__ b,
                                                        T* __restrict__ x, const int* done) {
  if (*done) return;
  const int64_t row = blockIdx.x * int64_t(BS / 64) + (threadIdx.x >> 6);
  if (row >= n) return;
  gemv_row<T>(A, np, n, b, x, row, threadIdx.x & 63);
}

template <typename T>
struct GemvArgs {
  const double* A;
  int64_t np, n;
  const T* b;
  T* x;
  const int* done;
};

template <typename T>
__global__ __launch_bounds__(BS) void gemv_multi_kernel(const GemvArgs<T>* __restrict__ descs,
                                                        const uint2* __restrict__ tiles) {
  const uint2 t = tiles[blockIdx.x];
  const GemvArgs<T> a = descs[t.x];
  if (*a.done) return;
  const int64_t row = int64_t(t.y) * (BS / 64) + (threadIdx.x >> 6);
  if (row >= a.n) return;
  gemv_row<T>(a.A, a.np, a.n, a.b, a.x, row, threadIdx.x & 63);
}

template <typename T>
int gemv_multi_launch(hipStream_t s, const void* descs, const uint2* tiles, unsigned nb,
                      size_t lds) {
  hipLaunchKernelGGL((gemv_multi_kernel<T>), dim3(nb), dim3(BS), lds, s,
                     static_cast<const GemvArgs<T>*>(descs), tiles);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

// Up to kRecPack systems multiplying one coarsest inverse (the load cases of
// one instance): the matrix is read once for all of them.
template <typename T>
struct GemvPackArgs {
  const double* A;
  int64_t np, n;
  int nv;
  const T* b[kRecPack];
  T* x[kRecPack];
  const int* done[kRecPack];
};

template <typename T>
__global__ __launch_bounds__(BS) void gemv_pack_kernel(const GemvPackArgs<T>* __restrict__ descs,
                                                       const uint2* __restrict__ tiles) {
  const uint2 t = tiles[blockIdx.x];
  const GemvPackArgs<T>& a = descs[t.x];
  bool on[kRecPack];
  bool any = false;
#pragma unroll
  for (int v = 0; v < kRecPack; ++v) {
    on[v] = v < a.nv && !*a.done[v];
    any = any || on[v];
  }
  if (!any) return;
  const int64_t row = int64_t(t.y) * (BS / 64) + (threadIdx.x >> 6);
  if (row >= a.n) return;
  const T* b[kRecPack];
  T* x[kRecPack];
#pragma unroll
  for (int v = 0; v < kRecPack; ++v) { b[v] = a.b[v]; x[v] = a.x[v]; }
  gemv_row_nv<T, kRecPack>(a.A, a.np, a.n, b, x, on, row, threadIdx.x & 63);
}

template <typename T>
int gemv_pack_launch(hipStream_t s, const void* descs, const uint2* tiles, unsigned nb,
                     size_t lds) {
  hipLaunchKernelGGL((gemv_pack_kernel<T>), dim3(nb), dim3(BS), lds, s,
                     static_cast<const GemvPackArgs<T>*>(descs), tiles);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

template <typename T>
unsigned gemv_pack(const RecOp* const* ops, int nops, std::vector<unsigned char>& out) {
  GemvPackArgs<T> p{};
  for (int k = 0; k < nops; ++k) {
    GemvArgs<T> g;
    std::memcpy(&g, ops[k]->args.data(), sizeof(g));
    if (k == 0) { p.A = g.A; p.np = g.np; p.n = g.n; }
    p.b[k] = g.b;
    p.x[k] = g.x;
    p.done[k] = g.done;
  }
  p.nv = nops;
  for (int k = nops; k < kRecPack; ++k) p.done[k] = p.done[0];
  const auto* c = reinterpret_cast<const unsigned char*>(&p);
  out.insert(out.end(), c, c + sizeof(p));
  return unsigned((p.n + 3) / 4);
}


// ---------------------------------------------------------------------------
// dense coarsest operator and its in-place blocked Gauss-Jordan inverse
// ---------------------------------------------------------------------------
__global__ void dense_from_stencil_kernel(GridD g, int ndpn, const double* __restrict__ st,
                                          double* __restrict__ A, int64_t np) {
  const int nq = nstencil(g.dim) * ndpn * ndpn;
  const int64_t tot = g.nnodes * nq;
  for (int64_t t = blockIdx.x * int64_t(BS) + threadIdx.x; t < tot;
       t += int64_t(gridDim.x) * BS) {
    const int64_t n = t % g.nnodes;
    const int q = int(t / g.nnodes);
    const int c = q % ndpn, d = (q / ndpn) % ndpn, k = q / (ndpn * ndpn);
    int rx, ry, rz;
    if (g.dim == 3) { rz = k / 9 - 1; rx = (k / 3) % 3 - 1; ry = k % 3 - 1; }
    else { rz = 0; rx = k / 3 - 1; ry = k % 3 - 1; }
    int ix, iy, iz;
    node_xyz(g, n, ix, iy, iz);
    if (!node_in(g, ix + rx, iy + ry, iz + rz)) continue;
    const int64_t m = node_id(g, ix + rx, iy + ry, iz + rz);
    A[(n * ndpn + d) * np + m * ndpn + c] = st[t];
  }
}

__global__ void pad_identity_kernel(double* A, int64_t n, int64_t np) {
  const int64_t i = n + blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < np) A[i * np + i] = 1.0;
}

// Blocked Gauss-Jordan inverse of the symmetric positive definite coarsest
// matrix, one launch per block step k (pivot block K = [kb, kb+GJB)):
//   new row panel   A[K][j] = X[j] = P^-1 R[j]      (j outside K, R = old A[K][j])
//   trailing update A[i][j] -= C[i] X[j]           (C = old A[i][K])
//   column panel    A[i][K] = -C[i] P^-1, pivot block A[K][K] = P^-1.
// Every intermediate satisfies A[i][j] = s(i) s(j) A[j][i] with s = -1 on the
// indices already swept (below kb), so only the lower 64x64 tiles (tile row
// >= tile column; diagonal tiles whole) are updated, and the column panel is
// C[i][t] = s(i) R[t][i].  Step k reads R (the full row panel of K) and P^-1
// from buffers that step k-1 wrote, so no tile reads a value that another
// tile of the same launch writes.  Each tile forms its own columns of X and
// writes its part of the next row panel: the tiles of the next pivot's tile
// row their rows of it, the tiles below the next pivot tile their columns of
// the next block, transposed (those rows are not swept, s = +1).  The tile
// that holds the next pivot block inverts it.  After the last step the lower
// tiles hold the (symmetric) inverse and are mirrored into the upper ones.

// In-place Gauss-Jordan inverse of the GJB x GJB LDS block by wave 0; every
// thread of the workgroup calls it.  Lane l keeps row l & 31, columns
// 16 (l >> 5) .. + 15 in registers; each step publishes the pivot row and
// column through `pub` (2 x 64 doubles).  One wave's LDS operations execute
// in order, so the steps need no workgroup barrier.
__device__ inline void gj_invert_pivot(double (*a)[GJB + 1], double* pub) {
  static_assert(GJB == 32, "one wave: 32 rows x 2 column halves");
  if (threadIdx.x < 64) {
    const int r = threadIdx.x & 31, h = threadIdx.x >> 5;
    double v[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) v[c] = a[r][16 * h + c];
#pragma unroll
    for (int t = 0; t < GJB; ++t) {
      double* rb = pub + (t & 1) * 64;  // row t
      double* cb = rb + 32;             // column t
      if (r == t) {
#pragma unroll
        for (int c = 0; c < 16; ++c) rb[16 * h + c] = v[c];
      }
      if (h == (t >> 4)) cb[r] = v[t & 15];
      __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the step's writes have landed
      __builtin_amdgcn_wave_barrier();
      const double att = rb[t], ait = cb[r];
      double rt[16];
#pragma unroll
      for (int c = 0; c < 16; ++c) rt[c] = rb[16 * h + c];
      const double inv = 1.0 / att;  // one reciprocal per step
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        const bool jt = h == (t >> 4) && c == (t & 15);
        const double atj = rt[c] * inv;
        double nv;
        if (r == t && jt) nv = inv;
        else if (r == t) nv = atj;
        else if (jt) nv = -ait * inv;
        else nv = v[c] - ait * atj;
        v[c] = nv;
      }
      __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) a[r][16 * h + c] = v[c];
  }
  __syncthreads();
}

// Step-0 buffers: the row panel R = A[0..GJB)[j] for the workgroup's 64
// columns; workgroup 0 also inverts the first pivot block.
__global__ __launch_bounds__(BS) void gj_init_kernel(const double* __restrict__ A, int64_t np,
                                                     double* __restrict__ rowp,
                                                     double* __restrict__ pinv) {
  __shared__ double a[GJB][GJB + 1];
  __shared__ double pub[128];
  const int64_t c0 = int64_t(blockIdx.x) * 64;
  for (int e = threadIdx.x; e < 64 * GJB; e += BS) {
    const int64_t j = c0 + e % 64;
    const int t = e / 64;
    if (j < np) rowp[t * np + j] = A[t * np + j];
  }
  if (blockIdx.x != 0) return;
  for (int e = threadIdx.x; e < GJB * GJB; e += BS) a[e / GJB][e % GJB] = A[(e / GJB) * np + e % GJB];
  __syncthreads();
  gj_invert_pivot(a, pub);
  for (int e = threadIdx.x; e < GJB * GJB; e += BS) pinv[e] = a[e / GJB][e % GJB];
}

// lower tile (I, J), I >= J, of linear index b = I (I + 1) / 2 + J
__device__ inline void gj_tile_of(unsigned b, unsigned& I, unsigned& J) {
  unsigned i = unsigned((sqrtf(8.0f * float(b) + 1.0f) - 1.0f) * 0.5f);
  while ((i + 1) * (i + 2) / 2 <= b) ++i;
  while (i * (i + 1) / 2 > b) --i;
  I = i;
  J = b - i * (i + 1) / 2;
}

// Step k on the lower 64x64 tiles.  S = C X on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64), wave w computing rows 16w..16w+15 as 4 blocks of
// 16x16, GJB/4 K slices each.  Lane layouts: lane l holds A[l%16][l/16] and
// B[l/16][l%16] of a K slice; output element i of lane l is C[l/16 + 4i][l%16].
// The diagonal tile that holds the next pivot block (rows and columns
// [kb + GJB, kb + 2 GJB)) runs first (hardware block 0) and inverts that
// block while the other tiles are updated.
typedef double gj_d4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(BS) void gj_step_kernel(double* A, int64_t np, int k, int nbk,
                                                     const double* __restrict__ rowp,
                                                     const double* __restrict__ pinv,
                                                     double* __restrict__ rowp_next,
                                                     double* __restrict__ pinv_next) {
  static_assert(BS == 256 && GJB % 4 == 0 && GJB == 32, "four waves, K slices of 4");
  __shared__ double cp[64][GJB + 1];   // C rows of the tile
  __shared__ double xp[GJB][64 + 1];   // X columns of the tile
  __shared__ double pb[GJB][GJB + 1];  // the next pivot block (its tile only)
  __shared__ double pub[128];
  const int64_t kb = int64_t(k) * GJB, kn = kb + GJB;
  const bool has_next = k + 1 < nbk;
  const unsigned tn = has_next ? unsigned(kn / 64) : ~0u;    // tile of the next pivot
  const unsigned own = has_next ? tn * (tn + 1) / 2 + tn : 0u;  // its linear index
  unsigned lb = blockIdx.x;
  if (lb == 0) lb = own;
  else if (lb == own) lb = 0;
  const bool own_next = has_next && lb == own;
  unsigned ti, tj;
  gj_tile_of(lb, ti, tj);
  const int64_t i0 = int64_t(ti) * 64, j0 = int64_t(tj) * 64;
  double aold[4][4];  // the tile's A values, requested first
  {
    const int ww = threadIdx.x >> 6, ll = threadIdx.x & 63;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t i = i0 + 16 * ww + (ll >> 4) + 4 * e, jj = j0 + 16 * b + (ll & 15);
        aold[b][e] = (i < np && jj < np) ? A[i * np + jj] : 0.0;
      }
  }
  // C[r][t] = s(i) R[t][i] for the tile's rows i = i0 + r (0 past np)
  for (int e = threadIdx.x; e < 64 * GJB; e += BS) {
    const int t = e / 64, r = e % 64;
    const int64_t i = i0 + r;
    const double v = i < np ? rowp[t * np + i] : 0.0;
    cp[r][t] = i < kb ? -v : v;
  }
  // R rows of the tile's columns (0 inside K and past np), staged in xp
  for (int e = threadIdx.x; e < GJB * 64; e += BS) {
    const int t = e / 64, c = e % 64;
    const int64_t jj = j0 + c;
    xp[t][c] = (jj < np && (jj < kb || jj >= kb + GJB)) ? rowp[t * np + jj] : 0.0;
  }
  __syncthreads();
  // X column j0 + c2, rows rb .. rb + 7 (one wave): P^-1 R[j] outside K,
  // P^-1 itself inside K, 0 past np.  P^-1 rows are wave-uniform (scalar loads).
  const int c2 = threadIdx.x & 63;
  const int rb = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * (GJB / 4);
  const int64_t j = j0 + c2;
  const bool jin = j >= kb && j < kb + GJB;
  const bool active = j < np && !jin;
  double col[GJB];
#pragma unroll
  for (int t = 0; t < GJB; ++t) col[t] = xp[t][c2];
  __syncthreads();
#pragma unroll
  for (int r = rb; r < rb + GJB / 4; ++r) {
    double sum = 0.0;
    if (active) {
#pragma unroll
      for (int t = 0; t < GJB; ++t) sum += pinv[r * GJB + t] * col[t];
    } else if (jin) {
      sum = pinv[r * GJB + (j - kb)];
    }
    xp[r][c2] = sum;
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, lr = l & 15, lk = l >> 4;
  gj_d4 acc[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) acc[b] = gj_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int sl = 0; sl < GJB / 4; ++sl) {
    const double av = cp[16 * w + lr][4 * sl + lk];
#pragma unroll
    for (int b = 0; b < 4; ++b)
      acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, xp[4 * sl + lk][16 * b + lr], acc[b], 0, 0, 0);
  }
  // next row panel: rows of the next block in its tile row (direct), its
  // columns in the tiles below its tile (transposed)
  const bool below = has_next && ti > tn;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int64_t jj = j0 + 16 * b + lr;
    if (jj >= np) continue;
    const bool jcol = jj >= kb && jj < kb + GJB;
    const bool jnext = has_next && jj >= kn && jj < kn + GJB;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t i = i0 + 16 * w + lk + 4 * e;
      if (i >= np) continue;
      double v;
      if (i >= kb && i < kb + GJB) v = xp[i - kb][16 * b + lr];  // row panel, P^-1 inside K
      else if (jcol) v = -acc[b][e];                              // column panel
      else v = aold[b][e] - acc[b][e];                            // trailing update
      A[i * np + jj] = v;
      const bool inext = has_next && i >= kn && i < kn + GJB;
      if (inext) rowp_next[(i - kn) * np + jj] = v;
      if (below && jnext) rowp_next[(jj - kn) * np + i] = v;
      if (own_next && inext && jnext) pb[i - kn][jj - kn] = v;
    }
  }
  if (!own_next) return;
  __syncthreads();
  gj_invert_pivot(pb, pub);
  for (int e = threadIdx.x; e < GJB * GJB; e += BS) pinv_next[e] = pb[e / GJB][e % GJB];
}

// upper tile (I, J), I < J, of linear index b = J (J - 1) / 2 + I: the
// transpose of lower tile (J, I)
__global__ __launch_bounds__(BS) void gj_mirror_kernel(double* A, int64_t np) {
  __shared__ double tl[64][64 + 1];
  unsigned J, I;
  gj_tile_of(blockIdx.x, J, I);  // lower index pair (J', I) with I <= J'
  ++J;                           // strictly lower: J' = J - 1 enumerates J >= I + 1
  const int64_t i0 = int64_t(I) * 64, j0 = int64_t(J) * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += BS) {
    const int r = e / 64, c = e % 64;  // row j0 + r, column i0 + c of the lower tile
    if (j0 + r < np && i0 + c < np) tl[r][c] = A[(j0 + r) * np + i0 + c];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += BS) {
    const int r = e / 64, c = e % 64;  // row i0 + r, column j0 + c of the upper tile
    if (i0 + r < np && j0 + c < np) A[(i0 + r) * np + j0 + c] = tl[c][r];
  }
}

unsigned gsz(int64_t work) {
  return unsigned(std::max<int64_t>(1, std::min<int64_t>((work + BS - 1) / BS, 1 << 20)));
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
// Child interpolation W_c[i][j]: weight of coarse local node j at fine local
// node i of child c (child offset (cx,cy,cz), c = (cz*2+cx)*2+cy).
void child_weights(int dim, std::vector<double>& W) {
  const int nn = dim == 3 ? 8 : 4, nch = nn;
  W.assign(size_t(nch) * nn * nn, 0.0);
  for (int ch = 0; ch < nch; ++ch) {
    const int cy = ch & 1, cx = (ch >> 1) & 1, cz = ch >> 2;
    for (int i = 0; i < nn; ++i) {
      const double px = 0.5 * (cx + loc_dx(i)), py = 0.5 * (cy + loc_dy(i)),
                   pz = 0.5 * (cz + loc_dz(i));
      for (int j = 0; j < nn; ++j) {
        double w = (loc_dx(j) ? px : 1 - px) * (loc_dy(j) ? py : 1 - py);
        if (dim == 3) w *= (loc_dz(j) ? pz : 1 - pz);
        W[(size_t(ch) * nn + i) * nn + j] = w;
      }
    }
  }
}

// per-node fixed bit masks of every level; fails if the Dirichlet set is not
// grid aligned (a fixed DOF interpolated from a free coarse DOF)
int level_masks(const tl_problem* P, int nlevels, std::vector<std::vector<uint8_t>>& masks) {
  masks.assign(nlevels, {});
  masks[0].assign(P->g.nnodes, 0);
  for (int64_t f : P->fixed_h) masks[0][f / P->ndpn] |= uint8_t(1u << (f % P->ndpn));
  Grid gf = P->g;
  for (int l = 1; l < nlevels; ++l) {
    Grid gc;
    gc.set(gf.dim, gf.nelx / 2, gf.nely / 2, gf.nelz / 2);
    auto& mc = masks[l];
    const auto& mf = masks[l - 1];
    mc.assign(gc.nnodes, 0);
    for (int iz = 0; iz < gc.nz; ++iz)
      for (int ix = 0; ix < gc.nx; ++ix)
        for (int iy = 0; iy < gc.ny; ++iy)
          mc[(int64_t(iz) * gc.nx + ix) * gc.ny + iy] =
              mf[(int64_t(2 * iz) * gf.nx + 2 * ix) * gf.ny + 2 * iy];
    for (int iz = 0; iz < gf.nz; ++iz)
      for (int ix = 0; ix < gf.nx; ++ix)
        for (int iy = 0; iy < gf.ny; ++iy) {
          const uint8_t m = mf[(int64_t(iz) * gf.nx + ix) * gf.ny + iy];
          if (!m) continue;
          for (int cz = iz >> 1; cz <= (iz + 1) >> 1; ++cz)
            for (int cx = ix >> 1; cx <= (ix + 1) >> 1; ++cx)
              for (int cy = iy >> 1; cy <= (iy + 1) >> 1; ++cy) {
                const uint8_t pc = mc[(int64_t(cz) * gc.nx + cx) * gc.ny + cy];
                if ((m & pc) != m) {
                  set_error("multigrid needs grid-aligned Dirichlet DOFs: a fixed DOF "
                            "interpolates from a free coarse DOF");
                  return TL_EINVAL;
                }
              }
        }
    gf = gc;
  }
  return TL_OK;
}

template <int NDPN, typename T>
int vcycle_level(tl_problem* P, int l, const int* done);

template <typename T>
T* vp(const DevBuf& b) { return static_cast<T*>(b.ptr); }

// Host-side guard: every buffer a V-cycle kernel dereferences must exist.
#define TL_NONNULL(p) TL_ARG((p) != nullptr, "internal: unallocated multigrid buffer " #p)

// Restriction of a residual into level l+1.  Below it a stencil level starts
// its cycle from x = omega D^-1 b, written by the same kernel; the coarsest
// level (dense solve) takes b alone.
template <int NDPN, typename T>
int restrict_to(tl_problem* P, int l, const Grid& gf, const T* rf, const int* done) {
  MgState& M = P->mg;
  Level& C = M.lv[l + 1];
  const bool seed = l + 1 < M.nlevels - 1;
  const T* dinvc = std::is_same<T, double>::value ? static_cast<const T*>(C.dinv64.ptr)
                                                   : static_cast<const T*>(C.dinv32.ptr);
  if (seed) { TL_NONNULL(dinvc); TL_NONNULL(C.vx.ptr); }
  return launch_restrict<NDPN, T>(P, gf, C.g, rf, C.mask.as<uint8_t>(), vp<T>(C.vb),
                                  seed ? dinvc : nullptr, T(M.omega), seed ? vp<T>(C.vx) : nullptr,
                                  done);
}

// Stencil form of a level: 3D levels take the grouped form while each
// forward-half slab (NDPN^2 FP64 arrays) fits one buffer (below 4 GiB, i.e.
// up to 59 million nodes), with 1 pair per group above 64k nodes (4) and 3
// pairs up to 64k (5); 2D levels of up to 64k nodes (a few hundred
// workgroups) take the unrolled form (2), the rest the looped form (0).  NDPN 2
// only occurs in 2D and NDPN 3 in 3D for the dimension-specific forms.
template <int NDPN>
int st_dimu(const Grid& g) {
  if (g.dim == 3 && NDPN != 2) {
    if (g.nnodes * NDPN * NDPN * 8 >= (int64_t(1) << 32)) return 0;
    return g.nnodes > 65536 ? 4 : 5;
  }
  if (g.dim == 2 && NDPN != 3 && g.nnodes <= 65536) return 2;
  return 0;
}

// Workgroup size of the grouped 3D form: 64 threads spread a level over 4x
// as many workgroups, which balances them across the 256 CUs (level 1 of the
// 192x96x96 problem: 3,639 workgroups instead of 910, residual 73 -> 69 us;
// levels 2-3: 120-480 workgroups instead of 17-120).
constexpr unsigned kStBlock3d = 64;

template <int NDPN, typename T>
int st_apply(tl_problem* P, int l, int mode, const T* x, const T* b, T* y, const int* done) {
  Level& L = P->mg.lv[l];
  TL_ARG(l >= 1 && l < P->mg.nlevels, "internal: stencil level out of range");
  TL_NONNULL(L.sth64.ptr); TL_NONNULL(L.st32.ptr); TL_NONNULL(L.dinv64.ptr);
  TL_NONNULL(L.dinv32.ptr); TL_NONNULL(L.mask.ptr); TL_NONNULL(x); TL_NONNULL(b); TL_NONNULL(y);
  const T* st = std::is_same<T, double>::value ? static_cast<const T*>(L.sth64.ptr)
                                                : static_cast<const T*>(L.st32.ptr);
  const T* dinv = std::is_same<T, double>::value ? static_cast<const T*>(L.dinv64.ptr)
                                                  : static_cast<const T*>(L.dinv32.ptr);
  const unsigned nb = unsigned((L.g.nnodes + BS - 1) / BS);
  // launch timing of the first coarse level: kind 3 residual, kind 4 Jacobi
  ProfEvent ev{};
  ev.kind = -1;
  if (l == 1) TL_TRY(prof_begin(P, 3 + mode, ev));
  const int dimu = st_dimu<NDPN>(L.g);
  const GridD gd = to_d(L.g);
  const T om = T(P->mg.omega);
  const uint8_t* mk = L.mask.as<uint8_t>();
  auto go = [&](auto mode_c, auto dim_c) {
    constexpr int MODE_ = decltype(mode_c)::value, DIM_ = decltype(dim_c)::value;
    if (P->rec) {
      P->rec->add(P->rec_phase, P->rec_level, &st_apply_multi_launch<NDPN, T, MODE_, DIM_, false>,
                  StArgs<T>{gd, st, x, b, dinv, om, mk, y, done, CoarseCorr<T>{}}, nb, 0);
      return;
    }
    const unsigned bs = DIM_ >= 4 ? kStBlock3d : unsigned(BS);
    hipLaunchKernelGGL((st_apply_kernel<NDPN, T, MODE_, DIM_>),
                       dim3(unsigned((L.g.nnodes + bs - 1) / bs)), dim3(bs), 0, P->stream, gd, st,
                       x, b, dinv, om, mk, y, done, CoarseCorr<T>{});
  };
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I4 = std::integral_constant<int, 4>;
  using I5 = std::integral_constant<int, 5>;
  if (dimu == 4) {
    if constexpr (NDPN != 2) { if (mode == 0) go(I0{}, I4{}); else go(I1{}, I4{}); }
  } else if (dimu == 5 && !P->rec && L.g.nnodes <= 16384) {
    // four lanes per node on the smallest levels (3D levels are never
    // recorded for a lockstep batch)
    if constexpr (NDPN != 2) {
      const unsigned nq = unsigned((4 * L.g.nnodes + kStBlock3d - 1) / kStBlock3d);
      if (mode == 0)
        hipLaunchKernelGGL((st_apply_quad_kernel<NDPN, T, 0>), dim3(nq), dim3(kStBlock3d), 0,
                           P->stream, gd, st, x, b, dinv, om, mk, y, done);
      else
        hipLaunchKernelGGL((st_apply_quad_kernel<NDPN, T, 1>), dim3(nq), dim3(kStBlock3d), 0,
                           P->stream, gd, st, x, b, dinv, om, mk, y, done);
    }
  } else if (dimu == 5) {
    if constexpr (NDPN != 2) { if (mode == 0) go(I0{}, I5{}); else go(I1{}, I5{}); }
  } else if (dimu == 2) {
    if constexpr (NDPN != 3) { if (mode == 0) go(I0{}, I2{}); else go(I1{}, I2{}); }
  } else {
    if (mode == 0) go(I0{}, I0{}); else go(I1{}, I0{});
  }
  TL_CHECK_LAUNCH();
  return prof_end(P, ev);
}

// Whether the prolongation into level l is fused with its first post-smoothing
// sweep: on the small (unrolled-form) 2D levels.  In 3D the 27 neighbours'
// extra coarse gathers exceed the registers (3 DOFs per node: 2 KB of scratch
// per lane in FP64; 1 DOF: 256 VGPRs, one wave per SIMD), so the two passes
// stay separate there.
template <int NDPN>
bool fuse_correction(const Grid& g) {
  return NDPN != 3 && st_dimu<NDPN>(g) == 2;
}

// y = x' + omega D^-1 (b - A x'), x' = x + P xc on the free DOFs: the
// prolongation from level l+1 and the first post-smoothing sweep of level l
// in one pass (fuse_correction levels only).
template <int NDPN, typename T>
int st_smooth_corrected(tl_problem* P, int l, const T* x, const T* xc, const T* b, T* y,
                        const int* done) {
  Level& L = P->mg.lv[l];
  TL_ARG(l >= 1 && l + 1 < P->mg.nlevels, "internal: stencil level out of range");
  const int dimu = st_dimu<NDPN>(L.g);
  TL_ARG(fuse_correction<NDPN>(L.g), "internal: fused correction on a level without it");
  TL_NONNULL(L.sth64.ptr); TL_NONNULL(L.st32.ptr); TL_NONNULL(L.dinv64.ptr);
  TL_NONNULL(L.dinv32.ptr); TL_NONNULL(L.mask.ptr); TL_NONNULL(x); TL_NONNULL(xc);
  TL_NONNULL(b); TL_NONNULL(y);
  const T* st = std::is_same<T, double>::value ? static_cast<const T*>(L.sth64.ptr)
                                                : static_cast<const T*>(L.st32.ptr);
  const T* dinv = std::is_same<T, double>::value ? static_cast<const T*>(L.dinv64.ptr)
                                                  : static_cast<const T*>(L.dinv32.ptr);
  const unsigned nb = unsigned((L.g.nnodes + BS - 1) / BS);
  ProfEvent ev{};
  ev.kind = -1;
  if (l == 1) TL_TRY(prof_begin(P, 4, ev));
  CoarseCorr<T> cc;
  cc.gc = to_d(P->mg.lv[l + 1].g);
  cc.xc = xc;
  cc.mask = L.mask.as<uint8_t>();
  const GridD gd = to_d(L.g);
  const T om = T(P->mg.omega);
  (void)dimu;  // 2: fuse_correction levels are 2D
  if constexpr (NDPN != 3) {
    if (P->rec) {
      P->rec->add(P->rec_phase, P->rec_level, &st_apply_multi_launch<NDPN, T, 1, 2, true>,
                  StArgs<T>{gd, st, x, b, dinv, om, cc.mask, y, done, cc}, nb, 0);
      return prof_end(P, ev);
    }
    hipLaunchKernelGGL((st_apply_kernel<NDPN, T, 1, 2, true>), dim3(nb), dim3(BS), 0, P->stream,
                       gd, st, x, b, dinv, om, cc.mask, y, done, cc);
  }
  TL_CHECK_LAUNCH();
  return prof_end(P, ev);
}

// Solve A_l x_l = b_l approximately (x_l := V-cycle(b_l) from the zero guess; the
// caller's restriction has already written the first sweep into x_l).
template <int NDPN, typename T>
int vcycle_level(tl_problem* P, int l, const int* done) {
  MgState& M = P->mg;
  Level& L = M.lv[l];
  const int64_t n = L.ndof;
  const unsigned nb = unsigned((n + BS - 1) / BS);
  TL_NONNULL(L.vb.ptr); TL_NONNULL(L.vx.ptr); TL_NONNULL(L.vt.ptr); TL_NONNULL(L.vr.ptr);
  TL_NONNULL(L.mask.ptr);
  P->rec_phase = 0;  // down the cycle at level l
  P->rec_level = l;
  if (l == M.nlevels - 1) {
    TL_NONNULL(M.cinv.ptr);
    TL_ARG(M.cn == n, "internal: coarsest size mismatch");
    if (P->rec) {
      P->rec->add(0, l, &gemv_multi_launch<T>,
                  GemvArgs<T>{M.cinv.as<double>(), M.cnp, n, vp<T>(L.vb), vp<T>(L.vx), done},
                  unsigned((n + 3) / 4), 0);
      RecOp& op = P->rec->ops.back();  // systems of one instance share the inverse
      op.share = M.cinv.ptr;
      op.pack = &gemv_pack<T>;
      op.pack_fn = &gemv_pack_launch<T>;
      return TL_OK;
    }
    hipLaunchKernelGGL(dense_gemv_kernel<T>, dim3(unsigned((n + 3) / 4)), dim3(BS), 0, P->stream,
                       M.cinv.as<double>(), M.cnp, n, vp<T>(L.vb), vp<T>(L.vx), done);
    TL_CHECK_LAUNCH();
    return TL_OK;
  }
  T* x = vp<T>(L.vx);
  T* t = vp<T>(L.vt);
  const T* b = vp<T>(L.vb);
  // gamma > 1 (W-cycle): the level's cycle repeats from the current x; each
  // repetition smooths, corrects from a fresh coarse cycle and smooths again
  const int gam = l >= M.gamma_from ? M.gamma : 1;
  for (int g = 0; g < gam; ++g) {
    // g == 0: the first sweep from the zero guess, x = omega D^-1 b, was
    // written by the restriction that produced b (restrict_to)
    if (g > 0) {
      TL_TRY((st_apply<NDPN, T>(P, l, 1, x, b, t, done)));
      std::swap(L.vx.ptr, L.vt.ptr);
      x = vp<T>(L.vx); t = vp<T>(L.vt);
    }
    for (int s = 1; s < M.nu1; ++s) {
      TL_TRY((st_apply<NDPN, T>(P, l, 1, x, b, t, done)));
      std::swap(L.vx.ptr, L.vt.ptr);
      x = vp<T>(L.vx); t = vp<T>(L.vt);
    }
    TL_TRY((st_apply<NDPN, T>(P, l, 0, x, b, vp<T>(L.vr), done)));
    Level& C = M.lv[l + 1];
    TL_TRY((restrict_to<NDPN, T>(P, l, L.g, vp<T>(L.vr), done)));
    TL_TRY((vcycle_level<NDPN, T>(P, l + 1, done)));
    P->rec_phase = 1;  // up the cycle at level l
    P->rec_level = l;
    // coarse correction and post-smoothing; on the fuse_correction levels the
    // correction is added inside the first sweep
    int s0 = 0;
    if (fuse_correction<NDPN>(L.g)) {
      TL_TRY((st_smooth_corrected<NDPN, T>(P, l, x, vp<T>(C.vx), b, t, done)));
      std::swap(L.vx.ptr, L.vt.ptr);
      x = vp<T>(L.vx); t = vp<T>(L.vt);
      s0 = 1;
    } else {
      TL_TRY((launch_prolong<NDPN, T>(P, L.g, C.g, vp<T>(C.vx), L.mask.as<uint8_t>(), x, done)));
    }
    for (int s = s0; s < M.nu2; ++s) {
      TL_TRY((st_apply<NDPN, T>(P, l, 1, x, b, t, done)));
      std::swap(L.vx.ptr, L.vt.ptr);
      x = vp<T>(L.vx); t = vp<T>(L.vt);
    }
  }
  (void)nb;
  return TL_OK;
}

// Fine level: r (fp64 PCG residual) -> z (fp64), rho = r.z into the PCG state.
template <int NDPN, typename T>
int vcycle_fine(tl_problem* P, const double* r, double* z, const int* done, bool x0_ready) {
  MgState& M = P->mg;
  TL_ARG(M.nlevels >= 2 && int(M.lv.size()) == M.nlevels, "internal: hierarchy not set up");
  Level& L = M.lv[0];
  Level& C = M.lv[1];
  const int64_t n = P->ndof;
  TL_NONNULL(L.vx.ptr); TL_NONNULL(L.vt.ptr); TL_NONNULL(L.vr.ptr); TL_NONNULL(C.vb.ptr);
  TL_NONNULL(C.vx.ptr); TL_NONNULL(C.mask.ptr); TL_NONNULL(P->op.mask.ptr);
  TL_NONNULL(P->op.dinv32.ptr); TL_NONNULL(P->op.dinv64.ptr); TL_NONNULL(r); TL_NONNULL(z);
  // fine-level operator: fp64, fp32, or MFMA with bf16/fp16 inputs (fp32 vectors)
  const int prec = std::is_same<T, double>::value
                       ? TL_PREC_FP64
                       : ((M.prec == TL_PREC_BF16 || M.prec == TL_PREC_FP16) ? M.prec : TL_PREC_FP32);
  const void* dinv = std::is_same<T, double>::value ? P->op.dinv64.ptr : P->op.dinv32.ptr;
  T* x = vp<T>(L.vx);
  T* t = vp<T>(L.vt);
  P->rec_phase = 0;
  P->rec_level = 0;
  if (!x0_ready) {
    if (P->rec) {  // no multi-instance form: the lockstep batch needs the fused first sweep
      P->rec->unsupported = true;
      return TL_OK;
    }
    hipLaunchKernelGGL((jacobi_zero_kernel<T, double>), dim3(gsz(n)), dim3(BS), 0, P->stream, r,
                       static_cast<const T*>(dinv), T(M.omega), x, n, done);
    TL_CHECK_LAUNCH();
  }
  for (int s = 1; s < M.nu1; ++s) {
    TL_TRY(fine_apply_ex(P, x, t, prec, KU_JACOBI, r, true, dinv, M.omega, nullptr, nullptr,
                         nullptr, nullptr, 0, done, false));
    std::swap(L.vx.ptr, L.vt.ptr);
    x = vp<T>(L.vx); t = vp<T>(L.vt);
  }
  TL_TRY(fine_apply_ex(P, x, vp<T>(L.vr), prec, KU_RESID, r, true, nullptr, 0.0, nullptr, nullptr,
                       nullptr, nullptr, 0, done, false));
  TL_TRY((restrict_to<NDPN, T>(P, 0, L.g, vp<T>(L.vr), done)));
  TL_TRY((vcycle_level<NDPN, T>(P, 1, done)));
  P->rec_phase = 1;
  P->rec_level = 0;
  // the fine grid's Dirichlet mask is the operator's (level 0 has no copy)
  TL_TRY((launch_prolong<NDPN, T>(P, L.g, C.g, vp<T>(C.vx), P->op.mask.as<uint8_t>(), x, done)));
  for (int s = 0; s + 1 < M.nu2; ++s) {
    TL_TRY(fine_apply_ex(P, x, t, prec, KU_JACOBI, r, true, dinv, M.omega, nullptr, nullptr,
                         nullptr, nullptr, 0, done, false));
    std::swap(L.vx.ptr, L.vt.ptr);
    x = vp<T>(L.vx); t = vp<T>(L.vt);
  }
  // last post-smoothing sweep writes z (fp64) and rho = r.z (FIN_RHO = 1)
  return fine_apply_ex(P, x, z, prec, KU_JACOBI_DOT, r, true, dinv, M.omega, nullptr, nullptr,
                       nullptr, nullptr, 1, done, true);
}

}  // namespace

int mg_vcycle(tl_problem* P, const double* r, double* z, const int* done, bool x0_ready) {
  const bool f64 = P->mg.prec == TL_PREC_FP64;
  TL_ARG(!x0_ready || (P->mg.nu1 == 1 && P->mg.nu2 == 1),
         "internal: a precomputed first sweep needs V(1,1) (the iterate buffer must not rotate)");
  switch (P->ndpn) {
    case 1: return f64 ? vcycle_fine<1, double>(P, r, z, done, x0_ready)
                       : vcycle_fine<1, float>(P, r, z, done, x0_ready);
    case 2: return f64 ? vcycle_fine<2, double>(P, r, z, done, x0_ready)
                       : vcycle_fine<2, float>(P, r, z, done, x0_ready);
    case 3: return f64 ? vcycle_fine<3, double>(P, r, z, done, x0_ready)
                       : vcycle_fine<3, float>(P, r, z, done, x0_ready);
  }
  set_error("unsupported DOFs per node");
  return TL_EINVAL;
}

}  // namespace tl

using namespace tl;

extern "C" int tl_setup_mg(tl_problem* P, const tl_mg_opts* o) {
  TL_ARG(P && o, "null argument");
  TL_ARG(P->op.have_density, "tl_set_density must be called first");
  TL_ARG(o->nlevels >= 2 && o->nlevels <= 12, "nlevels must be in [2, 12]");
  TL_ARG(o->nu1 >= 1 && o->nu2 >= 1, "nu1 and nu2 must be >= 1");
  TL_ARG(o->omega > 0.0, "omega must be positive");
  TL_ARG(o->precision == TL_PREC_FP64 || o->precision == TL_PREC_FP32 ||
             o->precision == TL_PREC_BF16 || o->precision == TL_PREC_FP16,
         "V-cycle precision must be FP64, FP32, BF16 or FP16");
  TL_ARG(o->precision == TL_PREC_FP64 || o->precision == TL_PREC_FP32 ||
             (P->g.dim == 3 && P->ndpn == 3),
         "BF16/FP16 (MFMA) V-cycles are implemented for 3D H8 elasticity only");
  const int L = o->nlevels;
  const int f = 1 << (L - 1);
  const Grid& g = P->g;
  TL_ARG(g.nelx % f == 0 && g.nely % f == 0 && (g.dim == 2 || g.nelz % f == 0),
         "element counts must be divisible by 2^(nlevels-1)");
  MgState& M = P->mg;
  const int ndpn = P->ndpn, nd = P->nd;
  const bool rebuild = M.nlevels != L || M.lv.empty();
  ++P->gen;  // captured iteration chunks refer to the previous hierarchy
  // (zero-initialised gamma fields mean a V-cycle)
  TL_ARG(o->gamma >= 0 && o->gamma <= 3 && o->gamma_from >= 0, "gamma must be 0..3, gamma_from >= 0");
  M.nu1 = o->nu1; M.nu2 = o->nu2; M.omega = o->omega; M.prec = o->precision;
  M.gamma = std::max(1, int(o->gamma)); M.gamma_from = std::max(1, int(o->gamma_from));
  hipStream_t s = P->stream;
  if (rebuild) {
    std::vector<std::vector<uint8_t>> masks;
    TL_TRY(level_masks(P, L, masks));
    for (auto& l : M.lv) l.release();
    M.lv.clear();
    M.lv.resize(L);
    M.nlevels = L;
    Grid gl = g;
    for (int l = 0; l < L; ++l) {
      Level& lv = M.lv[l];
      if (l > 0) gl.set(gl.dim, gl.nelx / 2, gl.nely / 2, gl.nelz / 2);
      lv.g = gl;
      lv.ndof = gl.nnodes * ndpn;
      const size_t vb = sizeof(double) * size_t(lv.ndof);
      TL_TRY(lv.vx.alloc(vb)); TL_TRY(lv.vt.alloc(vb)); TL_TRY(lv.vr.alloc(vb)); TL_TRY(lv.vb.alloc(vb));
      if (l == 0) continue;
      TL_TRY(lv.mask.alloc(gl.nnodes));
      TL_HIP(hipMemcpy(lv.mask.ptr, masks[l].data(), gl.nnodes, hipMemcpyHostToDevice));
      const size_t nst = size_t(nstencil(gl.dim)) * ndpn * ndpn * gl.nnodes;
      TL_TRY(lv.st64.alloc(sizeof(double) * nst));
      const size_t nsh = size_t(center_k(gl.dim) + 1) * ndpn * ndpn * gl.nnodes;
      TL_TRY(lv.sth64.alloc(sizeof(double) * nsh));
      TL_TRY(lv.st32.alloc(sizeof(float) * nsh));
      TL_TRY(lv.dinv64.alloc(sizeof(double) * lv.ndof));
      TL_TRY(lv.dinv32.alloc(sizeof(float) * lv.ndof));
    }
    // G1[c] = (W_c x I)^T KE (W_c x I)
    std::vector<double> W;
    child_weights(g.dim, W);
    const int nn = g.dim == 3 ? 8 : 4;
    std::vector<double> G(size_t(nn) * nd * nd, 0.0);
    std::vector<double> Q(size_t(nd) * nd), T(size_t(nd) * nd);
    for (int ch = 0; ch < nn; ++ch) {
      std::fill(Q.begin(), Q.end(), 0.0);
      for (int i = 0; i < nn; ++i)
        for (int j = 0; j < nn; ++j)
          for (int d = 0; d < ndpn; ++d)
            Q[size_t(i * ndpn + d) * nd + j * ndpn + d] = W[(size_t(ch) * nn + i) * nn + j];
      for (int a = 0; a < nd; ++a)  // T = KE Q
        for (int b = 0; b < nd; ++b) {
          double sum = 0.0;
          for (int k = 0; k < nd; ++k) sum += P->KEh[size_t(a) * nd + k] * Q[size_t(k) * nd + b];
          T[size_t(a) * nd + b] = sum;
        }
      for (int a = 0; a < nd; ++a)  // G = Q^T T
        for (int b = 0; b < nd; ++b) {
          double sum = 0.0;
          for (int k = 0; k < nd; ++k) sum += Q[size_t(k) * nd + a] * T[size_t(k) * nd + b];
          G[(size_t(ch) * nd + a) * nd + b] = sum;
        }
    }
    TL_TRY(M.G1.alloc(sizeof(double) * G.size()));
    TL_HIP(hipMemcpy(M.G1.ptr, G.data(), sizeof(double) * G.size(), hipMemcpyHostToDevice));
    const Level& Lc = M.lv[L - 1];
    M.cn = Lc.ndof;
    M.cnp = ((M.cn + GJB - 1) / GJB) * GJB;
    TL_TRY(M.cinv.alloc(sizeof(double) * size_t(M.cnp) * M.cnp));
  }
  // level 1 from the fine densities
  {
    Level& l1 = M.lv[1];
#define TL_SFF(D3, NDPN)                                                                      \
  hipLaunchKernelGGL((st_from_fine_kernel<D3, NDPN>), dim3(gsz(l1.g.nnodes), NDPN * NDPN),     \
                     dim3(BS), 0, s,                                                          \
                     to_d(l1.g), P->op.sp64.as<double>(), g.nelx, g.nely, M.G1.as<double>(),   \
                     l1.mask.as<uint8_t>(), l1.st64.as<double>())
    const bool d3 = g.dim == 3;
    switch (ndpn) {
      case 1: if (d3) TL_SFF(true, 1); else TL_SFF(false, 1); break;
      case 2: if (d3) TL_SFF(true, 2); else TL_SFF(false, 2); break;
      case 3: if (d3) TL_SFF(true, 3); else TL_SFF(false, 3); break;
      default: set_error("unsupported DOFs per node"); return TL_EINVAL;
    }
#undef TL_SFF
    TL_CHECK_LAUNCH();
  }
  for (int l = 2; l < L; ++l) {
    Level& lf = M.lv[l - 1];
    Level& lc = M.lv[l];
    const dim3 grid(gsz(lc.g.nnodes * ndpn * ndpn));
    if (g.dim == 3)
      hipLaunchKernelGGL(st_coarsen_kernel<true>, grid, dim3(BS), 0, s, to_d(lf.g), to_d(lc.g), ndpn,
                         lf.st64.as<double>(), lc.mask.as<uint8_t>(), lc.st64.as<double>());
    else
      hipLaunchKernelGGL(st_coarsen_kernel<false>, grid, dim3(BS), 0, s, to_d(lf.g), to_d(lc.g), ndpn,
                         lf.st64.as<double>(), lc.mask.as<uint8_t>(), lc.st64.as<double>());
    TL_CHECK_LAUNCH();
  }
  for (int l = 1; l < L; ++l) {
    Level& lv = M.lv[l];
    const int64_t tot = lv.g.nnodes * nstencil(g.dim) * ndpn * ndpn;
    hipLaunchKernelGGL(st_finish_kernel, dim3(gsz(tot)), dim3(BS), 0, s, to_d(lv.g), ndpn,
                       lv.st64.as<double>(), lv.sth64.as<double>(), lv.st32.as<float>(),
                       lv.mask.as<uint8_t>(), lv.dinv64.as<double>(), lv.dinv32.as<float>());
    TL_CHECK_LAUNCH();
  }
  // coarsest dense inverse
  {
    Level& lc = M.lv[L - 1];
    double* A = M.cinv.as<double>();
    TL_HIP(hipMemsetAsync(A, 0, sizeof(double) * size_t(M.cnp) * M.cnp, s));
    const int64_t tot = lc.g.nnodes * nstencil(g.dim) * ndpn * ndpn;
    hipLaunchKernelGGL(dense_from_stencil_kernel, dim3(gsz(tot)), dim3(BS), 0, s, to_d(lc.g), ndpn,
                       lc.st64.as<double>(), A, M.cnp);
    TL_CHECK_LAUNCH();
    if (M.cnp > M.cn) {
      hipLaunchKernelGGL(pad_identity_kernel, dim3(1), dim3(GJB), 0, s, A, M.cn, M.cnp);
      TL_CHECK_LAUNCH();
    }
    const int nbk = int(M.cnp / GJB);
    const unsigned nt = unsigned((M.cnp + 63) / 64);
    const size_t panel = size_t(M.cnp) * GJB;
    TL_TRY(M.gjw.alloc(sizeof(double) * (2 * panel + 2 * GJB * GJB)));
    double* w = M.gjw.as<double>();
    double* rowp[2] = {w, w + panel};
    double* pinv[2] = {w + 2 * panel, w + 2 * panel + GJB * GJB};
    hipLaunchKernelGGL(gj_init_kernel, dim3(nt), dim3(BS), 0, s, A, M.cnp, rowp[0], pinv[0]);
    for (int k = 0; k < nbk; ++k) {
      const int c = k & 1, n = c ^ 1;
      hipLaunchKernelGGL(gj_step_kernel, dim3(nt * (nt + 1) / 2), dim3(BS), 0, s, A, M.cnp, k, nbk,
                         rowp[c], pinv[c], rowp[n], pinv[n]);
    }
    if (nt > 1)
      hipLaunchKernelGGL(gj_mirror_kernel, dim3(nt * (nt - 1) / 2), dim3(BS), 0, s, A, M.cnp);
    TL_CHECK_LAUNCH();
  }
  M.ready = true;
  return TL_OK;
}
