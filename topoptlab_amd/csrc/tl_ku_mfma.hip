// V-cycle fine-grid operator of 3D H8 elasticity on MFMA (BF16 or FP16
// inputs, FP32 accumulation): the element contraction Y_e = s_e KE u_e of
// topoptlab's assembly (topology_optimization.py:419-431, fem.py:141-178),
// summed into the nodes without assembling anything.
//
// Formulation ("column walk").  A wavefront owns a chunk of 16 consecutive
// nodes along y (MFMA column m <-> node gy = y0-1+m; columns 0 and 15 are
// the chunk's halo, so 14 nodes are produced per chunk), a tile of TX node
// lines in x and a range of element layers in z.  For one element column
// (ex, ez) two v_mfma_f32_16x16x32 products
//     D0 = A0 * B0   (element (ex, gy, ez))
//     D1 = A1 * B1   (element (ex, gy-1, ez))
// give, for every column m, the contributions of the two elements that
// share node gy to the 4 node lines (ex+dx, ez+dz) of that node:
//   * rows of A/D: 4 node lines (dx, dz) x 3 components (+1 zero row),
//   * K of A/B (32): 4 node lines (dx', dz') x [2 nodes along y, 3
//     components] (6 entries, 2 zero); each lane's 8 B0 entries are the
//     values of one node line at nodes gy, gy+1, and B1 (nodes gy-1, gy) is
//     B0 of the neighbouring column, moved over by one DPP row shift;
//   * A0 / A1 are the KE rows of local node (dx, dy=0 / 1, dz).
// (BF16: s_e inside B, one accumulator; FP16: y = g0*D0 + g1*D1.)
// Each lane loads one node (12 bytes) and one element scale per step; the
// y neighbour comes from the adjacent lane.  The parts of a node plane
// produced by the 4 lane groups (dx, dz) are summed through LDS: dz = 0
// parts within the layer, dz = 1 parts one layer later.  A finished node
// plane goes to a coalesced node-major epilogue (residual / damped-Jacobi
// sweep) whose inputs were requested at the start of the layer.  No
// workgroup barrier is needed in the walk; load addresses are a scalar base
// per step plus a per-lane offset fixed for the task.
//
// Precision: BF16 has the FP32 exponent range, so B0 holds s_gy * u_e(gy)
// rounded per element (B1, the neighbouring column's B0, then carries
// s_gy-1).  FP16 columns are scaled by the reciprocal of the element's max
// |u| (and KE by 1/max|KE|) to stay inside the FP16 range; there the scales
// s_gy * max|u_e| are applied to the FP32 results of the two products.
#include "tl_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>

namespace tl {

namespace {

constexpr int BS = 256;           // 4 wavefronts, one task each
constexpr int WAVES = BS / 64;
constexpr int CH = 14;            // nodes produced per y chunk (16 MFMA columns)
constexpr int TX = 8;             // node lines per task in x
constexpr int NPL = TX * CH;      // nodes of one node plane of a task
// LDS per wave: parts [dx][c][NPL] of the current layer (dz = 0) and, by
// layer parity, of the previous one (dz = 1).  The part stride and the gap
// before the dz = 1 parts place the four lane groups of one write
// instruction on disjoint LDS banks.
constexpr int PART = 352;
constexpr int HI_OFF = 2 * PART + 32;
constexpr int WAVE_LDS = 2176;
static_assert(3 * NPL <= PART && HI_OFF + 4 * PART <= WAVE_LDS, "LDS layout");
constexpr int EPR = (NPL + 63) / 64;  // epilogue nodes per lane

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

template <typename IN>
struct Vec8;
template <>
struct Vec8<__bf16> {
  typedef bf16x8 type;
};
template <>
struct Vec8<_Float16> {
  typedef half8 type;
};

template <typename IN>
__device__ inline floatx4 mfma16x16x32(typename Vec8<IN>::type a, typename Vec8<IN>::type b,
                                       floatx4 c) {
  if constexpr (std::is_same<IN, __bf16>::value)
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// value of lane m-1 / m+1 within each row of 16 lanes (0 at the row ends)
__device__ inline float from_below(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                               0x111 /* row_shr:1 */, 0xf, 0xf,
                                                               true));
}
__device__ inline float from_above(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                               0x101 /* row_shl:1 */, 0xf, 0xf,
                                                               true));
}

// element of `base` at a 32-bit byte offset (uniform base + per-lane 32-bit
// offset: the scalar-base addressing form, no 64-bit per-lane arithmetic)
template <typename T>
__device__ inline T ldo(const T* base, unsigned byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}
struct F3 {
  float v[3];
};
template <typename T>
__device__ inline void sto(T* base, unsigned byte_off, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

struct ColArgs {
  const float* u;         // V-cycle iterate (fp32, full DOF vector)
  const float* sp;        // padded SIMP scale (fp32), (nelz+2) x (nelx+2) x (nely+2)
  const float* KEf;       // element matrix, fp32, 24x24 row-major
  float kinv, kmax;       // 1/max|KE| and max|KE|
  const uint8_t* mask;    // per node, bit c = DOF c fixed
  const double* b;        // right-hand side (fp64 PCG residual)
  const float* dinv;      // inverse diagonal (fp32)
  float omega;
  void* y;                // fp32 (residual, Jacobi) or fp64 (Jacobi + dot)
  double* partials;
  const int* done;
  int nelx, nely, nelz, nx, ny, nz;
  int zlen, ntx, nyc, ntask;
};

template <typename IN, int MODE>
__global__ __launch_bounds__(BS, 4) void
kucol_kernel(ColArgs a) {
  if (a.done && *a.done) return;
  __shared__ float lds[WAVES * WAVE_LDS];
  __shared__ double shred[WAVES + 2];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int dx = kq & 1, dz = kq >> 1;

  // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs
  // (each with its own L2); XCD x runs a contiguous range of tasks, so the
  // node lines and planes shared by neighbouring tasks hit its own L2.
  const unsigned nb = gridDim.x, hw = blockIdx.x;
  const unsigned xcd = hw & 7u, q = nb >> 3, r = nb & 7u;
  const unsigned logical = xcd * q + min(xcd, r) + (hw >> 3);
  // wave-uniform task index (scalar registers; lets the walk branch on SCC)
  const int task = __builtin_amdgcn_readfirstlane(int(logical) * WAVES + wave);

  double dot = 0.0;
  if (task < a.ntask) {
    const int cy = task % a.nyc;
    const int rest = task / a.nyc;
    const int tx = rest % a.ntx, cz = rest / a.ntx;
    const int x0 = tx * TX, y0 = cy * CH;
    const int z0 = cz * a.zlen, z1 = min(a.nz, z0 + a.zlen);
    const int nxl = min(TX, a.nx - x0);  // node lines of this task
    const int gy = y0 - 1 + m;
    float* lw = lds + wave * WAVE_LDS;

    // A operands: lane row m -> node line m>>2 (dx_o, dz_o), component m&3
    // (3: zero row); K slice kq -> node line (dx_i, dz_i) = (kq&1, kq>>1),
    // entries j = 3*dy_i + c_i (j >= 6: zero)
    typename Vec8<IN>::type A0, A1;
    {
      const int line = m >> 2, c = m & 3;
      const int dxo = line & 1, dzo = line >> 1;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        float v0 = 0.f, v1 = 0.f;
        if (c < 3 && j < 6) {
          const int col = 3 * loc_of(dx, j / 3, dz) + j % 3;
          v0 = a.KEf[(3 * loc_of(dxo, 0, dzo) + c) * 24 + col] * a.kinv;
          v1 = a.KEf[(3 * loc_of(dxo, 1, dzo) + c) * 24 + col] * a.kinv;
        }
        A0[j] = IN(v0);
        A1[j] = IN(v1);
      }
    }

    // Per-lane byte offsets, fixed for the task.  Every value read through a
    // clamped index below is only multiplied by the zero SIMP scale of an
    // element of the padded border.
    const int gyn = min(max(gy, 0), a.ny - 1);
    const int gye = min(gy, a.nely);  // >= -1; nely is the zero border
    const unsigned lane_u = (__umul24(__umul24(dz, a.nx) + dx, a.ny) + gyn) * 12u;
    const unsigned lane_s = unsigned(gye + 1) * 4u;
    // step inputs of element column (x0-1+i, layer L): node (ex+dx, gy, L+dz)
    // and the scale of element (ex, gy, L); scalar base + lane offset
    auto load_step = [&](int i, int L, F3& u3, float& sv) {
      const int ex = x0 - 1 + i;
      const int exc = min(max(ex, 0), a.nelx - 1), Lz = min(max(L, 0), a.nelz - 1);
      const unsigned su = unsigned(Lz * a.nx + exc) * unsigned(a.ny) * 12u;
      u3 = ldo(reinterpret_cast<const F3*>(reinterpret_cast<const char*>(a.u) + su), lane_u);
      const int exs = min(ex, a.nelx), Ls = min(L, a.nelz);
      const unsigned ss = unsigned((Ls + 1) * (a.nelx + 2) + exs + 1) * unsigned(a.nely + 2) * 4u;
      sv = ldo(reinterpret_cast<const float*>(reinterpret_cast<const char*>(a.sp) + ss), lane_s);
    };
    // epilogue nodes of this lane: node n = lane + 64k of the task's plane
    unsigned eo[EPR];
    bool eok[EPR];
#pragma unroll
    for (int k = 0; k < EPR; ++k) {
      const int n = lane + 64 * k, nn = min(n, NPL - 1);
      const int line = nn / CH, node = nn - line * CH;
      eok[k] = n < NPL && line < nxl && y0 + node < a.ny;
      eo[k] = __umul24(min(x0 + line, a.nx - 1), a.ny) + min(y0 + node, a.ny - 1);
    }
    const unsigned plane = unsigned(a.nx) * unsigned(a.ny);

    F3 cu[TX + 1];
    float cs[TX + 1];
#pragma unroll
    for (int i = 0; i <= TX; ++i) load_step(i, z0 - 1, cu[i], cs[i]);

    for (int ez = z0 - 1; ez < z1; ++ez) {
      const bool out_plane = ez >= z0;
      float* lo = lw;                                    // [dx][c][NPL], this layer
      float* hi_w = lw + HI_OFF + (ez & 1) * 2 * PART;    // dz = 1 parts of this layer
      const float* hi_r = lw + HI_OFF + ((ez + 1) & 1) * 2 * PART;  // of the previous layer
      // epilogue inputs of node plane ez (unconditional, clamped; requested
      // before the walk so the wait counts below are static)
      const unsigned zoff = unsigned(max(ez, 0)) * plane;
      double eb[EPR][3];
      F3 ed[EPR], exv[EPR];
      unsigned em[EPR];
#pragma unroll
      for (int k = 0; k < EPR; ++k) {
        const unsigned nd = zoff + eo[k];
#pragma unroll
        for (int c = 0; c < 3; ++c) eb[k][c] = ldo(a.b, nd * 24u + 8u * c);
        em[k] = ldo(a.mask, nd);
        if constexpr (MODE != KU_RESID) {
          ed[k] = ldo(reinterpret_cast<const F3*>(a.dinv), nd * 12u);
          exv[k] = ldo(reinterpret_cast<const F3*>(a.u), nd * 12u);
        }
      }
#pragma unroll
      for (int i = 0; i <= TX; ++i) {
        const F3 u3 = cu[i];
        const float s0 = cs[i];
        float up[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) up[c] = from_above(u3.v[c]);
        const float s1 = from_below(s0);
        typename Vec8<IN>::type B0;
        float g0, g1;
        if constexpr (std::is_same<IN, __bf16>::value) {
          // bf16: B0 = s_gy * u_e(gy) rounded per element (B1 = B0 of column
          // m-1 then carries s_gy-1); rounding the iterate once for all its
          // elements instead gives a ~1.5x larger operator error
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            B0[c] = IN(u3.v[c] * s0);
            B0[3 + c] = IN(up[c] * s0);
          }
          g0 = g1 = a.kmax;
        } else {
          // fp16: column scale 1/max|u_e| of element gy (4 node lines x 2 nodes)
          float am = fmaxf(fabsf(u3.v[0]), fmaxf(fabsf(u3.v[1]), fabsf(u3.v[2])));
          am = fmaxf(am, __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(
                                                       __builtin_bit_cast(int, am), 0x401F)));
          am = fmaxf(am, __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(
                                                       (lane ^ 32) << 2, __builtin_bit_cast(int, am))));
          const float e0 = fmaxf(am, from_above(am));
          const float inv = e0 > 0.f ? __builtin_amdgcn_rcpf(e0) : 0.f;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            B0[c] = IN(u3.v[c] * inv);
            B0[3 + c] = IN(up[c] * inv);
          }
          g0 = s0 * e0 * a.kmax;
          g1 = s1 * from_below(e0) * a.kmax;
        }
        B0[6] = B0[7] = IN(0.f);
        // B1 (element gy-1: nodes gy-1, gy) is B0 of column m-1
        typedef int int4v __attribute__((ext_vector_type(4)));
        int4v w = __builtin_bit_cast(int4v, B0);
#pragma unroll
        for (int j = 0; j < 3; ++j)
          w[j] = __builtin_amdgcn_update_dpp(0, w[j], 0x111 /* row_shr:1 */, 0xf, 0xf, true);
        const typename Vec8<IN>::type B1 = __builtin_bit_cast(typename Vec8<IN>::type, w);
        // next layer, same step: issued after the last use of this step's
        // inputs so the loads land in the same registers (no copies, no wait
        // at the layer boundary); past the last layer they load clamped dummies
        load_step(i, ez + 1, cu[i], cs[i]);
        const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
        float part[3];
        if constexpr (std::is_same<IN, __bf16>::value) {
          const floatx4 D = mfma16x16x32<IN>(A1, B1, mfma16x16x32<IN>(A0, B0, zero));
#pragma unroll
          for (int c = 0; c < 3; ++c) part[c] = g0 * D[c];
        } else {
          const floatx4 D0 = mfma16x16x32<IN>(A0, B0, zero);
          const floatx4 D1 = mfma16x16x32<IN>(A1, B1, zero);
#pragma unroll
          for (int c = 0; c < 3; ++c) part[c] = fmaf(g0, D0[c], g1 * D1[c]);
        }
        // lane (kq, m): part (dx, dz) of node (ex+dx, gy, ez+dz) -> line l = i-1+dx
        const int l = i - 1 + dx;
        if (m >= 1 && m <= CH && l >= 0 && l < nxl) {
          float* dst = (dz ? hi_w : lo) + dx * PART + l * CH + (m - 1);
#pragma unroll
          for (int c = 0; c < 3; ++c) dst[c * NPL] = part[c];
        }
      }
      if (out_plane) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int k = 0; k < EPR; ++k) {
          if (!eok[k]) continue;
          const int n = lane + 64 * k;
          const unsigned nd = zoff + eo[k];
          float acc[3];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const int o = c * NPL + n;
            // node plane ez: (dx 0 + dx 1) of layer ez + (dx 0 + dx 1) of layer ez-1
            acc[c] = (lo[o] + lo[PART + o]) + (hi_r[o] + hi_r[PART + o]);
          }
          if constexpr (MODE == KU_RESID) {
            F3 out;
#pragma unroll
            for (int c = 0; c < 3; ++c)
              out.v[c] = ((em[k] >> c) & 1u) ? 0.f : float(eb[k][c]) - acc[c];
            sto(reinterpret_cast<F3*>(a.y), nd * 12u, out);
          } else {
            float vj[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const float bb = float(eb[k][c]);
              vj[c] = ((em[k] >> c) & 1u) ? 0.f
                                          : exv[k].v[c] + a.omega * ed[k].v[c] * (bb - acc[c]);
              if constexpr (MODE == KU_JACOBI_DOT) dot += double(bb) * double(vj[c]);
            }
            if constexpr (MODE == KU_JACOBI_DOT) {
#pragma unroll
              for (int c = 0; c < 3; ++c) sto(static_cast<double*>(a.y), nd * 24u + 8u * c, double(vj[c]));
            } else {
              F3 out;
#pragma unroll
              for (int c = 0; c < 3; ++c) out.v[c] = vj[c];
              sto(reinterpret_cast<F3*>(a.y), nd * 12u, out);
            }
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
  }
  if constexpr (MODE == KU_JACOBI_DOT) {
    double vv[1] = {dot};
    block_sum<1>(vv, shred);
    store_partials<1>(vv, a.partials);
  }
}

// ----------------------------------------------------------------------------
// FP64 column walk on v_mfma_f64_16x16x4f64: the same walk as above with
// double operands.  Lane layouts of the f64 MFMA: lane l holds A[l%16][l/16]
// and B[l/16][l%16] of one K=4 slice; output element i of lane l is
// C[l/16 + 4i][l%16].  Hence rows are ordered row = line + 4*component
// (component 3: zero rows), so lane group kq receives node line kq's three
// components, and K slice s (0..5) is (node gy+j, component c) with
// s = 3j + c for input node line kq: each lane's B0 slices are its node
// line's values at nodes gy, gy+1 times the element scale, exactly the data
// of the BF16 lane.  B1 is B0 of column m-1 (two 32-bit DPP shifts per
// double).  Two accumulators (elements gy and gy-1), summed in FP64.
//
// Modes: KU_STORE (y = K x), KU_STORE_DOT with the fused PCG update (input p = z + beta*p_in,
// written to pout by the node's owner, q = K p, dot += p.q), KU_RESID
// (y = b - K x), KU_JACOBI / KU_JACOBI_DOT (y = x + omega dinv (b - K x),
// dot += b.y).  All vectors FP64.
// ----------------------------------------------------------------------------
typedef double doublex4 __attribute__((ext_vector_type(4)));

struct D3 {
  double v[3];
};

__device__ inline double dpp_f64(double v, int ctrl_shr) {
  typedef int int2v __attribute__((ext_vector_type(2)));
  int2v w = __builtin_bit_cast(int2v, v);
  if (ctrl_shr) {
    w[0] = __builtin_amdgcn_update_dpp(0, w[0], 0x111 /* row_shr:1 */, 0xf, 0xf, true);
    w[1] = __builtin_amdgcn_update_dpp(0, w[1], 0x111, 0xf, 0xf, true);
  } else {
    w[0] = __builtin_amdgcn_update_dpp(0, w[0], 0x101 /* row_shl:1 */, 0xf, 0xf, true);
    w[1] = __builtin_amdgcn_update_dpp(0, w[1], 0x101, 0xf, 0xf, true);
  }
  return __builtin_bit_cast(double, w);
}

struct Col64Args {
  const double* u;        // input: x (V-cycle modes) or z (fused PCG product)
  const double* pin;      // fused product: previous search direction
  double* pout;           // fused product: new search direction
  const double* beta;     // fused product: device scalar
  const double* sp;       // padded SIMP scale (fp64)
  const double* KE;       // element matrix, fp64, 24x24 row-major
  const uint8_t* mask;
  const double* b;        // right-hand side (residual, Jacobi)
  const double* dinv;     // inverse diagonal (Jacobi)
  double omega;
  double* y;
  double* partials;
  const int* done;
  int nelx, nely, nelz, nx, ny, nz;
  int zlen, ntx, nyc, ntask;
};

template <int TXN>
struct Col64Layout {
  static constexpr int NPL = TXN * CH;
  static constexpr int PART = 3 * NPL + 8;         // doubles
  static constexpr int HI_OFF = 2 * PART + 24;
  static constexpr int WAVE_LDS = HI_OFF + 4 * PART;
  static constexpr int EPR = (NPL + 63) / 64;
};

template <int MODE, int TXN>
__global__ __launch_bounds__(BS, 2) void
kucol64_kernel(Col64Args a) {
  if (a.done && *a.done) return;
  using LY = Col64Layout<TXN>;
  constexpr int NPL64 = LY::NPL, P64 = LY::PART, EPR64 = LY::EPR;
  constexpr bool FUSEP = MODE == KU_STORE_DOT;
  constexpr bool DOT = MODE == KU_STORE_DOT || MODE == KU_JACOBI_DOT;
  __shared__ double lds[WAVES * LY::WAVE_LDS];
  __shared__ double shred[WAVES + 2];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int dx = kq & 1, dz = kq >> 1;

  const unsigned nb = gridDim.x, hw = blockIdx.x;
  const unsigned xcd = hw & 7u, q = nb >> 3, r = nb & 7u;
  const unsigned logical = xcd * q + min(xcd, r) + (hw >> 3);
  const int task = __builtin_amdgcn_readfirstlane(int(logical) * WAVES + wave);

  double dot = 0.0;
  if (task < a.ntask) {
    const double beta = FUSEP ? *a.beta : 0.0;
    const int cy = task % a.nyc;
    const int rest = task / a.nyc;
    const int tx = rest % a.ntx, cz = rest / a.ntx;
    const int x0 = tx * TXN, y0 = cy * CH;
    const int z0 = cz * a.zlen, z1 = min(a.nz, z0 + a.zlen);
    const int nxl = min(TXN, a.nx - x0);
    const int gy = y0 - 1 + m;
    double* lw = lds + wave * LY::WAVE_LDS;

    // A operands: row m -> output node line m&3 (dx_o, dz_o), component m>>2
    // (3: zero row); K slice s of lane group kq -> input node line kq,
    // element-local y offset s/3, component s%3
    double A0[6], A1[6];
    {
      const int oline = m & 3, oc = m >> 2;
      const int dxo = oline & 1, dzo = oline >> 1;
#pragma unroll
      for (int s = 0; s < 6; ++s) {
        const int col = 3 * loc_of(dx, s / 3, dz) + s % 3;
        A0[s] = oc < 3 ? a.KE[(3 * loc_of(dxo, 0, dzo) + oc) * 24 + col] : 0.0;
        A1[s] = oc < 3 ? a.KE[(3 * loc_of(dxo, 1, dzo) + oc) * 24 + col] : 0.0;
      }
    }

    const int gyn = min(max(gy, 0), a.ny - 1);
    const int gye = min(gy, a.nely);
    const unsigned lane_u = (__umul24(__umul24(dz, a.nx) + dx, a.ny) + gyn) * 24u;
    const unsigned lane_s = unsigned(gye + 1) * 8u;
    auto load_step = [&](int i, int L, D3& u3, D3& p3, double& sv) {
      const int ex = x0 - 1 + i;
      const int exc = min(max(ex, 0), a.nelx - 1), Lz = min(max(L, 0), a.nelz - 1);
      const unsigned su = unsigned(Lz * a.nx + exc) * unsigned(a.ny) * 24u;
      u3 = ldo(reinterpret_cast<const D3*>(reinterpret_cast<const char*>(a.u) + su), lane_u);
      if constexpr (FUSEP)
        p3 = ldo(reinterpret_cast<const D3*>(reinterpret_cast<const char*>(a.pin) + su), lane_u);
      const int exs = min(ex, a.nelx), Ls = min(L, a.nelz);
      const unsigned ss = unsigned((Ls + 1) * (a.nelx + 2) + exs + 1) * unsigned(a.nely + 2) * 8u;
      sv = ldo(reinterpret_cast<const double*>(reinterpret_cast<const char*>(a.sp) + ss), lane_s);
    };
    unsigned eo[EPR64];
    bool eok[EPR64];
#pragma unroll
    for (int k = 0; k < EPR64; ++k) {
      const int n = lane + 64 * k, nn = min(n, NPL64 - 1);
      const int line = nn / CH, node = nn - line * CH;
      eok[k] = n < NPL64 && line < nxl && y0 + node < a.ny;
      eo[k] = __umul24(min(x0 + line, a.nx - 1), a.ny) + min(y0 + node, a.ny - 1);
    }
    const unsigned plane = unsigned(a.nx) * unsigned(a.ny);

    D3 cu[TXN + 1];
    D3 cp[FUSEP ? TXN + 1 : 1];
    double cs[TXN + 1];
#pragma unroll
    for (int i = 0; i <= TXN; ++i) load_step(i, z0 - 1, cu[i], cp[FUSEP ? i : 0], cs[i]);

    for (int ez = z0 - 1; ez < z1; ++ez) {
      const bool out_plane = ez >= z0;
      double* lo = lw;
      double* hi_w = lw + LY::HI_OFF + (ez & 1) * 2 * P64;
      const double* hi_r = lw + LY::HI_OFF + ((ez + 1) & 1) * 2 * P64;
      const unsigned zoff = unsigned(max(ez, 0)) * plane;
      D3 eb[EPR64], ed[EPR64], exv[EPR64], epv[EPR64];
      unsigned em[EPR64];
#pragma unroll
      for (int k = 0; k < EPR64; ++k) {
        const unsigned nd = zoff + eo[k];
        em[k] = ldo(a.mask, nd);
        if constexpr (MODE != KU_STORE_DOT && MODE != KU_STORE)
          eb[k] = ldo(reinterpret_cast<const D3*>(a.b), nd * 24u);
        if constexpr (MODE == KU_JACOBI || MODE == KU_JACOBI_DOT) {
          ed[k] = ldo(reinterpret_cast<const D3*>(a.dinv), nd * 24u);
          exv[k] = ldo(reinterpret_cast<const D3*>(a.u), nd * 24u);
        }
        if constexpr (FUSEP) {
          exv[k] = ldo(reinterpret_cast<const D3*>(a.u), nd * 24u);
          epv[k] = ldo(reinterpret_cast<const D3*>(a.pin), nd * 24u);
        }
      }
#pragma unroll
      for (int i = 0; i <= TXN; ++i) {
        double uv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) uv[c] = FUSEP ? fma(beta, cp[FUSEP ? i : 0].v[c], cu[i].v[c]) : cu[i].v[c];
        const double s0 = cs[i];
        double B0[6], B1[6];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          B0[c] = s0 * uv[c];
          B0[3 + c] = s0 * dpp_f64(uv[c], 0);  // node gy+1 from column m+1
        }
#pragma unroll
        for (int t = 0; t < 6; ++t) B1[t] = dpp_f64(B0[t], 1);  // element gy-1 = column m-1
        load_step(i, ez + 1, cu[i], cp[FUSEP ? i : 0], cs[i]);
        doublex4 D0 = {0.0, 0.0, 0.0, 0.0}, D1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < 6; ++t) {
          D0 = __builtin_amdgcn_mfma_f64_16x16x4f64(A0[t], B0[t], D0, 0, 0, 0);
          D1 = __builtin_amdgcn_mfma_f64_16x16x4f64(A1[t], B1[t], D1, 0, 0, 0);
        }
        const int l = i - 1 + dx;
        if (m >= 1 && m <= CH && l >= 0 && l < nxl) {
          double* dst = (dz ? hi_w : lo) + dx * P64 + l * CH + (m - 1);
#pragma unroll
          for (int c = 0; c < 3; ++c) dst[c * NPL64] = D0[c] + D1[c];
        }
      }
      if (out_plane) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int k = 0; k < EPR64; ++k) {
          if (!eok[k]) continue;
          const int n = lane + 64 * k;
          const unsigned nd = zoff + eo[k];
          double acc[3];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const int o = c * NPL64 + n;
            acc[c] = (lo[o] + lo[P64 + o]) + (hi_r[o] + hi_r[P64 + o]);
          }
          D3 out;
          if constexpr (MODE == KU_STORE_DOT) {
            D3 pv;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              pv.v[c] = fma(beta, epv[k].v[c], exv[k].v[c]);
              out.v[c] = ((em[k] >> c) & 1u) ? 0.0 : acc[c];
              dot += pv.v[c] * out.v[c];
            }
            sto(reinterpret_cast<D3*>(a.pout), nd * 24u, pv);
          } else if constexpr (MODE == KU_STORE) {
#pragma unroll
            for (int c = 0; c < 3; ++c) out.v[c] = ((em[k] >> c) & 1u) ? 0.0 : acc[c];
          } else if constexpr (MODE == KU_RESID) {
#pragma unroll
            for (int c = 0; c < 3; ++c) out.v[c] = ((em[k] >> c) & 1u) ? 0.0 : eb[k].v[c] - acc[c];
          } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              out.v[c] = ((em[k] >> c) & 1u)
                             ? 0.0
                             : exv[k].v[c] + a.omega * ed[k].v[c] * (eb[k].v[c] - acc[c]);
              if constexpr (MODE == KU_JACOBI_DOT) dot += eb[k].v[c] * out.v[c];
            }
          }
          sto(reinterpret_cast<D3*>(a.y), nd * 24u, out);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
  }
  if constexpr (DOT) {
    double vv[1] = {dot};
    block_sum<1>(vv, shred);
    store_partials<1>(vv, a.partials);
  }
}

int task_zlen(int ntx, int nyc, int nz) {
  // z range per task: about 4 tasks per SIMD (1024 SIMDs), at least 4 layers
  if (const char* e = std::getenv("TOPOPTLAB_AMD_MFMA_ZLEN")) {
    const int v = std::atoi(e);
    if (v >= 1) return std::min(v, nz);
  }
  const int64_t cols = int64_t(ntx) * nyc;
  const int zl = int((cols * nz + 4095) / 4096);
  return std::max(4, std::min(zl, nz));
}

}  // namespace

// Launch the MFMA operator for mode KU_RESID (y fp32), KU_JACOBI (y fp32) or
// KU_JACOBI_DOT (y fp64; one partial sum of b.y per workgroup in `partials`,
// *nblocks_out of them, to be folded by the caller).  b is the fp64 PCG
// residual, dinv the fp32 inverse diagonal, u the fp32 iterate (zero on fixed
// DOFs).
int fine_apply_mfma(tl_problem* P, const void* u, void* y, int prec, int mode, const void* b,
                    const void* dinv, double omega, double* partials, const int* done,
                    unsigned* nblocks_out) {
  const Grid& g = P->g;
  TL_ARG(g.dim == 3 && P->ndpn == 3, "the MFMA operator covers 3D H8 elasticity only");
  TL_ARG(prec == TL_PREC_BF16 || prec == TL_PREC_FP16, "MFMA operator needs BF16 or FP16");
  TL_ARG(P->op.KE32.ptr != nullptr && P->op.sp32.ptr != nullptr && P->op.mask.ptr != nullptr,
         "internal: operator buffers missing");
  TL_ARG(b != nullptr && (mode == KU_RESID || dinv != nullptr), "internal: MFMA operand missing");
  double kmax = 0.0;
  for (double v : P->KEh) kmax = std::max(kmax, std::fabs(v));
  TL_ARG(kmax > 0.0 && std::isfinite(kmax), "element matrix is zero or not finite");
  ColArgs a{};
  a.u = static_cast<const float*>(u);
  a.sp = P->op.sp32.as<float>();
  a.KEf = P->op.KE32.as<float>();
  a.kinv = float(1.0 / kmax);
  a.kmax = float(kmax);
  a.mask = P->op.mask.as<uint8_t>();
  a.b = static_cast<const double*>(b);
  a.dinv = static_cast<const float*>(dinv);
  a.omega = float(omega);
  a.y = y;
  a.partials = partials;
  a.done = done;
  a.nelx = g.nelx; a.nely = g.nely; a.nelz = g.nelz;
  a.nx = g.nx; a.ny = g.ny; a.nz = g.nz;
  a.ntx = (g.nx + TX - 1) / TX;
  a.nyc = (g.ny + CH - 1) / CH;
  a.zlen = task_zlen(a.ntx, a.nyc, g.nz);
  const int64_t ntask = int64_t(a.ntx) * a.nyc * ((g.nz + a.zlen - 1) / a.zlen);
  // 32-bit byte offsets and 24-bit index products in the kernel
  TL_ARG(ntask < (int64_t(1) << 30) && P->ndof * int64_t(sizeof(double)) < (int64_t(1) << 32) &&
             P->g.npad() < (int64_t(1) << 24) * 4,
         "grid too large for the MFMA operator (32-bit offsets)");
  a.ntask = int(ntask);
  const int64_t nblocks = (ntask + WAVES - 1) / WAVES;
  TL_ARG(nblocks <= kMaxRedBlocks || mode != KU_JACOBI_DOT,
         "MFMA operator grid exceeds the reduction slot count");
  if (nblocks_out) *nblocks_out = unsigned(nblocks);
  const dim3 grid{unsigned(nblocks), 1u, 1u}, block{unsigned(BS), 1u, 1u};
  const bool bf = prec == TL_PREC_BF16;
#define TL_KUM(IN, MODE) hipLaunchKernelGGL((kucol_kernel<IN, MODE>), grid, block, 0, P->stream, a)
  switch (mode) {
    case KU_RESID: if (bf) TL_KUM(__bf16, KU_RESID); else TL_KUM(_Float16, KU_RESID); break;
    case KU_JACOBI: if (bf) TL_KUM(__bf16, KU_JACOBI); else TL_KUM(_Float16, KU_JACOBI); break;
    case KU_JACOBI_DOT:
      if (bf) TL_KUM(__bf16, KU_JACOBI_DOT); else TL_KUM(_Float16, KU_JACOBI_DOT);
      break;
    default: set_error("MFMA operator: unsupported mode"); return TL_EINVAL;
  }
#undef TL_KUM
  TL_CHECK_LAUNCH();
  return TL_OK;
}

// FP64 MFMA operator (3D H8 elasticity): modes KU_STORE, KU_STORE_DOT with
// the fused update (pin, pout, beta given), KU_RESID, KU_JACOBI,
// KU_JACOBI_DOT.  DOT modes leave one partial per workgroup in `partials`
// (*nblocks_out of them) for the caller's fold.
int fine_apply_mfma64(tl_problem* P, const double* u, double* y, int mode, const double* b,
                      const double* dinv, double omega, const double* pin, double* pout,
                      const double* beta, double* partials, const int* done,
                      unsigned* nblocks_out) {
  const Grid& g = P->g;
  TL_ARG(g.dim == 3 && P->ndpn == 3, "the MFMA operator covers 3D H8 elasticity only");
  TL_ARG(P->op.KE64.ptr != nullptr && P->op.sp64.ptr != nullptr && P->op.mask.ptr != nullptr,
         "internal: operator buffers missing");
  TL_ARG(mode != KU_STORE_DOT || (pin && pout && beta), "internal: fused update operands missing");
  TL_ARG((mode != KU_RESID && mode != KU_JACOBI && mode != KU_JACOBI_DOT) || b,
         "internal: right-hand side missing");
  TL_ARG((mode != KU_JACOBI && mode != KU_JACOBI_DOT) || dinv, "internal: diagonal missing");
  constexpr int TXN = 8;
  Col64Args a{};
  a.u = u; a.pin = pin; a.pout = pout; a.beta = beta;
  a.sp = P->op.sp64.as<double>();
  a.KE = P->op.KE64.as<double>();
  a.mask = P->op.mask.as<uint8_t>();
  a.b = b; a.dinv = dinv; a.omega = omega; a.y = y;
  a.partials = partials;
  a.done = done;
  a.nelx = g.nelx; a.nely = g.nely; a.nelz = g.nelz;
  a.nx = g.nx; a.ny = g.ny; a.nz = g.nz;
  a.ntx = (g.nx + TXN - 1) / TXN;
  a.nyc = (g.ny + CH - 1) / CH;
  a.zlen = task_zlen(a.ntx, a.nyc, g.nz);
  const int64_t ntask = int64_t(a.ntx) * a.nyc * ((g.nz + a.zlen - 1) / a.zlen);
  TL_ARG(ntask < (int64_t(1) << 30) && P->ndof * int64_t(sizeof(double)) < (int64_t(1) << 32) &&
             P->g.npad() * int64_t(sizeof(double)) < (int64_t(1) << 32),
         "grid too large for the MFMA operator (32-bit offsets)");
  a.ntask = int(ntask);
  const int64_t nblocks = (ntask + WAVES - 1) / WAVES;
  const bool dotm = mode == KU_STORE_DOT || mode == KU_JACOBI_DOT;
  TL_ARG(nblocks <= kMaxRedBlocks || !dotm, "MFMA operator grid exceeds the reduction slot count");
  if (nblocks_out) *nblocks_out = unsigned(nblocks);
  const dim3 grid{unsigned(nblocks), 1u, 1u}, block{unsigned(BS), 1u, 1u};
#define TL_KUM64(MODE) hipLaunchKernelGGL((kucol64_kernel<MODE, TXN>), grid, block, 0, P->stream, a)
  switch (mode) {
    case KU_STORE: TL_KUM64(KU_STORE); break;
    case KU_STORE_DOT: TL_KUM64(KU_STORE_DOT); break;
    case KU_RESID: TL_KUM64(KU_RESID); break;
    case KU_JACOBI: TL_KUM64(KU_JACOBI); break;
    case KU_JACOBI_DOT: TL_KUM64(KU_JACOBI_DOT); break;
    default: set_error("MFMA operator: unsupported mode"); return TL_EINVAL;
  }
#undef TL_KUM64
  TL_CHECK_LAUNCH();
  return TL_OK;
}

}  // namespace tl
