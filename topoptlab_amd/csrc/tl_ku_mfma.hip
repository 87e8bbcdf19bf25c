// V-cycle fine-grid operator of 3D H8 elasticity on MFMA (BF16 or FP16
// inputs, FP32 accumulation): the per-element contraction
//   Y_e = KE (24x24) * u_e (24)
// of topoptlab's assembly (topology_optimization.py:419-431, fem.py:141-178)
// batched 16 elements at a time as the columns of
// v_mfma_f32_16x16x32_{bf16,f16} (KE padded to 32x32: two 16-row tiles, one
// 32-deep K step), then scaled by s_e and summed into the nodes.
//
// Precision: each element vector is normalised by its max |u| and KE by its
// max |entry| before the conversion, so the 1e9 SIMP contrast, the spread of
// the iterate's magnitude and the material's units stay out of the 8-/11-bit
// mantissas and the FP16 range; the FP32 factor s_e*max|u_e|*max|KE| is
// applied to the FP32 MFMA result.
//
// Tiling: a workgroup owns a TX x TY tile of nodes of every node plane in a z
// chunk and walks the element layers iz = z0-1 .. z1-1 (between node planes
// iz and iz+1).  Per layer the (TX+1)(TY+1) elements touching the tile are
// processed in 4 passes by element parity (ex%2, ey%2): elements of one
// parity share no node, so the scatter into the LDS node accumulators needs
// no atomics and the summation order is fixed.  Plane iz is complete after
// layer iz; its epilogue (residual / Jacobi sweep) runs from LDS.  The next
// node plane and element layer are prefetched into registers during the
// MFMA passes.  Workgroups are ordered XCD-aware like the FP32/FP64 kernel.
#include "tl_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

namespace tl {

namespace {

constexpr int BS = 256;
constexpr int TX = 8;    // owned nodes per tile in x
constexpr int TY = 49;   // owned nodes per tile in y
constexpr int WX = TX + 2, WY = TY + 2;          // node window (1-node halo)
constexpr int EX = TX + 1, EY = TY + 1;          // elements touching the tile
constexpr int NWIN = WX * WY;                    // window nodes per plane
constexpr int NOWN = TX * TY;                    // owned nodes per plane
constexpr int NELT = EX * EY;                    // elements per layer
constexpr int PLANE_VALS = NWIN * 3;
constexpr int PF_PLANE = (PLANE_VALS + BS - 1) / BS;  // prefetch slots per thread
constexpr int PF_DENS = (NELT + BS - 1) / BS;

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

struct MfmaArgs {
  const float* u;         // V-cycle iterate (fp32, full DOF vector)
  const float* sp;        // padded SIMP scale (fp32)
  const float* KEf;       // element matrix, fp32, 24x24 row-major
  float kinv, kmax;       // 1/max|KE| (applied before the conversion) and max|KE|
  const uint8_t* mask;    // per node
  const void* b;          // right-hand side (fp64 PCG residual)
  const float* dinv;      // inverse diagonal (fp32)
  float omega;
  void* y;                // output: fp32 (residual, Jacobi) or fp64 (Jacobi + dot)
  double* partials;
  const int* done;
  int nelx, nely, nelz, nx, ny, nz, plane, zchunk, ntx, nty;
};

// local node k of an H8 element: offsets from its (ix,iy,iz) corner node
__device__ constexpr int odx(int k) { return loc_dx(k); }
__device__ constexpr int ody(int k) { return loc_dy(k); }
__device__ constexpr int odz(int k) { return loc_dz(k); }

// element of parity class cls (0..3) with rank r -> local (lex, ley); false
// if r is past the class size
__device__ inline bool class_elem(int cls, int r, int& lex, int& ley) {
  const int px = cls & 1, py = cls >> 1;
  const int cx = (EX - px + 1) / 2, cy = (EY - py + 1) / 2;
  if (r >= cx * cy) return false;
  lex = px + 2 * (r / cy);
  ley = py + 2 * (r % cy);
  return true;
}

template <typename IN, int MODE>
__global__ __launch_bounds__(BS) void kum3d_kernel(MfmaArgs a) {
  if (a.done && *a.done) return;
  __shared__ float planes[2][PLANE_VALS];  // node planes by parity of iz
  __shared__ float acc[2][NOWN * 3];       // accumulators of owned nodes, by plane parity
  __shared__ float dens[NELT];             // SIMP scale of the current element layer
  __shared__ double shred[BS / 64 + 2];

  // XCD-aware order of (tile x, tile y, z chunk): each XCD gets a contiguous range
  const unsigned nb = gridDim.x;
  const unsigned hw = blockIdx.x;
  const unsigned xcd = hw & 7u, q = nb >> 3, r = nb & 7u;
  const unsigned logical = xcd * q + min(xcd, r) + (hw >> 3);
  const int ty = int(logical % a.nty);
  const int rest = int(logical / a.nty);
  const int tx = rest % a.ntx;
  const int tz = rest / a.ntx;
  const int x0 = tx * TX, y0 = ty * TY;
  const int z0 = tz * a.zchunk;
  const int z1 = min(a.nz, z0 + a.zchunk);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kq = lane >> 4, col = lane & 15;

  // A operands: KE rows (tile 0: 0-15, tile 1: 16-31), K slice 8*kq..8*kq+7, zero padded
  typename Vec8<IN>::type A0, A1;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = 8 * kq + j;
    const int r0 = col, r1 = 16 + col;
    A0[j] = IN(k < 24 ? a.KEf[r0 * 24 + k] * a.kinv : 0.f);
    A1[j] = IN((k < 24 && r1 < 24) ? a.KEf[r1 * 24 + k] * a.kinv : 0.f);
  }

  // staging helpers ---------------------------------------------------------
  auto plane_val = [&](int pz, int t) -> float {
    // window entry t = (wx * WY + wy) * 3 + c
    const int c = t % 3, wn = t / 3;
    const int wx = wn / WY, wy = wn % WY;
    const int gx = x0 - 1 + wx, gy = y0 - 1 + wy;
    if (pz < 0 || pz >= a.nz || gx < 0 || gx >= a.nx || gy < 0 || gy >= a.ny) return 0.f;
    return a.u[((int64_t(pz) * a.nx + gx) * a.ny + gy) * 3 + c];
  };
  auto dens_val = [&](int ez, int e) -> float {
    // element e = lex * EY + ley of layer ez; padded array covers ex,ey,ez in [-1, nel]
    const int lex = e / EY, ley = e % EY;
    const int ex = x0 - 1 + lex, ey = y0 - 1 + ley;
    if (ez < -1 || ez > a.nelz || ex > a.nelx || ey > a.nely) return 0.f;
    return a.sp[(int64_t(ez + 1) * (a.nelx + 2) + ex + 1) * (a.nely + 2) + ey + 1];
  };

  // prologue: planes z0-1, z0 and layer z0-1; clear accumulators
  for (int t = threadIdx.x; t < PLANE_VALS; t += BS) {
    planes[(z0 - 1) & 1][t] = plane_val(z0 - 1, t);
    planes[z0 & 1][t] = plane_val(z0, t);
  }
  for (int e = threadIdx.x; e < NELT; e += BS) dens[e] = dens_val(z0 - 1, e);
  for (int t = threadIdx.x; t < 2 * NOWN * 3; t += BS) (&acc[0][0])[t] = 0.f;

  double dot = 0.0;
  for (int iz = z0 - 1; iz < z1; ++iz) {
    __syncthreads();  // planes iz, iz+1, layer iz densities and cleared accs visible
    // prefetch plane iz+2 and layer iz+1 into registers
    float pfp[PF_PLANE], pfd[PF_DENS];
#pragma unroll
    for (int k = 0; k < PF_PLANE; ++k) {
      const int t = threadIdx.x + k * BS;
      pfp[k] = (t < PLANE_VALS) ? plane_val(iz + 2, t) : 0.f;
    }
#pragma unroll
    for (int k = 0; k < PF_DENS; ++k) {
      const int e = threadIdx.x + k * BS;
      pfd[k] = (e < NELT) ? dens_val(iz + 1, e) : 0.f;
    }
    const float* plo = planes[iz & 1];
    const float* phi = planes[(iz + 1) & 1];
    float* acc_lo = acc[iz & 1];        // node plane iz
    float* acc_hi = acc[(iz + 1) & 1];  // node plane iz+1
    const bool want_lo = iz >= z0;      // plane iz owned by this chunk
    const bool want_hi = iz + 1 < z1;   // plane iz+1 owned by this chunk
    // 4 parity passes of element groups (16 elements per MFMA column block)
#pragma unroll 1
    for (int cls = 0; cls < 4; ++cls) {
      const int px = cls & 1, py = cls >> 1;
      const int csize = ((EX - px + 1) / 2) * ((EY - py + 1) / 2);
      const int ngroups = (csize + 15) / 16;
#pragma unroll 1
      for (int g = wave; g < ngroups; g += BS / 64) {
        int lex = 0, ley = 0;
        const bool valid = class_elem(cls, g * 16 + col, lex, ley);
        // B operand: this lane's 8 entries of element column `col`
        float v[8];
        float m = 0.f;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int k = 8 * kq + j;
          float val = 0.f;
          if (valid && k < 24) {
            const int ln = k / 3, c = k % 3;
            const int wn = (lex + odx(ln)) * WY + ley + ody(ln);
            val = (odz(ln) ? phi : plo)[wn * 3 + c];
          }
          v[j] = val;
          m = fmaxf(m, fabsf(val));
        }
        m = fmaxf(m, __shfl_xor(m, 16, 64));
        m = fmaxf(m, __shfl_xor(m, 32, 64));
        const float inv = m > 0.f ? 1.f / m : 0.f;
        typename Vec8<IN>::type B;
#pragma unroll
        for (int j = 0; j < 8; ++j) B[j] = IN(v[j] * inv);
        floatx4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
        d0 = mfma16x16x32<IN>(A0, B, d0);
        d1 = mfma16x16x32<IN>(A1, B, d1);
        if (!valid) continue;
        const float scale = dens[lex * EY + ley] * m * a.kmax;
        // lane holds rows 4*kq+i (tile 0) and 16+4*kq+i (tile 1) of its column
#pragma unroll
        for (int tile = 0; tile < 2; ++tile) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int row = 16 * tile + 4 * kq + i;
            if (row >= 24) continue;
            const int ln = row / 3, c = row % 3;
            const int wx = lex + odx(ln), wy = ley + ody(ln);
            if (wx < 1 || wx > TX || wy < 1 || wy > TY) continue;  // not owned
            const bool hi = odz(ln) != 0;
            if (hi ? !want_hi : !want_lo) continue;
            float* dst = hi ? acc_hi : acc_lo;
            const int o = ((wx - 1) * TY + (wy - 1)) * 3 + c;
            dst[o] += (tile ? d1[i] : d0[i]) * scale;
          }
        }
      }
      __syncthreads();  // parity classes never share nodes: order the passes
    }
    // epilogue of node plane iz (complete after this layer)
    if (want_lo) {
      for (int t = threadIdx.x; t < NOWN * 3; t += BS) {
        const int c = t % 3, on = t / 3;
        const int owx = on / TY, owy = on % TY;
        const int gx = x0 + owx, gy = y0 + owy;
        if (gx >= a.nx || gy >= a.ny) continue;
        const int64_t node = (int64_t(iz) * a.nx + gx) * a.ny + gy;
        const int64_t g = node * 3 + c;
        const bool fixed = (a.mask[node] >> c) & 1;
        const float accv = acc_lo[t];
        const float bb = float(static_cast<const double*>(a.b)[g]);
        if constexpr (MODE == KU_RESID) {
          static_cast<float*>(a.y)[g] = fixed ? 0.f : bb - accv;
        } else {
          const float xc = plo[((owx + 1) * WY + owy + 1) * 3 + c];
          const float vj = fixed ? 0.f : xc + a.omega * a.dinv[g] * (bb - accv);
          if constexpr (MODE == KU_JACOBI_DOT) {
            static_cast<double*>(a.y)[g] = double(vj);
            dot += double(bb) * double(vj);
          } else {
            static_cast<float*>(a.y)[g] = vj;
          }
        }
      }
    }
    __syncthreads();  // epilogue read plane iz and acc_lo
    // recycle: plane iz's buffer receives plane iz+2, acc_lo becomes plane iz+2's
    float* pdst = planes[iz & 1];
#pragma unroll
    for (int k = 0; k < PF_PLANE; ++k) {
      const int t = threadIdx.x + k * BS;
      if (t < PLANE_VALS) pdst[t] = pfp[k];
    }
#pragma unroll
    for (int k = 0; k < PF_DENS; ++k) {
      const int e = threadIdx.x + k * BS;
      if (e < NELT) dens[e] = pfd[k];
    }
    for (int t = threadIdx.x; t < NOWN * 3; t += BS) acc_lo[t] = 0.f;
  }
  if constexpr (MODE == KU_JACOBI_DOT) {
    double vv[1] = {dot};
    block_sum<1>(vv, shred);
    store_partials<1>(vv, a.partials);
  }
}

}  // namespace

// Launch the MFMA operator for mode KU_RESID (y fp32), KU_JACOBI (y fp32) or
// KU_JACOBI_DOT (y fp64; one partial sum of b.y per workgroup in `partials`,
// *nblocks_out of them, to be folded by the caller).  b is the fp64 PCG
// residual, dinv the fp32 inverse diagonal.
int fine_apply_mfma(tl_problem* P, const void* u, void* y, int prec, int mode, const void* b,
                    const void* dinv, double omega, double* partials, const int* done,
                    unsigned* nblocks_out) {
  const Grid& g = P->g;
  TL_ARG(g.dim == 3 && P->ndpn == 3, "the MFMA operator covers 3D H8 elasticity only");
  TL_ARG(prec == TL_PREC_BF16 || prec == TL_PREC_FP16, "MFMA operator needs BF16 or FP16");
  TL_ARG(P->op.KE32.ptr != nullptr && P->op.sp32.ptr != nullptr && P->op.mask.ptr != nullptr,
         "internal: operator buffers missing");
  double kmax = 0.0;
  for (double v : P->KEh) kmax = std::max(kmax, std::fabs(v));
  TL_ARG(kmax > 0.0 && std::isfinite(kmax), "element matrix is zero or not finite");
  MfmaArgs a{};
  a.u = static_cast<const float*>(u);
  a.sp = P->op.sp32.as<float>();
  a.KEf = P->op.KE32.as<float>();
  a.kinv = float(1.0 / kmax);
  a.kmax = float(kmax);
  a.mask = P->op.mask.as<uint8_t>();
  a.b = b;
  a.dinv = static_cast<const float*>(dinv);
  a.omega = float(omega);
  a.y = y;
  a.partials = partials;
  a.done = done;
  a.nelx = g.nelx; a.nely = g.nely; a.nelz = g.nelz;
  a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.plane = g.nx * g.ny;
  a.ntx = (g.nx + TX - 1) / TX;
  a.nty = (g.ny + TY - 1) / TY;
  // z chunk: about 1024 workgroups (4 per CU), chunks of >= 2 node planes
  int zc = g.nz;
  while (zc > 2 && int64_t(a.ntx) * a.nty * ((g.nz + zc - 1) / zc) < 1024) zc = (zc + 1) / 2;
  a.zchunk = zc;
  const int64_t nblocks = int64_t(a.ntx) * a.nty * ((g.nz + zc - 1) / zc);
  TL_ARG(nblocks <= kMaxRedBlocks || mode != KU_JACOBI_DOT,
         "MFMA operator grid exceeds the reduction slot count");
  TL_ARG(nblocks < (int64_t(1) << 31), "grid too large");
  if (nblocks_out) *nblocks_out = unsigned(nblocks);
  const dim3 grid{unsigned(nblocks), 1u, 1u}, block{unsigned(BS), 1u, 1u};
  const bool bf = prec == TL_PREC_BF16;
#define TL_KUM(IN, MODE) hipLaunchKernelGGL((kum3d_kernel<IN, MODE>), grid, block, 0, P->stream, a)
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

}  // namespace tl
