// Problem handle, SIMP scaling, matrix-free K.u on structured Q4/H8 grids,
// diagonal of K_ff and the device-resident preconditioned CG of the C ABI.
//
// Matrix-free operator (replaces the COO->CSC assembly and sparse product of
// topoptlab/topology_optimization.py:419-451 and linear_solvers.py:114,130):
//   y_n = sum_{elements e containing node n} s_e * (KE u_e)_{local(n,e)}
// evaluated node-centrically, so every output DOF is written once (no
// atomics, deterministic).  A workgroup owns a contiguous run of BS nodes of
// one (x,y) node plane and walks a chunk of z-planes; each new z-plane (plus
// its +-(ny+1) halo) is staged once in LDS and every thread keeps the 3x3x3
// neighbourhood of its node in registers, rotating three plane slots so no
// register moves are needed.  KE indices are compile-time constants with
// wave-uniform addresses, so KE entries arrive as scalar operands.
#include "tl_internal.hpp"
#include "ke_pattern.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

namespace tl {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

int DevBuf::alloc(size_t nbytes) {
  if (ptr && bytes >= nbytes) return TL_OK;
  release();
  if (nbytes == 0) return TL_OK;
  TL_HIP(hipMalloc(&ptr, nbytes));
  bytes = nbytes;
  return TL_OK;
}

void Level::release() {
  mask.release(); st64.release(); sth64.release(); st32.release(); dinv64.release();
  dinv32.release(); em.release(); vb.release(); vx.release(); vr.release(); vt.release();
}

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

constexpr int BS = 256;  // threads per workgroup of the vector/stencil kernels

enum { FIN_STORE = 0, FIN_RHO = 1, FIN_PQ = 2, FIN_RR = 3, FIN_BB = 4, FIN_RR0 = 5 };

__device__ inline void finalize(int fin, PcgState* st, double tot, double* out) {
  switch (fin) {
    case FIN_STORE:
      *out = tot;
      break;
    case FIN_RHO:
      st->rho = tot;
      st->beta = (st->it == 0) ? 0.0 : tot / st->rho_prev;
      break;
    case FIN_PQ:
      st->pq = tot;
      st->alpha = st->rho / tot;
      break;
    case FIN_RR:
      st->rr = tot;
      st->rho_prev = st->rho;
      st->it += 1;
      st->done = (tot < st->tol2 || st->it >= st->maxiter) ? 1 : 0;
      break;
    case FIN_BB:
      st->bb = tot;
      break;
    case FIN_RR0:
      st->rr = tot;
      st->it = 0;
      st->done = (tot < st->tol2 || st->maxiter <= 0) ? 1 : 0;
      break;
  }
}

// ----------------------------------------------------------------------------
// SIMP scale into the zero-bordered element array
// ----------------------------------------------------------------------------
__global__ void simp_pad_kernel(const double* __restrict__ x, double* __restrict__ s64,
                                float* __restrict__ s32, int dim, int nelx, int nely,
                                int nelz, double eps, double penal) {
  const int64_t npad = int64_t(nelx + 2) * (nely + 2) * (dim == 3 ? nelz + 2 : 1);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < npad;
       i += int64_t(gridDim.x) * blockDim.x) {
    int64_t t = i;
    const int py = int(t % (nely + 2)); t /= (nely + 2);
    const int px = int(t % (nelx + 2)); t /= (nelx + 2);
    const int pz = int(t);
    const int ex = px - 1, ey = py - 1, ez = pz - 1;
    double v = 0.0;
    const bool in = ex >= 0 && ex < nelx && ey >= 0 && ey < nely &&
                    (dim == 2 || (ez >= 0 && ez < nelz));
    if (in) {
      const int64_t e = (dim == 3 ? (int64_t(ez) * nelx + ex) * nely + ey
                                  : int64_t(ex) * nely + ey);
      v = eps + (1.0 - eps) * pow(x[e], penal);
    }
    s64[i] = v;
    s32[i] = float(v);
  }
}

// ----------------------------------------------------------------------------
// fine operator kernels
// ----------------------------------------------------------------------------
struct KuArgs {
  const void* u;        // input (T); with FUSEP: z
  const void* pin;      // FUSEP: previous p (T)
  void* pout;           // FUSEP: new p (T)
  const double* beta;   // FUSEP: device scalar
  const void* sp;       // padded scale (T)
  const void* KE;       // element matrix (T), nd x nd row-major
  const uint8_t* mask;  // per node
  const void* b;        // RESID/JACOBI right-hand side (TB)
  const void* dinv;     // JACOBI (T)
  double omega;
  void* y;              // output (TO)
  double* partials;
  double* dot_out;
  PcgState* st;
  int fin;
  const int* done;
  int nelx, nely, nelz, nx, ny, nz, plane, ncol;  // ncol: 3D tile columns
  int ty;  // 2D workgroup shape the launch uses (0: node line), for the multi-instance form
  // group values of a pattern-matched element matrix (see ke_pattern.hpp)
  double kv64[8];
  float kv32[8];
};

// Compile-time (+-)(group+1) pattern of the unit-cell element matrix.
template <bool D3, int NDPN>
__host__ __device__ constexpr int ke_pat(int i) {
  if constexpr (D3 && NDPN == 3) return kPat_h8_elast(i);
  else if constexpr (D3 && NDPN == 1) return kPat_h8_poisson(i);
  else if constexpr (!D3 && NDPN == 2) return kPat_q4_elast(i);
  else if constexpr (!D3 && NDPN == 1) return kPat_q4_poisson(i);
  else return 0;
}

__device__ inline float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T>
__device__ inline T kv_at(const KuArgs& a, int g) {
  if constexpr (sizeof(T) == 8) return T(a.kv64[g]);
  else return T(a.kv32[g]);
}

template <typename T>
__device__ inline T ldval(const void* p, int64_t i) {
  return static_cast<const T*>(p)[i];
}

template <int NDPN, typename T, int MODE, typename TB, typename TO>
__device__ inline void ku_epilogue(const KuArgs& a, int64_t node, const T (&uc)[NDPN],
                                   const T (&acc)[NDPN], double& dot) {
  const uint8_t m = a.mask[node];
#pragma unroll
  for (int d = 0; d < NDPN; ++d) {
    const int64_t g = node * NDPN + d;
    const bool fixed = (m >> d) & 1;
    if constexpr (MODE == KU_STORE || MODE == KU_STORE_DOT) {
      const T v = fixed ? T(0) : acc[d];
      static_cast<TO*>(a.y)[g] = TO(v);
      if constexpr (MODE == KU_STORE_DOT) dot += double(uc[d]) * double(v);
    } else if constexpr (MODE == KU_RESID) {
      const T v = fixed ? T(0) : T(ldval<TB>(a.b, g)) - acc[d];
      static_cast<TO*>(a.y)[g] = TO(v);
    } else {
      const T bb = T(ldval<TB>(a.b, g));
      const T v = fixed ? T(0)
                        : uc[d] + T(a.omega) * ldval<T>(a.dinv, g) * (bb - acc[d]);
      static_cast<TO*>(a.y)[g] = TO(v);
      if constexpr (MODE == KU_JACOBI_DOT) dot += double(bb) * double(v);
    }
  }
}

template <typename T>
__device__ inline T stage_val(const KuArgs& a, int64_t g, bool fusep) {
  if (fusep) return ldval<T>(a.u, g) + T(*a.beta) * ldval<T>(a.pin, g);
  return ldval<T>(a.u, g);
}

// Contribution of the 4 (2D) / 8 (3D) elements around a node.
// U: neighbour values [3 slots][9][NDPN]; slot index of relative plane rz.
template <int NDPN, typename T, int SM, int SC, int SP, bool D3, bool FAST>
__device__ inline void node_sum(const KuArgs& a, const T (&U)[3][9][NDPN],
                                const T (&S)[3][4], int sdn, int sup, T (&acc)[NDPN]) {
  constexpr int ND = (D3 ? 8 : 4) * NDPN;
  // An opaque zero offset keeps the KE loads inside the z loop: hoisted out
  // of it, the 576 loop-invariant entries would be pinned in vector registers
  // and spill.  Inside the loop they are wave-uniform scalar loads.
  // KE is read through the constant address space: the buffer is never
  // written while a kernel runs, so wave-uniform reads become scalar loads
  // whose results feed the FMAs as SGPR operands.
  int opq;
  asm volatile("s_mov_b32 %0, 0" : "=s"(opq));
  const __attribute__((address_space(4))) T* KE =
      (const __attribute__((address_space(4))) T*)(static_cast<const T*>(a.KE)) + opq;
#pragma unroll
  for (int d = 0; d < NDPN; ++d) acc[d] = T(0);
#pragma unroll
  for (int oz = (D3 ? -1 : 0); oz <= 0; ++oz) {
#pragma unroll
    for (int ox = -1; ox <= 0; ++ox) {
#pragma unroll
      for (int oy = -1; oy <= 0; ++oy) {
        const T s = (oz < 0) ? S[sdn][(ox + 1) * 2 + (oy + 1)] : S[sup][(ox + 1) * 2 + (oy + 1)];
        const int la = loc_of(-ox, -oy, -oz);
        T t[NDPN];
#pragma unroll
        for (int d = 0; d < NDPN; ++d) t[d] = T(0);
#pragma unroll
        for (int lb = 0; lb < (D3 ? 8 : 4); ++lb) {
          const int rx = ox + loc_dx(lb), ry = oy + loc_dy(lb), rz = oz + loc_dz(lb);
          const int slot = (rz < 0) ? SM : (rz == 0 ? SC : SP);
          const int nb = (rx + 1) * 3 + (ry + 1);
#pragma unroll
          for (int d = 0; d < NDPN; ++d)
#pragma unroll
            for (int c = 0; c < NDPN; ++c) {
              const int e = (la * NDPN + d) * ND + lb * NDPN + c;
              // explicit fused multiply-adds: each term stays one FMA (the
              // compiler may otherwise share kv*U products between elements
              // as separate multiplies held in registers)
              if constexpr (FAST) {
                // group value resident in SGPRs; the sign folds into the FMA
                const int pg = ke_pat<D3, NDPN>(e);
                const T kv = kv_at<T>(a, (pg > 0 ? pg : -pg) - 1);
                t[d] = fma_t(pg > 0 ? kv : -kv, U[slot][nb][c], t[d]);
              } else {
                t[d] = fma_t(T(KE[e]), U[slot][nb][c], t[d]);
              }
            }
        }
#pragma unroll
        for (int d = 0; d < NDPN; ++d) acc[d] = fma_t(s, t[d], acc[d]);
      }
    }
  }
}

template <int NDPN, typename T>
__device__ inline void stage_plane(const KuArgs& a, T* lds, int pz, int64_t j0, bool fusep) {
  const int L = BS + 2 * a.ny + 2;
  const int64_t base = j0 - a.ny - 1;
  const bool zin = pz >= 0 && pz < a.nz;
  for (int t = threadIdx.x; t < L * NDPN; t += BS) {
    const int64_t jj = base + t / NDPN;
    T v = T(0);
    if (zin && jj >= 0 && jj < a.plane)
      v = stage_val<T>(a, (int64_t(pz) * a.plane + jj) * NDPN + (t % NDPN), fusep);
    lds[t] = v;
  }
}

template <int NDPN, typename T, int SLOT, int NS>
__device__ inline void read_nbrs(const KuArgs& a, const T* lds, T (&U)[NS][9][NDPN]) {
  const int c0 = threadIdx.x + a.ny + 1;
#pragma unroll
  for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int c = 0; c < NDPN; ++c)
        U[SLOT][(dx + 1) * 3 + (dy + 1)][c] = lds[(c0 + dx * a.ny + dy) * NDPN + c];
}

template <typename T>
__device__ inline void load_dens(const KuArgs& a, int ix, int iy, int ez, T (&S)[4]) {
  const T* sp = static_cast<const T*>(a.sp);
  const int64_t rowz = int64_t(ez + 1) * (a.nelx + 2);
#pragma unroll
  for (int ox = -1; ox <= 0; ++ox)
#pragma unroll
    for (int oy = -1; oy <= 0; ++oy)
      S[(ox + 1) * 2 + (oy + 1)] = sp[(rowz + ix + ox + 1) * (a.nely + 2) + iy + oy + 1];
}

// Plane staging of the 3D operator by LDS DMA (global_load_lds): the
// L*NDPN entries of node plane pz around tile j0 (nodes j0-ny-1 ..
// j0+BS+ny, L = BS+2ny+2) are copied from HBM straight into LDS, one dword
// per lane per copy, lane k of a copy writing dword k after the wave's base
// (the 12-byte form writes 16-byte lane slots, which would leave gaps in the
// dense plane).  No registers hold the plane in flight, so the copy of plane iz+2
// overlaps the FMAs of step iz in full (a register prefetch had to stop at
// 4 slots per thread to keep two waves per SIMD and staged the rest of each
// plane synchronously).  The compiler waits for the copies (vmcnt) before
// the next barrier.  Entries outside the array are clamped to its first or
// last unit: halo nodes outside the plane or the grid only meet elements of
// the zero-scaled padded border, so any finite value gives the same sums.
template <int NDPN, typename T>
struct PlaneDma {
  static constexpr int UB = 4;  // bytes per lane per copy
  // copies per plane and LDS bytes per plane buffer (padded to whole rounds)
  __host__ __device__ static int units(int ny) {
    return int((int64_t(BS + 2 * ny + 2) * NDPN * sizeof(T) + UB - 1) / UB);
  }
  __host__ __device__ static int rounds(int ny) { return (units(ny) + BS - 1) / BS; }
  __host__ __device__ static size_t buf_bytes(int ny) { return size_t(rounds(ny)) * BS * UB; }
};

template <int NDPN, typename T>
__device__ inline void dma_plane(const KuArgs& a, const void* src, T* lds, int pz, int64_t j0) {
  constexpr int UB = PlaneDma<NDPN, T>::UB;
  const int nu = PlaneDma<NDPN, T>::units(a.ny);
  const int64_t total = int64_t(a.plane) * a.nz * NDPN * int64_t(sizeof(T));
  const int64_t base = (int64_t(pz) * a.plane + j0 - a.ny - 1) * NDPN * int64_t(sizeof(T));
  const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const char* g = static_cast<const char*>(src);
  char* l8 = reinterpret_cast<char*>(lds) + wave * 64 * UB;
  if (base >= 0 && base + int64_t(nu) * UB <= total) {
    // the whole plane range lies inside the array (every tile but the
    // first and last of the grid): no clamping
    const char* gp = g + base + int64_t(threadIdx.x) * UB;
    for (int r = 0; r * BS < nu; ++r)
      if (r * BS + int(threadIdx.x) < nu)
        __builtin_amdgcn_global_load_lds(gp + r * BS * UB, l8 + r * BS * UB, 4, 0, 0);
    return;
  }
  for (int r = 0; r * BS < nu; ++r) {
    const int k = r * BS + int(threadIdx.x);
    if (k < nu) {
      int64_t o = base + int64_t(k) * UB;
      o = o < 0 ? int64_t(0) : (o > total - UB ? total - UB : o);
      __builtin_amdgcn_global_load_lds(g + o, l8 + r * BS * UB, 4, 0, 0);
    }
  }
}

// Wait until this wave's copies (and every other vector memory operation it
// issued) have completed: vmcnt(0) in the gfx9 s_waitcnt encoding, issued
// through the builtin so that the compiler's own wait tracking knows that no
// LDS copy is pending afterwards.
__device__ inline void wait_copies() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// LDS plane buffers of one workgroup, one per plane parity.  Addresses are
// formed arithmetically (a runtime-indexed pointer array would live in
// scratch).
template <typename T>
struct PlaneBufs {
  char* base;
  unsigned bb;  // bytes per buffer
  __device__ T* z(int parity) const { return reinterpret_cast<T*>(base + parity * bb); }
};

// Epilogue inputs of one node, fetched at the start of its step.
template <int NDPN, typename T, int MODE, typename TB>
struct EpiPrefetch {
  TB b[NDPN];
  T dinv[NDPN];
  uint8_t mask;
};

template <int NDPN, typename T, int MODE, typename TB>
__device__ inline void fetch_epi(const KuArgs& a, int64_t node, EpiPrefetch<NDPN, T, MODE, TB>& e) {
  e.mask = a.mask[node];
#pragma unroll
  for (int d = 0; d < NDPN; ++d) {
    if constexpr (MODE == KU_RESID || MODE == KU_JACOBI || MODE == KU_JACOBI_DOT)
      e.b[d] = ldval<TB>(a.b, node * NDPN + d);
    if constexpr (MODE == KU_JACOBI || MODE == KU_JACOBI_DOT)
      e.dinv[d] = ldval<T>(a.dinv, node * NDPN + d);
  }
}

template <int NDPN, typename T, int MODE, typename TB, typename TO>
__device__ inline void ku_epilogue_pf(const KuArgs& a, int64_t node,
                                      const EpiPrefetch<NDPN, T, MODE, TB>& e,
                                      const T (&uc)[NDPN], const T (&acc)[NDPN], double& dot) {
#pragma unroll
  for (int d = 0; d < NDPN; ++d) {
    const int64_t g = node * NDPN + d;
    const bool fixed = (e.mask >> d) & 1;
    if constexpr (MODE == KU_STORE || MODE == KU_STORE_DOT) {
      const T v = fixed ? T(0) : acc[d];
      static_cast<TO*>(a.y)[g] = TO(v);
      if constexpr (MODE == KU_STORE_DOT) dot += double(uc[d]) * double(v);
    } else if constexpr (MODE == KU_RESID) {
      const T v = fixed ? T(0) : T(e.b[d]) - acc[d];
      static_cast<TO*>(a.y)[g] = TO(v);
    } else {
      const T bb = T(e.b[d]);
      const T v = fixed ? T(0) : uc[d] + T(a.omega) * e.dinv[d] * (bb - acc[d]);
      static_cast<TO*>(a.y)[g] = TO(v);
      if constexpr (MODE == KU_JACOBI_DOT) dot += double(bb) * double(v);
    }
  }
}

// Contributions of the 4 elements of one element layer (between node planes
// iz and iz+1, neighbour values in U slots SA and SB) to the thread's node on
// plane iz (acc_top: the layer lies above it) and on plane iz+1 (acc_bot: the
// layer lies below it).  DO_TOP / DO_BOT select the halves (chunk ends).
template <int NDPN, typename T, int SA, int SB, bool FAST, bool DO_TOP, bool DO_BOT>
__device__ inline void layer_sum(const KuArgs& a, const T (&U)[2][9][NDPN], const T (&S)[4],
                                 T (&acc_top)[NDPN], T (&acc_bot)[NDPN]) {
  constexpr int ND = 8 * NDPN;
  int opq;
  asm volatile("s_mov_b32 %0, 0" : "=s"(opq));
  const __attribute__((address_space(4))) T* KE =
      (const __attribute__((address_space(4))) T*)(static_cast<const T*>(a.KE)) + opq;
#pragma unroll
  for (int d = 0; d < NDPN; ++d) acc_top[d] = acc_bot[d] = T(0);
#pragma unroll
  for (int ox = -1; ox <= 0; ++ox) {
#pragma unroll
    for (int oy = -1; oy <= 0; ++oy) {
      const T s = S[(ox + 1) * 2 + (oy + 1)];
      const int la_top = loc_of(-ox, -oy, 0), la_bot = loc_of(-ox, -oy, 1);
      T tt[NDPN], tb[NDPN];
#pragma unroll
      for (int d = 0; d < NDPN; ++d) tt[d] = tb[d] = T(0);
#pragma unroll
      for (int lb = 0; lb < 8; ++lb) {
        const int slot = loc_dz(lb) == 0 ? SA : SB;
        const int nb = (ox + loc_dx(lb) + 1) * 3 + (oy + loc_dy(lb) + 1);
#pragma unroll
        for (int d = 0; d < NDPN; ++d)
#pragma unroll
          for (int c = 0; c < NDPN; ++c) {
            const int et = (la_top * NDPN + d) * ND + lb * NDPN + c;
            const int eb = (la_bot * NDPN + d) * ND + lb * NDPN + c;
            const T uv = U[slot][nb][c];
            if constexpr (FAST) {
              const int pt = ke_pat<true, NDPN>(et), pb = ke_pat<true, NDPN>(eb);
              const T kt = kv_at<T>(a, (pt > 0 ? pt : -pt) - 1);
              const T kb = kv_at<T>(a, (pb > 0 ? pb : -pb) - 1);
              if constexpr (DO_TOP) tt[d] = fma_t(pt > 0 ? kt : -kt, uv, tt[d]);
              if constexpr (DO_BOT) tb[d] = fma_t(pb > 0 ? kb : -kb, uv, tb[d]);
            } else {
              if constexpr (DO_TOP) tt[d] = fma_t(T(KE[et]), uv, tt[d]);
              if constexpr (DO_BOT) tb[d] = fma_t(T(KE[eb]), uv, tb[d]);
            }
          }
      }
#pragma unroll
      for (int d = 0; d < NDPN; ++d) {
        if constexpr (DO_TOP) acc_top[d] = fma_t(s, tt[d], acc_top[d]);
        if constexpr (DO_BOT) acc_bot[d] = fma_t(s, tb[d], acc_bot[d]);
      }
    }
  }
}

// One step over element layer iz (node planes iz, iz+1).  On entry: U[SA]
// holds plane iz, the copy of plane iz+1 into its LDS buffer is in flight,
// S holds layer iz's densities and `carry` the layer-(iz-1) part of node iz.
// The step reads plane iz+1, starts the copy of plane iz+2 (when a later
// step needs it) into the buffer plane iz came from, prefetches layer iz+1's
// densities, completes and writes node iz and carries node iz+1's partial sum.
template <int NDPN, typename T, int MODE, typename TB, typename TO, bool FAST, int SA, int SB>
__device__ inline void z_step(const KuArgs& a, const PlaneBufs<T>& pb, int iz, int z1,
                              int64_t j0, int64_t j, int ix, int iy, bool active,
                              T (&U)[2][9][NDPN], T (&S)[4], T (&carry)[NDPN], double& dot) {
  const int pc = (iz + 1) & 1;  // buffer parity of plane iz+1 (plane iz+2: the other)
  __syncthreads();
  read_nbrs<NDPN, T, SB>(a, pb.z(pc), U);
  if (iz + 1 < z1) dma_plane<NDPN, T>(a, a.u, pb.z(pc ^ 1), iz + 2, j0);
  const int64_t node = int64_t(iz) * a.plane + (active ? j : 0);
  EpiPrefetch<NDPN, T, MODE, TB> ep;
  fetch_epi<NDPN, T, MODE, TB>(a, node, ep);
  // scales of element layer iz+1 for the next step; on the last node plane
  // (iz = nelz) there is no next step, and layer nelz+1 would lie past the
  // padded array: read the (zero) border layer nelz instead
  T Sn[4];
  load_dens<T>(a, ix, iy, min(iz + 1, a.nelz), Sn);
  T acc_top[NDPN], acc_bot[NDPN];
  layer_sum<NDPN, T, SA, SB, FAST, true, true>(a, U, S, acc_top, acc_bot);
  // Wait for the copy of plane iz+2 and the epilogue inputs here, after the
  // FMAs (the empty asm consumes the sums, so the FMAs cannot be sunk below
  // the wait) and before the stores: no copy is then pending at the next
  // barrier, and the stores drain across it.
#pragma unroll
  for (int d = 0; d < NDPN; ++d) asm volatile("" ::"v"(acc_top[d]), "v"(acc_bot[d]));
  wait_copies();
  if (active) {
    T uc[NDPN], tot[NDPN];
#pragma unroll
    for (int d = 0; d < NDPN; ++d) {
      uc[d] = U[SA][4][d];
      tot[d] = carry[d] + acc_top[d];
    }
    ku_epilogue_pf<NDPN, T, MODE, TB, TO>(a, node, ep, uc, tot, dot);
  }
#pragma unroll
  for (int d = 0; d < NDPN; ++d) carry[d] = acc_bot[d];
#pragma unroll
  for (int k = 0; k < 4; ++k) S[k] = Sn[k];
}

// Node planes z0..z1-1 of the column of tile bx (BS consecutive nodes of a
// plane, starting at node bx*BS).  Called with the LDS free (all threads
// past their last read of the previous segment, no copy in flight).
template <int NDPN, typename T, int MODE, typename TB, typename TO, bool FAST>
__device__ inline void ku3d_segment(const KuArgs& a, const PlaneBufs<T>& pb, unsigned bx, int z0,
                                    int z1, double& dot) {
  const int64_t j0 = int64_t(bx) * BS;
  const int64_t j = j0 + threadIdx.x;
  const bool inplane = j < a.plane;
  const int64_t jc = inplane ? j : a.plane - 1;
  // 32-bit division (planes hold fewer than 2^31 nodes, tl_create)
  const int ix = int(uint32_t(jc) / uint32_t(a.ny)), iy = int(jc) - ix * a.ny;
  T U[2][9][NDPN];
  T S[4];
  T carry[NDPN];
  // prologue: planes z0-1 -> U[0], z0 -> U[1]; layer z0-1 gives node z0's
  // "below" part; the copy of plane z0+1 for the first step; layer z0
  // densities
  const int pm = (z0 + 1) & 1;  // parity of z0-1 == parity of z0+1
  dma_plane<NDPN, T>(a, a.u, pb.z(pm), z0 - 1, j0);
  dma_plane<NDPN, T>(a, a.u, pb.z(pm ^ 1), z0, j0);
  load_dens<T>(a, ix, iy, z0 - 1, S);
  __syncthreads();
  read_nbrs<NDPN, T, 0>(a, pb.z(pm), U);
  read_nbrs<NDPN, T, 1>(a, pb.z(pm ^ 1), U);
  {
    T unused[NDPN];
    layer_sum<NDPN, T, 0, 1, FAST, false, true>(a, U, S, unused, carry);
  }
  __syncthreads();  // every wave has read plane z0-1 before its buffer is refilled
  dma_plane<NDPN, T>(a, a.u, pb.z(pm), z0 + 1, j0);
  load_dens<T>(a, ix, iy, z0, S);
  // no copy may be pending where the step loop starts: the compiler would
  // otherwise wait for every memory operation at the loop's barrier, the
  // previous step's stores included
  wait_copies();
  // alternating U slots: plane iz in slot SA, plane iz+1 read into SB
  for (int iz = z0; iz < z1; iz += 2) {
    z_step<NDPN, T, MODE, TB, TO, FAST, 1, 0>(a, pb, iz, z1, j0, j, ix, iy, inplane, U, S, carry, dot);
    if (iz + 1 >= z1) break;
    z_step<NDPN, T, MODE, TB, TO, FAST, 0, 1>(a, pb, iz + 1, z1, j0, j, ix, iy, inplane, U, S, carry, dot);
  }
}

// One workgroup per resident slot (the grid is sized by the launcher from
// the kernel's occupancy); the ncol*nz (tile column, node plane) units are
// split evenly over the grid in column-major order, so every workgroup
// walks about the same number of planes, in one or two column segments.
// A grid larger than the resident slots would leave a partial last round.
template <int NDPN, typename T, int MODE, typename TB, typename TO, bool FAST>
__global__ __launch_bounds__(BS) void ku3d_kernel(KuArgs a) {
  if (a.done && *a.done) return;
  extern __shared__ double lds_raw[];
  PlaneBufs<T> pb;
  pb.base = reinterpret_cast<char*>(lds_raw);
  pb.bb = unsigned(PlaneDma<NDPN, T>::buf_bytes(a.ny));
  // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs,
  // each with its own L2; remap so that XCD x runs a contiguous range of
  // units and the halos shared by neighbouring tiles are re-read from its
  // own L2 instead of HBM.
  const unsigned nb = gridDim.x;
  const unsigned hw = blockIdx.x;
  const unsigned xcd = hw & 7u, q = nb >> 3, r = nb & 7u;
  const unsigned w = xcd * q + min(xcd, r) + (hw >> 3);
  // unit range [u0, u1) of this workgroup (32-bit: the launcher checks that
  // units * gridDim.x fits); wave-uniform values kept in scalar registers
  const unsigned units = unsigned(a.ncol) * unsigned(a.nz);
  const unsigned u0 = __builtin_amdgcn_readfirstlane(units * w / nb);
  const unsigned u1 = __builtin_amdgcn_readfirstlane(units * (w + 1) / nb);
  double dot = 0.0;
  // the range spans at most two columns (the launcher keeps it <= nz units)
  const unsigned bx = __builtin_amdgcn_readfirstlane(u0 / unsigned(a.nz));
  const int z0 = int(u0 - bx * unsigned(a.nz));
  const int z1 = min(a.nz, z0 + int(u1 - u0));
  ku3d_segment<NDPN, T, MODE, TB, TO, FAST>(a, pb, bx, z0, z1, dot);
  const int rest = int(u1 - u0) - (z1 - z0);
  if (rest > 0) {
    __syncthreads();  // the next segment's prologue reuses both LDS buffers
    ku3d_segment<NDPN, T, MODE, TB, TO, FAST>(a, pb, bx + 1, 0, rest, dot);
  }
  if constexpr (MODE == KU_STORE_DOT || MODE == KU_JACOBI_DOT) {
    __shared__ double sh[BS / 64 + 2];
    double v[1] = {dot};
    block_sum<1>(v, sh);
    store_partials<1>(v, a.partials);
  }
}

// lds: (BS + 2 ny + 2) * NDPN values of T
template <int NDPN, typename T, int MODE, typename TB, typename TO, bool FUSEP, bool FAST>
__device__ __forceinline__ void ku2d_body(const KuArgs& a, unsigned bid, T* lds) {
  if (a.done && *a.done) return;
  const int64_t j0 = int64_t(bid) * BS;
  const int64_t j = j0 + threadIdx.x;
  const bool active = j < a.plane;
  const int64_t jc = active ? j : a.plane - 1;
  // 32-bit division (planes hold fewer than 2^31 nodes, tl_create)
  const int ix = int(uint32_t(jc) / uint32_t(a.ny)), iy = int(jc) - ix * a.ny;
  T U[3][9][NDPN];
  T S[3][4];
  stage_plane<NDPN, T>(a, lds, 0, j0, FUSEP);
  // 2D padded array: rows of (nely+2), one layer -> use ez = -1 addressing
  {
    const T* sp = static_cast<const T*>(a.sp);
#pragma unroll
    for (int ox = -1; ox <= 0; ++ox)
#pragma unroll
      for (int oy = -1; oy <= 0; ++oy)
        S[1][(ox + 1) * 2 + (oy + 1)] = sp[int64_t(ix + ox + 1) * (a.nely + 2) + iy + oy + 1];
  }
  __syncthreads();
  read_nbrs<NDPN, T, 1>(a, lds, U);
  T acc[NDPN];
  node_sum<NDPN, T, 0, 1, 2, false, FAST>(a, U, S, 0, 1, acc);
  double dot = 0.0;
  if (active) {
    T uc[NDPN];
#pragma unroll
    for (int d = 0; d < NDPN; ++d) uc[d] = U[1][4][d];
    if constexpr (FUSEP) {
#pragma unroll
      for (int d = 0; d < NDPN; ++d) static_cast<T*>(a.pout)[j * NDPN + d] = uc[d];
    }
    ku_epilogue<NDPN, T, MODE, TB, TO>(a, j, uc, acc, dot);
  }
  if constexpr (MODE == KU_STORE_DOT || MODE == KU_JACOBI_DOT) {
    __shared__ double sh[BS / 64 + 2];
    double v[1] = {dot};
    block_sum<1>(v, sh);
    store_partials_at<1>(v, a.partials, bid);
  }
}

template <int NDPN, typename T, int MODE, typename TB, typename TO, bool FUSEP, bool FAST>
__global__ __launch_bounds__(BS) void ku2d_kernel(KuArgs a) {
  extern __shared__ double lds_raw[];
  ku2d_body<NDPN, T, MODE, TB, TO, FUSEP, FAST>(a, blockIdx.x, reinterpret_cast<T*>(lds_raw));
}

static int launch_fold(tl_problem* P, unsigned nb, int fin, double* out, const int* done);

// Rectangular-tile form of the 2D operator: a workgroup owns a TX x TY node
// tile (TX = BS / TY, y fastest inside the tile) and stages its (TX+2) x
// (TY+2) node halo once.  The node-line kernel above stages BS + 2*ny + 2
// nodes for BS computed, a halo that grows with the grid height; the tile's
// does not, but its rows are shorter.  tl_autotune picks the faster shape
// per grid.
template <int TY>
constexpr int ku2t_halo_nodes() { return (BS / TY + 2) * (TY + 2); }

// lds: ku2t_halo_nodes<TY>() * NDPN values of T
template <int NDPN, typename T, int MODE, typename TB, typename TO, bool FUSEP, int TY>
__device__ __forceinline__ void ku2t_body(const KuArgs& a, unsigned bid, T* lds) {
  if (a.done && *a.done) return;
  constexpr int TX = BS / TY;
  constexpr int HY = TY + 2, HN = (TX + 2) * HY;
  static_assert(TX * TY == BS, "tile must hold BS nodes");
  const int ntx = (a.nx + TX - 1) / TX;
  const int x0 = int(bid % unsigned(ntx)) * TX, y0 = int(bid / unsigned(ntx)) * TY;
  const int lx = threadIdx.x / TY, ly = threadIdx.x % TY;
  const int ix = x0 + lx, iy = y0 + ly;
  const bool active = ix < a.nx && iy < a.ny;
  for (int t = threadIdx.x; t < HN * NDPN; t += BS) {
    const int h = t / NDPN, c = t % NDPN;
    const int hx = x0 - 1 + h / HY, hy = y0 - 1 + h % HY;
    T v = T(0);
    if (hx >= 0 && hx < a.nx && hy >= 0 && hy < a.ny)
      v = stage_val<T>(a, (int64_t(hx) * a.ny + hy) * NDPN + c, FUSEP);
    lds[t] = v;
  }
  T U[3][9][NDPN];
  T S[3][4];
  {
    // padded scale: (nelx+2) rows of (nely+2); inactive threads read node 0's
    const T* sp = static_cast<const T*>(a.sp);
    const int cx = active ? ix : 0, cy = active ? iy : 0;
#pragma unroll
    for (int ox = -1; ox <= 0; ++ox)
#pragma unroll
      for (int oy = -1; oy <= 0; ++oy)
        S[1][(ox + 1) * 2 + (oy + 1)] = sp[int64_t(cx + ox + 1) * (a.nely + 2) + cy + oy + 1];
  }
  __syncthreads();
#pragma unroll
  for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
      for (int c = 0; c < NDPN; ++c)
        U[1][(dx + 1) * 3 + (dy + 1)][c] = lds[((lx + 1 + dx) * HY + ly + 1 + dy) * NDPN + c];
  T acc[NDPN];
  node_sum<NDPN, T, 0, 1, 2, false, true>(a, U, S, 0, 1, acc);
  double dot = 0.0;
  if (active) {
    const int64_t j = int64_t(ix) * a.ny + iy;
    T uc[NDPN];
#pragma unroll
    for (int d = 0; d < NDPN; ++d) uc[d] = U[1][4][d];
    if constexpr (FUSEP) {
#pragma unroll
      for (int d = 0; d < NDPN; ++d) static_cast<T*>(a.pout)[j * NDPN + d] = uc[d];
    }
    ku_epilogue<NDPN, T, MODE, TB, TO>(a, j, uc, acc, dot);
  }
  if constexpr (MODE == KU_STORE_DOT || MODE == KU_JACOBI_DOT) {
    __shared__ double sh[BS / 64 + 2];
    double v[1] = {dot};
    block_sum<1>(v, sh);
    store_partials_at<1>(v, a.partials, bid);
  }
}

template <int NDPN, typename T, int MODE, typename TB, typename TO, bool FUSEP, int TY>
__global__ __launch_bounds__(BS) void ku2t_kernel(KuArgs a) {
  __shared__ T lds[ku2t_halo_nodes<TY>() * NDPN];
  ku2t_body<NDPN, T, MODE, TB, TO, FUSEP, TY>(a, blockIdx.x, lds);
}

// Multi-instance form of the 2D operator for lockstep batched solves (see
// tl::Recorder): block b runs local block tiles[b].y of system tiles[b].x in
// that system's workgroup shape (node line or 16/32/64-high tile), so one
// launch serves systems whatever shape their autotuning chose.  All shapes
// stage into one dynamic LDS buffer, sized for the largest among the systems.
template <int NDPN, typename T, int MODE, typename TB, typename TO, bool FUSEP, bool FAST>
__global__ __launch_bounds__(BS) void ku2_multi_kernel(const KuArgs* __restrict__ descs,
                                                       const uint2* __restrict__ tiles) {
  extern __shared__ double lds_raw[];
  T* lds = reinterpret_cast<T*>(lds_raw);
  const uint2 t = tiles[blockIdx.x];
  const KuArgs& a = descs[t.x];
  if constexpr (NDPN <= 2 && FAST) {
    switch (a.ty) {
      case 16: ku2t_body<NDPN, T, MODE, TB, TO, FUSEP, 16>(a, t.y, lds); return;
      case 32: ku2t_body<NDPN, T, MODE, TB, TO, FUSEP, 32>(a, t.y, lds); return;
      case 64: ku2t_body<NDPN, T, MODE, TB, TO, FUSEP, 64>(a, t.y, lds); return;
      default: break;
    }
  }
  ku2d_body<NDPN, T, MODE, TB, TO, FUSEP, FAST>(a, t.y, lds);
}

template <int NDPN, typename T, int MODE, typename TB, typename TO, bool FUSEP, bool FAST>
static int ku2_multi_launch(hipStream_t s, const void* descs, const uint2* tiles, unsigned nb,
                            size_t lds) {
  hipLaunchKernelGGL((ku2_multi_kernel<NDPN, T, MODE, TB, TO, FUSEP, FAST>), dim3(nb), dim3(BS),
                     lds, s, static_cast<const KuArgs*>(descs), tiles);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

// Workgroups of BS threads that the current device holds at once for
// `kernel` with `lds` bytes of dynamic LDS: CUs x resident workgroups per CU
// (cached per device, kernel and LDS size).
static int resident_slots(const void* kernel, size_t lds, int64_t& slots) {
  static std::mutex mu;
  static std::map<std::tuple<int, const void*, size_t>, int64_t> cache;
  int dev = 0;
  TL_HIP(hipGetDevice(&dev));
  const auto key = std::make_tuple(dev, kernel, lds);
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      slots = it->second;
      return TL_OK;
    }
  }
  int ncu = 0, per = 0;
  TL_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  TL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, BS, lds));
  slots = int64_t(std::max(ncu, 1)) * std::max(per, 1);
  std::lock_guard<std::mutex> lk(mu);
  cache[key] = slots;
  return TL_OK;
}

template <int NDPN, typename T, int MODE, typename TB, typename TO, bool FUSEP, int TY>
static int launch_ku2t(const tl_problem* P, const KuArgs& a) {
  constexpr int TX = BS / TY;
  const int64_t nb = int64_t((a.nx + TX - 1) / TX) * ((a.ny + TY - 1) / TY);
  if (nb > kMaxRedBlocks && (MODE == KU_STORE_DOT || MODE == KU_JACOBI_DOT)) {
    set_error("fine operator grid exceeds the reduction slot count");
    return TL_EINVAL;
  }
  if (P->rec) {
    KuArgs ar = a;
    ar.ty = TY;
    P->rec->add(P->rec_phase, P->rec_level, &ku2_multi_launch<NDPN, T, MODE, TB, TO, FUSEP, true>,
                ar, unsigned(nb), size_t(ku2t_halo_nodes<TY>()) * NDPN * sizeof(T));
  } else {
    hipLaunchKernelGGL((ku2t_kernel<NDPN, T, MODE, TB, TO, FUSEP, TY>), dim3(unsigned(nb)),
                       dim3(BS), 0, P->stream, a);
    TL_CHECK_LAUNCH();
  }
  if constexpr (MODE == KU_STORE_DOT || MODE == KU_JACOBI_DOT)
    return launch_fold(const_cast<tl_problem*>(P), unsigned(nb), a.fin, a.dot_out, a.done);
  return TL_OK;
}

template <int NDPN, typename T, int MODE, typename TB, typename TO, bool FUSEP, bool FAST>
static int launch_ku_impl(const tl_problem* P, KuArgs a) {

  const Grid& g = P->g;
  a.nelx = g.nelx; a.nely = g.nely; a.nelz = g.nelz;
  a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.plane = g.nx * g.ny;
  const int L = BS + 2 * g.ny + 2;
  const unsigned nbx = unsigned((int64_t(a.plane) + BS - 1) / BS);
  if (g.dim == 3 && P->rec) {  // no multi-instance form of the 3D operator
    P->rec->unsupported = true;
    return TL_OK;
  }
  if (g.dim == 3) {
    // p = z + beta p is a separate pass before the 3D product (the fused
    // form needed a second copy stream and an in-LDS combine and was
    // slower: 96 vs 64 + 19 us per product at 192x96x96)
    if constexpr (FUSEP) {
      set_error("internal: the 3D operator has no fused search-direction update");
      return TL_EINVAL;
    } else {
    // one workgroup per resident slot, at least 4 planes each
    const size_t lds = 2 * PlaneDma<NDPN, T>::buf_bytes(g.ny);
    TL_ARG(lds <= size_t(160) * 1024, "grid too wide in y for the 3D operator's LDS planes");
    const int64_t units = int64_t(nbx) * g.nz;
    int64_t slots = 0;
    TL_TRY(resident_slots(reinterpret_cast<const void*>(ku3d_kernel<NDPN, T, MODE, TB, TO, FAST>),
                          lds, slots));
    // at least one workgroup per column: a range of <= nz units spans at
    // most two columns, which is what the kernel walks
    const int64_t nwg =
        std::max<int64_t>(nbx, std::min<int64_t>(slots, (units + 3) / 4));
    // the kernel's unit arithmetic is 32-bit
    TL_ARG(units * (nwg + 1) < (int64_t(1) << 32), "grid too large for the 3D operator");
    if (nwg > kMaxRedBlocks && (MODE == KU_STORE_DOT || MODE == KU_JACOBI_DOT)) {
      set_error("fine operator grid exceeds the reduction slot count");
      return TL_EINVAL;
    }
    a.ncol = int(nbx);
    const dim3 grid(static_cast<unsigned>(nwg));
    hipLaunchKernelGGL((ku3d_kernel<NDPN, T, MODE, TB, TO, FAST>), grid, dim3(BS), lds,
                       P->stream, a);
    if constexpr (MODE == KU_STORE_DOT || MODE == KU_JACOBI_DOT) {
      TL_CHECK_LAUNCH();
      return launch_fold(const_cast<tl_problem*>(P), grid.x, a.fin, a.dot_out, a.done);
    }
    }
  } else {
    a.ncol = 0;
    if constexpr (NDPN <= 2 && FAST) {
      switch (P->op.tile2d[sizeof(T) == 8 ? 0 : 1]) {
        case 16: return launch_ku2t<NDPN, T, MODE, TB, TO, FUSEP, 16>(P, a);
        case 32: return launch_ku2t<NDPN, T, MODE, TB, TO, FUSEP, 32>(P, a);
        case 64: return launch_ku2t<NDPN, T, MODE, TB, TO, FUSEP, 64>(P, a);
        default: break;
      }
    }
    if (nbx > unsigned(kMaxRedBlocks) && (MODE == KU_STORE_DOT || MODE == KU_JACOBI_DOT)) {
      set_error("fine operator grid exceeds the reduction slot count");
      return TL_EINVAL;
    }
    if (P->rec) {
      KuArgs ar = a;
      ar.ty = 0;
      P->rec->add(P->rec_phase, P->rec_level,
                  &ku2_multi_launch<NDPN, T, MODE, TB, TO, FUSEP, FAST>, ar, nbx,
                  size_t(L) * NDPN * sizeof(T));
    } else {
      hipLaunchKernelGGL((ku2d_kernel<NDPN, T, MODE, TB, TO, FUSEP, FAST>), dim3(nbx), dim3(BS),
                         size_t(L) * NDPN * sizeof(T), P->stream, a);
    }
    if constexpr (MODE == KU_STORE_DOT || MODE == KU_JACOBI_DOT) {
      TL_CHECK_LAUNCH();
      return launch_fold(const_cast<tl_problem*>(P), nbx, a.fin, a.dot_out, a.done);
    }
  }
  TL_CHECK_LAUNCH();
  return TL_OK;
}

template <int NDPN, typename T, int MODE, typename TB, typename TO, bool FUSEP>
static int launch_ku(const tl_problem* P, KuArgs a) {
  if (P->op.fast_ke) return launch_ku_impl<NDPN, T, MODE, TB, TO, FUSEP, true>(P, a);
  return launch_ku_impl<NDPN, T, MODE, TB, TO, FUSEP, false>(P, a);
}

template <int NDPN, typename T>
static int dispatch_mode(const tl_problem* P, const KuArgs& a, int mode, bool fusep, bool b64,
                         bool out64) {
  // (mode, fusep, b type, out type) combinations used by the solver
  if (mode == KU_STORE) {
    if (out64) return launch_ku<NDPN, T, KU_STORE, double, double, false>(P, a);
    return launch_ku<NDPN, T, KU_STORE, T, T, false>(P, a);
  }
  if (mode == KU_STORE_DOT) {
    if (fusep) return launch_ku<NDPN, T, KU_STORE_DOT, T, T, true>(P, a);
    return launch_ku<NDPN, T, KU_STORE_DOT, T, T, false>(P, a);
  }
  if (mode == KU_RESID) {
    if (b64) return launch_ku<NDPN, T, KU_RESID, double, T, false>(P, a);
    return launch_ku<NDPN, T, KU_RESID, T, T, false>(P, a);
  }
  if (mode == KU_JACOBI) {
    if (b64) return launch_ku<NDPN, T, KU_JACOBI, double, T, false>(P, a);
    return launch_ku<NDPN, T, KU_JACOBI, T, T, false>(P, a);
  }
  if (mode == KU_JACOBI_DOT) {
    if (b64) return launch_ku<NDPN, T, KU_JACOBI_DOT, double, double, false>(P, a);
  }
  set_error("unsupported operator mode");
  return TL_EINVAL;
}

static int fine_apply_dispatch(tl_problem* P, KuArgs& a, int prec, int mode, bool fusep, bool b64,
                               bool out64);

// Generic entry used by the solver and the V-cycle.  prec selects T; b is
// fp64 when b64; the output type is T except for KU_JACOBI_DOT (fp64) and
// KU_STORE with prec fp32 and out64 (fp64 output).
int fine_apply_ex(tl_problem* P, const void* u, void* y, int prec, int mode, const void* b,
                  bool b64, const void* dinv, double omega, const void* zfuse_pin,
                  void* pout, const double* beta, double* dot_out, int fin, const int* done,
                  bool out64) {
  KuArgs a{};
  a.u = u; a.y = y; a.b = b; a.dinv = dinv; a.omega = omega;
  a.pin = zfuse_pin; a.pout = pout; a.beta = beta;
  a.mask = P->op.mask.as<uint8_t>();
  a.partials = P->partials.as<double>();
  a.dot_out = dot_out;
  a.st = P->state.as<PcgState>();
  a.fin = fin;
  a.done = done;
  for (int g = 0; g < 8; ++g) {
    a.kv64[g] = P->op.kv[g];
    a.kv32[g] = float(P->op.kv[g]);
  }
  const bool fusep = zfuse_pin != nullptr;
  // kind 0: products of the PCG loop (K p, true residual); kind 1: V-cycle
  // residual; kind 2: V-cycle Jacobi sweeps
  ProfEvent ev{};
  TL_TRY(prof_begin(P, (mode == KU_STORE || mode == KU_STORE_DOT) ? 0 : (mode == KU_RESID ? 1 : 2),
                    ev));
  const int rc = fine_apply_dispatch(P, a, prec, mode, fusep, b64, out64);
  TL_TRY(prof_end(P, ev));
  return rc;
}

int prof_begin(tl_problem* P, int kind, ProfEvent& ev) {
  ev.kind = -1;
  if (!P->prof_on) return TL_OK;
  if (P->prof_free.empty()) {
    TL_HIP(hipEventCreate(&ev.start));
    TL_HIP(hipEventCreate(&ev.stop));
  } else {
    ev = P->prof_free.back();
    P->prof_free.pop_back();
  }
  ev.kind = kind;
  TL_HIP(hipEventRecord(ev.start, P->stream));
  return TL_OK;
}

int prof_end(tl_problem* P, ProfEvent& ev) {
  if (ev.kind < 0) return TL_OK;
  TL_HIP(hipEventRecord(ev.stop, P->stream));
  P->prof_used.push_back(ev);
  return TL_OK;
}

static int fine_apply_dispatch(tl_problem* P, KuArgs& a, int prec, int mode, bool fusep, bool b64,
                               bool out64) {
  if (P->rec && (prec == TL_PREC_BF16 || prec == TL_PREC_FP16 || (prec == TL_PREC_FP64 &&
                                                                   P->op.mfma64))) {
    P->rec->unsupported = true;  // no multi-instance form of the MFMA operators
    return TL_OK;
  }
  if (prec == TL_PREC_BF16 || prec == TL_PREC_FP16) {
    // V-cycle smoothing and residual on the matrix cores
    TL_ARG(b64 && !fusep && (mode == KU_RESID || mode == KU_JACOBI || mode == KU_JACOBI_DOT),
           "internal: the MFMA operator serves the V-cycle modes with an fp64 right-hand side");
    unsigned nb = 0;
    TL_TRY(fine_apply_mfma(P, a.u, a.y, prec, mode, a.b, a.dinv, a.omega, a.partials, a.done, &nb));
    if (mode == KU_JACOBI_DOT) return launch_fold(P, nb, a.fin, a.dot_out, a.done);
    return TL_OK;
  }
  if (prec == TL_PREC_FP64) {
    if (P->op.mfma64 && P->g.dim == 3 && P->ndpn == 3 &&
        ((mode == KU_STORE && out64) || (mode == KU_STORE_DOT && fusep) ||
         (b64 && (mode == KU_RESID || mode == KU_JACOBI || mode == KU_JACOBI_DOT)))) {
      unsigned nb = 0;
      TL_TRY(fine_apply_mfma64(P, static_cast<const double*>(a.u), static_cast<double*>(a.y), mode,
                               static_cast<const double*>(a.b), static_cast<const double*>(a.dinv),
                               a.omega, static_cast<const double*>(a.pin),
                               static_cast<double*>(a.pout), a.beta, a.partials, a.done, &nb));
      if (mode == KU_STORE_DOT || mode == KU_JACOBI_DOT)
        return launch_fold(P, nb, a.fin, a.dot_out, a.done);
      return TL_OK;
    }
    a.sp = P->op.sp64.ptr; a.KE = P->op.KE64.ptr;
    switch (P->ndpn) {
      case 1: return dispatch_mode<1, double>(P, a, mode, fusep, b64, out64);
      case 2: return dispatch_mode<2, double>(P, a, mode, fusep, b64, out64);
      case 3: return dispatch_mode<3, double>(P, a, mode, fusep, b64, out64);
    }
  } else {
    a.sp = P->op.sp32.ptr; a.KE = P->op.KE32.ptr;
    switch (P->ndpn) {
      case 1: return dispatch_mode<1, float>(P, a, mode, fusep, b64, out64);
      case 2: return dispatch_mode<2, float>(P, a, mode, fusep, b64, out64);
      case 3: return dispatch_mode<3, float>(P, a, mode, fusep, b64, out64);
    }
  }
  set_error("unsupported DOFs per node");
  return TL_EINVAL;
}

// ----------------------------------------------------------------------------
// diagonal of K_ff
// ----------------------------------------------------------------------------
template <int NDPN>
__global__ void diag_kernel(const double* __restrict__ sp, const double* __restrict__ KE,
                            const uint8_t* __restrict__ mask, double* __restrict__ d64,
                            float* __restrict__ dinv32, double* __restrict__ dinv64, int dim,
                            int nelx, int nely, int nelz, int ny, int64_t nnodes) {
  const int64_t n = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (n >= nnodes) return;
  const int nd = (dim == 3 ? 8 : 4) * NDPN;
  const int plane = (nelx + 1) * ny;
  // 32-bit divisions (fewer than 2^31 nodes, tl_create)
  const int iz = int(uint32_t(n) / uint32_t(plane));
  const int j = int(n) - iz * plane;
  const int ix = int(uint32_t(j) / uint32_t(ny)), iy = j - ix * ny;
  double acc[NDPN];
  for (int d = 0; d < NDPN; ++d) acc[d] = 0.0;
  for (int oz = (dim == 3 ? -1 : 0); oz <= 0; ++oz)
    for (int ox = -1; ox <= 0; ++ox)
      for (int oy = -1; oy <= 0; ++oy) {
        const int64_t pidx =
            (dim == 3 ? (int64_t(iz + oz + 1) * (nelx + 2) + ix + ox + 1) * (nely + 2)
                      : int64_t(ix + ox + 1) * (nely + 2)) + iy + oy + 1;
        const double s = sp[pidx];
        const int la = loc_of(-ox, -oy, -oz);
        for (int d = 0; d < NDPN; ++d) acc[d] += s * KE[(la * NDPN + d) * nd + la * NDPN + d];
      }
  const uint8_t m = mask[n];
  for (int d = 0; d < NDPN; ++d) {
    const bool fixed = (m >> d) & 1;
    const int64_t g = n * NDPN + d;
    if (d64) d64[g] = fixed ? 0.0 : acc[d];
    const double inv = (fixed || acc[d] == 0.0) ? 0.0 : 1.0 / acc[d];
    if (dinv64) dinv64[g] = inv;
    if (dinv32) dinv32[g] = float(inv);
  }
}

int fine_diag_to(tl_problem* P, double* d64, double* dinv64, float* dinv32) {
  const Grid& g = P->g;
  const unsigned nb = unsigned((g.nnodes + BS - 1) / BS);
#define TL_DIAG(N)                                                                       \
  hipLaunchKernelGGL(diag_kernel<N>, dim3(nb), dim3(BS), 0, P->stream,                  \
                     P->op.sp64.as<double>(), P->op.KE64.as<double>(),                  \
                     P->op.mask.as<uint8_t>(), d64, dinv32, dinv64, g.dim, g.nelx, g.nely, \
                     g.nelz, g.ny, g.nnodes)
  switch (P->ndpn) {
    case 1: TL_DIAG(1); break;
    case 2: TL_DIAG(2); break;
    case 3: TL_DIAG(3); break;
    default: set_error("unsupported DOFs per node"); return TL_EINVAL;
  }
#undef TL_DIAG
  TL_CHECK_LAUNCH();
  return TL_OK;
}

int fine_diag(tl_problem* P) {
  return fine_diag_to(P, nullptr, P->op.dinv64.as<double>(), P->op.dinv32.as<float>());
}

// ----------------------------------------------------------------------------
// vector kernels of the PCG loop
// ----------------------------------------------------------------------------
template <int NV>
__device__ inline void reduce_store(double (&v)[NV], double* partials) {
  __shared__ double sh[NV * (BS / 64) + 4];
  block_sum<NV>(v, sh);
  store_partials<NV>(v, partials);
}

// One block: fold the partials of the preceding kernel (nb blocks) and apply
// the scalar update `fin`.  Skipped, like the kernel it follows, once the
// solve is done.
__global__ __launch_bounds__(BS) void fold_kernel(const double* __restrict__ partials, unsigned nb,
                                                  int fin, PcgState* st, double* out,
                                                  const int* done) {
  if (done && *done) return;
  __shared__ double sh[BS / 64 + 2];
  double tot[1];
  fold_partials<1>(partials, nb, tot, sh);
  if (threadIdx.x == 0) finalize(fin, st, tot[0], out);
}

struct FoldArgs {
  const double* partials;
  unsigned nb;
  int fin;
  PcgState* st;
  double* out;
  const int* done;
};

// one block per system (tiles[b].y == 0)
__global__ __launch_bounds__(BS) void fold_multi_kernel(const FoldArgs* __restrict__ descs,
                                                        const uint2* __restrict__ tiles) {
  const FoldArgs a = descs[tiles[blockIdx.x].x];
  if (a.done && *a.done) return;
  __shared__ double sh[BS / 64 + 2];
  double tot[1];
  fold_partials<1>(a.partials, a.nb, tot, sh);
  if (threadIdx.x == 0) finalize(a.fin, a.st, tot[0], a.out);
}

static int fold_multi_launch(hipStream_t s, const void* descs, const uint2* tiles, unsigned nb,
                             size_t lds) {
  hipLaunchKernelGGL(fold_multi_kernel, dim3(nb), dim3(BS), lds, s,
                     static_cast<const FoldArgs*>(descs), tiles);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

static unsigned red_blocks(int64_t n) {
  // enough workgroups to keep HBM busy (each thread streams ~5 elements at
  // the 5.4M-DOF size), bounded by the reduction slot count
  int64_t nb = (n + BS - 1) / BS;
  nb = std::min<int64_t>(nb, kMaxRedBlocks);
  return unsigned(std::max<int64_t>(nb, 1));
}

// out = a.b (fin selects where the total goes)
__global__ __launch_bounds__(BS) void dot_kernel(const double* __restrict__ a,
                                                 const double* __restrict__ b, int64_t n,
                                                 double* partials) {
  double s = 0.0;
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BS)
    s += a[i] * b[i];
  double v[1] = {s};
  reduce_store<1>(v, partials);
}

// r = b - y (masked by y's zeros already); rr
__global__ __launch_bounds__(BS) void resid_kernel(const double* __restrict__ b,
                                                   const double* __restrict__ y,
                                                   const uint8_t* __restrict__ mask, int ndpn,
                                                   double* __restrict__ r, int64_t n,
                                                   double* partials) {
  double s = 0.0;
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BS) {
    const bool fixed = (mask[i / ndpn] >> (i % ndpn)) & 1;
    const double v = fixed ? 0.0 : b[i] - y[i];
    if (r) r[i] = v;
    s += v * v;
  }
  double v[1] = {s};
  reduce_store<1>(v, partials);
}

// x += alpha p ; r -= alpha q ; r.r partials.  With xs (GMG with V(1,1)),
// also the next V-cycle's first fine-level sweep xs = omega * dinv * r from
// the new residual, in the V-cycle's precision TS (saves that sweep's pass).
// x, p, r and q move with non-temporal loads and stores: they stream through
// once per iteration, and leaving them out of the last-level cache keeps
// more of the level-1 stencil there between V-cycles (the stencil residual
// that follows runs 71 -> 68 us at 192x96x96; solve 39.1 -> 38.6 ms).
template <typename TS>
struct UpdateArgs {
  double* x;
  double* r;
  const double* p;
  const double* q;
  int64_t n;
  double* partials;
  const PcgState* st;
  TS* xs;
  const TS* dinv;
  TS omega;
  const int* done;
  unsigned nb;  // blocks of the system's grid-stride loop
};

template <typename TS>
__device__ __forceinline__ void update_body(const UpdateArgs<TS>& a, unsigned bid) {
  if (*a.done) return;
  const double alpha = a.st->alpha;
  double s = 0.0;
  for (int64_t i = bid * int64_t(BS) + threadIdx.x; i < a.n; i += int64_t(a.nb) * BS) {
    const double xv = __builtin_nontemporal_load(&a.x[i]) + alpha * __builtin_nontemporal_load(&a.p[i]);
    __builtin_nontemporal_store(xv, &a.x[i]);
    const double rv = __builtin_nontemporal_load(&a.r[i]) - alpha * __builtin_nontemporal_load(&a.q[i]);
    __builtin_nontemporal_store(rv, &a.r[i]);
    s += rv * rv;
    if (a.xs) a.xs[i] = zero_guess_sweep<TS, double>(a.omega, a.dinv[i], rv);
  }
  double v[1] = {s};
  __shared__ double sh[BS / 64 + 4];
  block_sum<1>(v, sh);
  store_partials_at<1>(v, a.partials, bid);
}

template <typename TS>
__global__ __launch_bounds__(BS) void update_kernel(UpdateArgs<TS> a) {
  update_body<TS>(a, blockIdx.x);
}

template <typename TS>
__global__ __launch_bounds__(BS) void update_multi_kernel(const UpdateArgs<TS>* __restrict__ descs,
                                                          const uint2* __restrict__ tiles) {
  const uint2 t = tiles[blockIdx.x];
  update_body<TS>(descs[t.x], t.y);
}

template <typename TS>
static int update_multi_launch(hipStream_t s, const void* descs, const uint2* tiles, unsigned nb,
                               size_t lds) {
  hipLaunchKernelGGL(update_multi_kernel<TS>, dim3(nb), dim3(BS), lds, s,
                     static_cast<const UpdateArgs<TS>*>(descs), tiles);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

// xs = omega * dinv * r (the first fine-level sweep of the first V-cycle)
template <typename TS>
__global__ __launch_bounds__(BS) void zero_sweep_kernel(const double* __restrict__ r,
                                                        const TS* __restrict__ dinv, TS omega,
                                                        TS* __restrict__ xs, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BS)
    xs[i] = zero_guess_sweep<TS, double>(omega, dinv[i], r[i]);
}

// z = dinv * r (or z = r when dinv == nullptr); rho = r.z
__global__ __launch_bounds__(BS) void jacobi_prec_kernel(const double* __restrict__ r,
                                                         const double* __restrict__ dinv,
                                                         double* __restrict__ z, int64_t n,
                                                         double* partials, const int* done) {
  if (*done) return;
  double s = 0.0;
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BS) {
    const double zv = dinv ? dinv[i] * r[i] : r[i];
    z[i] = zv;
    s += r[i] * zv;
  }
  double v[1] = {s};
  reduce_store<1>(v, partials);
}

// masked copy (fixed DOFs -> 0)
__global__ void masked_copy_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                   const uint8_t* __restrict__ mask, int ndpn, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const bool fixed = (mask[i / ndpn] >> (i % ndpn)) & 1;
    dst[i] = fixed ? 0.0 : src[i];
  }
}

__global__ void cast_kernel_d2f(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    b[i] = float(a[i]);
}

__global__ void cast_kernel_f2d(const float* __restrict__ a, double* __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    b[i] = double(a[i]);
}

__global__ void neg_cast_kernel_f2d(const float* __restrict__ a, double* __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    b[i] = -double(a[i]);
}

__global__ void init_state_kernel(PcgState* st, double rtol, double atol, int maxiter) {
  // tolerance as linear_solvers.py:107-109 (scipy _get_atol_rtol): max(atol, rtol*||b||)
  const double tol = fmax(atol, rtol * sqrt(st->bb));
  st->tol2 = tol * tol;
  st->maxiter = maxiter;
  st->it = 0;
  st->done = 0;
  st->rho = st->rho_prev = st->pq = st->alpha = st->beta = 0.0;
}

static unsigned grid_for(int64_t n) {
  return unsigned(std::max<int64_t>(1, std::min<int64_t>((n + BS - 1) / BS, 8192)));
}

static int launch_fold(tl_problem* P, unsigned nb, int fin, double* out, const int* done) {
  if (P->rec) {
    P->rec->add(P->rec_phase, P->rec_level, &fold_multi_launch,
                FoldArgs{P->partials.as<double>(), nb, fin, P->state.as<PcgState>(), out, done},
                1u, 0);
    return TL_OK;
  }
  hipLaunchKernelGGL(fold_kernel, dim3(1), dim3(BS), 0, P->stream, P->partials.as<double>(), nb,
                     fin, P->state.as<PcgState>(), out, done);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

}  // namespace tl

// ----------------------------------------------------------------------------
// handle lifetime
// ----------------------------------------------------------------------------
tl_problem::~tl_problem() {
  for (auto& ev : prof_used) { (void)hipEventDestroy(ev.start); (void)hipEventDestroy(ev.stop); }
  for (auto& ev : prof_free) { (void)hipEventDestroy(ev.start); (void)hipEventDestroy(ev.stop); }
  op.KE64.release(); op.KE32.release(); op.KE16.release();
  op.sp64.release(); op.sp32.release();
  op.dinv64.release(); op.dinv32.release(); op.mask.release();
  if (chunk_exec) (void)hipGraphExecDestroy(chunk_exec);
  for (auto& l : mg.lv) l.release();
  mg.G1.release(); mg.W.release(); mg.cinv.release(); mg.gjw.release();
  r.release(); z.release(); p.release(); p2.release(); q.release(); t.release(); xb.release();
  partials.release(); state.release();
}

namespace tl {
// Use the pattern path when every KE entry equals (+-) its group's first
// entry to 1e-13 of max|KE|; the kernels then use the group values.
static void detect_ke_pattern(tl_problem* P) {
  P->op.fast_ke = false;
  const int nd = P->nd;
  const bool d3 = P->g.dim == 3;
  int (*pat)(int) = nullptr;
  const int* first = nullptr;
  int ng = 0;
  if (d3 && P->ndpn == 3) { pat = [](int i) { return kPat_h8_elast(i); }; first = kFirst_h8_elast; ng = kGroups_h8_elast; }
  else if (d3 && P->ndpn == 1) { pat = [](int i) { return kPat_h8_poisson(i); }; first = kFirst_h8_poisson; ng = kGroups_h8_poisson; }
  else if (!d3 && P->ndpn == 2) { pat = [](int i) { return kPat_q4_elast(i); }; first = kFirst_q4_elast; ng = kGroups_q4_elast; }
  else if (!d3 && P->ndpn == 1) { pat = [](int i) { return kPat_q4_poisson(i); }; first = kFirst_q4_poisson; ng = kGroups_q4_poisson; }
  if (!pat || ng > 8) return;
  double mx = 0.0;
  for (double v : P->KEh) mx = std::max(mx, std::fabs(v));
  for (int g = 0; g < ng; ++g) P->op.kv[g] = P->KEh[first[g]];
  for (int i = 0; i < nd * nd; ++i) {
    const int pg = pat(i);
    const double v = (pg > 0 ? 1.0 : -1.0) * P->op.kv[(pg > 0 ? pg : -pg) - 1];
    if (std::fabs(v - P->KEh[i]) > 1e-13 * mx) return;
  }
  P->op.fast_ke = true;
}
}  // namespace tl

using namespace tl;

extern "C" {

const char* tl_last_error(void) { return tl::g_last_error.c_str(); }

int64_t tl_ndof(const tl_problem* p) { return p ? p->ndof : -1; }
int64_t tl_nel(const tl_problem* p) { return p ? p->g.nel : -1; }

int tl_create(const tl_grid_desc* d, tl_problem** out) {
  TL_ARG(d && out, "null argument");
  TL_ARG(d->ndim == 2 || d->ndim == 3, "ndim must be 2 or 3");
  TL_ARG(d->nelx > 0 && d->nely > 0 && (d->ndim == 2 || d->nelz > 0), "element counts must be positive");
  TL_ARG(d->ndpn == 1 || d->ndpn == d->ndim, "ndpn must be 1 or ndim");
  TL_ARG(d->KE != nullptr, "KE is null");
  TL_ARG(d->nfixed >= 0 && (d->nfixed == 0 || d->fixed), "fixed DOF list is null");
  const int64_t nmax = int64_t(1) << 31;
  const int64_t nodes = int64_t(d->nelx + 1) * (d->nely + 1) * (d->ndim == 3 ? d->nelz + 1 : 1);
  TL_ARG(nodes * d->ndpn < nmax, "grid too large for 32-bit plane indexing");
  auto* P = new tl_problem();
  P->g.set(d->ndim, d->nelx, d->nely, d->nelz);
  P->ndpn = d->ndpn;
  P->nd = (d->ndim == 3 ? 8 : 4) * d->ndpn;
  P->ndof = P->g.nnodes * d->ndpn;
  P->eps = d->eps;
  P->penal = d->penal;
  P->stream = static_cast<hipStream_t>(d->stream);
  const int nd = P->nd;
  P->KEh.assign(d->KE, d->KE + nd * nd);
  std::vector<uint8_t> mask(P->g.nnodes, 0);
  for (int64_t i = 0; i < d->nfixed; ++i) {
    const int64_t f = d->fixed[i];
    if (f < 0 || f >= P->ndof) {
      delete P;
      set_error("fixed DOF index out of range");
      return TL_EINVAL;
    }
    mask[f / d->ndpn] |= uint8_t(1u << (f % d->ndpn));
  }
  P->fixed_h.assign(d->fixed, d->fixed + d->nfixed);
  detect_ke_pattern(P);
  P->op.mfma64 = false;  // the vector kernel is faster (tl_set_fp64_operator selects)
  std::vector<float> ke32(nd * nd);
  for (int i = 0; i < nd * nd; ++i) ke32[i] = float(P->KEh[i]);
  auto fail = [&](int rc) { delete P; return rc; };
  int rc;
  if ((rc = P->op.KE64.alloc(sizeof(double) * nd * nd)) ||
      (rc = P->op.KE32.alloc(sizeof(float) * nd * nd)) ||
      (rc = P->op.mask.alloc(P->g.nnodes)) ||
      (rc = P->op.sp64.alloc(sizeof(double) * P->g.npad())) ||
      (rc = P->op.sp32.alloc(sizeof(float) * P->g.npad())) ||
      (rc = P->op.dinv64.alloc(sizeof(double) * P->ndof)) ||
      (rc = P->op.dinv32.alloc(sizeof(float) * P->ndof)) ||
      (rc = P->partials.alloc(sizeof(double) * 4 * kMaxRedBlocks)) ||
      (rc = P->state.alloc(sizeof(PcgState))))
    return fail(rc);
  hipError_t e;
  if ((e = hipMemcpy(P->op.KE64.ptr, P->KEh.data(), sizeof(double) * nd * nd, hipMemcpyHostToDevice)) ||
      (e = hipMemcpy(P->op.KE32.ptr, ke32.data(), sizeof(float) * nd * nd, hipMemcpyHostToDevice)) ||
      (e = hipMemcpy(P->op.mask.ptr, mask.data(), P->g.nnodes, hipMemcpyHostToDevice)) ||
      (e = hipMemset(P->state.ptr, 0, sizeof(PcgState))) ||
      (e = hipMemset(P->op.sp64.ptr, 0, sizeof(double) * P->g.npad())) ||
      (e = hipMemset(P->op.sp32.ptr, 0, sizeof(float) * P->g.npad()))) {
    set_error(std::string("tl_create: ") + hipGetErrorString(e));
    return fail(TL_EHIP);
  }
  *out = P;
  return TL_OK;
}

void tl_destroy(tl_problem* p) { delete p; }

int tl_set_stream(tl_problem* p, void* stream) {
  TL_ARG(p, "null handle");
  p->stream = static_cast<hipStream_t>(stream);
  return TL_OK;
}

int tl_set_fp64_operator(tl_problem* p, int32_t kind) {
  TL_ARG(p, "null handle");
  TL_ARG(kind == 0 || kind == 1, "kind must be 0 (vector FMA) or 1 (MFMA)");
  TL_ARG(kind == 0 || (p->g.dim == 3 && p->ndpn == 3),
         "the FP64 MFMA operator covers 3D H8 elasticity only");
  p->op.mfma64 = kind == 1;
  ++p->gen;  // captured iteration chunks hold the previous kernel
  return TL_OK;
}

int tl_set_density(tl_problem* P, const double* xPhys) {
  TL_ARG(P && xPhys, "null argument");
  const Grid& g = P->g;
  hipLaunchKernelGGL(simp_pad_kernel, dim3(grid_for(g.npad())), dim3(BS), 0, P->stream, xPhys,
                     P->op.sp64.as<double>(), P->op.sp32.as<float>(), g.dim, g.nelx, g.nely,
                     g.nelz, P->eps, P->penal);
  TL_CHECK_LAUNCH();
  TL_TRY(fine_diag(P));
  P->op.have_density = true;
  P->mg.ready = false;
  return TL_OK;
}

int tl_apply_K(tl_problem* P, const double* u, double* y, int32_t prec) {
  TL_ARG(P && u && y, "null argument");
  TL_ARG(P->op.have_density, "tl_set_density must be called first");
  TL_ARG(prec == TL_PREC_FP64 || prec == TL_PREC_FP32 || prec == TL_PREC_BF16 ||
             prec == TL_PREC_FP16,
         "prec must be FP64, FP32, BF16 or FP16");
  TL_ARG(prec == TL_PREC_FP64 || prec == TL_PREC_FP32 || (P->g.dim == 3 && P->ndpn == 3),
         "BF16/FP16 (MFMA) products are implemented for 3D H8 elasticity only");
  TL_TRY(P->t.alloc(sizeof(double) * P->ndof));
  const unsigned nb = grid_for(P->ndof);
  hipLaunchKernelGGL(masked_copy_kernel, dim3(nb), dim3(BS), 0, P->stream, u, P->t.as<double>(),
                     P->op.mask.as<uint8_t>(), P->ndpn, P->ndof);
  TL_CHECK_LAUNCH();
  if (prec == TL_PREC_FP64)
    return fine_apply_ex(P, P->t.ptr, y, TL_PREC_FP64, KU_STORE, nullptr, false, nullptr, 0.0,
                         nullptr, nullptr, nullptr, nullptr, FIN_STORE, nullptr, true);
  // fp32: cast input, apply, widen output
  TL_TRY(P->z.alloc(sizeof(double) * P->ndof));
  float* u32 = reinterpret_cast<float*>(P->z.ptr);
  float* y32 = u32 + P->ndof;  // z holds 2*ndof floats
  hipLaunchKernelGGL(cast_kernel_d2f, dim3(nb), dim3(BS), 0, P->stream, P->t.as<double>(), u32,
                     P->ndof);
  TL_CHECK_LAUNCH();
  if (prec == TL_PREC_FP32) {
    TL_TRY(fine_apply_ex(P, u32, y32, TL_PREC_FP32, KU_STORE, nullptr, false, nullptr, 0.0,
                         nullptr, nullptr, nullptr, nullptr, FIN_STORE, nullptr, false));
    hipLaunchKernelGGL(cast_kernel_f2d, dim3(nb), dim3(BS), 0, P->stream, y32, y, P->ndof);
    TL_CHECK_LAUNCH();
    return TL_OK;
  }
  // bf16/fp16: the MFMA operator's residual mode with b = 0 gives -K u
  TL_TRY(P->q.alloc(sizeof(double) * P->ndof));
  TL_HIP(hipMemsetAsync(P->q.ptr, 0, sizeof(double) * P->ndof, P->stream));
  TL_TRY(fine_apply_ex(P, u32, y32, prec, KU_RESID, P->q.ptr, true, nullptr, 0.0, nullptr,
                       nullptr, nullptr, nullptr, 0, nullptr, false));
  hipLaunchKernelGGL(neg_cast_kernel_f2d, dim3(nb), dim3(BS), 0, P->stream, y32, y, P->ndof);
  TL_CHECK_LAUNCH();
  return TL_OK;
}

int tl_profile(tl_problem* P, int32_t enable) {
  TL_ARG(P, "null handle");
  P->prof_on = enable != 0;
  return TL_OK;
}

int tl_profile_read(tl_problem* P, int32_t kind, double* total_ms, int64_t* launches) {
  TL_ARG(P && total_ms && launches, "null argument");
  double tot = 0.0;
  int64_t cnt = 0;
  std::vector<ProfEvent> keep;
  for (auto& ev : P->prof_used) {
    if (ev.kind != kind) {
      keep.push_back(ev);
      continue;
    }
    TL_HIP(hipEventSynchronize(ev.stop));
    float ms = 0.f;
    TL_HIP(hipEventElapsedTime(&ms, ev.start, ev.stop));
    tot += ms;
    ++cnt;
    P->prof_free.push_back(ev);
  }
  P->prof_used.swap(keep);
  *total_ms = tot;
  *launches = cnt;
  return TL_OK;
}

int tl_diag_K(tl_problem* P, double* d) {
  TL_ARG(P && d, "null argument");
  TL_ARG(P->op.have_density, "tl_set_density must be called first");
  return fine_diag_to(P, d, nullptr, nullptr);
}

}  // extern "C"

namespace tl {
namespace {

// Buffers and launch shapes of one PCG solve on a handle.
struct PcgCtx {
  int64_t n = 0;
  double *r = nullptr, *z = nullptr, *q = nullptr, *t = nullptr;
  double* pp[2] = {nullptr, nullptr};  // search directions, ping-pong
  PcgState* st = nullptr;
  const int* done = nullptr;
  double* partials = nullptr;
  const uint8_t* mask = nullptr;
  unsigned nb = 0, rb = 0;
};

int pcg_context(tl_problem* P, PcgCtx& c) {
  const int64_t n = P->ndof;
  for (DevBuf* b : {&P->r, &P->z, &P->p, &P->p2, &P->q, &P->t}) TL_TRY(b->alloc(sizeof(double) * n));
  c.n = n;
  c.r = P->r.as<double>(); c.z = P->z.as<double>(); c.q = P->q.as<double>(); c.t = P->t.as<double>();
  c.pp[0] = P->p.as<double>(); c.pp[1] = P->p2.as<double>();
  c.st = P->state.as<PcgState>();
  c.done = &c.st->done;
  c.partials = P->partials.as<double>();
  c.mask = P->op.mask.as<uint8_t>();
  c.nb = grid_for(n); c.rb = red_blocks(n);
  return TL_OK;
}

// t = masked f, bb, state, x = masked x0 or 0, r = t - K x, rr0, p = 0.
// GMG-PCG with the default V(1,1) cycle: the residual update writes the next
// V-cycle's first fine-level sweep (the fine iterate buffer never rotates
// with one pre- and one post-smoothing sweep).
bool pcg_fuses_x0(const tl_problem* P, int precond) {
  return precond == TL_PRECOND_GMG && P->mg.ready && P->mg.nu1 == 1 && P->mg.nu2 == 1 &&
         P->mg.nlevels >= 2 && !P->mg.lv.empty() && P->mg.lv[0].vx.ptr != nullptr;
}

int pcg_init(tl_problem* P, const PcgCtx& c, const double* f, const double* x0, double* x,
             const tl_solve_opts* o) {
  hipStream_t s = P->stream;
  const int64_t n = c.n;
  hipLaunchKernelGGL(masked_copy_kernel, dim3(c.nb), dim3(BS), 0, s, f, c.t, c.mask, P->ndpn, n);
  TL_CHECK_LAUNCH();
  hipLaunchKernelGGL(dot_kernel, dim3(c.rb), dim3(BS), 0, s, c.t, c.t, n, c.partials);
  TL_CHECK_LAUNCH();
  TL_TRY(launch_fold(P, c.rb, FIN_BB, nullptr, nullptr));
  hipLaunchKernelGGL(init_state_kernel, dim3(1), dim3(1), 0, s, c.st, o->rtol, o->atol, o->maxiter);
  TL_CHECK_LAUNCH();
  if (x0) {
    hipLaunchKernelGGL(masked_copy_kernel, dim3(c.nb), dim3(BS), 0, s, x0, x, c.mask, P->ndpn, n);
    TL_CHECK_LAUNCH();
    TL_TRY(fine_apply_ex(P, x, c.q, TL_PREC_FP64, KU_STORE, nullptr, false, nullptr, 0.0, nullptr,
                         nullptr, nullptr, nullptr, FIN_STORE, nullptr, true));
  } else {
    TL_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
    TL_HIP(hipMemsetAsync(c.q, 0, sizeof(double) * n, s));
  }
  hipLaunchKernelGGL(resid_kernel, dim3(c.rb), dim3(BS), 0, s, c.t, c.q, c.mask, P->ndpn, c.r, n,
                     c.partials);
  TL_CHECK_LAUNCH();
  TL_TRY(launch_fold(P, c.rb, FIN_RR0, nullptr, nullptr));
  TL_HIP(hipMemsetAsync(c.pp[0], 0, sizeof(double) * n, s));
  TL_HIP(hipMemsetAsync(c.pp[1], 0, sizeof(double) * n, s));
  if (pcg_fuses_x0(P, o->precond)) {
    const MgState& M = P->mg;
    if (M.prec == TL_PREC_FP64)
      hipLaunchKernelGGL(zero_sweep_kernel<double>, dim3(c.nb), dim3(BS), 0, s, c.r,
                         P->op.dinv64.as<double>(), M.omega, M.lv[0].vx.as<double>(), n);
    else
      hipLaunchKernelGGL(zero_sweep_kernel<float>, dim3(c.nb), dim3(BS), 0, s, c.r,
                         P->op.dinv32.as<float>(), float(M.omega), M.lv[0].vx.as<float>(), n);
    TL_CHECK_LAUNCH();
  }
  return TL_OK;
}

// p_out = z + beta p_in (the 3D search-direction update, before K p), with
// non-temporal loads and stores like the x/r update
__global__ __launch_bounds__(BS) void pupdate_kernel(const double* __restrict__ z,
                                                     const double* __restrict__ pin,
                                                     double* __restrict__ pout,
                                                     const double* __restrict__ beta, int64_t n,
                                                     const int* done) {
  if (*done) return;
  const double b = *beta;
  for (int64_t i = blockIdx.x * int64_t(BS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * BS)
    __builtin_nontemporal_store(__builtin_nontemporal_load(&z[i]) + b * __builtin_nontemporal_load(&pin[i]),
                                &pout[i]);
}

// One PCG iteration (k counts from 0 after pcg_init).  Every kernel returns
// at once when the device state says done, so iterations launched past
// convergence change nothing.  No allocation, memset or host transfer: the
// sequence can be captured into a graph.
int pcg_iteration(tl_problem* P, const PcgCtx& c, double* x, int precond, int64_t k) {
  hipStream_t s = P->stream;
  // z = M r ; rho = r.z ; beta
  const bool fx0 = pcg_fuses_x0(P, precond);
  if (precond == TL_PRECOND_GMG) {
    TL_TRY(mg_vcycle(P, c.r, c.z, c.done, fx0));
  } else {
    if (P->rec) {  // no multi-instance form of the Jacobi preconditioner
      P->rec->unsupported = true;
      return TL_OK;
    }
    hipLaunchKernelGGL(jacobi_prec_kernel, dim3(c.rb), dim3(BS), 0, s, c.r,
                       precond == TL_PRECOND_JACOBI ? P->op.dinv64.as<double>() : nullptr, c.z,
                       c.n, c.partials, c.done);
    TL_CHECK_LAUNCH();
    TL_TRY(launch_fold(P, c.rb, FIN_RHO, nullptr, c.done));
  }
  // p = z + beta p ; q = K p ; pq ; alpha.  The fused product (2D, and the
  // 3D MFMA operator) reads p_k (halo included) while it writes p_{k+1}:
  // two buffers.  The 3D vector operator takes the update as a separate pass.
  double* pin = c.pp[k & 1];
  double* pout = c.pp[(k + 1) & 1];
  P->rec_phase = 2;  // the PCG update after the preconditioner
  P->rec_level = 0;
  if (P->g.dim == 3 && P->rec) {
    P->rec->unsupported = true;
    return TL_OK;
  }
  if (P->g.dim == 3 && !(P->op.mfma64 && P->ndpn == 3)) {
    hipLaunchKernelGGL(pupdate_kernel, dim3(c.nb), dim3(BS), 0, s, c.z, pin, pout, &c.st->beta,
                       c.n, c.done);
    TL_CHECK_LAUNCH();
    TL_TRY(fine_apply_ex(P, pout, c.q, TL_PREC_FP64, KU_STORE_DOT, nullptr, false, nullptr, 0.0,
                         nullptr, nullptr, nullptr, nullptr, FIN_PQ, c.done, false));
  } else {
    TL_TRY(fine_apply_ex(P, c.z, c.q, TL_PREC_FP64, KU_STORE_DOT, nullptr, false, nullptr, 0.0,
                         pin, pout, &c.st->beta, nullptr, FIN_PQ, c.done, false));
  }
  // x += alpha p ; r -= alpha q ; rr ; done
  auto upd = [&](auto ua) -> int {
    using TS = std::remove_pointer_t<decltype(ua.xs)>;
    if (P->rec) {
      P->rec->add(2, 0, &update_multi_launch<TS>, ua, c.rb, 0);
      return TL_OK;
    }
    hipLaunchKernelGGL(update_kernel<TS>, dim3(c.rb), dim3(BS), 0, s, ua);
    TL_CHECK_LAUNCH();
    return TL_OK;
  };
  if (fx0 && P->mg.prec == TL_PREC_FP64)
    TL_TRY(upd(UpdateArgs<double>{x, c.r, pout, c.q, c.n, c.partials, c.st,
                                  P->mg.lv[0].vx.as<double>(), P->op.dinv64.as<double>(),
                                  P->mg.omega, c.done, c.rb}));
  else if (fx0)
    TL_TRY(upd(UpdateArgs<float>{x, c.r, pout, c.q, c.n, c.partials, c.st,
                                 P->mg.lv[0].vx.as<float>(), P->op.dinv32.as<float>(),
                                 float(P->mg.omega), c.done, c.rb}));
  else
    TL_TRY(upd(UpdateArgs<double>{x, c.r, pout, c.q, c.n, c.partials, c.st, nullptr, nullptr,
                                  0.0, c.done, c.rb}));
  return launch_fold(P, c.rb, FIN_RR, nullptr, c.done);
}

// True residual |t - K x|^2 into *tr2_dev (a device slot), state unchanged.
int pcg_true_residual(tl_problem* P, const PcgCtx& c, const double* x, double* tr2_dev) {
  TL_TRY(fine_apply_ex(P, x, c.q, TL_PREC_FP64, KU_STORE, nullptr, false, nullptr, 0.0, nullptr,
                       nullptr, nullptr, nullptr, FIN_STORE, nullptr, true));
  hipLaunchKernelGGL(resid_kernel, dim3(c.rb), dim3(BS), 0, P->stream, c.t, c.q, c.mask, P->ndpn,
                     static_cast<double*>(nullptr), c.n, c.partials);
  TL_CHECK_LAUNCH();
  return launch_fold(P, c.rb, FIN_STORE, tr2_dev, nullptr);
}

bool pcg_info(const PcgState& hs, double tr2, tl_solve_info* info) {
  const bool conv = hs.rr < hs.tol2;
  if (info) {
    const double bn = std::sqrt(hs.bb);
    info->iterations = hs.it;
    info->status = conv ? 0 : TL_ENOCONV;
    info->relres_recursive = bn > 0 ? std::sqrt(hs.rr) / bn : 0.0;
    info->relres_true = bn > 0 ? std::sqrt(tr2) / bn : 0.0;
  }
  return conv;
}

}  // namespace
}  // namespace tl

extern "C" {

int tl_solve(tl_problem* P, const double* f, double* u, const tl_solve_opts* o,
             tl_solve_info* info) {
  TL_ARG(P && f && u && o, "null argument");
  TL_ARG(P->op.have_density, "tl_set_density must be called first");
  TL_ARG(o->precond >= TL_PRECOND_NONE && o->precond <= TL_PRECOND_GMG, "unknown preconditioner");
  TL_ARG(o->precond != TL_PRECOND_GMG || P->mg.ready, "tl_setup_mg must be called first");
  TL_ARG(o->rtol >= 0.0 && o->atol >= 0.0, "tolerances must be non-negative");
  PcgCtx c;
  TL_TRY(pcg_context(P, c));
  hipStream_t s = P->stream;
  TL_TRY(pcg_init(P, c, f, o->x0_given ? u : nullptr, u, o));
  // Iterations run in chunks between host checks of the device state; after
  // the first chunk the chunk length follows the observed residual decay so
  // that few iterations are launched past convergence (they exit at once but
  // still cost a launch each).
  int chunk = std::max(1, o->check_every);
  PcgState hs{};
  TL_HIP(hipMemcpyAsync(&hs, c.st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
  TL_HIP(hipStreamSynchronize(s));
  double rr_prev = hs.rr;
  int it_prev = 0;
  int64_t kk = 0;
  for (int launched = 0; !hs.done;) {
    for (int k = 0; k < chunk; ++k) TL_TRY(pcg_iteration(P, c, u, o->precond, kk++));
    launched += chunk;
    TL_HIP(hipMemcpyAsync(&hs, c.st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    TL_HIP(hipStreamSynchronize(s));
    if (hs.done) break;
    if (launched > o->maxiter + chunk) break;  // defensive: done must have been set
    if (hs.it > it_prev && hs.rr > 0.0 && hs.rr < rr_prev && hs.tol2 > 0.0) {
      const double rate = std::log(hs.rr / rr_prev) / double(hs.it - it_prev);
      const double rem = std::log(hs.tol2 / hs.rr) / rate;
      chunk = int(std::min(64.0, std::max(1.0, std::floor(0.9 * rem))));
    } else {
      chunk = std::max(1, o->check_every);
    }
    rr_prev = hs.rr;
    it_prev = hs.it;
  }
  double* scal = c.partials + 3 * kMaxRedBlocks;  // scratch slot for the total
  TL_TRY(pcg_true_residual(P, c, u, scal));
  double tr2 = 0.0;
  TL_HIP(hipMemcpyAsync(&tr2, scal, sizeof(double), hipMemcpyDeviceToHost, s));
  TL_HIP(hipMemcpyAsync(&hs, c.st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
  TL_HIP(hipStreamSynchronize(s));
  if (!pcg_info(hs, tr2, info)) {
    set_error("pcg iteration did not converge");
    return TL_ENOCONV;
  }
  return TL_OK;
}


// ----------------------------------------------------------------------------
// autotuner of the 2D operator's workgroup shape
// ----------------------------------------------------------------------------
namespace {
struct TuneKey {
  int nx, ny, ndpn;
  bool operator<(const TuneKey& o) const {
    return std::tie(nx, ny, ndpn) < std::tie(o.nx, o.ny, o.ndpn);
  }
};
struct TuneVal {
  int tile[2];
  double us[2][4];
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;
constexpr int kTiles[4] = {0, 16, 32, 64};
}  // namespace

int tl_set_tile2d(tl_problem* P, int32_t prec, int32_t ty) {
  TL_ARG(P, "null argument");
  TL_ARG(P->g.dim == 2, "the tile shape applies to 2D grids");
  TL_ARG(prec == TL_PREC_FP64 || prec == TL_PREC_FP32, "tile shapes exist for FP64 and FP32");
  TL_ARG(ty == 0 || ty == 16 || ty == 32 || ty == 64, "tile height must be 0, 16, 32 or 64");
  TL_ARG(ty == 0 || P->op.fast_ke, "tiles need the pattern-matched element matrix");
  P->op.tile2d[prec == TL_PREC_FP64 ? 0 : 1] = ty;
  return TL_OK;
}

int tl_autotune(tl_problem* P, int32_t reps, int32_t* tile_out, double* us_out) {
  TL_ARG(P, "null argument");
  TL_ARG(P->op.have_density, "tl_set_density must be called first");
  TL_ARG(reps >= 1 && reps <= 1000, "reps must be in [1, 1000]");
  if (P->g.dim != 2 || !P->op.fast_ke) {  // one shape only: nothing to choose
    if (tile_out) tile_out[0] = tile_out[1] = 0;
    if (us_out) for (int k = 0; k < 8; ++k) us_out[k] = 0.0;
    return TL_OK;
  }
  const TuneKey key{P->g.nx, P->g.ny, P->ndpn};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      P->op.tile2d[0] = it->second.tile[0];
      P->op.tile2d[1] = it->second.tile[1];
      if (tile_out) { tile_out[0] = it->second.tile[0]; tile_out[1] = it->second.tile[1]; }
      if (us_out) for (int k = 0; k < 8; ++k) us_out[k] = it->second.us[k / 4][k % 4];
      return TL_OK;
    }
  }
  const int64_t n = P->ndof;
  TL_TRY(P->t.alloc(sizeof(double) * n));
  TL_TRY(P->q.alloc(sizeof(double) * n));
  hipStream_t s = P->stream;
  TL_HIP(hipMemsetAsync(P->t.ptr, 0, sizeof(double) * n, s));
  hipEvent_t e0, e1;
  TL_HIP(hipEventCreate(&e0));
  TL_HIP(hipEventCreate(&e1));
  TuneVal tv{};
  const bool prof = P->prof_on;
  P->prof_on = false;
  int rc = TL_OK;
  for (int pi = 0; pi < 2 && rc == TL_OK; ++pi) {
    const int prec = pi == 0 ? TL_PREC_FP64 : TL_PREC_FP32;
    double best = 1e300;
    for (int ci = 0; ci < 4 && rc == TL_OK; ++ci) {
      P->op.tile2d[pi] = kTiles[ci];
      auto run = [&]() {
        return fine_apply_ex(P, P->t.ptr, P->q.ptr, prec, KU_STORE, nullptr, false, nullptr, 0.0,
                             nullptr, nullptr, nullptr, nullptr, FIN_STORE, nullptr, false);
      };
      if ((rc = run()) != TL_OK) break;  // warm-up
      if (hipEventRecord(e0, s) != hipSuccess) { rc = TL_EHIP; break; }
      for (int k = 0; k < reps && rc == TL_OK; ++k) rc = run();
      if (rc != TL_OK) break;
      float ms = 0.f;
      if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        set_error("autotune timing failed");
        rc = TL_EHIP;
        break;
      }
      tv.us[pi][ci] = 1e3 * ms / reps;
      if (tv.us[pi][ci] < best) { best = tv.us[pi][ci]; tv.tile[pi] = kTiles[ci]; }
    }
    P->op.tile2d[pi] = tv.tile[pi];
  }
  P->prof_on = prof;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != TL_OK) {
    P->op.tile2d[0] = P->op.tile2d[1] = 0;
    return rc;
  }
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = tv;
  }
  if (tile_out) { tile_out[0] = tv.tile[0]; tile_out[1] = tv.tile[1]; }
  if (us_out) for (int k = 0; k < 8; ++k) us_out[k] = tv.us[k / 4][k % 4];
  return TL_OK;
}

void tl_autotune_clear(void) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
}

// ----------------------------------------------------------------------------
// batched solves
// ----------------------------------------------------------------------------
namespace {
// Capture `chunk` PCG iterations of handle P (iterate in P->xb) on stream s
// into an executable graph, or reuse the one captured with the same key.
int chunk_graph(tl_problem* P, const PcgCtx& c, int precond, int chunk) {
  const int64_t key[5] = {int64_t(P->gen), P->op.tile2d[0], P->op.tile2d[1], chunk, precond};
  if (P->chunk_exec && std::equal(key, key + 5, P->chunk_key)) return TL_OK;
  if (P->chunk_exec) {
    (void)hipGraphExecDestroy(P->chunk_exec);
    P->chunk_exec = nullptr;
  }
  hipStream_t s = P->stream;
  TL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = TL_OK;
  for (int k = 0; k < chunk && rc == TL_OK; ++k)
    rc = pcg_iteration(P, c, P->xb.as<double>(), precond, k);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(s, &g);
  if (rc != TL_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  TL_HIP(ec);
  const hipError_t ei = hipGraphInstantiate(&P->chunk_exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  TL_HIP(ei);
  std::copy(key, key + 5, P->chunk_key);
  return TL_OK;
}

struct BatchSlot {
  PcgState hs;
  double tr2;
  double pad[7];
};

// ---------------------------------------------------------------------------
// Lockstep batched solves.  Every (instance, right-hand side) pair is a
// system with its own PCG and V-cycle vectors (carved from one arena); the
// hierarchy (stencils, masks, coarsest inverse) stays shared per instance.
// One PCG iteration of each system is recorded once per parity of the
// iteration counter (the search directions and the levels' iterate buffers
// alternate), the recordings are merged across systems, and every iteration
// replays the merged schedule: one launch per (place in the cycle, position,
// kernel) covers the blocks of all systems that run that kernel there.
// ---------------------------------------------------------------------------
constexpr int kSysBufs = 9;  // r z p p2 q t xb partials state
struct SysBufs {
  DevBuf b[kSysBufs];
  std::vector<DevBuf> lv;  // 4 per level: vb vx vr vt
};

DevBuf* sys_slot(tl_problem* P, int k) {
  DevBuf* slots[kSysBufs] = {&P->r, &P->z, &P->p, &P->p2, &P->q, &P->t, &P->xb, &P->partials,
                             &P->state};
  return slots[k];
}

DevBuf* lv_slot(tl_problem* P, int l, int k) {
  Level& L = P->mg.lv[l];
  DevBuf* slots[4] = {&L.vb, &L.vx, &L.vr, &L.vt};
  return slots[k];
}

// Exchange the handle's work vectors with the system's (an involution).
void swap_sys(tl_problem* P, SysBufs& S) {
  for (int k = 0; k < kSysBufs; ++k) std::swap(*sys_slot(P, k), S.b[k]);
  for (int l = 0; l < P->mg.nlevels; ++l)
    for (int k = 0; k < 4; ++k) std::swap(*lv_slot(P, l, k), S.lv[size_t(l) * 4 + k]);
}

// Byte sizes of a system's buffers on handle P (same as the handle's own).
void sys_sizes(tl_problem* P, std::vector<size_t>& sz) {
  sz.clear();
  const size_t v = sizeof(double) * size_t(P->ndof);
  for (int k = 0; k < 7; ++k) sz.push_back(v);
  sz.push_back(sizeof(double) * 4 * kMaxRedBlocks);
  sz.push_back(sizeof(PcgState));
  for (int l = 0; l < P->mg.nlevels; ++l)
    for (int k = 0; k < 4; ++k) sz.push_back(std::max<size_t>(lv_slot(P, l, k)->bytes, 8));
}

struct MergedLaunch {
  MultiLaunchFn fn;
  size_t lds = 0;
  size_t desc_off = 0, tile_off = 0;  // byte offsets into the upload
  unsigned nblocks = 0;
};

// Merge the recordings of the active systems into one launch list.  Places
// run down the cycle by level, up the cycle by level, then the PCG update; at
// each place the systems' ops run position by position, one launch per
// kernel at each position, so every system's ops keep their order.
void merge_schedule(const std::vector<const Recorder*>& recs, std::vector<MergedLaunch>& out,
                    std::vector<unsigned char>& blob) {
  out.clear();
  blob.clear();
  int maxl = 0;
  for (const Recorder* r : recs)
    for (const RecOp& op : r->ops) maxl = std::max(maxl, op.level);
  std::vector<std::pair<int, int>> places;
  for (int l = 0; l <= maxl; ++l) places.push_back({0, l});
  for (int l = maxl; l >= 0; --l) places.push_back({1, l});
  places.push_back({2, 0});
  auto align = [&](size_t a) { blob.resize((blob.size() + a - 1) / a * a); };
  for (const auto& pl : places) {
    std::vector<std::vector<const RecOp*>> seq(recs.size());
    size_t npos = 0;
    for (size_t i = 0; i < recs.size(); ++i) {
      for (const RecOp& op : recs[i]->ops)
        if (op.phase == pl.first && op.level == pl.second) seq[i].push_back(&op);
      npos = std::max(npos, seq[i].size());
    }
    for (size_t pos = 0; pos < npos; ++pos) {
      std::vector<MultiLaunchFn> fns;  // kernels at this position, first-seen order
      for (size_t i = 0; i < recs.size(); ++i)
        if (pos < seq[i].size() && std::find(fns.begin(), fns.end(), seq[i][pos]->fn) == fns.end())
          fns.push_back(seq[i][pos]->fn);
      for (MultiLaunchFn fn : fns) {
        std::vector<std::pair<size_t, const RecOp*>> members;
        for (size_t i = 0; i < recs.size(); ++i)
          if (pos < seq[i].size() && seq[i][pos]->fn == fn) members.push_back({i, seq[i][pos]});
        MergedLaunch m;
        m.fn = fn;
        align(64);
        m.desc_off = blob.size();
        std::vector<unsigned> nbs;  // blocks per argument struct
        const RecOp* first = members[0].second;
        if (first->pack) {
          // systems with the same share key in packs of up to kRecPack
          m.fn = first->pack_fn;
          std::vector<const void*> keys;
          for (const auto& mb : members)
            if (std::find(keys.begin(), keys.end(), mb.second->share) == keys.end())
              keys.push_back(mb.second->share);
          for (const void* key : keys) {
            std::vector<const RecOp*> grp;
            for (const auto& mb : members)
              if (mb.second->share == key) grp.push_back(mb.second);
            for (size_t g0 = 0; g0 < grp.size(); g0 += kRecPack) {
              const int cnt = int(std::min<size_t>(kRecPack, grp.size() - g0));
              nbs.push_back(first->pack(grp.data() + g0, cnt, blob));
            }
          }
        } else {
          for (const auto& mb : members) {
            m.lds = std::max(m.lds, mb.second->lds);
            blob.insert(blob.end(), mb.second->args.begin(), mb.second->args.end());
            nbs.push_back(mb.second->nb);
          }
        }
        align(64);
        m.tile_off = blob.size();
        for (size_t j = 0; j < nbs.size(); ++j) {
          for (unsigned b = 0; b < nbs[j]; ++b) {
            const uint2 t{unsigned(j), b};
            const auto* pt = reinterpret_cast<const unsigned char*>(&t);
            blob.insert(blob.end(), pt, pt + sizeof(uint2));
          }
          m.nblocks += nbs[j];
        }
        if (m.nblocks) out.push_back(m);
      }
    }
  }
}

__global__ void gather_done_kernel(PcgState* const* __restrict__ st, int n, int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = st[i]->done;
}

// Whether the batch can run in lockstep: 2D handles, GMG V(1,1) cycles
// whose first fine sweep is fused into the PCG update (the launches that
// have multi-instance forms).
bool lockstep_eligible(int n, tl_problem* const* hs, int precond) {
  if (precond != TL_PRECOND_GMG) return false;
  for (int i = 0; i < n; ++i) {
    const tl_problem* P = hs[i];
    if (P->g.dim != 2 || !pcg_fuses_x0(P, precond) || P->mg.gamma > 1 ||
        P->mg.nlevels > 12)
      return false;
  }
  return true;
}

int batch_lockstep(int32_t n, tl_problem* const* hs, const double* const* f, double* const* u,
                   const int32_t* nrhs, const tl_solve_opts* o, int chunk, bool graphs,
                   tl_solve_info* infos) {
  // systems, their buffer layout in one arena
  struct Sys { int inst, rhs; SysBufs buf; PcgCtx ctx; Recorder rec[2]; };
  std::vector<Sys> sys;
  size_t total = 0;
  std::vector<std::vector<size_t>> sizes(n);
  for (int i = 0; i < n; ++i) {
    sys_sizes(hs[i], sizes[i]);
    for (int j = 0; j < nrhs[i]; ++j) {
      sys.push_back(Sys{i, j, {}, {}, {}});
      for (size_t b : sizes[i]) total += (b + 255) / 256 * 256;
    }
  }
  DevBuf arena;
  TL_TRY(arena.alloc(total));
  struct ArenaGuard { DevBuf& a; ~ArenaGuard() { a.release(); } } guard{arena};
  {
    size_t off = 0;
    for (Sys& sy : sys) {
      const auto& sz = sizes[sy.inst];
      auto carve = [&](DevBuf& d, size_t b) {
        d.ptr = static_cast<unsigned char*>(arena.ptr) + off;
        d.bytes = b;
        off += (b + 255) / 256 * 256;
      };
      for (int k = 0; k < kSysBufs; ++k) carve(sy.buf.b[k], sz[k]);
      sy.buf.lv.resize(sz.size() - kSysBufs);
      for (size_t k = 0; k < sy.buf.lv.size(); ++k) carve(sy.buf.lv[k], sz[kSysBufs + k]);
    }
  }
  // the batch runs on a stream of its own (a graph cannot be captured on the
  // null stream), after the work already queued on the first handle's
  // stream; it is synchronised before this function returns
  hipStream_t s = nullptr;
  TL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamDestroy(s); } } sguard0{s};
  {
    hipEvent_t e = nullptr;
    TL_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const hipError_t er = hipEventRecord(e, hs[0]->stream);
    const hipError_t ew = er == hipSuccess ? hipStreamWaitEvent(s, e, 0) : er;
    (void)hipEventDestroy(e);
    TL_HIP(ew);
  }
  // initialise every system and record its two iteration parities
  for (Sys& sy : sys) {
    tl_problem* P = hs[sy.inst];
    const int64_t nd = P->ndof;
    swap_sys(P, sy.buf);
    const hipStream_t saved = P->stream;
    P->stream = s;
    int rc = pcg_context(P, sy.ctx);
    const double* x0 = o->x0_given ? u[sy.inst] + int64_t(sy.rhs) * nd : nullptr;
    if (rc == TL_OK) rc = pcg_init(P, sy.ctx, f[sy.inst] + int64_t(sy.rhs) * nd, x0,
                                   P->xb.as<double>(), o);
    for (int k = 0; k < 2 && rc == TL_OK; ++k) {
      P->rec = &sy.rec[k];
      rc = pcg_iteration(P, sy.ctx, P->xb.as<double>(), o->precond, k);
      P->rec = nullptr;
    }
    P->stream = saved;
    swap_sys(P, sy.buf);
    TL_TRY(rc);
    TL_ARG(!sy.rec[0].unsupported && !sy.rec[1].unsupported,
           "internal: a lockstep system reached a launch without a multi-instance form");
  }
  // state pointers for the convergence checks
  const int ns = int(sys.size());
  DevBuf stp, dflags;
  struct Bufs { DevBuf& a; DevBuf& b; ~Bufs() { a.release(); b.release(); } } bufs{stp, dflags};
  TL_TRY(stp.alloc(sizeof(PcgState*) * ns));
  TL_TRY(dflags.alloc(sizeof(int) * ns));
  std::vector<PcgState*> sts(ns);
  for (int i = 0; i < ns; ++i) sts[i] = sys[i].ctx.st;
  TL_HIP(hipMemcpyAsync(stp.ptr, sts.data(), sizeof(PcgState*) * ns, hipMemcpyHostToDevice, s));
  int* hflags = nullptr;
  TL_HIP(hipHostMalloc(reinterpret_cast<void**>(&hflags), sizeof(int) * ns, hipHostMallocDefault));
  struct HostGuard { int* p; ~HostGuard() { (void)hipHostFree(p); } } hguard{hflags};
  // schedules (rebuilt without the converged systems as they drop out)
  std::vector<MergedLaunch> sched[2];
  std::vector<unsigned char> blob[2];
  std::vector<DevBuf> uploads;  // kept until the end: launches in flight read them
  struct UpGuard { std::vector<DevBuf>& v; ~UpGuard() { for (auto& d : v) d.release(); } } ug{uploads};
  std::vector<char> active(ns, 1);
  const unsigned char* dev_blob[2] = {nullptr, nullptr};
  auto build = [&]() -> int {
    for (int par = 0; par < 2; ++par) {
      std::vector<const Recorder*> recs;
      for (int i = 0; i < ns; ++i)
        if (active[i]) recs.push_back(&sys[i].rec[par]);
      merge_schedule(recs, sched[par], blob[par]);
      DevBuf d;
      TL_TRY(d.alloc(std::max<size_t>(blob[par].size(), 64)));
      uploads.push_back(d);
      // blob[par] stays unchanged until the stream is synchronised below
      TL_HIP(hipMemcpyAsync(d.ptr, blob[par].data(), blob[par].size(), hipMemcpyHostToDevice, s));
      dev_blob[par] = static_cast<const unsigned char*>(d.ptr);
    }
    return TL_OK;
  };
  TL_TRY(build());
  // one chunk of lockstep iterations; chunks start on an even iteration, so
  // every chunk replays the same launches until the schedule is rebuilt:
  // with graphs it is captured once per schedule and replayed
  auto launch_chunk = [&]() -> int {
    for (int c = 0; c < chunk; ++c) {
      const int par = c & 1;
      for (const MergedLaunch& m : sched[par])
        TL_TRY(m.fn(s, dev_blob[par] + m.desc_off,
                    reinterpret_cast<const uint2*>(dev_blob[par] + m.tile_off), m.nblocks, m.lds));
    }
    return TL_OK;
  };
  hipGraphExec_t gexec = nullptr;
  struct ExecGuard { hipGraphExec_t& g; ~ExecGuard() { if (g) (void)hipGraphExecDestroy(g); } } eg{gexec};
  int nactive = ns, built_for = ns;
  int64_t k = 0;
  for (;;) {
    if (graphs) {
      if (!gexec) {
        TL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = launch_chunk();
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(s, &g);
        if (rc != TL_OK) {
          if (g) (void)hipGraphDestroy(g);
          return rc;
        }
        TL_HIP(ec);
        const hipError_t ei = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        TL_HIP(ei);
      }
      TL_HIP(hipGraphLaunch(gexec, s));
    } else {
      TL_TRY(launch_chunk());
    }
    k += chunk;
    hipLaunchKernelGGL(gather_done_kernel, dim3(unsigned((ns + 255) / 256)), dim3(256), 0, s,
                       stp.as<PcgState*>(), ns, dflags.as<int>());
    TL_CHECK_LAUNCH();
    TL_HIP(hipMemcpyAsync(hflags, dflags.ptr, sizeof(int) * ns, hipMemcpyDeviceToHost, s));
    TL_HIP(hipStreamSynchronize(s));
    nactive = 0;
    for (int i = 0; i < ns; ++i) {
      active[i] = !hflags[i];
      nactive += active[i];
    }
    if (nactive == 0 || k > int64_t(o->maxiter) + chunk) break;
    // drop converged systems from the launches once a quarter of them is done
    if (nactive <= built_for * 3 / 4) {
      TL_TRY(build());
      built_for = nactive;
      if (gexec) {
        (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
      }
    }
  }
  // true residuals and results
  BatchSlot* slots = nullptr;
  TL_HIP(hipHostMalloc(reinterpret_cast<void**>(&slots), sizeof(BatchSlot) * ns,
                       hipHostMallocDefault));
  struct SlotGuard { BatchSlot* p; ~SlotGuard() { (void)hipHostFree(p); } } sguard{slots};
  for (int i = 0; i < ns; ++i) {
    Sys& sy = sys[i];
    tl_problem* P = hs[sy.inst];
    swap_sys(P, sy.buf);
    const hipStream_t saved = P->stream;
    P->stream = s;
    double* scal = sy.ctx.partials + 3 * kMaxRedBlocks;
    int rc = pcg_true_residual(P, sy.ctx, P->xb.as<double>(), scal);
    if (rc == TL_OK &&
        (hipMemcpyAsync(u[sy.inst] + int64_t(sy.rhs) * P->ndof, P->xb.ptr,
                        sizeof(double) * P->ndof, hipMemcpyDeviceToDevice, s) != hipSuccess ||
         hipMemcpyAsync(&slots[i].tr2, scal, sizeof(double), hipMemcpyDeviceToHost, s) !=
             hipSuccess ||
         hipMemcpyAsync(&slots[i].hs, sy.ctx.st, sizeof(PcgState), hipMemcpyDeviceToHost, s) !=
             hipSuccess)) {
      set_error("lockstep batch: result copies failed");
      rc = TL_EHIP;
    }
    P->stream = saved;
    swap_sys(P, sy.buf);
    TL_TRY(rc);
  }
  TL_HIP(hipStreamSynchronize(s));
  bool all_conv = true;
  std::vector<int64_t> info_at(n, 0);
  for (int i = 1; i < n; ++i) info_at[i] = info_at[i - 1] + nrhs[i - 1];
  for (int i = 0; i < ns; ++i) {
    tl_solve_info* inf = infos ? infos + info_at[sys[i].inst] + sys[i].rhs : nullptr;
    if (!pcg_info(slots[i].hs, slots[i].tr2, inf)) all_conv = false;
  }
  if (!all_conv) {
    set_error("pcg iteration did not converge for at least one right-hand side");
    return TL_ENOCONV;
  }
  return TL_OK;
}
}  // namespace

int tl_batch_solve(int32_t n, tl_problem* const* hs, const double* const* f, double* const* u,
                   const int32_t* nrhs, const tl_solve_opts* o, const tl_batch_opts* bo,
                   tl_solve_info* infos) {
  TL_ARG(n >= 1 && hs && f && u && nrhs && o, "null argument or empty batch");
  TL_ARG(o->precond >= TL_PRECOND_NONE && o->precond <= TL_PRECOND_GMG, "unknown preconditioner");
  TL_ARG(o->rtol >= 0.0 && o->atol >= 0.0, "tolerances must be non-negative");
  const int nstreams = std::max(1, std::min<int>(n, bo && bo->nstreams > 0 ? bo->nstreams : 4));
  int chunk = bo && bo->chunk > 0 ? bo->chunk : 8;
  chunk += chunk & 1;  // even: each chunk starts on the same search-direction buffer
  TL_ARG(chunk <= 256, "chunk must be <= 256");
  const bool graphs = !bo || bo->use_graphs;
  for (int i = 0; i < n; ++i) {
    TL_ARG(hs[i] && f[i] && u[i], "null handle or vector");
    TL_ARG(nrhs[i] >= 1, "every instance needs at least one right-hand side");
    TL_ARG(hs[i]->op.have_density, "tl_set_density must be called first");
    TL_ARG(o->precond != TL_PRECOND_GMG || hs[i]->mg.ready, "tl_setup_mg must be called first");
    for (int j = 0; j < i; ++j) TL_ARG(hs[i] != hs[j], "a handle appears twice in the batch");
  }
  if (bo && bo->lockstep && lockstep_eligible(n, hs, o->precond)) {
    // the systems start after the work already queued on each handle's stream
    for (int i = 1; i < n; ++i)
      if (hs[i]->stream != hs[0]->stream) {
        hipEvent_t e = nullptr;
        TL_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        const hipError_t er = hipEventRecord(e, hs[i]->stream);
        const hipError_t ew = er == hipSuccess ? hipStreamWaitEvent(hs[0]->stream, e, 0) : er;
        (void)hipEventDestroy(e);
        TL_HIP(ew);
      }
    std::vector<bool> prof(n);
    for (int i = 0; i < n; ++i) { prof[i] = hs[i]->prof_on; hs[i]->prof_on = false; }
    const int rc = batch_lockstep(n, hs, f, u, nrhs, o, chunk, graphs, infos);
    for (int i = 0; i < n; ++i) hs[i]->prof_on = prof[i];
    // later work on the other handles' streams follows the batch
    for (int i = 1; i < n; ++i)
      if (hs[i]->stream != hs[0]->stream) (void)hipStreamSynchronize(hs[0]->stream);
    return rc;
  }
  std::vector<hipStream_t> streams(nstreams, nullptr);
  std::vector<hipEvent_t> ev(n, nullptr);
  std::vector<hipStream_t> saved(n);
  std::vector<bool> saved_prof(n);
  std::vector<PcgCtx> ctx(n);
  BatchSlot* slots = nullptr;
  int rc = TL_OK;
  auto fail = [&](int code) { if (rc == TL_OK) rc = code; };
  for (int i = 0; i < n; ++i) {
    saved[i] = hs[i]->stream;
    saved_prof[i] = hs[i]->prof_on;
  }
  if (hipHostMalloc(reinterpret_cast<void**>(&slots), sizeof(BatchSlot) * n, hipHostMallocDefault) !=
      hipSuccess) {
    set_error("hipHostMalloc of the batch state slots failed");
    return TL_EHIP;
  }
  for (int k = 0; k < nstreams && rc == TL_OK; ++k)
    if (hipStreamCreateWithFlags(&streams[k], hipStreamNonBlocking) != hipSuccess) fail(TL_EHIP);
  for (int i = 0; i < n && rc == TL_OK; ++i)
    if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) fail(TL_EHIP);
  if (rc != TL_OK) set_error("stream or event creation failed");
  // per instance: r index, iterations launched, phase (0 iterate, 1 finish, 2 done)
  std::vector<int> rhs(n, 0), launched(n, 0), phase(n, 0);
  std::vector<int64_t> info_at(n, 0);
  for (int i = 1; i < n; ++i) info_at[i] = info_at[i - 1] + nrhs[i - 1];
  auto launch_chunk = [&](int i) -> int {
    tl_problem* P = hs[i];
    if (graphs) {
      TL_TRY(chunk_graph(P, ctx[i], o->precond, chunk));
      TL_HIP(hipGraphLaunch(P->chunk_exec, P->stream));
    } else {
      for (int k = 0; k < chunk; ++k) TL_TRY(pcg_iteration(P, ctx[i], P->xb.as<double>(), o->precond, k));
    }
    launched[i] += chunk;
    TL_HIP(hipMemcpyAsync(&slots[i].hs, ctx[i].st, sizeof(PcgState), hipMemcpyDeviceToHost, P->stream));
    TL_HIP(hipEventRecord(ev[i], P->stream));
    return TL_OK;
  };
  auto start_rhs = [&](int i) -> int {
    tl_problem* P = hs[i];
    const int64_t nd = P->ndof;
    const double* x0 = o->x0_given ? u[i] + int64_t(rhs[i]) * nd : nullptr;
    TL_TRY(pcg_init(P, ctx[i], f[i] + int64_t(rhs[i]) * nd, x0, P->xb.as<double>(), o));
    launched[i] = 0;
    phase[i] = 0;
    return launch_chunk(i);
  };
  auto finish_rhs = [&](int i) -> int {
    tl_problem* P = hs[i];
    double* scal = ctx[i].partials + 3 * kMaxRedBlocks;
    TL_TRY(pcg_true_residual(P, ctx[i], P->xb.as<double>(), scal));
    TL_HIP(hipMemcpyAsync(u[i] + int64_t(rhs[i]) * P->ndof, P->xb.ptr, sizeof(double) * P->ndof,
                          hipMemcpyDeviceToDevice, P->stream));
    TL_HIP(hipMemcpyAsync(&slots[i].tr2, scal, sizeof(double), hipMemcpyDeviceToHost, P->stream));
    TL_HIP(hipMemcpyAsync(&slots[i].hs, ctx[i].st, sizeof(PcgState), hipMemcpyDeviceToHost,
                          P->stream));
    TL_HIP(hipEventRecord(ev[i], P->stream));
    phase[i] = 1;
    return TL_OK;
  };
  bool all_conv = true;
  // batch streams start after the work already queued on each handle's stream
  for (int i = 0; i < n && rc == TL_OK; ++i)
    if (hipEventRecord(ev[i], saved[i]) != hipSuccess ||
        hipStreamWaitEvent(streams[i % nstreams], ev[i], 0) != hipSuccess) {
      set_error("ordering the batch streams after the handles' streams failed");
      fail(TL_EHIP);
    }
  if (rc == TL_OK) {
    for (int i = 0; i < n && rc == TL_OK; ++i) {
      tl_problem* P = hs[i];
      P->stream = streams[i % nstreams];
      P->prof_on = false;  // no event records inside captured chunks
      int e = pcg_context(P, ctx[i]);
      if (e == TL_OK) e = P->xb.alloc(sizeof(double) * P->ndof);
      if (e == TL_OK) e = start_rhs(i);
      fail(e);
    }
    int remaining = n;
    while (rc == TL_OK && remaining > 0) {
      for (int i = 0; i < n && rc == TL_OK; ++i) {
        if (phase[i] == 2) continue;
        const hipError_t q = hipEventQuery(ev[i]);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) {
          set_error(std::string("batch event: ") + hipGetErrorString(q));
          fail(TL_EHIP);
          break;
        }
        if (phase[i] == 0) {
          const PcgState& st = slots[i].hs;
          if (st.done || launched[i] > o->maxiter + chunk) fail(finish_rhs(i));
          else fail(launch_chunk(i));
        } else {
          tl_solve_info* inf = infos ? infos + info_at[i] + rhs[i] : nullptr;
          if (!pcg_info(slots[i].hs, slots[i].tr2, inf)) all_conv = false;
          if (++rhs[i] < nrhs[i]) {
            fail(start_rhs(i));
          } else {
            phase[i] = 2;
            --remaining;
          }
        }
      }
    }
  }
  for (int k = 0; k < nstreams; ++k)
    if (streams[k]) (void)hipStreamSynchronize(streams[k]);
  for (int i = 0; i < n; ++i) {
    hs[i]->stream = saved[i];
    hs[i]->prof_on = saved_prof[i];
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  }
  for (int k = 0; k < nstreams; ++k)
    if (streams[k]) (void)hipStreamDestroy(streams[k]);
  (void)hipHostFree(slots);
  if (rc != TL_OK) return rc;
  if (!all_conv) {
    set_error("pcg iteration did not converge for at least one right-hand side");
    return TL_ENOCONV;
  }
  return TL_OK;
}

}  // extern "C"
