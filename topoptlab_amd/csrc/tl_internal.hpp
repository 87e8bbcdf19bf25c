// Internal declarations shared by the topoptlab_amd HIP sources (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/topoptlab_amd.h"

namespace tl {

// ----------------------------------------------------------------------------
// error handling
// ----------------------------------------------------------------------------
void set_error(const std::string& msg);

struct Status {
  int code = TL_OK;
};

#define TL_HIP(expr)                                                         \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) {                                                  \
      ::tl::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));   \
      return TL_EHIP;                                                        \
    }                                                                        \
  } while (0)

#define TL_CHECK_LAUNCH() TL_HIP(hipGetLastError())

#define TL_ARG(cond, msg)              \
  do {                                 \
    if (!(cond)) {                     \
      ::tl::set_error(msg);            \
      return TL_EINVAL;                \
    }                                  \
  } while (0)

#define TL_TRY(expr)               \
  do {                             \
    int _rc = (expr);              \
    if (_rc != TL_OK) return _rc;  \
  } while (0)

// ----------------------------------------------------------------------------
// device buffers owned by a handle
// ----------------------------------------------------------------------------
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  int alloc(size_t nbytes);
  void release();
  template <typename T>
  T* as() const { return static_cast<T*>(ptr); }
};

// ----------------------------------------------------------------------------
// block reductions with a deterministic last-block finish
// ----------------------------------------------------------------------------
constexpr int kMaxRedBlocks = 4096;  // partial slots per reduction value

// Sum of v over the 64 lanes of the calling wavefront (fixed butterfly order).
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Block-wide sum of NV values per thread; result valid in every thread.
// sh must hold NV * (blockDim.x/64) doubles.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k * nw + wid] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += sh[k * nw + w];
    v[k] = s;
  }
  __syncthreads();
}

// Each block stores its NV block totals with plain stores; a separate
// one-block kernel folds them (kernel boundaries order the stores).  No
// per-block device-scope fence or atomic is used: on a multi-XCD GPU a
// device-scope release writes back the XCD's L2, which made the former
// "last block folds" scheme cost ~50 ns per block.
template <int NV>
__device__ inline void store_partials_at(const double (&part)[NV], double* partials,
                                         unsigned bid) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[k * kMaxRedBlocks + bid] = part[k];
  }
}

template <int NV>
__device__ inline void store_partials(const double (&part)[NV], double* partials) {
  store_partials_at<NV>(part, partials,
                        blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
}

// Fold nb partials per value in a fixed order (strided per thread, then the
// fixed block tree); every thread receives the totals.  sh: NV*(bs/64) doubles.
template <int NV>
__device__ inline void fold_partials(const double* partials, unsigned nb, double (&tot)[NV],
                                     double* sh) {
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double acc = 0.0;
    for (unsigned b = threadIdx.x; b < nb; b += blockDim.x) acc += partials[k * kMaxRedBlocks + b];
    tot[k] = acc;
  }
  block_sum<NV>(tot, sh);
}

// ----------------------------------------------------------------------------
// structured-grid geometry
// ----------------------------------------------------------------------------
// Local node k of a Q4/H8 element (reference order, bottom layer first):
// k = 4*dz + {(0,1)->0, (1,1)->1, (1,0)->2, (0,0)->3}[(dx,dy)], with (dx,dy,dz)
// the node's offset from the element's (ix,iy,iz) corner node.
__host__ __device__ constexpr int loc_dx(int k) { return ((k & 3) == 1 || (k & 3) == 2) ? 1 : 0; }
__host__ __device__ constexpr int loc_dy(int k) { return ((k & 3) <= 1) ? 1 : 0; }
__host__ __device__ constexpr int loc_dz(int k) { return k >> 2; }
__host__ __device__ constexpr int loc_of(int dx, int dy, int dz) {
  return 4 * dz + (dy == 1 ? dx : 3 - dx);
}

struct Grid {
  int dim = 3;
  int nelx = 0, nely = 0, nelz = 0;  // nelz = 0 for 2D
  int nx = 1, ny = 1, nz = 1;        // node counts
  int64_t nnodes = 0, nel = 0;
  void set(int d, int ex, int ey, int ez) {
    dim = d; nelx = ex; nely = ey; nelz = (d == 3) ? ez : 0;
    nx = ex + 1; ny = ey + 1; nz = (d == 3) ? ez + 1 : 1;
    nnodes = int64_t(nx) * ny * nz;
    nel = int64_t(ex) * ey * (d == 3 ? ez : 1);
  }
  int64_t npad() const {  // padded element array (zero border)
    return int64_t(nelx + 2) * (nely + 2) * (dim == 3 ? nelz + 2 : 1);
  }
};

// ----------------------------------------------------------------------------
// problem handle
// ----------------------------------------------------------------------------
struct PcgState {
  double rho, rho_prev, pq, rr, alpha, beta, tol2, bb;
  int it, maxiter, done, pad;
};

struct Level {
  void release();
  Grid g;
  int64_t ndof = 0;
  DevBuf mask;            // uint8 per node, bit d = DOF d fixed
  DevBuf st64;            // coarse stencil [ns*ndpn*ndpn][nnodes] (fp64, all blocks)
  DevBuf sth64, st32;     // forward half [(ns/2+1)*ndpn*ndpn][nnodes] (fp64, fp32) for the V-cycle
  DevBuf dinv64, dinv32;  // 1/diag (0 on fixed DOFs)
  DevBuf em;              // element matrices (levels >= 1, fp64)
  DevBuf vb, vx, vr, vt;  // V-cycle work vectors (V-cycle precision)
};

struct MgState {
  bool ready = false;
  int nlevels = 0, nu1 = 1, nu2 = 1, prec = TL_PREC_FP32;
  int gamma = 1, gamma_from = 1;  // cycle index on levels >= gamma_from
  double omega = 0.6;
  std::vector<Level> lv;  // lv[0] is the fine grid (no stencil)
  DevBuf G1;              // 8 x (nd*nd) fp64: child Galerkin matrices
  DevBuf W;               // 8 x 8 fp64: child interpolation weights
  DevBuf cinv;            // dense inverse of the coarsest operator (fp64)
  DevBuf gjw;             // Gauss-Jordan workspace: row panels and pivot inverses of two steps
  int64_t cn = 0, cnp = 0;  // coarsest ndof, padded size
};

struct FineOp {
  DevBuf KE64, KE32, KE16;     // element matrix (fp64, fp32, reserved)
  DevBuf sp64, sp32;           // padded SIMP scale
  DevBuf dinv64, dinv32;       // 1/diag(K_ff)
  DevBuf mask;                 // uint8 per node
  bool have_density = false;
  bool fast_ke = false;        // KE matches its unit-cell pattern
  double kv[8] = {};           // group values of the pattern
  // 2D operator workgroup shape per precision [fp64, fp32]: 0 = node line
  // (BS consecutive nodes), TY = 16/32/64: (BS/TY) x TY node tile
  int tile2d[2] = {0, 0};
  // FP64 3D elasticity operator: true = MFMA column walk, false = vector FMAs
  bool mfma64 = false;
};

}  // namespace tl

namespace tl {
struct ProfEvent {
  hipEvent_t start, stop;
  int kind;
};
}  // namespace tl

namespace tl {
// Launch recording for lockstep batched solves.  While a handle's `rec` is
// set, the launch helpers of the PCG iteration and the V-cycle do not launch:
// they append the kernel's multi-instance launcher, its argument struct, its
// block count and dynamic LDS, tagged with the op's place in the cycle
// (phase: 0 down the cycle at `level`, 1 up the cycle at `level`, 2 the PCG
// update after the preconditioner).  tl_batch_solve merges the recordings of
// many systems and replays each (place, position, launcher) group as one
// launch over all systems' blocks: multi-instance kernels read their
// system's argument struct and local block index from a tile table.
using MultiLaunchFn = int (*)(hipStream_t s, const void* descs, const uint2* tiles,
                              unsigned nblocks, size_t lds);
struct RecOp;
// Packs up to kRecPack ops with the same `share` key (e.g. systems reading
// one matrix) into one argument struct appended to out; returns its blocks.
constexpr int kRecPack = 4;
using PackFn = unsigned (*)(const RecOp* const* ops, int nops, std::vector<unsigned char>& out);
struct RecOp {
  int phase = 0, level = 0;
  MultiLaunchFn fn = nullptr;
  unsigned nb = 0;
  size_t lds = 0;
  std::vector<unsigned char> args;
  const void* share = nullptr;     // ops with the same key may be packed
  PackFn pack = nullptr;           // how (nullptr: never packed)
  MultiLaunchFn pack_fn = nullptr;  // launcher of the packed argument structs
};
struct Recorder {
  std::vector<RecOp> ops;
  bool unsupported = false;  // a launch site without a multi-instance form was reached
  template <class A>
  void add(int phase, int level, MultiLaunchFn fn, const A& a, unsigned nb, size_t lds) {
    RecOp op;
    op.phase = phase; op.level = level; op.fn = fn; op.nb = nb; op.lds = lds;
    op.args.resize(sizeof(A));
    std::memcpy(op.args.data(), &a, sizeof(A));
    ops.push_back(std::move(op));
  }
};
}  // namespace tl

struct tl_problem {
  tl::Recorder* rec = nullptr;  // lockstep batch recording (see Recorder)
  int rec_phase = 0, rec_level = 0;
  tl::Grid g;
  int ndpn = 1, nd = 8;   // DOFs per node, per element
  int64_t ndof = 0;
  double eps = 1e-9, penal = 3.0;
  hipStream_t stream = nullptr;
  std::vector<double> KEh;
  std::vector<int64_t> fixed_h;
  tl::FineOp op;
  tl::MgState mg;
  // PCG workspace (fp64, ndof each) and reduction scratch
  tl::DevBuf r, z, p, p2, q, t;
  tl::DevBuf xb;  // batched solves: iterate buffer captured by the chunk graph
  // batched solves: captured chunk of PCG iterations, reused while its key
  // (hierarchy generation, tile shapes, chunk length, preconditioner) holds
  uint64_t gen = 0;  // bumped by every tl_setup_mg
  hipGraphExec_t chunk_exec = nullptr;
  int64_t chunk_key[5] = {-1, -1, -1, -1, -1};
  tl::DevBuf partials, state;
  // launch timing of the fine operator (tl_profile)
  bool prof_on = false;
  std::vector<tl::ProfEvent> prof_used, prof_free;
  ~tl_problem();
};

namespace tl {
// operator entry points shared between translation units
int fine_apply_ex(tl_problem* P, const void* u, void* y, int prec, int mode, const void* b,
                  bool b64, const void* dinv, double omega, const void* zfuse_pin,
                  void* pout, const double* beta, double* dot_out, int fin, const int* done,
                  bool out64);
int fine_diag(tl_problem* P);
// launch timing (tl_profile): no-ops while profiling is off
int prof_begin(tl_problem* P, int kind, ProfEvent& ev);
int prof_end(tl_problem* P, ProfEvent& ev);
// x0_ready: the fine level's first (zero-guess) smoothing sweep is already
// in lv[0].vx (written by the PCG's residual update, see pcg_fuses_x0)
int mg_vcycle(tl_problem* P, const double* r, double* z, const int* done, bool x0_ready);
// V-cycle fine operator on MFMA (3D H8 elasticity, prec TL_PREC_BF16/FP16)
int fine_apply_mfma(tl_problem* P, const void* u, void* y, int prec, int mode, const void* b,
                    const void* dinv, double omega, double* partials, const int* done,
                    unsigned* nblocks_out);
// FP64 fine operator on MFMA (3D H8 elasticity)
int fine_apply_mfma64(tl_problem* P, const double* u, double* y, int mode, const double* b,
                      const double* dinv, double omega, const double* pin, double* pout,
                      const double* beta, double* partials, const int* done,
                      unsigned* nblocks_out);
}  // namespace tl

// First damped-Jacobi sweep from a zero guess, x = omega * dinv * b, in the
// V-cycle's precision T (the fine level's b is the fp64 PCG residual).  One
// definition for every kernel that forms it, so their results agree bit for bit.
template <typename T, typename TB>
__device__ inline T zero_guess_sweep(T omega, T dinv, TB b) {
  return omega * dinv * T(b);
}

// modes of the fine matrix-free operator epilogue
enum {
  KU_STORE = 0,       // y = K u
  KU_STORE_DOT = 1,   // y = K u ; dot += u.y  (pcg: p = z + beta p fused)
  KU_RESID = 2,       // y = b - K u
  KU_JACOBI = 3,      // y = u + omega*dinv*(b - K u)
  KU_JACOBI_DOT = 4,  // as KU_JACOBI, y written fp64, dot += b.y
};
