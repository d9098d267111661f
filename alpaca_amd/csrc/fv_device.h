// Device-side numerics of ALPACA's single-phase finite-volume right-hand side
// for gfx950.
//
// The arithmetic (fv_numerics.inl) and the fused stage kernel
// (stage_kernel.inl) are compiled twice, in two namespaces:
//   alpaca::exact  a*b+c never fused: reproduces the reference's x86 build
//                  (no FMA) bit for bit; used with R = double;
//   alpaca::fast   contraction to FMA allowed; used with R = double (FP64
//                  with FMA) and with R = float (the mixed form: FP32
//                  reconstruction and Riemann fluxes with hardware
//                  reciprocal/square root, FP64 pressure recovery, FP64 RK
//                  accumulation).
// Integer layout helpers shared by both live here.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/alpaca_gpu.h"

namespace alpaca {

// Largest block count (and link count) an entry point accepts: every grid
// size computed from it stays inside 32-bit launch dimensions (3D: 2^24
// blocks are 9 TB of haloed state, far beyond one GPU's HBM).
constexpr long long kMaxBlocks = 1LL << 24;

constexpr int kIC = 16;  // internal cells per block side (compile_time_constants.h:39)
constexpr int kHS = 4;   // halo width (compile_time_constants.h:41)
constexpr int kTC = 24;  // cells per side with halo

enum : int { kMass = 0, kEnergy = 1, kMomX = 2 };

namespace {
// Per-axis direction of each BoundaryLocation, in the enum's order
// (boundary_specifications.h:56-86): +1 East/North/Top, -1 West/South/Bottom,
// 0 along an axis the location does not name.
__constant__ signed char kLocationDir[26][3] = {
    {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},    {0, 0, -1},  {0, 1, -1},
    {0, -1, -1}, {0, 1, 1},   {0, -1, 1},  {1, 0, -1},  {-1, 0, -1},  {1, 0, 1},   {-1, 0, 1},
    {1, 1, 0},   {-1, 1, 0},  {1, -1, 0},  {-1, -1, 0}, {1, 1, 1},    {1, 1, -1},  {1, -1, 1},
    {1, -1, -1}, {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
}  // namespace

template <int DIM> struct Dims {
  static constexpr int NE = DIM + 2;             // equations
  static constexpr int TZ = DIM == 3 ? kTC : 1;  // halo cells along z
  static constexpr int ICZ = DIM == 3 ? kIC : 1;
  static constexpr int SX = kTC * TZ;  // element stride of i
  static constexpr int SY = TZ;        // element stride of j
  static constexpr int SZ = 1;         // element stride of k
  static constexpr int FIELD = kTC * kTC * TZ;   // elements per field
  static constexpr int BLOCK = NE * FIELD;       // elements per haloed block
  static constexpr int CELLS = kIC * kIC * ICZ;  // internal cells per block
};

template <int DIM, int DIR> __host__ __device__ constexpr int axis_stride() {
  return DIR == 0 ? Dims<DIM>::SX : (DIR == 1 ? Dims<DIM>::SY : Dims<DIM>::SZ);
}

// Minor directions of each principal direction (eigendecomposition.h:97-130).
template <int DIR> struct Minor;
template <> struct Minor<0> { static constexpr int m1 = 1, m2 = 2; };
template <> struct Minor<1> { static constexpr int m1 = 0, m2 = 2; };
template <> struct Minor<2> { static constexpr int m1 = 0, m2 = 1; };

// Projection order l of the characteristic decomposition per direction
// (state_reconstruction.h:37-57).
template <int DIM, int DIR> __device__ __forceinline__ constexpr int seq(int q) {
  if constexpr (DIM == 2) {
    constexpr int s[4] = {2, 3, 0, 1};
    return s[q];
  } else if constexpr (DIR == 0) {
    constexpr int s[5] = {3, 4, 2, 0, 1};
    return s[q];
  } else if constexpr (DIR == 1) {
    constexpr int s[5] = {4, 2, 3, 0, 1};
    return s[q];
  } else {
    constexpr int s[5] = {2, 3, 4, 0, 1};
    return s[q];
  }
}

constexpr int round_up64(int n) { return (n + 63) / 64 * 64; }

// Reads of one haloed block through a raw buffer resource: the block's base
// address lives in scalar registers, a thread passes one byte offset in a
// VGPR and the field/stencil part as a scalar offset, so a read costs no
// 64-bit address arithmetic.  The block pointer must be uniform over the
// wavefront.  num_records is the block's size in bytes; every offset the
// kernels form stays inside it, and voffset and soffset are never negative.
template <int DIM> struct BlockReader {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ explicit BlockReader(const double *blk)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(blk), (short)0,
                                            (int)(Dims<DIM>::BLOCK * sizeof(double)),
                                            0x00020000)) {}
  __device__ __forceinline__ double operator()(int voffset_bytes, int soffset_bytes) const {
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    u2 const v = __builtin_amdgcn_raw_buffer_load_b64(r, voffset_bytes, soffset_bytes, 0);
    return __builtin_bit_cast(double, v);
  }
};

// Logical workgroup index with consecutive logical groups kept on one XCD:
// the hardware deals workgroups round-robin over the 8 XCDs, so group g runs
// on XCD g % 8.  Remapping makes each XCD sweep a contiguous range of blocks,
// which keeps a block's stencil data in one XCD's L2.
__device__ __forceinline__ unsigned xcd_remap_group(unsigned g, unsigned n) {
  constexpr unsigned kXcd = 8;
  unsigned const per = n / kXcd;
  unsigned const full = per * kXcd;
  if (g >= full) return g;
  return (g % kXcd) * per + g / kXcd;
}

struct StageArgs {
  int nb[3];
  int bc[3];
  double gamma, b;
  double inv_dx;
  int nblocks;
  int stage;
  int signal;         // reduce the CFL signal of the new state into state->smax
  double *init_copy;  // stage 0: also copy the Average buffer here (or null)
  int guard;          // do nothing once the case record has reached its end time
};

// One tile of work for the fused stage: the case's buffers (block 0 of the
// case at pool index `base`), its block grid and side conditions, and the
// stage.  Shared by the single-case and the batched (persistent) kernels.
struct TileJob {
  const double *avg;
  const double *init;
  double *out;
  alpaca_case_state *state;
  double *init_copy;
  int base;
  int nb[3];
  int bc[3];
  double gamma, b, inv_dx;
  int stage;
  int signal;
};

// Device description of a case held in an ensemble pool.
struct CaseSlot {
  int nb[3];
  int bc[3];
  int base;      // pool index of the case's first block
  int nblocks;
  double gamma, b, inv_dx, cell_size;
};

// Ensemble case-record flags (alpaca_case_state::flags).
enum : int {
  kFlagNonFinite = 1,   // a non-finite value was written
  kFlagUnhealthy = 2,   // time step below DBL_EPSILON
  kFlagDone = 4,        // reached t_end
  kFlagCancelled = 8,   // cancelled by the host
  kFlagFresh = 16,      // dt already computed for the next step
  kFlagActive = 32      // slot holds a case
};
__host__ __device__ constexpr int kFlagsStop() {
  return kFlagNonFinite | kFlagUnhealthy | kFlagDone | kFlagCancelled;
}

struct WorkItem {
  int case_id;
  int block;  // block index within the case
};

namespace exact {
#pragma clang fp contract(off)
constexpr bool kFastArith = false;  // correctly rounded FP64 division and sqrt
#include "fv_numerics.inl"
#include "stage_kernel.inl"
}  // namespace exact

namespace fast {
#pragma clang fp contract(fast)
constexpr bool kFastArith = true;  // Newton-Raphson FP64 division and sqrt (<= 1 ulp)
#include "fv_numerics.inl"
#include "stage_kernel.inl"
}  // namespace fast

#pragma clang fp contract(off)

}  // namespace alpaca
