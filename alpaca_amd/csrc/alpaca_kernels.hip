// gfx950 kernels and the C ABI (include/alpaca_gpu.h) for ALPACA's uniform
// single-phase finite-volume right-hand side.
//
// Kernels (one launch each, all on the caller's stream):
//   face_flux_kernel<DIM,DIR,FLUX,R>  one thread per cell face; six-cell
//       stencil read straight from the haloed block (coalesced along k);
//       Roe eigensystem + characteristic WENO5 + HLLC or Roe splitting.
//   combine_kernel<DIM>  one thread per internal cell: flux divergence with
//       the reference's ConsistencyManagedSum, times 1/dx, fused with the
//       SSP-RK3 combination and the non-finite check.
//   halo_kernel<DIM>  face halo slabs from neighbours / zero-gradient / symmetry.
//   signal_kernel<DIM> + dt_kernel  CFL reduction and time step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/alpaca_gpu.h"
#include "fv_device.h"

// Opaque context of the per-block entries (include/alpaca_gpu.h): the HIP
// stream its calls are queued on and the face-flux scratch they use, so
// contexts on different streams never share scratch.
struct alpaca_context {
  int device;
  hipStream_t stream;
  std::mutex mutex;
  double *ptr = nullptr;
  size_t bytes = 0;
  std::vector<double *> retired;  // outgrown buffers, freed at destroy
  alpaca_context(int dev, hipStream_t s) : device(dev), stream(s) {}
};

namespace alpaca {
namespace {

constexpr int kThreads = 256;

// Faces per (block, direction): 17 along the direction, 16 across.
template <int DIM> constexpr int faces_per_dir() {
  return (kIC + 1) * kIC * (DIM == 3 ? kIC : 1);
}

struct CaseArgs {
  int nb[3];
  int bc[3];
  double gamma, b, cell_size;
};

// Which block of device memory the i-th block of a launch is: the i-th block
// of a contiguous array (list == nullptr), or list[i] of a pool of `capacity`
// blocks, an entry outside [0, capacity) being skipped (-1).
struct BlockList {
  const int *list = nullptr;
  int capacity = 0;
  __device__ __forceinline__ int block(int i) const {
    if (!list) return i;
    int const b = list[i];
    return (b >= 0 && b < capacity) ? b : -1;
  }
};

// ---------------------------------------------------------------- faces ---
// Output layout per (block, dir): [NE][n0][n1][n2] with n_dir = 17 and the
// other extents 16 (1 for z in 2D): face index a along DIR is the face
// between halo cells a+3 and a+4.
// Arithmetic modes: 0 = FP64 without contraction (bit-exact with the
// reference), 1 = FP64 with FMA contraction, 2 = mixed (FP32 faces).
enum : int { kModeExact = 0, kModeFma = 1, kModeMixed = 2 };

template <int DIM, int DIR, int FLUX, int MODE>
__device__ __forceinline__ void mode_stencil_flux(const double *bb, int i, int j, int k,
                                                  double gamma, double b, double *out) {
  constexpr int NE = Dims<DIM>::NE;
  if constexpr (MODE == kModeExact) {
    exact::stencil_flux<DIM, DIR, FLUX, double>(bb, i, j, k, gamma, b, out);
  } else if constexpr (MODE == kModeFma) {
    fast::stencil_flux<DIM, DIR, FLUX, double>(bb, i, j, k, gamma, b, out);
  } else {
    float f[NE];
    fast::stencil_flux<DIM, DIR, FLUX, float>(bb, i, j, k, gamma, b, f);
#pragma unroll
    for (int e = 0; e < NE; ++e) out[e] = double(f[e]);
  }
}

template <int DIM, int DIR, int FLUX, int MODE>
__global__ __launch_bounds__(kThreads) void face_flux_kernel(const double *__restrict__ in,
                                                             double *__restrict__ flux_out,
                                                             int nblocks, double gamma,
                                                             double b) {
  using D = Dims<DIM>;
  constexpr int NE = D::NE;
  constexpr int NF = faces_per_dir<DIM>();
  constexpr int n1 = DIR == 1 ? kIC + 1 : kIC;
  constexpr int n2 = DIM == 3 ? (DIR == 2 ? kIC + 1 : kIC) : 1;
  constexpr int groups_per_block = (NF + kThreads - 1) / kThreads;
  unsigned const g = xcd_remap_group(blockIdx.x, gridDim.x);
  int const blk = g / groups_per_block;
  int const f = (g % groups_per_block) * kThreads + threadIdx.x;
  if (blk >= nblocks || f >= NF) return;
  int const a2 = f % n2;
  int const a1 = (f / n2) % n1;
  int const a0 = f / (n2 * n1);
  // Halo index of the cell left of the face.
  int const i = a0 + (DIR == 0 ? kHS - 1 : kHS);
  int const j = a1 + (DIR == 1 ? kHS - 1 : kHS);
  int const k = DIM == 3 ? a2 + (DIR == 2 ? kHS - 1 : kHS) : 0;
  double fl[NE];
  mode_stencil_flux<DIM, DIR, FLUX, MODE>(in + (size_t)blk * D::BLOCK, i, j, k, gamma, b, fl);
  double *o = flux_out + ((size_t)blk * DIM + DIR) * NE * NF + f;
#pragma unroll
  for (int e = 0; e < NE; ++e) o[e * NF] = fl[e];
}

// ------------------------------------------------------- jump fluxes ---
// The boundary-flux save of SpaceSolver::UpdateFluxes (space_solver.cpp:
// 160-182): the fluxes of a block's six boundary faces ADDED into its jump
// surfaces [6][NE][IC][ICZ] (SurfaceBuffer: East, West, North, South, Top,
// Bottom).  One thread per surface entry; in 2D the z faces do not exist and
// Bottom/Top receive += 0.0, as from the reference's zeroed z face array.
template <int DIM>
__global__ __launch_bounds__(kThreads) void jump_flux_kernel(const double *__restrict__ flux,
                                                             double *__restrict__ jump,
                                                             int nblocks) {
  using D = Dims<DIM>;
  constexpr int NE = D::NE;
  constexpr int NF = faces_per_dir<DIM>();
  constexpr int ICZ = D::ICZ;
  constexpr int FACE = NE * kIC * ICZ;
  size_t const t = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= (size_t)nblocks * 6 * FACE) return;
  int const blk = int(t / (6 * FACE));
  int const loc = int((t / FACE) % 6);
  int const r = int(t % FACE);
  int const e = r / (kIC * ICZ), i = (r / ICZ) % kIC, j = r % ICZ;
  const double *fx = flux + ((size_t)blk * DIM + 0) * NE * NF + (size_t)e * NF;
  const double *fy = flux + ((size_t)blk * DIM + 1) * NE * NF + (size_t)e * NF;
  double v = 0.0;
  switch (loc) {
    case 0: v = fx[(kIC * kIC + i) * ICZ + j]; break;        // East: x face IC
    case 1: v = fx[i * ICZ + j]; break;                        // West: x face 0
    case 2: v = fy[(i * (kIC + 1) + kIC) * ICZ + j]; break;  // North: y face IC
    case 3: v = fy[(i * (kIC + 1)) * ICZ + j]; break;          // South: y face 0
    default:
      if constexpr (DIM == 3) {
        const double *fz = flux + ((size_t)blk * DIM + 2) * NE * NF + (size_t)e * NF;
        v = fz[(i * kIC + j) * (kIC + 1) + (loc == 4 ? kIC : 0)];  // Top / Bottom
      }
  }
  jump[t] += v;
}

// ------------------------------------------------------------- combine ---
struct StageCoef {
  double m0, m1, c;
};
__host__ __device__ inline StageCoef stage_coef(int stage) {
  // RungeKutta3TVD buffer and time-step multipliers (runge_kutta_3_TVD.h:39-59).
  if (stage == 0) return {1.0, 0.0, 1.0};
  if (stage == 1) return {0.25, 0.75, 0.25};
  return {2.0 / 3.0, 1.0 / 3.0, 2.0 / 3.0};
}

// Flux divergence (space_solver.cpp:131-153) fused with PrepareBufferFor-
// Integration and IntegrateConservatives (time_integrator.h:98-113, 434-457).
// MODE 0: RK stage; MODE 1: write the right-hand side itself.
template <int DIM, int MODE>
__global__ __launch_bounds__(kThreads) void combine_kernel(
    const double *__restrict__ flux, const double *__restrict__ avg, const double *init,
    double *out, int nblocks, int stage, double inv_dx, alpaca_case_state *state) {
  using D = Dims<DIM>;
  constexpr int NE = D::NE;
  constexpr int NF = faces_per_dir<DIM>();
  constexpr int ICZ = D::ICZ;
  constexpr int cells = D::CELLS;
  constexpr int groups_per_block = (cells + kThreads - 1) / kThreads;
  unsigned const g = xcd_remap_group(blockIdx.x, gridDim.x);
  int const blk = g / groups_per_block;
  int const c = (g % groups_per_block) * kThreads + threadIdx.x;
  if (blk >= nblocks || c >= cells) return;
  int const kk = c % ICZ;
  int const jj = (c / ICZ) % kIC;
  int const ii = c / (ICZ * kIC);
  const double *fx = flux + ((size_t)blk * DIM + 0) * NE * NF;
  const double *fy = flux + ((size_t)blk * DIM + 1) * NE * NF;
  const double *fz = DIM == 3 ? flux + ((size_t)blk * DIM + 2) * NE * NF : nullptr;
  // Face a along a direction lies between internal cells a-1 and a.
  int const x_lo = (ii * kIC + jj) * ICZ + kk;           // fx [17][16][16|1]
  int const x_hi = ((ii + 1) * kIC + jj) * ICZ + kk;
  int const y_lo = (ii * (kIC + 1) + jj) * ICZ + kk;     // fy [16][17][16|1]
  int const y_hi = (ii * (kIC + 1) + jj + 1) * ICZ + kk;
  int const z_lo = (ii * kIC + jj) * (kIC + 1) + kk;     // fz [16][16][17]
  int const z_hi = z_lo + 1;
  size_t const cell = (size_t)blk * D::BLOCK + (ii + kHS) * D::SX + (jj + kHS) * D::SY +
                      (DIM == 3 ? kk + kHS : 0);
  double dt = 0.0;
  StageCoef co{1.0, 0.0, 1.0};
  if constexpr (MODE == 0) {
    dt = state->dt;
    co = stage_coef(stage);
  }
  double const timestep = 0.0 + dt;
  double const cdt = co.c * timestep;
  bool bad = false;
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    double const dx_ = fx[e * NF + x_lo] - fx[e * NF + x_hi];
    double const dy_ = fy[e * NF + y_lo] - fy[e * NF + y_hi];
    double dz_ = 0.0;
    if constexpr (DIM == 3) dz_ = fz[e * NF + z_lo] - fz[e * NF + z_hi];
    double const rhs = 0.0 + (0.0 + exact::dams<DIM>(dx_, dy_, dz_) * inv_dx);
    size_t const idx = cell + (size_t)e * D::FIELD;
    if constexpr (MODE == 0) {
      double const u = avg[idx];
      double a = u;
      if (stage == 0) {
        if (init) const_cast<double *>(init)[idx] = u;
      } else {
        a = co.m0 * u + co.m1 * init[idx];
      }
      double const v = a + cdt * rhs;
      out[idx] = v;
      bad |= !__builtin_isfinite(v);
    } else {
      out[idx] = rhs;
    }
  }
  if constexpr (MODE == 0) {
    if (bad) atomicOr(&state->flags, 1);
  }
}

// ---------------------------------------------------------------- halos ---
// One thread per (block, halo slab cell): 2*DIM slabs of HS x 16 x 16(|1).
template <int DIM>
__global__ __launch_bounds__(kThreads) void halo_kernel(double *blocks, int nblocks, CaseArgs ca) {
  using D = Dims<DIM>;
  constexpr int NE = D::NE;
  constexpr int ICZ = D::ICZ;
  constexpr int slab = kHS * kIC * ICZ;
  constexpr int per_block = 2 * DIM * slab;
  size_t const t = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= (size_t)nblocks * per_block) return;
  int const blk = t / per_block;
  int const r = t % per_block;
  int const face = r / slab;  // 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z
  int const q = r % slab;
  int const d = face / 2;
  int const hi = face % 2;
  // Position within the slab: h across the slab (0..3), then the two
  // transverse internal indices.
  int const h = q / (kIC * ICZ);
  int const t1 = (q / ICZ) % kIC;
  int const t2 = q % ICZ;
  int bc3[3] = {blk / (ca.nb[1] * ca.nb[2]), (blk / ca.nb[2]) % ca.nb[1], blk % ca.nb[2]};
  // Halo cell (halo indices) and its source.
  int dst[3], src[3];
  int tr0 = d == 0 ? 1 : 0;
  int tr1 = d == 2 ? 1 : 2;
  dst[d] = hi ? kHS + kIC + h : h;
  dst[tr0] = t1 + kHS;
  dst[tr1] = DIM == 3 ? t2 + kHS : 0;
  int sblk = blk;
  src[0] = dst[0];
  src[1] = dst[1];
  src[2] = dst[2];
  bool const at_edge = hi ? (bc3[d] == ca.nb[d] - 1) : (bc3[d] == 0);
  bool const symmetry = at_edge && ca.bc[d] == ALPACA_BC_SYMMETRY;
  if (at_edge && ca.bc[d] == ALPACA_BC_ZERO_GRADIENT) {
    src[d] = hi ? kHS + kIC - 1 : kHS;
  } else if (symmetry) {
    src[d] = (hi ? 2 * (kHS + kIC) - 1 : 2 * kHS - 1) - dst[d];  // mirror about the face
  } else {
    int nb3[3] = {bc3[0], bc3[1], bc3[2]};
    nb3[d] = hi ? (bc3[d] + 1 == ca.nb[d] ? 0 : bc3[d] + 1) : (bc3[d] == 0 ? ca.nb[d] - 1 : bc3[d] - 1);
    sblk = (nb3[0] * ca.nb[1] + nb3[1]) * ca.nb[2] + nb3[2];
    src[d] = hi ? kHS + h : kIC + h;  // neighbour's first / last internal cells
  }
  size_t const di = (size_t)blk * D::BLOCK + dst[0] * D::SX + dst[1] * D::SY + dst[2];
  size_t const si = (size_t)sblk * D::BLOCK + src[0] * D::SX + src[1] * D::SY + src[2];
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    double const v = blocks[si + (size_t)e * D::FIELD];
    blocks[di + (size_t)e * D::FIELD] = (symmetry && e == kMomX + d) ? -1.0 * v : v;
  }
}

// Halo fill over a neighbour table: every halo cell (faces, edges, corners)
// of every block.  Its offset o in {-1,0,1}^DIM from the block selects the
// covering neighbour; a present neighbour supplies its internal cell.  When
// the offset leaves the domain, the axes whose face neighbour is absent are
// external: each maps its coordinate into the block as its side condition
// does (zero gradient: the first/last internal layer; symmetry: the mirror
// about the face, that axis's momentum negated), and the axes left over take
// the neighbour across them.  These maps act on different axes and commute,
// so the result equals the reference's order of internal copies followed by
// the external conditions over the full transverse extent.  Every source is
// an internal cell, so no thread reads a cell another writes.
template <int DIM> __device__ __forceinline__ int offset_slot(const int *o) {
  return DIM == 3 ? ((o[0] + 1) * 3 + (o[1] + 1)) * 3 + (o[2] + 1) : (o[0] + 1) * 3 + (o[1] + 1);
}

struct SideBc {
  int kind[6];  // -x, +x, -y, +y, -z, +z
};

template <int DIM>
__global__ __launch_bounds__(kThreads) void halo_table_kernel(double *blocks, int nblocks,
                                                              const int *__restrict__ nbr,
                                                              SideBc bc) {
  using D = Dims<DIM>;
  constexpr int NE = D::NE;
  constexpr int SLOTS = DIM == 3 ? 27 : 9;
  constexpr int CELLS = kTC * kTC * D::TZ;
  size_t const t = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= (size_t)nblocks * CELLS) return;
  int const blk = int(t / CELLS);
  int const r = int(t % CELLS);
  int c[3] = {r / (kTC * D::TZ), (r / D::TZ) % kTC, DIM == 3 ? r % kTC : 0};
  int const dst[3] = {c[0], c[1], c[2]};
  int o[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < DIM; ++a) o[a] = c[a] < kHS ? -1 : (c[a] >= kHS + kIC ? 1 : 0);
  if (o[0] == 0 && o[1] == 0 && o[2] == 0) return;  // internal cell
  const int *tb = nbr + (size_t)blk * SLOTS;
  int src_blk = tb[offset_slot<DIM>(o)];
  double sign[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) sign[e] = 1.0;
  if (src_blk < 0) {
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      if (o[a] == 0) continue;
      int f[3] = {0, 0, 0};
      f[a] = o[a];
      if (tb[offset_slot<DIM>(f)] >= 0) continue;  // an internal axis
      int const kind = bc.kind[2 * a + (o[a] > 0)];
      if (kind == ALPACA_BC_ZERO_GRADIENT) {
        c[a] = o[a] < 0 ? kHS : kHS + kIC - 1;
      } else if (kind == ALPACA_BC_SYMMETRY) {
        c[a] = (o[a] < 0 ? 2 * kHS - 1 : 2 * (kHS + kIC) - 1) - c[a];
        sign[kMomX + a] = -1.0 * sign[kMomX + a];
      } else {
        return;  // a periodic side must appear in the table
      }
      o[a] = 0;
    }
    src_blk = (o[0] == 0 && o[1] == 0 && o[2] == 0) ? blk : tb[offset_slot<DIM>(o)];
    if (src_blk < 0) return;  // no covering neighbour across the remaining axes
  }
  if (src_blk >= nblocks) return;
  size_t const di = (size_t)blk * D::BLOCK + dst[0] * D::SX + dst[1] * D::SY + dst[2];
  size_t const si = (size_t)src_blk * D::BLOCK + (c[0] - o[0] * kIC) * D::SX +
                    (c[1] - o[1] * kIC) * D::SY + (c[2] - o[2] * kIC);
#pragma unroll
  for (int e = 0; e < NE; ++e) blocks[di + (size_t)e * D::FIELD] = sign[e] * blocks[si + (size_t)e * D::FIELD];
}

// Halo update of listed blocks of a multi-level pool: the same-level part
// of MaterialHaloUpdateOnLevel (internal_halo_manager.cpp:53-97) and the
// external conditions after it (halo_manager.cpp:89-123).  Row i of the
// table belongs to block list[i]; an entry is the neighbour's pool block,
// -1 where the offset leaves the domain, or -2 where the neighbour does not
// exist on this level (a level jump: the region is predicted from the parent
// by the jump update).  Mode 0: internal copies and external conditions,
// leaving every cell whose source lies in a jump region.  Mode 1, after the
// jump update: only those cells -- the external condition's source is then
// the block's own halo cell in the jump region (the reference applies the
// external conditions last, over the full transverse extent, so they read
// the predicted values).  Writes and reads of one launch never meet: mode 0
// reads internal cells, mode 1 reads jump-region halo cells and writes cells
// of external slabs.
template <int DIM>
__global__ __launch_bounds__(kThreads) void halo_list_kernel(double *blocks, int nblocks,
                                                             const int *__restrict__ list, int n,
                                                             const int *__restrict__ nbr,
                                                             SideBc bc, int mode) {
  using D = Dims<DIM>;
  constexpr int NE = D::NE;
  constexpr int SLOTS = DIM == 3 ? 27 : 9;
  constexpr int CELLS = kTC * kTC * D::TZ;
  size_t const t = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= (size_t)n * CELLS) return;
  int const row_i = int(t / CELLS);
  int const blk = list[row_i];
  if (blk < 0 || blk >= nblocks) return;
  int const r = int(t % CELLS);
  int c[3] = {r / (kTC * D::TZ), (r / D::TZ) % kTC, DIM == 3 ? r % kTC : 0};
  int const dst[3] = {c[0], c[1], c[2]};
  int o[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < DIM; ++a) o[a] = c[a] < kHS ? -1 : (c[a] >= kHS + kIC ? 1 : 0);
  if (o[0] == 0 && o[1] == 0 && o[2] == 0) return;  // internal cell
  const int *tb = nbr + (size_t)row_i * SLOTS;
  int src_blk = tb[offset_slot<DIM>(o)];
  if (src_blk == -2) return;  // jump region
  bool external = false;
  double sign[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) sign[e] = 1.0;
  if (src_blk < 0) {
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      if (o[a] == 0) continue;
      int f[3] = {0, 0, 0};
      f[a] = o[a];
      if (tb[offset_slot<DIM>(f)] != -1) continue;  // a neighbour or a jump across this axis
      int const kind = bc.kind[2 * a + (o[a] > 0)];
      if (kind == ALPACA_BC_ZERO_GRADIENT) {
        c[a] = o[a] < 0 ? kHS : kHS + kIC - 1;
      } else if (kind == ALPACA_BC_SYMMETRY) {
        c[a] = (o[a] < 0 ? 2 * kHS - 1 : 2 * (kHS + kIC) - 1) - c[a];
        sign[kMomX + a] = -1.0 * sign[kMomX + a];
      } else {
        return;  // a periodic side must appear in the table
      }
      o[a] = 0;
      external = true;
    }
    src_blk = (o[0] == 0 && o[1] == 0 && o[2] == 0) ? blk : tb[offset_slot<DIM>(o)];
    if (src_blk == -2) {
      if (mode != 1 || !external) return;
      // The external condition reads the block's own predicted halo cell.
      size_t const di = (size_t)blk * D::BLOCK + dst[0] * D::SX + dst[1] * D::SY + dst[2];
      size_t const si = (size_t)blk * D::BLOCK + c[0] * D::SX + c[1] * D::SY + c[2];
#pragma unroll
      for (int e = 0; e < NE; ++e)
        blocks[di + (size_t)e * D::FIELD] = sign[e] * blocks[si + (size_t)e * D::FIELD];
      return;
    }
    if (src_blk < 0) return;  // no covering neighbour across the remaining axes
  }
  if (mode != 0 || src_blk >= nblocks) return;
  size_t const di = (size_t)blk * D::BLOCK + dst[0] * D::SX + dst[1] * D::SY + dst[2];
  size_t const si = (size_t)src_blk * D::BLOCK + (c[0] - o[0] * kIC) * D::SX +
                    (c[1] - o[1] * kIC) * D::SY + (c[2] - o[2] * kIC);
#pragma unroll
  for (int e = 0; e < NE; ++e) blocks[di + (size_t)e * D::FIELD] = sign[e] * blocks[si + (size_t)e * D::FIELD];
}

// ----------------------------------------------------------- time step ---

// Per block: max over internal cells of sum_d(|u_d| + c)
// (modular_algorithm_assembler.cpp:1840-1869), then an atomic max on the
// IEEE bits (non-negative doubles order like their bit patterns).
template <int DIM>
__global__ __launch_bounds__(kThreads) void signal_kernel(const double *__restrict__ in, int nblocks,
                                                          double gamma, double b,
                                                          alpaca_case_state *state, BlockList list) {
  using D = Dims<DIM>;
  constexpr int NE = D::NE;
  constexpr int ICZ = D::ICZ;
  __shared__ double part[kThreads / 64];
  int const blk = blockIdx.x;
  if (blk >= nblocks) return;
  int const mem = list.block(blk);  // uniform over the workgroup
  if (mem < 0) return;
  double s = 0.0;
  for (int c = threadIdx.x; c < D::CELLS; c += kThreads) {
    int const kk = c % ICZ, jj = (c / ICZ) % kIC, ii = c / (ICZ * kIC);
    const double *p = in + (size_t)mem * D::BLOCK + (ii + kHS) * D::SX + (jj + kHS) * D::SY +
                      (DIM == 3 ? kk + kHS : 0);
    double u[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) u[e] = p[e * D::FIELD];
    exact::Prim<double> pr;
    exact::cons_to_prim<DIM>(gamma, b, u, pr);
    double const cs = exact::eos_sos(gamma, b, pr.rho, pr.p);
    double sum;
    if constexpr (DIM == 2)
      sum = (__builtin_fabs(pr.vel[0]) + cs) + (__builtin_fabs(pr.vel[1]) + cs);
    else
      sum = exact::cms3(__builtin_fabs(pr.vel[0]) + cs, __builtin_fabs(pr.vel[1]) + cs,
                 __builtin_fabs(pr.vel[2]) + cs);
    s = exact::smax(s, sum);
  }
  s = exact::wave_max64(s);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = part[0];
#pragma unroll
    for (int w = 1; w < kThreads / 64; ++w) m = exact::smax(m, part[w]);
    atomicMax(reinterpret_cast<unsigned long long *>(&state->smax),
              (unsigned long long)__double_as_longlong(m));
  }
}

// dt = CFL / (S / dx), end-time clamp, minimum-step health check
// (modular_algorithm_assembler.cpp:1932-1982; ComputeLoop :137-140).
__global__ void dt_kernel(alpaca_case_state *state, double cell_size, int guard = 0) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (guard && (state->flags & kFlagDone)) return;
  double const s = __longlong_as_double((long long)state->smax);
  double const rate = exact::fdiv(s, cell_size);
  double dt = rate == 0.0 ? 1.7976931348623157e308 : exact::fdiv(state->cfl, rate);
  if (state->time + dt > state->t_end) dt = exact::smax(state->t_end - state->time, 0.0);
  state->dt = dt;
  state->smax = 0;
  if (dt < 2.220446049250313080847263336181640625e-16 && dt > 0.0) state->flags |= 2;
}

// Zero fill of n doubles (grid-stride): used in place of hipMemsetAsync so
// that a caller may capture the per-block entries in a HIP graph.
__global__ __launch_bounds__(kThreads) void zero_kernel(double *p, size_t n) {
  size_t const stride = (size_t)gridDim.x * kThreads;
  for (size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x; t < n; t += stride) p[t] = 0.0;
}

// The listed blocks of a pool, 16 bytes per thread and step (grid-stride; a
// haloed block is an even number of doubles, so every pair is 16-byte
// aligned): OP 1 exchanges the blocks of a and b
// (SwapConservativeBuffersForNode, buffer_operations_material.h:83-91), OP 2
// copies a's blocks into b's.
template <int OP>
__global__ __launch_bounds__(kThreads) void listed_blocks_kernel(double *a, double *b,
                                                                 size_t block_doubles,
                                                                 BlockList list, int n) {
  size_t const pairs = block_doubles / 2;
  size_t const total = (size_t)n * pairs;
  size_t const stride = (size_t)gridDim.x * kThreads;
  for (size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x; t < total; t += stride) {
    int const mem = list.block(int(t / pairs));
    if (mem < 0) continue;
    size_t const o = (size_t)mem * pairs + t % pairs;
    double2 *pa = reinterpret_cast<double2 *>(a) + o;
    if constexpr (OP == 2) {
      reinterpret_cast<double2 *>(b)[o] = *pa;
    } else {
      double2 *pb = reinterpret_cast<double2 *>(b) + o;
      double2 const va = *pa;
      double2 const vb = *pb;
      *pa = vb;
      *pb = va;
    }
  }
}

template <int OP>
void launch_listed_blocks(double *a, double *b, int dim, BlockList list, int n, hipStream_t s) {
  size_t const bd = alpaca_block_doubles(dim);
  size_t const groups = ((size_t)n * (bd / 2) + kThreads - 1) / kThreads;
  hipLaunchKernelGGL((listed_blocks_kernel<OP>), dim3((unsigned)std::min<size_t>(groups, 65536)),
                     dim3(kThreads), 0, s, a, b, bd, list, n);
}

// The exact form's FP64 division and square root as the face kernels
// evaluate them, elementwise (alpaca_exact_fp64_ops): op 0 the division of
// the physical quantities, op 1 the WENO5 weight division (its in-range core),
// op 2 the square root.
template <int OP>
__global__ __launch_bounds__(kThreads) void exact_ops_kernel(const double *__restrict__ a,
                                                             const double *__restrict__ b,
                                                             double *__restrict__ out, int n) {
  int const i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  if constexpr (OP == 0)
    out[i] = exact::fdiv(a[i], b[i]);
  else if constexpr (OP == 1)
    out[i] = exact::weno_div(a[i], b[i]);
  else if constexpr (OP == 2)
    out[i] = exact::dsqrt(a[i]);
  else if constexpr (OP == 3)
    out[i] = fast::fdiv(a[i], b[i]);
  else
    out[i] = fast::dsqrt(a[i]);
}

// FinishMacroTimestep (time_integrator.h:244-248).
// TimeIntegrator::FinishMacroTimestep (time := time + dt) and ComputeLoop's
// end test (modular_algorithm_assembler.cpp:137-140): the done flag once the
// time is no longer below t_end.  guard: a record already done is left alone.
__global__ void finish_step_kernel(alpaca_case_state *state, int guard = 0) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (guard && (state->flags & kFlagDone)) return;
  state->time += 0.0 + state->dt;
  state->steps += 1;
  if (!(state->time < state->t_end)) state->flags |= kFlagDone;
}

// finish_step_kernel followed by dt_kernel for the next step, both guarded,
// in one launch (the same operations on the record in the same order): the
// step boundary inside alpaca_advance.
__global__ void finish_and_dt_kernel(alpaca_case_state *state, double cell_size) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (state->flags & kFlagDone) return;
  state->time += 0.0 + state->dt;
  state->steps += 1;
  if (!(state->time < state->t_end)) {
    state->flags |= kFlagDone;
    return;
  }
  double const s = __longlong_as_double((long long)state->smax);
  double const rate = exact::fdiv(s, cell_size);
  double dt = rate == 0.0 ? 1.7976931348623157e308 : exact::fdiv(state->cfl, rate);
  if (state->time + dt > state->t_end) dt = exact::smax(state->t_end - state->time, 0.0);
  state->dt = dt;
  state->smax = 0;
  if (dt < 2.220446049250313080847263336181640625e-16 && dt > 0.0) state->flags |= 2;
}

__global__ void state_init_kernel(alpaca_case_state *state, double time, double t_end, double cfl) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  state->time = time;
  state->t_end = t_end;
  state->cfl = cfl;
  state->dt = 0.0;
  state->smax = 0;
  state->steps = 0;
  state->flags = 0;
}

// Reference face-array layout for the per-block convective-flux entry:
// three arrays [NE][17][17][17|2]; face (a0,a1,a2) of direction DIR maps to
// (i - 3, j - 3, k - 3) with k - (-1) in 2D (finite_volume_scheme.cpp:176-180).
template <int DIM>
__global__ __launch_bounds__(kThreads) void scatter_ref_fluxes(const double *__restrict__ flux,
                                                               double *out, int nblocks) {
  using D = Dims<DIM>;
  constexpr int NE = D::NE;
  constexpr int NF = faces_per_dir<DIM>();
  constexpr int nz = DIM == 3 ? kIC + 1 : 2;
  constexpr int ref_per_dir = NE * (kIC + 1) * (kIC + 1) * nz;
  size_t const t = (size_t)blockIdx.x * kThreads + threadIdx.x;
  size_t const total = (size_t)nblocks * DIM * NF;
  if (t >= total) return;
  int const f = t % NF;
  int const dir = (t / NF) % DIM;
  int const blk = t / ((size_t)NF * DIM);
  int const n0 = dir == 0 ? kIC + 1 : kIC;
  int const n1 = dir == 1 ? kIC + 1 : kIC;
  int const n2 = DIM == 3 ? (dir == 2 ? kIC + 1 : kIC) : 1;
  int const a2 = f % n2, a1 = (f / n2) % n1, a0 = f / (n2 * n1);
  (void)n0;
  int const ri = a0 + (dir == 0 ? 0 : 1);
  int const rj = a1 + (dir == 1 ? 0 : 1);
  int const rk = DIM == 3 ? a2 + (dir == 2 ? 0 : 1) : 1;
  const double *src = flux + ((size_t)blk * DIM + dir) * NE * NF + f;
  double *dst = out + ((size_t)blk * 3 + dir) * ref_per_dir;
#pragma unroll
  for (int e = 0; e < NE; ++e)
    dst[((e * (kIC + 1) + ri) * (kIC + 1) + rj) * nz + rk] += src[e * NF];
}

// --------------------------------------------------------------- host ---
// Optional kernel timing: HIP events recorded on the launch stream around
// each launch of a kernel class, summed on read.
struct ProfileSlot {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  double units = 0.0;  // algorithmic work units (faces or cells) launched
};
std::mutex g_prof_mutex;
bool g_prof_on = false;
// Set while an entry point that may not allocate runs on this thread:
// ProfileScope then records no events (hipEventCreate allocates).
thread_local bool g_prof_suppressed = false;
ProfileSlot g_prof[ALPACA_KERNEL_COUNT];

struct ProfileScope {
  int kernel;
  hipStream_t stream;
  hipEvent_t start = nullptr, stop = nullptr;
  ProfileScope(int k, hipStream_t s, double units) : kernel(k), stream(s) {
    if (!g_prof_on || g_prof_suppressed) return;
    if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess) {
      start = stop = nullptr;
      return;
    }
    (void)hipEventRecord(start, stream);
    std::lock_guard<std::mutex> lock(g_prof_mutex);
    g_prof[kernel].units += units;
  }
  ~ProfileScope() {
    if (!start) return;
    (void)hipEventRecord(stop, stream);
    std::lock_guard<std::mutex> lock(g_prof_mutex);
    g_prof[kernel].events.emplace_back(start, stop);
  }
};

int hip_status(hipError_t e) { return e == hipSuccess ? ALPACA_OK : ALPACA_ERR_HIP; }

// Implicit contexts of the stream-argument entries: one per (device, stream),
// created on first use and kept until the process ends.
std::mutex g_ctx_mutex;
std::map<std::pair<int, hipStream_t>, alpaca_context *> g_stream_ctx;

int stream_context(hipStream_t s, alpaca_context **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return ALPACA_ERR_HIP;
  std::lock_guard<std::mutex> lock(g_ctx_mutex);
  alpaca_context *&c = g_stream_ctx[{dev, s}];
  if (!c) c = new alpaca_context(dev, s);
  *out = c;
  return ALPACA_OK;
}

// Face-flux scratch of at least `bytes` from a context on the current
// device.  A buffer that is too small is retired, not freed: work already
// queued on the context's stream may still read it, and freeing it here would
// need a device synchronisation.  Growth is geometric, so the retired
// buffers together stay below the live one.
int context_scratch(alpaca_context *c, size_t bytes, double **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return ALPACA_ERR_HIP;
  if (dev != c->device) return ALPACA_ERR_INVALID;
  std::lock_guard<std::mutex> lock(c->mutex);
  if (c->bytes < bytes) {
    size_t const grown = std::max(bytes, c->bytes + c->bytes / 2);
    double *p = nullptr;
    if (hipMalloc(&p, grown) != hipSuccess) return ALPACA_ERR_HIP;
    if (c->ptr) c->retired.push_back(c->ptr);
    c->ptr = p;
    c->bytes = grown;
  }
  *out = c->ptr;
  return ALPACA_OK;
}

int validate(const alpaca_case_desc *d) {
  if (!d) return ALPACA_ERR_INVALID;
  if (d->dim != 2 && d->dim != 3) return ALPACA_ERR_INVALID;
  for (int a = 0; a < d->dim; ++a)
    if (d->nb[a] < 1 || (d->bc[a] != ALPACA_BC_ZERO_GRADIENT && d->bc[a] != ALPACA_BC_PERIODIC &&
                         d->bc[a] != ALPACA_BC_SYMMETRY))
      return ALPACA_ERR_INVALID;
  if (d->dim == 2 && d->nb[2] != 1) return ALPACA_ERR_INVALID;
  if ((long long)d->nb[0] * d->nb[1] * (d->dim == 3 ? d->nb[2] : 1) > kMaxBlocks)
    return ALPACA_ERR_INVALID;
  if (d->flux != ALPACA_FLUX_HLLC && d->flux != ALPACA_FLUX_ROE) return ALPACA_ERR_INVALID;
  if (d->precision != ALPACA_FP64 && d->precision != ALPACA_MIXED &&
      d->precision != ALPACA_FP64_FMA)
    return ALPACA_ERR_INVALID;
  if (!(d->cell_size > 0.0) || !(d->gamma > 1.0)) return ALPACA_ERR_INVALID;
  return ALPACA_OK;
}

int case_blocks(const alpaca_case_desc *d) { return d->nb[0] * d->nb[1] * (d->dim == 3 ? d->nb[2] : 1); }

template <int DIM>
constexpr size_t flux_doubles(int nblocks) {
  return (size_t)nblocks * DIM * Dims<DIM>::NE * faces_per_dir<DIM>();
}

template <int DIM, int DIR, int FLUX>
void launch_faces_dir(const alpaca_case_desc *d, int nblocks, const double *in, double *fl,
                      hipStream_t s) {
  constexpr int NF = faces_per_dir<DIM>();
  constexpr int gpb = (NF + kThreads - 1) / kThreads;
  dim3 grid((unsigned)(nblocks * gpb));
  ProfileScope prof(ALPACA_KERNEL_FACE_FLUX, s, double(NF) * nblocks);
  if (d->precision == ALPACA_MIXED)
    hipLaunchKernelGGL((face_flux_kernel<DIM, DIR, FLUX, kModeMixed>), grid, dim3(kThreads), 0, s,
                       in, fl, nblocks, d->gamma, d->background_pressure);
  else if (d->precision == ALPACA_FP64_FMA)
    hipLaunchKernelGGL((face_flux_kernel<DIM, DIR, FLUX, kModeFma>), grid, dim3(kThreads), 0, s,
                       in, fl, nblocks, d->gamma, d->background_pressure);
  else
    hipLaunchKernelGGL((face_flux_kernel<DIM, DIR, FLUX, kModeExact>), grid, dim3(kThreads), 0, s,
                       in, fl, nblocks, d->gamma, d->background_pressure);
}

template <int DIM, int FLUX>
void launch_faces_flux(const alpaca_case_desc *d, int nblocks, const double *in, double *fl,
                       hipStream_t s) {
  launch_faces_dir<DIM, 0, FLUX>(d, nblocks, in, fl, s);
  launch_faces_dir<DIM, 1, FLUX>(d, nblocks, in, fl, s);
  if constexpr (DIM == 3) launch_faces_dir<DIM, 2, FLUX>(d, nblocks, in, fl, s);
}

template <int DIM>
void launch_faces(const alpaca_case_desc *d, int nblocks, const double *in, double *fl,
                  hipStream_t s) {
  if (d->flux == ALPACA_FLUX_HLLC)
    launch_faces_flux<DIM, ALPACA_FLUX_HLLC>(d, nblocks, in, fl, s);
  else
    launch_faces_flux<DIM, ALPACA_FLUX_ROE>(d, nblocks, in, fl, s);
}

// Faces into the scratch fl (flux_doubles<DIM>(nblocks) doubles), then the
// divergence (mode 1) or divergence + RK combination (mode 0) into out.
template <int DIM>
int stage_impl(const alpaca_case_desc *d, int nblocks, int stage, int mode, const double *avg,
               double *init, double *out, alpaca_case_state *state, hipStream_t s, double *fl,
               double *jump = nullptr) {
  launch_faces<DIM>(d, nblocks, avg, fl, s);
  constexpr int cells = Dims<DIM>::CELLS;
  constexpr int gpb = (cells + kThreads - 1) / kThreads;
  dim3 grid((unsigned)(nblocks * gpb));
  double const inv_dx = 1.0 / d->cell_size;
  ProfileScope prof(ALPACA_KERNEL_COMBINE, s, double(cells) * nblocks);
  if (mode == 0) {
    double *ini = (init && init != avg) ? init : nullptr;
    if (stage > 0) ini = init;
    hipLaunchKernelGGL((combine_kernel<DIM, 0>), grid, dim3(kThreads), 0, s, fl, avg, ini, out,
                       nblocks, stage, inv_dx, state);
  } else {
    hipLaunchKernelGGL((combine_kernel<DIM, 1>), grid, dim3(kThreads), 0, s, fl, avg, nullptr,
                       out, nblocks, 0, inv_dx, state);
  }
  if (jump) {
    size_t const total = (size_t)nblocks * 6 * Dims<DIM>::NE * kIC * Dims<DIM>::ICZ;
    hipLaunchKernelGGL((jump_flux_kernel<DIM>), dim3((unsigned)((total + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, s, fl, jump, nblocks);
  }
  return hip_status(hipGetLastError());
}

// Tile shape of the fused stage kernel: 8x8x8 cells and 256 threads in 3D,
// the whole 16x16 block and 128 threads in 2D.
// Fused stage kernel of an arithmetic mode.
template <int DIM, int FLUX, int MODE, int TX, int TY, int TZ, int NT>
void launch_stage_kernel(dim3 grid, hipStream_t s, const double *avg, const double *init,
                         double *out, alpaca_case_state *state, const StageArgs &a) {
  if constexpr (MODE == kModeExact)
    hipLaunchKernelGGL((exact::stage_kernel<DIM, FLUX, double, TX, TY, TZ, NT>), grid, dim3(NT), 0,
                       s, avg, init, out, state, a);
  else if constexpr (MODE == kModeFma)
    hipLaunchKernelGGL((fast::stage_kernel<DIM, FLUX, double, TX, TY, TZ, NT>), grid, dim3(NT), 0,
                       s, avg, init, out, state, a);
  else
    hipLaunchKernelGGL((fast::stage_kernel<DIM, FLUX, float, TX, TY, TZ, NT>), grid, dim3(NT), 0,
                       s, avg, init, out, state, a);
}

// 2D cases with fewer leaf blocks than this run on 8x8 tiles.
constexpr int kSmallCaseBlocks = 256;

// Tile shape of the fused stage kernel: 8x8x8 cells and 256 threads in 3D,
// the whole 16x16 block and 128 threads in 2D (8x8 and 256 threads for a case
// of fewer than kSmallCaseBlocks blocks).
template <int DIM, int FLUX, int MODE>
void launch_stage_fused(const alpaca_case_desc *d, int stage, const double *avg,
                        const double *init, double *out, double *init_copy,
                        alpaca_case_state *state, int signal, hipStream_t s, bool guard) {
  int const nblocks = case_blocks(d);
  StageArgs a;
  a.nb[0] = d->nb[0];
  a.nb[1] = d->nb[1];
  a.nb[2] = DIM == 3 ? d->nb[2] : 1;
  a.bc[0] = d->bc[0];
  a.bc[1] = d->bc[1];
  a.bc[2] = DIM == 3 ? d->bc[2] : 0;
  a.gamma = d->gamma;
  a.b = d->background_pressure;
  a.inv_dx = 1.0 / d->cell_size;
  a.nblocks = nblocks;
  a.stage = stage;
  a.signal = signal;
  a.guard = guard;
  a.init_copy = init_copy;
  ProfileScope prof(ALPACA_KERNEL_STAGE, s, double(Dims<DIM>::CELLS) * nblocks);
  if constexpr (DIM == 3) {
    constexpr int TX = 8, TY = 8, TZ = 8, NT = 256;
    constexpr int tiles = (kIC / TX) * (kIC / TY) * (kIC / TZ);
    launch_stage_kernel<3, FLUX, MODE, TX, TY, TZ, NT>(dim3(nblocks * tiles), s, avg, init, out,
                                                       state, a);
  } else if (nblocks >= kSmallCaseBlocks) {
    launch_stage_kernel<2, FLUX, MODE, 16, 16, 1, 128>(dim3(nblocks), s, avg, init, out, state, a);
  } else {
    // Fewer blocks than CUs: four 8x8 tiles per block spread the case over
    // four times as many workgroups, and 256 threads give every thread at
    // most one of a tile's faces (the same faces and sums per cell).
    launch_stage_kernel<2, FLUX, MODE, 8, 8, 1, 256>(dim3(nblocks * 4), s, avg, init, out, state, a);
  }
}

// Right-hand side with the jump-flux save over listed pool blocks: the fused
// tile kernel (faces in LDS), tile shapes as for the fused stage kernel.
template <int DIM, int FLUX, int MODE>
void launch_rhs_tiles(const alpaca_case_desc *d, const double *avg, double *rhs, double *jump,
                      const int *list, int n, int capacity, hipStream_t s) {
  constexpr int TX = DIM == 3 ? 8 : 16, TY = DIM == 3 ? 8 : 16, TZ = DIM == 3 ? 8 : 1;
  constexpr int NT = DIM == 3 ? 256 : 128;
  constexpr int tiles = exact::TileShape<DIM, TX, TY, TZ>::TILES;
  dim3 const grid((unsigned)((size_t)n * tiles));
  double const inv_dx = 1.0 / d->cell_size;
  ProfileScope prof(ALPACA_KERNEL_STAGE, s, double(Dims<DIM>::CELLS) * n);
  if constexpr (MODE == kModeExact)
    hipLaunchKernelGGL((exact::rhs_tile_kernel<DIM, FLUX, double, TX, TY, TZ, NT>), grid, dim3(NT),
                       0, s, avg, rhs, jump, list, n, capacity, d->gamma, d->background_pressure,
                       inv_dx);
  else if constexpr (MODE == kModeFma)
    hipLaunchKernelGGL((fast::rhs_tile_kernel<DIM, FLUX, double, TX, TY, TZ, NT>), grid, dim3(NT),
                       0, s, avg, rhs, jump, list, n, capacity, d->gamma, d->background_pressure,
                       inv_dx);
  else
    hipLaunchKernelGGL((fast::rhs_tile_kernel<DIM, FLUX, float, TX, TY, TZ, NT>), grid, dim3(NT),
                       0, s, avg, rhs, jump, list, n, capacity, d->gamma, d->background_pressure,
                       inv_dx);
}

template <int DIM>
void rhs_tiles(const alpaca_case_desc *d, const double *avg, double *rhs, double *jump,
               const int *list, int n, int capacity, hipStream_t s) {
  int const p = d->precision;
  if (d->flux == ALPACA_FLUX_HLLC) {
    if (p == ALPACA_MIXED)
      launch_rhs_tiles<DIM, ALPACA_FLUX_HLLC, kModeMixed>(d, avg, rhs, jump, list, n, capacity, s);
    else if (p == ALPACA_FP64_FMA)
      launch_rhs_tiles<DIM, ALPACA_FLUX_HLLC, kModeFma>(d, avg, rhs, jump, list, n, capacity, s);
    else
      launch_rhs_tiles<DIM, ALPACA_FLUX_HLLC, kModeExact>(d, avg, rhs, jump, list, n, capacity, s);
  } else {
    if (p == ALPACA_MIXED)
      launch_rhs_tiles<DIM, ALPACA_FLUX_ROE, kModeMixed>(d, avg, rhs, jump, list, n, capacity, s);
    else if (p == ALPACA_FP64_FMA)
      launch_rhs_tiles<DIM, ALPACA_FLUX_ROE, kModeFma>(d, avg, rhs, jump, list, n, capacity, s);
    else
      launch_rhs_tiles<DIM, ALPACA_FLUX_ROE, kModeExact>(d, avg, rhs, jump, list, n, capacity, s);
  }
}

template <int DIM>
void stage_fused(const alpaca_case_desc *d, int stage, const double *avg, const double *init,
                 double *out, double *init_copy, alpaca_case_state *state, int signal,
                 hipStream_t s, bool guard = false) {
  int const p = d->precision;
  if (d->flux == ALPACA_FLUX_HLLC) {
    if (p == ALPACA_MIXED)
      launch_stage_fused<DIM, ALPACA_FLUX_HLLC, kModeMixed>(d, stage, avg, init, out, init_copy, state, signal, s, guard);
    else if (p == ALPACA_FP64_FMA)
      launch_stage_fused<DIM, ALPACA_FLUX_HLLC, kModeFma>(d, stage, avg, init, out, init_copy, state, signal, s, guard);
    else
      launch_stage_fused<DIM, ALPACA_FLUX_HLLC, kModeExact>(d, stage, avg, init, out, init_copy, state, signal, s, guard);
  } else {
    if (p == ALPACA_MIXED)
      launch_stage_fused<DIM, ALPACA_FLUX_ROE, kModeMixed>(d, stage, avg, init, out, init_copy, state, signal, s, guard);
    else if (p == ALPACA_FP64_FMA)
      launch_stage_fused<DIM, ALPACA_FLUX_ROE, kModeFma>(d, stage, avg, init, out, init_copy, state, signal, s, guard);
    else
      launch_stage_fused<DIM, ALPACA_FLUX_ROE, kModeExact>(d, stage, avg, init, out, init_copy, state, signal, s, guard);
  }
}

template <int DIM>
int halo_impl(const alpaca_case_desc *d, double *blocks, hipStream_t s) {
  int const nblocks = case_blocks(d);
  CaseArgs ca{{d->nb[0], d->nb[1], DIM == 3 ? d->nb[2] : 1},
              {d->bc[0], d->bc[1], DIM == 3 ? d->bc[2] : 0},
              d->gamma, d->background_pressure, d->cell_size};
  size_t const total = (size_t)nblocks * 2 * DIM * kHS * kIC * Dims<DIM>::ICZ;
  dim3 grid((unsigned)((total + kThreads - 1) / kThreads));
  ProfileScope prof(ALPACA_KERNEL_HALO, s, double(total));
  hipLaunchKernelGGL((halo_kernel<DIM>), grid, dim3(kThreads), 0, s, blocks, nblocks, ca);
  return hip_status(hipGetLastError());
}

template <int DIM>
int halo_table_impl(double *blocks, int nblocks, const int *nbr, const SideBc &bc, hipStream_t s) {
  size_t const total = (size_t)nblocks * kTC * kTC * Dims<DIM>::TZ;
  dim3 grid((unsigned)((total + kThreads - 1) / kThreads));
  ProfileScope prof(ALPACA_KERNEL_HALO, s, double(total));
  hipLaunchKernelGGL((halo_table_kernel<DIM>), grid, dim3(kThreads), 0, s, blocks, nblocks, nbr, bc);
  return hip_status(hipGetLastError());
}

template <int DIM>
int timestep_impl(const alpaca_case_desc *d, const double *avg, alpaca_case_state *state,
                  hipStream_t s, int nblocks, BlockList list = {}, int guard = 0) {
  ProfileScope prof(ALPACA_KERNEL_TIMESTEP, s, double(Dims<DIM>::CELLS) * nblocks);
  hipLaunchKernelGGL((signal_kernel<DIM>), dim3(nblocks), dim3(kThreads), 0, s, avg, nblocks,
                     d->gamma, d->background_pressure, state, list);
  hipLaunchKernelGGL(dt_kernel, dim3(1), dim3(64), 0, s, state, d->cell_size, guard);
  return hip_status(hipGetLastError());
}

}  // namespace
}  // namespace alpaca

using namespace alpaca;

extern "C" {

const char *alpaca_status_string(int status) {
  switch (status) {
    case ALPACA_OK: return "ok";
    case ALPACA_ERR_INVALID: return "invalid argument";
    case ALPACA_ERR_HIP: return "HIP runtime error";
    case ALPACA_ERR_NONFINITE: return "non-finite state";
    case ALPACA_ERR_UNSUPPORTED: return "unsupported configuration";
    case ALPACA_ERR_CAPACITY: return "capacity exhausted";
    default: return "unknown status";
  }
}

int alpaca_abi_version(void) { return ALPACA_GPU_ABI_VERSION; }

int alpaca_profile_enable(int on) {
  std::lock_guard<std::mutex> lock(g_prof_mutex);
  g_prof_on = on != 0;
  return ALPACA_OK;
}

int alpaca_profile_read(int kernel, int reset, double *total_ms, long long *launches,
                        double *units) {
  if (kernel < 0 || kernel >= ALPACA_KERNEL_COUNT || !total_ms || !launches || !units)
    return ALPACA_ERR_INVALID;
  std::lock_guard<std::mutex> lock(g_prof_mutex);
  ProfileSlot &p = g_prof[kernel];
  double ms = 0.0;
  for (auto &ev : p.events) {
    if (hipEventSynchronize(ev.second) != hipSuccess) return ALPACA_ERR_HIP;
    float t = 0.f;
    if (hipEventElapsedTime(&t, ev.first, ev.second) != hipSuccess) return ALPACA_ERR_HIP;
    ms += t;
  }
  *total_ms = ms;
  *launches = (long long)p.events.size();
  *units = p.units;
  if (reset) {
    for (auto &ev : p.events) {
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
    p.events.clear();
    p.units = 0.0;
  }
  return ALPACA_OK;
}

size_t alpaca_block_doubles(int dim) {
  if (dim == 2) return Dims<2>::BLOCK;
  if (dim == 3) return Dims<3>::BLOCK;
  return 0;
}

int alpaca_context_create(void *stream, alpaca_context **out) {
  if (!out) return ALPACA_ERR_INVALID;
  *out = nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return ALPACA_ERR_HIP;
  *out = new alpaca_context(dev, static_cast<hipStream_t>(stream));
  return ALPACA_OK;
}

int alpaca_context_destroy(alpaca_context *ctx) {
  if (!ctx) return ALPACA_ERR_INVALID;
  int st = hip_status(hipStreamSynchronize(ctx->stream));
  for (double *p : ctx->retired)
    if (hipFree(p) != hipSuccess) st = ALPACA_ERR_HIP;
  if (ctx->ptr && hipFree(ctx->ptr) != hipSuccess) st = ALPACA_ERR_HIP;
  delete ctx;
  return st;
}

int alpaca_context_reserve(alpaca_context *ctx, int dim, int nblocks) {
  if (!ctx || (dim != 2 && dim != 3) || nblocks < 1 || nblocks > kMaxBlocks)
    return ALPACA_ERR_INVALID;
  double *fl = nullptr;
  size_t const n = dim == 2 ? flux_doubles<2>(nblocks) : flux_doubles<3>(nblocks);
  return context_scratch(ctx, n * sizeof(double), &fl);
}

int alpaca_ctx_convective_fluxes(alpaca_context *ctx, const alpaca_case_desc *desc, int nblocks,
                                 const double *d_avg, double *d_fluxes) {
  int st = validate(desc);
  if (st != ALPACA_OK) return st;
  if (!ctx || nblocks < 1 || nblocks > kMaxBlocks || !d_avg || !d_fluxes) return ALPACA_ERR_INVALID;
  hipStream_t s = ctx->stream;
  double *fl = nullptr;
  if (desc->dim == 2) {
    if ((st = context_scratch(ctx, flux_doubles<2>(nblocks) * sizeof(double), &fl)) != ALPACA_OK)
      return st;
    launch_faces<2>(desc, nblocks, d_avg, fl, s);
    size_t total = (size_t)nblocks * 2 * faces_per_dir<2>();
    hipLaunchKernelGGL(scatter_ref_fluxes<2>, dim3((unsigned)((total + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, s, fl, d_fluxes, nblocks);
  } else {
    if ((st = context_scratch(ctx, flux_doubles<3>(nblocks) * sizeof(double), &fl)) != ALPACA_OK)
      return st;
    launch_faces<3>(desc, nblocks, d_avg, fl, s);
    size_t total = (size_t)nblocks * 3 * faces_per_dir<3>();
    hipLaunchKernelGGL(scatter_ref_fluxes<3>, dim3((unsigned)((total + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, s, fl, d_fluxes, nblocks);
  }
  return hip_status(hipGetLastError());
}

}  // extern "C"

// Face-flux scratch of alpaca_right_hand_side_noalloc: one array per GPU,
// part of the library's code object, so each GPU the library runs on holds
// it (64 MiB) until the process ends.  Host code never passes this name to a
// kernel (on the host it is only a stand-in for the symbol): it uses the
// device address that noalloc_scratch() obtains per device.
constexpr size_t kNoallocScratchDoubles = ALPACA_RHS_NOALLOC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__(256) double scratch[kNoallocScratchDoubles];
static_assert(ALPACA_RHS_NOALLOC_MAX_BLOCKS_3D * alpaca::flux_doubles<3>(1) <= kNoallocScratchDoubles &&
                  (ALPACA_RHS_NOALLOC_MAX_BLOCKS_3D + 1) * alpaca::flux_doubles<3>(1) >
                      kNoallocScratchDoubles,
              "3D block limit of the no-allocation scratch");
static_assert(ALPACA_RHS_NOALLOC_MAX_BLOCKS_2D * alpaca::flux_doubles<2>(1) <= kNoallocScratchDoubles &&
                  (ALPACA_RHS_NOALLOC_MAX_BLOCKS_2D + 1) * alpaca::flux_doubles<2>(1) >
                      kNoallocScratchDoubles,
              "2D block limit of the no-allocation scratch");

namespace alpaca {
namespace {

// Right-hand side of nblocks blocks with the face fluxes in fl: zero the
// output blocks, faces, divergence, and (jump != null) the boundary-flux save.
int rhs_with_scratch(const alpaca_case_desc *desc, int nblocks, const double *d_avg, double *d_rhs,
                     double *d_jump, double *fl, hipStream_t s) {
  size_t const n = alpaca_block_doubles(desc->dim) * (size_t)nblocks;
  unsigned const groups = (unsigned)std::min<size_t>((n + kThreads - 1) / kThreads, 65536);
  hipLaunchKernelGGL(zero_kernel, dim3(groups), dim3(kThreads), 0, s, d_rhs, n);
  if (desc->dim == 2)
    return stage_impl<2>(desc, nblocks, 0, 1, d_avg, nullptr, d_rhs, nullptr, s, fl, d_jump);
  return stage_impl<3>(desc, nblocks, 0, 1, d_avg, nullptr, d_rhs, nullptr, s, fl, d_jump);
}

int rhs_context(alpaca_context *ctx, const alpaca_case_desc *desc, int nblocks,
                const double *d_avg, double *d_rhs, double *d_jump) {
  double *fl = nullptr;
  size_t const n = desc->dim == 2 ? flux_doubles<2>(nblocks) : flux_doubles<3>(nblocks);
  int st = context_scratch(ctx, n * sizeof(double), &fl);
  if (st != ALPACA_OK) return st;
  return rhs_with_scratch(desc, nblocks, d_avg, d_rhs, d_jump, fl, ctx->stream);
}

// Device address of `scratch` on the current device, looked up once per
// device (hipGetSymbolAddress with that device current) and kept.
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mutex;
double *g_noalloc_ptr[kMaxDevices] = {};

int noalloc_scratch(double **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return ALPACA_ERR_HIP;
  if (dev < 0 || dev >= kMaxDevices) return ALPACA_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  if (!g_noalloc_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return ALPACA_ERR_HIP;
    g_noalloc_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_noalloc_ptr[dev];
  return ALPACA_OK;
}

}  // namespace
}  // namespace alpaca

extern "C" {

int alpaca_ctx_right_hand_side(alpaca_context *ctx, const alpaca_case_desc *desc, int nblocks,
                               const double *d_avg, double *d_rhs) {
  int st = validate(desc);
  if (st != ALPACA_OK) return st;
  if (!ctx || nblocks < 1 || nblocks > kMaxBlocks || !d_avg || !d_rhs || d_avg == d_rhs)
    return ALPACA_ERR_INVALID;
  return rhs_context(ctx, desc, nblocks, d_avg, d_rhs, nullptr);
}

int alpaca_ctx_right_hand_side_jump_fluxes(alpaca_context *ctx, const alpaca_case_desc *desc,
                                           int nblocks, const double *d_avg, double *d_rhs,
                                           double *d_jump_fluxes) {
  int st = validate(desc);
  if (st != ALPACA_OK) return st;
  if (!ctx || nblocks < 1 || nblocks > kMaxBlocks || !d_avg || !d_rhs || !d_jump_fluxes ||
      d_avg == d_rhs)
    return ALPACA_ERR_INVALID;
  return rhs_context(ctx, desc, nblocks, d_avg, d_rhs, d_jump_fluxes);
}

int alpaca_ctx_right_hand_side_jump_fluxes_list(alpaca_context *ctx, const alpaca_case_desc *desc,
                                                int capacity, const int *d_list, int n,
                                                const double *d_avg, double *d_rhs,
                                                double *d_jump_fluxes) {
  int st = validate(desc);
  if (st != ALPACA_OK) return st;
  if (!ctx || capacity < 1 || capacity > kMaxBlocks || n < 0 || n > kMaxBlocks || !d_avg ||
      !d_rhs || !d_jump_fluxes || d_avg == d_rhs)
    return ALPACA_ERR_INVALID;
  if (n == 0) return ALPACA_OK;
  if (!d_list) return ALPACA_ERR_INVALID;
  if (desc->dim == 2)
    rhs_tiles<2>(desc, d_avg, d_rhs, d_jump_fluxes, d_list, n, capacity, ctx->stream);
  else
    rhs_tiles<3>(desc, d_avg, d_rhs, d_jump_fluxes, d_list, n, capacity, ctx->stream);
  return hip_status(hipGetLastError());
}

int alpaca_swap_blocks_list(int dim, double *d_a, double *d_b, int capacity, const int *d_list,
                            int n, void *stream) {
  if ((dim != 2 && dim != 3) || !d_a || !d_b || d_a == d_b || capacity < 1 ||
      capacity > kMaxBlocks || n < 0 || n > kMaxBlocks)
    return ALPACA_ERR_INVALID;
  if (n == 0) return ALPACA_OK;
  if (!d_list) return ALPACA_ERR_INVALID;
  launch_listed_blocks<1>(d_a, d_b, dim, BlockList{d_list, capacity}, n,
                          static_cast<hipStream_t>(stream));
  return hip_status(hipGetLastError());
}

int alpaca_copy_blocks_list(int dim, const double *d_src, double *d_dst, int capacity,
                            const int *d_list, int n, void *stream) {
  if ((dim != 2 && dim != 3) || !d_src || !d_dst || d_src == d_dst || capacity < 1 ||
      capacity > kMaxBlocks || n < 0 || n > kMaxBlocks)
    return ALPACA_ERR_INVALID;
  if (n == 0) return ALPACA_OK;
  if (!d_list) return ALPACA_ERR_INVALID;
  launch_listed_blocks<2>(const_cast<double *>(d_src), d_dst, dim, BlockList{d_list, capacity}, n,
                          static_cast<hipStream_t>(stream));
  return hip_status(hipGetLastError());
}

int alpaca_convective_fluxes(const alpaca_case_desc *desc, int nblocks, const double *d_avg,
                             double *d_fluxes, void *stream) {
  int st = validate(desc);
  if (st != ALPACA_OK) return st;
  if (nblocks < 1 || nblocks > kMaxBlocks || !d_avg || !d_fluxes) return ALPACA_ERR_INVALID;
  alpaca_context *ctx = nullptr;
  if ((st = stream_context(static_cast<hipStream_t>(stream), &ctx)) != ALPACA_OK) return st;
  return alpaca_ctx_convective_fluxes(ctx, desc, nblocks, d_avg, d_fluxes);
}

int alpaca_right_hand_side(const alpaca_case_desc *desc, int nblocks, const double *d_avg,
                           double *d_rhs, void *stream) {
  int st = validate(desc);
  if (st != ALPACA_OK) return st;
  if (nblocks < 1 || nblocks > kMaxBlocks || !d_avg || !d_rhs || d_avg == d_rhs)
    return ALPACA_ERR_INVALID;
  alpaca_context *ctx = nullptr;
  if ((st = stream_context(static_cast<hipStream_t>(stream), &ctx)) != ALPACA_OK) return st;
  return rhs_context(ctx, desc, nblocks, d_avg, d_rhs, nullptr);
}

int alpaca_right_hand_side_jump_fluxes(const alpaca_case_desc *desc, int nblocks,
                                       const double *d_avg, double *d_rhs, double *d_jump_fluxes,
                                       void *stream) {
  int st = validate(desc);
  if (st != ALPACA_OK) return st;
  if (nblocks < 1 || nblocks > kMaxBlocks || !d_avg || !d_rhs || !d_jump_fluxes || d_avg == d_rhs)
    return ALPACA_ERR_INVALID;
  alpaca_context *ctx = nullptr;
  if ((st = stream_context(static_cast<hipStream_t>(stream), &ctx)) != ALPACA_OK) return st;
  return rhs_context(ctx, desc, nblocks, d_avg, d_rhs, d_jump_fluxes);
}

int alpaca_right_hand_side_noalloc(const alpaca_case_desc *desc, int nblocks, const double *d_avg,
                                   double *d_rhs, void *stream) {
  int st = validate(desc);
  if (st != ALPACA_OK) return st;
  if (nblocks < 1 || !d_avg || !d_rhs || d_avg == d_rhs) return ALPACA_ERR_INVALID;
  if (nblocks > (desc->dim == 2 ? ALPACA_RHS_NOALLOC_MAX_BLOCKS_2D : ALPACA_RHS_NOALLOC_MAX_BLOCKS_3D))
    return ALPACA_ERR_CAPACITY;
  double *fl = nullptr;
  if ((st = noalloc_scratch(&fl)) != ALPACA_OK) return st;
  g_prof_suppressed = true;
  st = rhs_with_scratch(desc, nblocks, d_avg, d_rhs, nullptr, fl, static_cast<hipStream_t>(stream));
  g_prof_suppressed = false;
  return st;
}

int alpaca_exact_fp64_ops(int op, const double *d_a, const double *d_b, double *d_out, int n,
                          void *stream) {
  if (op < 0 || op > 4 || n < 0 || n > kMaxBlocks * 64 || !d_a || !d_out ||
      ((op < 2 || op == 3) && !d_b))
    return ALPACA_ERR_INVALID;
  if (n == 0) return ALPACA_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  dim3 const grid((unsigned)((n + kThreads - 1) / kThreads));
  if (op == 0)
    hipLaunchKernelGGL(exact_ops_kernel<0>, grid, dim3(kThreads), 0, s, d_a, d_b, d_out, n);
  else if (op == 1)
    hipLaunchKernelGGL(exact_ops_kernel<1>, grid, dim3(kThreads), 0, s, d_a, d_b, d_out, n);
  else if (op == 2)
    hipLaunchKernelGGL(exact_ops_kernel<2>, grid, dim3(kThreads), 0, s, d_a, d_a, d_out, n);
  else if (op == 3)
    hipLaunchKernelGGL(exact_ops_kernel<3>, grid, dim3(kThreads), 0, s, d_a, d_b, d_out, n);
  else
    hipLaunchKernelGGL(exact_ops_kernel<4>, grid, dim3(kThreads), 0, s, d_a, d_a, d_out, n);
  return hip_status(hipGetLastError());
}

int alpaca_rk3_stage(const alpaca_case_desc *desc, int stage, const double *d_avg, double *d_init,
                     double *d_out, alpaca_case_state *d_state, void *stream) {
  int st = validate(desc);
  if (st != ALPACA_OK) return st;
  if (stage < 0 || stage > 2 || !d_avg || !d_out || !d_state || d_out == d_avg)
    return ALPACA_ERR_INVALID;
  if (stage > 0 && !d_init) return ALPACA_ERR_INVALID;
  hipStream_t s = static_cast<hipStream_t>(stream);
  double *copy = (stage == 0 && d_init && d_init != d_avg) ? d_init : nullptr;
  if (desc->dim == 2)
    stage_fused<2>(desc, stage, d_avg, d_init, d_out, copy, d_state, 0, s);
  else
    stage_fused<3>(desc, stage, d_avg, d_init, d_out, copy, d_state, 0, s);
  return hip_status(hipGetLastError());
}

int alpaca_fill_halos(const alpaca_case_desc *desc, double *d_blocks, void *stream) {
  int st = validate(desc);
  if (st != ALPACA_OK) return st;
  if (!d_blocks) return ALPACA_ERR_INVALID;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (desc->dim == 2) return halo_impl<2>(desc, d_blocks, s);
  return halo_impl<3>(desc, d_blocks, s);
}

int alpaca_fill_halos_table(int dim, double *d_blocks, int nblocks, const int *d_neighbors,
                            const int *side_bc, void *stream) {
  if ((dim != 2 && dim != 3) || !d_blocks || !d_neighbors || !side_bc || nblocks < 1 ||
      nblocks > kMaxBlocks)
    return ALPACA_ERR_INVALID;
  SideBc bc;
  for (int i = 0; i < 6; ++i) {
    bc.kind[i] = i < 2 * dim ? side_bc[i] : ALPACA_BC_ZERO_GRADIENT;
    if (bc.kind[i] != ALPACA_BC_ZERO_GRADIENT && bc.kind[i] != ALPACA_BC_SYMMETRY &&
        bc.kind[i] != ALPACA_BC_PERIODIC)
      return ALPACA_ERR_INVALID;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dim == 2) return halo_table_impl<2>(d_blocks, nblocks, d_neighbors, bc, s);
  return halo_table_impl<3>(d_blocks, nblocks, d_neighbors, bc, s);
}

int alpaca_fill_halos_list(int dim, double *d_blocks, int nblocks, const int *d_list, int n,
                           const int *d_neighbors, const int *side_bc, int mode, void *stream) {
  if ((dim != 2 && dim != 3) || !d_blocks || !side_bc || nblocks < 1 || nblocks > kMaxBlocks ||
      n < 0 || n > kMaxBlocks || (mode != 0 && mode != 1))
    return ALPACA_ERR_INVALID;
  if (n == 0) return ALPACA_OK;
  if (!d_list || !d_neighbors) return ALPACA_ERR_INVALID;
  SideBc bc;
  for (int i = 0; i < 6; ++i) {
    bc.kind[i] = i < 2 * dim ? side_bc[i] : ALPACA_BC_ZERO_GRADIENT;
    if (bc.kind[i] != ALPACA_BC_ZERO_GRADIENT && bc.kind[i] != ALPACA_BC_SYMMETRY &&
        bc.kind[i] != ALPACA_BC_PERIODIC)
      return ALPACA_ERR_INVALID;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  size_t const cells = (size_t)kTC * kTC * (dim == 3 ? kTC : 1);
  dim3 grid((unsigned)(((size_t)n * cells + kThreads - 1) / kThreads));
  ProfileScope prof(ALPACA_KERNEL_HALO, s, double(n) * cells);
  if (dim == 2)
    hipLaunchKernelGGL(halo_list_kernel<2>, grid, dim3(kThreads), 0, s, d_blocks, nblocks, d_list,
                       n, d_neighbors, bc, mode);
  else
    hipLaunchKernelGGL(halo_list_kernel<3>, grid, dim3(kThreads), 0, s, d_blocks, nblocks, d_list,
                       n, d_neighbors, bc, mode);
  return hip_status(hipGetLastError());
}

int alpaca_timestep(const alpaca_case_desc *desc, const double *d_avg, alpaca_case_state *d_state,
                    void *stream) {
  int st = validate(desc);
  if (st != ALPACA_OK) return st;
  if (!d_avg || !d_state) return ALPACA_ERR_INVALID;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (desc->dim == 2) return timestep_impl<2>(desc, d_avg, d_state, s, case_blocks(desc));
  return timestep_impl<3>(desc, d_avg, d_state, s, case_blocks(desc));
}

int alpaca_timestep_list(const alpaca_case_desc *desc, const double *d_avg, int capacity,
                         const int *d_list, int n, alpaca_case_state *d_state, void *stream) {
  int st = validate(desc);
  if (st != ALPACA_OK) return st;
  if (!d_avg || !d_state || !d_list || capacity < 1 || capacity > kMaxBlocks || n < 1 ||
      n > kMaxBlocks)
    return ALPACA_ERR_INVALID;
  hipStream_t s = static_cast<hipStream_t>(stream);
  BlockList const list{d_list, capacity};
  if (desc->dim == 2) return timestep_impl<2>(desc, d_avg, d_state, s, n, list);
  return timestep_impl<3>(desc, d_avg, d_state, s, n, list);
}

int alpaca_finish_step(alpaca_case_state *d_state, void *stream) {
  if (!d_state) return ALPACA_ERR_INVALID;
  hipLaunchKernelGGL(finish_step_kernel, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream),
                     d_state);
  return hip_status(hipGetLastError());
}

int alpaca_case_state_init(alpaca_case_state *d_state, double time, double t_end, double cfl,
                           void *stream) {
  if (!d_state) return ALPACA_ERR_INVALID;
  hipLaunchKernelGGL(state_init_kernel, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream),
                     d_state, time, t_end, cfl);
  return hip_status(hipGetLastError());
}

int alpaca_advance(const alpaca_case_desc *desc, double *d_u, double *d_s1, double *d_s2,
                   alpaca_case_state *d_state, int nsteps, void *stream) {
  int st = validate(desc);
  if (st != ALPACA_OK) return st;
  if (!d_u || !d_s1 || !d_s2 || !d_state || d_u == d_s1 || d_u == d_s2 || d_s1 == d_s2 ||
      nsteps < 0)
    return ALPACA_ERR_INVALID;
  hipStream_t s = static_cast<hipStream_t>(stream);
  // The first step reduces the CFL signal of d_u; later steps take it from the
  // last stage of the step before.
  if (nsteps > 0) {
    st = desc->dim == 2 ? timestep_impl<2>(desc, d_u, d_state, s, case_blocks(desc), {}, 1)
                        : timestep_impl<3>(desc, d_u, d_state, s, case_blocks(desc), {}, 1);
    if (st != ALPACA_OK) return st;
  }
  // Steps once the record is done (its time reached t_end) do nothing: a run
  // can be advanced in chunks of steps with one read-back per chunk.
  for (int n = 0; n < nsteps; ++n) {
    bool const signal = n + 1 < nsteps;
    if (desc->dim == 2) {
      stage_fused<2>(desc, 0, d_u, nullptr, d_s1, nullptr, d_state, 0, s, true);
      stage_fused<2>(desc, 1, d_s1, d_u, d_s2, nullptr, d_state, 0, s, true);
      stage_fused<2>(desc, 2, d_s2, d_u, d_u, nullptr, d_state, signal, s, true);
    } else {
      stage_fused<3>(desc, 0, d_u, nullptr, d_s1, nullptr, d_state, 0, s, true);
      stage_fused<3>(desc, 1, d_s1, d_u, d_s2, nullptr, d_state, 0, s, true);
      stage_fused<3>(desc, 2, d_s2, d_u, d_u, nullptr, d_state, signal, s, true);
    }
    if (signal)  // this step's end and the next step's dt
      hipLaunchKernelGGL(finish_and_dt_kernel, dim3(1), dim3(64), 0, s, d_state, desc->cell_size);
    else
      hipLaunchKernelGGL(finish_step_kernel, dim3(1), dim3(64), 0, s, d_state, 1);
  }
  return hip_status(hipGetLastError());
}

}  // extern "C"
