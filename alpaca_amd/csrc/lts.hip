// Level-wise (local time stepping) integration of ALPACA's
// TimeIntegrator<RungeKutta3TVD> on gfx950: the per-node operations a
// multi-level step applies to the leaves and jump halos of one level, over a
// list of pool blocks per launch.
//
//   IntegrateConservatives / IntegrateHalo (time_integrator.h:98-153):
//       u_new = u_old + timestep * u_new on the internal cells or a halo range;
//   IntegrateJumpConservatives (time_integrator.h:64-84):
//       jump conservatives += timestep * jump fluxes, jump fluxes = 0;
//   PrepareBufferForIntegration (time_integrator.h:434-457):
//       avg = m0 * avg + m1 * init for stages 1 and 2.
//
// The level's time step -- the last n micro time steps summed from the newest,
// times the stage multiplier (time_integrator.h:281-301, 312-340) -- is host
// arithmetic (alpaca_lts_timestep).  Every expression keeps the reference's
// operation order; the translation unit is built without contraction.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fv_device.h"

namespace alpaca {
namespace lts {
namespace {

constexpr int kThreads = 256;

struct Box {
  int start[3];
  int count[3];
};

__device__ __forceinline__ bool block_ok(int b, int nblocks) { return b >= 0 && b < nblocks; }

// u_old may be u_new (no restrict): each element is read, then written.
// INTERNAL: the box is the internal cells (its extents known at compile time).
template <int DIM, bool INTERNAL>
__global__ __launch_bounds__(kThreads) void integrate_kernel(const double *u_old,
                                                             double *u_new, int nblocks,
                                                             const int *__restrict__ list, int n,
                                                             Box box, double timestep) {
  using D = Dims<DIM>;
  if constexpr (INTERNAL) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      box.start[a] = a < DIM ? kHS : 0;
      box.count[a] = a < DIM ? kIC : 1;
    }
  }
  int const cells = box.count[0] * box.count[1] * box.count[2];
  size_t const total = (size_t)n * D::NE * cells;
  for (size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x; t < total;
       t += (size_t)gridDim.x * kThreads) {
    int const b = list[t / ((size_t)D::NE * cells)];
    if (!block_ok(b, nblocks)) continue;
    int const r = int(t % cells);
    int const e = int((t / cells) % D::NE);
    int const i = box.start[0] + r / (box.count[1] * box.count[2]);
    int const j = box.start[1] + (r / box.count[2]) % box.count[1];
    int const k = box.start[2] + r % box.count[2];
    size_t const c = (size_t)b * D::BLOCK + (size_t)e * D::FIELD + i * D::SX + j * D::SY + k;
    u_new[c] = u_old[c] + timestep * u_new[c];
  }
}

// IntegrateJumpHalos (time_integrator.h:281-301) for every (node, location)
// pair of a level at once: entries[i] = {pool block, BoundaryLocation}, and
// the location's halo slab (GetStartIndicesHaloRecv / GetHaloSize,
// communication_types.h:76-132,190-197: per axis [IC+HS, TC), [0, HS) or
// [HS, HS+IC) for direction +1, -1, 0) takes u_new = u_old + timestep * u_new.
// The slabs of distinct locations are disjoint, so the launch order of the
// reference's per-location calls does not matter.  One workgroup per entry
// (grid-stride); entries with a block outside the pool, a location outside
// the enum or, in 2D, a Top/Bottom part are skipped.
template <int DIM>
__global__ __launch_bounds__(kThreads) void jump_halo_integrate_kernel(
    const double *u_old, double *u_new, int nblocks, const int *__restrict__ entries, int n,
    double timestep) {
  using D = Dims<DIM>;
  for (int q = blockIdx.x; q < n; q += gridDim.x) {
    int const b = entries[2 * q];
    int const loc = entries[2 * q + 1];
    if (!block_ok(b, nblocks) || loc < 0 || loc >= 26) continue;  // uniform over the group
    Box box;
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      int const d = kLocationDir[loc][a];
      if (a >= DIM) {
        ok = ok && d == 0;
        box.start[a] = 0;
        box.count[a] = 1;
      } else {
        box.start[a] = d > 0 ? kIC + kHS : (d < 0 ? 0 : kHS);
        box.count[a] = d != 0 ? kHS : kIC;
      }
    }
    if (!ok) continue;
    int const cells = box.count[0] * box.count[1] * box.count[2];
    for (int t = threadIdx.x; t < D::NE * cells; t += kThreads) {
      int const r = t % cells;
      int const e = t / cells;
      int const i = box.start[0] + r / (box.count[1] * box.count[2]);
      int const j = box.start[1] + (r / box.count[2]) % box.count[1];
      int const k = box.start[2] + r % box.count[2];
      size_t const c = (size_t)b * D::BLOCK + (size_t)e * D::FIELD + i * D::SX + j * D::SY + k;
      u_new[c] = u_old[c] + timestep * u_new[c];
    }
  }
}

template <int DIM>
__global__ __launch_bounds__(kThreads) void jump_kernel(double *__restrict__ cons,
                                                        double *__restrict__ flux, int nblocks,
                                                        const int *__restrict__ list, int n,
                                                        double timestep) {
  constexpr int ICZ = DIM == 3 ? kIC : 1;
  constexpr int FACE = Dims<DIM>::NE * kIC * ICZ;
  constexpr int LIVE = 2 * DIM * FACE;  // the faces CC::ANBS() names
  constexpr int BLOCK = 6 * FACE;       // SurfaceBuffer: six faces
  size_t const total = (size_t)n * LIVE;
  for (size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x; t < total;
       t += (size_t)gridDim.x * kThreads) {
    int const b = list[t / LIVE];
    if (!block_ok(b, nblocks)) continue;
    size_t const c = (size_t)b * BLOCK + t % LIVE;
    cons[c] += timestep * flux[c];
    flux[c] = 0.0;
  }
}

template <int DIM>
__global__ __launch_bounds__(kThreads) void prepare_kernel(double *__restrict__ avg,
                                                           const double *__restrict__ init,
                                                           int nblocks, const int *__restrict__ list,
                                                           int n, double m0, double m1) {
  using D = Dims<DIM>;
  size_t const total = (size_t)n * D::BLOCK;
  for (size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x; t < total;
       t += (size_t)gridDim.x * kThreads) {
    int const b = list[t / D::BLOCK];
    if (!block_ok(b, nblocks)) continue;
    size_t const c = (size_t)b * D::BLOCK + t % D::BLOCK;
    avg[c] = m0 * avg[c] + m1 * init[c];
  }
}

// ConsistencyManagedSum of three values (mathematical_functions.h:79-91) with
// std::min/std::max picks (NaN never chosen by a comparison).
__device__ __forceinline__ double cms3_std(double a, double b, double c) {
  double const s1 = (a + b) + c, s2 = (a + c) + b, s3 = (b + c) + a;
  double mn = s1, mx = s1;
  mn = (s2 < mn) ? s2 : mn;
  mn = (s3 < mn) ? s3 : mn;
  mx = (mx < s2) ? s2 : mx;
  mx = (mx < s3) ? s3 : mx;
  return 0.5 * (mn + mx);
}

// JumpFluxAdjustment's exchange (modular_algorithm_assembler.cpp:1410-1620):
// one thread per (leaf, equation, internal cell).  entries[l] = {leaf block,
// neighbour block per face location East, West, North, South, Top, Bottom or
// -1}: a neighbour that exists and is not a leaf.  Boundary-layer cells take
// c_L = J_leaf[L] * s * dir_L and f_L = J_nb[opposite L] * s * dir_L (dir -1 on
// East/North/Top, +1 on West/South/Bottom), and every internal cell becomes
// (u - CMS(c_E + c_W, c_N + c_S, c_T + c_B)) + CMS(f_E + f_W, f_N + f_S, f_T + f_B).
template <int DIM>
__global__ __launch_bounds__(kThreads) void adjust_kernel(double *__restrict__ cells,
                                                          const double *__restrict__ jump,
                                                          int nblocks, const int *__restrict__ entries,
                                                          int n, double one_cell_size) {
  using D = Dims<DIM>;
  constexpr int ICZ = D::ICZ;
  constexpr int FACE = D::NE * kIC * ICZ;
  constexpr int SURF = 6 * FACE;
  size_t const total = (size_t)n * D::NE * D::CELLS;
  for (size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x; t < total;
       t += (size_t)gridDim.x * kThreads) {
    const int *en = entries + (t / ((size_t)D::NE * D::CELLS)) * 7;
    int const leaf = en[0];
    if (!block_ok(leaf, nblocks)) continue;
    int const r = int(t % D::CELLS);
    int const e = int((t / D::CELLS) % D::NE);
    int const x[3] = {r / (kIC * ICZ), (r / ICZ) % kIC, r % ICZ};
    double c[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, f[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int loc = 0; loc < 2 * DIM; ++loc) {
      int const nb = en[1 + loc];
      int const axis = loc / 2;
      bool const high = (loc % 2) == 0;
      if (!block_ok(nb, nblocks) || x[axis] != (high ? kIC - 1 : 0)) continue;
      int const t1 = axis == 0 ? 1 : 0, t2 = axis == 2 ? 1 : 2;
      int const idx = (e * kIC + x[t1]) * ICZ + (DIM == 3 ? x[t2] : 0);
      double const dir = high ? -1.0 : 1.0;
      int const opp = high ? loc + 1 : loc - 1;
      c[loc] = jump[(size_t)leaf * SURF + loc * FACE + idx] * one_cell_size * dir;
      f[loc] = jump[(size_t)nb * SURF + opp * FACE + idx] * one_cell_size * dir;
    }
    double *u = cells + (size_t)leaf * D::BLOCK + (size_t)e * D::FIELD + (x[0] + kHS) * D::SX +
                (x[1] + kHS) * D::SY + (DIM == 3 ? x[2] + kHS : 0);
    double v = *u;
    v -= cms3_std(c[0] + c[1], c[2] + c[3], c[4] + c[5]);
    v += cms3_std(f[0] + f[1], f[2] + f[3], f[4] + f[5]);
    *u = v;
  }
}

// The leaf's jump surface at each exchanged face takes the neighbour's
// opposite one (after adjust_kernel has read it).
template <int DIM>
__global__ __launch_bounds__(kThreads) void adjust_copy_kernel(double *__restrict__ jump, int nblocks,
                                                               const int *__restrict__ entries, int n) {
  constexpr int FACE = Dims<DIM>::NE * kIC * (DIM == 3 ? kIC : 1);
  constexpr int SURF = 6 * FACE;
  size_t const total = (size_t)n * 2 * DIM * FACE;
  for (size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x; t < total;
       t += (size_t)gridDim.x * kThreads) {
    const int *en = entries + (t / ((size_t)2 * DIM * FACE)) * 7;
    int const loc = int((t / FACE) % (2 * DIM));
    int const leaf = en[0], nb = en[1 + loc];
    if (!block_ok(leaf, nblocks) || !block_ok(nb, nblocks)) continue;
    int const opp = (loc % 2) == 0 ? loc + 1 : loc - 1;
    jump[(size_t)leaf * SURF + loc * FACE + t % FACE] = jump[(size_t)nb * SURF + opp * FACE + t % FACE];
  }
}

// ResetJumpConservatives on the faces CC::ANBS() names.
template <int DIM>
__global__ __launch_bounds__(kThreads) void reset_jump_kernel(double *__restrict__ jump, int nblocks,
                                                              const int *__restrict__ list, int n) {
  constexpr int FACE = Dims<DIM>::NE * kIC * (DIM == 3 ? kIC : 1);
  constexpr int LIVE = 2 * DIM * FACE;
  size_t const total = (size_t)n * LIVE;
  for (size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x; t < total;
       t += (size_t)gridDim.x * kThreads) {
    int const b = list[t / LIVE];
    if (!block_ok(b, nblocks)) continue;
    jump[(size_t)b * 6 * FACE + t % LIVE] = 0.0;
  }
}

int hip_status(hipError_t e) { return e == hipSuccess ? ALPACA_OK : ALPACA_ERR_HIP; }

bool bad_sizes(int dim, int nblocks, int n) {
  return (dim != 2 && dim != 3) || nblocks < 1 || nblocks > kMaxBlocks || n < 0 || n > kMaxBlocks;
}

// Grid of a grid-stride launch: one thread per element up to 2^22 groups.
dim3 grid_for(size_t total) {
  size_t const groups = (total + kThreads - 1) / kThreads;
  return dim3((unsigned)(groups < (size_t(1) << 22) ? groups : (size_t(1) << 22)));
}

}  // namespace
}  // namespace lts
}  // namespace alpaca

using namespace alpaca;

extern "C" {

int alpaca_lts_timestep(const double *micro, int count, int n, int stage, int jump, double *out) {
  // RungeKutta3TVD multipliers (runge_kutta_3_TVD.h:42-53).
  static const double mc[3] = {1.0, 0.25, 2.0 / 3.0};
  static const double mj[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
  if (!out || count < 0 || n < 0 || n > count || (count > 0 && !micro) || stage < 0 || stage > 2)
    return ALPACA_ERR_INVALID;
  double s = 0.0;
  for (int i = 0; i < n; ++i) s = s + micro[count - 1 - i];
  *out = s * (jump ? mj[stage] : mc[stage]);
  return ALPACA_OK;
}

int alpaca_lts_integrate(int dim, const double *d_old, double *d_new, int nblocks, const int *d_list,
                         int n, const int *range, double timestep, void *stream) {
  if (lts::bad_sizes(dim, nblocks, n) || !d_old || !d_new || (n > 0 && !d_list))
    return ALPACA_ERR_INVALID;
  lts::Box box{{kHS, kHS, dim == 3 ? kHS : 0}, {kIC, kIC, dim == 3 ? kIC : 1}};
  if (range) {
    int const tz = dim == 3 ? kTC : 1;
    for (int a = 0; a < 3; ++a) {
      int const start = a < dim ? range[2 * a] : 0, count = a < dim ? range[2 * a + 1] : 1;
      int const extent = a < 2 ? kTC : tz;
      if (start < 0 || count < 0 || start + count > extent) return ALPACA_ERR_INVALID;
      box.start[a] = start;
      box.count[a] = count;
    }
  }
  size_t const cells = (size_t)box.count[0] * box.count[1] * box.count[2];
  if (n == 0 || cells == 0) return ALPACA_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  size_t const total = (size_t)n * (dim + 2) * cells;
  dim3 const grid = lts::grid_for(total), block(lts::kThreads);
  if (dim == 2) {
    if (range)
      hipLaunchKernelGGL((lts::integrate_kernel<2, false>), grid, block, 0, s, d_old, d_new, nblocks,
                         d_list, n, box, timestep);
    else
      hipLaunchKernelGGL((lts::integrate_kernel<2, true>), grid, block, 0, s, d_old, d_new, nblocks,
                         d_list, n, box, timestep);
  } else {
    if (range)
      hipLaunchKernelGGL((lts::integrate_kernel<3, false>), grid, block, 0, s, d_old, d_new, nblocks,
                         d_list, n, box, timestep);
    else
      hipLaunchKernelGGL((lts::integrate_kernel<3, true>), grid, block, 0, s, d_old, d_new, nblocks,
                         d_list, n, box, timestep);
  }
  return lts::hip_status(hipGetLastError());
}

int alpaca_lts_integrate_jump_halos(int dim, const double *d_old, double *d_new, int nblocks,
                                    const int *d_entries, int n, double timestep, void *stream) {
  if (lts::bad_sizes(dim, nblocks, n) || !d_old || !d_new || (n > 0 && !d_entries))
    return ALPACA_ERR_INVALID;
  if (n == 0) return ALPACA_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  dim3 const grid((unsigned)std::min(n, 65536));
  if (dim == 2)
    hipLaunchKernelGGL((lts::jump_halo_integrate_kernel<2>), grid, dim3(lts::kThreads), 0, s, d_old,
                       d_new, nblocks, d_entries, n, timestep);
  else
    hipLaunchKernelGGL((lts::jump_halo_integrate_kernel<3>), grid, dim3(lts::kThreads), 0, s, d_old,
                       d_new, nblocks, d_entries, n, timestep);
  return lts::hip_status(hipGetLastError());
}

int alpaca_lts_integrate_jump_conservatives(int dim, double *d_cons, double *d_flux, int nblocks,
                                            const int *d_list, int n, double timestep,
                                            void *stream) {
  if (lts::bad_sizes(dim, nblocks, n) || !d_cons || !d_flux || d_cons == d_flux ||
      (n > 0 && !d_list))
    return ALPACA_ERR_INVALID;
  if (n == 0) return ALPACA_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  size_t const live = (size_t)2 * dim * (dim + 2) * kIC * (dim == 3 ? kIC : 1);
  if (dim == 2)
    hipLaunchKernelGGL((lts::jump_kernel<2>), lts::grid_for(n * live), dim3(lts::kThreads), 0, s,
                       d_cons, d_flux, nblocks, d_list, n, timestep);
  else
    hipLaunchKernelGGL((lts::jump_kernel<3>), lts::grid_for(n * live), dim3(lts::kThreads), 0, s,
                       d_cons, d_flux, nblocks, d_list, n, timestep);
  return lts::hip_status(hipGetLastError());
}

int alpaca_lts_prepare_buffer(int dim, double *d_avg, const double *d_init, int nblocks,
                              const int *d_list, int n, int stage, void *stream) {
  if (lts::bad_sizes(dim, nblocks, n) || !d_avg || !d_init || stage < 0 || stage > 2 ||
      (n > 0 && !d_list))
    return ALPACA_ERR_INVALID;
  if (n == 0 || stage == 0) return ALPACA_OK;  // only later stages combine the buffers
  // GetBufferMultiplier (runge_kutta_3_TVD.h:55-59).
  double const m0 = stage == 1 ? 0.25 : 2.0 / 3.0;
  double const m1 = stage == 1 ? 0.75 : 1.0 / 3.0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  size_t const per = (size_t)(dim + 2) * kTC * kTC * (dim == 3 ? kTC : 1);
  if (dim == 2)
    hipLaunchKernelGGL((lts::prepare_kernel<2>), lts::grid_for(n * per), dim3(lts::kThreads), 0, s,
                       d_avg, d_init, nblocks, d_list, n, m0, m1);
  else
    hipLaunchKernelGGL((lts::prepare_kernel<3>), lts::grid_for(n * per), dim3(lts::kThreads), 0, s,
                       d_avg, d_init, nblocks, d_list, n, m0, m1);
  return lts::hip_status(hipGetLastError());
}

int alpaca_lts_jump_flux_adjustment(int dim, double *d_cells, double *d_jump_cons, int nblocks,
                                    const int *d_entries, int n, double one_cell_size,
                                    void *stream) {
  if (lts::bad_sizes(dim, nblocks, n) || !d_cells || !d_jump_cons || (n > 0 && !d_entries))
    return ALPACA_ERR_INVALID;
  if (n == 0) return ALPACA_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  size_t const cells = (size_t)n * (dim + 2) * kIC * kIC * (dim == 3 ? kIC : 1);
  size_t const faces = (size_t)n * 2 * dim * (dim + 2) * kIC * (dim == 3 ? kIC : 1);
  if (dim == 2) {
    hipLaunchKernelGGL((lts::adjust_kernel<2>), lts::grid_for(cells), dim3(lts::kThreads), 0, s,
                       d_cells, d_jump_cons, nblocks, d_entries, n, one_cell_size);
    hipLaunchKernelGGL((lts::adjust_copy_kernel<2>), lts::grid_for(faces), dim3(lts::kThreads), 0,
                       s, d_jump_cons, nblocks, d_entries, n);
  } else {
    hipLaunchKernelGGL((lts::adjust_kernel<3>), lts::grid_for(cells), dim3(lts::kThreads), 0, s,
                       d_cells, d_jump_cons, nblocks, d_entries, n, one_cell_size);
    hipLaunchKernelGGL((lts::adjust_copy_kernel<3>), lts::grid_for(faces), dim3(lts::kThreads), 0,
                       s, d_jump_cons, nblocks, d_entries, n);
  }
  return lts::hip_status(hipGetLastError());
}

int alpaca_lts_reset_jump_conservatives(int dim, double *d_jump_cons, int nblocks,
                                        const int *d_list, int n, void *stream) {
  if (lts::bad_sizes(dim, nblocks, n) || !d_jump_cons || (n > 0 && !d_list))
    return ALPACA_ERR_INVALID;
  if (n == 0) return ALPACA_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  size_t const live = (size_t)n * 2 * dim * (dim + 2) * kIC * (dim == 3 ? kIC : 1);
  if (dim == 2)
    hipLaunchKernelGGL((lts::reset_jump_kernel<2>), lts::grid_for(live), dim3(lts::kThreads), 0, s,
                       d_jump_cons, nblocks, d_list, n);
  else
    hipLaunchKernelGGL((lts::reset_jump_kernel<3>), lts::grid_for(live), dim3(lts::kThreads), 0, s,
                       d_jump_cons, nblocks, d_list, n);
  return lts::hip_status(hipGetLastError());
}

}  // extern "C"
