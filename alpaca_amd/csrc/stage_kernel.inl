// Fused SSP-RK3 stage kernel, included into alpaca::exact and alpaca::fast
// (see fv_device.h).
//
// One workgroup per (leaf block, tile of TX x TY x TZ internal cells):
//   1. every face of the tile (all directions) is computed from the haloed
//      Average buffer — a flat loop whose per-direction ranges are padded to
//      64 so each wavefront works on a single direction — and its flux is
//      kept in LDS (in R: FP32 in the mixed form);
//   2. after one barrier each thread forms the flux divergence of its cells
//      in FP64 with the reference's ConsistencyManagedSum (space_solver.cpp:
//      131-153), applies the RK combination in FP64 (time_integrator.h:98-113,
//      434-457), writes the new state and pushes it into the face halo slabs
//      of the neighbouring blocks (internal_halo_manager.cpp:439-466) or, on a
//      zero-gradient or symmetry side, into its own halo
//      (boundary_constants.h:169-253);
//   3. on the last stage it also reduces the CFL signal of the new state
//      (modular_algorithm_assembler.cpp:1840-1869) for the next time step.

__device__ __forceinline__ double wave_max64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = smax(v, __shfl_xor(v, o, 64));
  return v;
}

// Writes the new value v of equation e of internal cell c3 of block blk of the
// job's case into the halo slabs that read it: the neighbours' slabs across
// each near face, or the block's own slab on a zero-gradient side (the first
// or last internal layer fills all four halo layers) or a symmetry side (the
// cell at internal depth c fills the halo layer at depth c, the face-normal
// momentum negated).
template <int DIM>
__device__ __forceinline__ void push_halo(const TileJob &j, int blk, const int *bc3, const int *c3,
                                          int e, double v) {
  using D = Dims<DIM>;
  double *__restrict__ out = j.out;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    int const c = c3[d];
    int pos[3] = {c3[0] + kHS, c3[1] + kHS, DIM == 3 ? c3[2] + kHS : 0};
    double const mirrored = (e == kMomX + d) ? -1.0 * v : v;
    if (c < kHS) {
      if (bc3[d] > 0 || j.bc[d] == ALPACA_BC_PERIODIC) {
        int nb3[3] = {bc3[0], bc3[1], bc3[2]};
        nb3[d] = bc3[d] > 0 ? bc3[d] - 1 : j.nb[d] - 1;
        int const nbk = j.base + (nb3[0] * j.nb[1] + nb3[1]) * j.nb[2] + nb3[2];
        pos[d] = c + kHS + kIC;
        out[(size_t)nbk * D::BLOCK + e * D::FIELD + pos[0] * D::SX + pos[1] * D::SY + pos[2]] = v;
      } else if (j.bc[d] == ALPACA_BC_SYMMETRY) {
        pos[d] = kHS - 1 - c;
        out[(size_t)blk * D::BLOCK + e * D::FIELD + pos[0] * D::SX + pos[1] * D::SY + pos[2]] =
            mirrored;
      } else if (c == 0) {
#pragma unroll
        for (int h = 0; h < kHS; ++h) {
          pos[d] = h;
          out[(size_t)blk * D::BLOCK + e * D::FIELD + pos[0] * D::SX + pos[1] * D::SY + pos[2]] = v;
        }
      }
    }
    if (c >= kIC - kHS) {
      if (bc3[d] < j.nb[d] - 1 || j.bc[d] == ALPACA_BC_PERIODIC) {
        int nb3[3] = {bc3[0], bc3[1], bc3[2]};
        nb3[d] = bc3[d] < j.nb[d] - 1 ? bc3[d] + 1 : 0;
        int const nbk = j.base + (nb3[0] * j.nb[1] + nb3[1]) * j.nb[2] + nb3[2];
        pos[d] = c - kIC + kHS;
        out[(size_t)nbk * D::BLOCK + e * D::FIELD + pos[0] * D::SX + pos[1] * D::SY + pos[2]] = v;
      } else if (j.bc[d] == ALPACA_BC_SYMMETRY) {
        pos[d] = kHS + 2 * kIC - 1 - c;
        out[(size_t)blk * D::BLOCK + e * D::FIELD + pos[0] * D::SX + pos[1] * D::SY + pos[2]] =
            mirrored;
      } else if (c == kIC - 1) {
#pragma unroll
        for (int h = 0; h < kHS; ++h) {
          pos[d] = kHS + kIC + h;
          out[(size_t)blk * D::BLOCK + e * D::FIELD + pos[0] * D::SX + pos[1] * D::SY + pos[2]] = v;
        }
      }
    }
  }
}

template <int DIM, int TX, int TY, int TZ> struct TileShape {
  static constexpr int TZe = DIM == 3 ? TZ : 1;
  static constexpr int NTX = kIC / TX, NTY = kIC / TY, NTZ = DIM == 3 ? kIC / TZ : 1;
  static constexpr int TILES = NTX * NTY * NTZ;
  static constexpr int NFX = (TX + 1) * TY * TZe;
  static constexpr int NFY = TX * (TY + 1) * TZe;
  static constexpr int NFZ = DIM == 3 ? TX * TY * (TZ + 1) : 0;
  static constexpr int PX = round_up64(NFX), PY = round_up64(NFY), PZ = round_up64(NFZ);
  static constexpr int CELLS = TX * TY * TZe;
  static constexpr int LDS = (DIM + 2) * (NFX + NFY + NFZ);  // reals
};

// Face fluxes of one tile (first cell x0, y0, z0 of block bb) into LDS:
// every face of the tile, a flat loop whose per-direction ranges are padded
// to 64 so each wavefront works on a single direction.  JUMP: a face on the
// block's boundary is also added into the block's jump-flux surfaces
// ([6][NE][IC][ICZ]: East, West, North, South, Top, Bottom; the boundary-flux
// save of SpaceSolver::UpdateFluxes, space_solver.cpp:160-182).
template <int DIM, int FLUX, typename R, int TX, int TY, int TZ, int NT, bool JUMP>
__device__ __forceinline__ void tile_faces(const double *__restrict__ bb, int x0, int y0, int z0,
                                           double gamma, double b, R *lds, double *jump) {
  using D = Dims<DIM>;
  using T = TileShape<DIM, TX, TY, TZ>;
  constexpr int NE = D::NE;
  constexpr int TZe = T::TZe;
  constexpr int NFX = T::NFX, NFY = T::NFY, NFZ = T::NFZ;
  constexpr int PX = T::PX, PY = T::PY, PZ = T::PZ;
  constexpr int ICZ = D::ICZ;
  R *FX = lds;
  R *FY = lds + NE * NFX;
  R *FZ = FY + NE * NFY;
  int const zh = DIM == 3 ? kHS : 0;  // halo offset along z
  // Adds a boundary face's flux into surface `loc` at entry (a, c).
  auto save = [jump](int loc, int a, int c, const R *fl) {
#pragma unroll
    for (int e = 0; e < NE; ++e) jump[((loc * NE + e) * kIC + a) * ICZ + c] += double(fl[e]);
  };
  for (int f = threadIdx.x; f < PX + PY + PZ; f += NT) {
    R fl[NE];
    if (f < PX) {
      if (f < NFX) {
        int const k = f % TZe, jj = (f / TZe) % TY, ia = f / (TZe * TY);
        stencil_flux<DIM, 0, FLUX, R>(bb, x0 + ia + kHS - 1, y0 + jj + kHS, z0 + k + zh, gamma, b,
                                      fl);
#pragma unroll
        for (int e = 0; e < NE; ++e) FX[e * NFX + f] = fl[e];
        if constexpr (JUMP) {
          int const gx = x0 + ia;
          if (gx == 0 || gx == kIC) save(gx == kIC ? 0 : 1, y0 + jj, z0 + k, fl);
        }
      }
    } else if (f < PX + PY) {
      int const r = f - PX;
      if (r < NFY) {
        int const k = r % TZe, ja = (r / TZe) % (TY + 1), i = r / (TZe * (TY + 1));
        stencil_flux<DIM, 1, FLUX, R>(bb, x0 + i + kHS, y0 + ja + kHS - 1, z0 + k + zh, gamma, b,
                                      fl);
#pragma unroll
        for (int e = 0; e < NE; ++e) FY[e * NFY + r] = fl[e];
        if constexpr (JUMP) {
          int const gy = y0 + ja;
          if (gy == 0 || gy == kIC) save(gy == kIC ? 2 : 3, x0 + i, z0 + k, fl);
        }
      }
    } else {
      if constexpr (DIM == 3) {
        int const r = f - PX - PY;
        if (r < NFZ) {
          int const ka = r % (TZ + 1), jj = (r / (TZ + 1)) % TY, i = r / ((TZ + 1) * TY);
          stencil_flux<DIM, 2, FLUX, R>(bb, x0 + i + kHS, y0 + jj + kHS, z0 + ka + kHS - 1, gamma,
                                        b, fl);
#pragma unroll
          for (int e = 0; e < NE; ++e) FZ[e * NFZ + r] = fl[e];
          if constexpr (JUMP) {
            int const gz = z0 + ka;
            if (gz == 0 || gz == kIC) save(gz == kIC ? 4 : 5, x0 + i, y0 + jj, fl);
          }
        }
      }
    }
  }
}

// One tile of one block: faces into LDS, barrier, divergence + RK + halo
// push (+ CFL signal).  blk_local is the block's index within its case.
// All threads of the workgroup must call it (it contains a barrier).
template <int DIM, int FLUX, typename R, int TX, int TY, int TZ, int NT>
__device__ __forceinline__ void stage_tile(const TileJob &j, int blk_local, int tile, R *lds) {
  using D = Dims<DIM>;
  using T = TileShape<DIM, TX, TY, TZ>;
  constexpr int NE = D::NE;
  constexpr int TZe = T::TZe, NTY = T::NTY, NTZ = T::NTZ;
  constexpr int NFX = T::NFX, NFY = T::NFY, NFZ = T::NFZ;
  R *FX = lds;
  R *FY = lds + NE * NFX;
  R *FZ = FY + NE * NFY;
  int const blk = j.base + blk_local;
  int const z0 = DIM == 3 ? (tile % NTZ) * TZ : 0;
  int const y0 = ((tile / NTZ) % NTY) * TY;
  int const x0 = (tile / (NTZ * NTY)) * TX;
  const double *bb = j.avg + (size_t)blk * D::BLOCK;

  // 1. Face fluxes of the tile into LDS.
  tile_faces<DIM, FLUX, R, TX, TY, TZ, NT, false>(bb, x0, y0, z0, j.gamma, j.b, lds, nullptr);
  __syncthreads();

  // 2. Divergence + RK combination + halo push (+ CFL signal).
  double const dt = j.state->dt;
  int const stage = j.stage;
  double const m0 = stage == 0 ? 1.0 : (stage == 1 ? 0.25 : 2.0 / 3.0);
  double const m1 = stage == 0 ? 0.0 : (stage == 1 ? 0.75 : 1.0 / 3.0);
  double const cs = stage == 0 ? 1.0 : (stage == 1 ? 0.25 : 2.0 / 3.0);
  double const cdt = cs * (0.0 + dt);
  int const bc3[3] = {blk_local / (j.nb[1] * j.nb[2]), (blk_local / j.nb[2]) % j.nb[1],
                      blk_local % j.nb[2]};
  double sig = 0.0;
  bool bad = false;
  for (int c = threadIdx.x; c < T::CELLS; c += NT) {
    int const kk = c % TZe, jj = (c / TZe) % TY, ii = c / (TZe * TY);
    int const c3[3] = {x0 + ii, y0 + jj, z0 + kk};
    size_t const cell = (size_t)blk * D::BLOCK + (c3[0] + kHS) * D::SX + (c3[1] + kHS) * D::SY +
                        (DIM == 3 ? c3[2] + kHS : 0);
    double v[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      double const dX = double(FX[e * NFX + (ii * TY + jj) * TZe + kk]) -
                        double(FX[e * NFX + ((ii + 1) * TY + jj) * TZe + kk]);
      double const dY = double(FY[e * NFY + (ii * (TY + 1) + jj) * TZe + kk]) -
                        double(FY[e * NFY + (ii * (TY + 1) + jj + 1) * TZe + kk]);
      double dZ = 0.0;
      if constexpr (DIM == 3)
        dZ = double(FZ[e * NFZ + (ii * TY + jj) * (TZ + 1) + kk]) -
             double(FZ[e * NFZ + (ii * TY + jj) * (TZ + 1) + kk + 1]);
      double const rhs = 0.0 + (0.0 + dams<DIM>(dX, dY, dZ) * j.inv_dx);
      size_t const idx = cell + (size_t)e * D::FIELD;
      double const u = j.avg[idx];
      if (stage == 0 && j.init_copy) j.init_copy[idx] = u;
      double const base = stage == 0 ? u : m0 * u + m1 * j.init[idx];
      v[e] = base + cdt * rhs;
      j.out[idx] = v[e];
      bad |= !__builtin_isfinite(v[e]);
      push_halo<DIM>(j, blk, bc3, c3, e, v[e]);
    }
    if (j.signal) {
      // The CFL signal is FP64 without contraction in every mode, so a step's
      // dt does not depend on which kernel reduced the signal.
      exact::Prim<double> p;
      exact::cons_to_prim<DIM>(j.gamma, j.b, v, p);
      double const cso = exact::eos_sos(j.gamma, j.b, p.rho, p.p);
      double sum;
      if constexpr (DIM == 2)
        sum = (__builtin_fabs(p.vel[0]) + cso) + (__builtin_fabs(p.vel[1]) + cso);
      else
        sum = exact::cms3(__builtin_fabs(p.vel[0]) + cso, __builtin_fabs(p.vel[1]) + cso,
                          __builtin_fabs(p.vel[2]) + cso);
      sig = smax(sig, sum);
    }
  }
  if (j.signal) {
    sig = wave_max64(sig);
    if ((threadIdx.x & 63) == 0)
      atomicMax(reinterpret_cast<unsigned long long *>(&j.state->smax),
                (unsigned long long)__double_as_longlong(sig));
  }
  if (bad) atomicOr(&j.state->flags, kFlagNonFinite);
}

// Waves per SIMD the register allocator is held to in the single-case
// kernel.  The exact FP64 2D HLLC kernel needs 178 VGPRs unconstrained (2
// waves); held to 168 it fits 3 waves with 12 spilled registers (2048^2
// cells: 2.16e9 -> 2.36e9 cell-updates/s).  The persistent batch kernel
// spills 34-64 registers under the same hold and runs slower with it, so it
// and every other variant keep the default.
template <int DIM, int FLUX, typename R>
constexpr int kMinWaves = (DIM == 2 && FLUX == 0 && !kFastArith && std::is_same<R, double>::value) ? 3
                          : (FLUX == 1 && !kFastArith && std::is_same<R, double>::value) ? 2 : 1;

// The persistent batch kernel holds only the exact flux-splitting variant (to
// 2 waves per SIMD, which its field-by-field face allows).
template <int FLUX, typename R>
constexpr int kBatchMinWaves = (FLUX == 1 && !kFastArith && std::is_same<R, double>::value) ? 2 : 1;

// Single case: one workgroup per (block, tile).
template <int DIM, int FLUX, typename R, int TX, int TY, int TZ, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(kMinWaves<DIM, FLUX, R>, 8))) void stage_kernel(const double *__restrict__ avg,
                                                   const double *init, double *out,
                                                   alpaca_case_state *state, StageArgs a) {
  using T = TileShape<DIM, TX, TY, TZ>;
  static_assert(NT % 64 == 0, "whole wavefronts");
  __shared__ R lds[T::LDS];
  unsigned const g = xcd_remap_group(blockIdx.x, gridDim.x);
  int const blk = g / T::TILES;
  if (blk >= a.nblocks) return;  // uniform over the workgroup
  if (a.guard && (state->flags & kFlagDone)) return;
  TileJob j;
  j.avg = avg;
  j.init = init;
  j.out = out;
  j.state = state;
  j.init_copy = a.init_copy;
  j.base = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    j.nb[d] = a.nb[d];
    j.bc[d] = a.bc[d];
  }
  j.gamma = a.gamma;
  j.b = a.b;
  j.inv_dx = a.inv_dx;
  j.stage = a.stage;
  j.signal = a.signal;
  stage_tile<DIM, FLUX, R, TX, TY, TZ, NT>(j, blk, g % T::TILES, lds);
}

// Right-hand side of listed blocks of a pool (SpaceSolver::UpdateFluxes in
// full, space_solver.cpp:50-182), one workgroup per (list entry, tile):
// the tile's faces into LDS, those on the block's boundary also added into
// its jump-flux surfaces; after the barrier each internal cell's flux
// divergence (ConsistencyManagedSum, space_solver.cpp:131-153) is written to
// the block's right-hand-side buffer, and the tile's share of the block's
// halo cells is set to zero (the whole output block is written, as by the
// unfused per-block entry).  In 2D the Top and Bottom surfaces receive += 0.0,
// as from the reference's zeroed z face array.  List entries outside
// [0, capacity) are skipped (uniformly over the workgroup).
template <int DIM, int FLUX, typename R, int TX, int TY, int TZ, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(kMinWaves<DIM, FLUX, R>, 8))) void rhs_tile_kernel(
    const double *__restrict__ avg, double *__restrict__ rhs, double *__restrict__ jump,
    const int *__restrict__ list, int n, int capacity, double gamma, double b, double inv_dx) {
  using D = Dims<DIM>;
  using T = TileShape<DIM, TX, TY, TZ>;
  constexpr int NE = D::NE;
  constexpr int TZe = T::TZe, NTY = T::NTY, NTZ = T::NTZ;
  constexpr int NFX = T::NFX, NFY = T::NFY, NFZ = T::NFZ;
  constexpr int ICZ = D::ICZ;
  constexpr int SURF = 6 * NE * kIC * ICZ;  // doubles per SurfaceBuffer
  __shared__ R lds[T::LDS];
  unsigned const g = xcd_remap_group(blockIdx.x, gridDim.x);
  int const item = int(g / T::TILES), tile = int(g % T::TILES);
  if (item >= n) return;
  int const blk = list[item];
  if (blk < 0 || blk >= capacity) return;  // uniform over the workgroup
  int const z0 = DIM == 3 ? (tile % NTZ) * TZ : 0;
  int const y0 = ((tile / NTZ) % NTY) * TY;
  int const x0 = (tile / (NTZ * NTY)) * TX;
  double *jb = jump + (size_t)blk * SURF;
  tile_faces<DIM, FLUX, R, TX, TY, TZ, NT, true>(avg + (size_t)blk * D::BLOCK, x0, y0, z0, gamma,
                                                 b, lds, jb);
  if constexpr (DIM == 2) {
    for (int t = threadIdx.x; t < 2 * NE * kIC; t += NT) jb[4 * NE * kIC + t] += 0.0;
  }
  __syncthreads();
  const R *FX = lds;
  const R *FY = lds + NE * NFX;
  const R *FZ = FY + NE * NFY;
  double *out = rhs + (size_t)blk * D::BLOCK;
  for (int c = threadIdx.x; c < T::CELLS; c += NT) {
    int const kk = c % TZe, jj = (c / TZe) % TY, ii = c / (TZe * TY);
    int const cell = (x0 + ii + kHS) * D::SX + (y0 + jj + kHS) * D::SY +
                     (DIM == 3 ? z0 + kk + kHS : 0);
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      double const dX = double(FX[e * NFX + (ii * TY + jj) * TZe + kk]) -
                        double(FX[e * NFX + ((ii + 1) * TY + jj) * TZe + kk]);
      double const dY = double(FY[e * NFY + (ii * (TY + 1) + jj) * TZe + kk]) -
                        double(FY[e * NFY + (ii * (TY + 1) + jj + 1) * TZe + kk]);
      double dZ = 0.0;
      if constexpr (DIM == 3)
        dZ = double(FZ[e * NFZ + (ii * TY + jj) * (TZ + 1) + kk]) -
             double(FZ[e * NFZ + (ii * TY + jj) * (TZ + 1) + kk + 1]);
      out[e * D::FIELD + cell] = 0.0 + (0.0 + dams<DIM>(dX, dY, dZ) * inv_dx);
    }
  }
  // The tile's share of the halo cells: positions [tile * FIELD / TILES, ...).
  constexpr int SHARE = D::FIELD / T::TILES;
  static_assert(SHARE * T::TILES == D::FIELD, "halo shares cover the block");
  for (int q = threadIdx.x; q < SHARE; q += NT) {
    int const r = tile * SHARE + q;
    int const i = r / D::SX, j = (r / D::SY) % kTC, k = DIM == 3 ? r % kTC : kHS;
    bool const inside = i >= kHS && i < kHS + kIC && j >= kHS && j < kHS + kIC &&
                        k >= kHS && k < kHS + kIC;
    if (!inside) {
#pragma unroll
      for (int e = 0; e < NE; ++e) out[e * D::FIELD + r] = 0.0;
    }
  }
}

// Every tile of every listed block of every case of an ensemble, one
// persistent kernel per RK stage: the grid is sized to the device and each
// workgroup walks the work list (read from device memory, so a captured
// graph of the step stays valid when the list changes).  Work units are
// split into 8 contiguous ranges, one per XCD (group g runs on XCD g % 8),
// so a case's blocks share an L2.  Units of stopped cases are skipped.
template <int DIM, int FLUX, typename R, int TX, int TY, int TZ, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(kBatchMinWaves<FLUX, R>, 8))) void batch_stage_kernel(
    const double *__restrict__ avg, const double *init, double *out, const CaseSlot *slots,
    alpaca_case_state *states, const WorkItem *items, const int *n_items, int stage) {
  using T = TileShape<DIM, TX, TY, TZ>;
  __shared__ R lds[T::LDS];
  __shared__ int skip;
  constexpr unsigned kXcd = 8;
  unsigned const n = unsigned(*n_items);
  unsigned const groups = gridDim.x;
  // Work items are dealt to the 8 XCDs round-robin (item k to XCD k % 8; the
  // hardware deals workgroups round-robin too, so blockIdx % 8 is the XCD),
  // all tiles of an item staying on one XCD's L2.  Round-robin rather than
  // contiguous ranges keeps the XCDs balanced when the active cases cluster
  // in the list (finished cases are skipped); within an XCD the workgroups
  // stride over its items' tiles.
  unsigned const xcd = blockIdx.x % kXcd;
  unsigned const per_xcd = groups / kXcd;  // grid is a multiple of 8
  unsigned const m = blockIdx.x / kXcd;
  unsigned const items_here = n > xcd ? (n - xcd + kXcd - 1) / kXcd : 0u;
  unsigned const units_here = items_here * T::TILES;
  for (unsigned w = m; w < units_here; w += per_xcd) {
    unsigned const u = (xcd + kXcd * (w / T::TILES)) * T::TILES + w % T::TILES;
    WorkItem const it = items[u / T::TILES];
    alpaca_case_state *st = states + it.case_id;
    // Another workgroup may flag the case during this stage: decide once per
    // workgroup so every thread takes the same branch around the barriers.
    if (threadIdx.x == 0) skip = (__atomic_load_n(&st->flags, __ATOMIC_RELAXED) & kFlagsStop()) != 0;
    __syncthreads();
    bool const stop = skip;
    __syncthreads();
    if (stop) continue;
    CaseSlot const &cs = slots[it.case_id];
    TileJob j;
    j.avg = avg;
    j.init = init;
    j.out = out;
    j.state = st;
    j.init_copy = nullptr;
    j.base = cs.base;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      j.nb[d] = cs.nb[d];
      j.bc[d] = cs.bc[d];
    }
    j.gamma = cs.gamma;
    j.b = cs.b;
    j.inv_dx = cs.inv_dx;
    j.stage = stage;
    j.signal = stage == 2;
    stage_tile<DIM, FLUX, R, TX, TY, TZ, NT>(j, it.block, u % T::TILES, lds);
    __syncthreads();  // LDS is reused by the next unit
  }
}
