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
//      zero-gradient side, into its own halo (boundary_constants.h:212-253);
//   3. on the last stage it also reduces the CFL signal of the new state
//      (modular_algorithm_assembler.cpp:1840-1869) for the next time step.

__device__ __forceinline__ double wave_max64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = smax(v, __shfl_xor(v, o, 64));
  return v;
}

// Writes value v of field e at internal cell c3 of block blk into the halo
// slabs that read it: the neighbours' slabs across each near face, or the
// block's own slab on a zero-gradient domain side.
template <int DIM>
__device__ __forceinline__ void push_halo(double *__restrict__ out, int blk, const int *bc3,
                                          const int *c3, const StageArgs &a, int e, double v) {
  using D = Dims<DIM>;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    int const c = c3[d];
    int pos[3] = {c3[0] + kHS, c3[1] + kHS, DIM == 3 ? c3[2] + kHS : 0};
    if (c < kHS) {
      if (bc3[d] > 0 || a.bc[d] == ALPACA_BC_PERIODIC) {
        int nb3[3] = {bc3[0], bc3[1], bc3[2]};
        nb3[d] = bc3[d] > 0 ? bc3[d] - 1 : a.nb[d] - 1;
        int const nbk = (nb3[0] * a.nb[1] + nb3[1]) * a.nb[2] + nb3[2];
        pos[d] = c + kHS + kIC;
        out[(size_t)nbk * D::BLOCK + e * D::FIELD + pos[0] * D::SX + pos[1] * D::SY + pos[2]] = v;
      } else if (c == 0) {
#pragma unroll
        for (int h = 0; h < kHS; ++h) {
          pos[d] = h;
          out[(size_t)blk * D::BLOCK + e * D::FIELD + pos[0] * D::SX + pos[1] * D::SY + pos[2]] = v;
        }
      }
    }
    if (c >= kIC - kHS) {
      if (bc3[d] < a.nb[d] - 1 || a.bc[d] == ALPACA_BC_PERIODIC) {
        int nb3[3] = {bc3[0], bc3[1], bc3[2]};
        nb3[d] = bc3[d] < a.nb[d] - 1 ? bc3[d] + 1 : 0;
        int const nbk = (nb3[0] * a.nb[1] + nb3[1]) * a.nb[2] + nb3[2];
        pos[d] = c - kIC + kHS;
        out[(size_t)nbk * D::BLOCK + e * D::FIELD + pos[0] * D::SX + pos[1] * D::SY + pos[2]] = v;
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

template <int DIM, int FLUX, typename R, int TX, int TY, int TZ, int NT>
__global__ __launch_bounds__(NT) void stage_kernel(const double *__restrict__ avg,
                                                   const double *init, double *out,
                                                   alpaca_case_state *state, StageArgs a) {
  using D = Dims<DIM>;
  constexpr int NE = D::NE;
  constexpr int TZe = DIM == 3 ? TZ : 1;
  constexpr int NTX = kIC / TX, NTY = kIC / TY, NTZ = DIM == 3 ? kIC / TZ : 1;
  constexpr int TILES = NTX * NTY * NTZ;
  constexpr int NFX = (TX + 1) * TY * TZe;
  constexpr int NFY = TX * (TY + 1) * TZe;
  constexpr int NFZ = DIM == 3 ? TX * TY * (TZ + 1) : 0;
  constexpr int PX = round_up64(NFX), PY = round_up64(NFY), PZ = round_up64(NFZ);
  constexpr int CELLS = TX * TY * TZe;
  static_assert(NT % 64 == 0, "whole wavefronts");
  __shared__ R lds[NE * (NFX + NFY + NFZ)];
  R *FX = lds;
  R *FY = lds + NE * NFX;
  R *FZ = FY + NE * NFY;

  unsigned const g = xcd_remap_group(blockIdx.x, gridDim.x);
  int const blk = g / TILES;
  if (blk >= a.nblocks) return;  // uniform over the workgroup
  int const tile = g % TILES;
  int const z0 = DIM == 3 ? (tile % NTZ) * TZ : 0;
  int const y0 = ((tile / NTZ) % NTY) * TY;
  int const x0 = (tile / (NTZ * NTY)) * TX;
  const double *bb = avg + (size_t)blk * D::BLOCK;
  int const zh = DIM == 3 ? kHS : 0;  // halo offset along z

  // 1. Face fluxes of the tile into LDS.
  for (int f = threadIdx.x; f < PX + PY + PZ; f += NT) {
    R fl[NE];
    if (f < PX) {
      if (f < NFX) {
        int const k = f % TZe, j = (f / TZe) % TY, ia = f / (TZe * TY);
        stencil_flux<DIM, 0, FLUX, R>(bb, x0 + ia + kHS - 1, y0 + j + kHS, z0 + k + zh, a.gamma, a.b, fl);
#pragma unroll
        for (int e = 0; e < NE; ++e) FX[e * NFX + f] = fl[e];
      }
    } else if (f < PX + PY) {
      int const r = f - PX;
      if (r < NFY) {
        int const k = r % TZe, ja = (r / TZe) % (TY + 1), i = r / (TZe * (TY + 1));
        stencil_flux<DIM, 1, FLUX, R>(bb, x0 + i + kHS, y0 + ja + kHS - 1, z0 + k + zh, a.gamma, a.b, fl);
#pragma unroll
        for (int e = 0; e < NE; ++e) FY[e * NFY + r] = fl[e];
      }
    } else {
      if constexpr (DIM == 3) {
        int const r = f - PX - PY;
        if (r < NFZ) {
          int const ka = r % (TZ + 1), j = (r / (TZ + 1)) % TY, i = r / ((TZ + 1) * TY);
          stencil_flux<DIM, 2, FLUX, R>(bb, x0 + i + kHS, y0 + j + kHS, z0 + ka + kHS - 1, a.gamma, a.b, fl);
#pragma unroll
          for (int e = 0; e < NE; ++e) FZ[e * NFZ + r] = fl[e];
        }
      }
    }
  }
  __syncthreads();

  // 2. Divergence + RK combination + halo push (+ CFL signal).
  double const dt = state->dt;
  int const stage = a.stage;
  double const m0 = stage == 0 ? 1.0 : (stage == 1 ? 0.25 : 2.0 / 3.0);
  double const m1 = stage == 0 ? 0.0 : (stage == 1 ? 0.75 : 1.0 / 3.0);
  double const cs = stage == 0 ? 1.0 : (stage == 1 ? 0.25 : 2.0 / 3.0);
  double const cdt = cs * (0.0 + dt);
  int const bc3[3] = {blk / (a.nb[1] * a.nb[2]), (blk / a.nb[2]) % a.nb[1], blk % a.nb[2]};
  double sig = 0.0;
  bool bad = false;
  for (int c = threadIdx.x; c < CELLS; c += NT) {
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
      double const rhs = 0.0 + (0.0 + dams<DIM>(dX, dY, dZ) * a.inv_dx);
      size_t const idx = cell + (size_t)e * D::FIELD;
      double const u = avg[idx];
      if (stage == 0 && a.init_copy) a.init_copy[idx] = u;
      double const base = stage == 0 ? u : m0 * u + m1 * init[idx];
      v[e] = base + cdt * rhs;
      out[idx] = v[e];
      bad |= !__builtin_isfinite(v[e]);
      push_halo<DIM>(out, blk, bc3, c3, a, e, v[e]);
    }
    if (a.signal) {
      Prim<double> p;
      cons_to_prim<DIM>(a.gamma, a.b, v, p);
      double const cso = eos_sos(a.gamma, a.b, p.rho, p.p);
      double sum;
      if constexpr (DIM == 2)
        sum = (__builtin_fabs(p.vel[0]) + cso) + (__builtin_fabs(p.vel[1]) + cso);
      else
        sum = cms3(__builtin_fabs(p.vel[0]) + cso, __builtin_fabs(p.vel[1]) + cso,
                   __builtin_fabs(p.vel[2]) + cso);
      sig = smax(sig, sum);
    }
  }
  if (a.signal) {
    sig = wave_max64(sig);
    if ((threadIdx.x & 63) == 0)
      atomicMax(reinterpret_cast<unsigned long long *>(&state->smax),
                (unsigned long long)__double_as_longlong(sig));
  }
  if (bad) atomicOr(&state->flags, 1);
}

