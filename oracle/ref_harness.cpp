// Test infrastructure only: never part of the product path.
//
// Driver that links ALPACA's own hot-path translation units (compiled in place
// from the reference source tree by oracle/Makefile into oracle/_ref/) and
// exposes them through a small C ABI.  It is used for two things:
//   * writing golden vectors (tests/golden/make_golden.py), which pin the CPU
//     restatement in oracle/alpaca_oracle.c;
//   * the "reference" CPU baseline leg of bench.py (ref_run with threads).
//
// What is the reference's own code here: the convective solvers
// (FiniteVolumeScheme + HllcRiemannSolver/Einfeldt, FluxSplittingScheme/Roe,
// EigenDecomposition, CharacteristicStateReconstruction, WENO5), the
// stiffened-gas EOS, ConservativesToPrimeStates, the RungeKutta3TVD time
// integrator on reference Nodes/Blocks, Zero-gradient/Symmetry boundary conditions,
// ConsistencyManagedSum, and the multiresolution operators (Multiresolution::
// Average/Prediction/ChildNeedsRemeshing, ThresholdComputer).  What this file restates around them (all cited):
// the SpaceSolver divergence loop (space_solver.cpp:50-153, source terms off),
// the uniform-level Advance choreography (modular_algorithm_assembler.cpp:
// 543-910), ComputeTimestepSize (:1777-1997, inviscid terms only) and the
// same-level halo copy (internal_halo_manager.cpp:439-466) because those live
// in TUs that need MPI and the topology manager.
#include "block_definitions/block.h"
#include "boundary_condition/symmetry_boundary_condition.h"
#include "boundary_condition/zero_gradient_boundary_condition.h"
#include "integrator/runge_kutta_3_TVD.h"
#include "materials/equations_of_state/stiffened_gas.h"
#include "materials/material_manager.h"
#include "multiresolution/multiresolution.h"
#include "prime_states/prime_state_handler.h"
#include "solvers/convective_term_contributions/finite_volume_scheme.h"
#include "solvers/convective_term_contributions/flux_splitting_scheme.h"
#include "solvers/convective_term_contributions/riemann_solvers/hllc_riemann_solver.h"
#include "solvers/eigendecomposition.h"
#include "stencils/spatial_reconstruction_stencils/weno5.h"
#include "stencils/stencil_utilities.h"
#include "topology/id_information.h"
#include "topology/node.h"
#include "unit_handler.h"
#include "utilities/mathematical_functions.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cstring>
#include <memory>
#include <pthread.h>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

constexpr unsigned int NE = MF::ANOE();
constexpr unsigned int NP = MF::ANOP();
constexpr unsigned int TX = CC::TCX(), TY = CC::TCY(), TZ = CC::TCZ();
constexpr unsigned int IX = CC::ICX(), IY = CC::ICY(), IZ = CC::ICZ();
constexpr unsigned int HX = CC::FICX();
constexpr unsigned int HY = CC::FICY();
constexpr unsigned int HZ = CC::FICZ();
constexpr int DIMS = DTI(CC::DIM());

struct Setup {
  std::unique_ptr<UnitHandler> units;
  std::unique_ptr<MaterialManager> materials;
  std::unique_ptr<EigenDecomposition> eigen;
  std::unique_ptr<FiniteVolumeScheme> fv;
  std::unique_ptr<FluxSplittingScheme> fs;
  std::unique_ptr<HllcRiemannSolver> hllc;
  double gamma = 0.0, b = 0.0;
};

Setup &S() {
  static Setup s;
  return s;
}

void EnsureSetup(double gamma, double b) {
  Setup &s = S();
  if (s.materials && s.gamma == gamma && s.b == b)
    return;
  s.hllc.reset();
  s.fs.reset();
  s.fv.reset();
  s.eigen.reset();
  s.materials.reset();
  s.units = std::make_unique<UnitHandler>(1.0, 1.0, 1.0, 1.0);
  std::unordered_map<std::string, double> eos_data{
      {"gamma", gamma}, {"backgroundPressure", b}};
  std::vector<std::tuple<MaterialType, Material>> mats;
  mats.emplace_back(
      MaterialType::Fluid,
      Material(std::make_unique<StiffenedGas const>(eos_data, *s.units), 0.0,
               0.0, 0.0, 0.0, nullptr, nullptr, *s.units));
  s.materials = std::make_unique<MaterialManager>(
      std::move(mats), std::vector<MaterialPairing>{});
  s.eigen = std::make_unique<EigenDecomposition>(*s.materials);
  s.fv = std::make_unique<FiniteVolumeScheme>(*s.materials, *s.eigen);
  s.fs = std::make_unique<FluxSplittingScheme>(*s.materials, *s.eigen);
  s.hllc = std::make_unique<HllcRiemannSolver>(*s.materials, *s.eigen);
  s.gamma = gamma;
  s.b = b;
}

EquationOfState const &Eos() {
  return S().materials->GetMaterial(MaterialName::MaterialOne)
      .GetEquationOfState();
}

// Prime states on all TC cells from the Average buffer, as
// ObtainPrimeStatesFromConservatives does (prime_state_handler.h:141-152).
void PrimesFromAverage(Block &block) {
  Conservatives const &avg = block.GetAverageBuffer();
  PrimeStates &prm = block.GetPrimeStateBuffer();
  EquationOfState const &eos = Eos();
  for (unsigned int i = 0; i < TX; ++i)
    for (unsigned int j = 0; j < TY; ++j)
      for (unsigned int k = 0; k < TZ; ++k) {
        std::array<double, NE> c;
        std::array<double, NP> p;
        for (unsigned int e = 0; e < NE; ++e)
          c[e] = avg[e][i][j][k];
        ConservativesToPrimeStates(eos, c, p);
        for (unsigned int q = 0; q < NP; ++q)
          prm[q][i][j][k] = p[q];
      }
}

using FaceFluxes = double[NE][IX + 1][IY + 1][IZ + 1];
using VolumeForces = double[NE][IX][IY][IZ];

struct FluxScratch {
  FaceFluxes fx, fy, fz;
  VolumeForces vf;
};

void ConvectiveFluxes(int family, std::pair<MaterialName const, Block> &mb,
                      double cell_size, FluxScratch &w) {
  std::memset(&w, 0, sizeof(FluxScratch));
  if (family == 0)
    S().fv->UpdateConvectiveFluxes(mb, cell_size, w.fx, w.fy, w.fz, w.vf);
  else
    S().fs->UpdateConvectiveFluxes(mb, cell_size, w.fx, w.fy, w.fz, w.vf);
}

// SpaceSolver::UpdateFluxes with all source terms off (space_solver.cpp:50-153).
void UpdateRhs(int family, std::pair<MaterialName const, Block> &mb,
               double cell_size, FluxScratch &w) {
  Block &block = mb.second;
  for (unsigned int e = 0; e < NE; ++e) {
    double(&rhs)[TX][TY][TZ] = block.GetRightHandSideBuffer(MF::ASOE()[e]);
    for (unsigned int i = 0; i < TX; ++i)
      for (unsigned int j = 0; j < TY; ++j)
        for (unsigned int k = 0; k < TZ; ++k)
          rhs[i][j][k] = 0.0;
  }
  double const one_cell_size = 1.0 / cell_size;
  ConvectiveFluxes(family, mb, cell_size, w);
  for (unsigned int e = 0; e < NE; ++e) {
    double(&cells)[TX][TY][TZ] = block.GetRightHandSideBuffer(MF::ASOE()[e]);
    for (unsigned int i = 0; i < IX; ++i)
      for (unsigned int j = 0; j < IY; ++j)
        for (unsigned int k = 0; k < IZ; ++k)
          cells[i + HX][j + HY][k + HZ] +=
              w.vf[e][i][j][k] +
              DimensionAwareConsistencyManagedSum(
                  w.fx[e][i][j + 1][k + 1] - w.fx[e][i + 1][j + 1][k + 1],
                  w.fy[e][i + 1][j][k + 1] - w.fy[e][i + 1][j + 1][k + 1],
                  w.fz[e][i + 1][j + 1][k] - w.fz[e][i + 1][j + 1][k + 1]) *
                  one_cell_size;
  }
}

void LoadAverage(Block &block, double const *avg) {
  Conservatives &a = block.GetAverageBuffer();
  for (unsigned int e = 0; e < NE; ++e)
    std::memcpy(&a[e][0][0][0], avg + std::size_t(e) * TX * TY * TZ,
                sizeof(double) * TX * TY * TZ);
}

// Runs fn(lo, hi) over [0, n) on `threads` threads with large stacks: the
// reference keeps ~3 MB of face arrays on the stack per UpdateFluxes call.
template <typename F> void ParallelFor(int n, int threads, F const &fn) {
  if (threads <= 1 || n <= 1) {
    fn(0, n);
    return;
  }
  threads = std::min(threads, n);
  struct Arg {
    F const *fn;
    int lo, hi;
  };
  std::vector<pthread_t> tids(threads);
  std::vector<Arg> args(threads);
  pthread_attr_t attr;
  pthread_attr_init(&attr);
  pthread_attr_setstacksize(&attr, std::size_t(64) << 20);
  for (int t = 0; t < threads; ++t) {
    args[t] = Arg{&fn, int(std::int64_t(n) * t / threads),
                  int(std::int64_t(n) * (t + 1) / threads)};
    pthread_create(
        &tids[t], &attr,
        [](void *p) -> void * {
          Arg *a = static_cast<Arg *>(p);
          (*a->fn)(a->lo, a->hi);
          return nullptr;
        },
        &args[t]);
  }
  for (int t = 0; t < threads; ++t)
    pthread_join(tids[t], nullptr);
  pthread_attr_destroy(&attr);
}

struct Grid {
  int nb[3];
  int bc[3];  // 0 zero-gradient, 1 periodic, 2 symmetry (both sides)
  std::vector<std::unique_ptr<Node>> nodes;
  int NumBlocks() const { return nb[0] * nb[1] * nb[2]; }
  int Index(int bx, int by, int bz) const {
    return (bx * nb[1] + by) * nb[2] + bz;
  }
  Block &Blk(int b) { return nodes[b]->GetSinglePhase(); }
};

// Same-level halo copy of the RightHandSide buffer from the owner block of
// each halo cell (internal_halo_manager.cpp:439-466; periodic partners via the
// wrapped neighbour, id_periodic_information.cpp:325).  Cells whose owner lies
// outside a non-periodic domain side are left for the external BC.
void HaloCopy(Grid &g, int b_lo, int b_hi) {
  for (int b = b_lo; b < b_hi; ++b) {
    int bx = b / (g.nb[1] * g.nb[2]);
    int by = (b / g.nb[2]) % g.nb[1];
    int bz = b % g.nb[2];
    Conservatives &dst = g.Blk(b).GetRightHandSideBuffer();
    for (unsigned int i = 0; i < TX; ++i)
      for (unsigned int j = 0; j < TY; ++j)
        for (unsigned int k = 0; k < TZ; ++k) {
          bool const halo_x = i < HX || i >= HX + IX;
          bool const halo_y = DIMS >= 2 && (j < HY || j >= HY + IY);
          bool const halo_z = DIMS == 3 && (k < HZ || k >= HZ + IZ);
          if (!(halo_x || halo_y || halo_z))
            continue;
          int idx[3] = {int(i) - int(HX), int(j) - int(HY), int(k) - int(HZ)};
          int blk[3] = {bx, by, bz};
          int ic[3] = {int(IX), int(IY), int(IZ)};
          bool outside = false;
          for (int d = 0; d < DIMS; ++d) {
            int gcell = blk[d] * ic[d] + idx[d];
            int n = g.nb[d] * ic[d];
            if (gcell < 0 || gcell >= n) {
              if (g.bc[d] != 1) {
                outside = true;
                break;
              }
              gcell = (gcell % n + n) % n;
            }
            blk[d] = gcell / ic[d];
            idx[d] = gcell % ic[d];
          }
          if (outside)
            continue;
          Conservatives const &src =
              g.Blk(g.Index(blk[0], blk[1], blk[2])).GetRightHandSideBuffer();
          for (unsigned int e = 0; e < NE; ++e)
            dst[e][i][j][k] =
                src[e][idx[0] + HX][DIMS >= 2 ? idx[1] + HY : 0]
                   [DIMS == 3 ? idx[2] + HZ : 0];
        }
  }
}

template <BoundaryLocation LOC> void External(Node &node, int kind) {
  if (kind == 2) {
    SymmetryBoundaryCondition<LOC> bc;
    bc.UpdateMaterialExternal(node, MaterialFieldType::Conservatives);
  } else {
    ZeroGradientBoundaryCondition<LOC> bc;
    bc.UpdateMaterialExternal(node, MaterialFieldType::Conservatives);
  }
}

// External boundaries in the order East, West, North, South, Top, Bottom.
void ExternalHalo(Grid &g, int b_lo, int b_hi) {
  for (int b = b_lo; b < b_hi; ++b) {
    int bx = b / (g.nb[1] * g.nb[2]);
    int by = (b / g.nb[2]) % g.nb[1];
    int bz = b % g.nb[2];
    Node &node = *g.nodes[b];
    if (g.bc[0] != 1 && bx == g.nb[0] - 1)
      External<BoundaryLocation::East>(node, g.bc[0]);
    if (g.bc[0] != 1 && bx == 0)
      External<BoundaryLocation::West>(node, g.bc[0]);
    if constexpr (DIMS >= 2) {
      if (g.bc[1] != 1 && by == g.nb[1] - 1)
        External<BoundaryLocation::North>(node, g.bc[1]);
      if (g.bc[1] != 1 && by == 0)
        External<BoundaryLocation::South>(node, g.bc[1]);
    }
    if constexpr (DIMS == 3) {
      if (g.bc[2] != 1 && bz == g.nb[2] - 1)
        External<BoundaryLocation::Top>(node, g.bc[2]);
      if (g.bc[2] != 1 && bz == 0)
        External<BoundaryLocation::Bottom>(node, g.bc[2]);
    }
  }
}

} // namespace

extern "C" {

int ref_dim() { return DIMS; }

// Face fluxes of one haloed block.  avg: [NE][TX][TY][TZ].  out: three arrays
// [NE][IX+1][IY+1][IZ+1] (x, y, z) as UpdateConvectiveFluxes fills them.
void ref_block_fluxes(int family, double gamma, double b, double const *avg,
                      double cell_size, double *out) {
  EnsureSetup(gamma, b);
  auto mb = std::make_unique<std::pair<MaterialName const, Block>>(
      std::piecewise_construct, std::forward_as_tuple(MaterialName::MaterialOne),
      std::forward_as_tuple());
  LoadAverage(mb->second, avg);
  PrimesFromAverage(mb->second);
  auto w = std::make_unique<FluxScratch>();
  ConvectiveFluxes(family, *mb, cell_size, *w);
  std::size_t const n = sizeof(FaceFluxes) / sizeof(double);
  std::memcpy(out, w->fx, sizeof(FaceFluxes));
  std::memcpy(out + n, w->fy, sizeof(FaceFluxes));
  std::memcpy(out + 2 * n, w->fz, sizeof(FaceFluxes));
}

// Right-hand side of one haloed block on its internal cells: out [NE][IX][IY][IZ].
void ref_block_rhs(int family, double gamma, double b, double const *avg,
                   double cell_size, double *out) {
  EnsureSetup(gamma, b);
  auto mb = std::make_unique<std::pair<MaterialName const, Block>>(
      std::piecewise_construct, std::forward_as_tuple(MaterialName::MaterialOne),
      std::forward_as_tuple());
  LoadAverage(mb->second, avg);
  PrimesFromAverage(mb->second);
  auto w = std::make_unique<FluxScratch>();
  UpdateRhs(family, *mb, cell_size, *w);
  Conservatives const &rhs = mb->second.GetRightHandSideBuffer();
  for (unsigned int e = 0; e < NE; ++e)
    for (unsigned int i = 0; i < IX; ++i)
      for (unsigned int j = 0; j < IY; ++j)
        for (unsigned int k = 0; k < IZ; ++k)
          out[((std::size_t(e) * IX + i) * IY + j) * IZ + k] =
              rhs[e][i + HX][j + HY][k + HZ];
}

// WENO5 on a 6-value stencil: out[0] = UpwindLeft, out[1] = UpwindRight.
void ref_weno5(double const *v, double *out) {
  std::array<double, 6> a{v[0], v[1], v[2], v[3], v[4], v[5]};
  out[0] = SU::Reconstruction<WENO5, SP::UpwindLeft>(a, 1.0);
  out[1] = SU::Reconstruction<WENO5, SP::UpwindRight>(a, 1.0);
}

// HLLC(Einfeldt) flux along direction dir for conservative states L and R.
void ref_hllc(int dir, double gamma, double b, double const *cl,
              double const *cr, double *out) {
  EnsureSetup(gamma, b);
  std::array<double, NE> l, r;
  std::array<double, NP> pl, pr;
  for (unsigned int e = 0; e < NE; ++e) {
    l[e] = cl[e];
    r[e] = cr[e];
  }
  ConservativesToPrimeStates(Eos(), l, pl);
  ConservativesToPrimeStates(Eos(), r, pr);
  std::array<double, NE> f;
  if (dir == 0)
    f = S().hllc->SolveRiemannProblem<Direction::X>(MaterialName::MaterialOne,
                                                    l, r, pl, pr);
  else if (dir == 1)
    f = S().hllc->SolveRiemannProblem<Direction::Y>(MaterialName::MaterialOne,
                                                    l, r, pl, pr);
#if DIMENSION == 3
  else
    f = S().hllc->SolveRiemannProblem<Direction::Z>(MaterialName::MaterialOne,
                                                    l, r, pl, pr);
#endif
  for (unsigned int e = 0; e < NE; ++e)
    out[e] = f[e];
}

// Stiffened-gas EOS: in (rho, mx, my, mz, E) -> out (p, H, c); in prime form
// (rho, u, v, w, p) -> out[3] = E, out[4] = psi(p, 1/rho).
void ref_eos(double gamma, double b, double const *cons, double const *prim,
             double *out) {
  EnsureSetup(gamma, b);
  EquationOfState const &eos = Eos();
  out[0] = eos.Pressure(cons[0], cons[1], cons[2], cons[3], cons[4]);
  out[1] = eos.Enthalpy(cons[0], cons[1], cons[2], cons[3], cons[4]);
  out[2] = eos.SpeedOfSound(prim[0], prim[4]);
  out[3] = eos.Energy(prim[0], prim[1], prim[2], prim[3], prim[4]);
  out[4] = eos.Psi(prim[4], 1.0 / prim[0]);
}

// Full uniform level-0 run.  state: internal conservatives of the whole domain
// [NE][NX][NY][NZ] (k fastest, NZ = 1 in 2D), overwritten with the final state.
// bc[d] per direction (both sides): 0 zero-gradient, 1 periodic, 2 symmetry.
// Stops after max_steps macro steps or at t_end (whichever first); the dt of
// each step is written to dts (capacity max_steps).  Returns the step count.
int ref_run(int family, double gamma, double b, int const *nb,
            int const *bc, double cell_size, double cfl, double t_end,
            int max_steps, int threads, double *state, double *dts,
            double *t_final) {
  EnsureSetup(gamma, b);
  Grid g;
  for (int d = 0; d < 3; ++d) {
    g.nb[d] = d < DIMS ? nb[d] : 1;
    g.bc[d] = d < DIMS ? bc[d] : 0;
  }
  int const nblk = g.NumBlocks();
  for (int i = 0; i < nblk; ++i)
    g.nodes.push_back(std::make_unique<Node>(
        IdSeed(), 1.0, std::vector<MaterialName>{MaterialName::MaterialOne}));
  std::size_t const NX = std::size_t(g.nb[0]) * IX;
  std::size_t const NY = std::size_t(g.nb[1]) * IY;
  std::size_t const NZ = std::size_t(g.nb[2]) * IZ;
  auto gidx = [&](unsigned e, std::size_t x, std::size_t y, std::size_t z) {
    return ((e * NX + x) * NY + y) * NZ + z;
  };
  // ImposeInitialCondition writes the internal cells of the RHS buffer, then
  // halo update, swap and primes (modular_algorithm_assembler.cpp:397-517).
  for (int b = 0; b < nblk; ++b) {
    int bx = b / (g.nb[1] * g.nb[2]);
    int by = (b / g.nb[2]) % g.nb[1];
    int bz = b % g.nb[2];
    Conservatives &r = g.Blk(b).GetRightHandSideBuffer();
    for (unsigned int e = 0; e < NE; ++e)
      for (unsigned int i = 0; i < IX; ++i)
        for (unsigned int j = 0; j < IY; ++j)
          for (unsigned int k = 0; k < IZ; ++k)
            r[e][i + HX][j + HY][k + HZ] =
                state[gidx(e, bx * IX + i, by * IY + j, bz * IZ + k)];
  }
  RungeKutta3TVD ti(0.0);
  auto halo_swap_primes = [&]() {
    ParallelFor(nblk, threads, [&](int lo, int hi) { HaloCopy(g, lo, hi); });
    ParallelFor(nblk, threads, [&](int lo, int hi) {
      ExternalHalo(g, lo, hi);
      for (int b = lo; b < hi; ++b) {
        ti.SwapBuffersForNextStage(*g.nodes[b]);
        PrimesFromAverage(g.Blk(b));
      }
    });
  };
  halo_swap_primes();

  EquationOfState const &eos = Eos();
  int steps = 0;
  double current = ti.CurrentRunTime();
  while (current < t_end && steps < max_steps) {
    // ComputeTimestepSize (modular_algorithm_assembler.cpp:1777-1997).
    std::vector<double> block_max(nblk, 0.0);
    ParallelFor(nblk, threads, [&](int lo, int hi) {
      for (int b = lo; b < hi; ++b) {
        PrimeStates const &p = g.Blk(b).GetPrimeStateBuffer();
        double s = 0.0;
        for (unsigned int i = HX; i < HX + IX; ++i)
          for (unsigned int j = HY; j < HY + IY; ++j)
            for (unsigned int k = HZ; k < HZ + IZ; ++k) {
              double const c =
                  eos.SpeedOfSound(p[PrimeState::Density][i][j][k],
                                   p[PrimeState::Pressure][i][j][k]);
              std::array<double, DTI(CC::DIM())> vps;
              for (int d = 0; d < DIMS; ++d)
                vps[d] = std::abs(p[MF::AV()[d]][i][j][k]) + c;
              s = std::max(s, ConsistencyManagedSum(vps));
            }
        block_max[b] = s;
      }
    });
    double sum_of_signalspeeds = 0.0;
    for (double s : block_max)
      sum_of_signalspeeds = std::max(sum_of_signalspeeds, s);
    double dt = sum_of_signalspeeds / cell_size;
    double local = dt == 0.0 ? std::numeric_limits<double>::max() : cfl / dt;
    double const run_time =
        std::accumulate(ti.MicroTimestepSizes().cbegin(),
                        ti.MicroTimestepSizes().cend(), ti.CurrentRunTime());
    if (run_time + local > t_end)
      local = std::max(t_end - run_time, 0.0);
    ti.AppendMicroTimestep(local);
    if (dts)
      dts[steps] = local;

    for (unsigned int stage = 0; stage < ti.NumberOfStages(); ++stage) {
      ParallelFor(nblk, threads, [&](int lo, int hi) {
        auto w = std::make_unique<FluxScratch>();
        for (int b = lo; b < hi; ++b) {
          Node &node = *g.nodes[b];
          ti.FillInitialBuffer(node, stage);
          UpdateRhs(family, *node.GetPhases().begin(), cell_size, *w);
          ti.PrepareBufferForIntegration(node, stage);
          ti.IntegrateNode(node, stage, 1);
        }
      });
      halo_swap_primes();
    }
    double const last = ti.MicroTimestepSizes().back();
    ti.FinishMacroTimestep();
    current = ti.CurrentRunTime();
    ++steps;
    if (last < CC::MTS() && last > 0.0)
      break;
  }
  for (int b = 0; b < nblk; ++b) {
    int bx = b / (g.nb[1] * g.nb[2]);
    int by = (b / g.nb[2]) % g.nb[1];
    int bz = b % g.nb[2];
    Conservatives const &a = g.Blk(b).GetAverageBuffer();
    for (unsigned int e = 0; e < NE; ++e)
      for (unsigned int i = 0; i < IX; ++i)
        for (unsigned int j = 0; j < IY; ++j)
          for (unsigned int k = 0; k < IZ; ++k)
            state[gidx(e, bx * IX + i, by * IY + j, bz * IZ + k)] =
                a[e][i + HX][j + HY][k + HZ];
  }
  if (t_final)
    *t_final = current;
  return steps;
}

// Level-0 node ids of a block grid in the order TopologyManager creates them
// (topology_manager.cpp:109-140: x fastest, then y, then z, each step taken
// with the reference's own neighbour-id arithmetic).  out[nx*ny*nz].
void ref_level0_ids(int const *nb, unsigned long long *out) {
  std::vector<nid_t> list;
  nid_t id = IdSeed();
  for (int i = 0; i < nb[2]; ++i) {
    for (int j = 0; j < nb[1]; ++j) {
      for (int k = 0; k < nb[0]; ++k) {
        list.push_back(id);
        id = EastNeighborOfNodeWithId(list.back());
      }
      id = NorthNeighborOfNodeWithId(list[nb[0] * (j + nb[1] * i)]);
    }
    id = TopNeighborOfNodeWithId(list[(nb[1] * nb[0]) * i]);
  }
  for (size_t n = 0; n < list.size(); ++n) out[n] = list[n];
}

} // extern "C"

// Multiresolution operators on the RHS buffers of two reference Blocks, the
// child at sibling position `position` of a level-0 parent (child id =
// parent id << 3 | position, id_information.h:158-160).  Blocks are
// [NE][TX][TY][TZ] arrays.
namespace {
void LoadRhs(Block &block, double const *in) {
  Conservatives &a = block.GetRightHandSideBuffer();
  for (unsigned int e = 0; e < NE; ++e)
    std::memcpy(&a[e][0][0][0], in + std::size_t(e) * TX * TY * TZ,
                sizeof(double) * TX * TY * TZ);
}
void StoreRhs(Block const &block, double *out) {
  Conservatives const &a = block.GetRightHandSideBuffer();
  for (unsigned int e = 0; e < NE; ++e)
    std::memcpy(out + std::size_t(e) * TX * TY * TZ, &a[e][0][0][0],
                sizeof(double) * TX * TY * TZ);
}
nid_t ChildId(int position) { return (IdSeed() << 3) | nid_t(position); }
} // namespace

extern "C" {

// Multiresolution::Average over the whole Conservatives buffer; parent is
// read and written (only the child's octant changes).
void ref_mr_average(double const *child, double *parent, int position) {
  auto c = std::make_unique<Block>();
  auto p = std::make_unique<Block>();
  LoadRhs(*c, child);
  LoadRhs(*p, parent);
  Multiresolution::Average(c->GetRightHandSideBuffer(),
                           p->GetRightHandSideBuffer(), ChildId(position));
  StoreRhs(*p, parent);
}

// Multiresolution::Prediction of one field [TX][TY][TZ]; child is read and
// written (cells outside the range keep their values).
void ref_mr_prediction(double const *parent, double *child, int position,
                       int const *range) {
  auto c = std::make_unique<Block>();
  auto p = std::make_unique<Block>();
  std::memcpy(&p->GetRightHandSideBuffer(Equation::Mass)[0][0][0], parent,
              sizeof(double) * TX * TY * TZ);
  std::memcpy(&c->GetRightHandSideBuffer(Equation::Mass)[0][0][0], child,
              sizeof(double) * TX * TY * TZ);
  Multiresolution::Prediction(p->GetRightHandSideBuffer(Equation::Mass),
                              c->GetRightHandSideBuffer(Equation::Mass),
                              ChildId(position), range[0], range[1], range[2],
                              range[3], range[4], range[5]);
  std::memcpy(child, &c->GetRightHandSideBuffer(Equation::Mass)[0][0][0],
              sizeof(double) * TX * TY * TZ);
}

// ChildNeedsRemeshing<N> (norm 0 Linfinity, 1 Lone, 2 Ltwo) of a level-1
// child under a Thresholder built from (maximum level 1, reference level 1,
// epsilon): the coarsening threshold on the child's level is then exactly
// `epsilon`.  Returns the RemeshIdentifier (0 Neutral, 1 Refine, 2 Coarse).
int ref_mr_child_needs_remeshing(double const *parent, double const *child,
                                 int position, double epsilon, int norm) {
  auto c = std::make_unique<Block>();
  auto p = std::make_unique<Block>();
  LoadRhs(*c, child);
  LoadRhs(*p, parent);
  Multiresolution mr(Thresholder(1, 1, epsilon));
  nid_t const id = ChildId(position);
  if (norm == 1)
    return int(mr.ChildNeedsRemeshing<Norm::Lone>(*p, *c, id));
  if (norm == 2)
    return int(mr.ChildNeedsRemeshing<Norm::Ltwo>(*p, *c, id));
  return int(mr.ChildNeedsRemeshing<Norm::Linfinity>(*p, *c, id));
}

// Multiresolution::AverageJumpBuffer on SurfaceBuffers given as
// [6][NE][IY][IZ] arrays; parent is read and written.
void ref_mr_average_jump(double const *child, double *parent, int position) {
  auto c = std::make_unique<SurfaceBuffer>();
  auto p = std::make_unique<SurfaceBuffer>();
  std::memcpy(c.get(), child, sizeof(SurfaceBuffer));
  std::memcpy(p.get(), parent, sizeof(SurfaceBuffer));
  Multiresolution::AverageJumpBuffer(*c, *p, ChildId(position));
  std::memcpy(parent, p.get(), sizeof(SurfaceBuffer));
}

// TimeIntegrator<RungeKutta3TVD> on one Node holding Average, RHS, Initial
// and the two jump surfaces (each given as the reference's arrays, read and
// written): the `count` micro time steps are appended, then
//   op 0: PrepareBufferForIntegration(stage) + IntegrateNode(stage, n)
//   op 1: IntegrateJumpHalos(stage, n, start, size)
void ref_lts_integrate(int op, int stage, int n, double const *micro, int count,
                       int const *start, int const *size, double *avg, double *rhs,
                       double *init, double *jump_cons, double *jump_flux) {
  RungeKutta3TVD ti(0.0);
  for (int i = 0; i < count; ++i)
    ti.AppendMicroTimestep(micro[i]);
  auto node = std::make_unique<Node>(IdSeed(), 1.0,
                                     std::vector<MaterialName>{MaterialName::MaterialOne});
  Block &b = node->GetSinglePhase();
  std::size_t const field = sizeof(double) * TX * TY * TZ;
  for (unsigned int e = 0; e < NE; ++e) {
    std::memcpy(&b.GetAverageBuffer()[e][0][0][0], avg + e * TX * TY * TZ, field);
    std::memcpy(&b.GetRightHandSideBuffer()[e][0][0][0], rhs + e * TX * TY * TZ, field);
    std::memcpy(&b.GetInitialBuffer()[e][0][0][0], init + e * TX * TY * TZ, field);
  }
  std::memcpy(&b.GetBoundaryJumpConservatives(), jump_cons, sizeof(SurfaceBuffer));
  std::memcpy(&b.GetBoundaryJumpFluxes(), jump_flux, sizeof(SurfaceBuffer));
  if (op == 0) {
    ti.PrepareBufferForIntegration(*node, unsigned(stage));
    ti.IntegrateNode(*node, unsigned(stage), unsigned(n));
  } else {
    ti.IntegrateJumpHalos(*node, unsigned(stage), unsigned(n), {start[0], start[1], start[2]},
                          {size[0], size[1], size[2]});
  }
  for (unsigned int e = 0; e < NE; ++e) {
    std::memcpy(avg + e * TX * TY * TZ, &b.GetAverageBuffer()[e][0][0][0], field);
    std::memcpy(rhs + e * TX * TY * TZ, &b.GetRightHandSideBuffer()[e][0][0][0], field);
    std::memcpy(init + e * TX * TY * TZ, &b.GetInitialBuffer()[e][0][0][0], field);
  }
  std::memcpy(jump_cons, &b.GetBoundaryJumpConservatives(), sizeof(SurfaceBuffer));
  std::memcpy(jump_flux, &b.GetBoundaryJumpFluxes(), sizeof(SurfaceBuffer));
}

// Reference ThresholdComputer::ThresholdOnLevel.
double ref_mr_threshold(int maximum_level, int reference_level,
                        double reference_epsilon, int level) {
  return Thresholder(unsigned(maximum_level), unsigned(reference_level),
                     reference_epsilon)
      .ThresholdOnLevel(unsigned(level));
}

} // extern "C"
