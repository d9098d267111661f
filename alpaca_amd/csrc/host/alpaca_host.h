// Host C++ adapters over the C ABI (include/alpaca_gpu.h) that keep the call
// shapes of ALPACA's convective-term solver and Runge-Kutta integrator, so a
// CPU call site can switch to the gfx950 path without changing its
// arguments.
//
//   GpuConvectiveTermSolver<DIM>  ConvectiveTermSolver<D>::UpdateConvectiveFluxes
//                                 (convective_term_solver.h:67-81)
//   GpuRungeKutta3TVD<DIM>        TimeIntegrator<RungeKutta3TVD> stage calls
//                                 (time_integrator.h:240-526)
//   GpuUniformSolver<DIM>         ModularAlgorithmAssembler::ComputeLoop /
//                                 Advance on a uniform level 0
//                                 (modular_algorithm_assembler.cpp:97-195, 543-910)
//   GpuThresholdComputer<DIM>     ThresholdComputer<DIM, STDO>
//                                 (threshold_computer.h:35-97)
//   GpuMultiresolution<DIM>       Multiresolution (multiresolution.h:59-188):
//                                 Average, AverageJumpBuffer, Prediction,
//                                 ChildNeedsRemeshing, per pair and batched
//   GpuLevelTimeIntegrator<DIM>   TimeIntegrator<RungeKutta3TVD> with micro
//                                 time steps on a level's block list
//                                 (time_integrator.h:240-526)
//
// Errors: the C ABI returns status codes; these adapters turn a non-OK status
// into AlpacaGpuError (the reference path has no error returns and traps FP
// faults instead).  Device memory is owned by the adapters (RAII).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstring>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/alpaca_gpu.h"

namespace alpaca_host {

class AlpacaGpuError : public std::runtime_error {
 public:
  AlpacaGpuError(int status, std::string const &what)
      : std::runtime_error(what + ": " + alpaca_status_string(status)), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

inline void Check(int status, char const *what) {
  if (status != ALPACA_OK) throw AlpacaGpuError(status, what);
}
inline void CheckHip(hipError_t e, char const *what) {
  if (e != hipSuccess) throw AlpacaGpuError(ALPACA_ERR_HIP, std::string(what) + " (" + hipGetErrorString(e) + ")");
}

// Compile-time block geometry of the reference (compile_time_constants.h:39-44).
template <int DIM> struct Geometry {
  static constexpr int IC = 16, HS = 4, TC = 24;
  static constexpr int NE = DIM + 2;                 // MF::ANOE()
  static constexpr int TCZ = DIM == 3 ? TC : 1;      // CC::TCZ()
  static constexpr int ICZ = DIM == 3 ? IC : 1;      // CC::ICZ()
  static constexpr std::size_t kBlockDoubles = std::size_t(NE) * TC * TC * TCZ;
  using Conservatives = double[NE][TC][TC][TCZ];           // FieldBuffer layout
  using FaceFluxes = double[NE][IC + 1][IC + 1][ICZ + 1];  // face arrays
  using VolumeForces = double[NE][IC][IC][ICZ];
};

// Device buffer with RAII.
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(std::size_t bytes) { Allocate(bytes); }
  DeviceBuffer(DeviceBuffer const &) = delete;
  DeviceBuffer &operator=(DeviceBuffer const &) = delete;
  DeviceBuffer(DeviceBuffer &&o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_) {
    o.ptr_ = nullptr;
    o.bytes_ = 0;
  }
  DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
    std::swap(ptr_, o.ptr_);
    std::swap(bytes_, o.bytes_);
    return *this;
  }
  ~DeviceBuffer() {
    if (ptr_) (void)hipFree(ptr_);
  }
  void Allocate(std::size_t bytes) {
    if (bytes <= bytes_) return;
    if (ptr_) CheckHip(hipFree(ptr_), "hipFree");
    ptr_ = nullptr;
    CheckHip(hipMalloc(&ptr_, bytes), "hipMalloc");
    bytes_ = bytes;
  }
  template <typename T = void> T *get() const { return static_cast<T *>(ptr_); }
  std::size_t bytes() const { return bytes_; }

 private:
  void *ptr_ = nullptr;
  std::size_t bytes_ = 0;
};

inline alpaca_case_desc MakeDesc(int dim, int const nb[3], int const bc[3], int flux,
                                 int precision, double gamma, double background_pressure,
                                 double cell_size) {
  alpaca_case_desc d{};
  d.dim = dim;
  for (int a = 0; a < 3; ++a) {
    d.nb[a] = (a < dim) ? nb[a] : 1;
    d.bc[a] = (a < dim) ? bc[a] : ALPACA_BC_ZERO_GRADIENT;
  }
  d.flux = flux;
  d.precision = precision;
  d.gamma = gamma;
  d.background_pressure = background_pressure;
  d.cell_size = cell_size;
  return d;
}

// ConvectiveTermSolver<D>::UpdateConvectiveFluxes with the reference's
// argument meaning: the block's Average conservatives (halos filled; prime
// states are recovered from them), the cell size, and the three face-flux
// arrays plus volume forces, which are ACCUMULATED into (+=) as
// finite_volume_scheme.cpp:244-246 / flux_splitting_scheme.cpp:309-311 do.
// Volume forces stay untouched (no convective volume term for these
// equations).  The flux family is fixed at construction, as the reference's
// Concretize<ConvectiveTermSolvers> selection fixes it at compile time.
// Owner of an alpaca_context (the stream and face-flux scratch of the
// per-block entries); movable, not copyable.
class Context {
 public:
  explicit Context(hipStream_t stream, int dim = 3, int reserve_blocks = 0) {
    Check(alpaca_context_create(stream, &ctx_), "alpaca_context_create");
    if (reserve_blocks > 0)
      Check(alpaca_context_reserve(ctx_, dim, reserve_blocks), "alpaca_context_reserve");
  }
  Context(Context const &) = delete;
  Context &operator=(Context const &) = delete;
  Context(Context &&o) noexcept : ctx_(o.ctx_) { o.ctx_ = nullptr; }
  Context &operator=(Context &&o) noexcept {
    std::swap(ctx_, o.ctx_);
    return *this;
  }
  ~Context() {
    if (ctx_) (void)alpaca_context_destroy(ctx_);
  }
  alpaca_context *get() const { return ctx_; }

 private:
  alpaca_context *ctx_ = nullptr;
};

// The reference's SpaceSolver owns its convective solver by value, one per
// thread of the rank; each GpuConvectiveTermSolver owns its context, so
// solvers on different streams never share scratch.
template <int DIM> class GpuConvectiveTermSolver {
  using G = Geometry<DIM>;

 public:
  GpuConvectiveTermSolver(int flux_family, double gamma, double background_pressure,
                          int precision = ALPACA_FP64, hipStream_t stream = nullptr)
      : stream_(stream), ctx_(stream, DIM, 1) {
    int const nb[3] = {1, 1, 1};
    int const bc[3] = {ALPACA_BC_ZERO_GRADIENT, ALPACA_BC_ZERO_GRADIENT,
                       ALPACA_BC_ZERO_GRADIENT};
    desc_ = MakeDesc(DIM, nb, bc, flux_family, precision, gamma, background_pressure, 1.0);
    block_.Allocate(sizeof(typename G::Conservatives));
    fluxes_.Allocate(3 * sizeof(typename G::FaceFluxes));
  }

  void UpdateConvectiveFluxes(typename G::Conservatives const &average, double cell_size,
                              typename G::FaceFluxes &fluxes_x, typename G::FaceFluxes &fluxes_y,
                              typename G::FaceFluxes &fluxes_z,
                              typename G::VolumeForces &volume_forces) const {
    (void)volume_forces;
    alpaca_case_desc d = desc_;
    d.cell_size = cell_size;
    std::size_t const fb = sizeof(typename G::FaceFluxes);
    CheckHip(hipMemcpyAsync(block_.get(), &average[0][0][0][0], sizeof(average),
                            hipMemcpyHostToDevice, stream_), "upload block");
    // Accumulate into the caller's arrays: start the device copy from them.
    CheckHip(hipMemcpyAsync(fluxes_.get<char>(), &fluxes_x[0][0][0][0], fb,
                            hipMemcpyHostToDevice, stream_), "upload fluxes");
    CheckHip(hipMemcpyAsync(fluxes_.get<char>() + fb, &fluxes_y[0][0][0][0], fb,
                            hipMemcpyHostToDevice, stream_), "upload fluxes");
    CheckHip(hipMemcpyAsync(fluxes_.get<char>() + 2 * fb, &fluxes_z[0][0][0][0], fb,
                            hipMemcpyHostToDevice, stream_), "upload fluxes");
    Check(alpaca_ctx_convective_fluxes(ctx_.get(), &d, 1, block_.get<double>(),
                                       fluxes_.get<double>()),
          "alpaca_ctx_convective_fluxes");
    CheckHip(hipMemcpyAsync(&fluxes_x[0][0][0][0], fluxes_.get<char>(), fb,
                            hipMemcpyDeviceToHost, stream_), "download fluxes");
    CheckHip(hipMemcpyAsync(&fluxes_y[0][0][0][0], fluxes_.get<char>() + fb, fb,
                            hipMemcpyDeviceToHost, stream_), "download fluxes");
    CheckHip(hipMemcpyAsync(&fluxes_z[0][0][0][0], fluxes_.get<char>() + 2 * fb, fb,
                            hipMemcpyDeviceToHost, stream_), "download fluxes");
    CheckHip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  }

 private:
  alpaca_case_desc desc_;
  hipStream_t stream_;
  Context ctx_;
  mutable DeviceBuffer block_, fluxes_;
};

// Device-resident leaf blocks of one uniform level-0 case: the Average,
// RightHandSide and Initial conservative buffers of every Block
// (block.h:57-60), block (bx,by,bz) at index (bx*nb[1]+by)*nb[2]+bz.
template <int DIM> class DeviceBlocks {
  using G = Geometry<DIM>;

 public:
  DeviceBlocks(alpaca_case_desc const &desc, hipStream_t stream = nullptr)
      : desc_(desc), stream_(stream) {
    nblocks_ = desc.nb[0] * desc.nb[1] * (DIM == 3 ? desc.nb[2] : 1);
    std::size_t const bytes = sizeof(double) * G::kBlockDoubles * nblocks_;
    average_.Allocate(bytes);
    rhs_.Allocate(bytes);
    initial_.Allocate(bytes);
    state_.Allocate(sizeof(alpaca_case_state));
    CheckHip(hipMemsetAsync(average_.get(), 0, bytes, stream_), "memset");
    CheckHip(hipMemsetAsync(rhs_.get(), 0, bytes, stream_), "memset");
    CheckHip(hipMemsetAsync(initial_.get(), 0, bytes, stream_), "memset");
  }
  int NumberOfBlocks() const { return nblocks_; }
  std::size_t BlockDoubles() const { return G::kBlockDoubles; }
  double *Average() const { return average_.get<double>(); }
  double *RightHandSide() const { return rhs_.get<double>(); }
  double *Initial() const { return initial_.get<double>(); }
  alpaca_case_state *State() const { return state_.get<alpaca_case_state>(); }
  alpaca_case_desc const &Desc() const { return desc_; }
  hipStream_t Stream() const { return stream_; }
  // SwapConservativeBuffers<RightHandSide, Average> (time_integrator.h:515-526).
  void SwapRightHandSideAndAverage() { std::swap(average_, rhs_); }

  // Upload/download the internal cells of a domain array [NE][NX][NY][NZ]
  // (k fastest) through host-side block staging; halos are then filled.
  void UploadDomain(std::vector<double> const &domain) {
    std::vector<double> host(G::kBlockDoubles * nblocks_, 0.0);
    ForEachInternal([&](std::size_t b, std::size_t g) { host[b] = domain[g]; });
    CheckHip(hipMemcpyAsync(Average(), host.data(), host.size() * sizeof(double),
                            hipMemcpyHostToDevice, stream_), "upload domain");
    Check(alpaca_fill_halos(&desc_, Average(), stream_), "alpaca_fill_halos");
    CheckHip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  }
  std::vector<double> DownloadDomain() const {
    std::vector<double> host(G::kBlockDoubles * nblocks_);
    CheckHip(hipMemcpyAsync(host.data(), Average(), host.size() * sizeof(double),
                            hipMemcpyDeviceToHost, stream_), "download domain");
    CheckHip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    std::vector<double> domain(DomainCells() * G::NE);
    const_cast<DeviceBlocks *>(this)->ForEachInternal(
        [&](std::size_t b, std::size_t g) { domain[g] = host[b]; });
    return domain;
  }
  std::size_t DomainCells() const {
    return std::size_t(desc_.nb[0]) * G::IC * desc_.nb[1] * G::IC * (DIM == 3 ? desc_.nb[2] * G::IC : 1);
  }
  alpaca_case_state ReadState() const {
    alpaca_case_state s{};
    CheckHip(hipMemcpyAsync(&s, State(), sizeof(s), hipMemcpyDeviceToHost, stream_), "read state");
    CheckHip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    return s;
  }

 private:
  template <typename F> void ForEachInternal(F const &f) {
    std::size_t const NX = std::size_t(desc_.nb[0]) * G::IC, NY = std::size_t(desc_.nb[1]) * G::IC,
                      NZ = DIM == 3 ? std::size_t(desc_.nb[2]) * G::IC : 1;
    int const nbz = DIM == 3 ? desc_.nb[2] : 1;
    for (int bx = 0; bx < desc_.nb[0]; ++bx)
      for (int by = 0; by < desc_.nb[1]; ++by)
        for (int bz = 0; bz < nbz; ++bz) {
          std::size_t const b = (std::size_t(bx) * desc_.nb[1] + by) * nbz + bz;
          for (int e = 0; e < G::NE; ++e)
            for (int i = 0; i < G::IC; ++i)
              for (int j = 0; j < G::IC; ++j)
                for (int k = 0; k < G::ICZ; ++k) {
                  std::size_t const bi =
                      b * G::kBlockDoubles +
                      ((std::size_t(e) * G::TC + i + G::HS) * G::TC + j + G::HS) * G::TCZ +
                      (DIM == 3 ? k + G::HS : 0);
                  std::size_t const gi =
                      ((std::size_t(e) * NX + bx * G::IC + i) * NY + by * G::IC + j) * NZ +
                      (DIM == 3 ? std::size_t(bz) * G::IC + k : 0);
                  f(bi, gi);
                }
        }
  }
  alpaca_case_desc desc_;
  hipStream_t stream_;
  int nblocks_ = 0;
  DeviceBuffer average_, rhs_, initial_, state_;
};

// TimeIntegrator<RungeKutta3TVD> over device blocks (time_integrator.h:
// 240-526, runge_kutta_3_TVD.h:39-59).  The reference's per-node calls
// FillInitialBuffer, UpdateFluxes, PrepareBufferForIntegration and
// IntegrateNode of one stage become one fused device stage; the time step
// is the reference's micro time step (AppendMicroTimestep) held in the
// case's device record.
template <int DIM> class GpuRungeKutta3TVD {
 public:
  static constexpr unsigned int number_of_stages_ = 3;
  constexpr unsigned int NumberOfStages() const { return number_of_stages_; }
  constexpr bool IsLastStage(unsigned int stage) const { return stage == number_of_stages_ - 1; }

  // One stage: out(RHS buffer) = (m0 Average + m1 Initial) + c_s dt RHS(Average),
  // including the halo fill of the RHS buffer.
  void IntegrateStage(DeviceBlocks<DIM> &blocks, unsigned int stage) const {
    if (stage >= number_of_stages_)
      throw std::invalid_argument("Stage is too large for Runge-Kutta-3 integration");
    Check(alpaca_rk3_stage(&blocks.Desc(), int(stage), blocks.Average(), blocks.Initial(),
                           blocks.RightHandSide(), blocks.State(), blocks.Stream()),
          "alpaca_rk3_stage");
  }
  void SwapBuffersForNextStage(DeviceBlocks<DIM> &blocks) const {
    blocks.SwapRightHandSideAndAverage();
  }
};

// Uniform level-0 ComputeLoop / Advance (modular_algorithm_assembler.cpp:
// 97-195, 543-910) with the end-time clamp and the minimum-step health check.
template <int DIM> class GpuUniformSolver {
 public:
  GpuUniformSolver(alpaca_case_desc const &desc, double cfl, double end_time,
                   hipStream_t stream = nullptr)
      : blocks_(desc, stream), cfl_(cfl), end_time_(end_time) {}

  DeviceBlocks<DIM> &Blocks() { return blocks_; }

  void Initialize(std::vector<double> const &domain_conservatives) {
    blocks_.UploadDomain(domain_conservatives);
    Check(alpaca_case_state_init(blocks_.State(), 0.0, end_time_, cfl_, blocks_.Stream()),
          "alpaca_case_state_init");
  }

  // One macro step: ComputeTimestepSize, then the three stages each followed
  // by the buffer swap (the stage kernels fill the halos of their output).
  void Advance() {
    Check(alpaca_timestep(&blocks_.Desc(), blocks_.Average(), blocks_.State(), blocks_.Stream()),
          "alpaca_timestep");
    for (unsigned int stage = 0; stage < integrator_.NumberOfStages(); ++stage) {
      integrator_.IntegrateStage(blocks_, stage);
      integrator_.SwapBuffersForNextStage(blocks_);
    }
    ++steps_;
  }

  // ComputeLoop: advance while time < end_time and the time step is healthy;
  // returns the number of macro steps.  Reads the device record each step.
  int ComputeLoop(int max_steps) {
    int n = 0;
    for (;;) {
      alpaca_case_state s = blocks_.ReadState();
      if (s.flags & 1) throw AlpacaGpuError(ALPACA_ERR_NONFINITE, "ComputeLoop");
      if (!(s.time < end_time_) || (s.flags & 2) || n >= max_steps) break;
      Advance();
      // TimeIntegrator::FinishMacroTimestep (time_integrator.h:244-248).
      Check(alpaca_finish_step(blocks_.State(), blocks_.Stream()), "alpaca_finish_step");
      ++n;
    }
    return n;
  }

  std::vector<double> State() const { return blocks_.DownloadDomain(); }

 private:
  DeviceBlocks<DIM> blocks_;
  GpuRungeKutta3TVD<DIM> integrator_;
  double cfl_, end_time_;
  int steps_ = 0;
};

// ThresholdComputer<DIM, STDO = 3> (threshold_computer.h:35-97).
template <int DIM> class GpuThresholdComputer {
 public:
  GpuThresholdComputer(unsigned int maximum_level, unsigned int user_reference_level,
                       double user_reference_epsilon)
      : maximum_level_(maximum_level),
        reference_epsilon_(user_reference_epsilon *
                           std::pow(2.0, -1.0 * (kAlpha + 1.0) *
                                             (double(maximum_level) - double(user_reference_level)))) {}
  double ThresholdOnLevel(unsigned int level) const {
    return reference_epsilon_ * std::pow(2.0, -1.0 * double(DIM) * (double(maximum_level_) - double(level)));
  }
  unsigned int MaximumLevel() const { return maximum_level_; }

 private:
  static constexpr double kAlpha = 3.0;  // CC::STDO() (compile_time_constants.h:617)
  unsigned int maximum_level_;
  double reference_epsilon_;
};

// RemeshIdentifier (remesh_identifier.h:29).
enum class RemeshIdentifier { Neutral, Refine, Coarse };

// LevelOfNode / PositionOfNodeAmongSiblings (id_information.h:59-69, 158-160).
inline unsigned int LevelOfNodeId(std::uint64_t id) {
  std::uint64_t w = id >> 21;
  unsigned int level = 0;
  while (w > 10) {
    w >>= 3;
    ++level;
  }
  return level;
}
inline int PositionOfNodeIdAmongSiblings(std::uint64_t id) { return int(id & 0x7); }

// Multiresolution (multiresolution.h:59-188) on the device.  The batched
// members take a pool of haloed blocks and a device array of links
// (child, parent, position, level) and run one launch over all of them; the
// reference-shaped members take one parent/child pair in host memory (one
// upload and download per call, for per-pair use and parity checks).
template <int DIM> class GpuMultiresolution {
  using G = Geometry<DIM>;

 public:
  using Field = double[G::TC][G::TC][G::TCZ];
  static constexpr std::size_t kSurfaceDoubles = std::size_t(6) * G::NE * G::IC * G::ICZ;

  explicit GpuMultiresolution(GpuThresholdComputer<DIM> thresholder, int norm = ALPACA_MR_NORM_LINF,
                              hipStream_t stream = nullptr)
      : thresholder_(thresholder), norm_(norm), stream_(stream) {}

  // ---- batched, device-resident
  void AverageLinks(double *d_pool, int nblocks, alpaca_mr_link const *d_links, int n) const {
    Check(alpaca_mr_average(DIM, d_pool, nblocks, d_links, n, stream_), "alpaca_mr_average");
  }
  void AverageJumpBufferLinks(double *d_surfaces, int nblocks, alpaca_mr_link const *d_links,
                              int n) const {
    Check(alpaca_mr_average_jump_buffers(DIM, d_surfaces, nblocks, d_links, n, stream_),
          "alpaca_mr_average_jump_buffers");
  }
  void PredictionLinks(double *d_pool, int nblocks, alpaca_mr_link const *d_links, int n,
                       int const range[6]) const {
    Check(alpaca_mr_prediction(DIM, d_pool, nblocks, d_links, n, range, stream_),
          "alpaca_mr_prediction");
  }
  // InternalHaloManager::UpdateMaterialJumpNoMpi for a batch of (child,
  // BoundaryLocation) pairs: d_locations (device, n ints) in the enum's order.
  void JumpHaloUpdateLinks(double *d_pool, int nblocks, alpaca_mr_link const *d_links,
                           int const *d_locations, int n) const {
    Check(alpaca_mr_jump_halo_update(DIM, d_pool, nblocks, d_links, d_locations, n, stream_),
          "alpaca_mr_jump_halo_update");
  }
  // Details (device, n doubles) and RemeshIdentifier values (device, n ints)
  // on the thresholds of levels 0..MaximumLevel.
  void ChildNeedsRemeshingLinks(double const *d_pool, int nblocks, alpaca_mr_link const *d_links,
                                int n, double *d_details, int *d_decisions) const {
    std::vector<double> eps(thresholder_.MaximumLevel() + 1);
    for (unsigned int l = 0; l < eps.size(); ++l) eps[l] = thresholder_.ThresholdOnLevel(l);
    Check(alpaca_mr_child_needs_remeshing(DIM, d_pool, nblocks, d_links, n, norm_, eps.data(),
                                          int(eps.size()), d_details, d_decisions, stream_),
          "alpaca_mr_child_needs_remeshing");
  }

  // ---- reference-shaped, one pair in host memory
  // Average(child_buffer, parent_buffer, child_id): overrides the child's
  // octant of the parent's internal cells.
  void Average(typename G::Conservatives const &child, typename G::Conservatives &parent,
               std::uint64_t child_id) const {
    PairCall(&parent, &child, child_id, [&](double *d, alpaca_mr_link const *l) { AverageLinks(d, 2, l, 1); },
             true, false);
  }
  // Prediction(parent_field, child_field, child_id, x_start, x_count, ...).
  void Prediction(Field const &parent, Field &child, std::uint64_t child_id, unsigned int x_start = 0,
                  unsigned int x_count = G::TC, unsigned int y_start = 0, unsigned int y_count = G::TC,
                  unsigned int z_start = 0, unsigned int z_count = G::TCZ) const {
    int const range[6] = {int(x_start), int(x_count), int(y_start), int(y_count), int(z_start), int(z_count)};
    std::vector<double> p(G::kBlockDoubles, 0.0), c(G::kBlockDoubles, 0.0);
    std::memcpy(p.data(), &parent[0][0][0], sizeof(Field));
    std::memcpy(c.data(), &child[0][0][0], sizeof(Field));
    PairCall(p.data(), c.data(), child_id,
             [&](double *d, alpaca_mr_link const *l) { PredictionLinks(d, 2, l, 1, range); }, false, true);
    std::memcpy(&child[0][0][0], c.data(), sizeof(Field));
  }
  // ChildNeedsRemeshing<norm>(parent, child, child_id) on the blocks given
  // (the reference passes the right-hand-side buffers).
  RemeshIdentifier ChildNeedsRemeshing(typename G::Conservatives const &parent,
                                       typename G::Conservatives const &child,
                                       std::uint64_t child_id, double *detail = nullptr) const {
    DeviceBuffer pool(2 * sizeof(double) * G::kBlockDoubles), link(sizeof(alpaca_mr_link)),
        out(sizeof(double) + sizeof(int));
    Upload(pool.get<double>(), &parent, G::kBlockDoubles);
    Upload(pool.get<double>() + G::kBlockDoubles, &child, G::kBlockDoubles);
    alpaca_mr_link const l{1, 0, PositionOfNodeIdAmongSiblings(child_id), int(LevelOfNodeId(child_id))};
    CheckHip(hipMemcpyAsync(link.get(), &l, sizeof(l), hipMemcpyHostToDevice, stream_), "upload link");
    ChildNeedsRemeshingLinks(pool.get<double>(), 2, link.get<alpaca_mr_link>(), 1, out.get<double>(),
                             reinterpret_cast<int *>(out.get<char>() + sizeof(double)));
    struct {
      double d;
      int r;
    } h{};
    CheckHip(hipMemcpyAsync(&h.d, out.get(), sizeof(double), hipMemcpyDeviceToHost, stream_), "download");
    CheckHip(hipMemcpyAsync(&h.r, out.get<char>() + sizeof(double), sizeof(int), hipMemcpyDeviceToHost,
                            stream_), "download");
    CheckHip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    if (detail) *detail = h.d;
    if (h.r < 0) throw AlpacaGpuError(ALPACA_ERR_INVALID, "ChildNeedsRemeshing: level beyond the thresholds");
    return RemeshIdentifier(h.r);
  }
  // AverageJumpBuffer(child_surfaces, parent_surfaces, child_id): surfaces of
  // kSurfaceDoubles doubles in SurfaceBuffer layout; adds into the parent.
  void AverageJumpBuffer(double const *child, double *parent, std::uint64_t child_id) const {
    DeviceBuffer pool(2 * sizeof(double) * kSurfaceDoubles), link(sizeof(alpaca_mr_link));
    Upload(pool.get<double>(), parent, kSurfaceDoubles);
    Upload(pool.get<double>() + kSurfaceDoubles, child, kSurfaceDoubles);
    alpaca_mr_link const l{1, 0, PositionOfNodeIdAmongSiblings(child_id), int(LevelOfNodeId(child_id))};
    CheckHip(hipMemcpyAsync(link.get(), &l, sizeof(l), hipMemcpyHostToDevice, stream_), "upload link");
    AverageJumpBufferLinks(pool.get<double>(), 2, link.get<alpaca_mr_link>(), 1);
    CheckHip(hipMemcpyAsync(parent, pool.get(), sizeof(double) * kSurfaceDoubles, hipMemcpyDeviceToHost,
                            stream_), "download");
    CheckHip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  }

 private:
  void Upload(double *d, void const *h, std::size_t n) const {
    CheckHip(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, stream_), "upload");
  }
  // Pool [parent, child] on the device, one link; downloads the parent or
  // the child afterwards.
  template <typename F>
  void PairCall(void *parent, void const *child, std::uint64_t child_id, F const &f, bool back_parent,
                bool back_child) const {
    DeviceBuffer pool(2 * sizeof(double) * G::kBlockDoubles), link(sizeof(alpaca_mr_link));
    Upload(pool.get<double>(), parent, G::kBlockDoubles);
    Upload(pool.get<double>() + G::kBlockDoubles, child, G::kBlockDoubles);
    alpaca_mr_link const l{1, 0, PositionOfNodeIdAmongSiblings(child_id), int(LevelOfNodeId(child_id))};
    CheckHip(hipMemcpyAsync(link.get(), &l, sizeof(l), hipMemcpyHostToDevice, stream_), "upload link");
    f(pool.get<double>(), link.get<alpaca_mr_link>());
    if (back_parent)
      CheckHip(hipMemcpyAsync(parent, pool.get(), sizeof(double) * G::kBlockDoubles, hipMemcpyDeviceToHost,
                              stream_), "download");
    if (back_child)
      CheckHip(hipMemcpyAsync(const_cast<void *>(child), pool.get<double>() + G::kBlockDoubles,
                              sizeof(double) * G::kBlockDoubles, hipMemcpyDeviceToHost, stream_),
               "download");
    CheckHip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  }

  GpuThresholdComputer<DIM> thresholder_;
  int norm_;
  hipStream_t stream_;
};

// TimeIntegrator<RungeKutta3TVD> for the multi-level mode
// (time_integrator.h:240-526): micro time steps appended on the finest
// level; a level advancing with n of them integrates with their sum (newest
// first, from 0.0) times the stage multiplier.  The per-node calls take a
// device list of the level's pool blocks and run one launch each.
template <int DIM> class GpuLevelTimeIntegrator {
 public:
  explicit GpuLevelTimeIntegrator(double start_time = 0.0, hipStream_t stream = nullptr)
      : current_run_time_(start_time), stream_(stream) {}
  static constexpr unsigned int NumberOfStages() { return 3; }
  static constexpr bool IsLastStage(unsigned int stage) { return stage == 2; }
  void AppendMicroTimestep(double dt) { micro_.push_back(dt); }
  std::vector<double> const &MicroTimestepSizes() const { return micro_; }
  double CurrentRunTime() const { return current_run_time_; }
  void FinishMacroTimestep() {
    double s = 0.0;
    for (double dt : micro_) s += dt;
    current_run_time_ += s;
    micro_.clear();
  }
  double Timestep(unsigned int stage, unsigned int n, bool jump = false) const {
    double out = 0.0;
    Check(alpaca_lts_timestep(micro_.data(), int(micro_.size()), int(n), int(stage), jump ? 1 : 0, &out),
          "alpaca_lts_timestep");
    return out;
  }
  // PrepareBufferForIntegration on the listed blocks.
  void PrepareBufferForIntegration(double *d_avg, double const *d_init, int nblocks, int const *d_list,
                                   int n, unsigned int stage) const {
    Check(alpaca_lts_prepare_buffer(DIM, d_avg, d_init, nblocks, d_list, n, int(stage), stream_),
          "alpaca_lts_prepare_buffer");
  }
  // IntegrateNode: jump conservatives (jump multiplier), then the internal
  // cells rhs = avg + timestep * rhs.
  void IntegrateNode(double const *d_avg, double *d_rhs, double *d_jump_cons, double *d_jump_flux,
                     int nblocks, int const *d_list, int n, unsigned int stage,
                     unsigned int number_of_timesteps) const {
    if (d_jump_cons)
      Check(alpaca_lts_integrate_jump_conservatives(DIM, d_jump_cons, d_jump_flux, nblocks, d_list, n,
                                                    Timestep(stage, number_of_timesteps, true), stream_),
            "alpaca_lts_integrate_jump_conservatives");
    Check(alpaca_lts_integrate(DIM, d_avg, d_rhs, nblocks, d_list, n, nullptr,
                               Timestep(stage, number_of_timesteps), stream_),
          "alpaca_lts_integrate");
  }
  // IntegrateJumpHalos: the halo range (start, size per axis) of the listed
  // blocks, with the conservatives' time step.
  void IntegrateJumpHalos(double const *d_avg, double *d_rhs, int nblocks, int const *d_list, int n,
                          unsigned int stage, unsigned int number_of_timesteps,
                          int const start[3], int const size[3]) const {
    int const range[6] = {start[0], size[0], start[1], size[1], start[2], size[2]};
    Check(alpaca_lts_integrate(DIM, d_avg, d_rhs, nblocks, d_list, n, range,
                               Timestep(stage, number_of_timesteps), stream_),
          "alpaca_lts_integrate (halo)");
  }

 private:
  double current_run_time_;
  hipStream_t stream_;
  std::vector<double> micro_;
};

}  // namespace alpaca_host
