// Ensemble of independent uniform cases resident in one HBM block pool
// (include/alpaca_gpu.h, alpaca_ensemble_*).
//
// Every admitted case owns a contiguous range of block slots in three pools
// (state u, stage buffers s1, s2).  Its blocks are listed in the work list of
// its class (flux family x arithmetic mode).  One macro step of the whole
// ensemble is a captured hipGraph:
//   ens_dt_kernel                 per-case CFL step (modular_algorithm_
//                                 assembler.cpp:1932-1982) from the signal the
//                                 previous step's last stage reduced;
//   3 x batch_stage_kernel/class  persistent kernels over every tile of every
//                                 listed block (fused faces + divergence + RK
//                                 + halo push; stage 2 also reduces the next
//                                 signal);
//   ens_finish_kernel             per-case FinishMacroTimestep and the stop
//                                 conditions (t_end reached, unhealthy step).
// Work lists and case slots live in device memory, so the graph stays valid
// while cases come and go; only the tuning (tile shape, groups per CU)
// re-captures it.  Each case's arithmetic depends only on its own blocks and
// record, so results are identical alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/alpaca_gpu.h"
#include "fv_device.h"

using namespace alpaca;

namespace {

constexpr int kModeExact = 0, kModeFma = 1, kModeMixed = 2;
constexpr int kClasses = 6;  // flux (2) x arithmetic mode (3)

int mode_of(int precision) {
  return precision == ALPACA_MIXED ? kModeMixed : (precision == ALPACA_FP64_FMA ? kModeFma : kModeExact);
}
int class_of(int flux, int precision) { return flux * 3 + mode_of(precision); }

__global__ void ens_dt_kernel(const CaseSlot *slots, alpaca_case_state *states, int ncases) {
  int const i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncases) return;
  alpaca_case_state &s = states[i];
  int const f = s.flags;
  if (!(f & kFlagActive) || (f & kFlagsStop())) return;
  if (f & kFlagFresh) {  // admission computed the first step already
    s.flags = f & ~kFlagFresh;
    return;
  }
  double const sig = __longlong_as_double((long long)s.smax);
  double const rate = exact::fdiv(sig, slots[i].cell_size);
  double dt = rate == 0.0 ? 1.7976931348623157e308 : exact::fdiv(s.cfl, rate);
  if (s.time + dt > s.t_end) dt = exact::smax(s.t_end - s.time, 0.0);
  s.dt = dt;
  s.smax = 0;
}

__global__ void ens_finish_kernel(alpaca_case_state *states, int ncases) {
  int const i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncases) return;
  alpaca_case_state &s = states[i];
  int f = s.flags;
  if (!(f & kFlagActive) || (f & kFlagsStop())) return;
  s.time += 0.0 + s.dt;
  s.steps += 1;
  if (!(s.time < s.t_end)) f |= kFlagDone;
  if (s.dt < 2.220446049250313080847263336181640625e-16 && s.dt > 0.0) f |= kFlagUnhealthy;
  s.flags = f;
}

__global__ void ens_set_flags_kernel(alpaca_case_state *states, int i, int set, int clear) {
  if (threadIdx.x == 0 && blockIdx.x == 0) states[i].flags = (states[i].flags | set) & ~clear;
}

// Tile configurations searched by the autotuner.
//   3D: 0 = 8x8x8 cells / 256 threads, 1 = 16x8x8 / 512
//   2D: 0 = 16x16 / 128, 1 = 16x16 / 256, 2 = 16x8 / 128
template <int DIM> constexpr int num_tiles_configs() { return DIM == 3 ? 2 : 3; }

template <int DIM, int FLUX, int TX, int TY, int TZ, int NT>
void launch_batch_tile(int groups, hipStream_t s, const double *avg, const double *init,
                       double *out, const CaseSlot *slots, alpaca_case_state *states,
                       const WorkItem *items, const int *n, int stage, int mode) {
  if (mode == kModeExact)
    hipLaunchKernelGGL((exact::batch_stage_kernel<DIM, FLUX, double, TX, TY, TZ, NT>), dim3(groups),
                       dim3(NT), 0, s, avg, init, out, slots, states, items, n, stage);
  else if (mode == kModeFma)
    hipLaunchKernelGGL((fast::batch_stage_kernel<DIM, FLUX, double, TX, TY, TZ, NT>), dim3(groups),
                       dim3(NT), 0, s, avg, init, out, slots, states, items, n, stage);
  else
    hipLaunchKernelGGL((fast::batch_stage_kernel<DIM, FLUX, float, TX, TY, TZ, NT>), dim3(groups),
                       dim3(NT), 0, s, avg, init, out, slots, states, items, n, stage);
}

template <int DIM, int FLUX>
void launch_batch(int tile_config, int groups, hipStream_t s, const double *avg, const double *init,
                  double *out, const CaseSlot *slots, alpaca_case_state *states,
                  const WorkItem *items, const int *n, int stage, int mode) {
  if constexpr (DIM == 3) {
    if (tile_config == 1)
      launch_batch_tile<3, FLUX, 16, 8, 8, 512>(groups, s, avg, init, out, slots, states,
                                                        items, n, stage, mode);
    else
      launch_batch_tile<3, FLUX, 8, 8, 8, 256>(groups, s, avg, init, out, slots, states,
                                                       items, n, stage, mode);
  } else {
    if (tile_config == 1)
      launch_batch_tile<2, FLUX, 16, 16, 1, 256>(groups, s, avg, init, out, slots, states,
                                                         items, n, stage, mode);
    else if (tile_config == 2)
      launch_batch_tile<2, FLUX, 16, 8, 1, 128>(groups, s, avg, init, out, slots, states,
                                                        items, n, stage, mode);
    else
      launch_batch_tile<2, FLUX, 16, 16, 1, 128>(groups, s, avg, init, out, slots, states,
                                                         items, n, stage, mode);
  }
}

}  // namespace

// The HBM block pool of one or two ensembles (a 2D and a 3D one admitted
// against one budget).  Its unit is one 2D haloed block (4 x 24^2 doubles,
// 18,432 B per buffer); a 3D haloed block (5 x 24^3 doubles) is exactly 30
// units, so a 3D case takes a 30-aligned unit range and indexes the same
// arrays in 3D blocks.  Three arrays (u, s1, s2) of `units` units each.
constexpr int kUnitDoubles = 4 * 24 * 24;
constexpr int kUnitsPerBlock3 = 30;
static_assert(5 * 24 * 24 * 24 == kUnitsPerBlock3 * kUnitDoubles, "3D block = 30 units");

struct BlockPool {
  double *u = nullptr, *s1 = nullptr, *s2 = nullptr;
  long long units = 0;
  std::map<long long, long long> free_ranges;  // unit start -> length, coalesced
  std::mutex mu;
  int refs = 0;
};

struct alpaca_ensemble {
  int dim = 0;
  int max_blocks = 0, max_cases = 0;
  int units_per_block = 1;
  BlockPool *pool = nullptr;
  hipStream_t stream = nullptr;
  hipStream_t aux[kClasses] = {};  // branch streams of the captured step
  hipEvent_t fork[3] = {}, join[3][kClasses] = {};
  double *u = nullptr, *s1 = nullptr, *s2 = nullptr;
  CaseSlot *d_slots = nullptr;
  alpaca_case_state *d_states = nullptr;
  WorkItem *d_items[kClasses] = {};
  int *d_counts = nullptr;  // one count per class
  unsigned class_mask = 0;
  std::vector<WorkItem> h_items[kClasses];
  bool items_dirty = false;
  std::vector<CaseSlot> h_slots;
  std::vector<alpaca_case_desc> h_desc;
  std::vector<int> h_active;
  int tile_config = 0;
  int groups_per_cu = 0;  // 0: default from occupancy
  int num_cus = 256;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  std::mutex mu;
};

namespace {

int hip_ok(hipError_t e) { return e == hipSuccess ? ALPACA_OK : ALPACA_ERR_HIP; }

size_t block_bytes(int dim) { return alpaca_block_doubles(dim) * sizeof(double); }

// First fit of n blocks of e's dimension in the shared pool (unit ranges
// aligned to e's block size); *start is in e's blocks.
int alloc_range(alpaca_ensemble *e, int n, int *start) {
  BlockPool &p = *e->pool;
  long long const upb = e->units_per_block;
  long long const need = (long long)n * upb;
  std::lock_guard<std::mutex> lock(p.mu);
  for (auto it = p.free_ranges.begin(); it != p.free_ranges.end(); ++it) {
    long long const rs = it->first, rl = it->second;
    long long const a = (rs + upb - 1) / upb * upb;
    if (a + need > rs + rl) continue;
    p.free_ranges.erase(it);
    if (a > rs) p.free_ranges[rs] = a - rs;
    if (a + need < rs + rl) p.free_ranges[a + need] = rs + rl - (a + need);
    *start = int(a / upb);
    return ALPACA_OK;
  }
  return ALPACA_ERR_CAPACITY;
}

void free_range(alpaca_ensemble *e, int start, int n) {
  BlockPool &p = *e->pool;
  long long const upb = e->units_per_block;
  std::lock_guard<std::mutex> lock(p.mu);
  auto it = p.free_ranges.emplace((long long)start * upb, (long long)n * upb).first;
  auto next = std::next(it);
  if (next != p.free_ranges.end() && it->first + it->second == next->first) {
    it->second += next->second;
    p.free_ranges.erase(next);
  }
  if (it != p.free_ranges.begin()) {
    auto prev = std::prev(it);
    if (prev->first + prev->second == it->first) {
      prev->second += it->second;
      p.free_ranges.erase(it);
    }
  }
}

// Free blocks of e's dimension and the largest range of them.
void pool_capacity(alpaca_ensemble *e, int *free_blocks, int *largest) {
  BlockPool &p = *e->pool;
  long long const upb = e->units_per_block;
  std::lock_guard<std::mutex> lock(p.mu);
  long long tot = 0, big = 0;
  for (auto const &r : p.free_ranges) {
    long long const a = (r.first + upb - 1) / upb * upb;
    long long const nb = a < r.first + r.second ? (r.first + r.second - a) / upb : 0;
    tot += nb;
    big = std::max(big, nb);
  }
  *free_blocks = (int)std::min<long long>(tot, INT32_MAX);
  *largest = (int)std::min<long long>(big, INT32_MAX);
}

int default_groups_per_cu(int dim, int tile_config) {
  if (dim == 3) return tile_config == 1 ? 1 : 2;
  return tile_config == 1 ? 3 : 6;
}

int upload_lists(alpaca_ensemble *e) {
  if (!e->items_dirty) return ALPACA_OK;
  int counts[kClasses];
  for (int c = 0; c < kClasses; ++c) {
    counts[c] = (int)e->h_items[c].size();
    if (counts[c] && hipMemcpyAsync(e->d_items[c], e->h_items[c].data(),
                                    counts[c] * sizeof(WorkItem), hipMemcpyHostToDevice,
                                    e->stream) != hipSuccess)
      return ALPACA_ERR_HIP;
  }
  if (hipMemcpyAsync(e->d_counts, counts, sizeof(counts), hipMemcpyHostToDevice, e->stream) !=
      hipSuccess)
    return ALPACA_ERR_HIP;
  // The host arrays must stay valid until the copies ran.
  if (hipStreamSynchronize(e->stream) != hipSuccess) return ALPACA_ERR_HIP;
  e->items_dirty = false;
  return ALPACA_OK;
}

// One class's persistent stage kernel on stream s.
void launch_class(alpaca_ensemble *e, int c, int groups, hipStream_t s, int stage) {
  double *const ins[3] = {e->u, e->s1, e->s2};
  double *const outs[3] = {e->s1, e->s2, e->u};
  int const flux = c / 3, mode = c % 3;
  if (e->dim == 3) {
    if (flux == 0)
      launch_batch<3, 0>(e->tile_config, groups, s, ins[stage], e->u, outs[stage], e->d_slots,
                         e->d_states, e->d_items[c], e->d_counts + c, stage, mode);
    else
      launch_batch<3, 1>(e->tile_config, groups, s, ins[stage], e->u, outs[stage], e->d_slots,
                         e->d_states, e->d_items[c], e->d_counts + c, stage, mode);
  } else {
    if (flux == 0)
      launch_batch<2, 0>(e->tile_config, groups, s, ins[stage], e->u, outs[stage], e->d_slots,
                         e->d_states, e->d_items[c], e->d_counts + c, stage, mode);
    else
      launch_batch<2, 1>(e->tile_config, groups, s, ins[stage], e->u, outs[stage], e->d_slots,
                         e->d_states, e->d_items[c], e->d_counts + c, stage, mode);
  }
}

// The macro step as captured: dt, then per stage the persistent kernel of
// every enabled class -- the classes of one stage as parallel branches (a
// fork event on the step's stream, one branch stream per further class, a
// join before the next stage), so one class's tail overlaps the others'
// work -- then the per-case finish.  Classes touch disjoint cases, so the
// branches never write the same block.
int record_step(alpaca_ensemble *e) {
  int const g0 = e->groups_per_cu > 0 ? e->groups_per_cu : default_groups_per_cu(e->dim, e->tile_config);
  int groups = g0 * e->num_cus;
  groups = (groups + 7) / 8 * 8;
  int const nc = e->max_cases;
  hipLaunchKernelGGL(ens_dt_kernel, dim3((nc + 255) / 256), dim3(256), 0, e->stream, e->d_slots,
                     e->d_states, nc);
  int cls[kClasses];
  int ncls = 0;
  for (int c = 0; c < kClasses; ++c)
    if (e->class_mask & (1u << c)) cls[ncls++] = c;
  for (int stage = 0; stage < 3; ++stage) {
    if (ncls > 1 && hipEventRecord(e->fork[stage], e->stream) != hipSuccess) return ALPACA_ERR_HIP;
    for (int i = 0; i < ncls; ++i) {
      int const c = cls[i];
      hipStream_t s = i == 0 ? e->stream : e->aux[c];
      if (i > 0 && hipStreamWaitEvent(s, e->fork[stage], 0) != hipSuccess) return ALPACA_ERR_HIP;
      launch_class(e, c, groups, s, stage);
      if (i > 0 && hipEventRecord(e->join[stage][c], s) != hipSuccess) return ALPACA_ERR_HIP;
    }
    for (int i = 1; i < ncls; ++i)
      if (hipStreamWaitEvent(e->stream, e->join[stage][cls[i]], 0) != hipSuccess)
        return ALPACA_ERR_HIP;
  }
  hipLaunchKernelGGL(ens_finish_kernel, dim3((nc + 255) / 256), dim3(256), 0, e->stream,
                     e->d_states, nc);
  return hip_ok(hipGetLastError());
}

void drop_graph(alpaca_ensemble *e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
}

int capture_step(alpaca_ensemble *e) {
  drop_graph(e);
  if (hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return ALPACA_ERR_HIP;
  int const st = record_step(e);
  hipGraph_t g = nullptr;
  hipError_t const end = hipStreamEndCapture(e->stream, &g);
  if (st != ALPACA_OK || end != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return st != ALPACA_OK ? st : ALPACA_ERR_HIP;
  }
  e->graph = g;
  if (hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0) != hipSuccess) {
    drop_graph(e);
    return ALPACA_ERR_HIP;
  }
  return ALPACA_OK;
}

}  // namespace

extern "C" {

size_t alpaca_ensemble_bytes(int dim, int max_blocks, int max_cases) {
  if ((dim != 2 && dim != 3) || max_blocks < 1 || max_cases < 1 || max_blocks > kMaxBlocks ||
      max_cases > kMaxBlocks)
    return 0;
  return 3 * block_bytes(dim) * (size_t)max_blocks +
         (size_t)max_cases * (sizeof(CaseSlot) + sizeof(alpaca_case_state)) +
         (size_t)kClasses * max_blocks * sizeof(WorkItem) + kClasses * sizeof(int);
}

}  // extern "C"

namespace {

void release_pool(BlockPool *p) {
  if (!p) return;
  bool last = false;
  {
    std::lock_guard<std::mutex> lock(p->mu);
    last = --p->refs == 0;
  }
  if (!last) return;
  (void)hipFree(p->u);
  (void)hipFree(p->s1);
  (void)hipFree(p->s2);
  delete p;
}

// Everything of an ensemble but its pool: case slots, records, work lists of
// the enabled classes, branch streams and events of the captured step.
int ensemble_init(alpaca_ensemble *e, BlockPool *pool, int dim, int max_cases,
                  unsigned class_mask, hipStream_t stream) {
  e->dim = dim;
  e->units_per_block = dim == 3 ? kUnitsPerBlock3 : 1;
  e->pool = pool;
  e->max_blocks = (int)std::min<long long>(pool->units / e->units_per_block, kMaxBlocks);
  e->max_cases = max_cases;
  e->class_mask = class_mask;
  e->stream = stream;
  e->u = pool->u;
  e->s1 = pool->s1;
  e->s2 = pool->s2;
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
    e->num_cus = prop.multiProcessorCount;
  bool ok = hipMalloc(&e->d_slots, sizeof(CaseSlot) * max_cases) == hipSuccess &&
            hipMalloc(&e->d_states, sizeof(alpaca_case_state) * max_cases) == hipSuccess &&
            hipMalloc(&e->d_counts, sizeof(int) * kClasses) == hipSuccess;
  for (int c = 0; ok && c < kClasses; ++c)
    if (class_mask & (1u << c))
      ok = hipMalloc(&e->d_items[c], sizeof(WorkItem) * e->max_blocks) == hipSuccess &&
           hipStreamCreateWithFlags(&e->aux[c], hipStreamNonBlocking) == hipSuccess;
  for (int st = 0; ok && st < 3; ++st) {
    ok = hipEventCreateWithFlags(&e->fork[st], hipEventDisableTiming) == hipSuccess;
    for (int c = 0; ok && c < kClasses; ++c)
      if (class_mask & (1u << c))
        ok = hipEventCreateWithFlags(&e->join[st][c], hipEventDisableTiming) == hipSuccess;
  }
  if (ok) {
    ok = hipMemsetAsync(e->d_states, 0, sizeof(alpaca_case_state) * max_cases, e->stream) ==
             hipSuccess &&
         hipMemsetAsync(e->d_counts, 0, sizeof(int) * kClasses, e->stream) == hipSuccess &&
         hipMemsetAsync(e->d_slots, 0, sizeof(CaseSlot) * max_cases, e->stream) == hipSuccess &&
         hipStreamSynchronize(e->stream) == hipSuccess;
  }
  if (!ok) return ALPACA_ERR_HIP;
  e->h_slots.assign(max_cases, CaseSlot{});
  e->h_desc.assign(max_cases, alpaca_case_desc{});
  e->h_active.assign(max_cases, 0);
  return ALPACA_OK;
}

}  // namespace

extern "C" {

int alpaca_ensemble_create(int dim, int max_blocks, int max_cases, unsigned class_mask,
                           void *stream, alpaca_ensemble **out) {
  if (!out || (dim != 2 && dim != 3) || max_blocks < 1 || max_cases < 1 || class_mask == 0 ||
      (class_mask >> kClasses) != 0 || max_blocks > kMaxBlocks || max_cases > kMaxBlocks)
    return ALPACA_ERR_INVALID;
  *out = nullptr;
  auto *pool = new (std::nothrow) BlockPool();
  auto *e = new (std::nothrow) alpaca_ensemble();
  if (!pool || !e) {
    delete pool;
    delete e;
    return ALPACA_ERR_CAPACITY;
  }
  pool->refs = 1;
  pool->units = (long long)max_blocks * (dim == 3 ? kUnitsPerBlock3 : 1);
  size_t const bb = sizeof(double) * kUnitDoubles * (size_t)pool->units;
  e->pool = pool;
  if (hipMalloc(&pool->u, bb) != hipSuccess || hipMalloc(&pool->s1, bb) != hipSuccess ||
      hipMalloc(&pool->s2, bb) != hipSuccess ||
      ensemble_init(e, pool, dim, max_cases, class_mask, static_cast<hipStream_t>(stream)) !=
          ALPACA_OK) {
    alpaca_ensemble_destroy(e);
    return ALPACA_ERR_HIP;
  }
  pool->free_ranges[0] = pool->units;
  *out = e;
  return ALPACA_OK;
}

int alpaca_ensemble_create_shared(alpaca_ensemble *peer, int dim, int max_cases,
                                  unsigned class_mask, void *stream, alpaca_ensemble **out) {
  if (!out || !peer || (dim != 2 && dim != 3) || max_cases < 1 || class_mask == 0 ||
      (class_mask >> kClasses) != 0 || max_cases > kMaxBlocks)
    return ALPACA_ERR_INVALID;
  *out = nullptr;
  auto *e = new (std::nothrow) alpaca_ensemble();
  if (!e) return ALPACA_ERR_CAPACITY;
  {
    std::lock_guard<std::mutex> lock(peer->pool->mu);
    peer->pool->refs += 1;
  }
  e->pool = peer->pool;
  if (ensemble_init(e, peer->pool, dim, max_cases, class_mask, static_cast<hipStream_t>(stream)) !=
      ALPACA_OK) {
    alpaca_ensemble_destroy(e);
    return ALPACA_ERR_HIP;
  }
  *out = e;
  return ALPACA_OK;
}

int alpaca_ensemble_destroy(alpaca_ensemble *e) {
  if (!e) return ALPACA_ERR_INVALID;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  drop_graph(e);
  (void)hipFree(e->d_slots);
  (void)hipFree(e->d_states);
  (void)hipFree(e->d_counts);
  for (int c = 0; c < kClasses; ++c) {
    (void)hipFree(e->d_items[c]);
    if (e->aux[c]) (void)hipStreamDestroy(e->aux[c]);
  }
  for (int st = 0; st < 3; ++st) {
    if (e->fork[st]) (void)hipEventDestroy(e->fork[st]);
    for (int c = 0; c < kClasses; ++c)
      if (e->join[st][c]) (void)hipEventDestroy(e->join[st][c]);
  }
  release_pool(e->pool);
  delete e;
  return ALPACA_OK;
}

int alpaca_ensemble_add_case(alpaca_ensemble *e, const alpaca_case_desc *desc,
                             const double *d_blocks, double t_end, double cfl, int *case_id) {
  if (!e || !desc || !d_blocks || !case_id || desc->dim != e->dim) return ALPACA_ERR_INVALID;
  int const cls = class_of(desc->flux, desc->precision);
  if (desc->flux < 0 || desc->flux > 1 || !(e->class_mask & (1u << cls))) return ALPACA_ERR_UNSUPPORTED;
  int const nb = desc->nb[0] * desc->nb[1] * (e->dim == 3 ? desc->nb[2] : 1);
  if (nb < 1 || !(cfl > 0.0)) return ALPACA_ERR_INVALID;
  std::lock_guard<std::mutex> lock(e->mu);
  int slot = -1;
  for (int i = 0; i < e->max_cases; ++i)
    if (!e->h_active[i]) {
      slot = i;
      break;
    }
  if (slot < 0) return ALPACA_ERR_CAPACITY;
  int base = 0;
  int st = alloc_range(e, nb, &base);
  if (st != ALPACA_OK) return st;
  double *const dst = e->u + (size_t)base * alpaca_block_doubles(e->dim);
  if (hipMemcpyAsync(dst, d_blocks, block_bytes(e->dim) * nb, hipMemcpyDeviceToDevice, e->stream) !=
      hipSuccess) {
    free_range(e, base, nb);
    return ALPACA_ERR_HIP;
  }
  alpaca_case_state *rec = e->d_states + slot;
  if ((st = alpaca_fill_halos(desc, dst, e->stream)) != ALPACA_OK ||
      (st = alpaca_case_state_init(rec, 0.0, t_end, cfl, e->stream)) != ALPACA_OK ||
      (st = alpaca_timestep(desc, dst, rec, e->stream)) != ALPACA_OK) {
    free_range(e, base, nb);
    return st;
  }
  hipLaunchKernelGGL(ens_set_flags_kernel, dim3(1), dim3(64), 0, e->stream, e->d_states, slot,
                     kFlagActive | kFlagFresh, 0);
  CaseSlot cs{};
  for (int a = 0; a < 3; ++a) {
    cs.nb[a] = a < e->dim ? desc->nb[a] : 1;
    cs.bc[a] = a < e->dim ? desc->bc[a] : ALPACA_BC_ZERO_GRADIENT;
  }
  cs.base = base;
  cs.nblocks = nb;
  cs.gamma = desc->gamma;
  cs.b = desc->background_pressure;
  cs.inv_dx = 1.0 / desc->cell_size;
  cs.cell_size = desc->cell_size;
  e->h_slots[slot] = cs;
  e->h_desc[slot] = *desc;
  e->h_active[slot] = 1;
  if (hipMemcpyAsync(e->d_slots + slot, &e->h_slots[slot], sizeof(CaseSlot),
                     hipMemcpyHostToDevice, e->stream) != hipSuccess)
    return ALPACA_ERR_HIP;
  for (int b = 0; b < nb; ++b) e->h_items[cls].push_back(WorkItem{slot, b});
  e->items_dirty = true;
  *case_id = slot;
  return hip_ok(hipGetLastError());
}

int alpaca_ensemble_remove_case(alpaca_ensemble *e, int case_id) {
  if (!e || case_id < 0 || case_id >= e->max_cases) return ALPACA_ERR_INVALID;
  std::lock_guard<std::mutex> lock(e->mu);
  if (!e->h_active[case_id]) return ALPACA_ERR_INVALID;
  int const cls = class_of(e->h_desc[case_id].flux, e->h_desc[case_id].precision);
  auto &v = e->h_items[cls];
  v.erase(std::remove_if(v.begin(), v.end(), [&](WorkItem const &w) { return w.case_id == case_id; }),
          v.end());
  e->items_dirty = true;
  // The work list must drop the case before its slot and blocks are reused.
  int st = upload_lists(e);
  if (st != ALPACA_OK) return st;
  hipLaunchKernelGGL(ens_set_flags_kernel, dim3(1), dim3(64), 0, e->stream, e->d_states, case_id, 0,
                     ~0);
  free_range(e, e->h_slots[case_id].base, e->h_slots[case_id].nblocks);
  e->h_active[case_id] = 0;
  return hip_ok(hipGetLastError());
}

int alpaca_ensemble_cancel_case(alpaca_ensemble *e, int case_id) {
  if (!e || case_id < 0 || case_id >= e->max_cases || !e->h_active[case_id]) return ALPACA_ERR_INVALID;
  hipLaunchKernelGGL(ens_set_flags_kernel, dim3(1), dim3(64), 0, e->stream, e->d_states, case_id,
                     kFlagCancelled, 0);
  return hip_ok(hipGetLastError());
}

int alpaca_ensemble_set_tuning(alpaca_ensemble *e, int tile_config, int groups_per_cu) {
  if (!e || groups_per_cu < 0 || groups_per_cu > 64) return ALPACA_ERR_INVALID;
  int const ntc = e->dim == 3 ? num_tiles_configs<3>() : num_tiles_configs<2>();
  if (tile_config < 0 || tile_config >= ntc) return ALPACA_ERR_INVALID;
  std::lock_guard<std::mutex> lock(e->mu);
  if (tile_config != e->tile_config || groups_per_cu != e->groups_per_cu) {
    e->tile_config = tile_config;
    e->groups_per_cu = groups_per_cu;
    drop_graph(e);
  }
  return ALPACA_OK;
}

int alpaca_ensemble_step(alpaca_ensemble *e, int nsteps) {
  if (!e || nsteps < 0) return ALPACA_ERR_INVALID;
  std::lock_guard<std::mutex> lock(e->mu);
  int st = upload_lists(e);
  if (st != ALPACA_OK) return st;
  if (!e->exec && (st = capture_step(e)) != ALPACA_OK) return st;
  for (int n = 0; n < nsteps; ++n)
    if (hipGraphLaunch(e->exec, e->stream) != hipSuccess) return ALPACA_ERR_HIP;
  return ALPACA_OK;
}

int alpaca_ensemble_read_states(alpaca_ensemble *e, alpaca_case_state *host_out, int n) {
  if (!e || !host_out || n < 0 || n > e->max_cases) return ALPACA_ERR_INVALID;
  if (hipMemcpyAsync(host_out, e->d_states, sizeof(alpaca_case_state) * n, hipMemcpyDeviceToHost,
                     e->stream) != hipSuccess)
    return ALPACA_ERR_HIP;
  return hip_ok(hipStreamSynchronize(e->stream));
}

int alpaca_ensemble_case_blocks(alpaca_ensemble *e, int case_id, double **d_blocks, int *nblocks) {
  if (!e || !d_blocks || !nblocks || case_id < 0 || case_id >= e->max_cases ||
      !e->h_active[case_id])
    return ALPACA_ERR_INVALID;
  *d_blocks = e->u + (size_t)e->h_slots[case_id].base * alpaca_block_doubles(e->dim);
  *nblocks = e->h_slots[case_id].nblocks;
  return ALPACA_OK;
}

int alpaca_ensemble_copy_case(alpaca_ensemble *e, int case_id, double *d_dst) {
  if (!e || !d_dst || case_id < 0 || case_id >= e->max_cases || !e->h_active[case_id])
    return ALPACA_ERR_INVALID;
  CaseSlot const &cs = e->h_slots[case_id];
  return hip_ok(hipMemcpyAsync(d_dst, e->u + (size_t)cs.base * alpaca_block_doubles(e->dim),
                               block_bytes(e->dim) * cs.nblocks, hipMemcpyDeviceToDevice,
                               e->stream));
}

int alpaca_ensemble_capacity(alpaca_ensemble *e, int *free_blocks, int *largest_range,
                             int *free_slots) {
  if (!e || !free_blocks || !largest_range || !free_slots) return ALPACA_ERR_INVALID;
  std::lock_guard<std::mutex> lock(e->mu);
  int tot = 0, big = 0;
  pool_capacity(e, &tot, &big);
  int slots = 0;
  for (int a : e->h_active) slots += a ? 0 : 1;
  *free_blocks = tot;
  *largest_range = big;
  *free_slots = slots;
  return ALPACA_OK;
}

}  // extern "C"
