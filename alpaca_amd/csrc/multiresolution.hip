// Multiresolution operators of ALPACA between parent and child blocks held in
// one pool of haloed blocks on gfx950: conservative averaging of a child into
// its parent (Multiresolution::Average, multiresolution.h:94-167; of its jump
// surfaces, AverageJumpBuffer, multiresolution.cpp:47-131), Harten's
// fifth-order prediction of a parent into a child (Multiresolution::
// Prediction, multiresolution.cpp:146-292) and the L-infinity wavelet detail
// with the remeshing decision (ChildNeedsRemeshing<Norm::Linfinity>,
// multiresolution.cpp:299-323; RemeshingDecision, :408-428).
//
// Every expression keeps the reference's operation order and the translation
// unit is built without contraction, so the results are the reference's bits;
// std::min/std::max picks (first extreme, NaN comparisons false) are spelled
// out so that non-finite inputs propagate as they do there.
//
// Work decomposition.  A link names (child block, parent block, sibling
// position, child level).  Averaging is one thread per two parent cells
// written (adjacent along the fastest axis, so the child reads of neighbouring
// threads are contiguous 16-byte accesses).  Prediction and details stage the parent's stencil window --
// the predicted parent cells plus two on each side, at most 16 per axis
// (32 KiB of doubles in 3D) -- in LDS once per field and let each thread
// predict the 2^DIM children of one parent cell from it: the 3D stencil reads
// 125 parent values per parent cell, which LDS serves instead of L1/L2.
#include <hip/hip_runtime.h>

#include <cstring>

#include "fv_device.h"

namespace alpaca {
namespace mr {
namespace {

constexpr int kThreads = 256;         // averaging
constexpr int kStencilThreads = 192;  // prediction and details: 1728 = 9 x 192 parent cells
constexpr int kMaxLevels = 32;

// LDS layout of a parent stencil window: [W][W][W] doubles in 3D, [W][W] in
// 2D, whatever part of it a range fills, so that every stencil offset is a
// compile-time constant.  W = 16 holds any window; W = 12 the windows of
// ranges of at most 8 predicted parent cells per axis (jump halo slabs, the
// internal cells).
template <int DIM, int W> struct LdsW {
  static constexpr int SX = DIM == 3 ? W * W : W;
  static constexpr int SY = DIM == 3 ? W : 1;
  static constexpr int SIZE = W * SX;
};
template <int DIM> using Lds = LdsW<DIM, 16>;

// ConsistencyManagedSum, three and four values (mathematical_functions.h:79-91,
// 115-124): 0.5 (min + max) of the three pairings s1, s2, s3, where std::min
// and std::max over the initializer list keep the first extreme and never
// pick a NaN by comparison.  The hardware minimum/maximum (one instruction
// each) pick the same value whenever s1 is not NaN:
//  * a NaN among s2, s3 is skipped by both (std: the comparison is false);
//  * ties: a rounded sum is -0 only when both addends are -0, so the s_i are
//    zeros of one sign or of none -- equal values are the same bits.
// When s1 is NaN, std keeps it for both extremes and the result is NaN.
__device__ __forceinline__ double cms_pick(double s1, double s2, double s3) {
  double const mn = __builtin_fmin(__builtin_fmin(s1, s2), s3);
  double const mx = __builtin_fmax(__builtin_fmax(s1, s2), s3);
  double const r = 0.5 * (mn + mx);
  return s1 != s1 ? s1 : r;
}
__device__ __forceinline__ double cms3(double a, double b, double c) {
  return cms_pick((a + b) + c, (a + c) + b, (b + c) + a);
}
__device__ __forceinline__ double cms4(double a, double b, double c, double d) {
  return cms_pick((a + b) + (c + d), (a + c) + (b + d), (a + d) + (b + c));
}
// DimensionAwareConsistencyManagedSum (mathematical_functions.h:98-108).
template <int DIM> __device__ __forceinline__ double dams(double a, double b, double c) {
  if constexpr (DIM == 2) {
    (void)c;
    return a + b;
  } else {
    return cms3(a, b, c);
  }
}

template <int DIM> __device__ __forceinline__ bool link_ok(const alpaca_mr_link &l, int nblocks) {
  return l.child >= 0 && l.child < nblocks && l.parent >= 0 && l.parent < nblocks &&
         l.child != l.parent && l.position >= 0 && l.position < (1 << DIM);
}

// ---------------------------------------------------------------- averaging

// Parent internal cells covered by one child: IC/2 per axis, starting at HS
// (low child) or TC/2 (high child) (xyz_look_up_table_, multiresolution.h:
// 35-45); child cells from HS in steps of two.  Each thread writes two parent
// cells adjacent along the fastest axis (k in 3D, j in 2D) and reads the four
// contiguous child cells under them per child row as two 16-byte loads; every
// such pair starts at an even element offset of a block (16-byte aligned).
template <int DIM> struct AvgShape {
  static constexpr int H = kIC / 2;                  // parent cells per axis
  static constexpr int PAIRS = DIM == 3 ? H * H * (H / 2) : H * (H / 2);  // threads per field
};

template <int DIM>
__global__ __launch_bounds__(kThreads) void average_kernel(double *__restrict__ blocks, int nblocks,
                                                           const alpaca_mr_link *__restrict__ links,
                                                           int nlinks) {
  using D = Dims<DIM>;
  constexpr int H = AvgShape<DIM>::H;
  constexpr int PAIRS = AvgShape<DIM>::PAIRS;
  constexpr int SX = D::SX, SY = D::SY;
  size_t const t = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= (size_t)nlinks * D::NE * PAIRS) return;
  int const r = int(t % PAIRS);
  int const e = int((t / PAIRS) % D::NE);
  alpaca_mr_link const l = links[t / ((size_t)PAIRS * D::NE)];
  if (!link_ok<DIM>(l, nblocks)) return;
  const double *c = blocks + (size_t)l.child * D::BLOCK + (size_t)e * D::FIELD;
  double *p = blocks + (size_t)l.parent * D::BLOCK + (size_t)e * D::FIELD;
  int const x0 = (l.position & 1) ? kTC / 2 : kHS;
  int const y0 = (l.position & 2) ? kTC / 2 : kHS;
  if constexpr (DIM == 2) {
    int const i = r / (H / 2), j = 2 * (r % (H / 2));
    const double *q = c + (kHS + 2 * i) * SX + (kHS + 2 * j) * SY;
    double2 const a0 = *reinterpret_cast<const double2 *>(q);           // (ci,   cj..cj+1)
    double2 const a1 = *reinterpret_cast<const double2 *>(q + 2);       // (ci,   cj+2..cj+3)
    double2 const b0 = *reinterpret_cast<const double2 *>(q + SX);      // (ci+1, cj..cj+1)
    double2 const b1 = *reinterpret_cast<const double2 *>(q + SX + 2);  // (ci+1, cj+2..cj+3)
    double2 out;
    out.x = ((a0.x + b0.y) + (b0.x + a0.y)) * 0.25;
    out.y = ((a1.x + b1.y) + (b1.x + a1.y)) * 0.25;
    *reinterpret_cast<double2 *>(p + (x0 + i) * SX + (y0 + j) * SY) = out;
  } else {
    int const i = r / (H * H / 2), j = (r / (H / 2)) % H, k = 2 * (r % (H / 2));
    int const z0 = (l.position & 4) ? kTC / 2 : kHS;
    const double *q = c + (kHS + 2 * i) * SX + (kHS + 2 * j) * SY + kHS + 2 * k;
    // Rows (dx, dy) of four child cells along k.
    double2 const r00a = *reinterpret_cast<const double2 *>(q);
    double2 const r00b = *reinterpret_cast<const double2 *>(q + 2);
    double2 const r01a = *reinterpret_cast<const double2 *>(q + SY);
    double2 const r01b = *reinterpret_cast<const double2 *>(q + SY + 2);
    double2 const r10a = *reinterpret_cast<const double2 *>(q + SX);
    double2 const r10b = *reinterpret_cast<const double2 *>(q + SX + 2);
    double2 const r11a = *reinterpret_cast<const double2 *>(q + SX + SY);
    double2 const r11b = *reinterpret_cast<const double2 *>(q + SX + SY + 2);
    double2 out;
    out.x = 0.125 * cms4(r00a.x + r11a.y, r00a.y + r11a.x, r01a.x + r10a.y, r01a.y + r10a.x);
    out.y = 0.125 * cms4(r00b.x + r11b.y, r00b.y + r11b.x, r01b.x + r10b.y, r01b.y + r10b.x);
    *reinterpret_cast<double2 *>(p + (x0 + i) * SX + (y0 + j) * SY + z0 + k) = out;
  }
}

// ------------------------------------------------------- jump surfaces

// Multiresolution::AverageJumpBuffer (multiresolution.cpp:47-131) on jump
// surfaces [6][NE][IC][ICZ] per block (SurfaceBuffer, block.h:34-41: East,
// West, North, South, Top, Bottom).  One thread per parent face entry of
// every face a child may touch; a child touches the high face of an axis when
// its bit of that axis is set, the low face otherwise, and covers the quarter
// (half in 2D) of it starting at IC/2 times its bits of the face's transverse
// axes.  The average is ADDED to the parent's value.
template <int DIM>
__global__ __launch_bounds__(kThreads) void average_jump_kernel(double *__restrict__ surfaces,
                                                                int nblocks,
                                                                const alpaca_mr_link *__restrict__ links,
                                                                int nlinks) {
  constexpr int NE = Dims<DIM>::NE;
  constexpr int ICZ = DIM == 3 ? kIC : 1;
  constexpr int HJ = DIM == 3 ? kIC / 2 : 1;      // parent entries along the second index
  constexpr int ENTRIES = NE * (kIC / 2) * HJ;    // per face
  constexpr int FACE = NE * kIC * ICZ;            // doubles per face
  constexpr int BLOCK = 6 * FACE;                 // doubles per block's surfaces
  size_t const t = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= (size_t)nlinks * DIM * ENTRIES) return;
  int const r = int(t % ENTRIES);
  int const axis = int((t / ENTRIES) % DIM);
  alpaca_mr_link const l = links[t / ((size_t)ENTRIES * DIM)];
  if (!link_ok<DIM>(l, nblocks)) return;
  // The face of this axis the child touches: East/North/Top (2 axis) when its
  // bit is set, West/South/Bottom (2 axis + 1) otherwise.
  int const loc = 2 * axis + (((l.position >> axis) & 1) ? 0 : 1);
  int const t1 = axis == 0 ? 1 : 0, t2 = axis == 2 ? 1 : 2;
  int const s1 = ((l.position >> t1) & 1) * (kIC / 2);
  int const s2 = DIM == 3 ? ((l.position >> t2) & 1) * (kIC / 2) : 0;
  int const e = r / ((kIC / 2) * HJ), i = (r / HJ) % (kIC / 2), j = r % HJ;
  const double *q = surfaces + (size_t)l.child * BLOCK + loc * FACE + (e * kIC + 2 * i) * ICZ +
                    (DIM == 3 ? 2 * j : 0);
  double v;
  if constexpr (DIM == 2)
    v = (q[0] + q[ICZ]) * 0.5;
  else
    v = ((q[0] + q[ICZ + 1]) + (q[ICZ] + q[1])) * 0.25;
  surfaces[(size_t)l.parent * BLOCK + loc * FACE + (e * kIC + s1 + i) * ICZ + s2 + j] += v;
}

// --------------------------------------------------------------- prediction

// Window of parent cells a prediction reads along one axis: the predicted
// parent cells [p0, p0 + n) (n = count / 2, p0 = start / 2 + 2 for the low
// child, start / 2 + TC / 2 - 2 for the high one; PIOLCH/PIOHCH,
// compile_time_constants.h:598-608) and two more on each side.
struct Range {
  int start[3];
  int count[3];  // predicted parent cells per axis (child count / 2)
};

// The 2^DIM predicted child values of the parent cell at LDS index `o` of a
// window (Lds layout): out[dx | dy << 1 | dz << 2].  In 2D the
// z terms are the reference's zero-initialised Qz, Qxz, Qyz, Qxyz, and enter
// the sums as such (their signs matter for signed zeros).
template <int DIM, typename L = Lds<DIM>>
__device__ __forceinline__ void predict_cell(const double *__restrict__ s, int o,
                                             double out[1 << DIM]) {
  constexpr int sx = L::SX, sy = L::SY;
  constexpr double c0 = -22.0 / 128.0, c1 = 3.0 / 128.0;
  constexpr double c00 = c0 * c0, c01 = c0 * c1, c10 = c01, c11 = c1 * c1;
  constexpr double c000 = c00 * c0, c001 = c00 * c1, c010 = c001, c011 = c0 * c11;
  constexpr double c100 = c001, c101 = c011, c110 = c011, c111 = c1 * c11;
#define P(a, b, c) s[o + (a) * sx + (b) * sy + (c)]
  double Qz = 0.0, Qxz = 0.0, Qyz = 0.0, Qxyz = 0.0;
  double const Qx = c0 * (P(1, 0, 0) - P(-1, 0, 0)) + c1 * (P(2, 0, 0) - P(-2, 0, 0));
  double const Qy = c0 * (P(0, 1, 0) - P(0, -1, 0)) + c1 * (P(0, 2, 0) - P(0, -2, 0));
  double const Qxy =
      (c00 * ((P(1, 1, 0) + P(-1, -1, 0)) - (P(-1, 1, 0) + P(1, -1, 0))) +
       c11 * ((P(2, 2, 0) + P(-2, -2, 0)) - (P(-2, 2, 0) + P(2, -2, 0)))) +
      (c01 * ((P(1, 2, 0) + P(-1, -2, 0)) - (P(-1, 2, 0) + P(1, -2, 0))) +
       c10 * ((P(2, 1, 0) + P(-2, -1, 0)) - (P(-2, 1, 0) + P(2, -1, 0))));
  if constexpr (DIM == 3) {
    Qz = c0 * (P(0, 0, 1) - P(0, 0, -1)) + c1 * (P(0, 0, 2) - P(0, 0, -2));
    Qxz = (c00 * ((P(1, 0, 1) + P(-1, 0, -1)) - (P(-1, 0, 1) + P(1, 0, -1))) +
           c11 * ((P(2, 0, 2) + P(-2, 0, -2)) - (P(-2, 0, 2) + P(2, 0, -2)))) +
          (c01 * ((P(1, 0, 2) + P(-1, 0, -2)) - (P(-1, 0, 2) + P(1, 0, -2))) +
           c10 * ((P(2, 0, 1) + P(-2, 0, -1)) - (P(-2, 0, 1) + P(2, 0, -1))));
    Qyz = (c00 * ((P(0, 1, 1) + P(0, -1, -1)) - (P(0, -1, 1) + P(0, 1, -1))) +
           c11 * ((P(0, 2, 2) + P(0, -2, -2)) - (P(0, -2, 2) + P(0, 2, -2)))) +
          (c01 * ((P(0, 1, 2) + P(0, -1, -2)) - (P(0, -1, 2) + P(0, 1, -2))) +
           c10 * ((P(0, 2, 1) + P(0, -2, -1)) - (P(0, -2, 1) + P(0, 2, -1))));
    // The four antisymmetric differences over the corners (+-a, +-b, +-c).
#define G(a, b, c)                                     \
  cms4(P(a, b, c) - P(-(a), -(b), -(c)), P(-(a), -(b), c) - P(a, b, -(c)), \
       P(a, -(b), -(c)) - P(-(a), b, c), P(-(a), b, -(c)) - P(a, -(b), c))
    Qxyz = cms4(c000 * G(1, 1, 1) + c111 * G(2, 2, 2), c001 * G(1, 1, 2) + c110 * G(2, 2, 1),
                c011 * G(1, 2, 2) + c100 * G(2, 1, 1), c010 * G(1, 2, 1) + c101 * G(2, 1, 2));
#undef G
  }
  double const p = P(0, 0, 0);
#undef P
  out[0] = p + dams<DIM>(Qx, Qy, Qz) + dams<DIM>(Qxy, Qxz, Qyz) + Qxyz;
  out[1] = p + dams<DIM>(-Qx, Qy, Qz) + dams<DIM>(-Qxy, -Qxz, Qyz) - Qxyz;
  out[2] = p + dams<DIM>(Qx, -Qy, Qz) + dams<DIM>(-Qxy, Qxz, -Qyz) - Qxyz;
  out[3] = p + dams<DIM>(-Qx, -Qy, Qz) + dams<DIM>(Qxy, -Qxz, -Qyz) + Qxyz;
  if constexpr (DIM == 3) {
    out[4] = p + cms3(Qx, Qy, -Qz) + cms3(Qxy, -Qxz, -Qyz) - Qxyz;
    out[5] = p + cms3(-Qx, Qy, -Qz) + cms3(-Qxy, Qxz, -Qyz) + Qxyz;
    out[6] = p + cms3(Qx, -Qy, -Qz) + cms3(-Qxy, -Qxz, Qyz) + Qxyz;
    out[7] = p + cms3(-Qx, -Qy, -Qz) + cms3(Qxy, Qxz, Qyz) - Qxyz;
  }
}

// Window geometry of one link: first parent cell of the window per axis and
// its extent (at most 16).
template <int DIM> struct Window {
  int p0[3];  // first window cell (predicted cell - 2) per axis
  int w[3];   // window extent per axis
  __device__ Window(const Range &r, int position) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      int const off = ((position >> a) & 1) ? kTC / 2 - kHS / 2 : kHS / 2;
      bool const live = a < DIM;
      p0[a] = live ? r.start[a] / 2 + off - 2 : 0;
      w[a] = live ? r.count[a] + 4 : 1;
    }
  }
  // LDS index of predicted parent cell (a, b, c) (counted from the range's
  // first predicted cell).
  template <typename L = Lds<DIM>> __device__ static int centre(int a, int b, int c) {
    return (a + 2) * L::SX + (b + 2) * L::SY + (DIM == 3 ? c + 2 : 0);
  }
};

template <int DIM>
__device__ __forceinline__ void stage_window(double *s, const double *__restrict__ field,
                                             const Window<DIM> &win) {
  using D = Dims<DIM>;
  int const n = win.w[0] * win.w[1] * win.w[2];
  for (int t = threadIdx.x; t < n; t += kStencilThreads) {
    int const a = t / (win.w[1] * win.w[2]), b = (t / win.w[2]) % win.w[1], c = t % win.w[2];
    s[a * Lds<DIM>::SX + b * Lds<DIM>::SY + c] =
        field[(win.p0[0] + a) * D::SX + (win.p0[1] + b) * D::SY + (win.p0[2] + c)];
  }
}

// Child cells (2a + dx, 2b + dy, 2c + dz) from `q`, the child's cell
// (2a, 2b, 2c): the two cells adjacent along the fastest axis (dz in 3D, dy in
// 2D) are one 16-byte access (the offset is even: aligned).
template <int DIM> __device__ __forceinline__ void store_children(double *q, const double v[1 << DIM]) {
  using D = Dims<DIM>;
  if constexpr (DIM == 3) {
    *reinterpret_cast<double2 *>(q) = make_double2(v[0], v[4]);
    *reinterpret_cast<double2 *>(q + D::SX) = make_double2(v[1], v[5]);
    *reinterpret_cast<double2 *>(q + D::SY) = make_double2(v[2], v[6]);
    *reinterpret_cast<double2 *>(q + D::SX + D::SY) = make_double2(v[3], v[7]);
  } else {
    *reinterpret_cast<double2 *>(q) = make_double2(v[0], v[2]);
    *reinterpret_cast<double2 *>(q + D::SX) = make_double2(v[1], v[3]);
  }
}
template <int DIM> __device__ __forceinline__ void load_children(const double *q, double x[1 << DIM]) {
  using D = Dims<DIM>;
  if constexpr (DIM == 3) {
    double2 const a = *reinterpret_cast<const double2 *>(q);
    double2 const b = *reinterpret_cast<const double2 *>(q + D::SX);
    double2 const c = *reinterpret_cast<const double2 *>(q + D::SY);
    double2 const d = *reinterpret_cast<const double2 *>(q + D::SX + D::SY);
    x[0] = a.x, x[4] = a.y, x[1] = b.x, x[5] = b.y, x[2] = c.x, x[6] = c.y, x[3] = d.x, x[7] = d.y;
  } else {
    double2 const a = *reinterpret_cast<const double2 *>(q);
    double2 const b = *reinterpret_cast<const double2 *>(q + D::SX);
    x[0] = a.x, x[2] = a.y, x[1] = b.x, x[3] = b.y;
  }
}

// One workgroup per (link, equation): the child cells of the range,
// predicted from the parent's field.  FULL: the range is the whole block
// (12 predicted parent cells per axis, known at compile time).
// Predicts range r of child l.child's equation e from its parent: stages the
// parent window in s, then two children per axis per task.  The caller's
// early exits are uniform over the workgroup, so every thread reaches the
// barrier.
template <int DIM>
__device__ __forceinline__ void predict_range(double *s, double *__restrict__ blocks,
                                              const alpaca_mr_link &l, int e, const Range &r) {
  using D = Dims<DIM>;
  Window<DIM> const win(r, l.position);
  stage_window<DIM>(s, blocks + (size_t)l.parent * D::BLOCK + (size_t)e * D::FIELD, win);
  __syncthreads();
  double *child = blocks + (size_t)l.child * D::BLOCK + (size_t)e * D::FIELD;
  int const nz = DIM == 3 ? r.count[2] : 1;
  int const tasks = r.count[0] * r.count[1] * nz;
  for (int t = threadIdx.x; t < tasks; t += kStencilThreads) {
    int const a = t / (r.count[1] * nz), b = (t / nz) % r.count[1], c = t % nz;
    double v[1 << DIM];
    predict_cell<DIM>(s, Window<DIM>::centre(a, b, c), v);
    int const ci = r.start[0] + 2 * a, cj = r.start[1] + 2 * b;
    int const ck = DIM == 3 ? r.start[2] + 2 * c : 0;
    store_children<DIM>(child + ci * D::SX + cj * D::SY + ck, v);
  }
}

// Predicts range r (at most 8 predicted parent cells per axis) of every
// equation of child l.child from its parent, in one workgroup: while the
// children of equation e are predicted from one LDS window, the parent
// window of equation e + 1 is loaded into registers and then written to the
// other window, so the loads overlap the stencil arithmetic.  One barrier
// per equation; the caller's early exits are uniform over the workgroup.
template <int DIM> using SmallLds = LdsW<DIM, 12>;

template <int DIM>
__device__ __forceinline__ void predict_range_equations(double *s, double *__restrict__ blocks,
                                                        const alpaca_mr_link &l, const Range &r) {
  using D = Dims<DIM>;
  using L = SmallLds<DIM>;
  constexpr int K = 12;  // window cells per axis at most
  static_assert(K * K <= kStencilThreads, "one thread per window column");
  Window<DIM> const win(r, l.position);
  const double *parent = blocks + (size_t)l.parent * D::BLOCK;
  double *child = blocks + (size_t)l.child * D::BLOCK;
  // Thread (b, c) of the window's w1 x w2 plane moves the column a = 0..w0-1:
  // field offset g + a SX, LDS offset q + a L::SX.
  int const plane = win.w[1] * win.w[2];
  bool const mover = int(threadIdx.x) < plane;
  int const b = mover ? int(threadIdx.x) / win.w[2] : 0, c = mover ? int(threadIdx.x) % win.w[2] : 0;
  int const g = win.p0[0] * D::SX + (win.p0[1] + b) * D::SY + (win.p0[2] + c);
  int const q = b * L::SY + c;
  int const w0 = win.w[0];
  double pre[K];
  auto load = [&](int e) {
    if (!mover) return;
    const double *src = parent + (size_t)e * D::FIELD + g;
#pragma unroll
    for (int a = 0; a < K; ++a)
      if (a < w0) pre[a] = src[a * D::SX];
  };
  auto put = [&](double *w) {
    if (!mover) return;
#pragma unroll
    for (int a = 0; a < K; ++a)
      if (a < w0) w[q + a * L::SX] = pre[a];
  };
  load(0);
  put(s);
  __syncthreads();
  int const nz = DIM == 3 ? r.count[2] : 1;
  int const tasks = r.count[0] * r.count[1] * nz;
#pragma unroll 1
  for (int e = 0; e < D::NE; ++e) {
    if (e + 1 < D::NE) load(e + 1);
    const double *cur = s + (e & 1) * L::SIZE;
    double *out = child + (size_t)e * D::FIELD;
    for (int t = threadIdx.x; t < tasks; t += kStencilThreads) {
      int const a = t / (r.count[1] * nz), b = (t / nz) % r.count[1], c = t % nz;
      double v[1 << DIM];
      predict_cell<DIM, L>(cur, Window<DIM>::template centre<L>(a, b, c), v);
      int const ci = r.start[0] + 2 * a, cj = r.start[1] + 2 * b;
      int const ck = DIM == 3 ? r.start[2] + 2 * c : 0;
      store_children<DIM>(out + ci * D::SX + cj * D::SY + ck, v);
    }
    // The other window was last read by equation e - 1, before the previous
    // barrier.
    if (e + 1 < D::NE) put(s + ((e + 1) & 1) * L::SIZE);
    __syncthreads();
  }
}

// Ranges whose window fits the 12-cell layout: at most 8 predicted parent
// cells per axis.
__host__ __device__ inline bool small_window(const Range &r) {
  return r.count[0] <= 8 && r.count[1] <= 8 && r.count[2] <= 8;
}

// prediction_kernel<DIM, false> over ranges with small windows: one
// workgroup per link, every equation.
template <int DIM>
__global__ __launch_bounds__(kStencilThreads) void prediction_small_kernel(
    double *__restrict__ blocks, int nblocks, const alpaca_mr_link *__restrict__ links, Range r) {
  __shared__ double s[2 * SmallLds<DIM>::SIZE];
  alpaca_mr_link const l = links[blockIdx.x];
  if (!link_ok<DIM>(l, nblocks)) return;
  predict_range_equations<DIM>(s, blocks, l, r);
}

template <int DIM, bool FULL>
__global__ __launch_bounds__(kStencilThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) void prediction_kernel(
    double *__restrict__ blocks, int nblocks, const alpaca_mr_link *__restrict__ links, Range r) {
  using D = Dims<DIM>;
  __shared__ double s[Lds<DIM>::SIZE];
  if constexpr (FULL) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      r.start[a] = 0;
      r.count[a] = a < DIM ? kTC / 2 : 1;
    }
  }
  int const e = blockIdx.x % D::NE;
  alpaca_mr_link const l = links[blockIdx.x / D::NE];
  if (!link_ok<DIM>(l, nblocks)) return;
  predict_range<DIM>(s, blocks, l, e, r);
}

// The halo slab of a location (CommunicationTypes::GetStartIndicesHaloRecv and
// GetHaloSize, communication_types.h:76-132,190-197): per axis, direction +1
// gives the cells [IC+HS, TC), -1 gives [0, HS), 0 the internal cells
// [HS, HS+IC) (in 2D the single z cell).  False for a location outside the
// enum or, in 2D, one with a Top/Bottom part.
template <int DIM> __device__ __forceinline__ bool location_range(int loc, Range &r) {
  if (loc < 0 || loc >= 26) return false;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int const d = kLocationDir[loc][a];
    if (DIM == 2 && a == 2) {
      if (d != 0) return false;
      r.start[2] = 0;
      r.count[2] = 1;
    } else {
      r.start[a] = d > 0 ? kIC + kHS : (d < 0 ? 0 : kHS);
      r.count[a] = (d != 0 ? kHS : kIC) / 2;
    }
  }
  return true;
}

// UpdateMaterialJumpNoMpi (internal_halo_manager.cpp:231-262) for a batch:
// link i's child gets the halo slab of location locations[i] predicted from
// its parent (every equation), one workgroup per link.
template <int DIM>
__global__ __launch_bounds__(kStencilThreads) void jump_halo_kernel(
    double *__restrict__ blocks, int nblocks, const alpaca_mr_link *__restrict__ links,
    const int *__restrict__ locations) {
  __shared__ double s[2 * SmallLds<DIM>::SIZE];
  int const i = blockIdx.x;
  alpaca_mr_link const l = links[i];
  Range r;
  if (!link_ok<DIM>(l, nblocks) || !location_range<DIM>(locations[i], r)) return;
  predict_range_equations<DIM>(s, blocks, l, r);
}

// ------------------------------------------------------------------ details

struct Thresholds {
  double eps[kMaxLevels];  // coarsening threshold per level
  int nlevels;
};

// One workgroup per link: over Mass and Energy (the wavelet equations,
// field_setup.h:243-247) and every child cell, halo included, the relative
// detail between the child and the parent's prediction over the whole block,
// ChildNeedsRemeshing<N> (multiresolution.cpp:299-398):
//   NORM 0 (Linfinity): the largest |exact - predicted| / |exact|;
//   NORM 1 (Lone): per equation the sum of |exact - predicted| / |exact|
//                  times 1/(TC^2 TCZ), the largest over the equations;
//   NORM 2 (Ltwo): per equation sqrt(sum of ((exact - predicted)/exact)^2)
//                  times 1/(TC^2 TCZ), the largest over the equations.
// std::max(max_detail, x) drops a NaN x, so every running maximum is a
// non-NaN value >= +0 and the hardware maximum (which skips a NaN operand)
// picks the same value; for L-infinity the order of the reduction does not
// matter (the quotients of absolute values are never -0).  The L-one and
// L-two sums are reduced as a tree (lanes, then waves) instead of the
// reference's cell order: they agree to rounding, and a NaN or infinite term
// gives a NaN or infinite sum either way.
template <int DIM, int NORM>
__global__ __launch_bounds__(kStencilThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) void details_kernel(
    const double *__restrict__ blocks, int nblocks, const alpaca_mr_link *__restrict__ links,
    Thresholds th, double *__restrict__ details, int *__restrict__ decisions) {
  using D = Dims<DIM>;
  constexpr int HZ = DIM == 3 ? kTC / 2 : 1;
  constexpr int TASKS = (kTC / 2) * (kTC / 2) * HZ;
  constexpr int kWaves = kStencilThreads / 64;
  constexpr double kOneNumberOfCells = 1.0 / double(kTC * kTC * D::TZ);
  __shared__ double s[Lds<DIM>::SIZE];
  // The window's first entries take the per-wave partial results once it is
  // consumed (one LDS allocation of 32 KiB keeps five workgroups per CU).
  double *part = s;
  alpaca_mr_link const l = links[blockIdx.x];
  if (!link_ok<DIM>(l, nblocks)) {
    if (threadIdx.x == 0) {
      details[blockIdx.x] = __builtin_nan("");
      if (decisions) decisions[blockIdx.x] = -1;
    }
    return;
  }
  Range const full{{0, 0, 0}, {kTC / 2, kTC / 2, HZ}};
  Window<DIM> const win(full, l.position);
  double m = 0.0;
  for (int e = kMass; e <= kEnergy; ++e) {
    if (e != kMass) __syncthreads();  // the previous field's window is consumed
    stage_window<DIM>(s, blocks + (size_t)l.parent * D::BLOCK + (size_t)e * D::FIELD, win);
    __syncthreads();
    const double *child = blocks + (size_t)l.child * D::BLOCK + (size_t)e * D::FIELD;
    double acc = 0.0;
    for (int t = threadIdx.x; t < TASKS; t += kStencilThreads) {
      int const a = t / ((kTC / 2) * HZ), b = (t / HZ) % (kTC / 2), c = t % HZ;
      double v[1 << DIM];
      predict_cell<DIM>(s, Window<DIM>::centre(a, b, c), v);
      double x[1 << DIM];
      load_children<DIM>(child + 2 * a * D::SX + 2 * b * D::SY + (DIM == 3 ? 2 * c : 0), x);
#pragma unroll
      for (int n = 0; n < (1 << DIM); ++n) {
        if constexpr (NORM == 0) {
          m = __builtin_fmax(m, __builtin_fabs(x[n] - v[n]) / __builtin_fabs(x[n]));
        } else if constexpr (NORM == 1) {
          acc += __builtin_fabs(x[n] - v[n]) / __builtin_fabs(x[n]);
        } else {
          double const q = (x[n] - v[n]) / x[n];
          acc = acc + q * q;
        }
      }
    }
    if constexpr (NORM != 0) {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
      __syncthreads();  // every thread is done with the window
      if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
      __syncthreads();
      if (threadIdx.x == 0) {
        double sum = part[0];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) sum += part[w];
        double const val =
            NORM == 1 ? sum * kOneNumberOfCells : __builtin_sqrt(sum) * kOneNumberOfCells;
        m = __builtin_fmax(m, val);
      }
    }
  }
  if constexpr (NORM == 0) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = __builtin_fmax(m, __shfl_xor(m, off, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
      for (int w = 1; w < kWaves; ++w) m = __builtin_fmax(m, part[w]);
      m = __builtin_fmax(part[0], m);
    }
  }
  if (threadIdx.x == 0) {
    details[blockIdx.x] = m;
    if (decisions) {
      int dec = -1;
      if (l.level >= 0 && l.level < th.nlevels) {
        // RemeshingDecision (multiresolution.cpp:408-428): Coarse 2,
        // Refine 1, Neutral 0 (remesh_identifier.h:29).
        double const eps_coarsen = th.eps[l.level];
        double const eps_refine = eps_coarsen * 32.0;
        dec = m <= eps_coarsen ? 2 : (m >= eps_refine ? 1 : 0);
      }
      decisions[blockIdx.x] = dec;
    }
  }
}

int hip_status(hipError_t e) { return e == hipSuccess ? ALPACA_OK : ALPACA_ERR_HIP; }

bool bad_sizes(int dim, int nblocks, int nlinks) {
  return (dim != 2 && dim != 3) || nblocks < 1 || nblocks > kMaxBlocks || nlinks < 0 ||
         nlinks > kMaxBlocks;
}

template <int DIM> int average_impl(double *blocks, int nblocks, const alpaca_mr_link *links,
                                    int nlinks, hipStream_t s) {
  size_t const total = (size_t)nlinks * Dims<DIM>::NE * AvgShape<DIM>::PAIRS;
  dim3 grid((unsigned)((total + kThreads - 1) / kThreads));
  hipLaunchKernelGGL((average_kernel<DIM>), grid, dim3(kThreads), 0, s, blocks, nblocks, links, nlinks);
  return hip_status(hipGetLastError());
}

template <int DIM> int average_jump_impl(double *surfaces, int nblocks, const alpaca_mr_link *links,
                                         int nlinks, hipStream_t s) {
  size_t const total = (size_t)nlinks * DIM * Dims<DIM>::NE * (kIC / 2) * (DIM == 3 ? kIC / 2 : 1);
  dim3 grid((unsigned)((total + kThreads - 1) / kThreads));
  hipLaunchKernelGGL((average_jump_kernel<DIM>), grid, dim3(kThreads), 0, s, surfaces, nblocks, links,
                     nlinks);
  return hip_status(hipGetLastError());
}

template <int DIM> int prediction_impl(double *blocks, int nblocks, const alpaca_mr_link *links,
                                       int nlinks, const Range &r, hipStream_t s) {
  dim3 grid((unsigned)nlinks * Dims<DIM>::NE);
  bool const full = r.start[0] == 0 && r.start[1] == 0 && r.start[2] == 0 &&
                    r.count[0] == kTC / 2 && r.count[1] == kTC / 2 &&
                    r.count[2] == (DIM == 3 ? kTC / 2 : 1);
  if (!full && small_window(r))
    hipLaunchKernelGGL((prediction_small_kernel<DIM>), dim3((unsigned)nlinks),
                       dim3(kStencilThreads), 0, s, blocks, nblocks, links, r);
  else if (full)
    hipLaunchKernelGGL((prediction_kernel<DIM, true>), grid, dim3(kStencilThreads), 0, s, blocks,
                       nblocks, links, r);
  else
    hipLaunchKernelGGL((prediction_kernel<DIM, false>), grid, dim3(kStencilThreads), 0, s, blocks,
                       nblocks, links, r);
  return hip_status(hipGetLastError());
}

template <int DIM> int jump_halo_impl(double *blocks, int nblocks, const alpaca_mr_link *links,
                                      const int *locations, int nlinks, hipStream_t s) {
  dim3 grid((unsigned)nlinks);
  hipLaunchKernelGGL((jump_halo_kernel<DIM>), grid, dim3(kStencilThreads), 0, s, blocks, nblocks,
                     links, locations);
  return hip_status(hipGetLastError());
}

template <int DIM> int details_impl(const double *blocks, int nblocks, const alpaca_mr_link *links,
                                    int nlinks, int norm, const Thresholds &th, double *details,
                                    int *decisions, hipStream_t s) {
  dim3 const grid((unsigned)nlinks), block(kStencilThreads);
  if (norm == ALPACA_MR_NORM_L1)
    hipLaunchKernelGGL((details_kernel<DIM, 1>), grid, block, 0, s, blocks, nblocks, links, th,
                       details, decisions);
  else if (norm == ALPACA_MR_NORM_L2)
    hipLaunchKernelGGL((details_kernel<DIM, 2>), grid, block, 0, s, blocks, nblocks, links, th,
                       details, decisions);
  else
    hipLaunchKernelGGL((details_kernel<DIM, 0>), grid, block, 0, s, blocks, nblocks, links, th,
                       details, decisions);
  return hip_status(hipGetLastError());
}

}  // namespace
}  // namespace mr
}  // namespace alpaca

using namespace alpaca;

extern "C" {

int alpaca_mr_average(int dim, double *d_blocks, int nblocks, const alpaca_mr_link *d_links,
                      int nlinks, void *stream) {
  if (mr::bad_sizes(dim, nblocks, nlinks) || !d_blocks || (nlinks > 0 && !d_links))
    return ALPACA_ERR_INVALID;
  if (reinterpret_cast<uintptr_t>(d_blocks) % 16) return ALPACA_ERR_INVALID;  // 16-byte accesses
  if (nlinks == 0) return ALPACA_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dim == 2) return mr::average_impl<2>(d_blocks, nblocks, d_links, nlinks, s);
  return mr::average_impl<3>(d_blocks, nblocks, d_links, nlinks, s);
}

int alpaca_mr_average_jump_buffers(int dim, double *d_surfaces, int nblocks,
                                   const alpaca_mr_link *d_links, int nlinks, void *stream) {
  if (mr::bad_sizes(dim, nblocks, nlinks) || !d_surfaces ||
      (nlinks > 0 && !d_links))
    return ALPACA_ERR_INVALID;
  if (nlinks == 0) return ALPACA_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dim == 2) return mr::average_jump_impl<2>(d_surfaces, nblocks, d_links, nlinks, s);
  return mr::average_jump_impl<3>(d_surfaces, nblocks, d_links, nlinks, s);
}

int alpaca_mr_prediction(int dim, double *d_blocks, int nblocks, const alpaca_mr_link *d_links,
                         int nlinks, const int *range, void *stream) {
  if (mr::bad_sizes(dim, nblocks, nlinks) || !d_blocks || !range ||
      (nlinks > 0 && !d_links))
    return ALPACA_ERR_INVALID;
  if (reinterpret_cast<uintptr_t>(d_blocks) % 16) return ALPACA_ERR_INVALID;  // 16-byte accesses
  mr::Range r{{0, 0, 0}, {1, 1, 1}};
  for (int a = 0; a < dim; ++a) {
    int const start = range[2 * a], count = range[2 * a + 1];
    // Even start (the reference's own precondition, multiresolution.cpp:154-
    // 159) and inside the block: every parent cell read, start / 2 + 2 - 2 up
    // to start / 2 + count / 2 + TC / 2 - 2 + 1, then lies in [0, TC).
    if (start < 0 || count < 0 || (start & 1) || start + count > kTC) return ALPACA_ERR_INVALID;
    r.start[a] = start;
    r.count[a] = count / 2;
  }
  if (nlinks == 0 || r.count[0] == 0 || r.count[1] == 0 || r.count[2] == 0)
    return ALPACA_OK;  // nothing to predict
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dim == 2) return mr::prediction_impl<2>(d_blocks, nblocks, d_links, nlinks, r, s);
  return mr::prediction_impl<3>(d_blocks, nblocks, d_links, nlinks, r, s);
}

int alpaca_mr_jump_halo_update(int dim, double *d_blocks, int nblocks,
                               const alpaca_mr_link *d_links, const int *d_locations, int nlinks,
                               void *stream) {
  if (mr::bad_sizes(dim, nblocks, nlinks) || !d_blocks ||
      (nlinks > 0 && (!d_links || !d_locations)))
    return ALPACA_ERR_INVALID;
  if (reinterpret_cast<uintptr_t>(d_blocks) % 16) return ALPACA_ERR_INVALID;  // 16-byte accesses
  if (nlinks == 0) return ALPACA_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dim == 2) return mr::jump_halo_impl<2>(d_blocks, nblocks, d_links, d_locations, nlinks, s);
  return mr::jump_halo_impl<3>(d_blocks, nblocks, d_links, d_locations, nlinks, s);
}

int alpaca_mr_child_needs_remeshing(int dim, const double *d_blocks, int nblocks,
                                    const alpaca_mr_link *d_links, int nlinks, int norm,
                                    const double *eps_coarsen, int nlevels, double *d_details,
                                    int *d_decisions, void *stream) {
  if (mr::bad_sizes(dim, nblocks, nlinks) || !d_blocks ||
      (norm != ALPACA_MR_NORM_LINF && norm != ALPACA_MR_NORM_L1 && norm != ALPACA_MR_NORM_L2) ||
      (nlinks > 0 && (!d_links || !d_details)) || nlevels < 0 || nlevels > mr::kMaxLevels ||
      (nlevels > 0 && !eps_coarsen) || (d_decisions && nlevels == 0))
    return ALPACA_ERR_INVALID;
  if (reinterpret_cast<uintptr_t>(d_blocks) % 16) return ALPACA_ERR_INVALID;  // 16-byte accesses
  if (nlinks == 0) return ALPACA_OK;
  mr::Thresholds th;
  std::memset(&th, 0, sizeof(th));
  th.nlevels = nlevels;
  for (int i = 0; i < nlevels; ++i) th.eps[i] = eps_coarsen[i];
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dim == 2)
    return mr::details_impl<2>(d_blocks, nblocks, d_links, nlinks, norm, th, d_details,
                               d_decisions, s);
  return mr::details_impl<3>(d_blocks, nblocks, d_links, nlinks, norm, th, d_details,
                             d_decisions, s);
}

}  // extern "C"
