// Face numerics, included into alpaca::exact and alpaca::fast (see
// fv_device.h).  Real type R is double (FP64) or float (mixed form).  For
// R = double every expression keeps the reference's operation order, so the
// exact namespace reproduces it bit for bit.  Citations name the reference
// file:line each function follows.

template <typename R> __device__ __forceinline__ R smin(R a, R b) { return (b < a) ? b : a; }
template <typename R> __device__ __forceinline__ R smax(R a, R b) { return (a < b) ? b : a; }

// FP64 division a/b by Newton-Raphson on the hardware reciprocal: two
// iterations on r ~ 1/b (r += r(1 - b r)), then one residual correction of
// the quotient (q += r(a - b q)); each residual is exact in an FMA, so the
// result is within 1 ulp of a/b for normal operands.
__device__ __forceinline__ double nr_div64(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  double const q = a * r;
  double const rem = __builtin_fma(-b, q, a);
  return __builtin_fma(rem, r, q);
}

// FP64 division of the fast namespace: one Newton step on the hardware
// reciprocal, then the residual correction of the quotient.  v_rcp_f64 is
// within 2^-24 of 1/b (largest relative error 2^-24.4 over 16.7 M operands of
// exponents -200..200, including divisors just below a power of two), the
// step leaves at most about 2^-48, and the correction (residual exact in an
// FMA) squares that again, so the result is within 1 ulp of a/b for normal
// operands.  Random pairs come out correctly rounded (all 16.7 M of them);
// quotients within about 2^-90 of a rounding midpoint can round the other
// way.
__device__ __forceinline__ double fast_div64(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  double const e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  double const q = a * r;
  double const rem = __builtin_fma(-b, q, a);
  return __builtin_fma(rem, r, q);
}

// FP64 square root by the coupled Newton iteration for g ~ sqrt(x) and
// h ~ 1/(2 sqrt(x)) from the hardware reciprocal square root, with a final
// residual correction (x - g^2 exact in an FMA).  sqrt(0) = 0.
__device__ __forceinline__ double nr_sqrt64(double x) {
  double const r = __builtin_amdgcn_rsq(x);
  double g = x * r;
  double h = 0.5 * r;
  double const e = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, e, g);
  h = __builtin_fma(h, e, h);
  double const d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  return x == 0.0 ? 0.0 : g;
}

// Divisors whose reciprocal lies within about 2^-100 (relative) of a rounding
// midpoint: b = 2^e (2 - k 2^-52) for small odd k, whose significand has its
// top 20 bits set (a density just below a power of two after round-off is
// one; so are the Roe-averaged density and sqrt(rho_i) + sqrt(rho_n) of two
// such cells).  For them the two Newton steps of the hardware division sequence can
// leave the reciprocal one ulp from RN(1/b), and a quotient that also lies
// within about 2^-105 of a midpoint then rounds the wrong way (0x1.6666666666663p-2
// / 0x1.ffffffffffffbp-4 gives 2.8 instead of x86's 2.8000000000000003).
// Other divisors' reciprocals come out of the sequence correctly rounded, and
// with a correctly rounded reciprocal its final residual correction rounds the
// quotient correctly.
__device__ __forceinline__ bool recip_hard(double b) {
  return ((unsigned)((unsigned long long)__double_as_longlong(b) >> 32) & 0xFFFFFu) == 0xFFFFFu;
}

// The correctly rounded quotient from q, a quotient within one ulp of a/b:
// of q and its neighbour toward a/b, the one whose residual a - b q' is
// smaller (both residuals are exact in an FMA when |a| >= 2^-960; the distance
// to a/b is |residual| / |b| for both).  A tie cannot occur: no quotient of
// two doubles lies exactly on a midpoint.
__device__ __forceinline__ double div_round(double a, double b, double q) {
  if (!(__builtin_fabs(q) <= 1.7976931348623157e308) || q == 0.0 ||
      !(__builtin_fabs(a) >= 0x1p-960))
    return q;
  double const r = __builtin_fma(-b, q, a);
  if (r == 0.0) return q;
  bool const above = (r > 0.0) == (b > 0.0);  // a/b > q
  long long const qi = __double_as_longlong(q);
  double const q2 = __longlong_as_double(above == (q > 0.0) ? qi + 1 : qi - 1);
  double const r2 = __builtin_fma(-b, q2, a);
  return __builtin_fabs(r2) < __builtin_fabs(r) ? q2 : q;
}

// The WENO5 weight divisions c / x in the exact form: x = (s + 1e-6)^2 in
// [1e-12, 2^700) and c in {0.1, 0.6, 0.3} (and 1 / (a1 + a2 + a3) >= 2^-704),
// where the hardware division sequence scales nothing and fixes up nothing,
// so it is exactly its core nr_div64.  These 40 divisions per face take no
// hard-reciprocal check (it would cost about 3% of the exact stage kernel's time): their
// divisors are smoothness indicators, not state values, and one of them would
// have to lie within 2^-48 below a power of two, with the quotient within
// about 2^-105 of a midpoint, to round differently from x86.
__device__ __forceinline__ double weno_div(double c, double x) { return nr_div64(c, x); }

// Division and square root: the compiler's FP64 sequences (what a/b and
// sqrt(x) compile to) in the exact namespace; Newton-Raphson FP64 in the fast
// namespace; hardware reciprocal and square root (about 1 ulp) for the FP32
// part of the mixed form.
template <typename R> __device__ __forceinline__ R fdiv(R a, R b) {
  if constexpr (std::is_same<R, float>::value)
    return a * __builtin_amdgcn_rcpf(b);
  else if constexpr (kFastArith)
    return fast_div64(a, b);
  else {
    // The compiler's sequence, correctly rounded also for the divisors with a
    // hard reciprocal.
    double const q = a / b;
    return __builtin_expect(recip_hard(b), 0) ? div_round(a, b, q) : q;
  }
}
template <typename R> __device__ __forceinline__ R dsqrt(R x) {
  if constexpr (std::is_same<R, float>::value) {
    return __builtin_amdgcn_sqrtf(x);
  } else if constexpr (kFastArith) {
    return nr_sqrt64(x);
  } else {
    // The compiler's square root.  For finite x >= 2^-767 (high word in
    // [0x10000000, 0x7FF00000)) the hardware sequence scales nothing and its
    // zero/infinity select passes the iterate through, so it reduces to this
    // core (v_rsq, then Goldschmidt refinement of g ~ sqrt(x) and
    // h ~ 1/(2 sqrt(x)) with two residual corrections): the same bits with 8
    // fewer instructions.  Other arguments (0, tiny, negative, inf, NaN) take
    // the full sequence.
    unsigned const hw = (unsigned)((unsigned long long)__double_as_longlong(x) >> 32);
    if (__builtin_expect(hw - 0x10000000u < 0x6FF00000u, 1)) {
      double const r = __builtin_amdgcn_rsq(x);
      double g = x * r;
      double h = r * 0.5;
      double const e = __builtin_fma(-h, g, 0.5);
      g = __builtin_fma(g, e, g);
      double d = __builtin_fma(-g, g, x);
      h = __builtin_fma(h, e, h);
      g = __builtin_fma(d, h, g);
      d = __builtin_fma(-g, g, x);
      return __builtin_fma(d, h, g);
    }
    return __builtin_sqrt(x);
  }
}

// ConsistencyManagedSum, full symmetry (mathematical_functions.h:79-91).
// For finite a, b, c the three orderings are never NaN (two roundings cannot
// meet an opposite infinity) and are -0 all together or not at all (a rounded
// sum is -0 only when both addends are -0), so the hardware minimum/maximum
// (one instruction each) pick exactly what std::min/std::max pick.
template <typename R> __device__ __forceinline__ R cms3(R a, R b, R c) {
  R const s1 = (a + b) + c;
  R const s2 = (a + c) + b;
  R const s3 = (b + c) + a;
  R mn, mx;
  if constexpr (std::is_same<R, float>::value) {
    mn = __builtin_fminf(__builtin_fminf(s1, s2), s3);
    mx = __builtin_fmaxf(__builtin_fmaxf(s1, s2), s3);
  } else {
    mn = __builtin_fmin(__builtin_fmin(s1, s2), s3);
    mx = __builtin_fmax(__builtin_fmax(s1, s2), s3);
  }
  return R(0.5) * (mn + mx);
}

// DimensionAwareConsistencyManagedSum (mathematical_functions.h:100-108).
template <int DIM, typename R> __device__ __forceinline__ R dams(R a, R b, R c) {
  if constexpr (DIM == 2) {
    (void)c;
    return a + b;
  } else {
    return cms3(a, b, c);
  }
}

// CalculatePressure<false> (generic_stiffened_gas.h:40-58).  The recovery
// p = -gamma*B + (gamma-1)(E - |m|^2/(2 rho)) always runs in FP64: the
// difference cancels at high Mach number.
template <int DIM>
__device__ __forceinline__ double eos_pressure64(double gamma, double b, double rho, double mx,
                                                 double my, double mz, double e) {
  double const q = 0.5 * dams<DIM>(mx * mx, my * my, mz * mz);
  return -gamma * b + (gamma - 1.0) * (e - fdiv(q, rho));
}
template <int DIM, typename R>
__device__ __forceinline__ R eos_pressure(R gamma, R b, R rho, R mx, R my, R mz, R e) {
  return R(eos_pressure64<DIM>(double(gamma), double(b), double(rho), double(mx), double(my),
                               double(mz), double(e)));
}

// CalculateSpeedOfSound<false> (generic_stiffened_gas.h:84-97).
template <typename R> __device__ __forceinline__ R eos_sos(R gamma, R b, R rho, R p) {
  return dsqrt(fdiv(gamma * (p + b), rho));
}

// ComputePsi (stiffened_gas.cpp:127-130).
template <typename R> __device__ __forceinline__ R eos_psi(R gamma, R b, R p, R one_rho) {
  return (p + gamma * b) * one_rho;
}

template <typename R> struct Prim {
  R rho, p, vel[3];
};

// ConservativesToPrimeStates (prime_state_handler.h:279-370).
template <int DIM, typename R>
__device__ __forceinline__ void cons_to_prim(R gamma, R b, const R *c, Prim<R> &o) {
  o.rho = c[kMass];
  R const one_density = fdiv(R(1.0), c[kMass]);
  o.vel[0] = c[kMomX] * one_density;
  o.vel[1] = c[kMomX + 1] * one_density;
  if constexpr (DIM == 3)
    o.vel[2] = c[kMomX + 2] * one_density;
  else
    o.vel[2] = R(0.0);
  R const mz = DIM == 3 ? c[kMomX + DIM - 1] : R(0.0);
  o.p = eos_pressure<DIM>(gamma, b, c[kMass], c[kMomX], c[kMomX + 1], mz, c[kEnergy]);
}

// Per-cell quantities the Roe average needs, from the cell's FP64
// conservatives: 1/rho, prime velocities and pressure
// (prime_state_handler.h:279-370), psi (stiffened_gas.cpp:127-130) and the
// enthalpy with its EOS-recomputed pressure (stiffened_gas.cpp:80-87), as
// eigendecomposition.h:266-283 reads them.
template <typename R> struct CellQ {
  R rho, one_rho, vel[3], p, psi, H;
};

template <int DIM, typename R>
__device__ __forceinline__ void cell_quantities(double gamma, double b, const double *c,
                                                CellQ<R> &q) {
  Prim<double> p;
  cons_to_prim<DIM, double>(gamma, b, c, p);
  double const one_rho = fdiv(1.0, c[kMass]);
  double const mz = DIM == 3 ? c[kMomX + DIM - 1] : 0.0;
  double const H = fdiv(c[kEnergy] + eos_pressure64<DIM>(gamma, b, c[kMass], c[kMomX],
                                                         c[kMomX + 1], mz, c[kEnergy]),
                        c[kMass]);
  q.rho = R(c[kMass]);
  q.one_rho = R(one_rho);
  q.vel[0] = R(p.vel[0]);
  q.vel[1] = R(p.vel[1]);
  q.vel[2] = R(p.vel[2]);
  q.p = R(p.p);
  q.psi = R(eos_psi<double>(gamma, b, p.p, one_rho));
  q.H = R(H);
}

// WENO5::ApplyImplementation (weno5.h:88-158) on ordered v1..v5.  ONE_DIV
// (FP64 with FMA only) forms the weights with one division, below.
template <typename R, bool ONE_DIV = false>
__device__ __forceinline__ R weno5(R v1, R v2, R v3, R v4, R v5) {
  R const c_s1 = R(13.0 / 12.0), c_s2 = R(0.25), eps = R(1.0e-6);
  R const s11 = R(1.0) * v1 + R(-2.0) * v2 + R(1.0) * v3;
  R const s12 = R(1.0) * v1 + R(-4.0) * v2 + R(3.0) * v3;
  R const s1 = c_s1 * s11 * s11 + c_s2 * s12 * s12;
  R const s21 = R(1.0) * v2 + R(-2.0) * v3 + R(1.0) * v4;
  R const s22 = R(1.0) * v2 + R(-1.0) * v4;
  R const s2 = c_s1 * s21 * s21 + c_s2 * s22 * s22;
  R const s31 = R(1.0) * v3 + R(-2.0) * v4 + R(1.0) * v5;
  R const s32 = R(3.0) * v3 + R(-4.0) * v4 + R(1.0) * v5;
  R const s3 = c_s1 * s31 * s31 + c_s2 * s32 * s32;
  R const x1 = (s1 + eps) * (s1 + eps);
  R const x2 = (s2 + eps) * (s2 + eps);
  R const x3 = (s3 + eps) * (s3 + eps);
  R const p1 = R(2.0 / 6.0) * v1 + R(-7.0 / 6.0) * v2 + R(11.0 / 6.0) * v3;
  R const p2 = R(-1.0 / 6.0) * v2 + R(5.0 / 6.0) * v3 + R(2.0 / 6.0) * v4;
  R const p3 = R(2.0 / 6.0) * v3 + R(5.0 / 6.0) * v4 + R(-1.0 / 6.0) * v5;
  R a1, a2, a3, one_a_sum;
  if constexpr (ONE_DIV) {
    static_assert(kFastArith && std::is_same<R, double>::value, "FP64 with FMA only");
    // FP64 with FMA: the weights d_k / x_k over the common denominator
    // x1 x2 x3, so that w_k = b_k / (b1 + b2 + b3) with b_k = d_k x_i x_j and
    // one division replaces four.  Each x_k (>= 1e-12) is first clamped to at
    // most 2^1000 (an infinite one then gets a weight about 1e-289 of the
    // others', as its d_k / x_k = 0 would; a NaN one still makes the result
    // NaN through its stencil polynomial), and all three are scaled by 2^-e,
    // e the exponent of the largest, into (2^-1041, 1), so the b_k stay at
    // most 0.6 and their sum positive.  Scaling by a power of two changes no
    // significand, so for x_k in the normal range the quotient has the bits
    // of the unscaled one.
    double const c1 = __builtin_fmin(x1, 0x1p1000), c2 = __builtin_fmin(x2, 0x1p1000),
                 c3 = __builtin_fmin(x3, 0x1p1000);
    int const e = __builtin_amdgcn_frexp_exp(__builtin_fmax(__builtin_fmax(c1, c2), c3));
    double const y1 = __builtin_amdgcn_ldexp(c1, -e), y2 = __builtin_amdgcn_ldexp(c2, -e),
                 y3 = __builtin_amdgcn_ldexp(c3, -e);
    double const b1 = 0.1 * y2 * y3, b2 = 0.6 * y1 * y3, b3 = 0.3 * y1 * y2;
    return fast_div64(b1 * p1 + b2 * p2 + b3 * p3, b1 + b2 + b3);
  }
  if constexpr (!kFastArith && std::is_same<R, double>::value) {
    // x_k = (s_k + 1e-6)^2 >= 1e-12 (s_k >= 0 or NaN).  While every x_k is
    // below 2^700 the hardware's correctly rounded division of 0.1, 0.6, 0.3
    // by x_k and of 1 by a1 + a2 + a3 (>= 2^-704) scales nothing and fixes up
    // nothing, so it is exactly its Newton-Raphson core nr_div64 (v_rcp, two
    // refinements, quotient, one residual correction): same bits, 3 fewer
    // instructions per division.  The positive x_k order like the high words
    // of their bit patterns; NaN and infinity fail the test and take the
    // full division.
    auto hi = [](double v) { return (unsigned)((unsigned long long)__double_as_longlong(v) >> 32); };
    unsigned const hmax = __builtin_elementwise_max(__builtin_elementwise_max(hi(x1), hi(x2)), hi(x3));
    if (__builtin_expect(hmax < 0x6BB00000u, 1)) {  // 2^700
      a1 = weno_div(0.1, x1);
      a2 = weno_div(0.6, x2);
      a3 = weno_div(0.3, x3);
      one_a_sum = weno_div(1.0, a1 + a2 + a3);
    } else {
      a1 = fdiv(0.1, x1);
      a2 = fdiv(0.6, x2);
      a3 = fdiv(0.3, x3);
      one_a_sum = fdiv(1.0, a1 + a2 + a3);
    }
  } else {
    a1 = fdiv(R(0.1), x1);
    a2 = fdiv(R(0.6), x2);
    a3 = fdiv(R(0.3), x3);
    one_a_sum = fdiv(R(1.0), a1 + a2 + a3);
  }
  R const w1 = a1 * one_a_sum;
  R const w2 = a2 * one_a_sum;
  R const w3 = a3 * one_a_sum;
  return w1 * p1 + w2 * p2 + w3 * p3;
}
// UpwindLeft {0,+1} and UpwindRight {1,-1} (apply_wrapper.h:48-58).
template <typename R, bool ONE_DIV = false> __device__ __forceinline__ R weno_left(const R *a) {
  return weno5<R, ONE_DIV>(a[0], a[1], a[2], a[3], a[4]);
}
template <typename R, bool ONE_DIV = false> __device__ __forceinline__ R weno_right(const R *a) {
  return weno5<R, ONE_DIV>(a[5], a[4], a[3], a[2], a[1]);
}
// The FP64-with-FMA form reconstructs with one division per WENO5.
template <typename R> constexpr bool kWenoOneDiv = kFastArith && std::is_same<R, double>::value;

template <int NE, typename R> struct Eig {
  R L[NE][NE];
  R Rm[NE][NE];
  R lam[NE];
  // Roe averages the right eigenvectors are built from when they are built
  // late (kDeferRight below).
  R u, v, w, c, H, g, rho_roe;
  bool ok;
};

// Structural zeros of the Roe eigenvectors (eigendecomposition.h:238-466):
// L rows 1 and 2 (the shear waves) hold only the mass and their own
// transverse momentum; R leaves out the shear columns of the mass, normal
// momentum and other transverse momentum rows.  The FP64-with-FMA form skips
// these products (0 * x changes a sum only for a non-finite x, which that
// form does not track); the exact form forms them, as the reference does,
// and the mixed form keeps its measured accuracy by keeping them too.
template <int DIM, int DIR> __device__ __forceinline__ constexpr bool left_zero(int n, int c) {
  constexpr int Mp = kMomX + DIR, M1 = kMomX + Minor<DIR>::m1, M2 = kMomX + Minor<DIR>::m2;
  if (n == 1) return c != kMass && c != M1;
  if (DIM == 3 && n == 2) return c != kMass && c != M2;
  (void)Mp;
  return false;
}
template <int DIM, int DIR> __device__ __forceinline__ constexpr bool right_zero(int r, int c) {
  constexpr int Mp = kMomX + DIR, M1 = kMomX + Minor<DIR>::m1, M2 = kMomX + Minor<DIR>::m2;
  bool const shear = c >= 1 && c < DIM;  // columns 1 .. DIM-1
  if (!shear) return false;
  if (r == kMass || r == Mp) return true;
  if (r == M1) return c != 1;
  if (DIM == 3 && r == M2) return c != 2;
  return false;
}
template <typename R> constexpr bool kSkipZeros = kFastArith && std::is_same<R, double>::value;

// The exact FP64 finite-volume (HLLC) kernels build the right eigenvectors
// only once the characteristic values are formed, so the 19 entries are not
// held in registers meanwhile (it saves 20 VGPRs in 3D); without contraction
// the position changes no bit.  (Flux splitting builds them with the left
// ones: deferring them there made the exact kernel slower.)  The contracting kernels build them beside
// the left eigenvectors: there the basic block they sit in decides which
// products fuse, and the mixed form's accuracy was measured that way.
template <typename R>
constexpr bool kDeferRight = std::is_same<R, double>::value && !kFastArith;

template <int DIM, int DIR, typename R>
__device__ __forceinline__ void set_right_eigenvectors(Eig<DIM + 2, R> &ev, R u, R v, R w, R c,
                                                       R H, R g, R rho_roe) {
  constexpr int NE = DIM + 2;
  constexpr int m1 = Minor<DIR>::m1, m2 = Minor<DIR>::m2;
  constexpr int evp = DIM;
  constexpr int Mp = kMomX + DIR, M1 = kMomX + m1, M2 = kMomX + m2;
  (void)m2;
  (void)M2;
  (void)w;
  ev.Rm[kMass][0] = R(1.0);
  ev.Rm[kMass][evp] = g;
  ev.Rm[kMass][NE - 1] = R(1.0);

  ev.Rm[Mp][0] = u - c;
  ev.Rm[Mp][evp] = g * u;
  ev.Rm[Mp][NE - 1] = u + c;

  ev.Rm[M1][0] = v;
  ev.Rm[M1][1] = -rho_roe;
  ev.Rm[M1][evp] = g * v;
  ev.Rm[M1][NE - 1] = v;

  if constexpr (DIM == 3) {
    ev.Rm[M2][0] = w;
    ev.Rm[M2][2] = rho_roe;
    ev.Rm[M2][evp] = g * w;
    ev.Rm[M2][NE - 1] = w;
  }

  ev.Rm[kEnergy][0] = H - u * c;
  ev.Rm[kEnergy][1] = -v * rho_roe;
  if constexpr (DIM == 3) ev.Rm[kEnergy][2] = rho_roe * w;
  ev.Rm[kEnergy][evp] = g * H - c * c;
  ev.Rm[kEnergy][NE - 1] = H + u * c;
}

// The right eigenvectors of a roe_eigensystem call that deferred them.
template <int DIM, int DIR, typename R, bool kDefer = kDeferRight<R>>
__device__ __forceinline__ void roe_right_eigenvectors(Eig<DIM + 2, R> &ev) {
  if constexpr (kDefer) {
    if (ev.ok) set_right_eigenvectors<DIM, DIR>(ev, ev.u, ev.v, ev.w, ev.c, ev.H, ev.g, ev.rho_roe);
  }
}

// ComputeRoeEigendecomposition for one face (eigendecomposition.h:238-466).
// t = cell i, n = cell i+1 along DIR.  Returns false (all zero) when a
// density is not positive.
template <int DIM, int DIR, typename R, bool kDefer = kDeferRight<R>>
__device__ __forceinline__ bool roe_eigensystem(R gamma, const CellQ<R> &t, const CellQ<R> &n,
                                                Eig<DIM + 2, R> &ev) {
  constexpr int NE = DIM + 2;
  constexpr int m1 = Minor<DIR>::m1, m2 = Minor<DIR>::m2;
  constexpr int evp = DIM;
  constexpr int Mp = kMomX + DIR, M1 = kMomX + m1, M2 = kMomX + m2;
#pragma unroll
  for (int r = 0; r < NE; ++r) {
#pragma unroll
    for (int c = 0; c < NE; ++c) {
      ev.L[r][c] = R(0.0);
      ev.Rm[r][c] = R(0.0);
    }
    ev.lam[r] = R(0.0);
  }
  ev.ok = false;
  if (t.rho <= R(0.0) || n.rho <= R(0.0)) return false;
  ev.ok = true;

  R const u_t = t.vel[DIR];
  R const v_t = t.vel[m1];
  R const w_t = DIM == 3 ? t.vel[m2] : R(0.0);
  R const u_n = n.vel[DIR];
  R const v_n = n.vel[m1];
  R const w_n = DIM == 3 ? n.vel[m2] : R(0.0);

  R const sq_t = dsqrt(t.rho);
  R const sq_n = dsqrt(n.rho);
  R const rho_div = fdiv(R(1.0), sq_t + sq_n);
  R const rho_roe = sq_t * sq_n;
  R const one_rho_roe = fdiv(R(1.0), rho_roe);

  R const u = ((u_t * sq_t) + (u_n * sq_n)) * rho_div;
  R const v = ((v_t * sq_t) + (v_n * sq_n)) * rho_div;
  R const w = ((w_t * sq_t) + (w_n * sq_n)) * rho_div;
  R const psi = ((t.psi * sq_t) + (n.psi * sq_n)) * rho_div;
  R const g = gamma - R(1.0);
  R const H = (t.H * sq_t + n.H * sq_n) * rho_div;

  R const q2 = dams<DIM>(u * u, v * v, w * w);
  R const tmp = dams<DIM>((u_t - u_n) * (u_t - u_n), (v_t - v_n) * (v_t - v_n),
                          (w_t - w_n) * (w_t - w_n));
  R const pod = ((t.p * t.one_rho) * sq_t + (n.p * n.one_rho) * sq_n) * rho_div +
                R(0.5) * rho_roe * (rho_div * rho_div) * tmp;
  R const cc = psi + g * pod;
  R const one_cc = fdiv(R(1.0), cc);
  R const c = dsqrt(cc);

  ev.L[0][kMass] = R(0.5) * one_cc * (g * q2 - g * H + (u + c) * c);
  ev.L[0][Mp] = R(0.5) * one_cc * (-u * g - c);
  ev.L[0][M1] = R(0.5) * one_cc * (-v * g);
  if constexpr (DIM == 3) ev.L[0][M2] = R(0.5) * one_cc * (-w * g);
  ev.L[0][kEnergy] = R(0.5) * one_cc * g;

  ev.L[1][kMass] = v * one_rho_roe;
  ev.L[1][M1] = -one_rho_roe;

  if constexpr (DIM == 3) {
    ev.L[2][kMass] = -w * one_rho_roe;
    ev.L[2][M2] = one_rho_roe;
  }

  ev.L[evp][kMass] = one_cc * (H - q2);
  ev.L[evp][Mp] = one_cc * u;
  ev.L[evp][M1] = one_cc * v;
  if constexpr (DIM == 3) ev.L[evp][M2] = one_cc * w;
  ev.L[evp][kEnergy] = -one_cc;

  ev.L[NE - 1][kMass] = R(0.5) * one_cc * (g * q2 - g * H - (u - c) * c);
  ev.L[NE - 1][Mp] = R(0.5) * one_cc * (-u * g + c);
  ev.L[NE - 1][M1] = R(0.5) * one_cc * (-v * g);
  if constexpr (DIM == 3) ev.L[NE - 1][M2] = R(0.5) * one_cc * (-w * g);
  ev.L[NE - 1][kEnergy] = R(0.5) * one_cc * g;

  if constexpr (kDefer) {
    ev.u = u;
    ev.v = v;
    ev.w = w;
    ev.c = c;
    ev.H = H;
    ev.g = g;
    ev.rho_roe = rho_roe;
  } else {
    set_right_eigenvectors<DIM, DIR>(ev, u, v, w, c, H, g, rho_roe);
  }

  ev.lam[0] = __builtin_fabs(u - c);
  ev.lam[NE - 1] = __builtin_fabs(u + c);
#pragma unroll
  for (int l = 1; l < NE - 1; ++l) ev.lam[l] = __builtin_fabs(u);
  return true;
}

// TransformToPhysicalSpace (eigendecomposition.h:144-165).  Every product is
// formed, the zero entries' too: 0 * x is NaN for a non-finite x, as in the
// reference.
template <int DIM, int DIR, typename R>
__device__ __forceinline__ void to_physical(const R *ch, const R (&Rm)[DIM + 2][DIM + 2], R *out) {
  constexpr int NE = DIM + 2;
#pragma unroll
  for (int l = 0; l < NE; ++l) {
    R s = R(0.0);
#pragma unroll
    for (int n = 1; n < NE - 1; ++n)
      if (!(kSkipZeros<R> && right_zero<DIM, DIR>(l, n))) s += ch[n] * Rm[l][n];
    s += (ch[0] * Rm[l][0] + ch[NE - 1] * Rm[l][NE - 1]);
    out[l] = s;
  }
}

// CalculateSignalSpeed, Einfeldt (hll_signal_speed_calculator.h:77-101).
template <typename R>
__device__ __forceinline__ void einfeldt(R rl, R rr, R ul, R ur, R cl, R cr, R &sl, R &sr) {
  R const sql = dsqrt(rl);
  R const sqr = dsqrt(rr);
  R const rho_div = fdiv(R(1.0), sql + sqr);
  R const u_roe = ((ul * sql) + (ur * sqr)) * rho_div;
  R const c_e = dsqrt(((cl * cl * sql) + (cr * cr * sqr)) * rho_div +
                      (R(0.5) * sql * sqr * rho_div * rho_div) * (ur - ul) * (ur - ul));
  sl = smin(u_roe - c_e, ul - cl);
  sr = smax(u_roe + c_e, ur + cr);
}

// HllcRiemannSolver::SolveRiemannProblemImplementation (hllc_riemann_solver.h:
// 52-198), including the clamped wave speeds in the energy star states.
template <int DIM, int DIR, typename R>
__device__ __forceinline__ void hllc(R gamma, R b, const R *cl, const R *cr, const Prim<R> &pl,
                                     const Prim<R> &pr, R *f) {
  constexpr int NE = DIM + 2;
  constexpr int Mp = kMomX + DIR;
  R qsl[NE], qsr[NE], fl[NE], fr[NE];
  R const p_l = pl.p, p_r = pr.p;
  R const one_rho_l = fdiv(R(1.0), cl[kMass]);
  R const one_rho_r = fdiv(R(1.0), cr[kMass]);
  R const u_l = pl.vel[DIR], u_r = pr.vel[DIR];
  R const c_l = eos_sos(gamma, b, cl[kMass], p_l);
  R const c_r = eos_sos(gamma, b, cr[kMass], p_r);
  R sls, srs;
  einfeldt(cl[kMass], cr[kMass], u_l, u_r, c_l, c_r, sls, srs);
  R const sc = fdiv((p_r - p_l) + (cl[Mp] * (sls - u_l) - cr[Mp] * (srs - u_r)),
                    cl[kMass] * (sls - u_l) - cr[kMass] * (srs - u_r));
  R const sl = smin(sls, R(0.0));
  R const sr = smax(srs, R(0.0));
  R const chi_l = fdiv(sls - u_l, sls - sc);
  R const chi_r = fdiv(srs - u_r, srs - sc);

  qsl[kMass] = cl[kMass] * chi_l;
  qsl[Mp] = cl[kMass] * chi_l * sc;
  qsl[kEnergy] = cl[kMass] * chi_l *
                 (cl[kEnergy] * one_rho_l + (sc - u_l) * (sc + fdiv(p_l, cl[kMass] * (sl - u_l))));
  qsr[kMass] = cr[kMass] * chi_r;
  qsr[Mp] = cr[kMass] * chi_r * sc;
  qsr[kEnergy] = cr[kMass] * chi_r *
                 (cr[kEnergy] * one_rho_r + (sc - u_r) * (sc + fdiv(p_r, cr[kMass] * (sr - u_r))));

  fl[kMass] = cl[Mp];
  fl[Mp] = ((cl[Mp] * cl[Mp]) * one_rho_l) + p_l;
  fl[kEnergy] = u_l * (cl[kEnergy] + p_l);
  fr[kMass] = cr[Mp];
  fr[Mp] = ((cr[Mp] * cr[Mp]) * one_rho_r) + p_r;
  fr[kEnergy] = u_r * (cr[kEnergy] + p_r);

#pragma unroll
  for (int d = 0; d < DIM - 1; ++d) {
    int const Mm = kMomX + (d == 0 ? Minor<DIR>::m1 : Minor<DIR>::m2);
    qsl[Mm] = chi_l * cl[Mm];
    qsr[Mm] = chi_r * cr[Mm];
    fl[Mm] = u_l * cl[Mm];
    fr[Mm] = u_r * cr[Mm];
  }
  int const sg = (R(0.0) < sc) - (sc < R(0.0));
  R const wl = R(0.5) * (R(1.0) + R(sg));
  R const wr = R(0.5) * (R(1.0) - R(sg));
#pragma unroll
  for (int n = 0; n < NE; ++n) {
    R const fsl = fl[n] + (sl * (qsl[n] - cl[n]));
    R const fsr = fr[n] + (sr * (qsr[n] - cr[n]));
    f[n] = (wl * fsl + wr * fsr);
  }
}

// ComputeAdvection (flux_splitting_scheme.cpp:42-76) of one cell in FP64.
template <int DIM, int DIR>
__device__ __forceinline__ void advection(const double *c, const Prim<double> &pd, double *a) {
  double const ud = pd.vel[DIR];
  a[kMass] = pd.rho * ud;
  a[kEnergy] = (c[kEnergy] + pd.p) * ud;
#pragma unroll
  for (int d = 0; d < DIM; ++d) a[kMomX + d] = pd.rho * (ud * pd.vel[d]);
  a[kMomX + DIR] += pd.p;
}

// FluxSplittingScheme::ComputeFluxes + ComputeAdvection for one face in the
// exact FP64 form (flux_splitting_scheme.cpp:42-76, 185-315), added into f.
// Every value is formed by the reference's operations in the reference's
// order; only the time at which each is formed differs, which without
// contraction changes no bit.  To keep two waves per SIMD the face is
// processed in a fixed sequence with few values live: the eigensystem of
// cells i and i+1, then the advection of the six stencil cells one cell at a
// time (kept in registers), then field by field the characteristic
// projections and WENO5, with the conservatives read again from the block for
// each field (the L1 and L2 serve the re-reads).  The reads go through a
// buffer resource (no address arithmetic).  Each group of reads (and each
// field's row of L) takes its inputs through an empty asm statement that
// also consumes the previous group's last result, so the compiler neither
// merges the re-reads nor hoists one field's work above the previous one's:
// left to itself it forms the products of all fields at once and needs
// about 340 registers (one wave per SIMD).
constexpr int kRoePrefetch = 2;  // stencil cells read ahead for the next field

template <int DIM, int DIR>
__device__ __forceinline__ void roe_split_flux_exact(double gamma, double b,
                                                     const double *__restrict__ blk, int off,
                                                     double *f) {
  constexpr int NE = DIM + 2;
  constexpr int S = axis_stride<DIM, DIR>();
  BlockReader<DIM> const rd(blk);
  // Byte offset of stencil cell 0 (cell i-2, inside the haloed block).
  int const off0 = (off - 2 * S) * (int)sizeof(double);
  auto at = [&rd](int o, int m, int e) {
    return rd(o, (e * Dims<DIM>::FIELD + m * S) * (int)sizeof(double));
  };
  Eig<NE, double> ev;
  {
    double Ui[NE], Un[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      Ui[e] = at(off0, 2, e);
      Un[e] = at(off0, 3, e);
    }
    CellQ<double> qt, qn;
    cell_quantities<DIM, double>(gamma, b, Ui, qt);
    cell_quantities<DIM, double>(gamma, b, Un, qn);
    roe_eigensystem<DIM, DIR, double, true>(gamma, qt, qn, ev);
  }
  double dep = ev.lam[0];
  double adv[6][NE];
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    int o = off0;
    asm volatile("" : "+v"(o) : "v"(dep));
    double U[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) U[e] = at(o, m, e);
    Prim<double> P;
    cons_to_prim<DIM, double>(gamma, b, U, P);
    advection<DIM, DIR>(U, P, adv[m]);
    dep = adv[m][NE - 1];
  }
  // Conservatives of the stencil for the field being projected; the first
  // kRoePrefetch cells of the next field are read while this field's WENO5
  // runs, the rest after it.
  double U[6][NE];
  {
    int o = off0;
    asm volatile("" : "+v"(o) : "v"(dep));
#pragma unroll
    for (int m = 0; m < 6; ++m)
#pragma unroll
      for (int e = 0; e < NE; ++e) U[m][e] = at(o, m, e);
  }
  double chf[NE];
#pragma unroll
  for (int n = 0; n < NE; ++n) {
    // The field's row of L and its eigenvalue pass through an asm statement
    // that consumes the previous field's result, so no product of this field
    // is formed before the previous field is done.
    double lam = ev.lam[n], Lr[5];
#pragma unroll
    for (int l = 0; l < NE; ++l) Lr[l] = ev.L[n][l];
    if constexpr (NE == 5)
      asm volatile("" : "+v"(lam), "+v"(Lr[0]), "+v"(Lr[1]), "+v"(Lr[2]), "+v"(Lr[3]),
                   "+v"(Lr[4]) : "v"(dep));
    else
      asm volatile("" : "+v"(lam), "+v"(Lr[0]), "+v"(Lr[1]), "+v"(Lr[2]), "+v"(Lr[3])
                   : "v"(dep));
    double pos[6], neg[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      double sp = 0.0, sn = 0.0;
#pragma unroll
      for (int q = 0; q < NE; ++q) {
        int const l = seq<DIM, DIR>(q);
        double const uct = U[m][l] * Lr[l];
        double const act = adv[m][l] * Lr[l];
        sp += (act + lam * uct);
        sn += (act - lam * uct);
      }
      pos[m] = sp;
      neg[m] = sn;
    }
    if (n + 1 < NE) {
      int o = off0;
      asm volatile("" : "+v"(o) : "v"(pos[5]), "v"(neg[5]));
#pragma unroll
      for (int m = 0; m < kRoePrefetch; ++m)
#pragma unroll
        for (int e = 0; e < NE; ++e) U[m][e] = at(o, m, e);
    }
    chf[n] = 0.5 * (weno_left(pos) + weno_right(neg));
    dep = chf[n];
    if (n + 1 < NE) {
      int o = off0;
      asm volatile("" : "+v"(o) : "v"(dep));
#pragma unroll
      for (int m = kRoePrefetch; m < 6; ++m)
#pragma unroll
        for (int e = 0; e < NE; ++e) U[m][e] = at(o, m, e);
    }
  }
  roe_right_eigenvectors<DIM, DIR, double, true>(ev);
  double phys[NE];
  to_physical<DIM, DIR>(chf, ev.Rm, phys);
#pragma unroll
  for (int e = 0; e < NE; ++e) f[e] += phys[e];
}

// FluxSplittingScheme::ComputeFluxes + ComputeAdvection for one face
// (flux_splitting_scheme.cpp:42-76, 185-315), added into f, in the
// contracting forms (FP64 with FMA, mixed).  The stencil is walked cell by
// cell -- load, advection, projection onto every field -- which keeps few
// values live: 2 waves per SIMD for the FP64-with-FMA kernel, 3 for the mixed
// one.  Each projection is the expression of the exact form.
template <int DIM, int DIR, typename R>
__device__ __forceinline__ void roe_split_flux(double gamma64, double b64,
                                               const double *__restrict__ base, R *f) {
  constexpr int NE = DIM + 2;
  constexpr bool kShift = std::is_same<R, float>::value;
  R const gamma = R(gamma64);
  auto load = [base](int m, double *c) {
#pragma unroll
    for (int e = 0; e < NE; ++e)
      c[e] = base[e * Dims<DIM>::FIELD + (m - 2) * axis_stride<DIM, DIR>()];
  };
  Eig<NE, R> ev;
  CellQ<R> qt, qn;
  double adv2[NE];  // advection of cell i (FP64; mixed form only)
  R chf[NE];
  {
    double Ui[NE], Un[NE];
    load(2, Ui);
    load(3, Un);
    cell_quantities<DIM, R>(gamma64, b64, Ui, qt);
    cell_quantities<DIM, R>(gamma64, b64, Un, qn);
    roe_eigensystem<DIM, DIR, R, false>(gamma, qt, qn, ev);
    if constexpr (kShift) {
      Prim<double> pd;
      cons_to_prim<DIM, double>(gamma64, b64, Ui, pd);
      advection<DIM, DIR>(Ui, pd, adv2);
    }
    R pos[NE][6], neg[NE][6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      double Um[NE];
      load(m, Um);
      R Uc[NE];
#pragma unroll
      for (int e = 0; e < NE; ++e) Uc[e] = kShift ? R(Um[e] - Ui[e]) : R(Um[e]);
      R adv[NE];
      Prim<double> pd;
      cons_to_prim<DIM, double>(gamma64, b64, Um, pd);
      if constexpr (kShift) {
        // ComputeAdvection in FP64, then the difference from cell i into
        // FP32 (see face_flux).
        double a64[NE];
        advection<DIM, DIR>(Um, pd, a64);
#pragma unroll
        for (int e = 0; e < NE; ++e) adv[e] = R(a64[e] - adv2[e]);
      } else {
        R const rho = R(pd.rho), p = R(pd.p), ud = R(pd.vel[DIR]);
        adv[kMass] = rho * ud;
        adv[kEnergy] = (Uc[kEnergy] + p) * ud;
#pragma unroll
        for (int d = 0; d < DIM; ++d) adv[kMomX + d] = rho * (ud * R(pd.vel[d]));
        adv[kMomX + DIR] += p;
      }
#pragma unroll
      for (int n = 0; n < NE; ++n) {
        R sp = R(0.0), sn = R(0.0);
#pragma unroll
        for (int q = 0; q < NE; ++q) {
          int const l = seq<DIM, DIR>(q);
          if (kSkipZeros<R> && left_zero<DIM, DIR>(n, l)) continue;
          R const uct = Uc[l] * ev.L[n][l];
          R const act = adv[l] * ev.L[n][l];
          sp += (act + ev.lam[n] * uct);
          sn += (act - ev.lam[n] * uct);
        }
        pos[n][m] = sp;
        neg[n][m] = sn;
      }
    }
#pragma unroll
    for (int n = 0; n < NE; ++n)
      chf[n] = R(0.5) * (weno_left<R, kWenoOneDiv<R>>(pos[n]) +
                         weno_right<R, kWenoOneDiv<R>>(neg[n]));
  }
  R phys[NE];
  to_physical<DIM, DIR>(chf, ev.Rm, phys);
  if constexpr (kShift) {
    // R L (adv of cell i) = adv of cell i: the half sums of the shifts.
#pragma unroll
    for (int e = 0; e < NE; ++e) f[e] += R(double(phys[e]) + adv2[e]);
  } else {
#pragma unroll
    for (int e = 0; e < NE; ++e) f[e] += phys[e];
  }
}

// The finite-volume face (FiniteVolumeScheme::ComputeFluxes,
// finite_volume_scheme.cpp:160-251, with CharacteristicStateReconstruction,
// characteristic_state_reconstruction.h:60-134) from its stencil Ud (FP64
// conservatives of cells i-2..i+3 along DIR).  Writes f (zero when the
// reference skips the face).
template <int DIM, int DIR, typename R>
__device__ __forceinline__ void fv_face(double gamma64, double b64, const double (&Ud)[6][DIM + 2],
                                        R *f) {
  constexpr int NE = DIM + 2;
  R const gamma = R(gamma64), b = R(b64);
  // Mixed form: the stencil enters the FP32 reconstruction as differences
  // from cell i (taken in FP64), and cell i is added back in FP64 after the
  // back-transform.  Exact arithmetic gives the same flux: the WENO5
  // smoothness indicators see only differences and its candidate stencils
  // sum to one, so WENO(v - c) = WENO(v) - c, and R L = I.  In FP32 it keeps
  // the rounding error proportional to the local variation instead of the
  // state's magnitude.
  constexpr bool kShift = std::is_same<R, float>::value;
  Eig<NE, R> ev;
#pragma unroll
  for (int e = 0; e < NE; ++e) f[e] = R(0.0);
  R U[6][NE];
#pragma unroll
  for (int m = 0; m < 6; ++m)
#pragma unroll
    for (int e = 0; e < NE; ++e) U[m][e] = kShift ? R(Ud[m][e] - Ud[2][e]) : R(Ud[m][e]);
  CellQ<R> qt, qn;
  cell_quantities<DIM, R>(gamma64, b64, Ud[2], qt);
  cell_quantities<DIM, R>(gamma64, b64, Ud[3], qn);
  roe_eigensystem<DIM, DIR>(gamma, qt, qn, ev);
  R minus[NE], plus[NE];
#pragma unroll
  for (int n = 0; n < NE; ++n) {
    R uc[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      R s = R(0.0);
#pragma unroll
      for (int q = 0; q < NE; ++q) {
        int const l = seq<DIM, DIR>(q);
        if (!(kSkipZeros<R> && left_zero<DIM, DIR>(n, l))) s += ev.L[n][l] * U[m][l];
      }
      uc[m] = s;
    }
    minus[n] = weno_left<R, kWenoOneDiv<R>>(uc);
    plus[n] = weno_right<R, kWenoOneDiv<R>>(uc);
  }
  R cl[NE], cr[NE];
  roe_right_eigenvectors<DIM, DIR>(ev);
  to_physical<DIM, DIR>(minus, ev.Rm, cl);
  to_physical<DIM, DIR>(plus, ev.Rm, cr);
  if constexpr (kShift) {
#pragma unroll
    for (int e = 0; e < NE; ++e) {
      cl[e] = R(double(cl[e]) + Ud[2][e]);
      cr[e] = R(double(cr[e]) + Ud[2][e]);
    }
  }
  R const eps = R(2.220446049250313080847263336181640625e-16);  // DBL_EPSILON
  if (cl[kMass] <= eps || cr[kMass] <= eps) return;
  Prim<R> pl, pr;
  cons_to_prim<DIM>(gamma, b, cl, pl);
  cons_to_prim<DIM>(gamma, b, cr, pr);
  if (pl.p <= -b || pr.p <= -b) return;
  R hf[NE];
  hllc<DIM, DIR>(gamma, b, cl, cr, pl, pr, hf);
#pragma unroll
  for (int e = 0; e < NE; ++e) f[e] += hf[e];
}

// The six-cell stencil of a face: cells off + (m - 2) * stride, m = 0..5,
// through the block's buffer resource (one 32-bit byte offset per cell, the
// field as a scalar offset).
template <int DIM>
__device__ __forceinline__ void load_stencil(const BlockReader<DIM> &rd, int off, int stride,
                                             double (&Ud)[6][DIM + 2]) {
#pragma unroll
  for (int m = 0; m < 6; ++m)
#pragma unroll
    for (int e = 0; e < DIM + 2; ++e)
      Ud[m][e] = rd((off + (m - 2) * stride) * (int)sizeof(double),
                    e * Dims<DIM>::FIELD * (int)sizeof(double));
}

// One face flux from the six-cell stencil Ud[m] (FP64 conservatives of cells
// i-2..i+3 along DIR); reconstruction and Riemann arithmetic in R.
// FV:  FiniteVolumeScheme::ComputeFluxes (finite_volume_scheme.cpp:160-251)
//      + CharacteristicStateReconstruction (characteristic_state_reconstruction.h:60-134).
// Roe: FluxSplittingScheme::ComputeFluxes + ComputeAdvection
//      (flux_splitting_scheme.cpp:42-76, 185-315).
// Writes f (zero when the reference skips the face).
template <int DIM, int DIR, int FLUX, typename R>
__device__ __forceinline__ void face_flux(double gamma64, double b64,
                                          const double *__restrict__ blk, int off, R *f) {
  constexpr int NE = DIM + 2;
  const double *__restrict__ base = blk + off;
#pragma unroll
  for (int e = 0; e < NE; ++e) f[e] = R(0.0);
  if constexpr (FLUX == 0) {
    BlockReader<DIM> const rd(blk);
    double Ud[6][NE];
    load_stencil<DIM>(rd, off, axis_stride<DIM, DIR>(), Ud);
    fv_face<DIM, DIR, R>(gamma64, b64, Ud, f);
  } else if constexpr (std::is_same<R, double>::value && !kFastArith) {
    roe_split_flux_exact<DIM, DIR>(gamma64, b64, blk, off, f);
  } else {
    roe_split_flux<DIM, DIR, R>(gamma64, b64, base, f);
  }
}

// Loads the six-cell stencil around the face (i,j,k)+1/2 along DIR (halo
// indices of the cell left of the face) and computes its flux.
template <int DIM, int DIR, int FLUX, typename R>
__device__ __forceinline__ void stencil_flux(const double *__restrict__ blk, int i, int j, int k,
                                             double gamma, double b, R *f) {
  using D = Dims<DIM>;
  face_flux<DIM, DIR, FLUX, R>(gamma, b, blk, i * D::SX + j * D::SY + k, f);
}
