/*
 * ORACLE — test infrastructure only (see alpaca_oracle.h).  CPU restatement
 * of ALPACA's single-phase FV right-hand side, the SSP-RK3 stage update, the
 * same-level halo fill with zero-gradient / symmetry / periodic sides and the CFL time
 * step.  Each function cites the reference file:line it restates.  Build with
 * -ffp-contract=off: the reference x86 build has no FMA.
 */
#include "alpaca_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define NEMAX 5
#define TC ORA_TC
#define IC ORA_IC
#define HS ORA_HS

enum { MASS = 0, ENERGY = 1, MOMX = 2 };

/* std::min / std::max semantics: min(a,b) = (b < a) ? b : a. */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

/* ConsistencyManagedSum, full-symmetry form (mathematical_functions.h:79-91). */
static double cms3(double a, double b, double c) {
  double const s1 = (a + b) + c;
  double const s2 = (a + c) + b;
  double const s3 = (b + c) + a;
  double mn = s1, mx = s1;
  mn = smin(mn, s2);
  mn = smin(mn, s3);
  mx = smax(mx, s2);
  mx = smax(mx, s3);
  return 0.5 * (mn + mx);
}

/* DimensionAwareConsistencyManagedSum (mathematical_functions.h:100-108). */
static inline double dams(int dim, double a, double b, double c) {
  if (dim == 2)
    return a + b;
  return cms3(a, b, c);
}

static inline int signum(double x) { return (0.0 < x) - (x < 0.0); }

/* ---------------------------------------------------------------- EOS ---- */
/* GenericStiffenedGas::CalculatePressure<false> (generic_stiffened_gas.h:40-58). */
static inline double eos_pressure(int dim, double gamma, double b, double rho,
                                  double mx, double my, double mz, double e) {
  double const q = 0.5 * dams(dim, mx * mx, my * my, mz * mz);
  return -gamma * b + (gamma - 1.0) * (e - q / rho);
}
/* StiffenedGas::ComputeEnthalpy (stiffened_gas.cpp:80-87). */
static inline double eos_enthalpy(int dim, double gamma, double b, double rho,
                                  double mx, double my, double mz, double e) {
  return (e + eos_pressure(dim, gamma, b, rho, mx, my, mz, e)) / rho;
}
/* CalculateSpeedOfSound<false> (generic_stiffened_gas.h:84-97). */
static inline double eos_sos(double gamma, double b, double rho, double p) {
  return sqrt(gamma * (p + b) / rho);
}
/* CalculateEnergy (generic_stiffened_gas.h:61-71). */
static inline double eos_energy(int dim, double gamma, double b, double rho,
                                double u, double v, double w, double p) {
  return (p + gamma * b) / (gamma - 1.0) +
         (0.5 * dams(dim, u * u, v * v, w * w) * rho);
}
/* StiffenedGas::ComputePsi (stiffened_gas.cpp:127-130). */
static inline double eos_psi(double gamma, double b, double p, double one_rho) {
  return (p + gamma * b) * one_rho;
}

void ora_eos(int dim, double gamma, double b, const double *cons,
             const double *prim, double *out) {
  out[0] = eos_pressure(dim, gamma, b, cons[0], cons[1], cons[2], cons[3],
                        cons[4]);
  out[1] = eos_enthalpy(dim, gamma, b, cons[0], cons[1], cons[2], cons[3],
                        cons[4]);
  out[2] = eos_sos(gamma, b, prim[0], prim[4]);
  out[3] = eos_energy(dim, gamma, b, prim[0], prim[1], prim[2], prim[3],
                      prim[4]);
  out[4] = eos_psi(gamma, b, prim[4], 1.0 / prim[0]);
}

/* Primitive state of one cell: rho, p, velocity[3] (prime_state_handler.h:
 * 279-370; velocity = m * (1/rho), pressure from the EOS, absent momenta 0). */
typedef struct {
  double rho, p, vel[3];
} prim_t;

static inline void cons_to_prim(int dim, double gamma, double b,
                                const double *c, prim_t *o) {
  o->rho = c[MASS];
  double const one_density = 1.0 / c[MASS];
  o->vel[0] = c[MOMX] * one_density;
  o->vel[1] = c[MOMX + 1] * one_density;
  o->vel[2] = dim == 3 ? c[MOMX + 2] * one_density : 0.0;
  o->p = eos_pressure(dim, gamma, b, c[MASS], c[MOMX], c[MOMX + 1],
                      dim == 3 ? c[MOMX + 2] : 0.0, c[ENERGY]);
}

/* ------------------------------------------------------------- WENO5 ---- */
/* WENO5::ApplyImplementation (weno5.h:88-158), v1..v5 already ordered. */
static double weno5_core(double v1, double v2, double v3, double v4,
                         double v5) {
  double const c_s1 = 13.0 / 12.0, c_s2 = 0.25, eps = 1.0e-6;
  double const s11 = 1.0 * v1 + -2.0 * v2 + 1.0 * v3;
  double const s12 = 1.0 * v1 + -4.0 * v2 + 3.0 * v3;
  double const s1 = c_s1 * s11 * s11 + c_s2 * s12 * s12;
  double const s21 = 1.0 * v2 + -2.0 * v3 + 1.0 * v4;
  double const s22 = 1.0 * v2 + -1.0 * v4;
  double const s2 = c_s1 * s21 * s21 + c_s2 * s22 * s22;
  double const s31 = 1.0 * v3 + -2.0 * v4 + 1.0 * v5;
  double const s32 = 3.0 * v3 + -4.0 * v4 + 1.0 * v5;
  double const s3 = c_s1 * s31 * s31 + c_s2 * s32 * s32;
  double const a1 = 0.1 / ((s1 + eps) * (s1 + eps));
  double const a2 = 0.6 / ((s2 + eps) * (s2 + eps));
  double const a3 = 0.3 / ((s3 + eps) * (s3 + eps));
  double const one_a_sum = 1.0 / (a1 + a2 + a3);
  double const w1 = a1 * one_a_sum;
  double const w2 = a2 * one_a_sum;
  double const w3 = a3 * one_a_sum;
  return w1 * ((2.0 / 6.0) * v1 + (-7.0 / 6.0) * v2 + (11.0 / 6.0) * v3) +
         w2 * ((-1.0 / 6.0) * v2 + (5.0 / 6.0) * v3 + (2.0 / 6.0) * v4) +
         w3 * ((2.0 / 6.0) * v3 + (5.0 / 6.0) * v4 + (-1.0 / 6.0) * v5);
}
/* UpwindLeft = {offset 0, sign +1}, UpwindRight = {1, -1} (apply_wrapper.h:
 * 48-58) on the 6-value stencil with downstream size 2. */
static inline double weno_left(const double *a) {
  return weno5_core(a[0], a[1], a[2], a[3], a[4]);
}
static inline double weno_right(const double *a) {
  return weno5_core(a[5], a[4], a[3], a[2], a[1]);
}

void ora_weno5(const double *v, double *out) {
  out[0] = weno_left(v);
  out[1] = weno_right(v);
}

/* ---------------------------------------------------- Roe eigensystem ---- */
/* Minor directions of each principal direction (eigendecomposition.h:97-130). */
static const int kMinor[3][2] = {{1, 2}, {0, 2}, {0, 1}};

/* Symmetry-preserving projection order (state_reconstruction.h:37-57). */
static const int kSeq2[2][4] = {{2, 3, 0, 1}, {2, 3, 0, 1}};
static const int kSeq3[3][5] = {
    {3, 4, 2, 0, 1}, {4, 2, 3, 0, 1}, {2, 3, 4, 0, 1}};

static inline const int *sum_sequence(int dim, int dir) {
  return dim == 2 ? kSeq2[dir] : kSeq3[dir];
}

typedef struct {
  double L[NEMAX][NEMAX];
  double R[NEMAX][NEMAX];
  double lam[NEMAX];
} eig_t;

/* EigenDecomposition::ComputeRoeEigendecomposition for one face
 * (eigendecomposition.h:238-466).  t = cell i, n = cell i+1 along dir; the
 * velocities and pressure are the cells' prime states. */
static int roe_eigensystem(int dim, int dir, double gamma, double b,
                           const double *ct, const prim_t *pt,
                           const double *cn, const prim_t *pn, eig_t *ev) {
  int const ne = dim + 2;
  memset(ev, 0, sizeof(*ev));
  if (ct[MASS] <= 0.0 || cn[MASS] <= 0.0)
    return 0;
  int const pr = dir, m1 = kMinor[dir][0], m2 = kMinor[dir][1];
  int const ev_p = dim; /* eigenvector index of the principal momentum */
  int const Mp = MOMX + pr, M1 = MOMX + m1, M2 = MOMX + m2;

  double const rho_t = ct[MASS];
  double const one_rho_t = 1.0 / ct[MASS];
  double const u_t = pt->vel[pr];
  double const v_t = pt->vel[m1];
  double const w_t = dim == 3 ? pt->vel[m2] : 0.0;
  double const p_t = pt->p;
  double const psi_t = eos_psi(gamma, b, p_t, one_rho_t);

  double const rho_n = cn[MASS];
  double const one_rho_n = 1.0 / cn[MASS];
  double const u_n = pn->vel[pr];
  double const v_n = pn->vel[m1];
  double const w_n = dim == 3 ? pn->vel[m2] : 0.0;
  double const p_n = pn->p;
  double const psi_n = eos_psi(gamma, b, p_n, one_rho_n);

  double const sq_t = sqrt(rho_t);
  double const sq_n = sqrt(rho_n);
  double const rho_div = 1.0 / (sq_t + sq_n);
  double const rho_roe = sq_t * sq_n;
  double const one_rho_roe = 1.0 / rho_roe;

  double const u = ((u_t * sq_t) + (u_n * sq_n)) * rho_div;
  double const v = ((v_t * sq_t) + (v_n * sq_n)) * rho_div;
  double const w = ((w_t * sq_t) + (w_n * sq_n)) * rho_div;
  double const psi = ((psi_t * sq_t) + (psi_n * sq_n)) * rho_div;
  double const g = gamma - 1.0;
  double const mz_t = dim == 3 ? ct[MOMX + 2] : 0.0;
  double const mz_n = dim == 3 ? cn[MOMX + 2] : 0.0;
  double const H =
      (eos_enthalpy(dim, gamma, b, rho_t, ct[MOMX], ct[MOMX + 1], mz_t,
                    ct[ENERGY]) *
           sq_t +
       eos_enthalpy(dim, gamma, b, rho_n, cn[MOMX], cn[MOMX + 1], mz_n,
                    cn[ENERGY]) *
           sq_n) *
      rho_div;

  double const q2 = dams(dim, u * u, v * v, w * w);
  double const tmp =
      dams(dim, (u_t - u_n) * (u_t - u_n), (v_t - v_n) * (v_t - v_n),
           (w_t - w_n) * (w_t - w_n));
  double const pod = ((p_t * one_rho_t) * sq_t + (p_n * one_rho_n) * sq_n) *
                         rho_div +
                     0.5 * rho_roe * (rho_div * rho_div) * tmp;
  double const cc = psi + g * pod;
  double const one_cc = 1.0 / cc;
  double const c = sqrt(cc);

  /* Left eigenvectors (eigendecomposition.h:316-359). */
  ev->L[0][MASS] = 0.5 * one_cc * (g * q2 - g * H + (u + c) * c);
  ev->L[0][Mp] = 0.5 * one_cc * (-u * g - c);
  ev->L[0][M1] = 0.5 * one_cc * (-v * g);
  if (dim == 3)
    ev->L[0][M2] = 0.5 * one_cc * (-w * g);
  ev->L[0][ENERGY] = 0.5 * one_cc * g;

  ev->L[1][MASS] = v * one_rho_roe;
  ev->L[1][Mp] = 0.0;
  ev->L[1][M1] = -one_rho_roe;
  if (dim == 3)
    ev->L[1][M2] = 0.0;
  ev->L[1][ENERGY] = 0.0;

  if (dim == 3) {
    ev->L[2][MASS] = -w * one_rho_roe;
    ev->L[2][Mp] = 0.0;
    ev->L[2][M1] = 0.0;
    ev->L[2][M2] = one_rho_roe;
    ev->L[2][ENERGY] = 0.0;
  }

  ev->L[ev_p][MASS] = one_cc * (H - q2);
  ev->L[ev_p][Mp] = one_cc * u;
  ev->L[ev_p][M1] = one_cc * v;
  if (dim == 3)
    ev->L[ev_p][M2] = one_cc * w;
  ev->L[ev_p][ENERGY] = -one_cc;

  ev->L[ne - 1][MASS] = 0.5 * one_cc * (g * q2 - g * H - (u - c) * c);
  ev->L[ne - 1][Mp] = 0.5 * one_cc * (-u * g + c);
  ev->L[ne - 1][M1] = 0.5 * one_cc * (-v * g);
  if (dim == 3)
    ev->L[ne - 1][M2] = 0.5 * one_cc * (-w * g);
  ev->L[ne - 1][ENERGY] = 0.5 * one_cc * g;

  /* Right eigenvectors (eigendecomposition.h:366-409). */
  ev->R[MASS][0] = 1.0;
  ev->R[MASS][1] = 0.0;
  if (dim == 3)
    ev->R[MASS][2] = 0.0;
  ev->R[MASS][ev_p] = g;
  ev->R[MASS][ne - 1] = 1.0;

  ev->R[Mp][0] = u - c;
  ev->R[Mp][1] = 0.0;
  if (dim == 3)
    ev->R[Mp][2] = 0.0;
  ev->R[Mp][ev_p] = g * u;
  ev->R[Mp][ne - 1] = u + c;

  ev->R[M1][0] = v;
  ev->R[M1][1] = -rho_roe;
  if (dim == 3)
    ev->R[M1][2] = 0.0;
  ev->R[M1][ev_p] = g * v;
  ev->R[M1][ne - 1] = v;

  if (dim == 3) {
    ev->R[M2][0] = w;
    ev->R[M2][1] = 0.0;
    ev->R[M2][2] = rho_roe;
    ev->R[M2][ev_p] = g * w;
    ev->R[M2][ne - 1] = w;
  }

  ev->R[ENERGY][0] = H - u * c;
  ev->R[ENERGY][1] = -v * rho_roe;
  if (dim == 3)
    ev->R[ENERGY][2] = rho_roe * w;
  ev->R[ENERGY][ev_p] = g * H - c * c;
  ev->R[ENERGY][ne - 1] = H + u * c;

  /* Roe flux-function eigenvalues (eigendecomposition.h:416-421). */
  ev->lam[0] = fabs(u - c);
  ev->lam[ne - 1] = fabs(u + c);
  for (int l = 1; l < ne - 1; ++l)
    ev->lam[l] = fabs(u);
  return 1;
}

int ora_roe_eigensystem(int dim, int dir, double gamma, double b,
                        const double *cons_t, const double *cons_n, double *L,
                        double *R, double *lambda) {
  int const ne = dim + 2;
  prim_t pt, pn;
  cons_to_prim(dim, gamma, b, cons_t, &pt);
  cons_to_prim(dim, gamma, b, cons_n, &pn);
  eig_t ev;
  int const ok =
      roe_eigensystem(dim, dir, gamma, b, cons_t, &pt, cons_n, &pn, &ev);
  for (int r = 0; r < ne; ++r) {
    for (int c = 0; c < ne; ++c) {
      L[r * ne + c] = ev.L[r][c];
      R[r * ne + c] = ev.R[r][c];
    }
    lambda[r] = ev.lam[r];
  }
  return ok;
}

/* TransformToPhysicalSpace (eigendecomposition.h:144-165). */
static void to_physical(int ne, const double *ch, const double R[][NEMAX],
                        double *out) {
  for (int l = 0; l < ne; ++l) {
    double s = 0.0;
    for (int n = 1; n < ne - 1; ++n)
      s += ch[n] * R[l][n];
    s += (ch[0] * R[l][0] + ch[ne - 1] * R[l][ne - 1]);
    out[l] = s;
  }
}

/* ------------------------------------------------------------- HLLC ---- */
/* CalculateSignalSpeed, Einfeldt branch (hll_signal_speed_calculator.h:
 * 77-101). */
static inline void einfeldt(double rl, double rr, double ul, double ur,
                            double cl, double cr, double *sl, double *sr) {
  double const sql = sqrt(rl);
  double const sqr = sqrt(rr);
  double const rho_div = 1.0 / (sql + sqr);
  double const u_roe = ((ul * sql) + (ur * sqr)) * rho_div;
  double const c_e = sqrt(((cl * cl * sql) + (cr * cr * sqr)) * rho_div +
                          (0.5 * sql * sqr * rho_div * rho_div) * (ur - ul) *
                              (ur - ul));
  *sl = smin(u_roe - c_e, ul - cl);
  *sr = smax(u_roe + c_e, ur + cr);
}

/* HllcRiemannSolver::SolveRiemannProblemImplementation (hllc_riemann_solver.h:
 * 52-198), including the clamped signal speed in the energy star state. */
static void hllc_flux(int dim, int dir, double gamma, double b,
                      const double *cl, const double *cr, const prim_t *pl,
                      const prim_t *pr, double *f) {
  int const ne = dim + 2;
  int const Mp = MOMX + dir;
  double qsl[NEMAX], qsr[NEMAX], fl[NEMAX], fr[NEMAX];
  double const p_l = pl->p, p_r = pr->p;
  double const one_rho_l = 1.0 / cl[MASS];
  double const one_rho_r = 1.0 / cr[MASS];
  double const u_l = pl->vel[dir], u_r = pr->vel[dir];
  double const c_l = eos_sos(gamma, b, cl[MASS], p_l);
  double const c_r = eos_sos(gamma, b, cr[MASS], p_r);
  double sls, srs;
  einfeldt(cl[MASS], cr[MASS], u_l, u_r, c_l, c_r, &sls, &srs);
  double const sc =
      ((p_r - p_l) + (cl[Mp] * (sls - u_l) - cr[Mp] * (srs - u_r))) /
      (cl[MASS] * (sls - u_l) - cr[MASS] * (srs - u_r));
  double const sl = smin(sls, 0.0);
  double const sr = smax(srs, 0.0);
  double const chi_l = (sls - u_l) / (sls - sc);
  double const chi_r = (srs - u_r) / (srs - sc);

  qsl[MASS] = cl[MASS] * chi_l;
  qsl[Mp] = cl[MASS] * chi_l * sc;
  qsl[ENERGY] = cl[MASS] * chi_l *
                (cl[ENERGY] * one_rho_l +
                 (sc - u_l) * (sc + p_l / (cl[MASS] * (sl - u_l))));
  qsr[MASS] = cr[MASS] * chi_r;
  qsr[Mp] = cr[MASS] * chi_r * sc;
  qsr[ENERGY] = cr[MASS] * chi_r *
                (cr[ENERGY] * one_rho_r +
                 (sc - u_r) * (sc + p_r / (cr[MASS] * (sr - u_r))));

  fl[MASS] = cl[Mp];
  fl[Mp] = ((cl[Mp] * cl[Mp]) * one_rho_l) + p_l;
  fl[ENERGY] = u_l * (cl[ENERGY] + p_l);
  fr[MASS] = cr[Mp];
  fr[Mp] = ((cr[Mp] * cr[Mp]) * one_rho_r) + p_r;
  fr[ENERGY] = u_r * (cr[ENERGY] + p_r);

  for (int d = 0; d < dim - 1; ++d) {
    int const Mm = MOMX + kMinor[dir][d];
    qsl[Mm] = chi_l * cl[Mm];
    qsr[Mm] = chi_r * cr[Mm];
    fl[Mm] = u_l * cl[Mm];
    fr[Mm] = u_r * cr[Mm];
  }
  int const sg = signum(sc);
  for (int n = 0; n < ne; ++n) {
    double const fsl = fl[n] + (sl * (qsl[n] - cl[n]));
    double const fsr = fr[n] + (sr * (qsr[n] - cr[n]));
    f[n] = (0.5 * (1.0 + sg) * fsl + 0.5 * (1.0 - sg) * fsr);
  }
}

void ora_hllc(int dim, int dir, double gamma, double b, const double *cl,
              const double *cr, double *out) {
  prim_t pl, pr;
  cons_to_prim(dim, gamma, b, cl, &pl);
  cons_to_prim(dim, gamma, b, cr, &pr);
  hllc_flux(dim, dir, gamma, b, cl, cr, &pl, &pr, out);
}

/* ------------------------------------------------------ face fluxes ---- */
typedef struct {
  int dim, ne;
  int tz;              /* TCZ */
  ptrdiff_t stride[3]; /* element strides of i, j, k */
  ptrdiff_t field;     /* elements per field */
} layout_t;

static layout_t make_layout(int dim) {
  layout_t L;
  L.dim = dim;
  L.ne = dim + 2;
  L.tz = dim == 3 ? TC : 1;
  L.stride[0] = (ptrdiff_t)TC * L.tz;
  L.stride[1] = L.tz;
  L.stride[2] = 1;
  L.field = (ptrdiff_t)TC * TC * L.tz;
  return L;
}

static inline void load_cell(const layout_t *L, const double *blk,
                             ptrdiff_t off, double *c) {
  for (int e = 0; e < L->ne; ++e)
    c[e] = blk[e * L->field + off];
}

/* Flux of one face (cell `off`, face +1/2 along dir) for either family.
 * FV: FiniteVolumeScheme::ComputeFluxes (finite_volume_scheme.cpp:160-251)
 *     with CharacteristicStateReconstruction (characteristic_state_
 *     reconstruction.h:60-134).
 * Roe: FluxSplittingScheme::ComputeFluxes (flux_splitting_scheme.cpp:185-315)
 *     with ComputeAdvection (:42-76).
 * Returns 0 if the face is skipped (flux left at zero). */
static int face_flux(const layout_t *L, int flux, int dir, double gamma,
                     double b, const double *blk, ptrdiff_t off, double *f) {
  int const dim = L->dim, ne = L->ne;
  ptrdiff_t const s = L->stride[dir];
  double U[6][NEMAX];
  prim_t P[6];
  for (int m = 0; m < 6; ++m) {
    load_cell(L, blk, off + (m - 2) * s, U[m]);
    cons_to_prim(dim, gamma, b, U[m], &P[m]);
  }
  eig_t ev;
  roe_eigensystem(dim, dir, gamma, b, U[2], &P[2], U[3], &P[3], &ev);
  const int *seq = sum_sequence(dim, dir);
  for (int e = 0; e < ne; ++e)
    f[e] = 0.0;

  if (flux == ORA_FLUX_HLLC) {
    double minus[NEMAX], plus[NEMAX];
    for (int n = 0; n < ne; ++n) {
      double uc[6];
      for (int m = 0; m < 6; ++m) {
        uc[m] = 0.0;
        for (int q = 0; q < ne; ++q) {
          int const l = seq[q];
          uc[m] += ev.L[n][l] * U[m][l];
        }
      }
      minus[n] = weno_left(uc);
      plus[n] = weno_right(uc);
    }
    double cl[NEMAX], cr[NEMAX];
    to_physical(ne, minus, (const double(*)[NEMAX])ev.R, cl);
    to_physical(ne, plus, (const double(*)[NEMAX])ev.R, cr);
    if (cl[MASS] <= DBL_EPSILON || cr[MASS] <= DBL_EPSILON)
      return 0;
    prim_t pl, pr;
    cons_to_prim(dim, gamma, b, cl, &pl);
    cons_to_prim(dim, gamma, b, cr, &pr);
    if (pl.p <= -b || pr.p <= -b)
      return 0;
    double hf[NEMAX];
    hllc_flux(dim, dir, gamma, b, cl, cr, &pl, &pr, hf);
    for (int e = 0; e < ne; ++e)
      f[e] += hf[e];
    return 1;
  }

  /* Roe flux splitting. */
  double adv[6][NEMAX];
  for (int m = 0; m < 6; ++m) {
    double const rho = P[m].rho, p = P[m].p, ud = P[m].vel[dir];
    adv[m][MASS] = rho * ud;
    adv[m][ENERGY] = (U[m][ENERGY] + p) * ud;
    for (int d = 0; d < dim; ++d)
      adv[m][MOMX + d] = rho * (ud * P[m].vel[d]);
    adv[m][MOMX + dir] += p;
  }
  double chf[NEMAX] = {0.0};
  for (int n = 0; n < ne; ++n) {
    double pos[6], neg[6];
    for (int m = 0; m < 6; ++m) {
      pos[m] = 0.0;
      neg[m] = 0.0;
      for (int q = 0; q < ne; ++q) {
        int const l = seq[q];
        double const uct = U[m][l] * ev.L[n][l];
        double const act = adv[m][l] * ev.L[n][l];
        pos[m] += (act + ev.lam[n] * uct);
        neg[m] += (act - ev.lam[n] * uct);
      }
    }
    chf[n] = 0.5 * (weno_left(pos) + weno_right(neg));
  }
  double phys[NEMAX];
  to_physical(ne, chf, (const double(*)[NEMAX])ev.R, phys);
  for (int e = 0; e < ne; ++e)
    f[e] += phys[e];
  return 1;
}

/* Face-flux arrays [NE][IC+1][IC+1][ICZ+1] as the reference's UpdateFluxes
 * holds them: the face i+1/2 of cell i (halo index) sits at i - (HS-1). */
static void block_face_fluxes(int dim, int flux, double gamma, double b,
                              const double *blk, double *fx, double *fy,
                              double *fz) {
  layout_t const L = make_layout(dim);
  int const icz = dim == 3 ? IC : 1;
  size_t const n1 = IC + 1, nz = (size_t)icz + 1;
  size_t const per_field = n1 * n1 * nz;
  double *out[3] = {fx, fy, fz};
  for (int d = 0; d < 3; ++d)
    if (out[d])
      memset(out[d], 0, sizeof(double) * (size_t)L.ne * per_field);
  for (int dir = 0; dir < dim; ++dir) {
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      if (a >= dim) {
        lo[a] = 0;
        hi[a] = 0;
      } else {
        lo[a] = (a == dir) ? HS - 1 : HS;
        hi[a] = HS + IC - 1;
      }
    }
    /* Offsets from halo indices to face-array indices
     * (finite_volume_scheme.cpp:176-180): HS-1 for active axes, -1 for z in 2D. */
    int const oz = dim == 3 ? HS - 1 : -1;
    for (int i = lo[0]; i <= hi[0]; ++i)
      for (int j = lo[1]; j <= hi[1]; ++j)
        for (int k = lo[2]; k <= hi[2]; ++k) {
          ptrdiff_t const off =
              i * L.stride[0] + j * L.stride[1] + k * L.stride[2];
          double f[NEMAX];
          face_flux(&L, flux, dir, gamma, b, blk, off, f);
          size_t const fi = (size_t)(i - (HS - 1)), fj = (size_t)(j - (HS - 1)),
                       fk = (size_t)(k - oz);
          for (int e = 0; e < L.ne; ++e)
            out[dir][(((size_t)e * n1 + fi) * n1 + fj) * nz + fk] += f[e];
        }
  }
}

void ora_block_fluxes(int dim, int flux, double gamma, double b,
                      const double *avg, double cell_size, double *out) {
  (void)cell_size;
  int const icz = dim == 3 ? IC : 1;
  size_t const n = (size_t)(dim + 2) * (IC + 1) * (IC + 1) * (icz + 1);
  block_face_fluxes(dim, flux, gamma, b, avg, out, out + n, out + 2 * n);
}

/* SpaceSolver::UpdateFluxes divergence (space_solver.cpp:131-153): RHS starts
 * at zero, volume forces are zero (sources off).  out: [NE][IC][IC][ICZ]. */
static void block_rhs(int dim, int flux, double gamma, double b,
                      const double *blk, double cell_size, double *out,
                      double *scratch) {
  int const ne = dim + 2, icz = dim == 3 ? IC : 1;
  size_t const n1 = IC + 1, nz = (size_t)icz + 1;
  size_t const nf = (size_t)ne * n1 * n1 * nz;
  double *fx = scratch, *fy = scratch + nf, *fz = scratch + 2 * nf;
  block_face_fluxes(dim, flux, gamma, b, blk, fx, fy, fz);
  double const one_cell_size = 1.0 / cell_size;
#define FI(a, e, i, j, k) a[(((size_t)(e)*n1 + (i)) * n1 + (j)) * nz + (k)]
  for (int e = 0; e < ne; ++e)
    for (int i = 0; i < IC; ++i)
      for (int j = 0; j < IC; ++j)
        for (int k = 0; k < icz; ++k) {
          double const s =
              dams(dim, FI(fx, e, i, j + 1, k + 1) - FI(fx, e, i + 1, j + 1, k + 1),
                   FI(fy, e, i + 1, j, k + 1) - FI(fy, e, i + 1, j + 1, k + 1),
                   FI(fz, e, i + 1, j + 1, k) - FI(fz, e, i + 1, j + 1, k + 1));
          out[(((size_t)e * IC + i) * IC + j) * icz + k] =
              0.0 + (0.0 + s * one_cell_size);
        }
#undef FI
}

static size_t rhs_scratch_doubles(int dim) {
  int const icz = dim == 3 ? IC : 1;
  return 3 * (size_t)(dim + 2) * (IC + 1) * (IC + 1) * (icz + 1);
}

void ora_block_rhs(int dim, int flux, double gamma, double b,
                   const double *avg, double cell_size, double *out) {
  double *scratch = (double *)malloc(sizeof(double) * rhs_scratch_doubles(dim));
  block_rhs(dim, flux, gamma, b, avg, cell_size, out, scratch);
  free(scratch);
}

/* ---------------------------------------------------------- full run ---- */
typedef struct {
  ora_grid g;
  layout_t L;
  int nblk;
  size_t bsize; /* doubles per haloed block */
  double *avg, *rhs, *init;
} run_t;

typedef void (*block_fn)(run_t *r, int lo, int hi, void *arg);
typedef struct {
  run_t *r;
  block_fn fn;
  int lo, hi;
  void *arg;
} job_t;

static void *job_main(void *p) {
  job_t *j = (job_t *)p;
  j->fn(j->r, j->lo, j->hi, j->arg);
  return NULL;
}

static void parallel_blocks(run_t *r, int threads, block_fn fn, void *arg) {
  int const n = r->nblk;
  if (threads <= 1 || n <= 1) {
    fn(r, 0, n, arg);
    return;
  }
  if (threads > n)
    threads = n;
  pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
  job_t *jobs = (job_t *)malloc(sizeof(job_t) * (size_t)threads);
  for (int t = 0; t < threads; ++t) {
    jobs[t].r = r;
    jobs[t].fn = fn;
    jobs[t].lo = (int)((long long)n * t / threads);
    jobs[t].hi = (int)((long long)n * (t + 1) / threads);
    jobs[t].arg = arg;
    pthread_create(&tid[t], NULL, job_main, &jobs[t]);
  }
  for (int t = 0; t < threads; ++t)
    pthread_join(tid[t], NULL);
  free(tid);
  free(jobs);
}

static inline void block_coords(const run_t *r, int b, int *c) {
  c[0] = b / (r->g.nb[1] * r->g.nb[2]);
  c[1] = (b / r->g.nb[2]) % r->g.nb[1];
  c[2] = b % r->g.nb[2];
}

/* Same-level halo copy (internal_halo_manager.cpp:439-466): every halo cell of
 * every block takes the value of the internal cell that owns its position,
 * wrapped across periodic sides; positions outside a zero-gradient or
 * symmetry side are left to the external BC below. */
static void halo_copy(run_t *r, int lo, int hi, void *arg) {
  (void)arg;
  layout_t const *L = &r->L;
  int const dim = L->dim;
  for (int b = lo; b < hi; ++b) {
    int bc[3];
    block_coords(r, b, bc);
    double *dst = r->rhs + (size_t)b * r->bsize;
    for (int i = 0; i < TC; ++i)
      for (int j = 0; j < (dim >= 2 ? TC : 1); ++j)
        for (int k = 0; k < L->tz; ++k) {
          int const hx = i < HS || i >= HS + IC;
          int const hy = j < HS || j >= HS + IC;
          int const hz = dim == 3 && (k < HS || k >= HS + IC);
          if (!(hx || hy || hz))
            continue;
          int idx[3] = {i - HS, j - HS, k - HS};
          int blk[3] = {bc[0], bc[1], bc[2]};
          int outside = 0;
          for (int d = 0; d < dim; ++d) {
            int gc = blk[d] * IC + idx[d];
            int const n = r->g.nb[d] * IC;
            if (gc < 0 || gc >= n) {
              if (r->g.bc[d] != ORA_BC_PERIODIC) {
                outside = 1;
                break;
              }
              gc = (gc % n + n) % n;
            }
            blk[d] = gc / IC;
            idx[d] = gc % IC;
          }
          if (outside)
            continue;
          int const src_b = (blk[0] * r->g.nb[1] + blk[1]) * r->g.nb[2] +
                            (dim == 3 ? blk[2] : 0);
          const double *src = r->rhs + (size_t)src_b * r->bsize;
          ptrdiff_t const so = (idx[0] + HS) * L->stride[0] +
                               (idx[1] + HS) * L->stride[1] +
                               (dim == 3 ? (idx[2] + HS) : 0);
          ptrdiff_t const doff =
              i * L->stride[0] + j * L->stride[1] + k * L->stride[2];
          for (int e = 0; e < L->ne; ++e)
            dst[e * L->field + doff] = src[e * L->field + so];
        }
  }
}

/* External boundary conditions on the halo slab over the full transverse
 * extent, in the order East, West, North, South, Top, Bottom:
 *  - ZeroGradientBoundaryCondition (zero_gradient_boundary_condition.h:45-57,
 *    boundary_constants.h:212-253): the first/last internal layer;
 *  - SymmetryBoundaryCondition (symmetry_boundary_condition.h:89-109,
 *    boundary_constants.h:169-208, mirror offsets :39-48): the internal cell
 *    mirrored about the domain face times a sign that is -1 for the
 *    face-normal momentum only (symmetry_boundary_condition.h:39-53,148-215). */
static void external_bc(run_t *r, int lo, int hi, void *arg) {
  (void)arg;
  layout_t const *L = &r->L;
  int const dim = L->dim;
  for (int b = lo; b < hi; ++b) {
    int bc[3];
    block_coords(r, b, bc);
    double *v = r->rhs + (size_t)b * r->bsize;
    for (int d = 0; d < dim; ++d) {
      int const kind = r->g.bc[d];
      if (kind == ORA_BC_PERIODIC)
        continue;
      for (int side = 1; side >= 0; --side) { /* high side first */
        if (side == 1 && bc[d] != r->g.nb[d] - 1)
          continue;
        if (side == 0 && bc[d] != 0)
          continue;
        int const src_layer = side ? HS + IC - 1 : HS;
        /* mirror: halo index h <-> internal index mirror - h */
        int const mirror = side ? 2 * (HS + IC) - 1 : 2 * HS - 1;
        int const h0 = side ? HS + IC : 0;
        int ext[3] = {TC, dim >= 2 ? TC : 1, L->tz};
        for (int a = 0; a < ext[0]; ++a)
          for (int c = 0; c < ext[1]; ++c)
            for (int z = 0; z < ext[2]; ++z) {
              int idx[3] = {a, c, z};
              if (idx[d] < h0 || idx[d] >= h0 + HS)
                continue;
              int sidx[3] = {a, c, z};
              sidx[d] = kind == ORA_BC_SYMMETRY ? mirror - idx[d] : src_layer;
              ptrdiff_t const doff = idx[0] * L->stride[0] +
                                     idx[1] * L->stride[1] + idx[2];
              ptrdiff_t const soff = sidx[0] * L->stride[0] +
                                     sidx[1] * L->stride[1] + sidx[2];
              for (int e = 0; e < L->ne; ++e) {
                double const val = v[e * L->field + soff];
                /* equation order Mass, Energy, MomX, MomY, MomZ */
                v[e * L->field + doff] =
                    (kind == ORA_BC_SYMMETRY && e == 2 + d) ? -1.0 * val : val;
              }
            }
      }
    }
  }
}

typedef struct {
  int stage;
  double dt;
  double *maxs; /* per block signal-speed maximum */
} stage_arg_t;

/* ComputeTimestepSize inner loop (modular_algorithm_assembler.cpp:1840-1869). */
static void block_signal(run_t *r, int lo, int hi, void *arg) {
  stage_arg_t *sa = (stage_arg_t *)arg;
  layout_t const *L = &r->L;
  int const dim = L->dim;
  for (int b = lo; b < hi; ++b) {
    const double *a = r->avg + (size_t)b * r->bsize;
    double s = 0.0;
    for (int i = HS; i < HS + IC; ++i)
      for (int j = HS; j < HS + IC; ++j)
        for (int k = (dim == 3 ? HS : 0); k < (dim == 3 ? HS + IC : 1); ++k) {
          double c[NEMAX];
          load_cell(L, a, i * L->stride[0] + j * L->stride[1] + k, c);
          prim_t p;
          cons_to_prim(dim, r->g.gamma, r->g.b, c, &p);
          double const cs = eos_sos(r->g.gamma, r->g.b, p.rho, p.p);
          double vps[3] = {0.0, 0.0, 0.0};
          for (int d = 0; d < dim; ++d)
            vps[d] = fabs(p.vel[d]) + cs;
          double const sum =
              dim == 2 ? vps[0] + vps[1] : cms3(vps[0], vps[1], vps[2]);
          s = smax(s, sum);
        }
    sa->maxs[b] = s;
  }
}

/* The same signal maximum for one haloed block (used by the multi-level
 * driver's checker, whose leaves have different cell sizes: the reference
 * divides the maximum over all leaves by the finest cell size,
 * modular_algorithm_assembler.cpp:1932). */
double ora_block_signal(int dim, double gamma, double b, const double *avg) {
  layout_t const L = make_layout(dim);
  double s = 0.0;
  for (int i = HS; i < HS + IC; ++i)
    for (int j = HS; j < HS + IC; ++j)
      for (int k = (dim == 3 ? HS : 0); k < (dim == 3 ? HS + IC : 1); ++k) {
        double c[NEMAX] = {0.0};
        load_cell(&L, avg, i * L.stride[0] + j * L.stride[1] + k, c);
        prim_t p;
        cons_to_prim(dim, gamma, b, c, &p);
        double const cs = eos_sos(gamma, b, p.rho, p.p);
        double vps[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < dim; ++d)
          vps[d] = fabs(p.vel[d]) + cs;
        double const sum = dim == 2 ? vps[0] + vps[1] : cms3(vps[0], vps[1], vps[2]);
        s = smax(s, sum);
      }
  return s;
}

/* One RK stage on blocks [lo, hi): FillInitialBuffer, UpdateFluxes,
 * PrepareBufferForIntegration, IntegrateNode (time_integrator.h:98-113,
 * 312-340, 369-457; runge_kutta_3_TVD.h:39-59). */
static void block_stage(run_t *r, int lo, int hi, void *arg) {
  stage_arg_t *sa = (stage_arg_t *)arg;
  layout_t const *L = &r->L;
  int const dim = L->dim, ne = L->ne, icz = dim == 3 ? IC : 1;
  static const double kConsMul[3] = {1.0, 0.25, 2.0 / 3.0};
  static const double kBufMul[2][2] = {{0.25, 0.75}, {2.0 / 3.0, 1.0 / 3.0}};
  double *scratch = (double *)malloc(sizeof(double) * rhs_scratch_doubles(dim));
  double *rhs_int =
      (double *)malloc(sizeof(double) * (size_t)ne * IC * IC * (size_t)icz);
  double const timestep = 0.0 + sa->dt; /* accumulate over one micro step */
  double const cdt = kConsMul[sa->stage] * timestep;
  for (int b = lo; b < hi; ++b) {
    double *a = r->avg + (size_t)b * r->bsize;
    double *u0 = r->init + (size_t)b * r->bsize;
    double *out = r->rhs + (size_t)b * r->bsize;
    if (sa->stage == 0)
      memcpy(u0, a, sizeof(double) * r->bsize);
    block_rhs(dim, r->g.flux, r->g.gamma, r->g.b, a, r->g.cell_size, rhs_int,
              scratch);
    if (sa->stage != 0) {
      double const m0 = kBufMul[sa->stage - 1][0], m1 = kBufMul[sa->stage - 1][1];
      for (size_t x = 0; x < r->bsize; ++x)
        a[x] = m0 * a[x] + m1 * u0[x];
    }
    for (int e = 0; e < ne; ++e)
      for (int i = 0; i < IC; ++i)
        for (int j = 0; j < IC; ++j)
          for (int k = 0; k < icz; ++k) {
            ptrdiff_t const off = e * L->field + (i + HS) * L->stride[0] +
                                  (j + HS) * L->stride[1] +
                                  (dim == 3 ? k + HS : 0);
            out[off] =
                a[off] + cdt * rhs_int[(((size_t)e * IC + i) * IC + j) * icz + k];
          }
  }
  free(scratch);
  free(rhs_int);
}

static void halo_and_swap(run_t *r, int threads) {
  parallel_blocks(r, threads, halo_copy, NULL);
  parallel_blocks(r, threads, external_bc, NULL);
  double *t = r->avg;
  r->avg = r->rhs;
  r->rhs = t;
}

int ora_run(const ora_grid *g, double cfl, double t_end, int max_steps,
            int threads, double *state, double *dts, double *t_final) {
  run_t r;
  r.g = *g;
  if (g->dim == 2) {
    r.g.nb[2] = 1;
    r.g.bc[2] = ORA_BC_ZERO_GRADIENT;
  }
  r.L = make_layout(g->dim);
  r.nblk = r.g.nb[0] * r.g.nb[1] * r.g.nb[2];
  r.bsize = (size_t)r.L.ne * (size_t)r.L.field;
  size_t const total = (size_t)r.nblk * r.bsize;
  r.avg = (double *)calloc(total, sizeof(double));
  r.rhs = (double *)calloc(total, sizeof(double));
  r.init = (double *)calloc(total, sizeof(double));
  int const dim = g->dim, ne = r.L.ne, icz = dim == 3 ? IC : 1;
  size_t const NX = (size_t)r.g.nb[0] * IC, NY = (size_t)r.g.nb[1] * IC,
               NZ = (size_t)r.g.nb[2] * icz;
  /* ImposeInitialCondition into the RHS buffer, halo update, swap
   * (modular_algorithm_assembler.cpp:397-517). */
  for (int b = 0; b < r.nblk; ++b) {
    int bc[3];
    block_coords(&r, b, bc);
    double *o = r.rhs + (size_t)b * r.bsize;
    for (int e = 0; e < ne; ++e)
      for (int i = 0; i < IC; ++i)
        for (int j = 0; j < IC; ++j)
          for (int k = 0; k < icz; ++k)
            o[e * r.L.field + (i + HS) * r.L.stride[0] + (j + HS) * r.L.stride[1] +
              (dim == 3 ? k + HS : 0)] =
                state[(((size_t)e * NX + (size_t)bc[0] * IC + i) * NY +
                       (size_t)bc[1] * IC + j) *
                          NZ +
                      (size_t)bc[2] * icz + k];
  }
  halo_and_swap(&r, threads);

  double *maxs = (double *)malloc(sizeof(double) * (size_t)r.nblk);
  double current = 0.0;
  int steps = 0;
  while (current < t_end && steps < max_steps) {
    stage_arg_t sa = {0, 0.0, maxs};
    parallel_blocks(&r, threads, block_signal, &sa);
    double s = 0.0;
    for (int b = 0; b < r.nblk; ++b)
      s = smax(s, maxs[b]);
    /* dt = CFL / (S / dx), end-time clamp (modular_algorithm_assembler.cpp:
     * 1932-1982). */
    double const rate = s / g->cell_size;
    double dt = rate == 0.0 ? DBL_MAX : cfl / rate;
    double const run_time = current;
    if (run_time + dt > t_end)
      dt = smax(t_end - run_time, 0.0);
    if (dts)
      dts[steps] = dt;
    for (int stage = 0; stage < 3; ++stage) {
      stage_arg_t st = {stage, dt, maxs};
      parallel_blocks(&r, threads, block_stage, &st);
      halo_and_swap(&r, threads);
    }
    current += 0.0 + dt; /* FinishMacroTimestep (time_integrator.h:244-248) */
    ++steps;
    if (dt < DBL_EPSILON && dt > 0.0)
      break;
  }
  for (int b = 0; b < r.nblk; ++b) {
    int bc[3];
    block_coords(&r, b, bc);
    const double *a = r.avg + (size_t)b * r.bsize;
    for (int e = 0; e < ne; ++e)
      for (int i = 0; i < IC; ++i)
        for (int j = 0; j < IC; ++j)
          for (int k = 0; k < icz; ++k)
            state[(((size_t)e * NX + (size_t)bc[0] * IC + i) * NY +
                   (size_t)bc[1] * IC + j) *
                      NZ +
                  (size_t)bc[2] * icz + k] =
                a[e * r.L.field + (i + HS) * r.L.stride[0] +
                  (j + HS) * r.L.stride[1] + (dim == 3 ? k + HS : 0)];
  }
  if (t_final)
    *t_final = current;
  free(maxs);
  free(r.avg);
  free(r.rhs);
  free(r.init);
  return steps;
}

void ora_domain_sums(int dim, const int *ncell, const double *state,
                     double *out) {
  size_t const n = (size_t)ncell[0] * (size_t)ncell[1] *
                   (size_t)(dim == 3 ? ncell[2] : 1);
  for (int e = 0; e < dim + 2; ++e) {
    double s = 0.0;
    for (size_t x = 0; x < n; ++x)
      s += state[(size_t)e * n + x];
    out[e] = s;
  }
}

/* ------------------------------------------------------------------------ */
/* Multiresolution operators between a parent block and one of its children.  */
/* Fields are [TC][TC][TCZ] (TCZ = 1 in 2D); blocks are [NE] such fields.      */
/* A child's position among its siblings is id & 7: bit 0 x, bit 1 y, bit 2 z  */
/* (id_information.h:158-160).                                                 */

/* ConsistencyManagedSum of four values, full symmetry
 * (mathematical_functions.h:115-124): the mean of the smallest and largest of
 * the three pairings, with std::min/std::max picking the first extreme. */
static double cms4(double a, double b, double c, double d) {
  double const s1 = (a + b) + (c + d);
  double const s2 = (a + c) + (b + d);
  double const s3 = (a + d) + (b + c);
  double mn = s1, mx = s1;
  mn = smin(mn, s2);
  mn = smin(mn, s3);
  mx = smax(mx, s2);
  mx = smax(mx, s3);
  return 0.5 * (mn + mx);
}

static inline size_t fidx(int dim, int i, int j, int k) {
  return ((size_t)i * TC + (size_t)j) * (size_t)(dim == 3 ? TC : 1) + (size_t)k;
}

/* Multiresolution::Average (multiresolution.h:94-167): the conservative mean
 * of the 2^dim child cells under each parent cell, written into the quarter
 * (octant) of the parent's internal cells that the child covers: parent start
 * HS or TC/2 per axis (xyz_look_up_table_, multiresolution.h:35-45), IC/2
 * cells per axis (PSOCIC, compile_time_constants.h:578-590). */
void ora_mr_average(int dim, const double *child, double *parent, int position) {
  int const ne = dim + 2, tz = dim == 3 ? TC : 1;
  int const x0 = (position & 1) ? TC / 2 : HS;
  int const y0 = (position & 2) ? TC / 2 : HS;
  int const z0 = dim == 3 ? ((position & 4) ? TC / 2 : HS) : 0;
  int const nz = dim == 3 ? IC / 2 : 1;
  size_t const field = (size_t)TC * TC * tz;
  for (int e = 0; e < ne; ++e) {
    const double *c = child + (size_t)e * field;
    double *p = parent + (size_t)e * field;
    for (int i = 0; i < IC / 2; ++i)
      for (int j = 0; j < IC / 2; ++j)
        for (int k = 0; k < nz; ++k) {
          int const ci = HS + 2 * i, cj = HS + 2 * j, ck = dim == 3 ? HS + 2 * k : 0;
          double v;
          if (dim == 2) {
            v = ((c[fidx(dim, ci, cj, 0)] + c[fidx(dim, ci + 1, cj + 1, 0)]) +
                 (c[fidx(dim, ci + 1, cj, 0)] + c[fidx(dim, ci, cj + 1, 0)])) *
                0.25;
          } else {
            v = 0.125 *
                cms4(c[fidx(dim, ci, cj, ck)] + c[fidx(dim, ci + 1, cj + 1, ck + 1)],
                     c[fidx(dim, ci, cj, ck + 1)] + c[fidx(dim, ci + 1, cj + 1, ck)],
                     c[fidx(dim, ci, cj + 1, ck)] + c[fidx(dim, ci + 1, cj, ck + 1)],
                     c[fidx(dim, ci, cj + 1, ck + 1)] + c[fidx(dim, ci + 1, cj, ck)]);
          }
          p[fidx(dim, x0 + i, y0 + j, z0 + k)] = v;
        }
  }
}

/* Multiresolution::Prediction (multiresolution.cpp:146-292): Harten's
 * fifth-order interpolation of the parent field into the child cells
 * [start, start + 2*(count/2)) per axis (range = x_start, x_count, y_start,
 * y_count, z_start, z_count; starts even; the z pair is ignored in 2D).  The
 * parent cell over child cell 2m+start is m + start/2 + 2 (low child) or
 * m + start/2 + TC/2 - 2 (high child) (PIOLCH/PIOHCH,
 * compile_time_constants.h:598-608).  In 2D the z terms are 0.0 and enter the
 * sums as the reference's zero-initialised variables do. */
void ora_mr_prediction(int dim, const double *parent, double *child, int position,
                       const int *range) {
  static const double c0 = -22.0 / 128.0, c1 = 3.0 / 128.0;
  double const c00 = c0 * c0, c01 = c0 * c1, c10 = c01, c11 = c1 * c1;
  double const c000 = c00 * c0, c001 = c00 * c1, c010 = c001, c011 = c0 * c11;
  double const c100 = c001, c101 = c011, c110 = c011, c111 = c1 * c11;
  int const ox = (position & 1) ? TC / 2 - HS / 2 : HS / 2;
  int const oy = (position & 2) ? TC / 2 - HS / 2 : HS / 2;
  int const oz = (position & 4) ? TC / 2 - HS / 2 : HS / 2;
  int const px0 = range[0] / 2 + ox, px1 = px0 + range[1] / 2;
  int const py0 = range[2] / 2 + oy, py1 = py0 + range[3] / 2;
  int const pz0 = dim == 3 ? range[4] / 2 + oz : 0;
  int const pz1 = dim == 3 ? pz0 + range[5] / 2 : 1;
#define P(a, b, c) parent[fidx(dim, (a), (b), (c))]
  int ci = range[0];
  for (int i = px0; i < px1; ++i, ci += 2) {
    int cj = range[2];
    for (int j = py0; j < py1; ++j, cj += 2) {
      int ck = dim == 3 ? range[4] : 0;
      for (int k = pz0; k < pz1; ++k, ck += 2) {
        double Qz = 0.0, Qxz = 0.0, Qyz = 0.0, Qxyz = 0.0;
        double const Qx = c0 * (P(i + 1, j, k) - P(i - 1, j, k)) + c1 * (P(i + 2, j, k) - P(i - 2, j, k));
        double const Qy = c0 * (P(i, j + 1, k) - P(i, j - 1, k)) + c1 * (P(i, j + 2, k) - P(i, j - 2, k));
        double const Qxy =
            (c00 * ((P(i + 1, j + 1, k) + P(i - 1, j - 1, k)) - (P(i - 1, j + 1, k) + P(i + 1, j - 1, k))) +
             c11 * ((P(i + 2, j + 2, k) + P(i - 2, j - 2, k)) - (P(i - 2, j + 2, k) + P(i + 2, j - 2, k)))) +
            (c01 * ((P(i + 1, j + 2, k) + P(i - 1, j - 2, k)) - (P(i - 1, j + 2, k) + P(i + 1, j - 2, k))) +
             c10 * ((P(i + 2, j + 1, k) + P(i - 2, j - 1, k)) - (P(i - 2, j + 1, k) + P(i + 2, j - 1, k))));
        if (dim == 3) {
          Qz = c0 * (P(i, j, k + 1) - P(i, j, k - 1)) + c1 * (P(i, j, k + 2) - P(i, j, k - 2));
          Qxz = (c00 * ((P(i + 1, j, k + 1) + P(i - 1, j, k - 1)) - (P(i - 1, j, k + 1) + P(i + 1, j, k - 1))) +
                 c11 * ((P(i + 2, j, k + 2) + P(i - 2, j, k - 2)) - (P(i - 2, j, k + 2) + P(i + 2, j, k - 2)))) +
                (c01 * ((P(i + 1, j, k + 2) + P(i - 1, j, k - 2)) - (P(i - 1, j, k + 2) + P(i + 1, j, k - 2))) +
                 c10 * ((P(i + 2, j, k + 1) + P(i - 2, j, k - 1)) - (P(i - 2, j, k + 1) + P(i + 2, j, k - 1))));
          Qyz = (c00 * ((P(i, j + 1, k + 1) + P(i, j - 1, k - 1)) - (P(i, j - 1, k + 1) + P(i, j + 1, k - 1))) +
                 c11 * ((P(i, j + 2, k + 2) + P(i, j - 2, k - 2)) - (P(i, j - 2, k + 2) + P(i, j + 2, k - 2)))) +
                (c01 * ((P(i, j + 1, k + 2) + P(i, j - 1, k - 2)) - (P(i, j - 1, k + 2) + P(i, j + 1, k - 2))) +
                 c10 * ((P(i, j + 2, k + 1) + P(i, j - 2, k - 1)) - (P(i, j - 2, k + 1) + P(i, j + 2, k - 1))));
          /* One corner group: the four antisymmetric differences of the
           * stencil points (i+-a, j+-b, k+-c). */
#define G(a, b, c)                                                                  \
  cms4(P(i + a, j + b, k + c) - P(i - a, j - b, k - c),                             \
       P(i - a, j - b, k + c) - P(i + a, j + b, k - c),                             \
       P(i + a, j - b, k - c) - P(i - a, j + b, k + c),                             \
       P(i - a, j + b, k - c) - P(i + a, j - b, k + c))
          Qxyz = cms4(c000 * G(1, 1, 1) + c111 * G(2, 2, 2), c001 * G(1, 1, 2) + c110 * G(2, 2, 1),
                      c011 * G(1, 2, 2) + c100 * G(2, 1, 1), c010 * G(1, 2, 1) + c101 * G(2, 1, 2));
#undef G
        }
        double const p = P(i, j, k);
        child[fidx(dim, ci, cj, ck)] = p + dams(dim, Qx, Qy, Qz) + dams(dim, Qxy, Qxz, Qyz) + Qxyz;
        child[fidx(dim, ci + 1, cj, ck)] = p + dams(dim, -Qx, Qy, Qz) + dams(dim, -Qxy, -Qxz, Qyz) - Qxyz;
        child[fidx(dim, ci, cj + 1, ck)] = p + dams(dim, Qx, -Qy, Qz) + dams(dim, -Qxy, Qxz, -Qyz) - Qxyz;
        child[fidx(dim, ci + 1, cj + 1, ck)] = p + dams(dim, -Qx, -Qy, Qz) + dams(dim, Qxy, -Qxz, -Qyz) + Qxyz;
        if (dim == 3) {
          child[fidx(dim, ci, cj, ck + 1)] = p + cms3(Qx, Qy, -Qz) + cms3(Qxy, -Qxz, -Qyz) - Qxyz;
          child[fidx(dim, ci + 1, cj, ck + 1)] = p + cms3(-Qx, Qy, -Qz) + cms3(-Qxy, Qxz, -Qyz) + Qxyz;
          child[fidx(dim, ci, cj + 1, ck + 1)] = p + cms3(Qx, -Qy, -Qz) + cms3(-Qxy, -Qxz, Qyz) + Qxyz;
          child[fidx(dim, ci + 1, cj + 1, ck + 1)] = p + cms3(-Qx, -Qy, -Qz) + cms3(Qxy, Qxz, Qyz) - Qxyz;
        }
      }
    }
  }
#undef P
}

/* ChildNeedsRemeshing<N> (multiresolution.cpp:299-398) without the decision:
 * over the wavelet equations -- Mass and Energy (MF::EWA(),
 * field_setup.h:243-247) -- and every cell of the child block (halo
 * included), the relative detail between the child's values and the
 * parent's prediction over the full block:
 *   norm 0, L-infinity (:299-323): max of |exact - predicted| / |exact|;
 *   norm 1, L-one (:329-361): per equation the sum over cells (i, j, k order)
 *           of |exact - predicted| / |exact|, times 1/(TC^2 TCZ);
 *   norm 2, L-two (:367-398): per equation sqrt(sum of ((exact - predicted) /
 *           exact)^2) times 1/(TC^2 TCZ);
 * the result is the largest over the equations, by std::max(max_detail, x),
 * which keeps max_detail when x is NaN. */
double ora_mr_detail(int dim, const double *parent, const double *child, int position, int norm) {
  int const tz = dim == 3 ? TC : 1;
  size_t const field = (size_t)TC * TC * tz;
  int const full[6] = {0, TC, 0, TC, 0, tz};
  double const one_number_of_cells = 1.0 / (double)(TC * TC * tz);
  double *pred = (double *)malloc(field * sizeof(double));
  double max_detail = 0.0;
  for (int e = MASS; e <= ENERGY; ++e) {
    ora_mr_prediction(dim, parent + (size_t)e * field, pred, position, full);
    const double *x = child + (size_t)e * field;
    if (norm == 0) {
      for (size_t c = 0; c < field; ++c)
        max_detail = smax(max_detail, fabs(x[c] - pred[c]) / fabs(x[c]));
    } else if (norm == 1) {
      double error_norm = 0.0;
      for (size_t c = 0; c < field; ++c)
        error_norm += fabs(x[c] - pred[c]) / fabs(x[c]);
      max_detail = smax(max_detail, error_norm * one_number_of_cells);
    } else {
      double error_norm = 0.0;
      for (size_t c = 0; c < field; ++c) {
        double const temp = (x[c] - pred[c]) / x[c];
        error_norm = error_norm + temp * temp;
      }
      max_detail = smax(max_detail, sqrt(error_norm) * one_number_of_cells);
    }
  }
  free(pred);
  return max_detail;
}

double ora_mr_detail_linf(int dim, const double *parent, const double *child, int position) {
  return ora_mr_detail(dim, parent, child, position, 0);
}

/* ThresholdComputer<DIM, STDO = 3> (threshold_computer.h:53-59, 91-94): the
 * coarsening threshold on a level. */
double ora_mr_threshold(int dim, int maximum_level, int reference_level,
                        double reference_epsilon, int level) {
  double const alpha = 3.0, d = (double)dim;
  double const eps_ref =
      reference_epsilon * pow(2.0, -1.0 * (alpha + 1.0) * ((double)maximum_level - (double)reference_level));
  return eps_ref * pow(2.0, -1.0 * d * ((double)maximum_level - (double)level));
}

/* Multiresolution::RemeshingDecision (multiresolution.cpp:408-428):
 * RemeshIdentifier Neutral 0, Refine 1, Coarse 2 (remesh_identifier.h:29). */
int ora_mr_decision(double detail, double eps_coarsen) {
  double const eps_refine = eps_coarsen * 32.0;
  if (detail <= eps_coarsen)
    return 2;
  if (detail >= eps_refine)
    return 1;
  return 0;
}

/* Multiresolution::AverageJumpBuffer (multiresolution.cpp:47-131): a block's
 * jump surfaces are [6][NE][IC][ICZ] doubles (SurfaceBuffer, block.h:34-41:
 * East, West, North, South, Top, Bottom; ICZ = 1 in 2D, where only the first
 * four are used).  For every face of the parent that the child touches (East
 * when the child's x bit is set, West when it is clear, and so on), the
 * child's face values averaged over 2x2 (3D) or 2 (2D) cells are ADDED to the
 * quarter (half) of the parent's face the child covers, which starts at
 * IC/2 times the child's bits of the face's two transverse axes in
 * increasing axis order (first_/second_index_look_up_table). */
void ora_mr_average_jump(int dim, const double *child, double *parent, int position) {
  int const ne = dim + 2, icz = dim == 3 ? IC : 1;
  size_t const face = (size_t)ne * IC * icz;
  for (int loc = 0; loc < 2 * dim; ++loc) {
    int const axis = loc / 2, high = (loc % 2) == 0; /* East, North, Top are high */
    if (((position >> axis) & 1) != high)
      continue;
    int const t1 = axis == 0 ? 1 : 0, t2 = axis == 2 ? 1 : 2;
    int const s1 = ((position >> t1) & 1) * (IC / 2);
    int const s2 = dim == 3 ? ((position >> t2) & 1) * (IC / 2) : 0;
    const double *c = child + (size_t)loc * face;
    double *p = parent + (size_t)loc * face;
    for (int e = 0; e < ne; ++e)
      for (int i = 0; i < IC / 2; ++i)
        for (int j = 0; j < (dim == 3 ? IC / 2 : 1); ++j) {
          const double *q = c + ((size_t)e * IC + 2 * i) * icz + (dim == 3 ? 2 * j : 0);
          double v;
          if (dim == 2)
            v = (q[0] + q[icz]) * 0.5;
          else
            v = ((q[0] + q[icz + 1]) + (q[icz] + q[1])) * 0.25;
          p[((size_t)e * IC + s1 + i) * icz + s2 + j] += v;
        }
  }
}

/* ------------------------------------------------------------------------ */
/* Level-wise (local time stepping) integration, TimeIntegrator with          */
/* RungeKutta3TVD (time_integrator.h, runge_kutta_3_TVD.h:39-59).             */

/* The time step of a level's stage: the last n micro time steps summed from
 * the newest, starting at 0.0 (std::accumulate over crbegin..crbegin+n),
 * times the stage multiplier -- of the conservatives (1, 1/4, 2/3) or, with
 * jump != 0, of the jump conservatives (1/6, 1/6, 2/3)
 * (time_integrator.h:281-301, 312-340). */
double ora_lts_timestep(const double *micro, int count, int n, int stage, int jump) {
  static const double mc[3] = {1.0, 0.25, 2.0 / 3.0};
  static const double mj[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
  double s = 0.0;
  for (int i = 0; i < n; ++i)
    s = s + micro[count - 1 - i];
  return s * (jump ? mj[stage] : mc[stage]);
}

/* IntegrateConservatives / IntegrateHalo (time_integrator.h:98-153):
 * u_new = u_old + timestep * u_new over the cells of range (x_start,
 * x_count, y_start, y_count, z_start, z_count) of every equation; range NULL
 * means the internal cells. */
void ora_lts_integrate(int dim, const double *u_old, double *u_new, const int *range,
                       double timestep) {
  int const ne = dim + 2, tz = dim == 3 ? TC : 1;
  int const internal[6] = {HS, IC, HS, IC, dim == 3 ? HS : 0, dim == 3 ? IC : 1};
  const int *r = range ? range : internal;
  size_t const field = (size_t)TC * TC * tz;
  for (int e = 0; e < ne; ++e)
    for (int i = r[0]; i < r[0] + r[1]; ++i)
      for (int j = r[2]; j < r[2] + r[3]; ++j)
        for (int k = r[4]; k < r[4] + r[5]; ++k) {
          size_t const c = (size_t)e * field + fidx(dim, i, j, k);
          u_new[c] = u_old[c] + timestep * u_new[c];
        }
}

/* IntegrateJumpConservatives (time_integrator.h:64-84) on jump surfaces
 * [6][NE][IC][ICZ]: conservatives += timestep * fluxes, fluxes = 0, on the
 * 2*dim faces (CC::ANBS()). */
void ora_lts_integrate_jump(int dim, double *cons, double *flux, double timestep) {
  size_t const n = (size_t)2 * dim * (dim + 2) * IC * (dim == 3 ? IC : 1);
  for (size_t c = 0; c < n; ++c) {
    cons[c] += timestep * flux[c];
    flux[c] = 0.0;
  }
}

/* PrepareBufferForIntegration (time_integrator.h:434-457) for stage 1 or 2:
 * avg = m0 * avg + m1 * init over every cell, (m0, m1) = (1/4, 3/4) and
 * (2/3, 1/3) (runge_kutta_3_TVD.h:55-59); stage 0 leaves avg unchanged. */
void ora_lts_prepare_buffer(int dim, double *avg, const double *init, int stage) {
  static const double m[2][2] = {{0.25, 0.75}, {2.0 / 3.0, 1.0 / 3.0}};
  if (stage == 0)
    return;
  size_t const n = (size_t)(dim + 2) * TC * TC * (dim == 3 ? TC : 1);
  for (size_t c = 0; c < n; ++c)
    avg[c] = m[stage - 1][0] * avg[c] + m[stage - 1][1] * init[c];
}

/* The boundary-flux save of SpaceSolver::UpdateFluxes (space_solver.cpp:
 * 160-182): the face fluxes of the block's six boundary faces ADDED into its
 * jump-flux surfaces [6][NE][IC][ICZ] (East, West, North, South, Top,
 * Bottom): West/East from the x faces 0 and IC, South/North from the y faces,
 * Bottom/Top from the z faces, at transverse face index (i+1, j+1) of the
 * reference's face arrays [NE][IC+1][IC+1][ICZ+1].  In 2D the z face arrays
 * are zero, so Bottom and Top receive += 0.0. */
void ora_block_jump_fluxes(int dim, int flux, double gamma, double b, const double *avg,
                           double *jump) {
  int const ne = dim + 2, icz = dim == 3 ? IC : 1;
  size_t const n1 = IC + 1, nz = (size_t)icz + 1, per = n1 * n1 * nz;
  double *f = (double *)calloc(3 * (size_t)ne * per, sizeof(double));
  block_face_fluxes(dim, flux, gamma, b, avg, f, f + ne * per, f + 2 * ne * per);
  const double *fx = f, *fy = f + ne * per, *fz = f + 2 * ne * per;
  size_t const face = (size_t)ne * IC * icz;
#define FA(a, e, i, j, k) a[(((size_t)(e) * n1 + (i)) * n1 + (j)) * nz + (k)]
#define JB(loc, e, i, j) jump[(loc) * face + ((size_t)(e) * IC + (i)) * icz + (j)]
  for (int e = 0; e < ne; ++e)
    for (int i = 0; i < IC; ++i)
      for (int j = 0; j < icz; ++j) {
        JB(1, e, i, j) += FA(fx, e, 0, i + 1, j + 1);  /* West */
        JB(0, e, i, j) += FA(fx, e, IC, i + 1, j + 1); /* East */
        JB(3, e, i, j) += FA(fy, e, i + 1, 0, j + 1);  /* South */
        JB(2, e, i, j) += FA(fy, e, i + 1, IC, j + 1); /* North */
        JB(5, e, i, j) += FA(fz, e, i + 1, j + 1, 0);  /* Bottom */
        JB(4, e, i, j) += FA(fz, e, i + 1, j + 1, icz); /* Top */
      }
#undef FA
#undef JB
  free(f);
}

/* The exchange part of JumpFluxAdjustment (modular_algorithm_assembler.cpp:
 * 1410-1620) for one leaf: for every face location L (East, West, North,
 * South, Top, Bottom; 2*dim of them) whose neighbour exists and is not a leaf
 * (nb[L] non-NULL: that neighbour's jump surfaces, holding its children's
 * averaged jump conservatives), the boundary layer of internal cells at L
 * takes the coarse contribution c_L = J_leaf[L] * one_cell_size * dir_L and
 * the fine one f_L = J_nb[opposite L] * one_cell_size * dir_L (dir = -1 on
 * East/North/Top, +1 on West/South/Bottom; transverse indices in increasing
 * axis order), and every internal cell becomes
 *   (u - CMS(c_E + c_W, c_N + c_S, c_T + c_B)) + CMS(f_E + f_W, f_N + f_S, f_T + f_B)
 * with zero contributions elsewhere (so every internal cell of the leaf is
 * rewritten, -0 becoming +0).  The leaf's jump surface at each such L is then
 * overwritten with the neighbour's (the copy the reference makes before
 * reading the fine fluxes). */
void ora_lts_jump_flux_adjustment(int dim, double *cells, double *jump_leaf,
                                  const double *const *jump_nb, double one_cell_size) {
  int const ne = dim + 2, icz = dim == 3 ? IC : 1, tz = dim == 3 ? TC : 1;
  size_t const face = (size_t)ne * IC * icz, field = (size_t)TC * TC * tz;
  for (int e = 0; e < ne; ++e)
    for (int i = 0; i < IC; ++i)
      for (int j = 0; j < IC; ++j)
        for (int k = 0; k < icz; ++k) {
          double c[6] = {0, 0, 0, 0, 0, 0}, f[6] = {0, 0, 0, 0, 0, 0};
          int const x[3] = {i, j, k};
          for (int loc = 0; loc < 2 * dim; ++loc) {
            if (!jump_nb[loc])
              continue;
            int const axis = loc / 2, high = (loc % 2) == 0;
            if (x[axis] != (high ? IC - 1 : 0))
              continue;
            int const t1 = axis == 0 ? 1 : 0, t2 = axis == 2 ? 1 : 2;
            int const a = x[t1], b = dim == 3 ? x[t2] : 0;
            double const dir = high ? -1.0 : 1.0;
            int const opp = high ? loc + 1 : loc - 1;
            c[loc] = jump_leaf[loc * face + ((size_t)e * IC + a) * icz + b] * one_cell_size * dir;
            f[loc] = jump_nb[loc][opp * face + ((size_t)e * IC + a) * icz + b] * one_cell_size * dir;
          }
          double *u = cells + (size_t)e * field + fidx(dim, i + HS, j + HS, dim == 3 ? k + HS : 0);
          *u -= cms3(c[0] + c[1], c[2] + c[3], c[4] + c[5]);
          *u += cms3(f[0] + f[1], f[2] + f[3], f[4] + f[5]);
        }
  for (int loc = 0; loc < 2 * dim; ++loc) {
    if (!jump_nb[loc])
      continue;
    int const opp = (loc % 2) == 0 ? loc + 1 : loc - 1;
    memcpy(jump_leaf + loc * face, jump_nb[loc] + opp * face, face * sizeof(double));
  }
}
