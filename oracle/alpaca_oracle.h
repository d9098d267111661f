/*
 * ORACLE — test infrastructure only.  Plain-C restatement of ALPACA's
 * single-phase finite-volume right-hand side on uniform level-0 blocks, used
 * as the checker for the HIP path (tests/, __graft_entry__.smoke() and the
 * "port" CPU-baseline leg of bench.py).  The product never links this.
 *
 * Every function follows the reference file:line cited at its definition in
 * alpaca_oracle.c, with the same operation order (compiled with
 * -ffp-contract=off, so a*b+c is never fused).
 *
 * Storage: one haloed block is [NE][TC][TC][TCZ] doubles, k fastest, with
 * NE = dim + 2 equations in the order Mass, Energy, MomX, MomY, (MomZ),
 * TC = 24 and TCZ = 24 in 3D, 1 in 2D (field_buffer.h:39, field_setup.h:101).
 */
#ifndef ALPACA_ORACLE_H
#define ALPACA_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORA_IC = 16, ORA_HS = 4, ORA_TC = 24 };
enum { ORA_FLUX_HLLC = 0, ORA_FLUX_ROE = 1 };
enum { ORA_BC_ZERO_GRADIENT = 0, ORA_BC_PERIODIC = 1, ORA_BC_SYMMETRY = 2 };

typedef struct {
  int dim;         /* 2 or 3 */
  int nb[3];       /* blocks per direction (nb[2] = 1 in 2D) */
  int bc[3];       /* per direction, both sides: ORA_BC_* */
  int flux;        /* ORA_FLUX_HLLC or ORA_FLUX_ROE */
  double gamma;    /* stiffened-gas gamma */
  double b;        /* stiffened-gas background pressure (0: ideal gas) */
  double cell_size;
} ora_grid;

/* WENO5-JS on six values: out[0] upwind-left, out[1] upwind-right. */
void ora_weno5(const double *v, double *out);

/* Stiffened-gas EOS, same argument/result layout as ref_eos. */
void ora_eos(int dim, double gamma, double b, const double *cons,
             const double *prim, double *out);

/* Roe eigensystem at the face between cells t (left) and n (right) along dir;
 * cons_* are NE conservatives; out L[NE][NE], R[NE][NE], lambda[NE].
 * Returns 0 if the face is skipped (a density <= 0, all outputs zero). */
int ora_roe_eigensystem(int dim, int dir, double gamma, double b,
                        const double *cons_t, const double *cons_n, double *L,
                        double *R, double *lambda);

/* HLLC (Einfeldt) flux along dir for conservative face states. */
void ora_hllc(int dim, int dir, double gamma, double b, const double *cl,
              const double *cr, double *out);

/* Face fluxes of one haloed block: out = three [NE][17][17][17 or 2] arrays. */
void ora_block_fluxes(int dim, int flux, double gamma, double b,
                      const double *avg, double cell_size, double *out);

/* Right-hand side of one haloed block on its internal cells [NE][16][16][16|1]. */
void ora_block_rhs(int dim, int flux, double gamma, double b,
                   const double *avg, double cell_size, double *out);

/* Full uniform run, same contract as ref_run: state is [NE][NX][NY][NZ]
 * internal conservatives of the domain, overwritten with the final state. */
int ora_run(const ora_grid *g, double cfl, double t_end, int max_steps,
            int threads, double *state, double *dts, double *t_final);

/* ComputeTimestepSize's signal maximum over one haloed block's internal
 * cells: max over cells of sum_d(|u_d| + c) (inviscid terms). */
double ora_block_signal(int dim, double gamma, double b, const double *avg);

/* Conservation sums over a domain state [NE][NX][NY][NZ]: out[e] = sum. */
void ora_domain_sums(int dim, const int *ncell, const double *state,
                     double *out);

/* Multiresolution operators between a parent block and the child at sibling
 * position `position` (bit 0 x, bit 1 y, bit 2 z).  Blocks are haloed blocks
 * as above; ora_mr_prediction works on one field [TC][TC][TCZ]. */
void ora_mr_average(int dim, const double *child, double *parent, int position);
/* range = x_start, x_count, y_start, y_count, z_start, z_count (child cells). */
void ora_mr_prediction(int dim, const double *parent, double *child, int position,
                       const int *range);
/* norm: 0 L-infinity, 1 L-one, 2 L-two (the reference's Norm enum order). */
double ora_mr_detail(int dim, const double *parent, const double *child, int position, int norm);
double ora_mr_detail_linf(int dim, const double *parent, const double *child, int position);
double ora_mr_threshold(int dim, int maximum_level, int reference_level,
                        double reference_epsilon, int level);
/* 0 Neutral, 1 Refine, 2 Coarse. */
int ora_mr_decision(double detail, double eps_coarsen);
/* Jump surfaces [6][NE][16][16|1] of a child added into its parent's. */
void ora_mr_average_jump(int dim, const double *child, double *parent, int position);

/* Boundary face fluxes of one haloed block added into its jump-flux surfaces
 * [6][NE][16][16|1] (SpaceSolver::UpdateFluxes). */
void ora_block_jump_fluxes(int dim, int flux, double gamma, double b, const double *avg,
                           double *jump);

/* Level-wise integration (RungeKutta3TVD on a level with n micro steps). */
double ora_lts_timestep(const double *micro, int count, int n, int stage, int jump);
void ora_lts_integrate(int dim, const double *u_old, double *u_new, const int *range,
                       double timestep);
void ora_lts_integrate_jump(int dim, double *cons, double *flux, double timestep);
void ora_lts_prepare_buffer(int dim, double *avg, const double *init, int stage);
/* JumpFluxAdjustment's exchange for one leaf block; jump_nb[6]: per face
 * location the jump surfaces of the non-leaf neighbour there, or NULL. */
void ora_lts_jump_flux_adjustment(int dim, double *cells, double *jump_leaf,
                                  const double *const *jump_nb, double one_cell_size);

#ifdef __cplusplus
}
#endif
#endif
