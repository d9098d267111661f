/*
 * alpaca_gpu.h — C ABI of the MI355X (gfx950) implementation of ALPACA's
 * single-phase finite-volume right-hand side on uniform level-0 leaf blocks.
 *
 * Plain pointers and sizes only.  Every device pointer is caller-owned HBM;
 * every call is asynchronous on the given HIP stream (NULL = default stream)
 * unless stated otherwise, and returns an alpaca_status code (no C++
 * exceptions cross this boundary; the reference has no error returns on this
 * path, invalid faces are skipped, FP faults trap — here non-finite states
 * raise ALPACA_ERR_NONFINITE through the case record).
 *
 * Block storage follows the reference Block/FieldBuffer contract
 * (block.h:54-161, field_buffer.h:39): one haloed block is
 * [NE][24][24][24] doubles in 3D, [NE][24][24] in 2D (k fastest), NE = dim+2
 * equations in the order Mass, Energy, MomX, MomY, (MomZ)
 * (field_setup.h:101-131).  A case is nb[0]*nb[1]*nb[2] consecutive blocks,
 * block (bx,by,bz) at index (bx*nb[1] + by)*nb[2] + bz.
 *
 * Counts of blocks, links and cases above 2^24 are refused with
 * ALPACA_ERR_INVALID (every launch grid derived from them then fits 32 bits).
 */
#ifndef ALPACA_GPU_H
#define ALPACA_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ALPACA_GPU_ABI_VERSION 1

typedef enum {
  ALPACA_OK = 0,
  ALPACA_ERR_INVALID = 1,     /* bad argument (null pointer, size, stage...) */
  ALPACA_ERR_HIP = 2,         /* a HIP runtime call failed */
  ALPACA_ERR_NONFINITE = 3,   /* a case produced a non-finite state */
  ALPACA_ERR_UNSUPPORTED = 4, /* configuration not built */
  ALPACA_ERR_CAPACITY = 5     /* serving slots / memory budget exhausted */
} alpaca_status;

/* Convective flux family: ConvectiveTermSolvers::FiniteVolume with
 * RiemannSolvers::Hllc + SignalSpeed::Einfeldt, or
 * ConvectiveTermSolvers::FluxSplitting with FluxSplitting::Roe
 * (riemann_solver_settings.h:29-54). */
enum { ALPACA_FLUX_HLLC = 0, ALPACA_FLUX_ROE = 1 };

/* Arithmetic of the reconstruction and Riemann part:
 *   ALPACA_FP64      FP64, a*b+c never fused: reproduces the reference's x86
 *                    build (no FMA) bit for bit;
 *   ALPACA_FP64_FMA  FP64 with fused multiply-add contraction, Newton-Raphson
 *                    division and square root (within 1 ulp), WENO5 weights
 *                    over a common denominator and the eigenvectors' zero
 *                    entries left out of the sums (a NaN or infinity can then
 *                    stay out of a sum it would reach in the reference);
 *                    within 1e-10 of the exact form after 100 steps;
 *   ALPACA_MIXED     FP32 reconstruction and Riemann fluxes (hardware
 *                    reciprocal and square root), FP64 pressure recovery
 *                    p = (gamma-1)(E - |m|^2/(2 rho)), FP64 flux divergence,
 *                    RK accumulation and sums. */
enum { ALPACA_FP64 = 0, ALPACA_MIXED = 1, ALPACA_FP64_FMA = 2 };

/* Side condition per direction: zero-gradient (ZeroGradientBoundaryCondition,
 * zero_gradient_boundary_condition.h:45-84), periodic (the neighbour wraps,
 * id_periodic_information.cpp:325) or symmetry (SymmetryBoundaryCondition,
 * symmetry_boundary_condition.h:89-109: the mirrored internal cell, with the
 * face-normal momentum negated). */
enum { ALPACA_BC_ZERO_GRADIENT = 0, ALPACA_BC_PERIODIC = 1, ALPACA_BC_SYMMETRY = 2 };

typedef struct alpaca_case_desc {
  int dim;            /* 2 or 3 */
  int nb[3];          /* leaf blocks per direction; nb[2] = 1 in 2D */
  int bc[3];          /* ALPACA_BC_* per direction (both sides alike) */
  int flux;           /* ALPACA_FLUX_* */
  int precision;      /* ALPACA_FP64 / ALPACA_MIXED */
  double gamma;       /* stiffened-gas gamma (StiffenedGas, stiffened_gas.cpp:41) */
  double background_pressure; /* stiffened-gas B; 0 = ideal gas */
  double cell_size;   /* dx of the level-0 cells (Node::GetCellSize) */
} alpaca_case_desc;

/* Device record of a running case (one per case, caller-allocated device
 * memory of sizeof(alpaca_case_state)).  Written by alpaca_timestep and the
 * stage kernels; read back by the host only at case completion. */
typedef struct alpaca_case_state {
  double time;     /* current simulation time (TimeIntegrator::CurrentRunTime) */
  double t_end;    /* end time (limit_end_time clamp) */
  double cfl;      /* CFL number */
  double dt;       /* time step of the running macro step */
  uint64_t smax;   /* max over internal cells of sum_d(|u_d|+c), as IEEE bits */
  int32_t steps;   /* completed macro steps */
  int32_t flags;   /* bit 0: non-finite state seen; bit 1: dt unhealthy;
                      bit 2: time reached t_end (set by the end of a step) */
} alpaca_case_state;

const char *alpaca_status_string(int status);
int alpaca_abi_version(void);

/* Kernel classes for the timing hooks below. */
enum {
  ALPACA_KERNEL_FACE_FLUX = 0, /* reconstruction + Riemann / splitting fluxes */
  ALPACA_KERNEL_COMBINE = 1,   /* flux divergence + RK stage combination */
  ALPACA_KERNEL_HALO = 2,      /* same-level halo fill + boundary conditions */
  ALPACA_KERNEL_TIMESTEP = 3,  /* CFL reduction */
  ALPACA_KERNEL_STAGE = 4,     /* fused stage: faces + divergence + RK + halo */
  ALPACA_KERNEL_COUNT = 5
};

/* Measurement hooks (no counterpart in the reference, which times macro steps
 * with MPI_Wtime): when enabled, every launch of a kernel class is bracketed
 * by HIP events on its launch stream.  alpaca_profile_read synchronises on
 * the recorded events and returns the summed milliseconds, the number of
 * launches and the algorithmic work units launched (faces for FACE_FLUX,
 * cells for COMBINE, TIMESTEP and STAGE, halo cells for HALO); reset != 0
 * clears. */
int alpaca_profile_enable(int on);
int alpaca_profile_read(int kernel, int reset, double *total_ms, long long *launches,
                        double *units);

/* Doubles per haloed block (NE * 24^dim). */
size_t alpaca_block_doubles(int dim);

/* Contexts of the per-block entries (i) and (ii).  The reference keeps its
 * face-flux arrays on the stack of SpaceSolver::UpdateFluxes
 * (space_solver.cpp:50-116), one set per call; here those faces go to device
 * scratch, which a context owns together with the HIP stream its calls are
 * queued on.  Use one context per host thread or stream: calls on one context
 * are ordered by its stream and share its scratch, calls on different
 * contexts never share scratch.  Scratch grows on demand without a device
 * synchronisation (outgrown buffers are kept until alpaca_context_destroy,
 * which synchronises the context's stream and frees them).  The
 * stream-argument forms (alpaca_convective_fluxes, alpaca_right_hand_side,
 * alpaca_right_hand_side_jump_fluxes) use an implicit context per (device,
 * stream), kept until the process ends; prefer an explicit context for
 * streams that are created and destroyed during the run. */
typedef struct alpaca_context alpaca_context;
/* Binds a new context to the current device and to `stream` (NULL = the
 * default stream). */
int alpaca_context_create(void *stream, alpaca_context **out);
int alpaca_context_destroy(alpaca_context *ctx);
/* Grows the context's scratch to what nblocks blocks of dimension dim need,
 * so later calls up to that size allocate nothing. */
int alpaca_context_reserve(alpaca_context *ctx, int dim, int nblocks);

/* (i) Per-block convective fluxes.
 * Replaces ConvectiveTermSolver<D>::UpdateConvectiveFluxes
 * (convective_term_solver.h:67-81) for the selected family.
 * d_avg: nblocks haloed Average buffers (halos filled).  Prime states are
 * recovered from d_avg on the fly (the reference reads them from the block's
 * PrimeStates buffer, which ObtainPrimeStatesFromConservatives fills from the
 * same Average values).  d_fluxes: per block three arrays x, y, z of
 * [NE][17][17][17] (3D) or [NE][17][17][2] (2D) doubles, indexed as the
 * reference's face arrays; fluxes are ACCUMULATED (+=) like the reference. */
int alpaca_convective_fluxes(const alpaca_case_desc *desc, int nblocks,
                             const double *d_avg, double *d_fluxes,
                             void *stream);

/* (ii) Per-block right-hand side: SpaceSolver::UpdateFluxes
 * (space_solver.cpp:50-153) with source terms off.  Writes the internal cells
 * of the nblocks haloed blocks d_rhs (halo cells are set to zero). */
int alpaca_right_hand_side(const alpaca_case_desc *desc, int nblocks,
                           const double *d_avg, double *d_rhs, void *stream);

/* (ii) with the rest of SpaceSolver::UpdateFluxes: also ADDS (+=) the fluxes
 * of each block's six boundary faces into its jump-flux surfaces
 * (space_solver.cpp:160-182), d_jump_fluxes = nblocks SurfaceBuffers
 * ([6][NE][16][16] doubles in 3D, [6][NE][16][1] in 2D: East, West, North,
 * South, Top, Bottom; in 2D Top/Bottom receive += 0.0). */
int alpaca_right_hand_side_jump_fluxes(const alpaca_case_desc *desc, int nblocks,
                                       const double *d_avg, double *d_rhs,
                                       double *d_jump_fluxes, void *stream);

/* (i) and (ii) on an explicit context (its stream and scratch); the same
 * results, bit for bit, as the stream-argument forms above. */
int alpaca_ctx_convective_fluxes(alpaca_context *ctx, const alpaca_case_desc *desc,
                                 int nblocks, const double *d_avg, double *d_fluxes);
int alpaca_ctx_right_hand_side(alpaca_context *ctx, const alpaca_case_desc *desc,
                               int nblocks, const double *d_avg, double *d_rhs);
int alpaca_ctx_right_hand_side_jump_fluxes(alpaca_context *ctx,
                                           const alpaca_case_desc *desc, int nblocks,
                                           const double *d_avg, double *d_rhs,
                                           double *d_jump_fluxes);

/* (ii) with the boundary-flux save over the listed leaves of a node pool:
 * the loop of ModularAlgorithmAssembler::ComputeRightHandSide
 * (modular_algorithm_assembler.cpp:990-1006) calling SpaceSolver::UpdateFluxes
 * for each leaf of a level.  d_avg, d_rhs and d_jump_fluxes are pools of
 * `capacity` haloed blocks / SurfaceBuffers; the n distinct device entries of
 * d_list name the blocks, each treated as alpaca_ctx_right_hand_side_
 * jump_fluxes treats its blocks (the same bits, in every arithmetic form), in
 * one launch with the face fluxes in LDS: no copy of the pool and no
 * face-flux scratch (the context gives the stream).  Entries outside
 * [0, capacity) are skipped. */
int alpaca_ctx_right_hand_side_jump_fluxes_list(alpaca_context *ctx,
                                                const alpaca_case_desc *desc, int capacity,
                                                const int *d_list, int n, const double *d_avg,
                                                double *d_rhs, double *d_jump_fluxes);

/* SwapBuffers (modular_algorithm_assembler.cpp:1044-1059) of a pool: the
 * haloed blocks d_list[i] of d_a and d_b exchanged
 * (SwapConservativeBuffersForNode, buffer_operations_material.h:83-91); n
 * distinct device entries, those outside [0, capacity) skipped. */
int alpaca_swap_blocks_list(int dim, double *d_a, double *d_b, int capacity, const int *d_list,
                            int n, void *stream);

/* FillInitialBuffer (time_integrator.h:369-406) of a pool, single phase: the
 * haloed blocks d_list[i] of d_src copied into d_dst (the Average buffers
 * into the Initial ones at stage 0); entries as for alpaca_swap_blocks_list. */
int alpaca_copy_blocks_list(int dim, const double *d_src, double *d_dst, int capacity,
                            const int *d_list, int n, void *stream);

/* (ii) without allocation.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same results, bit for bit, as
 * alpaca_right_hand_side; the face fluxes go to a fixed 64 MiB scratch array
 * that the library holds on each GPU it runs on until the process ends, so
 * the call makes no hipMalloc/hipFree and no other allocating call (the
 * profiling hooks record nothing for it).  At most
 * ALPACA_RHS_NOALLOC_MAX_BLOCKS_3D / _2D blocks per call; larger calls return
 * ALPACA_ERR_CAPACITY and launch nothing.  Every call on one GPU shares that
 * GPU's array, which keeps the last call's face fluxes (each call writes all
 * the faces it reads): calls that overlap on one GPU (different streams, or
 * host threads without ordering) race. */
#define ALPACA_RHS_NOALLOC_SCRATCH_BYTES (64u << 20)
#define ALPACA_RHS_NOALLOC_MAX_BLOCKS_3D 128
#define ALPACA_RHS_NOALLOC_MAX_BLOCKS_2D 3855
int alpaca_right_hand_side_noalloc(const alpaca_case_desc *desc, int nblocks,
                                   const double *d_avg, double *d_rhs, void *stream);

/* Verification hook, no counterpart in the reference: the FP64 operations
 * as the face kernels evaluate them, elementwise over n values on `stream`.
 * Exact form: op 0: d_out = d_a / d_b (the division of physical
 * quantities); op 1: the WENO5 weight division d_a / d_b (its core for
 * divisors in [1e-12, 2^700)); op 2: d_out = sqrt(d_a) (d_b unused).
 * FP64-with-FMA form: op 3: d_a / d_b; op 4: sqrt(d_a) (d_b unused).  Tests
 * compare the results with the x86 reference build's correctly rounded
 * operations (exact form: the same bits; FMA form: within 1 ulp). */
int alpaca_exact_fp64_ops(int op, const double *d_a, const double *d_b, double *d_out, int n,
                          void *stream);

/* (iii) One SSP-RK3 stage on every block of a case: FillInitialBuffer,
 * UpdateFluxes, PrepareBufferForIntegration and IntegrateNode
 * (time_integrator.h:98-113, 312-457; runge_kutta_3_TVD.h:39-59):
 *   out = (m0*avg + m1*init) + (c_s*dt)*RHS(avg)  on internal cells, which
 * are also written into the face halo slabs that read them (the same-level
 * halo fill of the output, so no alpaca_fill_halos call is needed after it),
 * with (m0, m1, c_s) = (1, 0, 1), (1/4, 3/4, 1/4), (2/3, 1/3, 2/3).
 * Stage 0 copies d_avg into d_init when d_init is non-null and differs from
 * d_avg; later stages read d_init.  d_out must not alias d_avg; it may alias
 * d_init (each cell reads its own init value before writing).  The time step
 * is read from the case's device record (d_state->dt). */
int alpaca_rk3_stage(const alpaca_case_desc *desc, int stage,
                     const double *d_avg, double *d_init, double *d_out,
                     alpaca_case_state *d_state, void *stream);

/* (iv) Same-level halo fill of a case: each block's face halo slabs take the
 * neighbour's internal cells (periodic sides wrap), zero-gradient sides copy
 * the first/last internal layer (internal_halo_manager.cpp:439-466,
 * boundary_constants.h:212-253).  Edge and corner halo cells, which no
 * direction-split stencil reads, are not written. */
int alpaca_fill_halos(const alpaca_case_desc *desc, double *d_blocks,
                      void *stream);

/* (iv) over a neighbour table, for block topologies that are not a full
 * grid: the same-level halo update the reference runs over its tree
 * (HaloManager::MaterialHaloUpdate, halo_manager.cpp:59-133;
 * InternalHaloManager::UpdateNoJumpLocal, internal_halo_manager.cpp:439-466;
 * external conditions boundary_constants.h:103-253).  Every halo cell of the
 * nblocks haloed blocks in d_blocks -- faces, edges and corners -- takes the
 * internal cell of the neighbour covering it.  d_neighbors (device,
 * int32[nblocks][27] in 3D, [nblocks][9] in 2D) holds for each offset
 * (dx, dy, dz) in {-1, 0, 1}^DIM, at index ((dx+1)*3 + (dy+1))*3 + (dz+1)
 * (2D: (dx+1)*3 + (dy+1)), the neighbour's block index, or -1 where the
 * offset leaves the domain; the centre entry is ignored.  Periodic sides
 * appear as neighbours.  side_bc (host, int[2*dim]: -x, +x, -y, +y, -z, +z)
 * gives the condition of external sides: ALPACA_BC_ZERO_GRADIENT (the
 * first/last internal layer) or ALPACA_BC_SYMMETRY (the mirror, normal
 * momentum negated), applied over the full transverse extent; a halo cell on
 * a side marked ALPACA_BC_PERIODIC without a neighbour is left unchanged. */
int alpaca_fill_halos_table(int dim, double *d_blocks, int nblocks,
                            const int *d_neighbors, const int *side_bc,
                            void *stream);

/* (iv) for a multi-level block pool: the halo update of the n blocks in
 * d_list (device int32) on one level (InternalHaloManager::
 * MaterialHaloUpdateOnLevel, internal_halo_manager.cpp:53-97, followed by the
 * external conditions, halo_manager.cpp:89-123).  d_neighbors (device,
 * int32[n][27] in 3D, [n][9] in 2D, row i for block d_list[i], slots as for
 * alpaca_fill_halos_table): the neighbour's pool block, -1 where the offset
 * leaves the domain (an external side), or -2 where no block of this level
 * exists there (a level jump, InternalBoundaryType::JumpBoundaryLocal: the
 * region is predicted from the parent by alpaca_mr_jump_halo_update).
 * mode 0: internal copies and external conditions, leaving jump regions and
 * every external-slab cell whose source lies in one; mode 1 (after the jump
 * update): exactly those external-slab cells, from the block's own predicted
 * halo cells, as the reference's external conditions over the full
 * transverse extent read them.  List entries outside [0, nblocks) are
 * skipped. */
int alpaca_fill_halos_list(int dim, double *d_blocks, int nblocks, const int *d_list,
                           int n, const int *d_neighbors, const int *side_bc, int mode,
                           void *stream);

/* (v) CFL time step of a case: ComputeTimestepSize
 * (modular_algorithm_assembler.cpp:1777-1997) for inviscid flow.
 * S = max over internal cells of sum_d(|u_d| + c); dt = cfl / (S / dx),
 * clamped to t_end - time.  Writes d_state->dt (and smax). */
int alpaca_timestep(const alpaca_case_desc *desc, const double *d_avg,
                    alpaca_case_state *d_state, void *stream);

/* (v) over the n listed leaves of a pool of `capacity` haloed blocks (every
 * leaf of a multi-level tree: ComputeTimestepSize's loop over
 * tree_.Leaves(), modular_algorithm_assembler.cpp:1797-1869); desc gives the
 * dimension, material and finest cell size (its nb is checked as for any
 * case, not used).  Entries outside [0, capacity) are skipped. */
int alpaca_timestep_list(const alpaca_case_desc *desc, const double *d_avg, int capacity,
                         const int *d_list, int n, alpaca_case_state *d_state, void *stream);

/* TimeIntegrator::FinishMacroTimestep (time_integrator.h:244-248) on the
 * device record: time += dt, steps += 1.  alpaca_advance does this itself;
 * callers that drive the stages one by one call it after the last stage. */
int alpaca_finish_step(alpaca_case_state *d_state, void *stream);

/* Initialise a device case record (time, t_end, cfl; counters zeroed). */
int alpaca_case_state_init(alpaca_case_state *d_state, double time,
                           double t_end, double cfl, void *stream);

/* nsteps full macro steps of a case, as ModularAlgorithmAssembler::Advance
 * does on a uniform level (modular_algorithm_assembler.cpp:543-910): time
 * step, then three stages whose kernels also fill the face halo slabs of
 * their output.  d_u holds u^n (halos filled) and receives u^(n+1) in place;
 * d_s1 and d_s2 are scratch blocks of the same size: stage 0 writes u1 to
 * d_s1, stage 1 writes u2 to d_s2 (d_u is the initial buffer), stage 2 writes
 * u^(n+1) over d_u.  The CFL signal of each new state is reduced in the last
 * stage, so only the first of the nsteps runs a separate reduction.  As
 * ComputeLoop stops at the end time (modular_algorithm_assembler.cpp:
 * 137-140), a step whose record already reached t_end (flags bit 2) does
 * nothing: a run can be advanced in chunks with one read-back per chunk. */
int alpaca_advance(const alpaca_case_desc *desc, double *d_u, double *d_s1,
                   double *d_s2, alpaca_case_state *d_state, int nsteps,
                   void *stream);

/* ------------------------------------------------------------ ensembles --
 * Many independent uniform cases of one dimension resident in one HBM block
 * pool (no counterpart in the reference, which runs one simulation per MPI
 * job).  Each case keeps its own time step, end time and stop flags; one
 * alpaca_ensemble_step advances every running case by one macro step with
 * one persistent kernel per RK stage and per enabled class (flux family x
 * arithmetic; the classes of one stage are parallel branches of the step),
 * replayed from a captured hipGraph.  A case's results do not
 * depend on what else runs: per-case reductions only, fixed arithmetic order.
 * The ensemble owns its device memory (allocated at create).
 *
 * Class bit for (flux, precision): 1 << (3*flux + m), m = 0 for ALPACA_FP64,
 * 1 for ALPACA_FP64_FMA, 2 for ALPACA_MIXED.
 * Case record flags: bit 0 non-finite state, bit 1 unhealthy time step,
 * bit 2 t_end reached, bit 3 cancelled, bit 5 slot in use; a case with any of
 * bits 0-3 set no longer advances. */
typedef struct alpaca_ensemble alpaca_ensemble;

#define ALPACA_CLASS_BIT(flux, m) (1u << (3 * (flux) + (m)))
enum {
  ALPACA_CASE_NONFINITE = 1,
  ALPACA_CASE_UNHEALTHY = 2,
  ALPACA_CASE_DONE = 4,
  ALPACA_CASE_CANCELLED = 8,
  ALPACA_CASE_ACTIVE = 32
};

/* HBM bytes an ensemble of this capacity allocates. */
size_t alpaca_ensemble_bytes(int dim, int max_blocks, int max_cases);
int alpaca_ensemble_create(int dim, int max_blocks, int max_cases, unsigned class_mask,
                           void *stream, alpaca_ensemble **out);
/* A second ensemble, of either dimension, that admits its cases into the
 * block pool of `peer` (and of every ensemble sharing it): one HBM budget
 * for the 2D and the 3D cases of a serving loop.  The pool's unit is one 2D
 * haloed block; a 3D block is exactly 30 units and takes a 30-aligned range.
 * Each ensemble keeps its own stream, case slots, work lists and captured
 * step; the pool is freed with the last ensemble that uses it. */
int alpaca_ensemble_create_shared(alpaca_ensemble *peer, int dim, int max_cases,
                                  unsigned class_mask, void *stream,
                                  alpaca_ensemble **out);
int alpaca_ensemble_destroy(alpaca_ensemble *e);
/* Admits a case: copies its nb[0]*nb[1]*nb[2] haloed blocks (internal cells
 * are used) from d_blocks into the pool, fills its halos and computes its
 * first time step; it advances from the next alpaca_ensemble_step.
 * ALPACA_ERR_CAPACITY when no contiguous block range or case slot is free. */
int alpaca_ensemble_add_case(alpaca_ensemble *e, const alpaca_case_desc *desc,
                             const double *d_blocks, double t_end, double cfl,
                             int *case_id);
/* Releases a case's slot and blocks (read its blocks first). */
int alpaca_ensemble_remove_case(alpaca_ensemble *e, int case_id);
/* Stops a case at the next stage boundary (the ABORTFILE analogue,
 * input_output_manager.h:197-199). */
int alpaca_ensemble_cancel_case(alpaca_ensemble *e, int case_id);
/* Tile shape (3D: 0 = 8x8x8/256 threads, 1 = 16x8x8/512; 2D: 0 = 16x16/128,
 * 1 = 16x16/256, 2 = 16x8/128) and persistent workgroups per CU (0 = the
 * occupancy default) of the stage kernels. */
int alpaca_ensemble_set_tuning(alpaca_ensemble *e, int tile_config, int groups_per_cu);
/* Advances every running case by nsteps macro steps (asynchronous). */
int alpaca_ensemble_step(alpaca_ensemble *e, int nsteps);
/* Synchronous copy of the first n case records to host memory. */
int alpaca_ensemble_read_states(alpaca_ensemble *e, alpaca_case_state *host_out, int n);
/* Device pointer to a case's current state blocks and their count. */
int alpaca_ensemble_case_blocks(alpaca_ensemble *e, int case_id, double **d_blocks,
                                int *nblocks);
/* Stream-ordered copy of a case's current state blocks into d_dst (caller
 * memory of nblocks * alpaca_block_doubles(dim) doubles). */
int alpaca_ensemble_copy_case(alpaca_ensemble *e, int case_id, double *d_dst);
/* Free pool blocks, the largest free contiguous range and free case slots. */
int alpaca_ensemble_capacity(alpaca_ensemble *e, int *free_blocks, int *largest_range,
                             int *free_slots);

/* ------------------------------------------------------------------------
 * Multiresolution operators between parent and child blocks of one pool of
 * haloed blocks (the building blocks of ALPACA's multi-level mode).  A link
 * names a child block, its parent block (pool indices), the child's position
 * among its siblings (id & 7, id_information.h:158-160: bit 0 x, bit 1 y,
 * bit 2 z; 0..3 in 2D) and the child's level.  Links are in device memory;
 * a link with an index outside [0, nblocks), child == parent or a position
 * outside [0, 2^dim) is skipped.  Within one call no block may be both a
 * parent and a child (the reference walks the levels one at a time).
 * d_blocks must be 16-byte aligned (ALPACA_ERR_INVALID otherwise): the
 * kernels move pairs of adjacent cells as one 16-byte access. */
typedef struct {
  int child;
  int parent;
  int position;
  int level;
} alpaca_mr_link;

/* Multiresolution::Average (multiresolution.h:94-167; used by
 * Averager::AverageMaterial, averager.cpp:62-140): for every link and every
 * equation, the parent internal cells the child covers -- IC/2 per axis from
 * HS or TC/2 -- become the conservative mean of the 2^dim child cells under
 * each.  Children of one parent may share a call. */
int alpaca_mr_average(int dim, double *d_blocks, int nblocks,
                      const alpaca_mr_link *d_links, int nlinks, void *stream);

/* Multiresolution::AverageJumpBuffer (multiresolution.cpp:47-131; used by
 * JumpFluxAdjustment, modular_algorithm_assembler.cpp:1347-1408): d_surfaces
 * holds nblocks jump-surface buffers laid out as the reference's
 * SurfaceBuffer (block.h:34-41), [6][NE][16][16] doubles in 3D and
 * [6][NE][16][1] in 2D (East, West, North, South, Top, Bottom; 2D uses the
 * first four).  For every link, on each face of the parent the child touches,
 * the child's values averaged over 2x2 (2D: 2) entries are ADDED (+=) to the
 * quarter (half) of the parent's face the child covers.  Children of one
 * parent may share a call (they cover disjoint entries). */
int alpaca_mr_average_jump_buffers(int dim, double *d_surfaces, int nblocks,
                                   const alpaca_mr_link *d_links, int nlinks,
                                   void *stream);

/* Multiresolution::Prediction (multiresolution.cpp:146-292) of every
 * equation field, for refinement (RefineNode, modular_algorithm_assembler.cpp:
 * 2547-2553: range = HS, IC per axis) and for halos across a level jump
 * (InternalHaloManager::UpdateMaterialJumpNoMpi, internal_halo_manager.cpp:
 * 231-284: the halo region).  range (host, int[2*dim]: x_start, x_count,
 * y_start, y_count, z_start, z_count) selects the child cells written:
 * [start, start + 2*(count/2)) per axis; starts must be even and
 * start + count <= 24 (ALPACA_ERR_INVALID otherwise). */
int alpaca_mr_prediction(int dim, double *d_blocks, int nblocks,
                         const alpaca_mr_link *d_links, int nlinks,
                         const int *range, void *stream);

/* UpdateMaterialJumpNoMpi (internal_halo_manager.cpp:231-262), the halo
 * update across a level jump, batched: for every link i the child's halo
 * slab at boundary location d_locations[i] (device int, the reference's
 * BoundaryLocation order, boundary_specifications.h:56-86: 0 East, 1 West,
 * 2 North, 3 South, 4 Top, 5 Bottom, 6-17 the sticks BottomNorth ...
 * SouthWest, 18-25 the cubes EastNorthTop ... WestSouthBottom) is predicted
 * from the parent on every equation, over GetStartIndicesHaloRecv(loc) and
 * GetHaloSize(loc) (communication_types.h:76-132,190-197): per axis the
 * cells [20, 24) for East/North/Top, [0, 4) for West/South/Bottom and the
 * internal [4, 20) otherwise.  The reference calls it on the right-hand-side
 * buffers for conservatives; pass the pool that holds them.  Links with an
 * invalid block index or location (in 2D: one with a Top/Bottom part) are
 * skipped. */
int alpaca_mr_jump_halo_update(int dim, double *d_blocks, int nblocks,
                               const alpaca_mr_link *d_links, const int *d_locations, int nlinks,
                               void *stream);

/* Norm of the wavelet details (the reference's Norm, norms.h; its default
 * CC::NFWA() is L-infinity, compile_time_constants.h:133). */
enum { ALPACA_MR_NORM_LINF = 0, ALPACA_MR_NORM_L1 = 1, ALPACA_MR_NORM_L2 = 2 };

/* ChildNeedsRemeshing<N> (multiresolution.cpp:299-398, called by
 * DetermineRemeshingNodes, modular_algorithm_assembler.cpp:2425-2480) on the
 * blocks given (the reference passes the right-hand-side buffers):
 * d_details[n] (device) over Mass and Energy (field_setup.h:243-247) and
 * every child cell, halo included, with q = (exact - predicted) / exact:
 * L-infinity the largest |q|; L-one per equation sum |q| / (TC^2 TCZ);
 * L-two per equation sqrt(sum q^2) / (TC^2 TCZ); the largest over the
 * equations, NaN values dropped (std::max), NaN for a skipped link.  The
 * L-one/L-two sums are tree reductions: equal to the reference's sequential
 * sums to rounding.  With d_decisions (device, may be NULL) also
 * RemeshingDecision (multiresolution.cpp:408-428):
 * 2 Coarse (detail <= eps), 1 Refine (detail >= 32 eps), 0 Neutral, -1 for a
 * skipped link or a level outside the table, where eps = eps_coarsen[level]
 * (host, nlevels <= 32 entries, ThresholdComputer::ThresholdOnLevel,
 * threshold_computer.h:91-94). */
int alpaca_mr_child_needs_remeshing(int dim, const double *d_blocks, int nblocks,
                                    const alpaca_mr_link *d_links, int nlinks, int norm,
                                    const double *eps_coarsen, int nlevels,
                                    double *d_details, int *d_decisions,
                                    void *stream);

/* ------------------------------------------------------------------------
 * Level-wise (local time stepping) integration: TimeIntegrator<RungeKutta3TVD>
 * operations on the leaves and jump halos of one level (Integrate,
 * modular_algorithm_assembler.cpp:1068-1098).  d_list (device, n int32) names
 * the pool blocks to treat; entries outside [0, nblocks) are skipped. */

/* The time step of a level's stage (host arithmetic): the last n of `count`
 * micro time steps summed from the newest, starting at 0.0 (std::accumulate
 * over crbegin..crbegin + n), times the stage multiplier of the
 * conservatives (1, 1/4, 2/3) or, with jump != 0, of the jump conservatives
 * (1/6, 1/6, 2/3) (time_integrator.h:281-301, 312-340; runge_kutta_3_TVD.h:
 * 42-53). */
int alpaca_lts_timestep(const double *micro, int count, int n, int stage, int jump,
                        double *out);

/* IntegrateConservatives / IntegrateHalo (time_integrator.h:98-153):
 * d_new = d_old + timestep * d_new on every equation of the listed blocks,
 * over range (host, int[6]: x_start, x_count, y_start, y_count, z_start,
 * z_count; the z pair is ignored in 2D) or, with range NULL, the internal
 * cells.  d_old may equal d_new. */
int alpaca_lts_integrate(int dim, const double *d_old, double *d_new, int nblocks,
                         const int *d_list, int n, const int *range, double timestep,
                         void *stream);

/* IntegrateJumpHalos (time_integrator.h:281-301) over the jump halos of a
 * level, the loop over (node, location) pairs of Integrate
 * (modular_algorithm_assembler.cpp:1084-1096), in one call: d_entries
 * (device, int32[n][2]) = {pool block, BoundaryLocation (East = 0 ...
 * WestSouthBottom = 25)}; the location's halo slab (GetStartIndicesHaloRecv /
 * GetHaloSize, communication_types.h:76-132,190-197) takes d_new = d_old +
 * timestep * d_new.  The slabs of different locations are disjoint.  Entries
 * with a block outside [0, nblocks), a location outside the enum or, in 2D,
 * one with a Top/Bottom part are skipped. */
int alpaca_lts_integrate_jump_halos(int dim, const double *d_old, double *d_new, int nblocks,
                                    const int *d_entries, int n, double timestep, void *stream);

/* IntegrateJumpConservatives (time_integrator.h:64-84) on jump surfaces in
 * SurfaceBuffer layout ([nblocks][6][NE][16][16|1]): d_cons += timestep *
 * d_flux, then d_flux = 0, on the faces the dimension has (4 in 2D, 6 in
 * 3D). */
int alpaca_lts_integrate_jump_conservatives(int dim, double *d_cons, double *d_flux,
                                            int nblocks, const int *d_list, int n,
                                            double timestep, void *stream);

/* PrepareBufferForIntegration (time_integrator.h:434-457): for stage 1 or 2,
 * d_avg = m0 * d_avg + m1 * d_init on every cell of the listed blocks,
 * (m0, m1) = (1/4, 3/4), (2/3, 1/3); stage 0 does nothing. */
int alpaca_lts_prepare_buffer(int dim, double *d_avg, const double *d_init, int nblocks,
                              const int *d_list, int n, int stage, void *stream);

/* The exchange of JumpFluxAdjustment (modular_algorithm_assembler.cpp:
 * 1410-1620) on the leaves of one level (one_cell_size = 1 / their cell
 * size).  d_entries (device, int32[n][7]): the leaf's pool block, then per
 * face location East, West, North, South, Top, Bottom the pool block of the
 * neighbour there when it exists and is not a leaf (its jump conservatives
 * hold its children's averages), else -1 (2D: the last two are ignored).
 * Every internal cell of each leaf becomes
 *   (u - CMS(c_E + c_W, c_N + c_S, c_T + c_B)) + CMS(f_E + f_W, f_N + f_S, f_T + f_B)
 * with c_L = J_leaf[L] * one_cell_size * dir_L, f_L = J_neighbour[opposite L]
 * * one_cell_size * dir_L on the boundary layer of a listed face (dir = -1 on
 * East/North/Top, +1 on West/South/Bottom) and 0 elsewhere; then the leaf's
 * jump conservatives at each listed face take the neighbour's.  d_cells is
 * the pool the reference corrects (its right-hand-side buffers); d_jump_cons
 * the jump-conservative surfaces ([nblocks][6][NE][16][16|1]).  Each leaf at
 * most once per call; no listed neighbour may be a listed leaf. */
int alpaca_lts_jump_flux_adjustment(int dim, double *d_cells, double *d_jump_cons, int nblocks,
                                    const int *d_entries, int n, double one_cell_size,
                                    void *stream);

/* ResetJumpConservativeBuffers (modular_algorithm_assembler.cpp:2021-2033):
 * the jump conservatives of the listed blocks (faces of CC::ANBS()) set to 0. */
int alpaca_lts_reset_jump_conservatives(int dim, double *d_jump_cons, int nblocks,
                                        const int *d_list, int n, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* ALPACA_GPU_H */
