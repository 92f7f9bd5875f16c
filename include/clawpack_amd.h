/*
 * clawpack_amd: C ABI of the MI355X (gfx950) 2D unsplit wave-propagation step.
 *
 * The entry points replace the native boundary of PyClaw's 2D classic solver:
 *
 *   f2py  classic2.step2(maxm, num_ghost, mx, my, qold, qnew, aux, dx, dy, dt, method, mthlim,
 *                        ..., use_fwave, rpn2, rpt2) -> (qnew, cfl)
 *         called from ClawSolver2D.step_hyperbolic
 *         (upstream pyclaw/src/pyclaw/classic/solver.py and classic/step2.f90; the pyclaw
 *          submodule is empty in the reference tree, .gitmodules:1-3; the f2py build rule is
 *          meson.build:29-35, FP64 via .f2py_f2cmap:1)
 *
 * plus the evolve_to_time accept/reject/dt loop that pyclaw runs in Python around it
 * (upstream pyclaw/src/pyclaw/solver.py), which here runs on the device.
 *
 * All pointers named *_dev are device pointers (HIP device memory).  `stream` is a hipStream_t
 * (NULL = default stream).  Every function returns 0 on success, a negative value for an invalid
 * argument and a positive HIP error code for a runtime failure; claw2d_last_error() describes the
 * last failure of the calling thread.  No function allocates device memory: the caller owns every
 * buffer (state planes, control block, workspaces), except the one fixed 16 KiB array that
 * claw2d_mass_sum_noalloc uses in place of a workspace.
 *
 * State layout on the device: structure of arrays, num_eqn planes of my rows of mx cells, x
 * fastest (plane m, row j, column i at index m*mx*my + j*mx + i), element type double (FP64) or
 * float (FP32).  Interior cells only; boundary conditions are applied by index mapping inside the
 * kernels (claw2d_step_padded instead reads a caller-filled ghost layer).  The transpose from
 * pyclaw's q(num_eqn, mx, my) happens once at upload.
 */
#ifndef CLAWPACK_AMD_H
#define CLAWPACK_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Riemann solvers: riemann.shallow_roe_with_efix_2D, riemann.euler_4wave_2D,
 * riemann.shallow_bathymetry_fwave_2D (f-waves, aux(1) = bathymetry) */
enum { CLAW_SOLVER_SHALLOW_ROE_EFIX = 0, CLAW_SOLVER_EULER_4WAVE = 1,
       CLAW_SOLVER_SHALLOW_BATHY_FWAVE = 2 };
enum { CLAW_FP64 = 0, CLAW_FP32 = 1 };
/* pyclaw BC codes: BC.custom=0, extrap=1, periodic=2, wall=3 (custom: the padded entry points only) */
enum { CLAW_BC_CUSTOM = 0, CLAW_BC_EXTRAP = 1, CLAW_BC_PERIODIC = 2, CLAW_BC_WALL = 3 };
/* claw2d_ctrl.status bits */
enum { CLAW_STATUS_BAD_STATE = 1, CLAW_STATUS_CFL_FIXED_DT = 2, CLAW_STATUS_MAX_STEPS = 4 };
/* claw2d_ctrl.flags bits */
enum { CLAW_FLAG_RUN_TO_TEND = 1, CLAW_FLAG_DT_VARIABLE = 2 };

/* Problem and method description (the Solver/State attributes pyclaw hands to step2). */
typedef struct claw2d_desc {
  int32_t solver;           /* CLAW_SOLVER_* */
  int32_t dtype;            /* CLAW_FP64 or CLAW_FP32 */
  int32_t mx, my;           /* interior cells, >= 2 each */
  int32_t num_eqn;          /* 3 (shallow water) or 4 (Euler) */
  int32_t num_waves;        /* 3 or 4 */
  int32_t num_ghost;        /* must be 2 */
  int32_t order;            /* method(2): 1 or 2 */
  int32_t transverse_waves; /* method(3): 0, 1 or 2 */
  int32_t mthlim[4];        /* limiter id per wave: 0 none, 1 minmod, 2 superbee, 3 vanleer, 4 MC, 5 BW */
  int32_t bc_lower[2];      /* x, y */
  int32_t bc_upper[2];
  int32_t reflect_index[2]; /* 0-based component negated by a wall in x, y (pyclaw default 1, 2) */
  int32_t seg_h;            /* rows per marching segment (0 = library default, sized from the
                               grid, the device's SIMD count and the kernel's occupancy; results
                               do not depend on it) */
  int32_t dt_variable;      /* solver.dt_variable */
  double dx, dy;
  double grav;              /* problem_data['grav'] (shallow water) */
  double gamma;             /* problem_data['gamma'] (Euler) */
  double cfl_max, cfl_desired, dt_max;
  /* aux arrays (method(7) = num_aux) and f-waves (solver.fwave).  CLAW_SOLVER_SHALLOW_BATHY_FWAVE
   * requires num_aux >= 1 (aux plane 0 is the bathymetry b) and fwave = 1; the other solvers take
   * aux arrays only with a capacity function (mcapa below) and fwave = 0.  aux_bc_* are
   * solver.aux_bc_lower/upper (extrap, periodic or wall; a wall mirrors aux without a sign
   * change), applied to every aux plane. */
  int32_t num_aux;
  int32_t fwave;
  int32_t aux_bc_lower[2];
  int32_t aux_bc_upper[2];
  double dry_tolerance;     /* problem_data['dry_tolerance']: h <= dry_tolerance has zero velocity */
  /* method(6) = mcapa: the 1-based aux plane holding the capacity function kappa (pyclaw's
   * state.index_capa + 1), 0 = none.  With a capacity function every Riemann solver takes aux
   * arrays (num_aux >= mcapa), the step is the capacity form of step2/step2ds (each cell's update
   * divided by its kappa, dt/dx per cell in qadd, the CFL number and the transverse terms, the
   * average of the two cells' dt/(dx kappa) in the second-order correction), and kappa must be
   * positive in every cell and ghost cell.  Accepted by the _aux entry points; the padded, batched
   * and adjoint entry points reject it. */
  int32_t mcapa;
  int32_t reserved_capa;
  /* Reference depth of an FP32 shallow-water state (CLAW_SOLVER_SHALLOW_ROE_EFIX, CLAW_FP32):
   * its depth plane holds h - depth_ref, which the kernels add back wherever the depth itself
   * enters.  FP32 rounding of the stored depth is then relative to the deviation, not to h (the
   * dominant error of the FP32 path).  0 = the plane holds h; FP64 states and the other solvers
   * take only 0.  claw2d_mass_sum returns the sum of h (depth_ref times the cell count added). */
  double depth_ref;
  /* Launch shape of the single-grid step kernels (results do not depend on it): wavefronts per
   * workgroup, 1, 2 or 4 (0 = 4), and the order in which workgroups take strips and segments,
   * 0 = each XCD takes a contiguous range (neighbouring strips share halo lines in one L2),
   * 1 = plain order (workgroup b to the XCD b mod 8, as dispatched). */
  int32_t wg_waves;
  int32_t launch_order;
} claw2d_desc;

/* Device-resident time-loop control block (one per grid). */
typedef struct claw2d_ctrl {
  double t, dt, tend, tstart;
  double cfl, cflmax, dtmin, dtmax;
  double smax_last[2];             /* max |s| of the last attempt, x and y */
  unsigned long long smax_bits[2]; /* running max |s| (IEEE bits of a non-negative double) */
  long long numsteps, attempts, rejected, target_steps, max_steps;
  int32_t cur;                     /* which state buffer holds q^n (0 or 1) */
  int32_t done;
  int32_t status;                  /* CLAW_STATUS_* bits */
  int32_t flags;                   /* CLAW_FLAG_* bits */
  int32_t arrive;                  /* workgroups of the running step that have finished; 0 between calls */
  int32_t reserved;
} claw2d_ctrl;

/* Checks a descriptor; 0 if the HIP path supports it. */
int claw2d_validate(const claw2d_desc* d);

/* Bytes of one state buffer (num_eqn * mx * my * element size). */
size_t claw2d_state_bytes(const claw2d_desc* d);

/* Fills a host control block for a run starting at t0 with step dt0 (pyclaw solver.dt), either
 * to tend (run_to_tend = 1, Solver.evolve_to_time(solution, tend)) or for target_steps accepted
 * steps.  The dt clip to tend of evolve_to_time is applied to dt0 here. */
int claw2d_ctrl_init(const claw2d_desc* d, claw2d_ctrl* host_ctrl, double t0, double dt0,
                     double tend, int32_t run_to_tend, long long target_steps, long long max_steps);

/* One step2 with a caller-chosen dt: q_in_dev -> q_out_dev (the f2py step2 call).  The CFL number
 * max(dt/dx*max|s_x|, dt/dy*max|s_y|) of the step is written to ctrl_dev->cfl, the speeds to
 * ctrl_dev->smax_last, and CLAW_STATUS_BAD_STATE is or-ed into ctrl_dev->status when a depth or
 * density is not positive (with bathymetry: a depth is negative; dry cells have depth 0) or NaN.
 * ctrl_dev->smax_bits must be zero on entry (claw2d_ctrl_init and every
 * call leave it zero).  Replaces classic2.step2 (upstream classic/step2.f90). */
int claw2d_step(const claw2d_desc* d, const void* q_in_dev, void* q_out_dev, double dt,
                claw2d_ctrl* ctrl_dev, void* stream);

/* Enqueues `attempts` step attempts of Solver.evolve_to_time (upstream pyclaw solver.py) on the
 * ping-pong pair (q0_dev, q1_dev): each attempt is one kernel launch that steps with ctrl->dt
 * from q[ctrl->cur] into the other buffer; its last wavefront to finish applies accept/reject
 * (cfl <= cfl_max), t += dt, cur ^= 1 on accept, and dt <- min(dt_max, dt*cfl_desired/cfl).
 * Once ctrl->done is set the remaining attempts are no-ops.  No host synchronisation.  Calls
 * that share one control block must not run concurrently. */
int claw2d_advance(const claw2d_desc* d, void* q0_dev, void* q1_dev, claw2d_ctrl* ctrl_dev,
                   long long attempts, void* stream);

/* Dimensionally split (Godunov, x then y) versions of claw2d_step / claw2d_advance: replace
 * classic2.step2ds (upstream classic/step2ds.f90), which ClawSolver2D.step_hyperbolic calls when
 * solver.dimensional_split is True (pyclaw's default in 2D).  The x-sweeps over all rows, a fresh
 * ghost fill of the intermediate state q* and the y-sweeps over all columns; no transverse
 * propagation (transverse_waves is ignored); cfl = max of the two sweeps' CFL numbers.  One kernel
 * per attempt: q* is formed on the fly (each marched row is x-swept before its y-interfaces), so it
 * never goes to memory. */
int claw2d_step_ds(const claw2d_desc* d, const void* q_in_dev, void* q_out_dev, double dt,
                   claw2d_ctrl* ctrl_dev, void* stream);
int claw2d_advance_ds(const claw2d_desc* d, void* q0_dev, void* q1_dev, claw2d_ctrl* ctrl_dev,
                      long long attempts, void* stream);

/* The four entry points above with aux arrays, for solvers with num_aux > 0 (the entry points
 * without aux reject them).  aux_dev: num_aux planes of my rows of mx cells, x fastest, element
 * type as the state (the transpose of pyclaw's state.aux(num_aux, mx, my)); read only.  aux ghost
 * values come from aux_bc_lower/upper by index mapping, as pyclaw's _apply_bcs fills auxbc.
 * Replace classic2.step2 / step2ds called with auxbc and use_fwave (upstream classic/step2.f90,
 * step2ds.f90, flux2.f90: the f-wave correction uses sign(s) where the wave form uses |s|). */
int claw2d_step_aux(const claw2d_desc* d, const void* q_in_dev, void* q_out_dev, const void* aux_dev,
                    double dt, claw2d_ctrl* ctrl_dev, void* stream);
int claw2d_advance_aux(const claw2d_desc* d, void* q0_dev, void* q1_dev, const void* aux_dev,
                       claw2d_ctrl* ctrl_dev, long long attempts, void* stream);
int claw2d_step_ds_aux(const claw2d_desc* d, const void* q_in_dev, void* q_out_dev,
                       const void* aux_dev, double dt, claw2d_ctrl* ctrl_dev, void* stream);
int claw2d_advance_ds_aux(const claw2d_desc* d, void* q0_dev, void* q1_dev, const void* aux_dev,
                          claw2d_ctrl* ctrl_dev, long long attempts, void* stream);

/* One unsplit step2 from a ghost-padded state, exactly the f2py call's form: qbc_in_dev holds
 * num_eqn planes of my+4 rows of mx+4 cells (x fastest) whose two ghost layers the caller has
 * filled, as pyclaw's _apply_bcs fills qbc (any rule, BC.custom user callbacks included); q_out_dev
 * receives the interior planes (num_eqn, my, mx).  The descriptor's q BC codes are not used (0 =
 * custom is accepted); aux_dev (num_aux > 0) is as for claw2d_step_aux, its BCs by index mapping.
 * CFL number and status as claw2d_step.  Replaces classic2.step2(..., qold, qnew, ...) called after
 * _apply_bcs (upstream classic/solver.py step_hyperbolic, classic/step2.f90). */
int claw2d_step_padded(const claw2d_desc* d, const void* qbc_in_dev, void* q_out_dev,
                       const void* aux_dev, double dt, claw2d_ctrl* ctrl_dev, void* stream);

/* One sweep of the dimensionally split step from a ghost-padded state, for ghost cells filled by
 * the caller between the sweeps (BC.custom with dimensional_split).  ids = 1: x-sweeps of
 * qbc_in_dev (padded as for claw2d_step_padded) write q* into the interior cells of q_out_dev,
 * which is padded the same way (its ghost cells are not written: the caller fills them from q*);
 * ids = 2: y-sweeps of the padded q* in qbc_in_dev write the interior planes (num_eqn, my, mx) of
 * q_out_dev.  The x call folds its speed maximum into ctrl_dev and leaves the step open; the y
 * call closes it: ctrl_dev->cfl is then the maximum of both sweeps' CFL numbers and the status
 * bits are those of the stored q^{n+1}.  Call them in pairs, in that order, on one stream.  Other
 * arguments as claw2d_step_padded (-27: ids not 1 or 2).  Replaces classic2.step2ds(..., ids)
 * called from ClawSolver2D.step_hyperbolic with ghost cells refilled between the two directions
 * (upstream classic/solver.py, classic/step2ds.f90). */
int claw2d_sweep_padded(const claw2d_desc* d, const void* qbc_in_dev, void* q_out_dev,
                        const void* aux_dev, double dt, int32_t ids, claw2d_ctrl* ctrl_dev,
                        void* stream);

/* Bytes of the workspace claw2d_mass_sum needs. */
size_t claw2d_mass_sum_workspace_bytes(const claw2d_desc* d);

/* Compensated (Neumaier) FP64 sum of component m over the interior of q_dev (FP32 or FP64 state),
 * in a fixed reduction order (bitwise reproducible).  Writes the sum to out_dev[0]; for m = 0 of a
 * state with a depth_ref, the sum of h (depth_ref times mx*my added as one more compensated term).
 * No counterpart in the reference (pyclaw sums with numpy in FP64). */
int claw2d_mass_sum(const claw2d_desc* d, const void* q_dev, int32_t m, double* out_dev,
                    void* workspace_dev, void* stream);

/* claw2d_mass_sum without the workspace argument.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The partial sums live in a 16 KiB array of file-scope device
 * memory that the library holds on each GPU it runs on until the process ends; this function
 * allocates nothing.  Calls that overlap on one GPU (other streams, other threads) share that
 * array and race.  Grids up to claw2d_mass_sum_noalloc_max_cells() cells (2^31); larger ones are
 * rejected.  Same result, bit for bit, as claw2d_mass_sum. */
int claw2d_mass_sum_noalloc(const claw2d_desc* d, const void* q_dev, int32_t m, double* out_dev,
                            void* stream);
size_t claw2d_mass_sum_noalloc_max_cells(void);

/* ----------------------------------------------------------------------------------------
 * Adjoint (no counterpart in the reference: PyClaw has no adjoint).  Shallow water, FP64.
 * ---------------------------------------------------------------------------------------- */

/* Bytes of the workspace claw2d_adjoint_step needs: 6*(mx+4)*(my+4) doubles. */
size_t claw2d_adjoint_workspace_bytes(const claw2d_desc* d);

/* Reverse of one fixed-dt step: lam_out = (d q^{n+1} / d q^n)^T lam_in, where q^{n+1} is what
 * claw2d_step makes from q_dev = q^n with this dt (same method, limiters, BCs, physics).  All
 * arrays are SoA planes (3, my, mx) of doubles in device memory; lam_out may alias lam_in.
 * Branches (entropy fix, limiter case, upwind side, transverse speed signs) are those the forward
 * takes at q^n.  Two kernels on `stream`; the workspace is overwritten. */
int claw2d_adjoint_step(const claw2d_desc* d, const void* q_dev, const void* lam_in_dev,
                        void* lam_out_dev, void* workspace_dev, double dt, void* stream);

/* ----------------------------------------------------------------------------------------
 * Batched grids (no counterpart in the reference: pyclaw advances one grid per solver).
 * All grids share one claw2d_desc (solver, dtype, method, BCs, physics; its mx/my/dx/dy are
 * ignored) and one ping-pong pair of state buffers; grid g's planes start at element offset
 * grids[g].offset in both buffers.  Each grid has its own claw2d_ctrl (own CFL-chosen dt,
 * accept/reject, done flag).  One persistent kernel per attempt advances every grid.
 * ---------------------------------------------------------------------------------------- */
typedef struct claw2d_batch_grid {
  int64_t offset;           /* element offset of this grid's plane 0 (caller) */
  int32_t mx, my;           /* interior cells (caller) */
  int32_t nstrips, nsegs;   /* work decomposition (claw2d_batch_prepare) */
  int32_t seg_h, item0;     /* rows per item, first work item (claw2d_batch_prepare) */
  int32_t reserved[2];
  double dx, dy;            /* (caller) */
} claw2d_batch_grid;

/* Fills the decomposition fields of grids[0..ngrids) (host memory) for rows-per-item seg_h
 * (0 = library default) and returns the total number of work items. */
int claw2d_batch_prepare(const claw2d_desc* d, claw2d_batch_grid* grids, int32_t ngrids,
                         int32_t seg_h, int64_t* total_items);

/* Enqueues `attempts` evolve_to_time attempts of every grid: one persistent batched step kernel
 * and one per-grid control kernel per attempt.  grids_dev is the prepared table copied to the
 * device; ctrls_dev holds ngrids control blocks; counter_dev is one int, zero before the first
 * call (the control kernel rearms it). */
int claw2d_batch_advance(const claw2d_desc* d, const claw2d_batch_grid* grids_dev, int32_t ngrids,
                         int64_t total_items, void* q0_dev, void* q1_dev, claw2d_ctrl* ctrls_dev,
                         int32_t* counter_dev, long long attempts, void* stream);

/* Captures `attempts` batched attempts (arguments as claw2d_batch_advance) into a HIP graph and
 * returns its executable instance; launch it with claw2d_graph_launch, free with
 * claw2d_graph_destroy.  The captured sequence contains kernels only (no memset nodes). */
int claw2d_batch_graph_create(const claw2d_desc* d, const claw2d_batch_grid* grids_dev,
                              int32_t ngrids, int64_t total_items, void* q0_dev, void* q1_dev,
                              claw2d_ctrl* ctrls_dev, int32_t* counter_dev, long long attempts,
                              void** graph_exec_out);
int claw2d_graph_launch(void* graph_exec, void* stream);
int claw2d_graph_destroy(void* graph_exec);

/* sizeof(claw2d_desc) for which = 0, sizeof(claw2d_ctrl) for which = 1,
 * sizeof(claw2d_batch_grid) for which = 2 (binding checks). */
size_t claw2d_abi_struct_size(int32_t which);

/* Text of the last error on the calling thread ("" if none). */
const char* claw2d_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CLAWPACK_AMD_H */
