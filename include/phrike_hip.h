/*
 * phrike_hip.h -- C ABI of the MI355X (gfx950) backend for the pseudo-spectral
 * Euler right-hand side and RK4 step.
 *
 * The reference has no FFI: its backend switch is the string `backend`
 * threaded from phrike.run_simulation to Grid*(backend=...)
 * (phrike/__init__.py:14-56, phrike/grid.py:1424-1426, 924) and the solver is
 * plain array algebra on whatever the grid returns (phrike/solver.py:1177-1337).
 * Each entry point below replaces one reference routine, cited per function;
 * the Python host layer (phrike_amd) binds them with ctypes.
 *
 * Conventions
 *   - All array arguments are device pointers (HIP device memory) in the
 *     reference layout: (ncomp, Nz, Ny, Nx) for 3D, (batch, ncomp, Ny, Nx)
 *     for batched 2D, x fastest, C order, element type float64 or float32 as
 *     chosen at plan creation (PH_F64 / PH_F32).
 *   - `stream` is a hipStream_t (NULL = the null stream).  Every call is
 *     asynchronous on that stream; nothing synchronises the device.
 *   - `ws` is caller-owned device scratch of at least *_workspace_bytes();
 *     concurrent calls need distinct workspaces.
 *   - Return value 0 = success, negative = error; ph_last_error() returns a
 *     thread-local message for the last failure.
 *   - Grid sizes are powers of two in [8, 1024] along every axis.
 */
#ifndef PHRIKE_HIP_H
#define PHRIKE_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PH_F64 0
#define PH_F32 1

#define PH_OK 0
#define PH_ERR_ARG (-2)
#define PH_ERR_HIP (-5)
#define PH_ERR_INTERNAL (-9)

typedef struct ph_plan3d_s* ph_plan3d_t;
typedef struct ph_plan2d_s* ph_plan2d_t;
typedef struct ph_planh_s* ph_planh_t;
typedef struct ph_planmhd3d_s* ph_planmhd3d_t;

/* Library ABI version (no reference counterpart): PH_ABI_VERSION of the
 * header the library was built with. */
#define PH_ABI_VERSION 1
int ph_abi_version(void);

/* Per-kernel timing (no reference counterpart; measurement only).  While
 * enabled, every kernel launch of the library is bracketed by HIP events on
 * its stream.  ph_timing_collect() waits for the recorded events and writes
 * lines "<kernel-tag> <launches> <total-ms>\n" into buf, then resets. */
int ph_timing_enable(int on);
int ph_timing_collect(char* buf, size_t len);

/* Message for the last failed call on this thread.  Replaces the reference's
 * Python exceptions (ValueError/RuntimeError, phrike/grid.py:66-84, 308). */
const char* ph_last_error(void);

/* ------------------------------------------------------------------------
 * 3D periodic grid.  Replaces Grid3D.__post_init__ (phrike/grid.py:1378-1415:
 * wavenumbers, 2/3 mask, exponential filter tables).  filter_p/filter_alpha
 * follow integration.spectral_filter {enabled, p, alpha}
 * (phrike/problems/base.py:103-105; defaults p=8, alpha=36 at grid.py:1411).
 * ------------------------------------------------------------------------ */
int ph_plan3d_create(ph_plan3d_t* plan, int nx, int ny, int nz, double lx, double ly, double lz,
                     int dtype, int dealias, int filter_enabled, int filter_p, double filter_alpha,
                     double gamma);
int ph_plan3d_destroy(ph_plan3d_t plan);
int ph_plan3d_workspace_bytes(ph_plan3d_t plan, size_t* bytes);

/* rhs = -(dx Fx + dy Fy + dz Fz) from U (5 fields).
 * Replaces _compute_rhs_3d (phrike/solver.py:1177-1184) with
 * EulerEquations3D.flux (phrike/equations.py:582-607) and Grid3D.dx1/dy1/dz1
 * (phrike/grid.py:1705-1721). */
int ph_rhs3d(ph_plan3d_t plan, const void* U, void* rhs, void* ws, void* stream);

/* out = d/d(axis) f for nfields (1..5) fields; axis 0 = x, 1 = y, 2 = z;
 * axis -1 = Grid3D.apply_spectral_filter.  Replaces Grid3D.dx1/dy1/dz1 and
 * apply_spectral_filter (phrike/grid.py:1705-1726).  in == out is allowed. */
int ph_deriv3d(ph_plan3d_t plan, int axis, int nfields, const void* f, void* out, void* ws,
               void* stream);

/* One RK4 step followed by the post-step spectral filter:
 * out = filter(RK4(U, dt)).  dt is taken from dt_dev[0] when dt_dev is not
 * NULL (device FP64), else from `dt`.  U and out must not overlap.
 * Replaces SpectralSolver3D.step/_fixed_step/_rk4_step_3d/
 * _apply_physical_filters_3d (phrike/solver.py:1220-1234, 1313-1337). */
int ph_step3d(ph_plan3d_t plan, const void* U, void* out, double dt, const double* dt_dev,
              void* ws, void* stream);

/* ph_step3d with flags.
 *
 * With the 2/3 rule on (every workload), a 3D Euler step keeps its RK4 sum in
 * spectral space: it starts from S = the masked spectrum of U, each stage
 * subtracts w D (the stage's RHS spectrum is -D; w = dt/6, dt/3, dt/3, dt/6)
 * and stage 4 leaves sigma_x sigma_y sigma_z S, the spectrum of the filtered
 * new state, which is transformed back.  That is the reference's post-step
 * filter of U + dt/6 (k1 + 2 k2 + 2 k3 + k4) (phrike/solver.py:1220-1234),
 * since each k is band-limited, to rounding.
 *
 * PH_STEP_BANDLIMITED: U has no content beyond the 2/3 cutoff (true of every
 * output of a previous step).  Plans without the 2/3 rule, and a library
 * built with PH_SPECACC=0, then apply the FP32 post-step filter in
 * correction form, out = U' + F^-1((sx sy sz - 1) F U') on the kept modes
 * (the FP32 transforms carry only the small correction); other plans ignore
 * the flag.  No reference counterpart: a precision choice. */
#define PH_STEP_BANDLIMITED 1
/* PH_STEP_CHAIN: end the step with the x transform of its output's fluxes
 * (the next step's first pass, formed while the output is in registers) and
 * keep them, and the output's spectrum S, in the workspace.
 * PH_STEP_CHAINED: U is the unmodified output of a PH_STEP_CHAIN step of this
 * plan (or the final state of a PH_STEP_CHAIN ph_run3d call); if the
 * workspace still holds U's fluxes and spectrum the step skips its own first
 * passes (else it transforms U itself; the two agree to rounding).  That is
 * checked by address (the plan, the workspace, U equal to that step's output
 * or one of that call's buffers, and no call since that wrote the
 * workspace's spectral buffers).  The caller promises that U's contents are
 * unchanged and, after ph_run3d, that U is the buffer holding its final
 * state.  No reference counterpart: a fusion across the step boundary. */
#define PH_STEP_CHAIN 4
#define PH_STEP_CHAINED 8
int ph_step3d_ex(ph_plan3d_t plan, const void* U, void* out, double dt, const double* dt_dev, int flags,
                 void* ws, void* stream);

/* Device-resident run loop.  Per-run scalars live on the device:
 *   rs_dev[f * nruns + run], FP64, f = PH_RS_T (time), PH_RS_DT (next dt),
 *     PH_RS_TEND (t_end), PH_RS_TSTOP (stop time: the caller's next output or
 *     checkpoint), PH_RS_DTNEXT, PH_RS_DTUSED (dt of the last step taken);
 *   rf_dev[f * nruns + run], int32, f = PH_RF_ACTIVE, PH_RF_STEPS;
 * (nruns = 1 for ph_run3d).  Each of the nsteps steps, for every active run:
 *   dt = min(dt, t_end - t); U' = step(U, dt); dt_next = cfl*h/max_speed(U');
 *   t += dt; dt_used = dt; dt = dt_next; steps += 1;
 *   active = t < t_end - 1e-12 && t + 1e-12 < t_stop
 * (a non-finite dt retires the run with t = NaN).  Step k reads Ua (k even)
 * or Ub (k odd) and writes the other; steps of an inactive run do nothing, so
 * a run's state ends in Ua if it took an even number of steps in this call,
 * else in Ub.  The caller sets dt to the CFL dt of the starting state
 * (ph_max_wave_speed*), t_stop, and active = 1 before the call; nothing
 * synchronises the host, which reads t, steps and active back only when it
 * wants the state (outputs, checkpoints, the end).  `flags` (ph_run3d):
 * PH_STEP_BANDLIMITED for the first step when Ua is a previous step's
 * output (later steps always are); the steps of one call are chained;
 * PH_STEP_CHAIN also chains the call's last step to the next call (which
 * passes PH_STEP_CHAINED with the buffer holding this call's final state as
 * its Ua), so that a run in several calls computes what one call would.  ph_run2d applies the post-step filter
 * every step (filter interval 1).  Replaces the run loops of
 * SpectralSolver3D.run (phrike/solver.py:1475-1484) and SpectralSolver2D.run
 * (phrike/solver.py:764-779). */
#define PH_RS_T 0
#define PH_RS_DT 1
#define PH_RS_TEND 2
#define PH_RS_TSTOP 3
#define PH_RS_DTNEXT 4
#define PH_RS_DTUSED 5
#define PH_RS_N 6
#define PH_RF_ACTIVE 0
#define PH_RF_STEPS 1
#define PH_RF_N 2
int ph_run3d(ph_plan3d_t plan, void* Ua, void* Ub, int nsteps, double* rs_dev, int* rf_dev, double cfl, int flags,
             void* ws, void* stream);

/* max over cells of max(|ux|,|uy|,|uz|) + sqrt(gamma p / rho) -> max_dev[0]
 * (FP64), and when dt_dev is not NULL dt_dev[0] = cfl*min(dx,dy,dz)/max (1e-6
 * if max <= 0).  Replaces EulerEquations3D.max_wave_speed
 * (phrike/equations.py:609-620, 199-211) and SpectralSolver3D.compute_dt
 * (phrike/solver.py:1307-1311).  Deterministic two-level reduction. */
int ph_max_wave_speed3d(ph_plan3d_t plan, const void* U, double* max_dev, double* dt_dev,
                        double cfl, void* ws, void* stream);

/* FP64 sums over cells of rho, rho*ux, rho*uy, rho*uz, p/(gamma-1)+rho|u|^2/2
 * -> out5_dev.  Replaces the record() sums of SpectralSolver3D.run
 * (phrike/solver.py:1439-1468) and EulerEquations3D.conserved_quantities
 * (phrike/equations.py:622-637).  Deterministic. */
int ph_conserved_sums3d(ph_plan3d_t plan, const void* U, double* out5_dev, void* ws, void* stream);

/* 3D transforms of a plan created with dealias = 0 and filter_enabled = 0
 * (every mode kept): ph_fft3d writes the unnormalised forward transform of
 * nfields (1..5) real fields as half spectra (nfields, nz, ny, nx/2 + 1) of
 * interleaved complex values; ph_ifft3d takes such half spectra, read as the
 * kx >= 0 half of a Hermitian spectrum, and writes the real fields of the
 * normalised inverse.  The reference's Grid3D.fftn / ifftn (phrike/grid.py:
 * 1497-1513, full complex spectra; ifftn returns the real part) are these
 * with the kx < 0 half filled by Hermitian symmetry (fftn) or the input
 * first reduced to its Hermitian part, whose inverse is that real part
 * (ifftn); the host layer does that bookkeeping. */
int ph_fft3d(ph_plan3d_t plan, int nfields, const void* f, void* spec, void* ws, void* stream);
int ph_ifft3d(ph_plan3d_t plan, int nfields, const void* spec, void* f, void* ws, void* stream);

/* Monitoring statistics of a state, FP64, out_dev[PH_MON_*] (per run at
 * out_dev[run * PH_MON_N + f] for ph_monitor2d):
 *   PH_MON_MASS sum rho, PH_MON_RV2 sum rho |u|^2, PH_MON_VMAX / VMIN max / min
 *   |u|, PH_MON_RMAX / RMIN max / min rho, PH_MON_RS1 / RS2 sum (rho - c) and
 *   sum (rho - c)^2 with c = PH_MON_SHIFT (rho of the first cell; mean and
 *   variance without cancellation), PH_MON_PMAX / PMIN / PSUM max, min, sum
 *   of p, PH_MON_B2 sum |B|^2 (MHD, else 0).  A NaN cell makes the max / min
 *   entries NaN.  Deterministic.  The reductions behind
 *   BaseProblem.compute_velocity_stats / compute_density_stats /
 *   compute_mhd_stats (phrike/problems/base.py:666-822); the host forms
 *   v_rms, means and standard deviations from them. */
#define PH_MON_MASS 0
#define PH_MON_RV2 1
#define PH_MON_VMAX 2
#define PH_MON_VMIN 3
#define PH_MON_RMAX 4
#define PH_MON_RMIN 5
#define PH_MON_RS1 6
#define PH_MON_RS2 7
#define PH_MON_PMAX 8
#define PH_MON_PMIN 9
#define PH_MON_PSUM 10
#define PH_MON_B2 11
#define PH_MON_SHIFT 12
#define PH_MON_N 13
int ph_monitor3d(ph_plan3d_t plan, const void* U, double* out_dev, void* ws, void* stream);

/* ------------------------------------------------------------------------
 * Batched 2D Fourier x Fourier grid: `batch` independent runs of one shape,
 * state (batch, 4, ny, nx).  Replaces Grid2D Fourier mode
 * (phrike/grid.py:846-911, 1157-1287) for every run of a batch.
 * ------------------------------------------------------------------------ */
int ph_plan2d_create(ph_plan2d_t* plan, int nx, int ny, double lx, double ly, int dtype,
                     int dealias, int filter_enabled, int filter_p, double filter_alpha,
                     double gamma, int batch);
int ph_plan2d_destroy(ph_plan2d_t plan);
int ph_plan2d_workspace_bytes(ph_plan2d_t plan, size_t* bytes);

/* rhs = -(dx Fx + dy Fy) for the first nbatch runs.  Replaces
 * _compute_rhs_2d (phrike/solver.py:428-443). */
int ph_rhs2d(ph_plan2d_t plan, int nbatch, const void* U, void* rhs, void* ws, void* stream);

/* 2D derivative / filter of the 4 fields of each run (axis 0, 1, or -1 for
 * the filter).  Replaces Grid2D.dx1/dy1/apply_spectral_filter
 * (phrike/grid.py:1226-1287). */
int ph_deriv2d(ph_plan2d_t plan, int nbatch, int axis, const void* f, void* out, void* ws,
               void* stream);

/* One RK4 step with the positivity clamp after every stage and at the end,
 * then the filter if apply_filter != 0.  Per-run dt from dt_dev[run] (FP64)
 * when dt_dev is not NULL, else `dt` for all runs.  Runs with
 * active_dev[run] == 0 are left untouched in `out` (active_dev may be NULL).
 * U and out must not overlap.  Replaces SpectralSolver2D._fixed_step,
 * _rk4_step_2d, _positivity_clamp_2d, _apply_physical_filters_2d
 * (phrike/solver.py:467-524, 639-648). */
int ph_step2d(ph_plan2d_t plan, int nbatch, const void* U, void* out, double dt,
              const double* dt_dev, const int* active_dev, int apply_filter, void* ws,
              void* stream);

/* ph_step2d without a workspace: the scratch memory its kernels use (the
 * spectral buffer of the batch, 8 * nbatch * ny * mxp complex values, mxp =
 * nx/3 + 1 rounded up to 8 (FP64) or 16 (FP32)) is a static 32 MiB device
 * array of the library, so the call allocates nothing.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  A batch needing more than 32 MiB is an argument error (nothing
 * is launched).  Every call on one GPU shares that GPU's array and leaves its
 * contents behind: calls that overlap on one GPU (other streams, other
 * threads) race, so serialise them.  The array's device address is resolved
 * at 2D plan creation on the current device (or at the first call on a
 * device); arguments and results otherwise as ph_step2d. */
int ph_step2d_noalloc(ph_plan2d_t plan, int nbatch, const void* U, void* out, double dt,
                      const double* dt_dev, const int* active_dev, int apply_filter, void* stream);

/* Per-run max of |ux| + sqrt(gamma p / rho) -> max_dev[run] and, when dt_dev
 * is not NULL, dt_dev[run] = cfl*min(dx,dy)/max (1e-6 if max <= 0).  Runs with
 * active_dev[run] == 0 are not read (their outputs are meaningless);
 * active_dev may be NULL.
 * Replaces EulerEquations2D.max_wave_speed (phrike/equations.py:365-370,
 * 137-147) and SpectralSolver2D.compute_dt (phrike/solver.py:613-620). */
int ph_max_wave_speed2d(ph_plan2d_t plan, int nbatch, const void* U, double* max_dev,
                        double* dt_dev, double cfl, const int* active_dev, void* ws, void* stream);

/* The device-resident run loop (see ph_run3d) for nbatch runs of a batched
 * 2D plan, each with its own scalars. */
int ph_run2d(ph_plan2d_t plan, int nbatch, void* Ua, void* Ub, int nsteps, double* rs_dev, int* rf_dev, double cfl,
             void* ws, void* stream);

/* Per-run FP64 sums of rho, rho*ux, rho*uy, energy -> out4_dev[run*4 + i].
 * Replaces the record() sums of SpectralSolver2D.run
 * (phrike/solver.py:734-757). */
int ph_conserved_sums2d(ph_plan2d_t plan, int nbatch, const void* U, double* out4_dev, void* ws,
                        void* stream);

/* ph_monitor3d for the first nbatch runs (4-field 2D states). */
int ph_monitor2d(ph_plan2d_t plan, int nbatch, const void* U, double* out_dev, void* ws, void* stream);

/* Per-run time bookkeeping for device-resident run loops (ensembles): for
 * every run with active_dev[run] != 0,
 *   ph_runs_clip:    dt[run] = min(dt[run], t_end[run] - t[run])
 *   ph_runs_advance: t[run] += dt[run]; steps[run] += 1;
 *                    active[run] = t[run] < t_end[run] - 1e-12
 * Replaces the host loop arithmetic of SpectralSolver2D.run / SpectralSolver3D.run
 * (phrike/solver.py:764-779, 1475-1484).  All FP64 / int32 device arrays. */
int ph_runs_clip(int nbatch, const double* t_dev, const double* t_end_dev, double* dt_dev,
                 const int* active_dev, void* stream);
int ph_runs_advance(int nbatch, double* t_dev, const double* t_end_dev, const double* dt_dev,
                    int* active_dev, int* steps_dev, void* stream);

/* ------------------------------------------------------------------------
 * Dense LGL collocation derivative on matrix cores:
 *   out[b, i, x] = sum_j D[i, j] F[b, j, x]      (b < nbatch)
 * and, when add != NULL, the hybrid-RHS epilogue out = -(add + D.F).
 * F, add, out are (nbatch, ny, nx); D/Dhi/Dlo are row-major ny x ny.  out
 * must not overlap F (every output row reads the whole F line; an overlap is
 * an argument error); add may be out.
 * mode PH_GLL_F64: FP64 MFMA with D (float64); data float64.
 * mode PH_GLL_SPLIT_F32: FP32 MFMA, D split as Dhi + Dlo (float32 each,
 *   Dhi = fp32(D), Dlo = fp32(D - Dhi)); data float32.
 * mode PH_GLL_SPLIT_F64IO: FP32 MFMA on float64 data split as hi + lo:
 *   Dhi*Fhi + Dhi*Flo + Dlo*Fhi, result in float64.
 * mode | PH_GLL_OFFSET (split modes): apply D to F - F[row 0] per column,
 *   which equals D.F when the rows of D sum to zero (a derivative matrix) and
 *   keeps FP32 rounding proportional to the variation of F along the line.
 * ny, nx multiples of 16.  Replaces LegendreLobattoBasis2D.dy /
 * LegendreLobattoBasis1D.dx applies (phrike/basis/legendre.py:787-802,
 * 1020-1028; torch matmul at 1027). */
#define PH_GLL_F64 0
#define PH_GLL_SPLIT_F32 1
#define PH_GLL_SPLIT_F64IO 2
#define PH_GLL_OFFSET 8
int ph_gll_apply(int mode, int ny, int nx, int nbatch, const void* D, const void* Dhi, const void* Dlo,
                 const void* F, const void* add, void* out, void* stream);

/* ------------------------------------------------------------------------
 * 2D Euler on the hybrid grid: Fourier along x (periodic, length lx, nx
 * points), Legendre-Gauss-Lobatto collocation along y (ny nodes).  State U is
 * (nbatch, 4, ny, nx): rho, mx, my, E.
 *   rhs = -(dx Fx + D Fy), dx the masked/filtered 1D spectral derivative of
 *   Grid1D.dx1 (phrike/grid.py:512-519) per row, D the LGL derivative matrix
 *   along y (LegendreLobattoBasis1D.D_x, phrike/basis/legendre.py:714-717);
 *   step: RK4 with the positivity clamp (phrike/solver.py:467-513), then,
 *   with apply_filter, x mask*sigma (Grid1D.apply_spectral_filter,
 *   phrike/grid.py:533-538) and the LGL nodal filter along y
 *   (phrike/basis/legendre.py:817-872);
 *   dt = cfl * min(lx/nx, dy_min) / max(|ux| + a) (phrike/solver.py:613-620).
 * y_bc 1 (wall): along y, the 1D Legendre grid's wall rule (phrike/grid.py:
 *   621-738): my = 0 on the boundary rows, rho and E copied from the adjacent
 *   row; applied to every RHS input (the step's base state included, as the
 *   reference's in-place BC at phrike/solver.py:64-68) and after the filter.
 *   y_bc 0: none (the reference hybrid grid's no-op, phrike/grid.py:1082-1084).
 * D_host and Fn_host are row-major ny x ny float64 (derivative, nodal filter;
 * pass the identity for no y filter).  gll_mode: PH_GLL_F64 or
 * PH_GLL_SPLIT_F64IO for float64 data, PH_GLL_SPLIT_F32 for float32.
 * nx a power of two in [16, 1024], ny a multiple of 16.  The reference's
 * own Legendre Grid2D path has no working counterpart (Gauss nodes, FFTs
 * across the non-periodic axis: phrike/grid.py:738-772, 852-891). */
int ph_planh_create(ph_planh_t* plan, int nx, int ny, double lx, double dy_min, int dtype, int dealias,
                    int filter_enabled, int filter_p, double filter_alpha, double gamma, int batch,
                    const double* D_host, const double* Fn_host, int gll_mode, int y_bc);
int ph_planh_destroy(ph_planh_t plan);
int ph_planh_workspace_bytes(ph_planh_t plan, size_t* bytes);
int ph_rhs2dh(ph_planh_t plan, int nbatch, const void* U, void* rhs, void* ws, void* stream);
int ph_step2dh(ph_planh_t plan, int nbatch, const void* U, void* out, double dt, const double* dt_dev,
               const int* active_dev, int apply_filter, void* ws, void* stream);
int ph_max_wave_speed2dh(ph_planh_t plan, int nbatch, const void* U, double* max_dev, double* dt_dev,
                         double cfl, void* ws, void* stream);

/* ------------------------------------------------------------------------
 * 3D ideal MHD, Fourier in x, y, z.  State U is (8, Nz, Ny, Nx): rho, rho ux,
 * rho uy, rho uz, E, Bx, By, Bz.  Replaces SpectralSolverMHD3D with the
 * reference defaults (resistivity and viscosity 0, divergence_clean
 * "projection", projection after every substage, filter interval 1;
 * phrike/solver.py:840-1050, 1173), MHDEquations3D (phrike/equations.py:
 * 911-1100) and Grid3D.project_solenoidal (phrike/grid.py:1648-1684).  Same
 * grid arguments as ph_plan3d_create.
 * ------------------------------------------------------------------------ */
int ph_planmhd3d_create(ph_planmhd3d_t* plan, int nx, int ny, int nz, double lx, double ly, double lz,
                        int dtype, int dealias, int filter_enabled, int filter_p, double filter_alpha,
                        double gamma);
int ph_planmhd3d_destroy(ph_planmhd3d_t plan);
int ph_planmhd3d_workspace_bytes(ph_planmhd3d_t plan, size_t* bytes);

/* rhs = -(dx Fx + dy Fy + dz Fz) with the 8-component ideal-MHD fluxes.
 * Replaces SpectralSolverMHD._compute_rhs / _flux_divergence
 * (phrike/solver.py:932-954) and MHDEquations3D.flux (equations.py:949-1039). */
int ph_rhs_mhd3d(ph_planmhd3d_t plan, const void* U, void* rhs, void* ws, void* stream);

/* In place: B <- Helmholtz projection of B (Nyquist planes zeroed, every mode
 * kept), E recomputed from the projected B at the thermal pressure of the
 * unprojected state.  Replaces SpectralSolverMHD._project_B
 * (phrike/solver.py:966-986) with Grid3D.project_solenoidal. */
int ph_project_mhd3d(ph_planmhd3d_t plan, void* U, void* ws, void* stream);

/* One RK4 step, each stage state clamped (rho >= 1e-10, p >= 1e-12 through
 * prim -> cons) and projected, then the spectral filter and the projection:
 * out = proj(filter(proj(clamp(RK4(U, dt))))).  U and out must not overlap.
 * Replaces SpectralSolverMHD.step/_fixed_step/_rk_substages/_clamp
 * (phrike/solver.py:956-1006, 1037-1050). */
int ph_step_mhd3d(ph_planmhd3d_t plan, const void* U, void* out, double dt, void* ws, void* stream);

/* ph_step_mhd3d with flags.  PH_STEP_SOLENOIDAL: B of U is solenoidal (U is
 * the output of a previous step, or a projected initial state).  The RHS of B
 * is then solenoidal as well (the induction flux is antisymmetric and every
 * spectral factor is common to the three components), so every stage state
 * keeps a solenoidal B and the projection after each substage is the
 * identity up to round-off: the step runs its stage boundaries fused, as the
 * Euler step does, and projects once, inside the filter's z pass (with the
 * 2/3 rule; else after the filter).  Results equal the unflagged step to
 * round-off.  Without the flag: ph_step_mhd3d.  Replaces
 * the same routines as ph_step_mhd3d (phrike/solver.py:956-1006, 1037-1050). */
#define PH_STEP_SOLENOIDAL 2
int ph_step_mhd3d_ex(ph_planmhd3d_t plan, const void* U, void* out, double dt, int flags, void* ws,
                     void* stream);

/* max(|ux|,|uy|,|uz|) + sqrt(max(a^2 + b^2, 0)) over the grid to max_dev[0]
 * and/or dt_dev[0] = cfl*min(dx,dy,dz)/max (1e-6 if max <= 0).  Replaces
 * MHDEquations3D.max_wave_speed (phrike/equations.py:1041-1061) and
 * SpectralSolverMHD.compute_dt without diffusion (phrike/solver.py:1018-1035). */
int ph_max_wave_speed_mhd3d(ph_planmhd3d_t plan, const void* U, double* max_dev, double* dt_dev, double cfl,
                            void* ws, void* stream);

/* ph_monitor3d for the 8-field MHD state (PH_MON_B2 = sum |B|^2, p the
 * thermal pressure). */
int ph_monitor_mhd3d(ph_planmhd3d_t plan, const void* U, double* out_dev, void* ws, void* stream);

/* FP64 sums to out7_dev[0..6]: mass, momentum x/y/z, energy, magnetic
 * energy, cross helicity.  Replaces MHDEquations3D.conserved_quantities
 * (phrike/equations.py:1083-1100). */
int ph_conserved_sums_mhd3d(ph_planmhd3d_t plan, const void* U, double* out7_dev, void* ws, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PHRIKE_HIP_H */
