/*
 * gpe_hip.h -- C ABI of the MI355X (gfx950) split-step Fourier library.
 *
 * The reference (GPE-MATLAB) has no FFI: its boundary is MATLAB method dispatch on
 * handle classes.  These entry points are exactly what a MEX gateway behind those
 * methods binds; each comment names the reference method (file:line) whose body
 * the call replaces.  Plain pointers and sizes only; no exceptions cross the ABI.
 *
 * Conventions
 *   - Arrays use MATLAB's column-major layout [ny, nx(, nz)] (y contiguous), the
 *     layout of grid.mesh.* (grid2d/init.m:34, grid3d/init.m:41).
 *   - Complex arrays are interleaved complex128 (re, im), as MATLAB R2018a+
 *     mxGetComplexDoubles hands them out.
 *   - Every call returns 0 on success and a negative GPE_E* code on failure;
 *     gpe_last_error(h) then holds a message.  A handle is not thread-safe; calls
 *     are synchronous at return.
 *   - Supported axis lengths: 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512.
 */
#ifndef GPE_HIP_H
#define GPE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define GPE_OK 0
#define GPE_EINVAL (-1)   /* bad argument / unsupported size */
#define GPE_EHIP (-2)     /* HIP runtime error */
#define GPE_ENOMEM (-3)   /* device allocation failed */
#define GPE_ESTATE (-4)   /* call out of order (e.g. outer step before begin) */
#define GPE_ENONFINITE (-5) /* a block reduction produced NaN/Inf */

typedef struct gpe_handle gpe_handle;

/* Grid: grid2d(xmax,nx,ymax,ny) / grid3d(xmax,nx,ymax,ny,zmax,nz)
 * (lib/@grid2d/init.m:5-37, lib/@grid3d/init.m:5-46).  n = {nx, ny, nz},
 * xmax = {xmax, ymax, zmax}; for ndims == 2, n[2] and xmax[2] are ignored. */
typedef struct {
  int ndims;
  int n[3];
  double xmax[3];
} gpe_grid_desc;

/* Task properties read by the hot path (lib/@GPEtask/GPEtask.m:4-34;
 * GPEtaskSpinor adds the 2x2 g matrix, lib/@GPEtaskSpinor/GPEtaskSpinor.m:5-9). */
typedef struct {
  int ncomp;          /* 1: GPEtask, 2: GPEtaskSpinor                              */
  double g[4];        /* g, or the 2x2 matrix column-major {g11, g21, g12, g22}    */
  double gamma;       /* GPEtask.m:10 */
  double decay_rate;  /* GPEtask.m:11 */
  double Ntotal;      /* GPEtask.m:12 */
  double mu_init;     /* GPEtask.m:13 */
  int n_recalc;       /* GPEtask.m:9  */
  double nlincpl;     /* GPEtaskSpinor.m:7 (density-dependent coupling), 0 otherwise */
} gpe_params;

/* grid2d/grid3d construction + GPEtask(grid, V) (GPEtask.m:41-60): allocates the
 * device state, twiddle and k tables.  Selects device 0 unless GPE_DEVICE is set. */
int gpe_create(const gpe_grid_desc* desc, gpe_handle** out);
int gpe_destroy(gpe_handle* h);
const char* gpe_last_error(const gpe_handle* h);

/* Property setters (GPEtask.m:4-34). */
int gpe_set_params(gpe_handle* h, const gpe_params* p);

/* Vtrap / getVtotal(time) result for component comp (GPEtask.m:44-48, 62-71;
 * GPEtaskSpinor.m:24-44).  V is real float64, npts values. */
int gpe_set_potential(gpe_handle* h, int comp, const double* V);

/* Potential read by the block-end chemical-potential reductions of the
 * two-component solve_split.  Those apply applyham(phi) without a time argument
 * (lib/@GPEtaskSpinor/solve_split.m:109-114), i.e. at task.current_time
 * (GPEtaskSpinor.m:56-60), which with a time-dependent potential is the time of
 * the last callback rather than the block time.  V = NULL returns the reductions
 * to the step potential. */
int gpe_set_potential_mu(gpe_handle* h, int comp, const double* V);

/* init_state / current_state upload and download (GPEtask.m:18-19, 119).  The
 * download applies a pending renormalisation (solve_split.m:72). */
int gpe_set_state(gpe_handle* h, int comp, const double* psi);
int gpe_get_state(gpe_handle* h, int comp, double* psi);

/* The same transfers for arrays in C order, [ny][nx] (2D) or [ny][nx][nz] (3D) with
 * the last index fastest, as numpy and C callers hold them.  No reference
 * counterpart: MATLAB arrays are column-major and use the functions above.  The array
 * is copied to a device staging buffer (npts complex values, allocated on first use
 * and kept until gpe_destroy) and reordered on the GPU, so the caller needs no
 * host-side transpose.  Results are identical to the column-major functions applied
 * to the transposed array. */
int gpe_set_state_c(gpe_handle* h, int comp, const double* psi);
int gpe_get_state_c(gpe_handle* h, int comp, double* psi);
int gpe_set_potential_c(gpe_handle* h, int comp, const double* V);
int gpe_set_coupling_c(gpe_handle* h, const double* omega, int is_array);

/* Rotation frequency of the rotating frame (GPEtask.m:7, scalar omega): the kinetic
 * step becomes the axis-split step of set_kinop.m:5-7 / ssft_kin_step.m:22-25 and
 * applyham's kinetic term lap(phi, omega) (grid2d/lap.m, grid3d/lap.m).  One
 * component only.  Takes effect at the next gpe_split_begin / gpe_itp. */
int gpe_set_omega(gpe_handle* h, double omega);

/* Device-evaluated time-dependent potential (no reference counterpart; replaces the
 * per-block host evaluation and upload of getVtotal(time2) = Vtrap + Vtd(task, time2),
 * GPEtask.m:62-71 / GPEtaskSpinor.m:24-44, solve_split.m:53-54, for a Vtd that is a
 * linear combination of fixed fields).  gpe_vtd_set stores nbasis fields G (nbasis x
 * npts) and the trap potential of each component (ncomp x npts); gpe_vtd_coeffs
 * rebuilds every component's potential on the device as
 * Vtrap_j + (c_0 G_0 + c_1 G_1 + ...), summed in that order.  At most 16 fields. */
int gpe_vtd_set(gpe_handle* h, int nbasis, const double* G, const double* Vtrap);
/* target 0: the step potential (gpe_set_potential); 1: the reduction potential
 * (gpe_set_potential_mu) */
int gpe_vtd_coeffs(gpe_handle* h, const double* c, int target);

/* Spinor Rabi coupling (GPEtaskSpinor.m:6): scalar (is_array = 0, omega = {re, im})
 * or space-dependent complex array of npts values (is_array = 1). */
int gpe_set_coupling(gpe_handle* h, const double* omega, int is_array);

/* solve_split prologue (lib/@GPEtask/solve_split.m:13-46;
 * lib/@GPEtaskSpinor/solve_split.m:13-57): NN0, n_rec fallback, complex dt,
 * kinetic operator, initial chemical potential.  muc_out receives muc (the
 * value the reference computes at line 41).  On a one-component 3D grid with a
 * 256-point y axis the step runs as two sweeps; the first such call allocates their
 * work arrays (npts complex values and npts doubles, kept until gpe_destroy). */
int gpe_split_begin(gpe_handle* h, double ddt, int niter_inner, double* muc_out);

/* One iteration j of the outer loop with a static potential
 * (solve_split.m:49-79; spinor solve_split.m:60-115): runs niter_inner/n_rec
 * blocks and returns the values passed to ext_callback (solve_split.m:80):
 * time2 (start time of the last block), muc and ncur. */
int gpe_split_outer(gpe_handle* h, int j, double* time2, double* muc, double* ncur);

/* One block (solve_split.m:53-78) at block time time2 with whatever potential is
 * currently uploaded; used when Vtd is a host function (GPEtask.m:65-67). */
int gpe_split_block(gpe_handle* h, double time2, double* muc, double* ncur);

/* Reductions behind applyham/get_energy on the current state (GPEtask.m:73-107):
 * out = { sum|psi|^2 w, sum V|psi|^2 w, sum |psi|^4 w, Re<psi|lap psi> }
 * (spinor: summed over components; out[4] = 2 Re sum conj(psi1) Omega psi2 w). */
int gpe_measure(gpe_handle* h, double out[5]);

/* applyham(phi) / applyh0(phi) as arrays (GPEtask.m:73-95; GPEtaskSpinor.m:45-86:
 * applyham, applyham1c of each component, applyh0).  phi and out hold ncomp
 * column-major complex arrays of npts values, component after component; phi = NULL
 * applies the operator to the current state (with its pending renormalisation).
 * h0_only = 0: lap(phi[, omega]) + V phi + the interaction (g|phi|^2 phi, or
 * sum_k g_jk |phi_k|^2 phi_j) + the Rabi coupling (Omega phi2 / conj(Omega) phi1);
 * h0_only = 1: lap(phi[, omega]) + V phi.  V is the step potential (gpe_set_potential).
 * Work buffers are allocated for the call ((2 ncomp + 1) npts complex values). */
int gpe_applyham(gpe_handle* h, const double* phi, double* out, int h0_only);

/* The grid's own transforms on a column-major complex array of npts values:
 * op GPE_GRID_FFT = fftn (grid2d/fft.m:11, grid3d/fft.m:11), GPE_GRID_IFFT = ifftn with
 * its 1/npts (grid2d/ifft.m:11, grid3d/ifft.m:11), GPE_GRID_LAP = lap(phi, omega)
 * (grid2d/lap.m:5-9, grid3d/lap.m:5-10: ifftn(kk .* fftn(phi)), or the per-axis
 * rotating-frame form for omega != 0).  Needs no parameters, potential or state; work
 * buffers are allocated for the call (1 or 3 npts complex values). */
#define GPE_GRID_FFT 0
#define GPE_GRID_IFFT 1
#define GPE_GRID_LAP 2
int gpe_grid_transform(gpe_handle* h, int op, double omega, const double* in, double* out);

/* groundstate_itp after process_init_state (lib/@GPEtask/groundstate_itp.m:20-112).
 * The initial state must be uploaded (gpe_set_state).  MU, MU2, EE are caller
 * arrays of max_iter entries (MU2/EE may be NULL when nargout < 3/4).  Returns
 * the iteration count in n_iter, the final dt (itp_adjust_stepsize halves it),
 * and converged = 0 when itp_max_iter stopped the loop (the reference warns). */
typedef struct {
  double dt;
  double eps;
  int max_iter;          /* itp_max_iter, GPEtask.m:26 */
  int min_iter;          /* itp_min_iter, GPEtask.m:27 */
  int adjust_stepsize;   /* itp_adjust_stepsize, GPEtask.m:29 */
  int nargout;           /* 1..4, as MATLAB's nargout */
} gpe_itp_args;
int gpe_itp(gpe_handle* h, const gpe_itp_args* a, double* MU, double* MU2, double* EE, int* n_iter,
            double* dt_final, int* converged);

/* Two-component ground states after the initial state is uploaded (both
 * components): GPEtaskSpinor.groundstate_itp (lib/@GPEtaskSpinor/groundstate_itp.m:1-145,
 * mixture = 0: Rabi coupling from gpe_set_coupling, total norm Ntotal) and
 * GPEtask2comp.groundstate_itp (lib/@GPEtask2comp/groundstate_itp.m:1-157,
 * mixture = 1: no coupling, norms N1 / N2, masses M1 / M2).  omega = 0.  Caller
 * arrays of max_iter entries receive the histories, entries 1..n_iter-1 as the
 * reference returns them: MU (spinor MU, mixture MU1), MU2 (spinor: applyham-based,
 * filled when nargout >= 3; mixture: log(mu2)/dt), EE (nargout >= 4; may be NULL
 * otherwise).  converged = 0 when max_iter stopped the loop (the reference caps at
 * 500000 and warns). */
typedef struct {
  double dt;
  double eps;
  int max_iter;
  int nargout;     /* 1..4, as MATLAB's nargout */
  int mixture;
  double N1, N2;   /* mixture: particle numbers (GPEtask2comp.m:6-7) */
  double M1, M2;   /* mixture: masses (GPEtask2comp.m:10-11) */
} gpe_itp2_args;
int gpe_itp2(gpe_handle* h, const gpe_itp2_args* a, double* MU, double* MU2, double* EE, int* n_iter,
             double* dt_final, int* converged);

/* Stochastic GPE: SGPEtask.solve_split (lib/@SGPEtask/solve_split.m:1-81), one
 * component, omega = 0, from the uploaded state and potential.  begin sets the
 * step (dt = ddt i/(1+i gamma)), the kinetic factor (masked to kk < ecut when
 * ecut > 0 and T > 0), the noise amplitude sqrt(T gamma ddt / weight) and mu =
 * mu_init.  The noise of global step s at grid point p (column-major index) is one
 * complex normal variate from Philox4x32-10 with counter (p, s_lo, s_hi, 0) and key
 * (seed_lo, seed_hi), Box-Muller on two 53-bit uniforms (no reference counterpart:
 * MATLAB's randn stream).  gpe_sgpe_block runs one block whose first step is global
 * step step0 (a time-dependent potential is uploaded before it); gpe_sgpe_outer runs
 * outer iteration j with steps (j-1)*niter_inner + ...  ncur = integrate(|psi|^2)
 * after the block (solve_split.m:75). */
typedef struct {
  double ddt;
  int niter_inner;
  double T;                 /* SGPEtask.m:5 */
  double ecut;              /* SGPEtask.m:8 */
  unsigned long long seed;  /* noise key */
} gpe_sgpe_args;
int gpe_sgpe_begin(gpe_handle* h, const gpe_sgpe_args* a);
int gpe_sgpe_block(gpe_handle* h, long long step0, double* ncur);
int gpe_sgpe_outer(gpe_handle* h, int j, double* time2, double* ncur);

/* ------------------------------------------------------------------ diagnostics
 * Vortex cores of a 2D slice, lib/util/detect_core.m:1-66 called as
 * detect_core(phi(:,:,zslice+1), mesh.x(:,:,zslice+1), mesh.y(:,:,zslice+1), refine),
 * computed on the device so that a user_callback (GPEtask.m:128-135) needs no copy of
 * the state.  x[nx] = rx(1,:) and y[ny] = ry(:,1) are the coordinates of the columns
 * and rows.  resp / resm receive up to `cap` rows [x, y] (row-major) of the positively
 * / negatively charged cores in MATLAB's column-major find order; *np / *nm are the
 * numbers found (rows beyond cap are not written).  With refine != 0 each position is
 * the zero of the bilinear interpolants of Re and Im on the plaquette
 * (detect_core.m:54-64), [0, 0] when none lies inside.  Needs even nx, ny >= 4.
 * gpe_detect_core reads component comp of the resident state and leaves the solver
 * state untouched; gpe_detect_core_array takes a host array phi[2*ny*nx] (interleaved
 * complex, column-major) and uses h only for its device, stream and error slot. */
int gpe_detect_core(gpe_handle* h, int comp, int zslice, const double* x, const double* y, int refine, int cap,
                    double* resp, int* np, double* resm, int* nm);
int gpe_detect_core_array(gpe_handle* h, const double* phi, int ny, int nx, const double* x, const double* y,
                          int refine, int cap, double* resp, int* np, double* resm, int* nm);

/* No reference counterpart.  For callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc): gpe_detect_core with the same
 * arguments and results, whose intermediate fields live in a fixed device array of
 * GPE_DC_NOALLOC_SCRATCH_MIB MiB per GPU (held until the process ends) instead of the
 * handle's grown buffer; it calls no hipMalloc/hipFree.  A slice whose fields and
 * output rows need more than that (about 17 bytes per point + 32 * cap bytes; a
 * 512 x 512 slice fits with cap up to 300 000) returns GPE_EINVAL.  Calls that
 * overlap on one GPU share the array and race. */
#define GPE_DC_NOALLOC_SCRATCH_MIB 16
int gpe_detect_core_noalloc(gpe_handle* h, int comp, int zslice, const double* x, const double* y, int refine,
                            int cap, double* resp, int* np, double* resm, int* nm);

/* ------------------------------------------------------------------ adjoint
 * (no reference counterpart: the differentiation form of the split step, a gradient
 * entry point that plays the torch.autograd.Function.backward role).
 * Control potential V(x, t) = V0(x) + sum_k c[k + nbasis*m] G_k(x), where m is the
 * interval of the block holding t (V is evaluated at block starts,
 * solve_split.m:54; intervals are steps_per_interval steps long).  The forward
 * model is GPEtask.solve_split with gamma = 0; the loss is
 * 1 - |<psi_target|psi(T)>|^2.  Both calls start from the uploaded state
 * (gpe_set_state) and leave it in place. */
typedef struct {
  double ddt;
  int nsteps;              /* T, a multiple of n_recalc                        */
  int steps_per_interval;  /* a multiple of n_recalc                           */
  int nintervals;          /* number of coefficient columns m                  */
  int checkpoint_blocks;   /* blocks between checkpoints, 0: about sqrt(T) steps */
} gpe_control_args;
int gpe_control_set(gpe_handle* h, int nbasis, const double* V0, const double* G);
int gpe_control_forward(gpe_handle* h, const gpe_control_args* a, const double* coeffs, double* psi_out);
/* loss, grad[nbasis*nintervals] (dL/dc, same layout as coeffs), and the relative
 * L2 difference between the backward-reconstructed initial state and the stored
 * one (recon_err, may be NULL). */
int gpe_adjoint_grad(gpe_handle* h, const gpe_control_args* a, const double* coeffs, const double* psi_target,
                     double* loss, double* grad, double* recon_err);

/* ------------------------------------------------------------------ Rabi adjoint
 * (no reference counterpart: the differentiation form on the two-component step).
 * Forward model: GPEtaskSpinor.solve_split (lib/@GPEtaskSpinor/solve_split.m) with
 * gamma = 0, nlincpl = 0, Ntotal > 0, nsteps steps of ddt in blocks of n_recalc,
 * and the coupling replaced by Omega(x) = omega0 exp(i k x).  Loss: the final
 * population imbalance (N1 - N2) / Ntotal.  Both calls start from the uploaded
 * two-component state and leave it in place; they leave the handle's coupling set
 * to that Omega(x), and a later solve_split starts with gpe_split_begin.
 * The gradient cuts the run into segments (about sqrt(T) steps each).  It stores
 * the state before every step of as many trailing segments as tape_gib holds and
 * recomputes the others from per-segment checkpoints; device memory: about
 * 2 * (segments + 1) state arrays plus the tape (2 state arrays per stored step). */
typedef struct {
  double ddt;
  int nsteps;             /* a multiple of n_recalc                               */
  int checkpoint_blocks;  /* blocks per segment, 0: about sqrt(nsteps) steps       */
  double omega0;
  double k;
  double tape_gib;        /* tape budget in GiB; 0: half the free device memory, < 0:
                             one segment (every other segment is recomputed)         */
} gpe_rabi_args;
/* psi_out (may be NULL): 2 * npts complex values, component 1 then component 2 */
int gpe_rabi_forward(gpe_handle* h, const gpe_rabi_args* a, double* loss, double* psi_out);
/* grad[0] = dL/d omega0, grad[1] = dL/d k */
int gpe_rabi_grad(gpe_handle* h, const gpe_rabi_args* a, double* loss, double* grad);

/* ------------------------------------------------------------------ ensemble
 * (no reference counterpart: the batched form).  A batch advances many
 * independent 2D GPEtask.solve_split runs together; each trajectory has its own
 * grid (lengths may differ), potential, state and gpe_params (g, gamma, Ntotal > 0,
 * mu_init, decay_rate; one common n_recalc).  Every sweep is one launch per
 * distinct pencil length for the whole ensemble; with use_graph and no
 * decay_rate, a whole outer iteration is replayed from one captured hipGraph.
 * Trajectory i follows solve_split.m:13-84 exactly as a single handle would. */
typedef struct gpe_batch gpe_batch;
int gpe_batch_create(int ntraj, const gpe_grid_desc* grids, gpe_batch** out);
int gpe_batch_destroy(gpe_batch* b);
const char* gpe_batch_last_error(const gpe_batch* b);
int gpe_batch_set(gpe_batch* b, int i, const gpe_params* p, const double* V, const double* psi);
int gpe_batch_get_state(gpe_batch* b, int i, double* psi);
/* muc_out[ntraj] receives each trajectory's initial muc (solve_split.m:41). */
int gpe_batch_split_begin(gpe_batch* b, double ddt, int niter_inner, int use_graph, double* muc_out);
/* One outer iteration j for every trajectory; muc[ntraj], ncur[ntraj] as passed to ext_callback. */
int gpe_batch_split_outer(gpe_batch* b, int j, double* time2, double* muc, double* ncur);

/* ------------------------------------------------------------------ plan tuning
 * Replaces the FFTW planning the GPEtask constructor does when fftw_set_wisdom is
 * true (lib/@GPEtask/GPEtask.m:28 and 55-58: fftw('dwisdom',[]);
 * fftw('planner','measure'); fftn(grid.mesh.x)).  Pencil lengths 64, 128, 256, 384
 * and 512 have two FFT plans (variant 0 and 1, with different radices, elements per
 * thread and pencils per work-group).  The wisdom is a process-wide table keyed by
 * (ndims, axis, n); gpe_create and gpe_batch_create adopt the variants it holds
 * (variant 0 where it holds none).
 *
 * gpe_tune times the steady-state solve_split step of the handle's grid (x phase
 * sweep; 3D: y forward, z kinetic, y inverse; 2D: y kinetic) for every combination
 * of variants of the axes the wisdom has no plan for (all axes with remeasure != 0),
 * records the fastest combination and adopts the recorded variants.  The state is
 * preserved.  times_us[2*axis + variant] receives the best step time in microseconds
 * among the combinations with that axis at that variant (NaN where nothing was
 * measured); it may be null.  gpe_batch_tune tunes every (axis, length) pair of an
 * ensemble in turn, timing the whole ensemble's step with the pair at each variant
 * (times_us, if not null, holds 2 entries per pair in ascending (axis, length) order,
 * up to cap entries; *npairs receives the number of pairs); call it before
 * gpe_batch_split_begin. */
int gpe_tune(gpe_handle* h, int remeasure, double* times_us);
int gpe_get_plan(gpe_handle* h, int axis, int* variant);
int gpe_batch_tune(gpe_batch* b, int remeasure, double* times_us, int cap, int* npairs);
/* Plan variant recorded for (ndims, axis, n), or -1 if none. */
int gpe_wisdom_get(int ndims, int axis, int n);
int gpe_wisdom_set(int ndims, int axis, int n, int variant);
/* fftw('dwisdom', []): forget every recorded plan (handles keep the plans they hold). */
int gpe_wisdom_forget(void);
/* fftw('wisdom') export / import: one line "ndims axis n variant" per recorded plan.
 * gpe_wisdom_export writes the NUL-terminated text into buf (len bytes) and the size it
 * needs into *needed (buf may be NULL to query it); gpe_wisdom_import records every
 * line of text, or none if any line is malformed. */
int gpe_wisdom_export(char* buf, int len, int* needed);
int gpe_wisdom_import(const char* text);

/* Measurement hooks (no reference counterpart): time every sweep launch with
 * HIP events recorded on the handle's own stream.  gpe_profile_read fills
 * ms[seq*3 + axis] (summed milliseconds) and count[seq*3 + axis] for
 * seq < 11 (the sweep sequences) and axis 0..2 (x, y, z); n >= 33. */
int gpe_profile_enable(gpe_handle* h, int enable);
int gpe_profile_read(gpe_handle* h, double* ms, int* count, int n);

/* Self-test hook (no reference counterpart): out[2i], out[2i+1] =
 * exp(er[i]) * (cos(ei[i]), sin(ei[i])) computed by the device phase code the
 * sweep kernels use, for checking it against a host math library. */
int gpe_selftest_phase(int n, const double* er, const double* ei, double* out);

/* Self-test hook (no reference counterpart): the 16-point factor of the two-sweep
 * 3D step on one tile of 16 pencils x 256 points, moved from the sweeps' pencil layout
 * to the row layout and back.  in, out_f, out_i hold 4096 interleaved complex values,
 * entry a + 16 z = point z of pencil a.  out_f[k + 16 z] = sum_a in[a + 16 z]
 * e^{-2 pi i a k/16}; out_i the same with e^{+2 pi i a k/16} (no 1/16). */
int gpe_selftest_rowdft(const double* in, double* out_f, double* out_i);

/* Library/device information for diagnostics: writes a NUL-terminated string. */
int gpe_device_info(char* buf, int len);

#ifdef __cplusplus
}
#endif

#endif /* GPE_HIP_H */
