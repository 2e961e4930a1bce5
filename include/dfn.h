/*
 * dfn.h -- C ABI of the MI355X (gfx950) per-Runge-Kutta-stage hot path of
 * DemystiFicatioN (2D quasi-dynamic earthquake cycles on fault networks).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (DFN_OK == 0) and never aborts.  All arrays are FP64 unless noted,
 * element-indexed in the reference's concatenated order: N = sum_i(nb_el_i + 1) - 1
 * elements, i.e. the faults one after another with one fake element between
 * consecutive faults (src/configuration.f90:126,136).  Y = [V(1:N) | psi(1:N) |
 * sigmaN(1:N)] as one contiguous 3N array (src/configuration.f90:185-237), with
 * psi = ln(theta V0 / Dc) the log state.  Indices that the reference keeps as
 * Fortran integer(kind=8) are int64_t here, default integers are int32_t.
 *
 * One handle owns one problem: geometry, friction parameters, the stress-transfer
 * operator(s) (device-resident block lists), the fluid-diffusion state and the
 * ODE state.  Calls on one handle must not overlap; several handles may coexist.
 * The same declarations, with the prefix dfno_ instead of dfn_, are implemented
 * by the CPU oracle used in the tests.
 */
#ifndef DFN_H
#define DFN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (reference: ier flags and STOPs, see the notes per call) -- */
enum {
  DFN_OK = 0,
  DFN_ERR_ARG = 1,        /* bad argument / size (reference: assert_eq STOP, nrutil.f90:404-458) */
  DFN_ERR_HIP = 2,        /* HIP runtime error, or no gfx950 device */
  DFN_ERR_CG = 3,         /* compute_P did not converge in 200 iterations (fluid_diffusion.f90:120-125) */
  DFN_ERR_UNDERFLOW = 4,  /* 'stepsize underflow in rkqs' (ode_bs.f90:163-164) */
  DFN_ERR_TENSILE = 5,    /* sigmaN > 0 somewhere (solver.f90:412-420) */
  DFN_ERR_STATE = 6,      /* call out of order (e.g. H-matrix kernel before dfn_build_hmat) */
  DFN_ERR_NOMEM = 7,      /* host or device allocation failed */
  DFN_ERR_UNSUPPORTED = 8, /* option the library does not implement */
  DFN_ERR_CG_TIMEOUT = 9   /* the workgroups of a multi-workgroup compute_P lost contact (a bounded
                              message wait gave up): a device fault, not non-convergence; the
                              handle's state is left at the start of the failed step */
};

enum { DFN_MODE_II = 2, DFN_MODE_III = 3 };               /* hyperparameter/fracture_mode */
enum { DFN_KERNEL_DIRECT = 0, DFN_KERNEL_HMATRIX = 1 };   /* hyperparameter/static_kernel */
enum { DFN_OP_T = 0, DFN_OP_N = 1 };                      /* shear (Mode II/III) and normal (Mode II) operator */
/* storage of the operator blocks: all FP64; all FP32 (power-of-two block scales,
 * FP32 accumulation, compensated near-field sums); FP32 low-rank factors beside
 * FP64 near-field (dense) blocks */
enum { DFN_PREC_FP64 = 0, DFN_PREC_FP32 = 1, DFN_PREC_MIXED = 2, DFN_PREC_FP32_STORE = 3, DFN_PREC_MIXED_SPLIT = 4 };

/* Problem description: the content of problems/<name>/config
 * (src/configuration.f90:70-379) as flat arrays.  Per-element arrays have
 * N entries and include the fake slots between faults; the library overwrites
 * the fake slots with the reference's values (a=b=Dc=f0=V0=Snk=Lk=Tk=1,
 * src/configuration.f90:199-203,:226-232; everything else 0). */
typedef struct dfn_problem {
  int32_t fracture_mode;             /* DFN_MODE_II or DFN_MODE_III */
  int32_t static_kernel;             /* DFN_KERNEL_DIRECT or DFN_KERNEL_HMATRIX */
  int32_t nb_fault;
  int32_t nb_element;                /* N, fake elements included */
  const int64_t *fault_nb_element;   /* [nb_fault] faults(i)%nb_element */
  const double *node;                /* [2*(N+1)] node(2,N+1), (x,y) pairs */
  const double *a, *b, *Dc, *f0, *V0;                    /* [N] */
  const double *normal_loading_dot, *shear_loading_dot;  /* [N] */
  const double *V, *theta, *sigmaN, *P;                  /* [N] initial state, theta = psi (log state) */
  const double *permeability, *kmin, *kmax, *Snk, *Lk, *Tk; /* [N] */
  double mu, cs, cp;
  double Sigma11_dot, Sigma12_dot, Sigma22_dot, Sigma31_dot, Sigma32_dot;
  double porosity, fluid_comp, rock_comp, dyn_viscosity, fluid_density;
  double tol_solver;
  int32_t permeability_coupling;     /* 0 or 1 */
  int32_t nb_source;                 /* total number of fluid sources */
  const int64_t *source_fault;       /* [nb_source] 1-based fault id */
  const int64_t *source_index;       /* [nb_source] 1-based element index within that fault */
  const double *source_Q, *source_t_beg, *source_t_end; /* [nb_source] */
  /* H-matrix parameters (reference hard-codes nleaf=100, eta=3, eps_ACA=1e-8,
   * src/variables.f90:122-124, and max rank 20, H_matrix_construction.f90:59) */
  double eta, eps_aca;
  int32_t nleaf, max_rank;
} dfn_problem;

typedef struct dfn_handle_s *dfn_handle;

/* Statistics of a dfn_solve call (accumulated over the handle's life). */
typedef struct dfn_run_stats {
  int64_t accepted_steps;   /* it_count of src/solver.f90:60 */
  int64_t rejected_attempts;
  int64_t rhs_evaluations;  /* friction_law calls */
  int64_t cg_solves;        /* compute_P calls */
  int64_t cg_iterations;    /* CG iterations summed over all compute_P calls */
  int64_t cg_halvings;      /* t_try halvings after CG non-convergence (solver.f90:88-92) */
  double time;              /* simulated time reached (s) */
  double t_try;             /* next trial step (s) */
  double max_V;             /* max over elements of V at the end */
} dfn_run_stats;

/* ---- lifecycle ----------------------------------------------------------- */
/* Replaces config()+initialise() (src/configuration.f90:19, src/initialisation.f90:24):
 * copies the problem, computes element midpoints/ds/tangent/normal and the loading
 * projections (initialisation.f90:139-206), fault crossings (:279-344), the initial
 * permeability_star/permeability_x (:252-266), and uploads the state to the device. */
int dfn_create(const dfn_problem *problem, dfn_handle *out);
int dfn_destroy(dfn_handle h);

/* Replaces initiate_hmat (src/hmatrix/initialisation_hmat.f90:13-100) for the
 * problem's mode: Mode III builds the shear operator, Mode II the normal then the
 * shear operator (src/initialisation.f90:210-248).  No-op for the direct kernel. */
int dfn_build_hmat(dfn_handle h);
/* Operator storage precision for subsequent kernel calls (DFN_PREC_FP64 default):
 * DFN_PREC_FP32: factors and near-field blocks in FP32 (per-block power-of-two
 *   scales), FP32 products and sums, compensated on near-field columns;
 * DFN_PREC_MIXED: FP64 near-field blocks beside FP32 low-rank factors (stored as
 *   for DFN_PREC_FP32) read into FP64 products and sums;
 * DFN_PREC_FP32_STORE: the FP32 storage of DFN_PREC_FP32 (same bytes) read into
 *   FP64 products and sums, so that only the storage rounds;
 * DFN_PREC_MIXED_SPLIT: as DFN_PREC_MIXED, with each near-field entry stored in
 *   5 bytes instead of 8: its FP32 rounding and an 8-bit residual times a scale
 *   per panel task (about 31 significant bits for a task's largest entries, at
 *   least FP32's 24 for every entry).
 * No reference counterpart (the reference is FP64 throughout). */
int dfn_set_precision(dfn_handle h, int32_t precision);
/* stats[0..9] = {nb_dense, nb_lowrank, S_dense, S_lowrank_U, S_lowrank_V, max_rank,
 *                nbclus, nbblclus, n, sum_rank} of operator op. */
int dfn_hmat_stats(dfn_handle h, int32_t op, int64_t *stats);
/* Export an operator as flat block lists (H order, 0-based).  Buffers sized from
 * dfn_hmat_stats.  perm[k] = original element index of H position k (mat_corr - 1).
 * Dense block b: rows dense_row0[b]..+dense_m[b], cols dense_col0[b]..+dense_n[b],
 * column-major at dense_data + dense_off[b].  Low-rank block b = U V^T with
 * U (m x k) at lr_u + lr_uoff[b], V (n x k) at lr_v + lr_voff[b], column-major. */
int dfn_hmat_export(dfn_handle h, int32_t op, int32_t *perm,
                    int32_t *dense_row0, int32_t *dense_col0, int32_t *dense_m, int32_t *dense_n,
                    int64_t *dense_off, double *dense_data,
                    int32_t *lr_row0, int32_t *lr_col0, int32_t *lr_m, int32_t *lr_n, int32_t *lr_k,
                    int64_t *lr_uoff, int64_t *lr_voff, double *lr_u, double *lr_v);
/* Replace operator op by caller-provided block lists (same layout as the export). */
int dfn_hmat_import(dfn_handle h, int32_t op, int32_t n, const int32_t *perm,
                    int32_t nb_dense, const int32_t *dense_row0, const int32_t *dense_col0,
                    const int32_t *dense_m, const int32_t *dense_n, const int64_t *dense_off,
                    const double *dense_data,
                    int32_t nb_lowrank, const int32_t *lr_row0, const int32_t *lr_col0,
                    const int32_t *lr_m, const int32_t *lr_n, const int32_t *lr_k,
                    const int64_t *lr_uoff, const int64_t *lr_voff,
                    const double *lr_u, const double *lr_v);

/* ---- state ---------------------------------------------------------------- */
int dfn_set_state(dfn_handle h, const double *Y, const double *P, double time);
int dfn_get_state(dfn_handle h, double *Y, double *P, double *P_dot, double *time);
int dfn_get_stats(dfn_handle h, dfn_run_stats *stats);

/* ---- the reference's hot-path interfaces ------------------------------------ */
/* kernel procedure pointer (src/variables.f90:17-26): stress rates from slip rate.
 * Replaces hmatIII (hmatrixModeIII.f90:9-60), hmatII (hmatrixModeII.f90:9-74),
 * directIII (directModeIII.f:7-74), directII (directModeII.f90:7-107).
 * stressN may be NULL in Mode III (the reference leaves it untouched there). */
int dfn_kernel(dfn_handle h, const double *slip, double *stressT, double *stressN);

/* dfn_kernel over nrhs slip-rate vectors at once (H-matrix handles, FP64 operators
 * whatever dfn_set_precision chose): slip holds nrhs vectors of N doubles one after
 * another, stressT (and stressN in Mode II) likewise.  The reference applies its
 * H-matrix to one vector per call (Hgemv_right, src/hmatrix/calculation_hmat.f90:
 * 122-159, ZGEMV per block); here both phases are dense products on the matrix
 * cores over tiles of 16 vectors, in batches of up to 1024 vectors.  Each output
 * agrees with dfn_kernel's for the same vector to rounding (the sums run in
 * another order).  A dense direct handle gives DFN_ERR_STATE. */
int dfn_kernel_multi(dfn_handle h, int32_t nrhs, const double *slip, double *stressT, double *stressN);

/* dfn_kernel without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops); no counterpart in
 * the reference.  Dense direct kernel handles only (DFN_KERNEL_DIRECT; an H-matrix
 * handle gives DFN_ERR_UNSUPPORTED).  slip, stressT and stressN (Mode II; may be
 * NULL in Mode III) are DEVICE pointers of N doubles; the kernels are queued on
 * `stream` (a hipStream_t; NULL = the default stream) and the call returns without
 * waiting for them.  The per-tile partial sums live in a fixed 16 MiB device array
 * per GPU (ceil((N+1)/64) * N doubles per operator must fit in 2^21 doubles:
 * N <= 11 583 in Mode III, N <= 8 191 in Mode II; larger N gives DFN_ERR_ARG and a
 * message on stderr).  Every call on one GPU shares that array: calls that overlap
 * on one GPU (other streams, other threads, other handles) race. */
int dfn_kernel_noalloc(dfn_handle h, const double *slip, double *stressT, double *stressN, void *stream);

/* y <- alpha*H*x + beta*y for operator op in H order (Hgemv_right,
 * src/hmatrix/calculation_hmat.f90:3-165), x and y of length N. */
int dfn_hgemv(dfn_handle h, int32_t op, double alpha, double beta, const double *x, double *y);

/* friction_law procedure pointer (src/variables.f90:9-14): Y_dot = f(x, Y) with the
 * pore pressure of compute_P(x) (RateStateAgingModeIII.f90:19-121, ModeII.f90:19-92). */
int dfn_friction_law(dfn_handle h, double x, const double *Y, double *Y_dot);

/* compute_P (src/fluid_diffusion.f90:24-152): backward-Euler pressure after dt from the
 * handle's step-start pressure.  P_temp/P_dot may be NULL.  *ier = 1 on CG failure
 * (then P_temp and P_dot are not written, as in the reference). */
int dfn_compute_P(dfn_handle h, double dt, double *P_temp, double *P_dot, int32_t *ier);

/* rkqs (src/ode_solver/ode_bs.f90:125-174) with the handle's friction law as derivs.
 * Returns DFN_ERR_UNDERFLOW on 'stepsize underflow'. */
int dfn_rkqs(dfn_handle h, double *y, const double *dydx, double *x, double htry,
             double eps, const double *yscal, double *hdid, double *hnext);

/* The solver loop of src/solver.f90:59-198 on the handle's device-resident state:
 * runs until nsteps more accepted steps are done, or until time reaches t_final
 * (t_final <= 0: no time limit; the reference stops on max_it only).  With a time
 * limit, a trial step that would pass t_final is shortened to end on it, so the
 * run lands on t_final exactly.  The loop state (time, t_try, counters) persists
 * between calls. */
int dfn_solve(dfn_handle h, int64_t nsteps, double t_final, dfn_run_stats *stats);

/* Launch shapes the autotuner picked for the active operator layout (no reference
 * counterpart): choice = {phase-1 tasks per workgroup (2, 4, 8), phase-1 task order
 * (0 as built, 1 grouped by rank bucket, largest ranks first, 2 largest tasks
 * first), panel rows per wave step (2, 4), panel task order (0 as built, 1 largest
 * panels first), rows per panel task (0 whole panels, 32, 16), phase 1 and the
 * FP64 near-field panels of a DFN_PREC_MIXED layout in one launch (1) or two (0),
 * the rank ceiling of the phase-1 kernel instance (16, 20 or 32), phase-1 tasks of
 * at most 32 row vectors two to a wave (1) or one (0)} (8 entries);
 * ms = {matvec time of the default shapes, of the chosen ones}.  None of the
 * choices changes a result.  DFN_AUTOTUNE=0 keeps the defaults, DFN_AUTOTUNE=alt
 * takes a set of non-default shapes without timing; DFN_FUSE_NEAR=1 or 0 takes or
 * refuses the single launch of entry 5 whatever its timing, DFN_VTX_PAIRS=1 or 0
 * the wave items of entry 7. */
int dfn_tuning(dfn_handle h, int32_t *choice, double *ms);

/* ---- Newton-V form ------------------------------------------------------------
 * No counterpart in the reference, which integrates the slip rate V through its
 * time derivative (src/friction/RateStateAgingModeII.f90:61-66).  The Newton-V
 * form integrates slip and log state instead and solves, at every evaluation and
 * element, the stress balance whose derivative that is,
 *   tau_ref + shear_loading_dot t + K slip = eta V - (sigma_N + P) a asinh(V C / (2 V0)),
 * for V (C = exp((f0 + b psi) / a); sigma_N = sigma_ref + normal_loading_dot t, plus
 * K_N slip in Mode II; t and slip counted from the switch, whose state fixes
 * tau_ref and sigma_ref).  Same step controller and pressure solves; the k1
 * evaluation is at x = t_try as in the reference.  dfn_get_state returns the state
 * in the reference layout [V | psi | sigma_N]; dfn_set_state takes one.  Not with
 * the event monitor, solver output, dfn_friction_law or dfn_rkqs (DFN_ERR_STATE). */
#define DFN_FORM_REFERENCE 0
#define DFN_FORM_NEWTON_V 1
int dfn_set_formulation(dfn_handle h, int32_t form);

/* ---- event catalog and moment ------------------------------------------------
 * The inputs of the reference's post-processing, gathered on the device as the
 * solver runs (the reference writes V every stride_time steps and post-processes
 * in MATLAB).  With the monitor on, every accepted step of dfn_solve records the
 * time and, per fault, max |V| over the fault's elements (the series that
 * wrapper/microseismicity.m:447-472 thresholds into an event catalog), and adds
 * (V_old + V_new) h / 2 to each element's slip (cumtrapz of V over the accepted
 * steps, wrapper/network.m:474-476).  Turning it on zeroes the slip and drops
 * unread records. */
int dfn_monitor(dfn_handle h, int32_t on);
/* Moves up to cap records out: times[i] and maxV[i*nb_fault + f]; with times or maxV
 * NULL, *count = records available. */
int dfn_monitor_take(dfn_handle h, int64_t cap, double *times, double *maxV, int64_t *count);
/* Seismic moment per fault M0[f] = sum over its elements of mu * slip * ds
 * (wrapper/network.m:478-483) and, if slip is not NULL, the slip of every element. */
int dfn_moment(dfn_handle h, double *M0, double *slip);

/* ---- reading the reference's HDF5 input files -------------------------------
 * The reference reads its configuration through libhdf5 (src/configuration.f90:
 * 70-379).  This is a reader, without libhdf5, for the subset its input files use:
 * superblock version 0, symbol-table groups, version-1 object headers,
 * contiguous or compact integer, floating-point and string (fixed or variable
 * length) datasets.  Paths are like "fault1/a"; other layouts return
 * DFN_ERR_UNSUPPORTED, a missing path DFN_ERR_ARG. */
typedef struct dfn_h5_s *dfn_h5;
int dfn_h5_open(const char *path, dfn_h5 *out);
int dfn_h5_close(dfn_h5 f);
/* type_class: 0 integer, 1 floating point, 2 string; dims holds up to 8 extents */
int dfn_h5_info(dfn_h5 f, const char *path, int32_t *type_class, int32_t *rank, int64_t *dims);
/* n must equal the dataset's element count; integers and floats convert */
int dfn_h5_read_f64(dfn_h5 f, const char *path, double *buf, int64_t n);
int dfn_h5_read_i64(dfn_h5 f, const char *path, int64_t *buf, int64_t n);
/* the (first) string of a string dataset, NUL-terminated, at most cap bytes */
int dfn_h5_read_string(dfn_h5 f, const char *path, char *buf, int32_t cap);
/* Writes a new file at path whose root group holds n FP64 datasets (n <= 8,
 * distinct names, rank 1 or 2, C order; dims holds two extents per dataset, the
 * second unused at rank 1), in the subset above and with the header messages
 * libhdf5 writes for them: the format of the reference's output files
 * (write_data, src/solver.f90:227-317). */
int dfn_h5_write_f64(const char *path, int32_t n, const char *const *names, const int32_t *ranks,
                     const int64_t *dims, const double *const *data);

/* ---- solver output ----------------------------------------------------------
 * save_data / write_data of src/solver.f90:180-193, 211-317.  With an output
 * directory set, dfn_solve saves, after each accepted step it_count (the handle's
 * accepted-step count) with mod(it_count - 1, stride_time) == 0, V, theta =
 * Dc/V0 exp(psi), P, the last shear stress rate of the kernel (SigmaT_dot),
 * sigma_N, the permeability and the time, on the device; once freq_writing_file
 * saves are held it writes dir/outputNNNNNNN (NNNNNNN = it_count, 7 digits) with
 * datasets V, theta, P, tractionTel, tractionNel, permeability of C-order shape
 * (N, freq_writing_file) and time (freq_writing_file), as the reference does.
 * dir = NULL ends the output and drops unwritten saves. */
int dfn_set_output(dfn_handle h, const char *dir, int64_t stride_time, int64_t freq_writing_file);
/* The write_data after the loop (src/solver.f90:202): a file with the saves held,
 * shape (N, saves); nothing when none are held. */
int dfn_flush_output(dfn_handle h);

/* ---- measurement --------------------------------------------------------- */
/* Time `reps` applications of the stress-transfer kernel on device-resident input
 * (the handle's current V) with HIP events on the handle's stream.  ms[0] = mean ms
 * per application of the whole kernel (all operators), ms[1] = mean ms of the
 * dominant device kernel alone (H-matrix: the phase-2 panel kernel, or phase 1
 * with the near-field panels when they run as one launch), ms[2] =
 * algorithmic bytes per application, ms[3] = algorithmic bytes of the dominant
 * kernel per launch.  ms must hold 4 doubles. */
int dfn_bench_kernel(dfn_handle h, int32_t reps, double *ms);
/* Per-launch timing of the dominant kernel of every stress evaluation (the
 * phase-2 panel kernel of the H-matrix path, or phase 1 with the near-field panels
 * when they run as one launch, or the direct kernel) while the
 * solver runs, with a HIP event pair around each launch on the handle's stream (no
 * reference counterpart).  dfn_kernel_timing(h, 1) clears the sums and starts,
 * 0 stops; dfn_kernel_timing_read gives the mean duration in ms and the number of
 * launches timed since the last start. */
int dfn_kernel_timing(dfn_handle h, int32_t on);
/* Times dfn_kernel_multi's device work on nrhs (<= one batch) host vectors already
 * copied to the device: out[0] = ms per whole multi-RHS application, out[1] = ms
 * per phase-2 (panel product) launch, out[2] = its flops (2 x panel entries, the
 * dense and U ones, x vectors rounded up to 64), out[3] = algorithmic HBM bytes of one application
 * (every stored entry once, the vectors read and written once).  No reference
 * counterpart. */
int dfn_bench_kernel_multi(dfn_handle h, int32_t nrhs, const double *slip, int32_t reps, double *out);
int dfn_kernel_timing_read(dfn_handle h, double *ms_mean, int64_t *count);
/* Measurement helper: a plain device-to-device copy kernel over `bytes` bytes,
 * run `reps` times on a private stream; *gbps = 2*bytes / mean time (read + write).
 * Used to scale hardware traffic counters and to state the achievable HBM rate. */
int dfn_bench_stream(int64_t bytes, int32_t reps, double *gbps);
/* Device/library description string (NUL-terminated, at most len bytes). */
int dfn_device_info(char *buf, int32_t len);

/* ---- step-size proposal network ------------------------------------------
 * No counterpart in the reference (its step size comes from rkqs alone,
 * src/ode_solver/ode_bs.f90:125-174).  An 8 -> 64 -> 64 -> 1 perceptron with
 * GELU hidden layers, BF16 weights and FP32 accumulation.  Parameters (FP32
 * master copy, DFN_MLP_PARAMS values): W1 (64 x 8, row-major), b1 (64), W2
 * (64 x 64), b2 (64), w3 (64), b3.  norm (18 values): feature means (8), feature
 * scales (8), output mean, output scale; the network sees (x - mean) / scale and
 * its output y maps to mean_y + scale_y * y.  Loss: mean of (y - (t - mean_y) /
 * scale_y)^2. */
#define DFN_MLP_INPUTS 8
#define DFN_MLP_PARAMS 4801
typedef struct dfn_mlp_s *dfn_mlp;
int dfn_mlp_create(dfn_mlp *out);  /* He-initialised with a fixed seed, identity normalisation */
int dfn_mlp_destroy(dfn_mlp m);
int dfn_mlp_set_params(dfn_mlp m, const float *params, const float *norm); /* norm may be NULL; resets Adam */
int dfn_mlp_get_params(dfn_mlp m, float *params, float *norm);
/* Y[i] = network(X[i*8 .. i*8+7]) for host arrays, n samples */
int dfn_mlp_forward(dfn_mlp m, const float *X, int64_t n, float *Y);
/* loss and its gradient (DFN_MLP_PARAMS values, with respect to the BF16 weights
 * the forward pass uses) over the n samples (X, targets T) */
int dfn_mlp_loss_grad(dfn_mlp m, const float *X, const float *T, int64_t n, double *loss, float *grad);
/* Adam (beta 0.9 / 0.999, eps 1e-8) over `epochs` shuffled passes of whole
 * mini-batches of `batch` samples; *final_loss (may be NULL) = loss over all n */
int dfn_mlp_train(dfn_mlp m, const float *X, const float *T, int64_t n, int32_t epochs, int32_t batch, float lr,
                  uint64_t seed, double *final_loss);

/* ---- ensembles of independent networks ------------------------------------
 * No single counterpart in the reference, which runs one network per process
 * (src/main.f90); each network of an ensemble follows the step of
 * src/solver.f90:59-198 exactly as a dfn_handle would, with its own clock,
 * trial step and accept/reject decisions, while the stages of all networks run
 * as shared launches on one GPU.  All networks must use the H-matrix kernel,
 * the same fracture mode and elastic constants (one operator scale), and
 * permeability_coupling != 1. */
typedef struct dfn_ensemble_s *dfn_ensemble;

/* problems[0..nnet-1] as for dfn_create; the arrays may be freed after the call. */
int dfn_ensemble_create(const dfn_problem *problems, int32_t nnet, dfn_ensemble *out);
int dfn_ensemble_destroy(dfn_ensemble e);
/* Builds every network's operators (as dfn_build_hmat) and the block-diagonal
 * device layout of the whole ensemble. */
int dfn_ensemble_build_hmat(dfn_ensemble e);
/* Operator storage precision (DFN_PREC_*), as dfn_set_precision. */
int dfn_ensemble_set_precision(dfn_ensemble e, int32_t precision);
/* 1 (default): the per-attempt launch sequences are captured as HIP graphs and
 * replayed; 0: the same launches are issued one by one.  Ensembles with networks
 * above 4096 elements solve their pressure systems with cooperative launches,
 * which are always issued one by one (the request is ignored there). */
int dfn_ensemble_set_graph(dfn_ensemble e, int32_t use_graph);
/* Every network advances nsteps accepted steps. */
int dfn_ensemble_solve(dfn_ensemble e, int64_t nsteps);
/* Steps every network until its time reaches t_end exactly: a network's trial
 * step is clipped to land on t_end (a retried attempt stops short and the
 * network steps again), and networks already at t_end sit out.  Fails with
 * DFN_ERR_STATE if max_steps ensemble steps do not bring every network there. */
int dfn_ensemble_solve_until(dfn_ensemble e, double t_end, int64_t max_steps);
/* State of network `net` in its own layout: Y = [V | theta | sigmaN] (3*N), P, P_dot
 * (N each), time.  NULL pointers are skipped. */
int dfn_ensemble_get_state(dfn_ensemble e, int32_t net, double *Y, double *P, double *P_dot, double *time);
int dfn_ensemble_get_stats(dfn_ensemble e, int32_t net, dfn_run_stats *stats);
/* Step-size control of every network after an accepted step:
 * DFN_STEP_RKQS (default) rkqs's proposal (src/ode_solver/ode_bs.f90:165-169),
 * the reference's; DFN_STEP_PI a PI controller (exponents 0.7/5 and 0.4/5 on the
 * current and previous scaled error, growth within [0.2, 5]); DFN_STEP_LEARNED
 * rkqs's proposal times 10^y, y from the proposal network `m` (clamped to
 * [-1, log10 2]), with the PI controller for 3 steps after a learned proposal
 * was rejected.  Every step is still accepted or rejected by the Cash-Karp error
 * test, and the one-day cap and the injection guard of src/solver.f90 apply. */
#define DFN_STEP_RKQS 0
#define DFN_STEP_PI 1
#define DFN_STEP_LEARNED 2
int dfn_ensemble_set_controller(dfn_ensemble e, int32_t mode, dfn_mlp m);
/* Records (features, target) pairs for the proposal network after every step:
 * features of network i after accepted step n: max over its real elements of
 * log10 |V|, their mean log10 |V|, the rate of change of max log10 |V| over the
 * step (per second), log10 of the scaled error (errmax) of steps n, n-1 and n-2,
 * log10 of the step, and the fraction of elements with |V| > 1e-3 m/s; target
 * log10(accepted step n+1 / rkqs proposal after step n).  Steps set by the
 * one-day cap, the injection guard or a final time are left out. */
int dfn_ensemble_set_recording(dfn_ensemble e, int32_t on);
/* Moves up to cap recorded pairs (X: 8 floats each, T: 1) out of the ensemble;
 * with X or T NULL, *count = pairs available. */
int dfn_ensemble_take_records(dfn_ensemble e, float *X, float *T, int64_t cap, int64_t *count);
/* info[0..8] = networks, total elements, largest network, matrix entries streamed
 * per stress evaluation, algorithmic bytes per stress evaluation, graphs on, phase 1
 * and the FP64 near-field panels in one launch (1) or two (0), the rank ceiling of
 * the phase-1 kernel instance in use (16, 20 or 32), short phase-1 tasks two to a
 * wave (1) or one (0). */
int dfn_ensemble_info(dfn_ensemble e, int64_t *info);
/* Per-launch timing of the dominant kernel while the ensemble steps (no reference
 * counterpart): the phase-2 panels, or phase 1 with the near-field panels when
 * they run as one launch (dfn_ensemble_info entry 6).  With timing on, every
 * stress evaluation in which all networks take part records a HIP event pair
 * around that launch on the ensemble's stream (the launches are then issued one
 * by one).  1 clears the
 * sums and starts, 0 stops; _read gives the mean duration in ms, the launches
 * timed and the kernel's algorithmic bytes per such launch. */
int dfn_ensemble_kernel_timing(dfn_ensemble e, int32_t on);
int dfn_ensemble_kernel_timing_read(dfn_ensemble e, double *ms_mean, int64_t *count, double *bytes);

#ifdef __cplusplus
}
#endif
#endif /* DFN_H */
