/* diskfft on MI355X -- C ABI of the HIP time-step library (libdiskfft.so).
 *
 * Plain C: opaque handle, plain pointers and sizes, int status (DFK_OK = 0),
 * no exceptions or torch types across the boundary.  Host pointers are copied
 * in or out during the call and never retained.  One handle owns one HIP
 * stream on one device and holds one or more disks ("an ensemble"; a single
 * disk is an ensemble of one).  A handle must not be used from two host
 * threads at once.  Every compute call needs a HIP device; without one the
 * call fails with DFK_EHIP and dfk_last_error() says why (there is no CPU
 * fallback).
 *
 * Reference interfaces replaced (paths relative to the oklo/heavy-disks tree):
 *   DiskHydro                 diskfft/src/hydro.hpp:20-180
 *   KalnajsSolver             diskfft/src/kalnajs.hpp:19-117
 *   the diskdisk driver loop  diskfft/src/diskmain.cpp:16-115
 */
#ifndef DISKFFT_H
#define DISKFFT_H

#ifdef __cplusplus
extern "C" {
#endif

#define DFK_OK 0
#define DFK_EINVAL 1        /* bad argument (sizes, null pointers, disk index) */
#define DFK_EHIP 2          /* HIP runtime error or no device                  */
#define DFK_ENOMEM 3        /* device allocation failed                        */
#define DFK_EUNSUPPORTED 4  /* transform length with more than 16 prime factors */
#define DFK_ESTATE 5        /* call out of order (e.g. advance before run_init)*/

/* Constructor arguments of DiskHydro (hydro.hpp:28-29) and KalnajsSolver
 * (kalnajs.hpp:25-26) for one disk. */
typedef struct {
  int Nu, Nphi;          /* radial (log-spaced) and azimuthal zones; 3 <= Nu <= 2048, 4 <= Nphi <= 4096,
                            any factorisation (large prime factors make the transforms slow) */
  double Rin, Rout;      /* grid radii */
  double K, gamma;       /* barotropic EOS P = K Sigma^gamma */
  double gmstar;         /* G M_star */
  double eps, G;         /* Kalnajs softening (scale-free) and G */
  int selfgrav;          /* 0: star only; 1: Kalnajs self-gravity (diskmain.cpp:22,103-108) */
} dfk_disk_params;

/* Loop parameters for one disk (diskmain.cpp:88,92-101). */
typedef struct {
  double tlim;           /* loop ends when t >= tlim; last step clipped to land on tlim */
  double dtout;          /* output interval (reference: 0.25) */
  double cfl;            /* CFL number (reference: 0.3) */
  double mD;             /* mass normaliser of |C_m| (diskmain.cpp:65-70, 82) */
  long long max_steps;   /* stop after this many steps, <= 0: no cap */
  int stop_on_nonfinite; /* 1: end the disk's loop when a NaN/Inf zone appears */
} dfk_run_params;

/* One output record, as diskmain.cpp:93-98 prints it: t, |C0|, |C2|, arg C2. */
typedef struct { double t, c0, c2, ph2; } dfk_sample;

typedef struct {
  double t;              /* current time */
  long long steps;       /* steps taken */
  int done;              /* loop ended */
  int flags;             /* bit 0: a non-finite zone was seen */
  double c2_final;       /* |C2| at loop end (diskmain.cpp:111) */
  double ph2_final;
} dfk_status;

typedef struct dfk_handle dfk_handle;

/* Last error message of the calling thread ("" if none). */
const char *dfk_last_error(void);

/* Number of visible HIP devices (0 when none; never fails). */
int dfk_device_count(void);

/* Creates ndisks disks on `device` (DiskHydro ctor hydro.hpp:28-43 and
 * KalnajsSolver ctor kalnajs.hpp:25-42, including the K-hat build
 * kalnajs.hpp:105-116).  State is zeroed. */
int dfk_create(const dfk_disk_params *params, int ndisks, int device, dfk_handle **out);
void dfk_destroy(dfk_handle *h);
int dfk_num_disks(const dfk_handle *h);

/* Changes the EOS constants (P = K Sigma^gamma) of one disk after creation;
 * used when K comes out of an equilibrium solve that needed this handle's
 * gravity first.  Takes effect for the next dfk_run_init / unit call. */
int dfk_set_eos(dfk_handle *h, int disk, double K, double gamma);

/* Geometry tables of DiskHydro: R (Nu), Rf (Nu+1), Vol (Nu) (hydro.hpp:31-38). */
int dfk_geometry(dfk_handle *h, int disk, double *R, double *Rf, double *Vol);

/* Public state vectors Sig, vR, vp of DiskHydro (hydro.hpp:24), Nu*Nphi each. */
int dfk_set_state(dfk_handle *h, int disk, const double *Sig, const double *vR, const double *vp);
int dfk_get_state(dfk_handle *h, int disk, double *Sig, double *vR, double *vp);

/* DiskHydro::timestep (hydro.hpp:48-61).  Like dfk_modes and dfk_disk_mass
 * it reads the current state only: called during or after a run, it leaves
 * the disk's loop clock, step count and later steps as they are. */
int dfk_timestep(dfk_handle *h, int disk, double cfl, double *dt);

/* KalnajsSolver::potential (kalnajs.hpp:48-63): Phi from Sigma (Nu*Nphi). */
int dfk_potential(dfk_handle *h, int disk, const double *Sigma, double *Phi);

/* dfk_potential on device memory, allocating nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Sigma and Phi are device pointers on the handle's device (Nu*Nphi doubles,
 * row-major); stream is a hipStream_t (NULL: the device's null stream).  The call
 * only enqueues the three transform kernels on that stream and returns; Phi
 * is ready when the stream reaches that point.  Same kernels and bits as
 * dfk_potential.  The half spectrum lives in a scratch area of the library
 * (one per device, held until the process ends) of DFK_NOALLOC_MAX_SPECTRUM
 * complex entries: a disk with Nu*(Nphi/2+1) above that gets DFK_EINVAL.
 * Calls that overlap on one device (two streams, two threads, two handles)
 * race on that scratch; calls ordered on one stream do not.  Needs a handle
 * whose last call succeeded (shapes set, K-hat built); it does not touch the
 * handle's run, state or loop clock. */
#define DFK_NOALLOC_MAX_SPECTRUM 1048576
int dfk_potential_noalloc(dfk_handle *h, int disk, const double *Sigma, double *Phi, void *stream);

/* KalnajsSolver::forces (kalnajs.hpp:65-80). */
int dfk_forces(dfk_handle *h, int disk, const double *Phi, double *gR, double *gphi);

/* DiskHydro::step (hydro.hpp:63-70); gR / gphi may be NULL as in the reference. */
int dfk_step(dfk_handle *h, int disk, double dt, const double *gR, const double *gphi);

/* One stage of DiskHydro::step: 1 source (hydro.hpp:73-96, gR/gphi may be NULL),
 * 2 transportR (:99-140), 3 transportPhi (:143-179), 4 floor (:68-69). */
int dfk_stage(dfk_handle *h, int disk, int stage, double dt, const double *gR, const double *gphi);

/* Complex K-hat (M x (Nphi/2+1), row-major, re/im split) as built by
 * KalnajsSolver::build_kernel (kalnajs.hpp:105-116); for checks. */
int dfk_khat(dfk_handle *h, int disk, double *re, double *im);

/* The C_m diagnostic of diskmain.cpp:72-84 on the current state. */
int dfk_modes(dfk_handle *h, int disk, int m, double mD, double *mag, double *phase);

/* Disk mass sum Sigma R dR dphi of the current state (diskmain.cpp:65-70). */
int dfk_disk_mass(dfk_handle *h, int disk, double *mD);

/* The time loop of diskmain.cpp:92-113, device resident.  run_init resets
 * t = 0 and the output schedule for every disk (params: one per disk).
 * advance runs up to `nsteps` loop iterations on every unfinished disk, as a
 * captured HIP graph, and sets *all_done.  samples drains the records produced
 * since the last call (at most one per iteration, so drain after every
 * advance of more than `cap_hint` steps; see dfk_run_init). */
int dfk_run_init(dfk_handle *h, const dfk_run_params *params, int sample_cap);
int dfk_advance(dfk_handle *h, int nsteps, int *all_done);
int dfk_samples(dfk_handle *h, int disk, dfk_sample *buf, int cap, int *n);
int dfk_get_status(dfk_handle *h, int disk, dfk_status *st);

/* Snapshot / resume (no counterpart in the reference, which writes only the
 * diagnostic series).  save writes disk `disk`'s constructor parameters, loop
 * clock (t, next output time, steps), loop parameters (tlim, dtout, cfl, mD,
 * max_steps, stop_on_nonfinite) and current state (Sigma, vR, vphi) to a
 * binary file.  load, called after dfk_run_init (at any point of a run, on
 * any disk of a handle), restores state, clock and loop parameters of a disk
 * with the same grid and constructor parameters (DFK_EINVAL otherwise), in
 * place of those dfk_run_init set for it; the run then continues exactly as
 * the saved one would have (dt is a function of the state only).  The disk's
 * undrained samples are discarded by load: its sample series restarts with
 * the first output after the restored time. */
int dfk_snapshot_save(dfk_handle *h, int disk, const char *path);
int dfk_snapshot_load(dfk_handle *h, int disk, const char *path);

/* Tile / transform-batch shapes (the autotuner's knobs): update tile ti x tj,
 * complex rows per workgroup in the row transforms (pr) and columns per
 * workgroup in the column convolution (cb).  0 keeps the current value.
 * Takes effect for the next dfk_run_init / unit call. */
int dfk_set_shapes(dfk_handle *h, int disk, int ti, int tj, int pr, int cb);
int dfk_get_shapes(dfk_handle *h, int disk, int *ti, int *tj, int *pr, int *cb);

/* Transform options of a disk (autotuner knobs like dfk_set_shapes; no
 * counterpart in the reference, whose FFTW plans make these choices):
 *   rows_tw_lds, cols_tw_lds  where the row transforms / the column
 *       convolution read their twiddle table: 1 = staged in LDS per
 *       workgroup (default; fastest when few workgroups run), 0 = from HBM
 *       through L1/L2 (less LDS per workgroup, more workgroups per CU);
 *       transforms too long for a table beside their buffers in LDS
 *       (Nphi or 2Nu above 3413 on gfx950) always read it through the caches;
 *   cols_wave  how the column convolution is spread over threads:
 *       0 = the workgroup cooperates on its columns with two LDS buffers per
 *       column (default); 1 = each column is transformed by one wavefront,
 *       in place in one LDS buffer, without workgroup barriers; 2 = the
 *       workgroup transforms its columns one after another, in place in a
 *       single LDS buffer.  1 and 2 need 2Nu <= 1024 with radices 2, 3, 4,
 *       8 (otherwise the value is ignored and read back as 0); 3 = a two-pass
 *       register transform per direction (2Nu = N1 N2, in-register N1- and
 *       N2-point DFTs, one LDS transpose between them, 64 / N2 columns per
 *       wavefront), for 2Nu in {128, 192, 256, 384, 512, 768, 1024} only
 *       (otherwise read back as 0); it ignores cb and cols_tw_lds; 4 = three
 *       passes of at most 8-point in-register DFTs per direction, one column
 *       per wavefront (shorter per-lane chains for small, latency-bound
 *       grids), for 2Nu in {256, 384, 512, 768, 1024} only (otherwise read
 *       back as 0), ignoring cb and cols_tw_lds.
 * The twiddle placement does not change a bit of the results; the in-place
 * and register kernels agree with the cooperative one to round-off. */
int dfk_set_fft_options(dfk_handle *h, int disk, int rows_tw_lds, int cols_tw_lds, int cols_wave);
int dfk_get_fft_options(dfk_handle *h, int disk, int *rows_tw_lds, int *cols_tw_lds,
                        int *cols_wave);

/* Threads per workgroup (64, 128, 192 or 256; 0 keeps the current value) of
 * the row transforms, the column convolution and the update kernel, for all
 * disks of the handle (a batched launch has one workgroup size). */
int dfk_set_threads(dfk_handle *h, int rows, int cols, int update);
int dfk_get_threads(dfk_handle *h, int *rows, int *cols, int *update);

/* Row transforms of all disks of the handle (an autotuner knob; no
 * counterpart in the reference): 0 = the workgroup cooperates on its pr
 * complex rows with two LDS buffers (default); 1 = in place in one LDS buffer
 * (half the LDS per row; pr x Nphi is then kept <= max(8 x threads, 1024));
 * 2 = two-pass register transforms (one 64-thread wavefront per 4 row pairs
 * for Nphi <= 256, per 2 above; pr and the row threads do not apply);
 * 3 = three-pass register transforms (one wavefront per row pair, at most
 * 8-point DFTs per lane: small, latency-bound grids).  In place needs every
 * self-gravitating disk's Nphi to factor into 2, 3, 4 and 8, register every
 * such Nphi to be 128, 192, 256, 384, 512, 768 or 1024, three-pass register
 * 128, 192, 256, 384 or 512; otherwise the handle takes the next of three-pass ->
 * register -> in place -> cooperative that all its disks allow, and
 * dfk_get_row_transform reports the variant in use.  All agree to
 * round-off. */
int dfk_set_row_transform(dfk_handle *h, int inplace);
int dfk_get_row_transform(dfk_handle *h, int *inplace);

/* Update-kernel variant for all disks of the handle: 0 = LDS tile (ti x tj
 * zones per workgroup; suits small, latency-bound runs), 1 = register
 * streaming (one wavefront per 58 columns x ti rows; suits large or batched,
 * bandwidth-bound runs).  Both compute the same step (to round-off; dt to the
 * bit in both).  Kind 1 addresses a state buffer with 31-bit offsets: a
 * handle of more than 89 478 485 zones (3 fields x 8 bytes each within
 * 2^31 - 16 bytes) gets DFK_EUNSUPPORTED and keeps kind 0. */
int dfk_set_update_kernel(dfk_handle *h, int kind);
int dfk_get_update_kernel(dfk_handle *h, int *kind);

/* Fused row kernels in the time loop (an autotuner knob; no counterpart in
 * the reference, whose loop runs potential, forces and step one after
 * another, diskmain.cpp:100-108).  threads = 256, 512 or 1024: each
 * iteration is two launches, the column convolution (carrying the loop
 * control) and one workgroup of that many threads per block of ti full rows
 * that runs the row c2r of the rows its forces read, the update (LDS tile of
 * ti x Nphi) and the row r2c of its new Sigma, so Phi and the row spectra
 * never reach HBM; 0 = four launches (default).  Used when the update kernel
 * is the LDS tile (kind 0) and every disk has self-gravity, Nphi <= 8 x
 * threads factoring into 2, 3, 4 and 8, and a tile that fits the LDS (ti is
 * rounded down to an even number, at least 2); dfk_get_fused_rows reports
 * the threads in use (0: four launches).  Same step to round-off.  Changing
 * it needs a new dfk_run_init. */
int dfk_set_fused_rows(dfk_handle *h, int threads);
int dfk_get_fused_rows(dfk_handle *h, int *threads);

/* Graph chains (1 .. 16; no counterpart in the reference, which steps one
 * disk): the disks of the handle are split into that many contiguous groups
 * of about equal zone counts, and each group is stepped by its own chain of
 * launches in the loop's graphs; group g runs on stream g mod min(k, 4), the
 * groups of one stream one after the other within each step, so that the
 * kernels of one group fill the GPU while those of another drain.  Results
 * are the same bits as with one chain.  Default 3 (at most one group per
 * disk); only
 * handles whose disks all have self-gravity use more than one chain;
 * dfk_get_chains reports the number in use. */
int dfk_set_chains(dfk_handle *h, int k);
int dfk_get_chains(dfk_handle *h, int *k);

/* Device time in ms of the last dfk_advance (HIP events around the graph
 * launch on the handle's stream). */
int dfk_last_advance_ms(dfk_handle *h, double *ms);

/* Loop iterations captured per HIP graph in dfk_advance (default 128). */
int dfk_set_graph_steps(dfk_handle *h, int nsteps);

/* Profiling aid (no counterpart in the reference): launches each kernel of
 * one loop iteration `nrep` times back to back for every disk, bracketed by
 * HIP events on the handle's stream, and returns the average duration in ms
 * of one launch of: ms[0] control, ms[1] row r2c, ms[2] column convolution,
 * ms[3] row c2r, ms[4] fused update.  The state and loop position are left
 * as they were (the update rewrites the same output buffer each time).
 * Needs dfk_run_init first. */
int dfk_profile_step(dfk_handle *h, int nrep, double *ms);

#ifdef __cplusplus
}
#endif
#endif
