/* nsp.h -- C ABI of the MI355X nsPipe time step (libnsp.so).
 *
 * This boundary replaces the reference's compile-time configuration and the
 * free functions its driver calls (Mordered/nsPipe-GPU, Main_Code/src).  Each
 * entry point names the reference interface it stands in for.  Plain C types
 * only: opaque handle, int status codes, pointers and sizes.
 *
 * Field layout (both directions, host or device memory): three arrays
 * ur, ut, uz of interleaved (re, im) doubles, i.e. double2[NR][NT][NZ] with
 * h = i*NT*NZ + it*NZ + l (head.h:73 vfield, main.cu:30-33).  Azimuthal
 * wavenumber m = it for it < NT/2, it - NT otherwise (mod_deriv.cu:15); axial
 * wavenumber kz = 2*pi*l/LZ (mod_deriv.cu:31).  Velocities are stored in the
 * radial/azimuthal form (not u+/-), as in the reference's restart files.
 *
 * Errors are returned, never exit()ed (the reference exits, head.h:64-70).
 * One handle = one HIP stream; handles are independent (thread-compatible,
 * not thread-safe per handle).  The library owns all device memory; caller
 * buffers are only copied.
 */
#ifndef NSP_H
#define NSP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  NSP_OK = 0,
  NSP_ERR_INVALID = 1,   /* bad argument or unsupported size */
  NSP_ERR_HIP = 2,       /* HIP runtime error (no device, launch failure, ...) */
  NSP_ERR_IO = 3,        /* file missing, short, or not writable */
  NSP_ERR_NOMEM = 4      /* device or host allocation failed */
} nsp_status;

typedef enum {
  NSP_FP64 = 0,   /* every stage in FP64 (parity form) */
  NSP_MIXED = 1,  /* FP32 transforms/products; the velocity solves of |m| >= 4 read FP32 LU factors
                     and keep their column scratch (right-hand side, eliminated rows, explicit part,
                     pressure) in FP32, with FP64 arithmetic: those modes' new state carries the FP32
                     rounding of the right-hand side (~6e-8 of |u|/dt times dt, i.e. of |u|) each
                     step.  The state array, pressure factors and the |m| < 4 solves are FP64 */
  NSP_MIXED_REFINED = 2  /* NSP_MIXED with the velocity solves of every mode made increment
                     solves: the FP32-stored right-hand side is rhs - H u^n (formed in FP64 from
                     small terms), the FP32 solve gives delta = u - u^n and u = u^n + delta is formed
                     in FP64: one FP64 residual-refinement step from u^n, so the FP32 rounding falls
                     on delta only.  Single pipes only (not in ensembles) */
} nsp_precision;

typedef enum { NSP_HOST = 0, NSP_DEVICE = 1 } nsp_where;

/* Runtime parameters; replaces the #defines of head.h:14-61. */
typedef struct {
  int nr;          /* NR  radial points (head.h:16), >= 8 */
  int nt;          /* NT  = 2*m_theta azimuthal modes, even (head.h:17) */
  int nz;          /* NZ  = m_z+1 axial modes, odd (head.h:18) */
  double lz;       /* LZ  pipe length in R (head.h:22) */
  double re;       /* Re  (head.h:25) */
  double dt;       /* dt  constant time step (head.h:48) */
  double d_im;     /* implicit weight, 0.51 in the reference (head.h:49) */
  int maxit;       /* corrector iterations cap, 10 (head.h:51) */
  double tol;      /* corrector tolerance, 5e-8 (head.h:52) */
  double qdes;     /* target bulk flux, 0.5 (mod_cublasflux.cu:92) */
  int precision;   /* nsp_precision */
} nsp_params;

typedef struct nsp_handle nsp_handle;

/* Per-step report (the reference prints err every dt_frc, mod_integrate.cu:136). */
typedef struct {
  int n_it;        /* corrector iterations used (mod_integrate.cu:120) */
  double err;      /* final convergence measure err[0] (mod_cublasflux.cu:115) */
  int flags;       /* bit0: err was NaN; bit1: stopped at maxit without err <= tol */
} nsp_step_info;

#define NSP_FLAG_NAN 1
#define NSP_FLAG_MAXIT 2

/* Fills the reference defaults (head.h:14-61: sten 7, d_im 0.51, maxit 10,
 * tol 5e-8; Qdes 0.5 from mod_cublasflux.cu:92) for the given sizes. */
nsp_status nsp_default_params(nsp_params* p, int nr, int nt, int nz, double lz, double re,
                              double dt);

/* Replaces main.cu:19-28 (init_fd, setFft, setLinear, setCublasFlux,
 * setDeriv, setNonlinear, setIO, setInt, setBCon) and main.cu:30-33 (state
 * allocation).  Selects HIP device `device`.  The state starts at zero. */
nsp_status nsp_create(const nsp_params* p, int device, nsp_handle** out);

/* Replaces the *Destroy functions (main.cu:46-50). */
nsp_status nsp_destroy(nsp_handle* h);

/* State in/out in the reference layout (rBufferBinary/wBufferBinary payload,
 * mod_io_nonlinear.cu:528-565).  where = NSP_HOST or NSP_DEVICE. */
nsp_status nsp_set_state(nsp_handle* h, const double* ur, const double* ut, const double* uz,
                         int where);
nsp_status nsp_get_state(nsp_handle* h, double* ur, double* ut, double* uz, int where);

/* Reference initial condition for restart == 0 (initField, mod_io_nonlinear.cu:424-449):
 * Hagen-Poiseuille uz00 = 1 - r^2 plus the A_P perturbation (head.h:26). */
nsp_status nsp_reference_ic(nsp_handle* h, double a_p);

/* Advances nsteps time steps: the body of integrate()'s loop
 * (mod_integrate.cu:112-131), one nonlinear + predictor, then correctors
 * until err <= tol or maxit.  info may be NULL, else receives nsteps entries. */
nsp_status nsp_step(nsp_handle* h, int nsteps, nsp_step_info* info);

/* nonlinear(u, du) (mod_io_nonlinear.cu:380-409) on caller fields:
 * du = -(u.grad)u in r/theta form plus the mean pressure gradient on mode (0,0). */
nsp_status nsp_nonlinear(nsp_handle* h, const double* ur, const double* ut, const double* uz,
                         double* dr, double* dt, double* dz, int where);

/* Diagnostics of the current state (wrt_frc_k, mod_io_nonlinear.cu:40-52, plus
 * checks): out[0] centre-line velocity, out[1] friction velocity u_tau,
 * out[2] kinetic energy 1/2 sum_modes w_l sum_r rdr |u|^2 (w_0 = 1, else 2),
 * out[3] L2 divergence (weights |rdr|), out[4] max |u| on the wall,
 * out[5] max |wall divergence|, out[6] bulk flux sum rdr Re uz00. */
nsp_status nsp_diagnostics(nsp_handle* h, double out[7]);

/* Allocation-free diagnostics.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same seven
 * values as nsp_diagnostics, written to out: 7 doubles in DEVICE memory.  The
 * call only enqueues two kernels on the handle's stream (nsp_stream): it
 * allocates nothing, does not synchronise (wait on the stream before reading
 * out) and uses no memset.  The reduction's block partials live in the
 * library's static per-device scratch array (512 KiB, held until the process
 * ends), which limits NT*NZ to 4,194,304 modes: larger grids fail with
 * NSP_ERR_INVALID.  Every call on a device shares that device's array, so
 * calls that overlap on one device (several streams, threads or handles at
 * once) race.  Sums are combined in another order than nsp_diagnostics', so
 * KE and the divergence norm agree to rounding, not bit for bit. */
nsp_status nsp_diagnostics_noalloc(nsp_handle* h, double* out);

/* Radial tables (init_fd, mod_radialFD.cu:199-312; get_dr0 :316-350;
 * get_dr1 :355-358): r[NR], D[2][NR][7], Lu[2][NR][7], Lp[2][NR][7],
 * rdr[NR], dr0[4], dr1[7].  Any pointer may be NULL. */
nsp_status nsp_tables(nsp_handle* h, double* r, double* D, double* Lu, double* Lp, double* rdr,
                      double* dr0, double* dr1);

/* Same tables without a handle or device (host-only setup code). */
nsp_status nsp_radial_tables(int nr, double* r, double* D, double* Lu, double* Lp, double* rdr,
                             double* dr0, double* dr1);

/* Restart files, raw little-endian double2[NR][NT][NZ] with no header
 * (rBufferBinary/wBufferBinary, mod_io_nonlinear.cu:528-565; the reference
 * reads ur_s.bin/ut_s.bin/uz_s.bin, :426-429, and writes ur.bin/ut.bin/uz.bin,
 * main.cu:42-44). */
nsp_status nsp_read_state_files(nsp_handle* h, const char* ur_path, const char* ut_path,
                                const char* uz_path);
nsp_status nsp_write_state_files(nsp_handle* h, const char* ur_path, const char* ut_path,
                                 const char* uz_path);

/* Restart directory with metadata.  The reference keeps no time, step or
 * parameters with its state files (main.cu:42-44 writes the three raw
 * files only).  nsp_write_restart writes dir/ur.bin, ut.bin, uz.bin in the
 * reference's raw format and dir/restart.txt ("key value" lines: format,
 * time, step, nr, nt, nz, lz, re, dt, d_im, maxit, tol, qdes, precision).
 * nsp_read_restart reads them back into the state and returns time and step;
 * it fails with NSP_ERR_INVALID if the metadata's grid differs from the
 * handle's and with NSP_ERR_IO if restart.txt is missing (the raw files alone
 * carry no grid: read a reference restart with nsp_read_state_files).  t and
 * step may be NULL. */
nsp_status nsp_write_restart(nsp_handle* h, const char* dir, double t, long long step);
nsp_status nsp_read_restart(nsp_handle* h, const char* dir, double* t, long long* step);

/* Remesh (tools/Re_Mesh/src/mod_remesh.cu: re_mesh :50-124, rmsh_k :8-29,
 * modeZZ_k :34-40): the state of src, on its (NR, NT, NZ), interpolated onto
 * the grid of dst and made dst's state.  Radially each new point is the
 * Lagrange interpolant of 2*3 old points of the grid extended through the
 * axis (mirrored values carry the parity of the field: ur, ut odd in r for
 * even m, uz even); the wall value is copied.  Modes present on both grids
 * are carried over, new modes are zero (padding), others dropped
 * (truncation); ur(0,0) and the imaginary parts of ut(0,0), uz(0,0) are set
 * to zero.  No projection follows, as in the reference.  flags:
 * NSP_REMESH_REFERENCE keeps the reference's column assignment of the
 * mirrored points (exact only for rows whose stencil starts at the outermost
 * mirrored point, :97-107); NSP_REMESH_MIRROR gives every mirrored point its
 * own column.  src and dst are different handles on one device. */
#define NSP_REMESH_REFERENCE 0
#define NSP_REMESH_MIRROR 1
nsp_status nsp_remesh(nsp_handle* dst, nsp_handle* src, int flags);

/* The radial remesh matrix [2][nr_new][nr_old] (half 0 for fields even in r,
 * half 1 odd; mod_remesh.cu:80-111) without a handle or device. */
nsp_status nsp_remesh_weights(int nr_old, int nr_new, int flags, double* RM);

/* Physical-space velocity component comp (0 ur, 1 ut, 2 uz) of the current
 * state: padForward + fftBackward of the reference (padForward_k, mod_io_nonlinear.cu:173-192;
 * mod_fft.cu:43-49), the field its commented HDF5 writer stores (:486-521).
 * out[NR][NTP][NZF+2] doubles, NTP = 3NT/2, NZF = 3(NZ-1): row (i, j) holds
 * u(r_i, theta_j = 2 pi j/NTP, z_n = n LZ/NZF) for n < NZF, then the two
 * padding doubles of the in-place real transform (zero).  where: NSP_HOST or
 * NSP_DEVICE.  Uses the nonlinear-term workspace (nothing else changes). */
nsp_status nsp_physical_field(nsp_handle* h, int comp, double* out, int where);

/* Cross-section integrals at the NZF axial points (wrt_qcr, wrt_ucr,
 * mod_io_nonlinear.cu:576-654: q0kernel_k/u0kernel_k :55-78, the periodic
 * three-point rule of q1kernel_k :81-103, intq mod_cublasflux.cu:119-130):
 * q[n] = sum_i rdr_i sum_j 0.5/(3 NTP) (f_{j-1} + 4 f_j + f_{j+1}) with
 * f = ur^2 + ut^2, u2[n] the same with f = uz^2.  Host arrays; either may be
 * NULL. */
nsp_status nsp_cross_sections(nsp_handle* h, double* q, double* u2);

/* Real part of uz at mode (0,0) on every level, prof[NR] (wrt_mnp_k,
 * mod_io_nonlinear.cu:119-130). */
nsp_status nsp_mean_profile(nsp_handle* h, double* prof);

/* Diagnostics files in the reference's formats (setIO, mod_io_nonlinear.cu:261-317;
 * wrt_frc :460-479; wrt_qcr/wrt_ucr :576-654; wrt_mpr :663-687).
 * nsp_io_open creates (truncates) io_friction.txt, io_qcross.txt,
 * io_ucross.txt and io_meanpr.txt in dir and writes the header rows (axial
 * positions, radii).  nsp_io_write appends, for time t, the lines selected by
 * mask: NSP_IO_FRICTION (t, u_cl, u_tau), NSP_IO_CROSS (t, q[NZF]; t,
 * u2[NZF]), NSP_IO_MEANPR (t, uz00[NR]); values as "%.8e" separated by two
 * spaces.  Lines are written at once (the reference writes batches of dc_*
 * lines and drops a final partial batch).  nsp_io_close closes the files
 * (nsp_destroy does too). */
#define NSP_IO_FRICTION 1
#define NSP_IO_CROSS 2
#define NSP_IO_MEANPR 4
nsp_status nsp_io_open(nsp_handle* h, const char* dir);
nsp_status nsp_io_write(nsp_handle* h, double t, int mask);
nsp_status nsp_io_close(nsp_handle* h);

/* Seeded synthetic initial condition (no reference counterpart: the
 * reference's only IC is nsp_reference_ic).  Poiseuille alpha*(1-r^2) scaled
 * to bulk flux qdes, plus an analytically divergence-free, no-slip
 * perturbation on modes |m| <= mmax, 0 <= l <= lmax with complex Gaussian
 * amplitudes (mt19937_64(seed), Box-Muller) times the envelope
 * (kind 0: flat, kind 1: |k|^(-5/3) with |k|^2 = kz^2 + m^2; kind 2: a
 * localized puff, one amplitude per m shared by all l with the spectrum of a
 * Gaussian axial window of width 5R centred at LZ/2), rescaled so the
 * perturbation rms velocity is `rms`.  Written into the state; with
 * kind | NSP_IC_ADD the perturbation alone (no base flow) is added to the
 * current state (e.g. member-specific perturbations of an ensemble). */
#define NSP_IC_ADD 16
nsp_status nsp_seeded_ic(nsp_handle* h, unsigned long long seed, int mmax, int lmax, int kind,
                         double rms);

/* Per-kernel timing with HIP events recorded on the handle's stream around
 * every launch of nsp_step (no reference counterpart: the reference only
 * reports total process CPU time, main.cu:16,54-55).  nsp_profile_enable
 * turns it on/off and clears the totals; nsp_profile_read fills
 * nsp_profile_count() entries of accumulated milliseconds and launch counts,
 * indexed like nsp_profile_name(i). */
nsp_status nsp_profile_enable(nsp_handle* h, int enable);
int nsp_profile_count(void);
const char* nsp_profile_name(int i);
nsp_status nsp_profile_read(nsp_handle* h, double* total_ms, long long* launches);

/* Autotuner (no reference counterpart): times every compiled kernel variant of
 * the nonlinear stages (radial derivative fused or separate, azimuthal tile
 * width, threads per axial row) on this handle's sizes, device and current
 * state, `reps` launches each, and keeps the fastest per stage.
 * nsp_tuning returns the timings and the choice as JSON text (valid until
 * the next call on the handle). */
nsp_status nsp_autotune(nsp_handle* h, int reps);
const char* nsp_tuning(nsp_handle* h);

/* Restores a choice made by nsp_autotune (no reference counterpart): json is
 * nsp_tuning's text, or just its "chosen" object.  Fails with
 * NSP_ERR_INVALID, keeping the current variants, if a named variant is not
 * compiled for this handle's sizes.  Runs one nonlinear evaluation of the
 * current state as the check. */
nsp_status nsp_set_tuning(nsp_handle* h, const char* json);

/* Ensembles of independent pipes (no reference counterpart: the reference runs
 * one pipe per process).  nsp_ensemble_create builds count pipes from
 * params[0..count-1] (sizes may differ; every member needs compiled transform
 * plans, and all share maxit and tol).  Members with identical parameters
 * share their operator data (LU factors, influence data).  nsp_ensemble_step
 * advances every member nsteps time steps: each stage is one launch per class
 * of members with the same transform plans (one for all members in the radial
 * derivative and the solves), the step captured in hipGraphs; members converge
 * independently.  info, if not NULL, receives nsteps*count entries, member
 * fastest.  nsp_ensemble_member returns a member's handle for the per-pipe
 * calls (state, ICs, diagnostics, restart files); it is owned by the
 * ensemble.  nsp_ensemble_autotune tunes one member per class and applies
 * the choice to the class. */
typedef struct nsp_ensemble nsp_ensemble;
nsp_status nsp_ensemble_create(const nsp_params* params, int count, int device, nsp_ensemble** out);
nsp_status nsp_ensemble_destroy(nsp_ensemble* e);
int nsp_ensemble_size(nsp_ensemble* e);
nsp_handle* nsp_ensemble_member(nsp_ensemble* e, int i);
nsp_status nsp_ensemble_autotune(nsp_ensemble* e, int reps);
nsp_status nsp_ensemble_step(nsp_ensemble* e, int nsteps, nsp_step_info* info);

/* HIP stream of the handle (a hipStream_t), for callers that order their own
 * work against it. */
void* nsp_stream(nsp_handle* h);

/* Waits for all work of the handle's device (hipDeviceSynchronize). */
nsp_status nsp_synchronize(nsp_handle* h);

/* Human-readable text of the last error on this thread. */
const char* nsp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* NSP_H */
