/*
 * psdns_gpu.h - C ABI of the MI355X (gfx950) implementation of the PSDNS
 * RK2 pseudo-spectral step.  Plain C types only, callable from Fortran
 * through ISO_C_BINDING (see fortran/psdns_gpu_mod.f90 in the package) and
 * from C/C++/Python (ctypes).
 *
 * The reference (deepanjan90/NASA-Bluewater-PSDNS-Benchmark-Minsky) has no
 * FFI: its hot path is a set of Fortran subroutines with implicit
 * interfaces working on module-global state, called from rksubstep
 * (PSDNS/timestep_module.f:152-317).  Each entry point below names the
 * reference routine it replaces.  Module globals (grid, shifts, integrating
 * factors, dt2) become arguments or context state.
 *
 * Conventions
 *  - nx = ny = nz = n, a power of two, 32 <= n <= 1024; one GPU plays the
 *    role of one task with iproc = jproc = 1 (xisz = n/2).
 *  - Host arrays use the reference layouts (Fortran column-major shown):
 *      cylinder  uy(ny, num_al, nv)            complex, precision prec
 *      physical  ux(nx, zisz=nz, yjsz=ny, nv)  real,    precision prec
 *    Cylinder column a enumerates (x, z) as next_xz (mpisetup.f:710-732).
 *  - prec = 8 (FP64) or 4 (FP32) for state and transforms; velmax, shell
 *    sums, dt and integrating factors are always FP64.
 *  - Every call is synchronous at return, runs on the context's own HIP
 *    stream, and returns 0 on success or a negative PSDNS_E* code; the
 *    message of the last failure on the calling thread is returned by
 *    psdns_last_error().  The Fortran caller maps non-zero to abrt().
 *  - One host thread per context.  Host pointers are borrowed for the call
 *    only.  The context owns all device memory.
 *  - stage_update / rk4_update need the nonlinear terms of the
 *    stage_nonlinear just before them: any call in between that writes
 *    the work arrays or the shift tables (set_shifts, put_field,
 *    vg_moments, the per-routine entry points, put_state, init_field,
 *    inpen) discards them, and the update then fails with PSDNS_ESTATE.
 */
#ifndef PSDNS_GPU_H
#define PSDNS_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define PSDNS_OK 0
#define PSDNS_EINVAL (-1)   /* bad argument */
#define PSDNS_EHIP (-2)     /* HIP runtime error */
#define PSDNS_ENOMEM (-3)   /* device allocation failed */
#define PSDNS_ESTATE (-4)   /* call out of order (e.g. update before nonlinear) */
#define PSDNS_EIO (-5)      /* file open, read or write failed */

typedef struct psdns_ctx psdns_ctx;

typedef struct {
  int nx, ny, nz;    /* grid, must be equal */
  int nc;            /* passive scalars, 0..8 (state variables 4 .. 3 + nc) */
  int prec;          /* 4 = FP32, 8 = FP64 */
  int kshift;        /* informational (shifts come in through psdns_set_shifts) */
  int mask_kind;     /* 0 = reference sphere |k|/n <= sqrt(2)/3 (masks.f:67-77) */
  double viscos;     /* kinematic viscosity (input, PSDNS/input) */
  double pr[8];      /* Prandtl/Schmidt numbers of the scalars (> 0), step.f:387-407 */
  double grad[8][3]; /* mean scalar gradients G (cb), itransform.f:522-531 */
} psdns_params;

/* Set-up: mpisetup2 cylinder (mpisetup.f:292-358), masks (masks.f:46-83),
 * init/waveno wavenumbers (init.f:15-37, waveno.f:16-38), FFT set-up in
 * place of epfftw (epfftw.f:27-221). */
int psdns_create(const psdns_params *p, psdns_ctx **out);
int psdns_destroy(psdns_ctx *c);
/* num_al and max_al_x of the cylinder (mpisetup.f:311-354) */
int psdns_geometry(const psdns_ctx *c, int *num_al, int *max_al_x);
/* column map: col_x[a], col_z[a] 0-based (next_xz order); cut_z[max_al_x] */
int psdns_columns(const psdns_ctx *c, int *col_x, int *col_z, int *cut_z);

/* State transfer (host <-> device), layout 0 = reference cylinder
 * uny(ny, num_al, 3 + nc) (velocity, then the scalars).  put_state also
 * sets u = un (main.f:357-359). */
int psdns_put_state(psdns_ctx *c, const void *uny, int layout);
int psdns_get_state(psdns_ctx *c, void *uny, int layout);

/* Checkpoint of un in the reference's pencil file format, replacing
 * outpen (homogeneous/outpen.f:1-229, record layout outpen_write.f:59-76)
 * and inpen (homogeneous/inpen.f, records read as inpen_read.f with
 * slabpenc = 2) for one task.  The file is Fortran sequential-unformatted,
 * big-endian (-fconvert=big-endian, makefile:251); per variable
 * i = 1 .. 3 + nc:
 *   record  i, nx, ny, nz                            4 x int32
 *   record  xist, xisz, zjst, zjsz, 1, ny            6 x int32 = 1, n/2, 1, n, 1, n
 *   for z = 1 .. n except n/2+1, two records of complex(file_prec):
 *     buf(1:n/2, 1:n/2, z), then buf(1:n/2, n/2+2:n, z)   (x fastest)
 * with buf(x, y, z) = uny(y, a) on cylinder column a = (x, z) and 0 off the
 * cylinder (kxcomm1.f:586-612).  file_prec 8 = complex*16 (the shipped
 * DOUBLE_PREC build), 4 = complex*8.  The reference's path for one task is
 * outpen/0/outpenc.p0 (outpen_write.f:45-60, main.f:196); the caller names
 * the path.  inpen reads either precision (from the record length), needs
 * the same grid and 3 + nc variables (no resolution change, no kinit
 * subset), zeroes the modes the file does not hold (ky = n/2 row, kz = n/2
 * plane), and sets u = un as put_state does.  Bad headers or record
 * lengths give PSDNS_EINVAL, open/read/write failures PSDNS_EIO; a failed
 * inpen may have replaced the variables before the one that failed (put a
 * state or read a good file before stepping again). */
int psdns_outpen(psdns_ctx *c, const char *path, int file_prec);
int psdns_inpen(psdns_ctx *c, const char *path);

/* Phase-shift offsets gsh(3,2) in grid cells, as produced by shifts.f
 * (shifts.f:65-84); the factors sx, sy, sz (shifts.f:146-178) are formed
 * in FP64 inside. */
int psdns_set_shifts(psdns_ctx *c, const double gsh[6]);

/* One RK2 stage, split where the reference calls step (timestep_module.f:261):
 * stage_nonlinear = phase1 + itransform_vel + realspace_vel + transform up to
 * the convective combination (timestep_module.f:152-244); on kstep 1 it
 * returns velmax (FP64) for the caller's dt = cfl/velmax (step.f:64-78).
 * stage_update = the y-forward part of transform + wavespace_vel + proc4a
 * (kstep 1) or proc4b (kstep 2) (timestep_module.f:244-317).  xfac(n/2),
 * yfac(n), zfac(n) are the integrating factors of step.f:366-385, used by
 * kstep 1 only (proc4b has none; they may be NULL on kstep 2).  forcing
 * is NULL or the EPFOR increments for(kfor, k2fo, k2fo, 3), k2fo =
 * 2 kfor - 1, as the reference's force.f forms them (FP64 complex,
 * Fortran order; needs psdns_set_epfor first): added after the projection,
 * on kstep 1 where |k| <= kforce with the efkz tally (proc4a.f:588-640),
 * on kstep 2 in the whole box before un <- u (proc4b.f:118-163).  FEK
 * forcing is psdns_ek_force.
 * stage_update(1) without forcing only records its arguments: the update
 * itself runs in the next call.  When that call is stage_nonlinear(2), the update and stage
 * 2's phase shift and inverse y transform are one kernel; any other call
 * runs the update on its own first, so every result the caller can observe
 * is the same either way. */
int psdns_stage_nonlinear(psdns_ctx *c, int kstep, double *velmax);
int psdns_stage_update(psdns_ctx *c, int kstep, double dt, const double *xfac,
                       const double *yfac, const double *zfac, const void *forcing);

/* RK4 (rkmethod = 4, timestep_module.f:293-310):
 * per substep kstep = 1..4, stage_nonlinear(kstep) (stages 3 and 4 use the
 * shift offsets of stages 1 and 2, shifts.f:133-137; velmax on kstep 1) and
 * then rk4_update(kstep), which replaces the y part of transform +
 * wavespace_vel + advanc_kstep (homogeneous/advanc_1.f .. advanc_4.f,
 * non-EPFOR, modpress = 1).  With N the nonlinear term (dt/2 attached), un
 * the state at the start of the step and u1 an accumulator the context
 * allocates on first use (zeroed, comp_set.f:72-74):
 *   1: u1 = F (un + N/3), u = P F (un + N), F = xfac yfac zfac (full dt)
 *   2: u1 += 2N/3, u = P (un + N)
 *   3: u1 += 2N/3, u = P (un + 2N)
 *   4: u1 = P (u1 + N/3); un = u = u1; zero mean
 * on the mask, P the projection with k = 0 excluded.  Passive scalars take
 * the same steps without the projection, with xdif ydif zdif in place of F
 * (advanc_1.f:101-111).  The factors are read on kstep 1 only (NULL
 * allowed after). */
int psdns_rk4_update(psdns_ctx *c, int kstep, double dt, const double *xfac,
                     const double *yfac, const double *zfac);

/* EPFOR forcing (the -DEPFOR build, makefile:69): set_epfor gives the
 * forcing box kfor (param_set.f:33: kfor = kforce + 1.5 for a cube) and the
 * radius kforce (input.f:186); kfor = 0 turns it off.  The increments
 * come with each stage_update (RK2) or, for RK4, with rk4_forcing, which
 * keeps them for the following rk4_update calls until replaced or cleared
 * (NULL): advanc_1 adds them with the tally, advanc_2 and advanc_3 in the
 * whole box; advanc_4's add is overwritten by u1 on the mask in the
 * reference (advanc_4.f:96-170), so substep 4 has none.
 * epfor_tally returns efkz(nxh, 3) of the last proc4a / advanc_1 (zeroed
 * when that stage had no increments, proc4a.f:66): efkz[(i-1) nxh + k-1]
 * = sum over shell k of 2 Re(u conj(for_i)) tfact, FP64, summed in a fixed
 * order (the caller's forcdt.f divides by dt). */
int psdns_set_epfor(psdns_ctx *c, int kfor, double kforce);
int psdns_epfor_tally(psdns_ctx *c, double *efkz);
int psdns_rk4_forcing(psdns_ctx *c, const void *forcing);

/* Lagrangian particles, RK2 (hooks/LAG_LOCAL: lagpart.f, partic.f,
 * partsp.f:217-299).  pp(np, 9) is the reference's particle array, FP64,
 * column-major (pp[(j - 1) np + i]): columns 1-3 positions in grid units
 * (the node i of the stage-1 grid is at gx(i) = 1 + (i - 1) + gsh(1,1),
 * partsp.f:44-53), 4-6 velocities, 7-9 the half-step positions rk_advpos
 * keeps.  The reference's order, from realspace at kstep 1 (before step, so
 * with the previous step's dt): rk_advpos(1, 1/2, 1/2); the spline of the
 * state on the stage-1 grid of the step and the interpolation at the
 * predicted positions; rk_advpos(2, 0, 1/2) (the first step only
 * interpolates, partic(4)).
 * lag_put: np particles, the first ncols (3, 6 or 9) columns given, the
 *   rest zero (device memory: pp, 72 bytes per particle, plus the
 *   interpolation's sort buffers, 16 bytes per particle and the radix
 *   sort's temporary storage); lag_get: all 9 columns.
 * lag_spline (spxyz_m.f with splx/splz/sply, solsp1.f): the tricubic
 *   spline coefficients of the velocity of un on the phase-shifted grid of
 *   stage 1 of the last set_shifts: the periodic (1,4,1) systems along x,
 *   y, z on 216 x the grid values (splz.f:22-28), solved exactly in Fourier
 *   space (eigenvalue 4 + 2 cos(2 pi k / n) per direction) and transformed
 *   to the grid; kept in ux(1:3) (allocated on first use).  Any call that
 *   writes the work arrays afterwards discards them.
 * lag_interp (spcal_local.f:119-364): pp(:,4:6) = the spline velocity at
 *   pp(:,1:3), B-spline weights of nodes ib-1 .. ib+2 (periodic), FP64 sums
 *   in the reference's order.  The particles are processed in the order of
 *   their 4 x 4 x 4-cell bin (a GPU sort each call); the order does not
 *   change any particle's result.
 * lag_advpos (rk_advpos.f:15-52), fact = cof dt n / (2 pi): pstep 1
 *   pp(:,7:9) = x + fact1 u, x = pp(:,7:9) + fact2 u; pstep 2 x = pp(:,7:9)
 *   + fact2 u. */
int psdns_lag_put(psdns_ctx *c, const double *pp, int np, int ncols);
int psdns_lag_get(psdns_ctx *c, double *pp, int np);
int psdns_lag_spline(psdns_ctx *c);
int psdns_lag_interp(psdns_ctx *c);
int psdns_lag_advpos(psdns_ctx *c, int pstep, double dt, double cof1, double cof2);

/* FEK forcing (ek_force.f:1-76 with ek_shell.f:81-176): shell energies of
 * un for ik = int(|k| + 1.5) <= kfor, then u and un scaled by
 * sqrt(ekinf(ik)/ek(ik)) for ik <= kf_shell (ratio 0 for ik = 1).
 * ek_out (nullable) receives the shell energies after the rescale. */
int psdns_ek_shell(psdns_ctx *c, int kfor, double *ek);
int psdns_ek_force(psdns_ctx *c, int kf_shell, int kfor, const double *ekinf, double *ek_out);

/* FP64 shell statistics of un (homogeneous/sptr.f:139-176,235):
 * ek[ik-1] = sum tfact |u|^2/2, dk[ik-1] = 2 viscos k^2 ek, ik = int(|k|+1.5),
 * nshell shells; tke and eps over all modes. */
int psdns_shell_stats(psdns_ctx *c, double *ek, double *dk, int nshell, double *tke, double *eps);

/* psdns_shell_stats without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results (the same kernels, bit for bit); the partial
 * rows live in a fixed device array of the library sized for n <= 1024
 * (7.9 MB per GPU, held until the process ends): a larger grid is refused
 * with PSDNS_EINVAL.  Calls that overlap on one GPU share that array and
 * race. */
int psdns_shell_stats_noalloc(psdns_ctx *c, double *ek, double *dk, int nshell, double *tke, double *eps);

/* Velocity-gradient moments (vgmomt.f:68-143): du/dx and dw/dz of the state
 * un (on the mask, unshifted grid) in physical space, dudx[m-1] and
 * dwdz[m-1] = sum of the m-th power over the grid / (nx ny nz), m = 1..8,
 * FP64 sums.  Skewness = [2] / [1]^1.5, flatness = [3] / [1]^2 (the
 * normalised moments of vgmomt.f:152-156).  Uses the u slots 4 and 5
 * (1-based) and the physical array ux(1:2) as work space. */
int psdns_vg_moments(psdns_ctx *c, double *dudx, double *dwdz);

/* Random-phase solenoidal initial field on the device (homogeneous/field.f:
 * 172-257 construction, consym.f symmetry), E(k) ~ k^4 exp(-2 (k/kpeak)^2),
 * rescaled to the given TKE; sets un and u. */
int psdns_init_field(psdns_ctx *c, unsigned long long seed, double kpeak, double tke);

/* Per-routine entry points, each the device counterpart of one reference
 * subroutine, on the context's work arrays:
 *   slot "u"  = uy(ny, num_al, 5) cylinder   (PSDNS_FIELD_U)
 *   slot "un" = uny(ny, num_al, 3 + nc) cylinder  (PSDNS_FIELD_UN)
 *   slot "ux" = ux(nx, nz, ny, 5) physical   (PSDNS_FIELD_UX) */
#define PSDNS_FIELD_U 0
#define PSDNS_FIELD_UN 1
#define PSDNS_FIELD_UX 2
int psdns_put_field(psdns_ctx *c, int which, int nv, const void *host);
int psdns_get_field(psdns_ctx *c, int which, int nv, void *host);

int psdns_phase1(psdns_ctx *c, int kstep);                 /* homogeneous/phase1.f:136-216 */
int psdns_itransform_vel(psdns_ctx *c);                    /* itransform_vel.f:11-129: u -> ux(1:3) */
int psdns_realspace_vel(psdns_ctx *c, int rkstep, double *velmax); /* realspace_vel.f:7-187 */
int psdns_transform(psdns_ctx *c);                         /* transform.f:10-301: ux(1:5) -> u(1:4) */
int psdns_wavespace_vel(psdns_ctx *c, int kstep, double dt2); /* wavespace_vel.f:11-235 */
int psdns_proc4a(psdns_ctx *c, const double *xfac, const double *yfac, const double *zfac); /* proc4a.f */
int psdns_proc4b(psdns_ctx *c);                            /* homogeneous/proc4b.f:3-208 */
int psdns_kxtran(psdns_ctx *c, int nv);                    /* kxtran.f:17-93: u(1:nv) -> ux(1:nv) */
int psdns_xktran(psdns_ctx *c, int nv);                    /* xktran.f:1-88: ux(1:nv) -> u(1:nv) */

/* Device time of the last stage calls, per pass (ms), the reference's
 * t_itrans/t_trans split (timestep_module.f:98-437); times[0..5] =
 * y-inverse, z-inverse, x-fused, z-forward, y-forward+update, total, for the
 * velocity passes (psdns_pass_times has the scalar passes too).  A stage 2
 * that follows stage_update(1) directly has no y-inverse of its own: that
 * transform is timed with the y-forward+update of stage 1. */
int psdns_stage_times(psdns_ctx *c, int enable, double *times);

/* Device time of the passes per RK substep, the reference's per-substep
 * tt(1..10) / t_itrans / t_trans timers (timestep_module.f:98-437) at the
 * granularity of this library's passes, scalars included: times[11 st + p]
 * for substep st = 0..3 (RK2 uses 0 and 1) and pass p =
 *   0 y-inverse (+ phase1)   1 z-inverse   2 x pass (+ realspace_vel)
 *   3 z-forward   4 y-forward + wavespace + update (RK4: + advanc_k)
 *   5 scalar gradient y pass   6 scalar z-inverse   7 scalar x pass
 *   8 scalar z-forward   9 scalar y-forward   10 scalar updates
 * summed over the scalars and over the calls since the last read (the x
 * pass of a stage serves the velocity and every scalar in one kernel, so
 * with scalars pass 2 holds their x work too and pass 7 stays 0; each
 * scalar's y forward pass is part of its update, so pass 9 stays 0 and
 * pass 10 holds both; a stage 2
 * that follows stage_update(1) directly has its y-inverse inside substep
 * 0's pass 4).  ntimes >= 44.  Reading clears the table; enable turns the
 * timing on or off (psdns_stage_times shares the table: its times[0..4]
 * are passes 0..4 summed over the substeps, times[5] their total). */
int psdns_pass_times(psdns_ctx *c, int enable, double *times, int ntimes);

const char *psdns_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
