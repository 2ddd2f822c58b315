/*
 * CPU FP64 oracle for the PSDNS RK2 pseudo-spectral step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (deepanjan90/NASA-Bluewater-PSDNS-Benchmark-Minsky, Fortran) used
 * by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py as
 * the checker.  The shipped GPU library never links, loads or calls it.
 *
 * Layouts (0-based indices, all arrays contiguous, fastest index last):
 *   cylinder   uy[v][a][y]      == reference uy(ny, num_al, nv)      complex
 *   z-pencils  uz[v][y][kx][z]  == reference uz(nz, xisz, yjsz, nv)  complex
 *   physical   ux[v][y][z][x]   == reference ux(nx, zisz, yjsz, nv)  real
 * with nx = ny = nz = n, xisz = n/2, one task (iproc = jproc = 1).
 */
#ifndef PSDNS_ORACLE_H
#define PSDNS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid orc_grid;

/* grid set-up: wavenumbers (init.f), cylinder (mpisetup.f mpisetup2/next_xz) */
orc_grid *orc_grid_create(int n);
void orc_grid_destroy(orc_grid *g);
int orc_num_al(const orc_grid *g);
int orc_max_al_x(const orc_grid *g);
void orc_columns(const orc_grid *g, int *col_x, int *col_z, int *cut_z);
long orc_mask_count(const orc_grid *g);
int orc_mask(const orc_grid *g, int y, int a);

/* shifts.f: phase-shift factors from grid shifts gsh[stage][dir] (cells) */
void orc_set_shifts(orc_grid *g, const double gsh[6]);
/* shifts.f kshift rules: advance (gsh1,gsh2) given fresh randoms r[3] */
void orc_shift_rule(int kshift, double gsh[6], const double r[3], int first);

/* 1-D transform contract (epfftw.f): unnormalised, sign -1 = FFTW_FORWARD */
void orc_fft_c2c(double *data, int len, int count, int sign);
void orc_fft_r2c(const double *in, double *out, int len, int count);
void orc_fft_c2r(const double *in, double *out, int len, int count);

/* hot path, one routine per reference subroutine */
void orc_phshift_inplace(const orc_grid *g, double *uy, int nv, int kstep);
void orc_itransform_vel(const orc_grid *g, const double *uy, double *ux);
void orc_realspace_vel(const orc_grid *g, double *ux, int rkstep, double *velmax);
void orc_transform(const orc_grid *g, const double *ux, double *uy);
void orc_wavespace_vel(const orc_grid *g, double *uy, int kstep, double dt2);
void orc_step_factors(const orc_grid *g, double viscos, double dt,
                      double *xfac, double *yfac, double *zfac);
void orc_proc4a(const orc_grid *g, double *uy, double *uny,
                const double *xfac, const double *yfac, const double *zfac);
void orc_proc4b(const orc_grid *g, double *uy, double *uny);
void orc_kxtran(const orc_grid *g, const double *uy, double *ux, int nv);
void orc_xktran(const orc_grid *g, const double *ux, double *uy, int nv);

/* EPFOR forcing increments for(kfor, 2 kfor - 1, 2 kfor - 1, 3) (complex)
 * after proc4a (mode 1, efkz tally), proc4b (mode 2) or advanc_2/3 (mode 3) */
void orc_epfor(const orc_grid *g, double *uy, double *uny, int kfor, double kforce,
               const double *forc, int mode, double *efkz);

/* FEK forcing (ek_shell.f / ek_force.f) and statistics (sptr.f) */
void orc_ek_shell(const orc_grid *g, const double *uny, int kfor, double *ek);
void orc_ek_force(const orc_grid *g, double *uy, double *uny, int kf_shell,
                  int kfor, const double *ekinf);
void orc_spectrum(const orc_grid *g, const double *uny, double viscos,
                  int nshell, double *ek, double *dk, double *tke, double *eps);
/* moments 1..8 of du/dx and dw/dz over the grid / n^3 (vgmomt.f:68-143) */
void orc_vg_moments(const orc_grid *g, const double *uny, double *dudx, double *dwdz);

/* Lagrangian particles: tricubic-spline coefficients of a periodic field
 * f[y][z][x] in place (spxyz_m.f: splx, splz x 216, sply with the solsp1
 * factors), interpolation of nv fields at pp(np, 1:3) into pp(np, 4:3+nv)
 * (spcal_local.f), the RK2 position update (rk_advpos.f); pp(np, 9) is
 * column-major as the reference's */
void orc_spline3(int n, double *f);
void orc_lag_interp(int n, const double *c, int nv, const double gsh[3], long np, double *pp);
void orc_lag_advpos(int n, long np, double *pp, int pstep, double dt, double cof1, double cof2);

/* initial field: random-phase solenoidal construction of field.f with a
 * counter-based generator, conjugate symmetry on kx=0 (consym.f) */
void orc_field(const orc_grid *g, double *uny, unsigned long long seed,
               double kpeak, double tke);

#ifdef __cplusplus
}
#endif
#endif
