/*
 * CPU FP64 oracle for the PSDNS RK2 pseudo-spectral step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU library, never part of
 * it.  Each function restates one reference subroutine; the citation above
 * each one names the file:line it follows (paths relative to the reference
 * repository).  Velocity-only path (nc = 0), unit metric (b11=b22=b33=1,
 * beta = 1), one MPI task (iproc = jproc = 1), no padding (nxpad = nx).
 *
 * Pins (see tests/test_oracle_*.py): the 1-D transforms against numpy's
 * pocketfft; the cylinder and mask against the counts restated from
 * mpisetup.f/masks.f; the inverse/forward transforms against single Fourier
 * modes in closed form and the shift theorem; the products pointwise; the
 * nonlinear term against the analytic Taylor-Green term; the update against
 * projection (k.u = 0), energy conservation of the band-limited nonlinear
 * term and the exact Stokes decay of a Beltrami field.
 */
#include "psdns_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

struct orc_grid {
  int n, nxh, max_al_x, num_al;
  int *cut_z;     /* [max_al_x]: positive kz outside the cylinder          */
  int *col_x;     /* [num_al]  : 0-based x (== kx) of cylinder column a     */
  int *col_z;     /* [num_al]  : 0-based z index of column a                */
  cplx *sx;       /* [2][nxh]  shift factors, RK stage 1 and 2              */
  cplx *sy;       /* [2][n]                                                  */
  cplx *sz;       /* [2][n]                                                  */
};

/* init.f:105-117: kx = 0..n/2-1; ky, kz in FFT order */
static inline int kwave(int i, int n) { return i <= n / 2 ? i : i - n; }

/* ---------------------------------------------------------------------- */
/* grid: mpisetup.f:311-354 (max_al_x, cut_z, num_al), mpisetup.f:710-732  */
/* (next_xz column order: x outer; z = 1..nzhp-c-1 then nzhp+c+1..nz)      */
/* ---------------------------------------------------------------------- */
orc_grid *orc_grid_create(int n) {
  orc_grid *g = (orc_grid *)calloc(1, sizeof(orc_grid));
  g->n = n;
  g->nxh = n / 2;
  /* mpisetup.f:311-312: ikmax = nx*sqrt(2.)/3. ; max_al_x = min(xisz, ikmax+2-xist) */
  int ikmax = (int)(n * sqrt(2.0) / 3.0);
  g->max_al_x = ikmax + 1 < g->nxh ? ikmax + 1 : g->nxh;
  g->cut_z = (int *)malloc(sizeof(int) * g->max_al_x);
  double rnxi = 1.0 / n;
  long tot = 0;
  for (int x = 0; x < g->max_al_x; x++) {
    /* mpisetup.f:331-332 */
    double kxr = (double)x * rnxi;
    double rc = n * (1.0 - 2.0 * sqrt(2.0 / 9.0 - kxr * kxr));
    int c = (int)(rc / 2);
    g->cut_z[x] = c;
    tot += n - 1 - 2 * c;
  }
  g->num_al = (int)tot;
  g->col_x = (int *)malloc(sizeof(int) * g->num_al);
  g->col_z = (int *)malloc(sizeof(int) * g->num_al);
  int a = 0;
  for (int x = 0; x < g->max_al_x; x++) {
    int c = g->cut_z[x];
    for (int z = 0; z < n / 2 - c; z++) { g->col_x[a] = x; g->col_z[a] = z; a++; }
    for (int z = n / 2 + c + 1; z < n; z++) { g->col_x[a] = x; g->col_z[a] = z; a++; }
  }
  g->sx = (cplx *)malloc(sizeof(cplx) * 2 * g->nxh);
  g->sy = (cplx *)malloc(sizeof(cplx) * 2 * n);
  g->sz = (cplx *)malloc(sizeof(cplx) * 2 * n);
  double zero[6] = {0, 0, 0, 0, 0, 0};
  orc_set_shifts(g, zero);
  return g;
}

void orc_grid_destroy(orc_grid *g) {
  if (!g) return;
  free(g->cut_z); free(g->col_x); free(g->col_z);
  free(g->sx); free(g->sy); free(g->sz);
  free(g);
}

int orc_num_al(const orc_grid *g) { return g->num_al; }
int orc_max_al_x(const orc_grid *g) { return g->max_al_x; }

void orc_columns(const orc_grid *g, int *col_x, int *col_z, int *cut_z) {
  if (col_x) memcpy(col_x, g->col_x, sizeof(int) * g->num_al);
  if (col_z) memcpy(col_z, g->col_z, sizeof(int) * g->num_al);
  if (cut_z) memcpy(cut_z, g->cut_z, sizeof(int) * g->max_al_x);
}

/* masks.f:67-77: (kx/nx)^2+(ky/ny)^2+(kz/nz)^2 <= 2/9, y != nyhp, z != nzhp */
int orc_mask(const orc_grid *g, int y, int a) {
  int n = g->n;
  double rn = 1.0 / n;
  double kx = g->col_x[a] * rn, ky = kwave(y, n) * rn, kz = kwave(g->col_z[a], n) * rn;
  double sqk = ky * ky + (kz * kz + kx * kx);
  if (sqk > 2.0 / 9.0 || y == n / 2 || g->col_z[a] == n / 2) return 0;
  return 1;
}

long orc_mask_count(const orc_grid *g) {
  long c = 0;
  for (int a = 0; a < g->num_al; a++)
    for (int y = 0; y < g->n; y++) c += orc_mask(g, y, a);
  return c;
}

/* shifts.f:146-178: sx(x,k) = exp(i 2 pi kx gsh(1,k) / nx), etc.
 * gsh layout: gsh[3*k + dir], k = stage 0/1, dir = 0 (x), 1 (y), 2 (z). */
void orc_set_shifts(orc_grid *g, const double gsh[6]) {
  int n = g->n;
  for (int k = 0; k < 2; k++) {
    for (int x = 0; x < g->nxh; x++) {
      double arg = 2.0 * M_PI / n * x * gsh[3 * k + 0];
      g->sx[k * g->nxh + x] = cos(arg) + I * sin(arg);
    }
    for (int y = 0; y < n; y++) {
      double arg = 2.0 * M_PI / n * kwave(y, n) * gsh[3 * k + 1];
      g->sy[k * n + y] = cos(arg) + I * sin(arg);
    }
    for (int z = 0; z < n; z++) {
      double arg = 2.0 * M_PI / n * kwave(z, n) * gsh[3 * k + 2];
      g->sz[k * n + z] = cos(arg) + I * sin(arg);
    }
  }
}

/* shifts.f:40-84: cancel(d) = mod(d + 0.5, 1); kshift 1..4 rules.
 * `first` reproduces the one-time gsh(:,2) initialisation (shifts.f:47-51),
 * which consumes r[] itself; the caller then calls again with fresh r[]. */
void orc_shift_rule(int kshift, double gsh[6], const double r[3], int first) {
  for (int i = 0; i < 3; i++) {
    if (first) { gsh[3 + i] = r[i]; continue; }
    if (kshift == 1) {
      gsh[i] = fmod(gsh[3 + i] + 0.5, 1.0);
      gsh[3 + i] = r[i];
    } else if (kshift == 2) {
      gsh[i] = r[i];
      gsh[3 + i] = fmod(gsh[i] + 0.5, 1.0);
    } else if (kshift == 3) {
      gsh[i] = fmod(gsh[3 + i] + 0.5, 1.0);
      gsh[3 + i] = fmod(gsh[i] + 0.5, 1.0);
    } else {
      gsh[i] = 0.0;
      gsh[3 + i] = 0.0;
    }
  }
}

/* ---------------------------------------------------------------------- */
/* 1-D transforms: iterative radix-2, twiddles from cos/sin in FP64.        */
/* Sign convention of epfftw.f:77,135,146 (FFTW_BACKWARD = +i for k->x,     */
/* FFTW_FORWARD = -i), unnormalised like FFTW.                              */
/* ---------------------------------------------------------------------- */
typedef struct { int n; int sign; cplx *w; int *rev; } fftplan;

static fftplan plan_make(int n, int sign) {
  fftplan p;
  p.n = n; p.sign = sign;
  p.w = (cplx *)malloc(sizeof(cplx) * (n / 2 > 0 ? n / 2 : 1));
  p.rev = (int *)malloc(sizeof(int) * n);
  for (int k = 0; k < n / 2; k++) {
    double ang = 2.0 * M_PI * k / n;
    p.w[k] = cos(ang) + I * (sign * sin(ang));
  }
  int lg = 0;
  while ((1 << lg) < n) lg++;
  for (int i = 0; i < n; i++) {
    int r = 0;
    for (int b = 0; b < lg; b++) r |= ((i >> b) & 1) << (lg - 1 - b);
    p.rev[i] = r;
  }
  return p;
}

static void plan_free(fftplan *p) { free(p->w); free(p->rev); }

static void fft_line(const fftplan *p, cplx *a) {
  int n = p->n;
  for (int i = 0; i < n; i++) {
    int r = p->rev[i];
    if (r > i) { cplx t = a[i]; a[i] = a[r]; a[r] = t; }
  }
  for (int len = 2; len <= n; len <<= 1) {
    int half = len >> 1, step = n / len;
    for (int i = 0; i < n; i += len)
      for (int j = 0; j < half; j++) {
        cplx t = p->w[j * step] * a[i + j + half];
        a[i + j + half] = a[i + j] - t;
        a[i + j] = a[i + j] + t;
      }
  }
}

void orc_fft_c2c(double *data, int len, int count, int sign) {
  fftplan p = plan_make(len, sign);
  cplx *d = (cplx *)data;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < count; c++) fft_line(&p, d + (size_t)c * len);
  plan_free(&p);
}

/* r2c: X[k], k = 0..len/2 (len/2+1 outputs per line) */
static void r2c_line(const fftplan *p, const double *in, cplx *out, cplx *tmp) {
  int n = p->n;
  for (int i = 0; i < n; i++) tmp[i] = in[i];
  fft_line(p, tmp);
  for (int k = 0; k <= n / 2; k++) out[k] = tmp[k];
}

/* c2r: Hermitian extension; imaginary parts of X[0] and X[len/2] ignored */
static void c2r_line(const fftplan *p, const cplx *in, double *out, cplx *tmp) {
  int n = p->n;
  tmp[0] = creal(in[0]);
  for (int k = 1; k < n / 2; k++) { tmp[k] = in[k]; tmp[n - k] = conj(in[k]); }
  tmp[n / 2] = creal(in[n / 2]);
  fft_line(p, tmp);
  for (int i = 0; i < n; i++) out[i] = creal(tmp[i]);
}

void orc_fft_r2c(const double *in, double *out, int len, int count) {
  fftplan p = plan_make(len, -1);
#pragma omp parallel
  {
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * len);
#pragma omp for schedule(static)
    for (int c = 0; c < count; c++)
      r2c_line(&p, in + (size_t)c * len, (cplx *)out + (size_t)c * (len / 2 + 1), tmp);
    free(tmp);
  }
  plan_free(&p);
}

void orc_fft_c2r(const double *in, double *out, int len, int count) {
  fftplan p = plan_make(len, +1);
#pragma omp parallel
  {
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * len);
#pragma omp for schedule(static)
    for (int c = 0; c < count; c++)
      c2r_line(&p, (const cplx *)in + (size_t)c * (len / 2 + 1), out + (size_t)c * len, tmp);
    free(tmp);
  }
  plan_free(&p);
}

/* ---------------------------------------------------------------------- */
/* hot path                                                                 */
/* ---------------------------------------------------------------------- */

/* homogeneous/phshift.f:71-135 (phshift_inplace, called from phase1.f:160):
 * u *= sy(y)*(sx(x)*sz(z)) on the mask, else 0. kstep is 1-based. */
void orc_phshift_inplace(const orc_grid *g, double *uy_, int nv, int kstep) {
  int n = g->n;
  cplx *uy = (cplx *)uy_;
  const cplx *sx = g->sx + (kstep - 1) * g->nxh;
  const cplx *sy = g->sy + (kstep - 1) * n;
  const cplx *sz = g->sz + (kstep - 1) * n;
#pragma omp parallel for schedule(static)
  for (int a = 0; a < g->num_al; a++) {
    cplx sxz = sx[g->col_x[a]] * sz[g->col_z[a]];
    for (int y = 0; y < n; y++) {
      if (orc_mask(g, y, a)) {
        cplx shifty = sy[y] * sxz;
        for (int i = 0; i < nv; i++) uy[((size_t)i * g->num_al + a) * n + y] *= shifty;
      } else {
        for (int i = 0; i < nv; i++) uy[((size_t)i * g->num_al + a) * n + y] = 0;
      }
    }
  }
}

/* inverse 3-D transform of one variable:
 * kxcomm1_clean.f:118 (y c2c +i over the cylinder), kxcomm1_clean.f:205-255
 * (unpack to z-pencils, zero z in (a1,a2) and x > max_al_x),
 * itransform_vel.f:80-94 (zero z Nyquist, z c2c +i),
 * kxcomm2.f:278-333 (zero kx >= nxh, c2r).  zero_ynyq reproduces kxtran.f:36. */
static void inverse_one(const orc_grid *g, const cplx *uy, double *ux, int zero_ynyq,
                        const fftplan *py, const fftplan *pz, const fftplan *px) {
  int n = g->n, nxh = g->nxh, na = g->num_al;
  cplx *yc = (cplx *)malloc(sizeof(cplx) * (size_t)na * n);
  memcpy(yc, uy, sizeof(cplx) * (size_t)na * n);
  if (zero_ynyq)
    for (int a = 0; a < na; a++) yc[(size_t)a * n + n / 2] = 0;
#pragma omp parallel for schedule(static)
  for (int a = 0; a < na; a++) fft_line(py, yc + (size_t)a * n);
  cplx *uz = (cplx *)calloc((size_t)n * nxh * n, sizeof(cplx));
#pragma omp parallel for schedule(static)
  for (int y = 0; y < n; y++)
    for (int a = 0; a < na; a++)
      uz[((size_t)y * nxh + g->col_x[a]) * n + g->col_z[a]] = yc[(size_t)a * n + y];
  free(yc);
#pragma omp parallel for schedule(static)
  for (int l = 0; l < n * nxh; l++) {
    uz[(size_t)l * n + n / 2] = 0;
    fft_line(pz, uz + (size_t)l * n);
  }
#pragma omp parallel
  {
    cplx *line = (cplx *)malloc(sizeof(cplx) * (n / 2 + 1));
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * n);
#pragma omp for schedule(static)
    for (int y = 0; y < n; y++)
      for (int z = 0; z < n; z++) {
        for (int x = 0; x < nxh; x++) line[x] = uz[((size_t)y * nxh + x) * n + z];
        line[nxh] = 0;
        c2r_line(px, line, ux + ((size_t)y * n + z) * n, tmp);
      }
    free(line); free(tmp);
  }
  free(uz);
}

/* itransform_vel.f:11-129 with CLEAN + KXCOMM2: 3 velocity components */
void orc_itransform_vel(const orc_grid *g, const double *uy, double *ux) {
  orc_kxtran(g, uy, ux, -3);
}

/* kxtran.f:17-93 (nv > 0: zero y Nyquist first, kxtran.f:36).
 * nv < 0 is the in-step itransform_vel variant (no y Nyquist zeroing). */
void orc_kxtran(const orc_grid *g, const double *uy, double *ux, int nv) {
  int n = g->n, zero_ynyq = nv > 0;
  if (nv < 0) nv = -nv;
  fftplan py = plan_make(n, +1), pz = plan_make(n, +1), px = plan_make(n, +1);
  for (int v = 0; v < nv; v++)
    inverse_one(g, (const cplx *)uy + (size_t)v * g->num_al * n,
                ux + (size_t)v * n * n * n, zero_ynyq, &py, &pz, &px);
  plan_free(&py); plan_free(&pz); plan_free(&px);
}

/* realspace_vel.f:48-130 (b11=b22=b33=1): slots (1..5) become
 * u^2-v^2, uv, w^2-v^2, vw, uw; velmax = max(nx|u|+ny|v|+nz|w|) on stage 1 */
void orc_realspace_vel(const orc_grid *g, double *ux, int rkstep, double *velmax) {
  size_t n3 = (size_t)g->n * g->n * g->n;
  double xnorm = g->n, vmax = 0.0;
  double *u = ux, *v = ux + n3, *w = ux + 2 * n3, *p4 = ux + 3 * n3, *p5 = ux + 4 * n3;
#pragma omp parallel for schedule(static) reduction(max : vmax)
  for (size_t i = 0; i < n3; i++) {
    double bv2 = v[i] * v[i];
    p5[i] = u[i] * w[i];
    p4[i] = v[i] * w[i];
    if (rkstep == 1) {
      double vel = xnorm * fabs(u[i]) + xnorm * fabs(v[i]) + xnorm * fabs(w[i]);
      if (vel > vmax) vmax = vel;
    }
    w[i] = w[i] * w[i] - bv2;
    v[i] = u[i] * v[i];
    u[i] = u[i] * u[i] - bv2;
  }
  if (velmax && rkstep == 1) *velmax = vmax;
}

/* forward: xkcomm1.f:97,135 (r2c, x 1/nxpad, keep kx < nxh),
 * transform.f:129-138 (z c2c -i, 5 variables), transform.f:169-176,229-234
 * (convective combination, z Nyquist skipped), xkcomm2.f:96-125,190-232
 * (cylinder pack, y c2c -i) for the first nvout variables.
 * mode 0: transform (nvin = 5, nvout = 4, combination on);
 * mode 1: xktran (no combination, extra 1/(ny nz), xktran.f:121-157). */
static void forward_impl(const orc_grid *g, const double *ux, double *uy_, int nvin,
                         int nvout, int mode) {
  int n = g->n, nxh = g->nxh, na = g->num_al;
  cplx *uy = (cplx *)uy_;
  fftplan px = plan_make(n, -1), pz = plan_make(n, -1), py = plan_make(n, -1);
  size_t vsz = (size_t)n * nxh * n;
  cplx *uz = (cplx *)malloc(sizeof(cplx) * vsz * nvin);
  double factor = 1.0 / (double)n;
  double factor_yz = 1.0 / ((double)n * (double)n);
  for (int v = 0; v < nvin; v++) {
    const double *src = ux + (size_t)v * n * n * n;
    cplx *dst = uz + v * vsz;
#pragma omp parallel
    {
      cplx *out = (cplx *)malloc(sizeof(cplx) * (n / 2 + 1));
      cplx *tmp = (cplx *)malloc(sizeof(cplx) * n);
#pragma omp for schedule(static)
      for (int y = 0; y < n; y++)
        for (int z = 0; z < n; z++) {
          r2c_line(&px, src + ((size_t)y * n + z) * n, out, tmp);
          for (int x = 0; x < nxh; x++) dst[((size_t)y * nxh + x) * n + z] = out[x] * factor;
        }
      free(out); free(tmp);
    }
#pragma omp parallel for schedule(static)
    for (int l = 0; l < n * nxh; l++) {
      fft_line(&pz, dst + (size_t)l * n);
      if (mode == 1)
        for (int z = 0; z < n; z++) dst[(size_t)l * n + z] *= factor_yz;
    }
  }
  if (mode == 0) {
    cplx *u1 = uz, *u3 = uz + 2 * vsz, *u5 = uz + 4 * vsz;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < n; y++)
      for (int x = 0; x < nxh; x++) {
        cplx bk1i = I * (double)x;
        for (int z = 0; z < n; z++) {
          if (z == n / 2) continue;
          cplx bk3i = I * (double)kwave(z, n);
          size_t i = ((size_t)y * nxh + x) * n + z;
          u1[i] = bk1i * u1[i] + bk3i * u5[i];
          u3[i] = bk1i * u5[i] + bk3i * u3[i];
        }
      }
  }
  for (int v = 0; v < nvout; v++) {
    cplx *dst = uy + (size_t)v * na * n;
    const cplx *src = uz + v * vsz;
#pragma omp parallel for schedule(static)
    for (int a = 0; a < na; a++) {
      for (int y = 0; y < n; y++)
        dst[(size_t)a * n + y] = src[((size_t)y * nxh + g->col_x[a]) * n + g->col_z[a]];
      fft_line(&py, dst + (size_t)a * n);
    }
  }
  free(uz);
  plan_free(&px); plan_free(&pz); plan_free(&py);
}

/* transform.f:10-301 (nc = 0): 5 products -> 4 cylinder variables
 * (C1 = i kx P1 + i kz P5, P2, C3 = i kx P5 + i kz P3, P4) */
void orc_transform(const orc_grid *g, const double *ux, double *uy) {
  forward_impl(g, ux, uy, 5, 4, 0);
}

/* xktran.f:94-181: fully normalised forward transform of nv variables */
void orc_xktran(const orc_grid *g, const double *ux, double *uy, int nv) {
  forward_impl(g, ux, uy, nv, nv, 1);
}

/* wavespace_vel.f:80,162-174 (non-ROTD branch, b=1):
 * norm = dt2/ny/nz; sxyz = -norm conj(sxz sy);
 * N1 = sxyz C1 + i ky sxyz P2; N2 = sxyz i (kx P2 + kz P4); N3 = sxyz C3 + i ky sxyz P4;
 * off-mask: all 5 slots zeroed. uy holds 5 variable slots. */
void orc_wavespace_vel(const orc_grid *g, double *uy_, int kstep, double dt2) {
  int n = g->n, na = g->num_al;
  cplx *uy = (cplx *)uy_;
  double norm = dt2 / n / n;
  const cplx *sx = g->sx + (kstep - 1) * g->nxh;
  const cplx *sy = g->sy + (kstep - 1) * n;
  const cplx *sz = g->sz + (kstep - 1) * n;
  size_t vs = (size_t)na * n;
#pragma omp parallel for schedule(static)
  for (int a = 0; a < na; a++) {
    int x = g->col_x[a], z = g->col_z[a];
    cplx sxz = sz[z] * sx[x];
    double bk1 = x, bk3 = kwave(z, n);
    for (int y = 0; y < n; y++) {
      size_t i = (size_t)a * n + y;
      if (orc_mask(g, y, a)) {
        cplx sxyz = -norm * conj(sxz * sy[y]);
        cplx utmp1 = (I * (double)kwave(y, n)) * sxyz;
        cplx c1 = uy[i], p2 = uy[vs + i], c3 = uy[2 * vs + i], p4 = uy[3 * vs + i];
        uy[i] = sxyz * c1 + utmp1 * p2;
        uy[vs + i] = (sxyz * I) * (bk1 * p2 + bk3 * p4);
        uy[2 * vs + i] = sxyz * c3 + utmp1 * p4;
      } else {
        for (int v = 0; v < 5; v++) uy[v * vs + i] = 0;
      }
    }
  }
}

/* step.f:366-385 (a11=a22=a33=0, unit metric): xfac = exp(-viscos kx^2 dt) */
void orc_step_factors(const orc_grid *g, double viscos, double dt, double *xfac,
                      double *yfac, double *zfac) {
  int n = g->n;
  for (int x = 0; x < g->nxh; x++) xfac[x] = exp(-viscos * (double)x * x * dt);
  for (int y = 0; y < n; y++) { double k = kwave(y, n); yfac[y] = exp(-viscos * k * k * dt); }
  for (int z = 0; z < n; z++) { double k = kwave(z, n); zfac[z] = exp(-viscos * k * k * dt); }
}

/* homogeneous/proc4a.f:540-571 / proc4b.f:85-106: psi = -(k.u)/k^2, u += k psi */
static inline void project(cplx *u1, cplx *u2, cplx *u3, double kx, double ky, double kz) {
  double bk12 = kx * kx + ky * ky;
  cplx psi = -(kx * *u1 + ky * *u2 + kz * *u3) / (bk12 + kz * kz);
  *u1 = *u1 + kx * psi;
  *u2 = *u2 + ky * psi;
  *u3 = *u3 + kz * psi;
}

/* homogeneous/proc4a.f:418-487 (integrating factor, predictor) and
 * proc4a.f:535-573 (projection of the new u; k = 0 left alone) */
void orc_proc4a(const orc_grid *g, double *uy_, double *uny_, const double *xfac,
                const double *yfac, const double *zfac) {
  int n = g->n, na = g->num_al;
  cplx *uy = (cplx *)uy_, *uny = (cplx *)uny_;
  size_t vs = (size_t)na * n;
#pragma omp parallel for schedule(static)
  for (int a = 0; a < na; a++) {
    int x = g->col_x[a], z = g->col_z[a];
    for (int y = 0; y < n; y++) {
      if (!orc_mask(g, y, a)) continue;
      size_t i = (size_t)a * n + y;
      double facty = xfac[x] * yfac[y] * zfac[z];
      for (int v = 0; v < 3; v++) {
        cplx untemp = uny[v * vs + i];
        uny[v * vs + i] = facty * (untemp + uy[v * vs + i]);
        uy[v * vs + i] = facty * (untemp + uy[v * vs + i] * 2.0);
      }
      if (x == 0 && z == 0 && y == 0) continue;
      project(uy + i, uy + vs + i, uy + 2 * vs + i, x, kwave(y, n), kwave(z, n));
    }
  }
}

/* homogeneous/proc4b.f:54-107 (corrector + projection), 176-205
 * (uny <- uy on the mask, zero mean) */
void orc_proc4b(const orc_grid *g, double *uy_, double *uny_) {
  int n = g->n, na = g->num_al;
  cplx *uy = (cplx *)uy_, *uny = (cplx *)uny_;
  size_t vs = (size_t)na * n;
#pragma omp parallel for schedule(static)
  for (int a = 0; a < na; a++) {
    int x = g->col_x[a], z = g->col_z[a];
    for (int y = 0; y < n; y++) {
      if (!orc_mask(g, y, a)) continue;
      size_t i = (size_t)a * n + y;
      for (int v = 0; v < 3; v++) uy[v * vs + i] = uny[v * vs + i] + uy[v * vs + i];
      if (!(x == 0 && z == 0 && y == 0))
        project(uy + i, uy + vs + i, uy + 2 * vs + i, x, kwave(y, n), kwave(z, n));
      for (int v = 0; v < 3; v++) uny[v * vs + i] = uy[v * vs + i];
    }
  }
  for (int v = 0; v < 3; v++) { uy[v * vs] = 0; uny[v * vs] = 0; }
}

/* EPFOR forcing increments (the -DEPFOR build, PSDNS/makefile:69).
 * forc = for(kfor, k2fo, k2fo, 3) complex, k2fo = 2 kfor - 1, as force.f
 * fills it: index x = kx + 1 <= kfor; iy = ky + 1 for 0 <= ky < kfor and
 * iy = ky + 2 kfor for -kfor < ky < 0 (proc4a.f:605-626: y <= kfor -> y,
 * y >= ny + 2 - kfor -> y - ny + k2fo), iz from kz the same way.  Entries
 * of the cylinder on the mask whose (kx, ky, kz) fall in that box take
 * for(x, iy, iz, i), added after the projection:
 *  mode 1  proc4a.f:588-640 (= advanc_1.f:165-210): only where
 *          rks = |k|^2 <= kforce^2; efkz(k, i) += 2 Re(u conj(for)) tfact(x)
 *          (u before the add, k = int(sqrt(rks) + 1.5), ssq = 1 for a
 *          cube), efkz zeroed first (proc4a.f:66); u <- u + for
 *  mode 2  proc4b.f:118-163: u <- u + for in the whole box, then (as
 *          proc4b.f:176-205 goes on) un <- u and zero mean
 *  mode 3  advanc_2.f:97-141 / advanc_3.f:98-142: u <- u + for in the box
 * efkz as the reference's efkz(nxh, 3): efkz[i * nxh + k - 1]. */
void orc_epfor(const orc_grid *g, double *uy_, double *uny_, int kfor, double kforce,
               const double *forc_, int mode, double *efkz) {
  const int n = g->n, na = g->num_al, nxh = n / 2, k2fo = 2 * kfor - 1;
  cplx *uy = (cplx *)uy_, *uny = (cplx *)uny_;
  const cplx *forc = (const cplx *)forc_;
  const size_t vs = (size_t)na * n;
  if (mode == 1)
    for (int e = 0; e < 3 * nxh; e++) efkz[e] = 0.0;
  for (int a = 0; a < na; a++) {
    const int x = g->col_x[a], z = g->col_z[a];  /* 0-based: reference x - 1, z - 1 */
    int iz = 0;
    if (z + 1 <= kfor) iz = z + 1;
    else if (z + 1 >= n + 2 - kfor) iz = z + 1 - n + k2fo;
    if (iz == 0 || x + 1 > kfor) continue;
    for (int i = 0; i < 3; i++) {
      for (int y = 0; y < n; y++) {
        if (!orc_mask(g, y, a)) continue;
        int iy = 0;
        if (y + 1 <= kfor) iy = y + 1;
        else if (y + 1 >= n + 2 - kfor) iy = y + 1 - n + k2fo;
        if (iy == 0) continue;
        const cplx f = forc[(((size_t)i * k2fo + (iz - 1)) * k2fo + (iy - 1)) * kfor + x];
        const size_t j = (size_t)i * vs + (size_t)a * n + y;
        if (mode == 1) {
          const double ky = kwave(y, n), kz = kwave(z, n);
          const double rks = (double)x * x + ky * ky + kz * kz;
          if (!(rks <= kforce * kforce)) continue;
          const int k = (int)(sqrt(rks) + 1.5);
          const double term = creal(uy[j] * conj(f));
          if (k <= nxh) efkz[i * nxh + k - 1] += 2.0 * term * (x == 0 ? 1.0 : 2.0);
        }
        uy[j] += f;
      }
    }
  }
  if (mode == 2) {
    for (int i = 0; i < 3; i++)
      for (size_t j = 0; j < vs; j++)
        if (orc_mask(g, (int)(j % n), (int)(j / n))) uny[i * vs + j] = uy[i * vs + j];
    for (int i = 0; i < 3; i++) { uy[i * vs] = 0; uny[i * vs] = 0; }
  }
}

/* ek_shell.f:131-168: shell energies (ik = int(|k| + 1.5), 1-based shells,
 * returned as ek[ik-1]) for ik <= kfor, z != nzhp, on the mask,
 * tfact = 1 at kx = 0 else 2 (comsp_set.f:85-87) */
void orc_ek_shell(const orc_grid *g, const double *uny_, int kfor, double *ek) {
  int n = g->n, na = g->num_al;
  const cplx *uny = (const cplx *)uny_;
  size_t vs = (size_t)na * n;
  for (int k = 0; k < kfor; k++) ek[k] = 0.0;
  for (int a = 0; a < na; a++) {
    int x = g->col_x[a], z = g->col_z[a];
    if (z == n / 2) continue;
    double tf = x == 0 ? 1.0 : 2.0;
    for (int y = 0; y < n; y++) {
      if (!orc_mask(g, y, a)) continue;
      double kz = kwave(z, n), ky = kwave(y, n);
      double rk2 = (double)x * x + ky * ky + kz * kz;
      int ik = (int)(sqrt(rk2) + 1.5);
      if (ik > kfor) continue;
      size_t i = (size_t)a * n + y;
      double t1 = creal(uny[i] * conj(uny[i]));
      double t2 = creal(uny[vs + i] * conj(uny[vs + i]));
      double t3 = creal(uny[2 * vs + i] * conj(uny[2 * vs + i]));
      ek[ik - 1] += (t1 + t2 + t3) / 2.0 * tf;
    }
  }
}

/* ek_force.f:27-71: ratio(1) = 0, ratio(ik) = sqrt(ekinf(ik)/ek(ik)),
 * uny and uy scaled in shells ik <= kf_shell.  ekinf[ik-1] per shell. */
void orc_ek_force(const orc_grid *g, double *uy_, double *uny_, int kf_shell, int kfor,
                  const double *ekinf) {
  int n = g->n, na = g->num_al;
  cplx *uy = (cplx *)uy_, *uny = (cplx *)uny_;
  size_t vs = (size_t)na * n;
  double *ek = (double *)calloc(kfor, sizeof(double));
  double *ratio = (double *)calloc(kfor, sizeof(double));
  orc_ek_shell(g, uny_, kfor, ek);
  ratio[0] = 0.0;
  for (int ik = 2; ik <= kfor; ik++) ratio[ik - 1] = sqrt(ekinf[ik - 1] / ek[ik - 1]);
  for (int a = 0; a < na; a++) {
    int x = g->col_x[a], z = g->col_z[a];
    if (z == n / 2) continue;
    for (int y = 0; y < n; y++) {
      if (!orc_mask(g, y, a)) continue;
      double kz = kwave(z, n), ky = kwave(y, n);
      double rk2 = (double)x * x + ky * ky + kz * kz;
      int ik = (int)(sqrt(rk2) + 1.5);
      if (ik > kf_shell) continue;
      size_t i = (size_t)a * n + y;
      for (int v = 0; v < 3; v++) {
        uny[v * vs + i] *= ratio[ik - 1];
        uy[v * vs + i] *= ratio[ik - 1];
      }
    }
  }
  free(ek); free(ratio);
}

/* homogeneous/sptr.f:139-176,235: E(ik) = sum tfact |u|^2 / 2,
 * D(ik) = 2 viscos sum k^2 tfact |u|^2 / 2, eps = sum D (ik = int(|k|+1.5)) */
void orc_spectrum(const orc_grid *g, const double *uny_, double viscos, int nshell,
                  double *ek, double *dk, double *tke, double *eps) {
  int n = g->n, na = g->num_al;
  const cplx *uny = (const cplx *)uny_;
  size_t vs = (size_t)na * n;
  for (int k = 0; k < nshell; k++) { ek[k] = 0; dk[k] = 0; }
  double t = 0, e = 0;
  for (int a = 0; a < na; a++) {
    int x = g->col_x[a], z = g->col_z[a];
    double tf = x == 0 ? 1.0 : 2.0;
    for (int y = 0; y < n; y++) {
      if (!orc_mask(g, y, a)) continue;
      double kz = kwave(z, n), ky = kwave(y, n);
      double rk2 = (double)x * x + ky * ky + kz * kz;
      int ik = (int)(sqrt(rk2) + 1.5);
      size_t i = (size_t)a * n + y;
      double en = 0.5 * tf * (creal(uny[i] * conj(uny[i])) + creal(uny[vs + i] * conj(uny[vs + i])) +
                              creal(uny[2 * vs + i] * conj(uny[2 * vs + i])));
      t += en;
      e += 2.0 * viscos * rk2 * en;
      if (ik <= nshell) { ek[ik - 1] += en; dk[ik - 1] += 2.0 * viscos * rk2 * en; }
    }
  }
  *tke = t;
  *eps = e;
}

/* velocity-gradient moments (vgmomt.f:68-143): du/dx = i kx u and dw/dz =
 * i kz w of the state on the mask, inverse transformed (kxtran), then the
 * sums of the powers 1..8 over the grid divided by n^3 (lines 70-87 tally,
 * 140-143 normalise). */
void orc_vg_moments(const orc_grid *g, const double *uny_, double *dudx, double *dwdz) {
  const int n = g->n, na = g->num_al;
  const size_t vs = (size_t)na * n, n3 = (size_t)n * n * n;
  const cplx *uny = (const cplx *)uny_;
  cplx *d = calloc(2 * vs, sizeof(cplx));
  double *f = malloc(2 * n3 * sizeof(double));
#pragma omp parallel for schedule(static)
  for (int a = 0; a < na; a++) {
    const double kx = g->col_x[a], kz = kwave(g->col_z[a], n);
    for (int y = 0; y < n; y++) {
      const size_t i = (size_t)a * n + y;
      if (!orc_mask(g, y, a)) continue;
      d[i] = I * kx * uny[i];
      d[vs + i] = I * kz * uny[2 * vs + i];
    }
  }
  orc_kxtran(g, (const double *)d, f, 2);
  for (int v = 0; v < 2; v++) {
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const double *fv = f + v * n3;
#pragma omp parallel for schedule(static) reduction(+ : s[:8])
    for (size_t i = 0; i < n3; i++) {
      double p = fv[i];
      for (int m = 0; m < 8; m++) {
        s[m] += p;
        p *= fv[i];
      }
    }
    double *out = v == 0 ? dudx : dwdz;
    for (int m = 0; m < 8; m++) out[m] = s[m] / (double)n3;
  }
  free(d);
  free(f);
}

/* ---------------------------------------------------------------------- */
/* initial field                                                            */
/* ---------------------------------------------------------------------- */
/* splitmix64 finaliser: uniform doubles keyed by (seed, x, y, z, stream) */
static inline uint64_t mix64(uint64_t v) {
  v += 0x9E3779B97F4A7C15ULL;
  v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ULL;
  v = (v ^ (v >> 27)) * 0x94D049BB133111EBULL;
  return v ^ (v >> 31);
}
static inline double uni(uint64_t seed, int x, int y, int z, int s) {
  uint64_t key = ((uint64_t)x << 42) ^ ((uint64_t)y << 21) ^ (uint64_t)z;
  uint64_t r = mix64(mix64(seed ^ mix64(key)) + (uint64_t)s);
  return (double)(r >> 11) * (1.0 / 9007199254740992.0);
}

/* homogeneous/field.f:172-257 (kinit = -1, unit metric): random-phase
 * solenoidal modes with |u|^2 = ef^2, ef = sqrt(E(k))/(sqrt(2 pi) k),
 * E(k) = k^4 exp(-2 (k/kpeak)^2); conjugate symmetry on kx = 0 as
 * homogeneous/consym.f; rescaled so that the sptr-style TKE equals `tke`. */
void orc_field(const orc_grid *g, double *uny_, unsigned long long seed, double kpeak,
               double tke) {
  int n = g->n, na = g->num_al;
  cplx *uny = (cplx *)uny_;
  size_t vs = (size_t)na * n;
  memset(uny_, 0, sizeof(cplx) * 3 * vs);
  double rpid = 1.0 / sqrt(2.0 * M_PI);
#pragma omp parallel for schedule(static)
  for (int a = 0; a < na; a++) {
    int x = g->col_x[a], z = g->col_z[a];
    for (int y = 0; y < n; y++) {
      if (!orc_mask(g, y, a)) continue;
      if (x == 0 && z == 0 && y == 0) continue;
      double k1 = x, k2 = kwave(y, n), k3 = kwave(z, n);
      double k = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
      double ekx = k * k * k * k * exp(-2.0 * (k / kpeak) * (k / kpeak));
      double ef = sqrt(ekx) * rpid / k;
      double th1 = uni(seed, x, y, z, 0), th2 = uni(seed, x, y, z, 1), ph = uni(seed, x, y, z, 2);
      double angphi = 2.0 * M_PI * ph;
      cplx alpha = ef * (cos(2.0 * M_PI * th1) + I * sin(2.0 * M_PI * th1)) * cos(angphi);
      cplx beta = ef * (cos(2.0 * M_PI * th2) + I * sin(2.0 * M_PI * th2)) * sin(angphi);
      double k12 = sqrt(k1 * k1 + k2 * k2);
      size_t i = (size_t)a * n + y;
      if (k12 == 0.0) {
        uny[i] = alpha; uny[vs + i] = beta; uny[2 * vs + i] = 0;
      } else {
        cplx c1 = alpha * k, c2 = beta * k3;
        double rt = 1.0 / k / k12;
        uny[i] = (c1 * k2 + c2 * k1) * rt;
        uny[vs + i] = (c2 * k2 - c1 * k1) * rt;
        uny[2 * vs + i] = -beta * k12 / k;
      }
    }
  }
  /* consym.f (FRESH branch): on kx = 0, the column kz = 0 gets
   * u(-ky) = conj(u(ky)); columns with kz < 0 get u(-ky,-kz) = conj(u(ky,kz)).
   * Column a for (x=0, z) is its index among x=0 columns. */
  int *colz0 = (int *)malloc(sizeof(int) * n);
  for (int z = 0; z < n; z++) colz0[z] = -1;
  for (int a = 0; a < na && g->col_x[a] == 0; a++) colz0[g->col_z[a]] = a;
  for (int v = 0; v < 3; v++) {
    cplx *u = uny + v * vs;
    int a0 = colz0[0];
    u[(size_t)a0 * n] = creal(u[(size_t)a0 * n]);
    for (int y = n / 2 + 1; y < n; y++) u[(size_t)a0 * n + y] = conj(u[(size_t)a0 * n + (n - y)]);
    for (int z = n / 2 + 1; z < n; z++) {
      int a = colz0[z], ao = colz0[n - z];
      if (a < 0 || ao < 0) continue;
      for (int y = 0; y < n; y++)
        u[(size_t)a * n + y] = conj(u[(size_t)ao * n + ((n - y) % n)]);
    }
  }
  free(colz0);
  double ek[1], dk[1], t, e;
  orc_spectrum(g, uny_, 0.0, 0, ek, dk, &t, &e);
  double s = sqrt(tke / t);
  for (size_t i = 0; i < 3 * vs; i++) uny[i] *= s;
}

/* ------------------------------------------------------------------------
 * Lagrangian particles (hooks/LAG_LOCAL), one task.
 * ---------------------------------------------------------------------- */

/* solsp1.f:9-43: factors of the periodic tridiagonal system with 4 on the
 * diagonal and 1 off it (Ahlberg et al. 1967), p already inverted */
typedef struct { int n; double *p, *q, *t, d; } spl_fac;

static spl_fac solsp1(int n) {
  spl_fac f = {n, malloc(sizeof(double) * (n + 1)), malloc(sizeof(double) * (n + 1)),
               malloc(sizeof(double) * (n + 1)), 0.0};
  double *s = malloc(sizeof(double) * (n + 1));
  f.q[0] = 0.0;
  for (int i = 1; i <= n; i++) {
    f.p[i] = f.q[i - 1] + 4.0;
    f.q[i] = -1.0 / f.p[i];
  }
  s[0] = 1.0;
  for (int i = 1; i <= n; i++) s[i] = -s[i - 1] / f.p[i];
  f.t[n] = 1.0;
  for (int i = n - 1; i >= 1; i--) f.t[i] = f.q[i] * f.t[i + 1] + s[i];
  f.d = 1.0 / (f.t[1] + f.t[n - 1] + 4.0);
  for (int i = 1; i <= n; i++) f.p[i] = 1.0 / f.p[i];
  free(s);
  return f;
}

static void spl_free(spl_fac *f) { free(f->p); free(f->q); free(f->t); }

/* one line as splx.f:24-50 (sply.f, splz.f with the data scaled by 216):
 * data fv[(i-1) * st], i = 1..n (node i-1) -> coefficients of nodes 0..n-1
 * in place (the reference's bs(i+1), with bs(1), bs(n+2), bs(n+3) their
 * periodic copies) */
static void spl_line(const spl_fac *f, double *fv, long st, double scale, double *u, double *v, double *bs) {
  const int n = f->n;
  u[0] = 0.0;
  for (int x = 1; x <= n; x++) u[x] = (fv[(x - 1) * st] * scale - u[x - 1]) * f->p[x];
  v[n] = 0.0;
  for (int x = n - 1; x >= 1; x--) v[x] = f->q[x] * v[x + 1] + u[x];
  bs[n + 1] = f->d * (fv[(n - 1) * st] * scale - v[1] - v[n - 1]);
  for (int x = n - 1; x >= 1; x--) bs[x + 1] = f->t[x] * bs[n + 1] + v[x];
  for (int x = 1; x <= n; x++) fv[(x - 1) * st] = bs[x + 1];
}

/* spxyz_m.f: the tricubic-spline coefficients of a periodic field
 * f[y][z][x] (n^3, in place): x lines (splx.f), z lines with the data
 * scaled by 216 (splz.f:22-28), y lines (sply.f) */
void orc_spline3(int n, double *f) {
  spl_fac fac = solsp1(n);
#pragma omp parallel
  {
    double *u = malloc(sizeof(double) * (n + 1)), *v = malloc(sizeof(double) * (n + 1)),
           *bs = malloc(sizeof(double) * (n + 4));
#pragma omp for schedule(static)
    for (long yz = 0; yz < (long)n * n; yz++) spl_line(&fac, f + yz * n, 1, 1.0, u, v, bs);
#pragma omp for schedule(static)
    for (long yx = 0; yx < (long)n * n; yx++) {
      const long y = yx / n, x = yx % n;
      spl_line(&fac, f + y * n * n + x, n, 216.0, u, v, bs);
    }
#pragma omp for schedule(static)
    for (long zx = 0; zx < (long)n * n; zx++) spl_line(&fac, f + zx, (long)n * n, 1.0, u, v, bs);
    free(u); free(v); free(bs);
  }
  spl_free(&fac);
}

/* spcal_local.f:119-364 (iod = 0, sf = 1): for particle m at pp(m, 1:3)
 * (grid units, gx(i) = 1 + (i - 1) + gsh), the value of nv spline fields
 * c[v][y][z][x] into pp(m, 4:3+nv); pp is the reference's column-major
 * pp(np, 9): pp[(j - 1) np + m].  Same basis functions, stencil (nodes
 * ibx - 1 .. ibx + 2, periodic) and summation order as the reference. */
void orc_lag_interp(int n, const double *c, int nv, const double gsh[3], long np, double *pp) {
  const double d6 = 1.0 / 6.0, tt = 2.0 / 3.0;
  const size_t n3 = (size_t)n * n * n;
#pragma omp parallel for schedule(static)
  for (long m = 0; m < np; m++) {
    double bff[4][3];
    int ib[3];
    for (int j = 0; j < 3; j++) {
      const double px1 = pp[j * np + m] - (gsh[j] + 1.0);
      const double px = px1 - (double)n;
      const long ipx = (long)floor(px);
      long jpx = ipx;
      if (ipx < 0) jpx = ipx + 1000L * n;
      const long jloc = jpx % n;
      const double bfp = (double)jloc + px - (double)ipx;
      ib[j] = (int)bfp;
      const double pos = bfp - ib[j], ym = 1.0 - pos, yms = ym * ym;
      bff[0][j] = ym * yms * d6;
      bff[1][j] = tt + 0.5 * pos * (yms - 1.0);
      bff[3][j] = pos * pos * pos * d6;
      bff[2][j] = 1.0 - bff[0][j] - bff[1][j] - bff[3][j];
    }
    double func[3] = {0, 0, 0};
    for (int k = 0; k < 4; k++) {
      const int z = ((ib[2] + k - 1) % n + n) % n;
      double sw[3] = {0, 0, 0};
      for (int i = 0; i < 4; i++) {
        const int x = ((ib[0] + i - 1) % n + n) % n;
        const double cst = bff[i][0];
        for (int iv = 0; iv < nv; iv++) {
          double bt[4];
          for (int jj = 0; jj < 4; jj++) {
            const int y = ((ib[1] + jj - 1) % n + n) % n;
            bt[jj] = c[iv * n3 + ((size_t)y * n + z) * n + x];
          }
          const double term1 = bt[0] * bff[0][1] + bt[2] * bff[2][1];
          const double term2 = bt[1] * bff[1][1] + bt[3] * bff[3][1];
          sw[iv] += (term1 + term2) * cst;
        }
      }
      for (int iv = 0; iv < nv; iv++) func[iv] += sw[iv] * bff[k][2];
    }
    for (int iv = 0; iv < nv; iv++) pp[(3 + iv) * np + m] = func[iv];
  }
}

/* rk_advpos.f:15-52 (RK2, b = 1): fact = cof dt / (2 pi) n = cof dt / dx;
 * pstep 1: pp(:,7:9) = x + fact1 u, x = pp(:,7:9) + fact2 u;
 * pstep 2: x = pp(:,7:9) + fact2 u */
void orc_lag_advpos(int n, long np, double *pp, int pstep, double dt, double cof1, double cof2) {
  const double pi = 3.14159265358979323846;
  for (int k = 0; k < 3; k++) {
    const double fact1 = cof1 * dt / 2.0 / pi * n, fact2 = cof2 * dt / 2.0 / pi * n;
    double *x = pp + k * np, *u = pp + (3 + k) * np, *h = pp + (6 + k) * np;
    for (long i = 0; i < np; i++) {
      if (pstep == 1) {
        h[i] = x[i] + fact1 * u[i];
        x[i] = h[i] + fact2 * u[i];
      } else {
        x[i] = h[i] + fact2 * u[i];
      }
    }
  }
}
