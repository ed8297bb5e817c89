/*
 * orac_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the ORAC
 * nonbonded force path: an FP64 restatement of the reference routines.
 *
 * Only tests, smoke() and bench.py's cpu_baseline leg use this file, as
 * the checker.  Expressions keep the reference's association order and
 * the build disables FMA contraction, so the serial path reproduces the
 * reference's serial build bit for bit wherever the reference is
 * deterministic.  Every function cites the reference lines it restates.
 *
 * The reference's 3-D FFT is FFTW3 (pme_fftw3.f90:114,116), which is not
 * part of the reference tree; this file has its own mixed-radix FFT with
 * the same conventions (unnormalised, forward e^{-i}, r2c/c2r over a
 * K1-fastest grid, c2r ignoring the imaginary parts of the k1 = 0 and
 * Nyquist planes).
 */
#include "orac_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#pragma STDC FP_CONTRACT OFF

/* A&S 7.1.26 coefficients, force.f:123-125 (also ferrf.f, cself.f, fnbgrp.f) */
static const double A1 = 0.2548296, A2 = -0.28449674, A3 = 1.4214137, A4 = -1.453152, A5 = 1.0614054,
                    QP = 0.3275911;
/* the reference's default-real pi literal (scalar_sum.f:18, pme_fftw3.f90:214) */
static const double PI_F32 = (double)3.14159265358979323846f;

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static double fnint(double x) { return round(x); } /* Fortran ANINT/DNINT: half away from zero */

/* ------------------------------------------------------------------ */
/* B-splines: pme_fftw3.f90:537-576                                    */
/* ------------------------------------------------------------------ */
static void one_pass(double *c, double x, int k) {
  double div = 1.0 / (k - 1);
  c[k - 1] = div * x * c[k - 2];
  for (int j = 1; j <= k - 2; ++j)
    c[k - j - 1] = div * ((x + j) * c[k - j - 2] + ((double)(k - j) - x) * c[k - j - 1]);
  c[0] = div * (1 - x) * c[0];
}

void orc_fill_bspline(double w, int order, double *array, double *darray) {
  array[order - 1] = 0.0; /* init */
  array[1] = w;
  array[0] = 1.0 - w;
  for (int k = 3; k <= order - 1; ++k) one_pass(array, w, k);
  darray[0] = -array[0]; /* diff */
  for (int j = 2; j <= order; ++j) darray[j - 1] = array[j - 2] - array[j - 1];
  one_pass(array, w, order);
}

/* pme_fftw3.f90:165-230 */
void orc_load_bsp_moduli(int order, int nfft, double *bsp_mod) {
  double array[32], darray[32];
  double *bsp_arr = (double *)calloc((size_t)nfft + order + 2, sizeof(double));
  orc_fill_bspline(0.0, order, array, darray);
  for (int i = 2; i <= order + 1; ++i) bsp_arr[i - 1] = array[i - 2];
  const double twopi = 2.0 * PI_F32;
  for (int k = 1; k <= nfft; ++k) {
    double s1 = 0.0, s2 = 0.0;
    for (int j = 1; j <= nfft; ++j) {
      double arg = twopi * (k - 1) * (j - 1) / nfft;
      s1 = s1 + bsp_arr[j - 1] * cos(arg);
      s2 = s2 + bsp_arr[j - 1] * sin(arg);
    }
    bsp_mod[k - 1] = s1 * s1 + s2 * s2;
  }
  for (int k = 1; k <= nfft; ++k)
    if (bsp_mod[k - 1] < 1.0e-7) bsp_mod[k - 1] = 0.5 * (bsp_mod[k - 2] + bsp_mod[k]);
  free(bsp_arr);
}

/* pme_fftw3.f90:132-163 */
void orc_scaled_fractionals(int n, const double *x, const double *y, const double *z, const double *recip,
                            int k1, int k2, int k3, double *fr1, double *fr2, double *fr3) {
  for (int i = 0; i < n; ++i) {
    double w = x[i] * recip[0] + y[i] * recip[1] + z[i] * recip[2];
    fr1[i] = k1 * (w - fnint(w) + 0.5);
    w = x[i] * recip[3] + y[i] * recip[4] + z[i] * recip[5];
    fr2[i] = k2 * (w - fnint(w) + 0.5);
    w = x[i] * recip[6] + y[i] * recip[7] + z[i] * recip[8];
    fr3[i] = k3 * (w - fnint(w) + 0.5);
  }
}

/* pme_fftw3.f90:487-512 */
void orc_bspline_coeffs(int n, const double *fr1, const double *fr2, const double *fr3, int order,
                        double *th1, double *th2, double *th3, double *dth1, double *dth2, double *dth3) {
  for (int i = 0; i < n; ++i) {
    orc_fill_bspline(fr1[i] - (int)fr1[i], order, th1 + (size_t)i * order, dth1 + (size_t)i * order);
    orc_fill_bspline(fr2[i] - (int)fr2[i], order, th2 + (size_t)i * order, dth2 + (size_t)i * order);
    orc_fill_bspline(fr3[i] - (int)fr3[i], order, th3 + (size_t)i * order, dth3 + (size_t)i * order);
  }
}

/* 0-based grid index for the 0-based offset k0 (pme_fftw3.f90:274: k0+1+(K-isign(K,k0))/2, 1-based) */
static inline int wrapidx(int k0, int K) { return k0 >= 0 ? k0 : k0 + K; }

/* pme_fftw3.f90:234-298; the OpenMP build reduces private grids (REDUCTION(+:Q)) */
void orc_fill_charge_grid(int n, const double *q, const double *th1, const double *th2, const double *th3,
                          const double *fr1, const double *fr2, const double *fr3, int order,
                          int k1, int k2, int k3, double *Q, int nthreads) {
  size_t G = (size_t)k1 * k2 * k3;
  memset(Q, 0, G * sizeof(double));
  if (nthreads <= 1) {
    for (int a = 0; a < n; ++a) {
      int kk0 = (int)fr3[a] - order;
      for (int i3 = 1; i3 <= order; ++i3) {
        kk0 += 1;
        int kk = wrapidx(kk0, k3);
        int jj0 = (int)fr2[a] - order;
        for (int i2 = 1; i2 <= order; ++i2) {
          jj0 += 1;
          int jj = wrapidx(jj0, k2);
          double prod = th2[(size_t)a * order + i2 - 1] * th3[(size_t)a * order + i3 - 1] * q[a];
          int ii0 = (int)fr1[a] - order;
          for (int i1 = 1; i1 <= order; ++i1) {
            ii0 += 1;
            int ii = wrapidx(ii0, k1);
            size_t g = (size_t)ii + (size_t)k1 * (jj + (size_t)k2 * kk);
            Q[g] = Q[g] + th1[(size_t)a * order + i1 - 1] * prod;
          }
        }
      }
    }
    return;
  }
#ifdef _OPENMP
  double *priv = (double *)calloc(G * (size_t)nthreads, sizeof(double));
#pragma omp parallel num_threads(nthreads)
  {
    int t = omp_get_thread_num();
    double *Qt = priv + G * (size_t)t;
#pragma omp for schedule(static)
    for (int a = 0; a < n; ++a) {
      int kk0 = (int)fr3[a] - order;
      for (int i3 = 1; i3 <= order; ++i3) {
        kk0 += 1;
        int kk = wrapidx(kk0, k3);
        int jj0 = (int)fr2[a] - order;
        for (int i2 = 1; i2 <= order; ++i2) {
          jj0 += 1;
          int jj = wrapidx(jj0, k2);
          double prod = th2[(size_t)a * order + i2 - 1] * th3[(size_t)a * order + i3 - 1] * q[a];
          int ii0 = (int)fr1[a] - order;
          for (int i1 = 1; i1 <= order; ++i1) {
            ii0 += 1;
            int ii = wrapidx(ii0, k1);
            size_t g = (size_t)ii + (size_t)k1 * (jj + (size_t)k2 * kk);
            Qt[g] += th1[(size_t)a * order + i1 - 1] * prod;
          }
        }
      }
    }
#pragma omp for schedule(static)
    for (long g = 0; g < (long)G; ++g) {
      double s = 0.0;
      for (int tt = 0; tt < nthreads; ++tt) s += priv[G * (size_t)tt + g];
      Q[g] = s;
    }
  }
  free(priv);
#endif
}

/* ------------------------------------------------------------------ */
/* mixed-radix complex FFT (recursive, radix 2/3/4/5 + generic prime)  */
/* ------------------------------------------------------------------ */
typedef struct { double re, im; } cpx;

static void fft_rec(int n, const cpx *in, int stride, cpx *out, int sign, const cpx *tw, int twstride, cpx *scratch) {
  if (n == 1) { out[0] = in[0]; return; }
  int r;
  if (n % 4 == 0) r = 4;
  else if (n % 2 == 0) r = 2;
  else if (n % 3 == 0) r = 3;
  else if (n % 5 == 0) r = 5;
  else {
    for (r = 7; r * r <= n; r += 2) if (n % r == 0) break;
    if (r * r > n) r = n;
  }
  int m = n / r;
  /* sub-transforms of the r decimated sequences */
  for (int s = 0; s < r; ++s)
    fft_rec(m, in + (size_t)s * stride, stride * r, out + (size_t)s * m, sign, tw, twstride * r, scratch);
  /* butterflies: X[k + q m] = sum_s W_n^{s(k+q m)} Y_s[k] */
  cpx y[64];
  cpx *yy = r <= 64 ? y : scratch;
  for (int k = 0; k < m; ++k) {
    for (int s = 0; s < r; ++s) yy[s] = out[(size_t)s * m + k];
    for (int qq = 0; qq < r; ++qq) {
      double sr = 0.0, si = 0.0;
      for (int s = 0; s < r; ++s) {
        long e = ((long)s * (k + (long)qq * m)) % n;
        cpx w = tw[e * twstride];
        double wi = sign < 0 ? w.im : -w.im;
        sr += yy[s].re * w.re - yy[s].im * wi;
        si += yy[s].re * wi + yy[s].im * w.re;
      }
      out[(size_t)qq * m + k].re = sr;
      out[(size_t)qq * m + k].im = si;
    }
  }
}

/* tw[e] = exp(-2 pi i e / n) */
static cpx *make_twiddles(int n) {
  cpx *tw = (cpx *)malloc(sizeof(cpx) * (size_t)n);
  for (int e = 0; e < n; ++e) {
    double a = -2.0 * M_PI * (double)e / (double)n;
    tw[e].re = cos(a);
    tw[e].im = sin(a);
  }
  return tw;
}

static void fft_line(int n, cpx *data, int sign, const cpx *tw, cpx *work, cpx *scratch) {
  fft_rec(n, data, 1, work, sign, tw, 1, scratch);
  memcpy(data, work, sizeof(cpx) * (size_t)n);
}

void orc_fft1d(int n, double *data, int sign) {
  cpx *tw = make_twiddles(n);
  cpx *work = (cpx *)malloc(sizeof(cpx) * (size_t)n * 2);
  fft_line(n, (cpx *)data, sign, tw, work, work + n);
  free(work);
  free(tw);
}

/* transform the complex array a(n1,n2,n3) (n1 fastest) along one axis */
static void fft_axis(cpx *a, int n1, int n2, int n3, int axis, int sign, int nthreads) {
  int n = axis == 0 ? n1 : (axis == 1 ? n2 : n3);
  cpx *tw = make_twiddles(n);
  long nlines = (long)n1 * n2 * n3 / n;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
    cpx *buf = (cpx *)malloc(sizeof(cpx) * (size_t)n * 3);
#pragma omp for schedule(static)
    for (long l = 0; l < nlines; ++l) {
      size_t base, step;
      if (axis == 0) { base = (size_t)l * n1; step = 1; }
      else if (axis == 1) { long i = l % n1, k = l / n1; base = (size_t)i + (size_t)n1 * n2 * k; step = n1; }
      else { base = (size_t)l; step = (size_t)n1 * n2; }
      for (int t = 0; t < n; ++t) buf[t] = a[base + step * t];
      fft_line(n, buf, sign, tw, buf + n, buf + 2 * n);
      for (int t = 0; t < n; ++t) a[base + step * t] = buf[t];
    }
    free(buf);
  }
  free(tw);
}

void orc_fft_r2c(int k1, int k2, int k3, const double *Q, double *FQ, int nthreads) {
  int h = k1 / 2 + 1;
  cpx *line = NULL;
  cpx *F = (cpx *)FQ;
  /* x: full complex transform of each real line, keep k1 = 0..K1/2 */
  cpx *tw = make_twiddles(k1);
  line = (cpx *)malloc(sizeof(cpx) * (size_t)k1 * 3);
  for (long l = 0; l < (long)k2 * k3; ++l) {
    for (int i = 0; i < k1; ++i) { line[i].re = Q[(size_t)l * k1 + i]; line[i].im = 0.0; }
    fft_line(k1, line, -1, tw, line + k1, line + 2 * k1);
    for (int i = 0; i < h; ++i) F[(size_t)l * h + i] = line[i];
  }
  free(line);
  free(tw);
  fft_axis(F, h, k2, k3, 1, -1, nthreads);
  fft_axis(F, h, k2, k3, 2, -1, nthreads);
}

void orc_fft_c2r(int k1, int k2, int k3, const double *FQ, double *Q, int nthreads) {
  int h = k1 / 2 + 1;
  size_t nc = (size_t)h * k2 * k3;
  cpx *F = (cpx *)malloc(sizeof(cpx) * nc);
  memcpy(F, FQ, sizeof(cpx) * nc);
  fft_axis(F, h, k2, k3, 2, +1, nthreads);
  fft_axis(F, h, k2, k3, 1, +1, nthreads);
  cpx *tw = make_twiddles(k1);
  cpx *line = (cpx *)malloc(sizeof(cpx) * (size_t)k1 * 3);
  for (long l = 0; l < (long)k2 * k3; ++l) {
    for (int i = 0; i < h; ++i) line[i] = F[(size_t)l * h + i];
    line[0].im = 0.0;
    if (k1 % 2 == 0) line[k1 / 2].im = 0.0;
    for (int i = h; i < k1; ++i) { line[i].re = line[k1 - i].re; line[i].im = -line[k1 - i].im; }
    fft_line(k1, line, +1, tw, line + k1, line + 2 * k1);
    for (int i = 0; i < k1; ++i) Q[(size_t)l * k1 + i] = line[i].re;
  }
  free(line);
  free(tw);
  free(F);
}

/* ------------------------------------------------------------------ */
/* scalar_sum.f:1-135 (FFTW variant)                                   */
/* ------------------------------------------------------------------ */
void orc_scalar_sum(double *FQ, double ewaldcof, double volume, const double *recip, const double *bsp1,
                    const double *bsp2, const double *bsp3, int nfft1, int nfft2, int nfft3, double rkcut,
                    double *eer, double *vir, int nthreads) {
  const double pi = PI_F32;
  const int ndim1 = 1 + nfft1 / 2;
  const long indtop = (long)ndim1 * nfft2 * nfft3;
  const double rkcut2 = 0.25 * rkcut * rkcut / (pi * pi);
  const double fac = pi * pi / (ewaldcof * ewaldcof);
  const long nff = (long)ndim1 * nfft2;
  int nf1 = nfft1 / 2; if (2 * nf1 < nfft1) nf1 = nf1 + 1;
  int nf2 = nfft2 / 2; if (2 * nf2 < nfft2) nf2 = nf2 + 1;
  int nf3 = nfft3 / 2; if (2 * nf3 < nfft3) nf3 = nf3 + 1;
  double energy = 0.0, v11 = 0, v22 = 0, v33 = 0, v12 = 0, v13 = 0, v23 = 0;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) \
    reduction(+ : energy, v11, v22, v33, v12, v13, v23) if (nthreads > 1)
  for (long ind = 1; ind <= indtop - 1; ++ind) {
    int k3 = (int)(ind / nff) + 1;
    long jnd = ind - (long)(k3 - 1) * nff;
    int k2 = (int)(jnd / ndim1) + 1;
    int k1 = (int)(jnd - (long)(k2 - 1) * ndim1) + 1;
    int m1 = k1 - 1; if (k1 > nf1) m1 = k1 - 1 - nfft1;
    int m2 = k2 - 1; if (k2 > nf2) m2 = k2 - 1 - nfft2;
    int m3 = k3 - 1; if (k3 > nf3) m3 = k3 - 1 - nfft3;
    double mhat1 = recip[0] * m1 + recip[3] * m2 + recip[6] * m3;
    double mhat2 = recip[1] * m1 + recip[4] * m2 + recip[7] * m3;
    double mhat3 = recip[2] * m1 + recip[5] * m2 + recip[8] * m3;
    double msq = mhat1 * mhat1 + mhat2 * mhat2 + mhat3 * mhat3;
    double eterm;
    double *qc = FQ + 2 * ind;
    if (msq > rkcut2) {
      eterm = 0.0;
    } else {
      double denom = pi * volume * bsp1[k1 - 1] * bsp2[k2 - 1] * bsp3[k3 - 1] * msq;
      eterm = exp(-fac * msq) / denom;
      double vterm = 2.0 * (fac * msq + 1.0) / msq;
      double qr = qc[0], qi = qc[1];
      double struc2, vvir, fac12, fac13, fac23;
      if (k1 == 1 || (k1 == ndim1 && nfft1 % 2 == 0)) {
        struc2 = (eterm * qr) * qr - (eterm * qi) * (-qi);
        vvir = struc2 * vterm;
        fac12 = 1.0; fac13 = 1.0; fac23 = 1.0;
      } else {
        double e2 = 2.0 * eterm;
        struc2 = (e2 * qr) * qr - (e2 * qi) * (-qi);
        vvir = struc2 * vterm;
        fac12 = (mhat1 * mhat2 < 0.0 && k2 <= ndim1) ? 0.0 : 1.0;
        fac13 = (mhat1 * mhat3 < 0.0 && k3 <= ndim1) ? 0.0 : 1.0;
        int neg = mhat2 * mhat3 < 0.0;
        if (k2 <= ndim1 && k3 <= ndim1) fac23 = neg ? 0.0 : 1.0;
        else if (k2 > ndim1 && k3 <= ndim1) fac23 = neg ? 1.0 : 0.0;
        else if (k3 > ndim1 && k2 <= ndim1) fac23 = neg ? 1.0 : 0.0;
        else fac23 = neg ? 0.0 : 1.0;
      }
      energy = energy + struc2;
      v11 = v11 + struc2 * (vterm * mhat1 * mhat1 - 1.0);
      v22 = v22 + struc2 * (vterm * mhat2 * mhat2 - 1.0);
      v33 = v33 + struc2 * (vterm * mhat3 * mhat3 - 1.0);
      v12 = v12 + vvir * mhat1 * mhat2 * fac12;
      v13 = v13 + vvir * mhat1 * mhat3 * fac13;
      v23 = v23 + vvir * mhat2 * mhat3 * fac23;
    }
    qc[0] = eterm * qc[0];
    qc[1] = eterm * qc[1];
  }
  *eer = 0.5 * energy;
  vir[0] = 0.5 * v11; vir[4] = 0.5 * v22; vir[8] = 0.5 * v33;
  vir[3] = 0.5 * v12; vir[6] = 0.5 * v13; vir[7] = 0.5 * v23;
  vir[1] = vir[3]; vir[2] = vir[6]; vir[5] = vir[7];
}

/* scalar_sum.f:137-217 (fftpack variant): every point of the full complex
 * grid Q(2,K1,K2,K3), weight 1, no virial masks. */
void orc_scalar_sum_full(double *Qc, double ewaldcof, double volume, const double *recip, const double *bsp1,
                         const double *bsp2, const double *bsp3, int nfft1, int nfft2, int nfft3, double rkcut,
                         double *eer, double *vir) {
  const double pi = PI_F32;
  const long indtop = (long)nfft1 * nfft2 * nfft3;
  const double rkcut2 = 0.25 * rkcut * rkcut / (pi * pi);
  const double fac = pi * pi / (ewaldcof * ewaldcof);
  const long nff = (long)nfft1 * nfft2;
  int nf1 = nfft1 / 2; if (2 * nf1 < nfft1) nf1 = nf1 + 1;
  int nf2 = nfft2 / 2; if (2 * nf2 < nfft2) nf2 = nf2 + 1;
  int nf3 = nfft3 / 2; if (2 * nf3 < nfft3) nf3 = nf3 + 1;
  double energy = 0.0, v[9] = {0};
  for (long ind = 1; ind <= indtop - 1; ++ind) {
    int k3 = (int)(ind / nff) + 1;
    long jnd = ind - (long)(k3 - 1) * nff;
    int k2 = (int)(jnd / nfft1) + 1;
    int k1 = (int)(jnd - (long)(k2 - 1) * nfft1) + 1;
    int m1 = k1 - 1; if (k1 > nf1) m1 = k1 - 1 - nfft1;
    int m2 = k2 - 1; if (k2 > nf2) m2 = k2 - 1 - nfft2;
    int m3 = k3 - 1; if (k3 > nf3) m3 = k3 - 1 - nfft3;
    double mhat1 = recip[0] * m1 + recip[3] * m2 + recip[6] * m3;
    double mhat2 = recip[1] * m1 + recip[4] * m2 + recip[7] * m3;
    double mhat3 = recip[2] * m1 + recip[5] * m2 + recip[8] * m3;
    double msq = mhat1 * mhat1 + mhat2 * mhat2 + mhat3 * mhat3;
    double eterm;
    double *qc = Qc + 2 * ind;
    if (msq > rkcut2) {
      eterm = 0.0;
    } else {
      double denom = pi * volume * bsp1[k1 - 1] * bsp2[k2 - 1] * bsp3[k3 - 1] * msq;
      eterm = exp(-fac * msq) / denom;
      double vterm = 2.0 * (fac * msq + 1.0) / msq;
      double struc2 = qc[0] * qc[0] + qc[1] * qc[1];
      energy = energy + eterm * struc2;
      v[0] = v[0] + eterm * struc2 * (vterm * mhat1 * mhat1 - 1.0);
      v[3] = v[3] + eterm * struc2 * (vterm * mhat1 * mhat2);
      v[6] = v[6] + eterm * struc2 * (vterm * mhat1 * mhat3);
      v[4] = v[4] + eterm * struc2 * (vterm * mhat2 * mhat2 - 1.0);
      v[7] = v[7] + eterm * struc2 * (vterm * mhat2 * mhat3);
      v[8] = v[8] + eterm * struc2 * (vterm * mhat3 * mhat3 - 1.0);
    }
    qc[0] = eterm * qc[0];
    qc[1] = eterm * qc[1];
  }
  *eer = 0.5 * energy;
  v[1] = v[3]; v[2] = v[6]; v[5] = v[7];
  for (int k = 0; k < 9; ++k) vir[k] = 0.5 * v[k];
}

/* pme_fftw3.f90:311-372 */
void orc_grad_sum(int n, const double *q, const double *recip, const double *th1, const double *th2,
                  const double *th3, const double *dth1, const double *dth2, const double *dth3,
                  double *fx, double *fy, double *fz, const double *fr1, const double *fr2, const double *fr3,
                  int order, int k1, int k2, int k3, const double *Q, int nthreads) {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
  for (int a = 0; a < n; ++a) {
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    const double *t1 = th1 + (size_t)a * order, *t2 = th2 + (size_t)a * order, *t3 = th3 + (size_t)a * order;
    const double *d1 = dth1 + (size_t)a * order, *d2 = dth2 + (size_t)a * order, *d3 = dth3 + (size_t)a * order;
    int kk0 = (int)fr3[a] - order;
    for (int i3 = 0; i3 < order; ++i3) {
      kk0 += 1;
      int kk = wrapidx(kk0, k3);
      int jj0 = (int)fr2[a] - order;
      for (int i2 = 0; i2 < order; ++i2) {
        jj0 += 1;
        int jj = wrapidx(jj0, k2);
        int ii0 = (int)fr1[a] - order;
        for (int i1 = 0; i1 < order; ++i1) {
          ii0 += 1;
          int ii = wrapidx(ii0, k1);
          double term = q[a] * Q[(size_t)ii + (size_t)k1 * (jj + (size_t)k2 * kk)];
          f1 = f1 - k1 * term * d1[i1] * t2[i2] * t3[i3];
          f2 = f2 - k2 * term * t1[i1] * d2[i2] * t3[i3];
          f3 = f3 - k3 * term * t1[i1] * t2[i2] * d3[i3];
        }
      }
    }
    fx[a] = fx[a] + recip[0] * f1 + recip[3] * f2 + recip[6] * f3;
    fy[a] = fy[a] + recip[1] * f1 + recip[4] * f2 + recip[7] * f3;
    fz[a] = fz[a] + recip[2] * f1 + recip[5] * f2 + recip[8] * f3;
  }
}

static double g_pme_times[5];
void orc_last_pme_times(double *t5) { memcpy(t5, g_pme_times, sizeof g_pme_times); }

/* fft_pme.f:1-115 around do_pmesh_kspace (pme_fftw3.f90:56-130, FFTW branch) */
int orc_fft_pme(int ntap, const double *xp, const double *yp, const double *zp, const double *xpcm,
                const double *ypcm, const double *zpcm, const double *pmechg, const double *co,
                const double *oc, double volume, double alphal, int order, int nfft1, int nfft2, int nfft3,
                double *eer, double *fpx, double *fpy, double *fpz, double *stressc, const int32_t *atomp,
                int32_t pressure, double rkcut, const double *bsp1, const double *bsp2, const double *bsp3,
                int nthreads, int variant) {
  size_t n = (size_t)ntap;
  double *x = malloc(n * 8), *y = malloc(n * 8), *z = malloc(n * 8);
  double *fx = calloc(n, 8), *fy = calloc(n, 8), *fz = calloc(n, 8);
  double *fr1 = malloc(n * 8), *fr2 = malloc(n * 8), *fr3 = malloc(n * 8);
  double *th = malloc(n * 8 * order * 6);
  double *th1 = th, *th2 = th + n * order, *th3 = th + 2 * n * order;
  double *dth1 = th + 3 * n * order, *dth2 = th + 4 * n * order, *dth3 = th + 5 * n * order;
  size_t G = (size_t)nfft1 * nfft2 * nfft3, Gc = (size_t)(nfft1 / 2 + 1) * nfft2 * nfft3;
  double *Q = malloc(G * 8), *FQ = malloc(Gc * 16);
  double recip[9], vir[9];
  for (size_t i = 0; i < n; ++i) {
    x[i] = co[0] * xp[i] + co[3] * yp[i] + co[6] * zp[i];
    y[i] = co[1] * xp[i] + co[4] * yp[i] + co[7] * zp[i];
    z[i] = co[2] * xp[i] + co[5] * yp[i] + co[8] * zp[i];
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) recip[i + 3 * j] = 0.5 * oc[j + 3 * i];
  double t0 = now_s();
  orc_scaled_fractionals(ntap, x, y, z, recip, nfft1, nfft2, nfft3, fr1, fr2, fr3);
  orc_bspline_coeffs(ntap, fr1, fr2, fr3, order, th1, th2, th3, dth1, dth2, dth3);
  orc_fill_charge_grid(ntap, pmechg, th1, th2, th3, fr1, fr2, fr3, order, nfft1, nfft2, nfft3, Q, nthreads);
  double t1 = now_s(), t2, t3, t4;
  if (variant == 0) {
    orc_fft_r2c(nfft1, nfft2, nfft3, Q, FQ, nthreads);
    t2 = now_s();
    orc_scalar_sum(FQ, alphal, volume, recip, bsp1, bsp2, bsp3, nfft1, nfft2, nfft3, rkcut, eer, vir, nthreads);
    t3 = now_s();
    orc_fft_c2r(nfft1, nfft2, nfft3, FQ, Q, nthreads);
    t4 = now_s();
  } else {
    /* fftpack branch (pme_fftw3.f90:119-127): full complex grid, fft_back
     * (forward e^{-i}), scalar_sum.f:137-217, fft_forward (e^{+i}) */
    cpx *Qc = (cpx *)malloc(sizeof(cpx) * G);
    for (size_t g = 0; g < G; ++g) { Qc[g].re = Q[g]; Qc[g].im = 0.0; }
    for (int ax = 0; ax < 3; ++ax) fft_axis(Qc, nfft1, nfft2, nfft3, ax, -1, nthreads);
    t2 = now_s();
    orc_scalar_sum_full((double *)Qc, alphal, volume, recip, bsp1, bsp2, bsp3, nfft1, nfft2, nfft3, rkcut, eer,
                        vir);
    t3 = now_s();
    for (int ax = 0; ax < 3; ++ax) fft_axis(Qc, nfft1, nfft2, nfft3, ax, +1, nthreads);
    for (size_t g = 0; g < G; ++g) Q[g] = Qc[g].re;
    free(Qc);
    t4 = now_s();
  }
  orc_grad_sum(ntap, pmechg, recip, th1, th2, th3, dth1, dth2, dth3, fx, fy, fz, fr1, fr2, fr3, order, nfft1,
               nfft2, nfft3, Q, nthreads);
  double t5 = now_s();
  g_pme_times[0] = t1 - t0; g_pme_times[1] = t2 - t1; g_pme_times[2] = t3 - t2;
  g_pme_times[3] = t4 - t3; g_pme_times[4] = t5 - t4;
  if (pressure) {
    double s[9] = {0};
    for (size_t i = 0; i < n; ++i) {
      int j = atomp[i] - 1;
      double sx = xp[i] - xpcm[j], sy = yp[i] - ypcm[j], sz = zp[i] - zpcm[j];
      double gx = sx * co[0] + sy * co[3] + sz * co[6];
      double gy = sx * co[1] + sy * co[4] + sz * co[7];
      double gz = sx * co[2] + sy * co[5] + sz * co[8];
      s[0] = s[0] - gx * fx[i]; s[1] = s[1] - gy * fx[i]; s[2] = s[2] - gz * fx[i];
      s[3] = s[3] - gx * fy[i]; s[4] = s[4] - gy * fy[i]; s[5] = s[5] - gz * fy[i];
      s[6] = s[6] - gx * fz[i]; s[7] = s[7] - gy * fz[i]; s[8] = s[8] - gz * fz[i];
    }
    /* stressc(r,c) with stc numbered row-major: stc1 = (1,1), stc2 = (1,2), ... */
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        stressc[r + 3 * c] = stressc[r + 3 * c] - vir[r + 3 * c] + s[3 * r + c];
  }
  for (size_t i = 0; i < n; ++i) {
    fpx[i] = fpx[i] + fx[i];
    fpy[i] = fpy[i] + fy[i];
    fpz[i] = fpz[i] + fz[i];
  }
  free(x); free(y); free(z); free(fx); free(fy); free(fz); free(fr1); free(fr2); free(fr3);
  free(th); free(Q); free(FQ);
  return 0;
}

/* ------------------------------------------------------------------ */
/* linked cells: dist_ijk.f, lc_index.f, lc_list.f                     */
/* ------------------------------------------------------------------ */
double orc_dist_ijk(int ni, int nj, int nk, double dx, double dy, double dz, const double *co) {
  static const int nv[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1},
                               {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};
  const int ndtmax = 500;
  double dmin = 1.0e8, d;
  int imin = 0, jmin = 0;
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j) {
      double lx = (ni + nv[j][0] - nv[i][0]) * dx;
      double ly = (nj + nv[j][1] - nv[i][1]) * dy;
      double lz = (nk + nv[j][2] - nv[i][2]) * dz;
      double mx = co[0] * lx + co[3] * ly + co[6] * lz;
      double my = co[1] * lx + co[4] * ly + co[7] * lz;
      double mz = co[2] * lx + co[5] * ly + co[8] * lz;
      d = mx * mx + my * my + mz * mz;
      if (d < dmin) { dmin = d; imin = i; jmin = j; }
    }
  double lx = (ni + nv[jmin][0] - nv[imin][0]) * dx;
  double ly = (nj + nv[jmin][1] - nv[imin][1]) * dy;
  double lz = (nk + nv[jmin][2] - nv[imin][2]) * dz;
  double mx = co[0] * lx + co[3] * ly + co[6] * lz;
  double my = co[1] * lx + co[4] * ly + co[7] * lz;
  double mz = co[2] * lx + co[5] * ly + co[8] * lz;
  double dcell[3] = {dx, dy, dz};
  dmin = 1.0e8;
  for (int ax = 0; ax < 3; ++ax) {
    /* sign(1., 0.5-nv) is a default-real +-1 */
    double sgn = (0.5f - (float)nv[jmin][ax]) >= 0.0f ? 1.0 : -1.0;
    double dt = sgn * dcell[ax] / ndtmax;
    for (int i = 1; i <= ndtmax; ++i) {
      double dmx = i * co[0 + 3 * ax] * dt;
      double dmy = i * co[1 + 3 * ax] * dt;
      double dmz = i * co[2 + 3 * ax] * dt;
      d = (mx + dmx) * (mx + dmx) + (my + dmy) * (my + dmy) + (mz + dmz) * (mz + dmz);
      if (d < dmin) dmin = d;
    }
  }
  return dmin;
}

/* lc_index.f:3-125; indx arrays are (2, indmax) column-major */
int orc_lc_index(int indmax, int ncx, int ncy, int ncz, int32_t *nind, int32_t *indxi, int32_t *indxj,
                 int32_t *indxk, double ctoff, const double *co) {
  double sqcut = ctoff * ctoff;
  double dx = 2.0 / ncx, dy = 2.0 / ncy, dz = 2.0 / ncz;
  int imax = 0, jmax = 0, kmax = 0;
  nind[0] = 1; indxi[0] = 0; indxj[0] = 0; indxk[0] = 0;
  nind[1] = 1; indxi[1] = 0; indxj[1] = 0; indxk[1] = 0;
  for (int n = 1; n <= 2; ++n) {
    int istart = (n == 1) ? 0 : 1 - ncx;
    for (int i = istart; i <= ncx - 1; ++i) {
      int jstart = (n == 1 && i == 0) ? 0 : 1 - ncy;
      for (int j = jstart; j <= ncy - 1; ++j) {
        int kstart = (n == 1 && i == 0 && j == 0) ? 0 : 1 - ncz;
        for (int k = kstart; k <= ncz - 1; ++k) {
          double rmin = orc_dist_ijk(i, j, k, dx, dy, dz, co);
          if (rmin < sqcut) {
            nind[n - 1] += 1;
            if (nind[n - 1] > indmax) return 1;
            int m = nind[n - 1];
            indxi[(n - 1) + 2 * (m - 1)] = i;
            indxj[(n - 1) + 2 * (m - 1)] = j;
            indxk[(n - 1) + 2 * (m - 1)] = k;
            if (imax < abs(i)) imax = abs(i);
            if (jmax < abs(j)) jmax = abs(j);
            if (kmax < abs(k)) kmax = abs(k);
            if (i == 0 && j == 0 && k == 0) nind[n - 1] -= 1;
          }
        }
      }
    }
  }
  if (imax >= (ncx + 1) / 2 || jmax >= (ncy + 1) / 2 || kmax >= (ncz + 1) / 2) return 1;
  return 0;
}

static inline int fmodp(int a, int n) { return ((a % n) + n) % n; }

/* cell index along one axis, lc_list.f:94-103 */
static inline int cell_of(double xr, double d, int nc) {
  double x1 = xr / d;
  int ix = (int)x1;
  double sg = (x1 - ix) >= 0.0 ? 1.0 : -1.0;
  int nx = (int)((double)ix + (sg - 1.0) / 2.0);
  return fmodp(nx, nc);
}

long orc_lc_list(int ncx, int ncy, int ncz, const int32_t *nind, const int32_t *indxi, const int32_t *indxj,
                 const int32_t *indxk, double rcut, const double *co, const double *xp, const double *yp,
                 const double *zp, int natp, int32_t *nnlpp, long capacity) {
  double sqcut = rcut * rcut;
  double dx = 2.0 / ncx, dy = 2.0 / ncy, dz = 2.0 / ncz;
  int ncell = ncx * ncy * ncz;
  int *headp = calloc((size_t)ncell, sizeof(int));
  int *chainp = malloc(sizeof(int) * ((size_t)natp + 1));
  int *ci = malloc(sizeof(int) * (size_t)natp * 3);
  for (int n = 1; n <= natp; ++n) {
    int nx = cell_of(xp[n - 1], dx, ncx), ny = cell_of(yp[n - 1], dy, ncy), nz = cell_of(zp[n - 1], dz, ncz);
    ci[3 * (n - 1)] = nx; ci[3 * (n - 1) + 1] = ny; ci[3 * (n - 1) + 2] = nz;
    int numcell = nz + ncz * (ny + ncy * nx);
    chainp[n] = headp[numcell];
    headp[numcell] = n;
  }
  long nvtot = 1; /* 1-based position of the next count slot */
  for (int n = 1; n <= natp; ++n) {
    double x1 = xp[n - 1], y1 = yp[n - 1], z1 = zp[n - 1];
    int nv = 0;
    int i = ci[3 * (n - 1)], j = ci[3 * (n - 1) + 1], k = ci[3 * (n - 1) + 2];
    for (int m = 1; m <= nind[1]; ++m) {
      int iv = indxi[1 + 2 * (m - 1)], jv = indxj[1 + 2 * (m - 1)], kv = indxk[1 + 2 * (m - 1)];
      int nx = fmodp(i + iv, ncx), ny = fmodp(j + jv, ncy), nz = fmodp(k + kv, ncz);
      int l = headp[nz + ncz * (ny + ncy * nx)];
      int nmin = (m == 1) ? n : 0;
      while (l > nmin) {
        if (l > n) {
          double x2 = x1 - xp[l - 1], y2 = y1 - yp[l - 1], z2 = z1 - zp[l - 1];
          x2 = x2 - 2.0 * fnint(0.5 * x2);
          y2 = y2 - 2.0 * fnint(0.5 * y2);
          z2 = z2 - 2.0 * fnint(0.5 * z2);
          double xx = co[0] * x2 + co[3] * y2 + co[6] * z2;
          double yy = co[1] * x2 + co[4] * y2 + co[7] * z2;
          double zz = co[2] * x2 + co[5] * y2 + co[8] * z2;
          double d = xx * xx + yy * yy + zz * zz;
          if (d < sqcut) {
            nv += 1;
            if (nvtot + nv > capacity) { free(headp); free(chainp); free(ci); return -1; }
            nnlpp[nvtot + nv - 1] = l;
          }
        }
        l = chainp[l];
      }
    }
    nnlpp[nvtot - 1] = nv;
    nvtot = nvtot + nv + 1;
  }
  free(headp); free(chainp); free(ci);
  return nvtot - 1;
}

/* ------------------------------------------------------------------ */
/* force.f:1-917                                                       */
/* ------------------------------------------------------------------ */
typedef struct {
  double st[9];
  double uconf[3], ucoul[3];
  /* force_REM.f only: the segment energy, and the last values its locals
   * coul and rsp hold (both carry over from one pair to the next there) */
  double seg, coul_last, rsp_last;
} force_acc;

/* One owner group i (1-based), restating force.f:235-835 for that group.
 * `fi` receives the forces on atoms of group i, `fj` on the neighbours
 * (the same array in the serial build). Scratch arrays are caller-owned. */
typedef struct {
  int *index0, *index1;
  double *xmap0, *ymap0, *zmap0, *xmap1, *ymap1, *zmap1, *xmap2, *ymap2, *zmap2;
  double *xmap3, *ymap3, *zmap3, *cmap2, *swrs, *dswrs;
  char *maplg, *mapag;
} force_scratch;

static force_scratch *scratch_new(int nato, int ngrp, int cap) {
  force_scratch *s = calloc(1, sizeof *s);
  s->index0 = malloc(sizeof(int) * cap);
  s->index1 = malloc(sizeof(int) * cap);
  double **d[] = {&s->xmap0, &s->ymap0, &s->zmap0, &s->xmap1, &s->ymap1, &s->zmap1, &s->xmap2, &s->ymap2,
                  &s->zmap2, &s->xmap3, &s->ymap3, &s->zmap3, &s->cmap2, &s->swrs, &s->dswrs};
  for (size_t k = 0; k < sizeof d / sizeof d[0]; ++k) *d[k] = malloc(sizeof(double) * cap);
  s->maplg = malloc((size_t)nato + 1);
  s->mapag = malloc((size_t)ngrp + 1);
  memset(s->maplg, 1, (size_t)nato + 1);
  memset(s->mapag, 1, (size_t)ngrp + 1);
  return s;
}

static void scratch_free(force_scratch *s) {
  free(s->index0); free(s->index1);
  double *d[] = {s->xmap0, s->ymap0, s->zmap0, s->xmap1, s->ymap1, s->zmap1, s->xmap2, s->ymap2,
                 s->zmap2, s->xmap3, s->ymap3, s->zmap3, s->cmap2, s->swrs, s->dswrs};
  for (size_t k = 0; k < sizeof d / sizeof d[0]; ++k) free(d[k]);
  free(s->maplg); free(s->mapag); free(s);
}

typedef struct {
  const int32_t *ss_index; const double *xp0, *yp0, *zp0, *xpg, *ypg, *zpg, *charge; const int32_t *nbtype;
  const int32_t *type; int ma; const int32_t *atomg; const double *xpcm, *ypcm, *zpcm; const int32_t *atomp;
  const double *co, *ecc12, *ecc6; double alphal; const int32_t *grppt; const double *gmass, *mass;
  double massinfty; int lskip_ewald, ewald, iz;
  double r2inn, r2out, rinn0, r2inn0, rout0, r2out0, r2neigh, twrtpi, arsout1, arsout2, arsinn1, arsinn2;
  double rinn, rout;
  /* force_REM.f: rem = 1 selects the segment pairs (seg_select, rem.f90:631-645) */
  int rem, rem_kind; const int32_t *rematom; double rem_f3;
} force_par;

/* rem.f90:631-645: kind 0 either atom in the segment, 1 both, 2 exactly one */
static int seg_select(int kind, const int32_t *rematom, int a, int b) {
  const int ra = rematom[a - 1] != 0, rb = rematom[b - 1] != 0;
  if (kind == 0) return ra || rb;
  if (kind == 1) return ra && rb;
  if (kind == 2) return ra != rb;
  return 0;
}

/* pair body shared by the four pair loops; returns qforce, updates energies */
#define PAIR_GEOM_AND_LJ_EWALD(SWITCHED)                                                     \
  lij = P->type[(nbti - 1) + (size_t)P->ma * (P->nbtype[j - 1] - 1)];                        \
  if (P->lskip_ewald && (P->ecc6[lij - 1] < 1e-12) && P->ewald) continue;                     \
  p2 = P->atomp[j - 1];

static void force_group(const force_par *P, int i, const int32_t *nbr, int m, int mbeg, const int32_t *mapnl_i,
                        int32_t *nnl1_out, int *map_out, force_scratch *S, double *fpx, double *fpy,
                        double *fpz, double *fjx, double *fjy, double *fjz, force_acc *A) {
  const double *co = P->co;
  double xpgi = P->xpg[i - 1], ypgi = P->ypg[i - 1], zpgi = P->zpg[i - 1];
  int map = 0, mp0 = 0, mp1 = 0;
  int typei = P->ss_index[P->grppt[2 * (i - 1)] - 1];
  for (int jj = mbeg; jj <= m; ++jj) {
    int j = nbr[jj - 1];
    double xgg = xpgi - P->xpg[j - 1], ygg = ypgi - P->ypg[j - 1], zgg = zpgi - P->zpg[j - 1];
    double xmap0j = 2.0 * fnint(0.5 * xgg), ymap0j = 2.0 * fnint(0.5 * ygg), zmap0j = 2.0 * fnint(0.5 * zgg);
    xgg = xgg - xmap0j; ygg = ygg - ymap0j; zgg = zgg - zmap0j;
    double xc = co[0] * xgg + co[3] * ygg + co[6] * zgg;
    double yc = co[4] * ygg + co[7] * zgg;
    double zc = co[8] * zgg;
    double drj = xc * xc + yc * yc + zc * zc;
    if (P->r2inn > 1.0e-5 && drj < P->r2neigh) {
      map += 1;
      if (nnl1_out) nnl1_out[map - 1] = j;
    }
    if (P->rinn0 > P->rout || P->iz == 0) continue;
    if (drj >= P->r2inn0 && drj < P->r2out) {
      mp0 += 1;
      S->xmap0[mp0 - 1] = xmap0j; S->ymap0[mp0 - 1] = ymap0j; S->zmap0[mp0 - 1] = zmap0j;
      S->index0[mp0 - 1] = j;
    } else {
      if (drj >= P->r2out && drj < P->r2out0) {
        mp1 += 1;
        S->xmap1[mp1 - 1] = xmap0j; S->ymap1[mp1 - 1] = ymap0j; S->zmap1[mp1 - 1] = zmap0j;
        S->xmap2[mp1 - 1] = xc; S->ymap2[mp1 - 1] = yc; S->zmap2[mp1 - 1] = zc;
        S->index1[mp1 - 1] = j;
        double rsp = sqrt(drj);
        A->rsp_last = rsp;
        double auxa = (P->arsout1 + 2.0 * rsp) / P->arsout2;
        double auxb = P->rout0 - rsp;
        S->swrs[mp1 - 1] = auxa * (auxb * auxb);
        double dsw = -(2.0 * auxa * auxb) + 2.0 * (auxb * auxb) / P->arsout2;
        S->dswrs[mp1 - 1] = dsw / rsp;
      } else if (drj >= P->r2inn && drj < P->r2inn0) {
        mp1 += 1;
        S->xmap1[mp1 - 1] = xmap0j; S->ymap1[mp1 - 1] = ymap0j; S->zmap1[mp1 - 1] = zmap0j;
        S->xmap2[mp1 - 1] = xc; S->ymap2[mp1 - 1] = yc; S->zmap2[mp1 - 1] = zc;
        S->index1[mp1 - 1] = j;
        double rsp = sqrt(drj);
        A->rsp_last = rsp;
        double auxa = (P->arsinn1 + 2.0 * rsp) / P->arsinn2;
        double auxb = P->rinn0 - rsp;
        S->swrs[mp1 - 1] = 1.0 - auxa * (auxb * auxb);
        double dsw = 2.0 * auxa * auxb - 2.0 * (auxb * auxb) / P->arsinn2;
        S->dswrs[mp1 - 1] = dsw / rsp;
      }
    }
  }
  *map_out = map;
  int mapa = mp0, mapb = mp1;
  for (int jj = 0; jj < mapb; ++jj) S->cmap2[jj] = 0.0;
  if (mapa == 0 && mapb == 0) return;
  const int32_t *mp = mapnl_i;
  for (int i1 = P->grppt[2 * (i - 1)]; i1 <= P->grppt[2 * (i - 1) + 1]; ++i1) {
    double xpi = P->xp0[i1 - 1], ypi = P->yp0[i1 - 1], zpi = P->zp0[i1 - 1];
    int nbti = P->nbtype[i1 - 1];
    double chrgei = P->charge[i1 - 1];
    int la = mp[0];
    for (int t = 1; t <= la; ++t) {
      int mm = mp[t];
      S->maplg[mm] = 0;
      S->mapag[P->atomg[mm - 1]] = 0;
    }
    int p1 = P->atomp[i1 - 1];
    double xpcmp1 = P->xpcm[p1 - 1], ypcmp1 = P->ypcm[p1 - 1], zpcmp1 = P->zpcm[p1 - 1];
    /* normal (non switched) pairs, force.f:396-565 */
    for (int jj = 1; jj <= mapa; ++jj) {
      int j1 = S->index0[jj - 1];
      int typej = P->ss_index[P->grppt[2 * (j1 - 1)] - 1];
      int typeij = typei + typej - 1;
      double xd = -S->xmap0[jj - 1], yd = -S->ymap0[jj - 1], zd = -S->zmap0[jj - 1];
      double xd1 = xpi + xd, yd1 = ypi + yd, zd1 = zpi + zd;
      xd = xpcmp1 + xd; yd = ypcmp1 + yd; zd = zpcmp1 + zd;
      double uconfa = 0.0, ucoula = 0.0;
      int whole = S->mapag[j1];
      for (int j = P->grppt[2 * (j1 - 1)]; j <= P->grppt[2 * (j1 - 1) + 1]; ++j) {
        int lij, p2;
        if (!whole) {
          if (!(S->maplg[j] && (P->mass[j - 1] < P->massinfty || P->mass[i1 - 1] < P->massinfty))) continue;
        }
        PAIR_GEOM_AND_LJ_EWALD(0)
        double xxcm = xd - P->xpcm[p2 - 1], yycm = yd - P->ypcm[p2 - 1], zzcm = zd - P->zpcm[p2 - 1];
        double xg = xd1 - P->xp0[j - 1], yg = yd1 - P->yp0[j - 1], zg = zd1 - P->zp0[j - 1];
        double xc = co[0] * xg + co[3] * yg + co[6] * zg;
        double yc = co[4] * yg + co[7] * zg;
        double zc = co[8] * zg;
        double rsq = xc * xc + yc * yc + zc * zc;
        double rsqi = 1.0 / rsq;
        double qforce = 0.0, conf = 0.0;
        if (P->ecc6[lij - 1] > 1e-12) {
          double r6 = rsqi * rsqi * rsqi, r12 = r6 * r6;
          double ssvir = 12.0 * P->ecc12[lij - 1] * r12 - 6.0 * P->ecc6[lij - 1] * r6;
          qforce = ssvir * rsqi;
          conf = P->ecc12[lij - 1] * r12 - P->ecc6[lij - 1] * r6;
          uconfa = uconfa + conf;
        }
        if (!P->lskip_ewald) {
          double rsp = sqrt(rsq), rspqi = rsqi / rsp, alphar = P->alphal * rsp;
          double qt = 1.0 / (1.0 + QP * alphar);
          double expcst = exp(-(alphar * alphar));
          double erfcst = ((((A5 * qt + A4) * qt + A3) * qt + A2) * qt + A1) * qt * expcst;
          double furpar = chrgei * P->charge[j - 1];
          double coul = furpar * erfcst / rsp;
          ucoula = ucoula + coul;
          double aux1 = furpar * (erfcst + P->twrtpi * alphar * expcst) * rspqi;
          qforce = qforce + aux1;
          A->rsp_last = rsp;
          A->coul_last = coul;
        } else if (!P->ewald) {
          double rsp = sqrt(rsq);
          double ssvir = chrgei * P->charge[j - 1] / rsp;
          A->ucoul[typeij - 1] = A->ucoul[typeij - 1] + ssvir;
          qforce = qforce + ssvir * rsqi;
          A->rsp_last = rsp;
          A->coul_last = ssvir;
        }
        if (P->rem && seg_select(P->rem_kind, P->rematom, i1, j)) {
          /* force_REM.f:491-494: coul as left by this pair or an earlier one */
          A->seg = A->seg + conf + A->coul_last;
          qforce = qforce * P->rem_f3;
        }
        double emvir = qforce;
        fpx[i1 - 1] = fpx[i1 - 1] + qforce * xc;
        fpy[i1 - 1] = fpy[i1 - 1] + qforce * yc;
        fpz[i1 - 1] = fpz[i1 - 1] + qforce * zc;
        fjx[j - 1] = fjx[j - 1] - qforce * xc;
        fjy[j - 1] = fjy[j - 1] - qforce * yc;
        fjz[j - 1] = fjz[j - 1] - qforce * zc;
        double qfx = emvir * xc, qfy = emvir * yc, qfz = emvir * zc;
        A->st[0] += qfx * xxcm; A->st[1] += qfx * yycm; A->st[2] += qfx * zzcm;
        A->st[3] += qfy * xxcm; A->st[4] += qfy * yycm; A->st[5] += qfy * zzcm;
        A->st[6] += qfz * xxcm; A->st[7] += qfz * yycm; A->st[8] += qfz * zzcm;
      }
      A->ucoul[typeij - 1] = A->ucoul[typeij - 1] + ucoula;
      A->uconf[typeij - 1] = A->uconf[typeij - 1] + uconfa;
    }
    /* switched pairs, force.f:569-731 */
    for (int jj = 1; jj <= mapb; ++jj) {
      int j1 = S->index1[jj - 1];
      int typej = P->ss_index[P->grppt[2 * (j1 - 1)] - 1];
      int typeij = typei + typej - 1;
      int whole = S->mapag[j1];
      double sw = S->swrs[jj - 1];
      for (int j = P->grppt[2 * (j1 - 1)]; j <= P->grppt[2 * (j1 - 1) + 1]; ++j) {
        int lij, p2;
        if (!whole) {
          if (!(S->maplg[j] && (P->mass[j - 1] < P->massinfty || P->mass[i1 - 1] < P->massinfty))) continue;
        }
        PAIR_GEOM_AND_LJ_EWALD(1)
        double xxcm = xpcmp1 - P->xpcm[p2 - 1] - S->xmap1[jj - 1];
        double yycm = ypcmp1 - P->ypcm[p2 - 1] - S->ymap1[jj - 1];
        double zzcm = zpcmp1 - P->zpcm[p2 - 1] - S->zmap1[jj - 1];
        double xg = xpi - P->xp0[j - 1] - S->xmap1[jj - 1];
        double yg = ypi - P->yp0[j - 1] - S->ymap1[jj - 1];
        double zg = zpi - P->zp0[j - 1] - S->zmap1[jj - 1];
        double xc = co[0] * xg + co[3] * yg + co[6] * zg;
        double yc = co[4] * yg + co[7] * zg;
        double zc = co[8] * zg;
        double rsq = xc * xc + yc * yc + zc * zc;
        double rsqi = 1.0 / rsq;
        double qforce = 0.0, conf = 0.0;
        const int sel = P->rem && seg_select(P->rem_kind, P->rematom, i1, j);
        if (P->ecc6[lij - 1] > 1e-12) {
          double r6 = rsqi * rsqi * rsqi, r12 = r6 * r6;
          double ssvir = 12.0 * P->ecc12[lij - 1] * r12 - 6.0 * P->ecc6[lij - 1] * r6;
          qforce = ssvir * rsqi * sw;
          conf = P->ecc12[lij - 1] * r12 - P->ecc6[lij - 1] * r6;
          S->cmap2[jj - 1] = S->cmap2[jj - 1] + (sel ? conf * P->rem_f3 : conf);
          A->uconf[typeij - 1] = A->uconf[typeij - 1] + sw * conf;
        }
        if (!P->lskip_ewald) {
          double rsp = sqrt(rsq), rspqi = rsqi / rsp, alphar = P->alphal * rsp;
          double qt = 1.0 / (1.0 + QP * alphar);
          double expcst = exp(-(alphar * alphar));
          double erfcst = ((((A5 * qt + A4) * qt + A3) * qt + A2) * qt + A1) * qt * expcst;
          double furpar = chrgei * P->charge[j - 1];
          if (P->rem) {
            /* force_REM.f:649-664 */
            double coul = furpar * erfcst / rsp;
            A->ucoul[typeij - 1] = A->ucoul[typeij - 1] + coul * sw;
            S->cmap2[jj - 1] = S->cmap2[jj - 1] + (sel ? coul * P->rem_f3 : coul);
            A->coul_last = coul;
          } else {
            A->ucoul[typeij - 1] = A->ucoul[typeij - 1] + sw * furpar * erfcst / rsp;
            S->cmap2[jj - 1] = S->cmap2[jj - 1] + furpar * erfcst / rsp;
          }
          double aux1 = furpar * (erfcst + P->twrtpi * alphar * expcst) * rspqi * sw;
          qforce = qforce + aux1;
          A->rsp_last = rsp;
        } else if (!P->ewald) {
          if (P->rem) {
            /* force_REM.f:666-678: the distance is the routine's rsp as an
             * earlier pair or group left it (not recomputed for this pair) */
            double coul = chrgei * P->charge[j - 1] / A->rsp_last;
            A->ucoul[typeij - 1] = A->ucoul[typeij - 1] + coul * sw;
            qforce = qforce + coul * rsqi * sw;
            S->cmap2[jj - 1] = S->cmap2[jj - 1] + (sel ? coul * P->rem_f3 : coul);
            A->coul_last = coul;
          } else {
            double rsp = sqrt(rsq);
            double ssvir = chrgei * P->charge[j - 1] / rsp;
            S->cmap2[jj - 1] = S->cmap2[jj - 1] + ssvir;
            A->ucoul[typeij - 1] = A->ucoul[typeij - 1] + ssvir * sw;
            qforce = qforce + ssvir * rsqi * sw;
            A->rsp_last = rsp;
          }
        }
        if (sel) {
          /* force_REM.f:679-682 */
          A->seg = A->seg + sw * conf + sw * A->coul_last;
          qforce = qforce * P->rem_f3;
        }
        double emvir = qforce;
        fpx[i1 - 1] = fpx[i1 - 1] + qforce * xc;
        fpy[i1 - 1] = fpy[i1 - 1] + qforce * yc;
        fpz[i1 - 1] = fpz[i1 - 1] + qforce * zc;
        fjx[j - 1] = fjx[j - 1] - qforce * xc;
        fjy[j - 1] = fjy[j - 1] - qforce * yc;
        fjz[j - 1] = fjz[j - 1] - qforce * zc;
        double qfx = emvir * xc, qfy = emvir * yc, qfz = emvir * zc;
        A->st[0] += qfx * xxcm; A->st[1] += qfx * yycm; A->st[2] += qfx * zzcm;
        A->st[3] += qfy * xxcm; A->st[4] += qfy * yycm; A->st[5] += qfy * zzcm;
        A->st[6] += qfz * xxcm; A->st[7] += qfz * yycm; A->st[8] += qfz * zzcm;
      }
    }
    for (int t = 1; t <= la; ++t) {
      int mm = mp[t];
      S->maplg[mm] = 1;
      S->mapag[P->atomg[mm - 1]] = 1;
    }
    mp += la + 1;
  }
  /* S'(r) V redistribution by gmass, force.f:756-817 */
  for (int jj = 0; jj < mapb; ++jj) {
    S->xmap3[jj] = -(S->dswrs[jj] * S->cmap2[jj] * S->xmap2[jj]);
    S->ymap3[jj] = -(S->dswrs[jj] * S->cmap2[jj] * S->ymap2[jj]);
    S->zmap3[jj] = -(S->dswrs[jj] * S->cmap2[jj] * S->zmap2[jj]);
  }
  if (mapb != 0) {
    for (int i1 = P->grppt[2 * (i - 1)]; i1 <= P->grppt[2 * (i - 1) + 1]; ++i1) {
      double massi = P->gmass[i1 - 1];
      int p1 = P->atomp[i1 - 1];
      double xpcmp1 = P->xpcm[p1 - 1], ypcmp1 = P->ypcm[p1 - 1], zpcmp1 = P->zpcm[p1 - 1];
      for (int jj = 0; jj < mapb; ++jj) {
        fpx[i1 - 1] = fpx[i1 - 1] + massi * S->xmap3[jj];
        fpy[i1 - 1] = fpy[i1 - 1] + massi * S->ymap3[jj];
        fpz[i1 - 1] = fpz[i1 - 1] + massi * S->zmap3[jj];
        A->st[0] += massi * S->xmap3[jj] * xpcmp1; A->st[1] += massi * S->xmap3[jj] * ypcmp1;
        A->st[2] += massi * S->xmap3[jj] * zpcmp1; A->st[3] += massi * S->ymap3[jj] * xpcmp1;
        A->st[4] += massi * S->ymap3[jj] * ypcmp1; A->st[5] += massi * S->ymap3[jj] * zpcmp1;
        A->st[6] += massi * S->zmap3[jj] * xpcmp1; A->st[7] += massi * S->zmap3[jj] * ypcmp1;
        A->st[8] += massi * S->zmap3[jj] * zpcmp1;
      }
    }
    for (int jj = 0; jj < mapb; ++jj) {
      int j = S->index1[jj];
      double xxcm = S->xmap1[jj], yycm = S->ymap1[jj], zzcm = S->zmap1[jj];
      for (int j1 = P->grppt[2 * (j - 1)]; j1 <= P->grppt[2 * (j - 1) + 1]; ++j1) {
        double massj = P->gmass[j1 - 1];
        fjx[j1 - 1] = fjx[j1 - 1] - massj * S->xmap3[jj];
        fjy[j1 - 1] = fjy[j1 - 1] - massj * S->ymap3[jj];
        fjz[j1 - 1] = fjz[j1 - 1] - massj * S->zmap3[jj];
        int p2 = P->atomp[j1 - 1];
        double xq = -((P->xpcm[p2 - 1] + xxcm) * massj);
        double yq = -((P->ypcm[p2 - 1] + yycm) * massj);
        double zq = -((P->zpcm[p2 - 1] + zzcm) * massj);
        A->st[0] += S->xmap3[jj] * xq; A->st[1] += S->xmap3[jj] * yq; A->st[2] += S->xmap3[jj] * zq;
        A->st[3] += S->ymap3[jj] * xq; A->st[4] += S->ymap3[jj] * yq; A->st[5] += S->ymap3[jj] * zq;
        A->st[6] += S->zmap3[jj] * xq; A->st[7] += S->zmap3[jj] * yq; A->st[8] += S->zmap3[jj] * zq;
      }
    }
  }
}

static int force_impl(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
                      const double *xpg, const double *ypg, const double *zpg, const double *charge,
                      const int32_t *nbtype, const int32_t *type, int ma, int nato, const int32_t *atomg,
                      const double *xpcm, const double *ypcm, const double *zpcm, const int32_t *atomp,
                      const double *co, const double *ecc12, const double *ecc6, double alphal,
                      const int32_t *mapnl, int ngrp, const int32_t *grppt, double *uconf_slt, double *uconf_slv,
                      double *uconf_ss, double *ucoul_slt, double *ucoul_slv, double *ucoul_ss, double *fpx,
                      double *fpy, double *fpz, double *stress, const int32_t *nnlpp, int32_t *nnlpp1,
                      long nnlpp1_cap, double rneigh, double rinn, double rout, double *rtolinn, double rtolout,
                      const double *gmass, int iz, int32_t *ncount1, const double *mass, double massinfty,
                      int32_t lskip_ewald, int32_t ewald, int nthreads, int rem, int rem_kind,
                      const int32_t *rematom, double rem_f3, double *seg_ene);

int orc_force(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
              const double *xpg, const double *ypg, const double *zpg, const double *charge,
              const int32_t *nbtype, const int32_t *type, int ma, int nato, const int32_t *atomg,
              const double *xpcm, const double *ypcm, const double *zpcm, const int32_t *atomp,
              const double *co, const double *ecc12, const double *ecc6, double alphal, const int32_t *mapnl,
              int ngrp, const int32_t *grppt, double *uconf_slt, double *uconf_slv, double *uconf_ss,
              double *ucoul_slt, double *ucoul_slv, double *ucoul_ss, double *fpx, double *fpy, double *fpz,
              double *stress, const int32_t *nnlpp, int32_t *nnlpp1, long nnlpp1_cap, double rneigh,
              double rinn, double rout, double *rtolinn, double rtolout, const double *gmass, int iz,
              int32_t *ncount1, const double *mass, double massinfty, int32_t lskip_ewald, int32_t ewald,
              int nthreads) {
  return force_impl(ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, nato, atomg, xpcm, ypcm,
                    zpcm, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, uconf_slv, uconf_ss,
                    ucoul_slt, ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, nnlpp1_cap, rneigh, rinn,
                    rout, rtolinn, rtolout, gmass, iz, ncount1, mass, massinfty, lskip_ewald, ewald, nthreads, 0, 0,
                    NULL, 1.0, NULL);
}

/* force_REM.f:1-995 (serial): force.f with the pairs seg_select picks
 * (rem.f90:631-645) scaled by rem_f3 = rem_factor(3) in force, and their
 * energy (unscaled, switched) summed into seg_ene; lskip_force without
 * ljintra returns at once, with ljintra the charges count as zero
 * (force_REM.f:148-158).  The locals coul and rsp carry over from pair to
 * pair as in the reference (the segment energy of an lskip_ewald call and
 * the switched plain-Coulomb branch, force_REM.f:667, read them); they start
 * at zero here, where the reference's start values are undefined. */
int orc_force_rem(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
                  const double *xpg, const double *ypg, const double *zpg, const double *charge,
                  const int32_t *nbtype, const int32_t *type, int ma, int nato, const int32_t *atomg,
                  const double *xpcm, const double *ypcm, const double *zpcm, const int32_t *atomp,
                  const double *co, const double *ecc12, const double *ecc6, double alphal, const int32_t *mapnl,
                  int ngrp, const int32_t *grppt, double *uconf_slt, double *uconf_slv, double *uconf_ss,
                  double *ucoul_slt, double *ucoul_slv, double *ucoul_ss, double *fpx, double *fpy, double *fpz,
                  double *stress, const int32_t *nnlpp, int32_t *nnlpp1, long nnlpp1_cap, double rneigh,
                  double rinn, double rout, double *rtolinn, double rtolout, const double *gmass, int iz,
                  int32_t *ncount1, const double *mass, double massinfty, double *seg_ene,
                  const int32_t *rematom, int rem_kind, double rem_f3, int32_t lskip_force, int32_t ljintra,
                  int32_t lskip_ewald, int32_t ewald) {
  if (lskip_force && !ljintra) return 0;
  double *q0 = NULL;
  if (lskip_force) {
    q0 = calloc((size_t)nato, sizeof(double));
    charge = q0;
  }
  *seg_ene = 0.0;
  int st = force_impl(ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, nato, atomg, xpcm, ypcm,
                      zpcm, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, uconf_slv, uconf_ss,
                      ucoul_slt, ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, nnlpp1_cap, rneigh, rinn,
                      rout, rtolinn, rtolout, gmass, iz, ncount1, mass, massinfty, lskip_ewald, ewald, 1, 1,
                      rem_kind, rematom, rem_f3, seg_ene);
  free(q0);
  return st;
}

static int force_impl(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
                      const double *xpg, const double *ypg, const double *zpg, const double *charge,
                      const int32_t *nbtype, const int32_t *type, int ma, int nato, const int32_t *atomg,
                      const double *xpcm, const double *ypcm, const double *zpcm, const int32_t *atomp,
                      const double *co, const double *ecc12, const double *ecc6, double alphal,
                      const int32_t *mapnl, int ngrp, const int32_t *grppt, double *uconf_slt, double *uconf_slv,
                      double *uconf_ss, double *ucoul_slt, double *ucoul_slv, double *ucoul_ss, double *fpx,
                      double *fpy, double *fpz, double *stress, const int32_t *nnlpp, int32_t *nnlpp1,
                      long nnlpp1_cap, double rneigh, double rinn, double rout, double *rtolinn, double rtolout,
                      const double *gmass, int iz, int32_t *ncount1, const double *mass, double massinfty,
                      int32_t lskip_ewald, int32_t ewald, int nthreads, int rem, int rem_kind,
                      const int32_t *rematom, double rem_f3, double *seg_ene) {
  force_par P;
  P.rem = rem;
  P.rem_kind = rem_kind;
  P.rematom = rematom;
  P.rem_f3 = rem_f3;
  P.ss_index = ss_index; P.xp0 = xp0; P.yp0 = yp0; P.zp0 = zp0; P.xpg = xpg; P.ypg = ypg; P.zpg = zpg;
  P.charge = charge; P.nbtype = nbtype; P.type = type; P.ma = ma; P.atomg = atomg; P.xpcm = xpcm;
  P.ypcm = ypcm; P.zpcm = zpcm; P.atomp = atomp; P.co = co; P.ecc12 = ecc12; P.ecc6 = ecc6;
  P.alphal = alphal; P.grppt = grppt; P.gmass = gmass; P.mass = mass; P.massinfty = massinfty;
  P.lskip_ewald = lskip_ewald; P.ewald = ewald; P.iz = iz; P.rinn = rinn; P.rout = rout;
  /* force.f:145-166 */
  P.r2inn = rinn * rinn;
  P.r2out = rout * rout;
  P.rinn0 = rinn + *rtolinn;
  P.r2inn0 = P.rinn0 * P.rinn0;
  if (rinn < 1.0e-5) *rtolinn = 0.0;
  P.rout0 = rout + rtolout;
  P.r2out0 = P.rout0 * P.rout0;
  P.r2neigh = (P.rinn0 + rneigh) * (P.rinn0 + rneigh);
  P.twrtpi = 2.0 / sqrt(4.0 * atan(1.0));
  P.arsout1 = P.rout0 - 3.0 * rout;
  P.arsout2 = rtolout * rtolout * rtolout;
  P.arsinn1 = P.rinn0 - 3.0 * rinn;
  P.arsinn2 = *rtolinn * *rtolinn * *rtolinn;

  /* segment offsets of nnlpp and mapnl per group (serial scan, as the reference walks them) */
  long *noff = malloc(sizeof(long) * ((size_t)ngrp + 1));
  long *moff = malloc(sizeof(long) * ((size_t)ngrp + 1));
  long n = 0, na = 0;
  int maxm = 0;
  for (int i = 1; i <= ngrp; ++i) {
    noff[i - 1] = n;
    moff[i - 1] = na;
    int m = nnlpp[n];
    if (m > maxm) maxm = m;
    for (int i1 = grppt[2 * (i - 1)]; i1 <= grppt[2 * (i - 1) + 1]; ++i1) na += mapnl[na] + 1;
    if (iz == 1) n += m + 1;
  }
  if (iz == 0) maxm = nnlpp[0];
  int cap = maxm + 1;
  int nt = (nthreads > 1 && !rem) ? nthreads : 1;  /* the segment pass follows the serial order */
  int status = 0;
  if (nt == 1) {
    force_scratch *S = scratch_new(nato, ngrp, cap);
    force_acc A;
    memset(&A, 0, sizeof A);
    long c1 = 0;
    int mbeg = 1;
    for (int i = 1; i <= ngrp; ++i) {
      const int32_t *seg = nnlpp + (iz == 1 ? noff[i - 1] : 0);
      int m = seg[0];
      if (iz == 0) mbeg = i + 1;
      int map = 0;
      int32_t *out1 = NULL;
      if (P.r2inn > 1.0e-5) {
        if (c1 + 1 + m > nnlpp1_cap) { status = 1; break; }
        out1 = nnlpp1 + c1 + 1;
      }
      force_group(&P, i, seg + 1, m, mbeg, mapnl + moff[i - 1], out1, &map, S, fpx, fpy, fpz, fpx, fpy, fpz, &A);
      if (rinn > 1.0e-5) {
        nnlpp1[c1] = map;
        c1 += map + 1;
      }
    }
    *ncount1 = (int32_t)c1;
    scratch_free(S);
    if (seg_ene) *seg_ene = A.seg;
    if (!(iz == 0 || P.rinn0 >= rout)) {
      *ucoul_slt = A.ucoul[0]; *ucoul_ss = A.ucoul[1]; *ucoul_slv = A.ucoul[2];
      *uconf_slt = A.uconf[0]; *uconf_ss = A.uconf[1]; *uconf_slv = A.uconf[2];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
          for (int k = 0; k < 3; ++k) stress[r + 3 * c] = stress[r + 3 * c] + A.st[3 * r + k] * co[c + 3 * k];
    }
    free(noff); free(moff);
    return status;
  }
#ifdef _OPENMP
  /* OpenMP variant: static blocks of groups, private force/stress/energy
   * accumulators reduced in thread order (force.f:189-212, 842-846). */
  size_t N = (size_t)nato;
  double *priv = calloc(N * 3 * (size_t)nt, sizeof(double));
  force_acc *accs = calloc((size_t)nt, sizeof(force_acc));
  int32_t *inner = NULL;
  long *ioff = NULL;
  int *imap = calloc((size_t)ngrp, sizeof(int));
  if (P.r2inn > 1.0e-5) {
    ioff = malloc(sizeof(long) * ((size_t)ngrp + 1));
    long s = 0;
    for (int i = 1; i <= ngrp; ++i) {
      ioff[i - 1] = s;
      s += (iz == 1 ? nnlpp[noff[i - 1]] : nnlpp[0]) + 1;
    }
    inner = malloc(sizeof(int32_t) * (size_t)(s + 1));
  }
#pragma omp parallel num_threads(nt)
  {
    int t = omp_get_thread_num();
    force_scratch *S = scratch_new(nato, ngrp, cap);
    double *fx = priv + N * 3 * t, *fy = fx + N, *fz = fy + N;
#pragma omp for schedule(static)
    for (int i = 1; i <= ngrp; ++i) {
      const int32_t *seg = nnlpp + (iz == 1 ? noff[i - 1] : 0);
      int m = seg[0];
      int mbeg = iz == 0 ? i + 1 : 1;
      int map = 0;
      force_group(&P, i, seg + 1, m, mbeg, mapnl + moff[i - 1], inner ? inner + ioff[i - 1] : NULL, &map, S,
                  fx, fy, fz, fx, fy, fz, &accs[t]);
      imap[i - 1] = map;
    }
    scratch_free(S);
  }
  long c1 = 0;
  if (rinn > 1.0e-5) {
    for (int i = 1; i <= ngrp; ++i) {
      int map = imap[i - 1];
      if (c1 + 1 + map > nnlpp1_cap) { status = 1; break; }
      nnlpp1[c1] = map;
      memcpy(nnlpp1 + c1 + 1, inner + ioff[i - 1], sizeof(int32_t) * (size_t)map);
      c1 += map + 1;
    }
  }
  *ncount1 = (int32_t)c1;
  if (!(iz == 0 || P.rinn0 >= rout)) {
#pragma omp parallel for schedule(static) num_threads(nt)
    for (long a = 0; a < (long)N; ++a) {
      double sx = 0, sy = 0, sz = 0;
      for (int t = 0; t < nt; ++t) {
        sx += priv[N * 3 * t + a];
        sy += priv[N * 3 * t + N + a];
        sz += priv[N * 3 * t + 2 * N + a];
      }
      fpx[a] += sx; fpy[a] += sy; fpz[a] += sz;
    }
    force_acc A;
    memset(&A, 0, sizeof A);
    for (int t = 0; t < nt; ++t) {
      for (int k = 0; k < 9; ++k) A.st[k] += accs[t].st[k];
      for (int k = 0; k < 3; ++k) { A.uconf[k] += accs[t].uconf[k]; A.ucoul[k] += accs[t].ucoul[k]; }
    }
    *ucoul_slt = A.ucoul[0]; *ucoul_ss = A.ucoul[1]; *ucoul_slv = A.ucoul[2];
    *uconf_slt = A.uconf[0]; *uconf_ss = A.uconf[1]; *uconf_slv = A.uconf[2];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 3; ++k) stress[r + 3 * c] = stress[r + 3 * c] + A.st[3 * r + k] * co[c + 3 * k];
  }
  free(priv); free(accs); free(inner); free(ioff); free(imap);
#endif
  free(noff); free(moff);
  return status;
}

/* mts_forces.f:212-241 */
int32_t orc_decide_force(double rinn, double alphal) {
  double qt = 1.0 / (1.0 + QP * alphal * rinn);
  double expcst = exp(-alphal * rinn * alphal * rinn);
  double erfcst = ((((A5 * qt + A4) * qt + A3) * qt + A2) * qt + A1) * qt * expcst;
  return erfcst < 10.0e-4;
}

/* ferrf.f:1-244; with erf_corr the ERF_CORR branches (ferrf.f:68-105, 170-193)
 * read the spline table erf_arr_corr(4, *) at r = rlew + (ic - 1) delew */
int orc_ferrf(const int32_t *ss_index, double alphal, const double *charge, double fudge, const double *x0,
              const double *y0, const double *z0, int iz, const int32_t *list, int nlist, double *fs_slt,
              double *fs_slv, double *fpx, double *fpy, double *fpz, int32_t erf_corr, int nbonds_added,
              const double *erf_arr_corr, double delew, double rlew) {
  double twrtpi = 2.0 / sqrt(4.0 * atan(1.0));
  double fs[2] = {0.0, 0.0};
  if (erf_corr) {
    for (int i = 0; i < nlist - nbonds_added; ++i) {
      const int ia = list[2 * i], ib = list[2 * i + 1];
      const double qq = charge[ia - 1] * charge[ib - 1];
      const double xab = x0[ia - 1] - x0[ib - 1], yab = y0[ia - 1] - y0[ib - 1], zab = z0[ia - 1] - z0[ib - 1];
      const double rsq = xab * xab + yab * yab + zab * zab;
      const double rsp = sqrt(rsq);
      const int ic = (int)(1 + (rsp - rlew) / delew);
      const double h = rsp - (rlew + (ic - 1) * delew);
      const double *c = erf_arr_corr + 4 * (ic - 1);
      const double alphar = alphal * rsp;
      const double qt = 1.0 / (1.0 + QP * alphar);
      const double expcst = exp(-alphar * alphar);
      const double erfcst = ((((A5 * qt + A4) * qt + A3) * qt + A2) * qt + A1) * qt * expcst;
      const double erfst = 1.0 - erfcst;
      const int type = ss_index[ia - 1];
      if (iz == 1) {
        const double furpar = fudge * qq;
        const double corr = qq * (c[0] + h * (c[1] + h * (c[2] + h * c[3] / 3) / 2));
        const double dcorr = qq * (c[1] + h * (c[2] + h * c[3] / 2)) / rsp;
        const double qforce = -dcorr - furpar * (erfst - twrtpi * alphar * expcst) / (rsp * rsq);
        fs[type - 1] = fs[type - 1] - (furpar * erfst / rsp - corr);
        fpx[ia - 1] = fpx[ia - 1] + qforce * xab;
        fpy[ia - 1] = fpy[ia - 1] + qforce * yab;
        fpz[ia - 1] = fpz[ia - 1] + qforce * zab;
        fpx[ib - 1] = fpx[ib - 1] - qforce * xab;
        fpy[ib - 1] = fpy[ib - 1] - qforce * yab;
        fpz[ib - 1] = fpz[ib - 1] - qforce * zab;
      } else {
        const double furpar = qq;
        const double corr = c[0] + h * (c[1] + h * (c[2] + h * c[3] / 3) / 2);
        fs[type - 1] = fs[type - 1] - furpar * (erfst / rsp - corr);
      }
    }
    *fs_slt = fs[0];
    *fs_slv = fs[1];
    return 0;
  }
  for (int i = 0; i < nlist - nbonds_added; ++i) {
    int ia = list[2 * i], ib = list[2 * i + 1];
    int type = ss_index[ia - 1];
    double furpar = iz == 1 ? fudge * charge[ia - 1] * charge[ib - 1] : charge[ia - 1] * charge[ib - 1];
    double xab = x0[ia - 1] - x0[ib - 1], yab = y0[ia - 1] - y0[ib - 1], zab = z0[ia - 1] - z0[ib - 1];
    double rsq = xab * xab + yab * yab + zab * zab;
    double rsp = sqrt(rsq);
    double alphar = alphal * rsp;
    double qt = 1.0 / (1.0 + QP * alphar);
    double expcst = exp(-(alphar * alphar));
    double erfcst = ((((A5 * qt + A4) * qt + A3) * qt + A2) * qt + A1) * qt * expcst;
    double erfst = 1.0 - erfcst;
    if (iz == 1) {
      double qforce = -(furpar * (erfst - twrtpi * alphar * expcst) / (rsp * rsq));
      fs[type - 1] = fs[type - 1] - furpar * erfst / rsp;
      fpx[ia - 1] = fpx[ia - 1] + qforce * xab;
      fpy[ia - 1] = fpy[ia - 1] + qforce * yab;
      fpz[ia - 1] = fpz[ia - 1] + qforce * zab;
      fpx[ib - 1] = fpx[ib - 1] - qforce * xab;
      fpy[ib - 1] = fpy[ib - 1] - qforce * yab;
      fpz[ib - 1] = fpz[ib - 1] - qforce * zab;
    } else {
      fs[type - 1] = fs[type - 1] - furpar * erfst / rsp;
    }
  }
  *fs_slt = fs[0];
  *fs_slv = fs[1];
  return 0;
}

/* cself.f:1-64 */
void orc_cself(const int32_t *ss_index, int nato, double alphal, double kcut, const double *charge,
               double *self_slt, double *self_slv) {
  double fs[2] = {0.0, 0.0};
  double pi = 4.0 * atan(1.0);
  for (int n = 0; n < nato; ++n) {
    int i = ss_index[n];
    fs[i - 1] = fs[i - 1] + charge[n] * charge[n];
  }
  for (int n = 0; n < 2; ++n) {
    double alphar = kcut / (2.0 * alphal);
    double qt = 1.0 / (1.0 + QP * alphar);
    double expcst = exp(-(alphar * alphar));
    double erfcst = ((((A5 * qt + A4) * qt + A3) * qt + A2) * qt + A1) * qt * expcst;
    fs[n] = -fs[n] * alphal / sqrt(pi) * (1.0 - erfcst);
  }
  *self_slt = fs[0];
  *self_slv = fs[1];
}

/* fnbgrp.f:1-213 */
void orc_fnbgrp(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
                const double *charge, const int32_t *type, const double *ecc12, const double *ecc6,
                int32_t ewald, double alphal, const int32_t *ingrp, int ingrpp, double *fpx, double *fpy,
                double *fpz, double *uconf_slt, double *uconf_slv, double *ucoul_slt, double *ucoul_slv) {
  double uconf[2] = {0, 0}, ucoul[2] = {0, 0};
  *uconf_slt = *uconf_slv = *ucoul_slt = *ucoul_slv = 0.0;
  if (ingrpp == 0) return;
  double twrtpi = 2.0 / sqrt(4.0 * atan(1.0));
  for (int i = 0; i < ingrpp; ++i) {
    int i1 = ingrp[2 * i], i2 = ingrp[2 * i + 1];
    int typei = ss_index[i1 - 1];
    int lj = type[i1 - 1] > type[i2 - 1] ? type[i1 - 1] : type[i2 - 1];
    int li = type[i1 - 1] < type[i2 - 1] ? type[i1 - 1] : type[i2 - 1];
    int lij = lj * (lj - 1) / 2 + li;
    double xa = xp0[i1 - 1] - xp0[i2 - 1], ya = yp0[i1 - 1] - yp0[i2 - 1], za = zp0[i1 - 1] - zp0[i2 - 1];
    double rsq = xa * xa + ya * ya + za * za;
    double rsp = sqrt(rsq);
    double rsqi = 1.0 / rsq;
    double chi = charge[i1 - 1], chj = charge[i2 - 1];
    double qforce;
    if (!ewald) {
      double r6 = rsqi * rsqi * rsqi, r12 = r6 * r6;
      double ssvir = 12.0 * ecc12[lij - 1] * r12 - 6.0 * ecc6[lij - 1] * r6;
      qforce = ssvir * rsqi;
      uconf[typei - 1] = uconf[typei - 1] + (ecc12[lij - 1] * r12 - ecc6[lij - 1] * r6);
      ssvir = chi * chj / rsp;
      ucoul[typei - 1] = ucoul[typei - 1] + ssvir;
      qforce = qforce + ssvir * rsqi;
    } else {
      double rspqi = rsqi / rsp;
      double r6 = rsqi * rsqi * rsqi, r12 = r6 * r6;
      double ssvir = 12.0 * ecc12[lij - 1] * r12 - 6.0 * ecc6[lij - 1] * r6;
      qforce = ssvir * rsqi;
      uconf[typei - 1] = uconf[typei - 1] + (ecc12[lij - 1] * r12 - ecc6[lij - 1] * r6);
      double alphar = alphal * rsp;
      double qt = 1.0 / (1.0 + QP * alphar);
      double expcst = exp(-(alphar * alphar));
      double erfcst = ((((A5 * qt + A4) * qt + A3) * qt + A2) * qt + A1) * qt * expcst;
      double furpar = chi * chj;
      ucoul[typei - 1] = ucoul[typei - 1] + furpar * erfcst / rsp;
      qforce = qforce + furpar * (erfcst + twrtpi * alphar * expcst) * rspqi;
    }
    fpx[i1 - 1] = fpx[i1 - 1] + qforce * xa; fpy[i1 - 1] = fpy[i1 - 1] + qforce * ya;
    fpz[i1 - 1] = fpz[i1 - 1] + qforce * za;
    fpx[i2 - 1] = fpx[i2 - 1] - qforce * xa; fpy[i2 - 1] = fpy[i2 - 1] - qforce * ya;
    fpz[i2 - 1] = fpz[i2 - 1] - qforce * za;
  }
  *uconf_slt = uconf[0]; *uconf_slv = uconf[1];
  *ucoul_slt = ucoul[0]; *ucoul_slv = ucoul[1];
}

/* ------------------------------------------------------------------ */
/* Alchemy: force_alchemy.f (soft-core GROMOS), alch_set                */
/* (force_alchemy_shifted.f:732-790), ferrf_alchemy.f, cself_alchemy.f  */
/* ------------------------------------------------------------------ */

/* alch_set: fact[2 * (typeaij + 4) + {0, 1}] = alch_fact(1:2, typeaij) for
 * typeaij = s_i + s_j (s = 1 normal, 2 added, -2 removed); entries the
 * reference leaves untouched stay 0 (force_alchemy.f:193-194 zeroes the LJ
 * tables before the call). */
void orc_alch_set(double *fact, double ladd, double lrm, int intramol, int ljintra, int qq) {
  memset(fact, 0, sizeof(double) * 18);
#define AF(t, k) fact[2 * ((t) + 4) + (k)]
  const double a = fabs(ladd), r = fabs(lrm);
  AF(2, 0) = 0.0; AF(2, 1) = 1.0;
  AF(3, 0) = a; AF(3, 1) = 1.0 - a;
  AF(-1, 0) = r; AF(-1, 1) = 1.0 - r;
  if (intramol && qq) {
    AF(4, 0) = a * a; AF(4, 1) = 1.0 - a * a;
    AF(-4, 0) = r * r; AF(-4, 1) = 1.0 - r * r;
  } else if (intramol && !ljintra) {
    AF(4, 0) = a; AF(4, 1) = 1.0 - a;
    AF(-4, 0) = r; AF(-4, 1) = 1.0 - r;
  } else {
    AF(4, 0) = 0.0; AF(4, 1) = 1.0;
    AF(-4, 0) = 0.0; AF(-4, 1) = 1.0;
  }
  AF(0, 0) = 0.0; AF(0, 1) = 0.0;
#undef AF
}

/* x**6 with an integer exponent, as the compiled Fortran evaluates it (five multiplications, left to right) */
static double ipow6(double x) {
  return x * x * x * x * x * x;
}

/* force_alchemy.f:147-730, serial build.  The module state the routine reads
 * arrives as arguments: ss_alchemy (cpropar), lam = {ladd, lrm, ladd0, lrm0,
 * laddq, lrmq, laddq0, lrmq0}, intramol, ljintra, opls (ffield_type = 'opls').
 * e6 = uconf_slt, uconf_slv, uconf_ss, ucoul_slt, ucoul_slv, ucoul_ss.  As in
 * the serial build, dwrk is added to (not reset) and energies are written only
 * when iz = 1 and rinn0 < rout. */
static int force_alchemy_impl(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
                      const double *xpg, const double *ypg, const double *zpg, const double *charge,
                      const int32_t *ss_alchemy, const double *rmin, const double *epsm, int nato,
                      const double *xpcm, const double *ypcm, const double *zpcm, const int32_t *atomp,
                      const double *co, double alphal, const int32_t *mapnl, int ngrp, const int32_t *grppt,
                      double *e6, double *fpx, double *fpy, double *fpz, double *stress, const int32_t *nnlpp,
                      int32_t *nnlpp1, long nnlpp1_cap, double rneigh, double rinn, double rout, double *rtolinn,
                      double rtolout, const double *gmass, int iz, int32_t *ncount1, const double *mass,
                      double massinfty, const double *lam, int intramol, int ljintra, int opls, double *dwrk,
                      int32_t lskip_ewald, int32_t ewald, double qsoft, double ljsoft, int shifted) {
  /* shifted: force_alchemy_shift (force_alchemy_shifted.f:1-730), with qsoft/ljsoft
   * standing for qshift/ljshift; two6 = 2**(1./6.) with the REAL(4) exponent
   * (force_alchemy_shifted.f:178) */
  const double two6 = pow(2.0, (double)(1.0f / 6.0f));
  const double r2inn = rinn * rinn, r2out = rout * rout;
  const double rinn0 = rinn + *rtolinn, r2inn0 = rinn0 * rinn0;
  if (rinn < 1.0e-5) *rtolinn = 0.0;
  const double rout0 = rout + rtolout, r2out0 = rout0 * rout0;
  const double r2neigh = (rinn0 + rneigh) * (rinn0 + rneigh);
  const double twrtpi = 2.0 / sqrt(4.0 * atan(1.0));
  const double arsout1 = rout0 - 3.0 * rout, arsout2 = rtolout * rtolout * rtolout;
  const double arsinn1 = rinn0 - 3.0 * rinn, arsinn2 = *rtolinn * *rtolinn * *rtolinn;
  double fl[18], fl0[18], fq[18], fq0[18];
  orc_alch_set(fl, lam[0], lam[1], intramol, ljintra, 0);
  orc_alch_set(fl0, lam[2], lam[3], intramol, ljintra, 0);
  orc_alch_set(fq, lam[4], lam[5], intramol, ljintra, 1);
  orc_alch_set(fq0, lam[6], lam[7], intramol, ljintra, 1);
  double st[9] = {0}, uconf[3] = {0}, ucoul[3] = {0}, uconf0[3] = {0}, ucoul0[3] = {0};
  int cap = 1;
  {
    long n = 0;
    for (int i = 1; i <= ngrp; ++i) {
      int m = nnlpp[n];
      if (m + 1 > cap) cap = m + 1;
      if (iz == 1) n += m + 1; else break;
    }
  }
  int *index1 = malloc(sizeof(int) * cap);
  double *xm1 = malloc(sizeof(double) * cap * 9);
  double *ym1 = xm1 + cap, *zm1 = ym1 + cap, *xm2 = zm1 + cap, *ym2 = xm2 + cap, *zm2 = ym2 + cap;
  double *dr = zm2 + cap, *swrs = dr + cap, *dswrs = swrs + cap;
  double *cmap2 = malloc(sizeof(double) * cap);
  char *maplg = malloc((size_t)nato + 1);
  memset(maplg, 1, (size_t)nato + 1);
  long n = 0, na = 0, c1 = 0;
  int mbeg = 1, status = 0;
  for (int i = 1; i <= ngrp; ++i) {
    const double xpgi = xpg[i - 1], ypgi = ypg[i - 1], zpgi = zpg[i - 1];
    const int m = nnlpp[n];
    int map = 0, mp = 0;
    const int typei = ss_index[grppt[2 * (i - 1)] - 1];
    if (iz == 0) mbeg = i + 1;
    if (r2inn > 1.0e-5 && c1 + 1 + m > nnlpp1_cap) { status = 1; break; }
    for (int jj = mbeg; jj <= m; ++jj) {
      const int j = nnlpp[jj + n];
      double xgg = xpgi - xpg[j - 1], ygg = ypgi - ypg[j - 1], zgg = zpgi - zpg[j - 1];
      const double sx = 2.0 * fnint(0.5 * xgg), sy = 2.0 * fnint(0.5 * ygg), sz = 2.0 * fnint(0.5 * zgg);
      xgg = xgg - sx; ygg = ygg - sy; zgg = zgg - sz;
      const double xc = co[0] * xgg + co[3] * ygg + co[6] * zgg;
      const double yc = co[4] * ygg + co[7] * zgg;
      const double zc = co[8] * zgg;
      const double drj = xc * xc + yc * yc + zc * zc;
      if (r2inn > 1.0e-5 && drj < r2neigh) {
        map += 1;
        nnlpp1[c1 + map] = j;
      }
      if (rinn0 > rout) continue;
      if (drj >= r2inn && drj < r2out0) {
        xm1[mp] = sx; ym1[mp] = sy; zm1[mp] = sz;
        xm2[mp] = xc; ym2[mp] = yc; zm2[mp] = zc;
        dr[mp] = drj;
        index1[mp] = j;
        mp += 1;
      }
    }
    if (rinn > 1.0e-5) {
      nnlpp1[c1] = map;
      c1 += map + 1;
    }
    const int mapb = mp;
    /* force_alchemy.f:344-368 */
    for (int jj = 0; jj < mapb; ++jj) {
      if (dr[jj] < r2out && dr[jj] >= r2inn0) {
        swrs[jj] = 1.0;
        dswrs[jj] = 0.0;
      } else if (dr[jj] >= r2out) {
        const double rsp = sqrt(dr[jj]);
        const double auxa = (arsout1 + 2.0 * rsp) / arsout2, auxb = rout0 - rsp;
        swrs[jj] = auxa * (auxb * auxb);
        dswrs[jj] = (-2.0 * auxa * auxb + 2.0 * (auxb * auxb) / arsout2) / rsp;
      } else {
        const double rsp = sqrt(dr[jj]);
        const double auxa = (arsinn1 + 2.0 * rsp) / arsinn2, auxb = rinn0 - rsp;
        swrs[jj] = 1.0 - auxa * (auxb * auxb);
        dswrs[jj] = (2.0 * auxa * auxb - 2.0 * (auxb * auxb) / arsinn2) / rsp;
      }
    }
    for (int jj = 0; jj < mapb; ++jj) cmap2[jj] = 0.0;
    const int ga = grppt[2 * (i - 1)], gb = grppt[2 * (i - 1) + 1];
    if (mapb == 0) {
      for (int i1 = ga; i1 <= gb; ++i1) na += mapnl[na] + 1;
    } else {
      for (int i1 = ga; i1 <= gb; ++i1) {
        const double xpi = xp0[i1 - 1], ypi = yp0[i1 - 1], zpi = zp0[i1 - 1];
        const double chrgei = charge[i1 - 1];
        const int p1 = atomp[i1 - 1];
        const double xpcmp1 = xpcm[p1 - 1], ypcmp1 = ypcm[p1 - 1], zpcmp1 = zpcm[p1 - 1];
        const int la = mapnl[na];
        for (int t = 1; t <= la; ++t) maplg[mapnl[na + t]] = 0;
        /* pairs in group-then-atom order (force_alchemy.f:411-424), computed in place */
        for (int jj = 0; jj < mapb; ++jj) {
          const int jg = index1[jj];
          const int typej = ss_index[grppt[2 * (jg - 1)] - 1];
          const int typeij = typei + typej - 1;
          for (int j = grppt[2 * (jg - 1)]; j <= grppt[2 * (jg - 1) + 1]; ++j) {
            if (!(maplg[j] && (mass[j - 1] < massinfty || mass[i1 - 1] < massinfty))) continue;
            const double sigma = opls ? 2.0 * pow(rmin[j - 1] * rmin[i1 - 1], 0.5) : rmin[j - 1] + rmin[i1 - 1];
            const double epsmin = 4.0 * sqrt(epsm[i1 - 1] * epsm[j - 1]);
            if (epsmin < 1.0e-12 && lskip_ewald && ewald) continue;
            const int ta = ss_alchemy[i1 - 1] + ss_alchemy[j - 1];
            const double *F = fl + 2 * (ta + 4), *F0 = fl0 + 2 * (ta + 4);
            const double *Q = fq + 2 * (ta + 4), *Q0 = fq0 + 2 * (ta + 4);
            const int p2 = atomp[j - 1];
            const double xxcm = xpcmp1 - xpcm[p2 - 1] - xm1[jj];
            const double yycm = ypcmp1 - ypcm[p2 - 1] - ym1[jj];
            const double zzcm = zpcmp1 - zpcm[p2 - 1] - zm1[jj];
            const double xg = xpi - xp0[j - 1] - xm1[jj];
            const double yg = ypi - yp0[j - 1] - ym1[jj];
            const double zg = zpi - zp0[j - 1] - zm1[jj];
            const double xc = co[0] * xg + co[3] * yg + co[6] * zg;
            const double yc = co[4] * yg + co[7] * zg;
            const double zc = co[8] * zg;
            const double rsq = xc * xc + yc * yc + zc * zc;
            const double sw = swrs[jj];
            double qforce;
            if (shifted) {
              /* force_alchemy_shifted.f:469-509 */
              const double r = sqrt(rsq);
              if (epsmin > 1.0e-12) {
                const double rl = r + ljsoft * sigma * two6 * F[0];
                const double rl0 = r + ljsoft * sigma * two6 * F0[0];
                const double r6 = ipow6(sigma / rl), r60 = ipow6(sigma / rl0);
                const double ssvirr = F[1] * epsmin * (12.0 * r6 * r6 - 6.0 * r6);
                const double auxc = epsmin * (r6 * r6 - r6), auxc0 = epsmin * (r60 * r60 - r60);
                const double ucon = F[1] * auxc, ucon0 = F0[1] * auxc0;
                qforce = ssvirr * sw / (r * rl);
                cmap2[jj] = cmap2[jj] + ucon;
                uconf[typeij - 1] = uconf[typeij - 1] + sw * ucon;
                uconf0[typeij - 1] = uconf0[typeij - 1] + sw * ucon0;
              } else {
                qforce = 0.0;
              }
              const double rq = r + qsoft * Q[0], rq0 = r + qsoft * Q0[0];
              if (!lskip_ewald) {
                const double alphar = alphal * r;
                const double qt = 1.0 / (1.0 + QP * alphar);
                const double expcst = exp(-alphar * alphar);
                const double erfcst = ((((A5 * qt + A4) * qt + A3) * qt + A2) * qt + A1) * qt * expcst;
                const double furpar = Q[1] * chrgei * charge[j - 1];
                const double furpar0 = Q0[1] * chrgei * charge[j - 1];
                ucoul[typeij - 1] = ucoul[typeij - 1] + sw * furpar * erfcst / rq;
                ucoul0[typeij - 1] = ucoul0[typeij - 1] + sw * furpar0 * erfcst / rq0;
                cmap2[jj] = cmap2[jj] + furpar * erfcst / rq;
                qforce = qforce + furpar * (erfcst * sw / (rq * rq * r) + twrtpi * alphar * expcst * sw / (r * r * rq));
              } else if (!ewald) {
                const double ssvir = Q[1] * chrgei * charge[j - 1] / rq;
                const double ssvir0 = Q0[1] * chrgei * charge[j - 1] / rq0;
                cmap2[jj] = cmap2[jj] + ssvir;
                ucoul[typeij - 1] = ucoul[typeij - 1] + ssvir * sw;
                ucoul0[typeij - 1] = ucoul0[typeij - 1] + ssvir0 * sw;
                qforce = qforce + ssvir * sw / (rq * r);
              }
              goto accumulate;
            }
            const double auxa = 1.0 / sqrt(rsq);
            const double auxaa = 1.0 / sqrt(qsoft * pow(Q[0], 2.0) + rsq);
            const double auxa0 = 1.0 / sqrt(qsoft * pow(Q0[0], 2.0) + rsq);
            const double auxbb = auxaa * auxaa * auxaa;
            if (epsmin > 1.0e-12) {
              const double r6 = rsq * rsq * rsq / pow(sigma, 6.0);
              const double gp = ljsoft * F[0] + r6, gp1 = 1.0 / gp, gp2 = gp1 * gp1;
              const double gp0 = ljsoft * F0[0] + r6, gp10 = 1.0 / gp0, gp20 = gp10 * gp10;
              const double gr = 6.0 * r6 * auxa;
              const double ssvir = F[1] * epsmin * (2.0 * gp1 * gp2 - gp2);
              const double ssvirr = gr * ssvir;
              const double auxc = epsmin * (gp2 - gp1), auxc0 = epsmin * (gp20 - gp10);
              const double ucon = F[1] * auxc, ucon0 = F0[1] * auxc0;
              qforce = ssvirr * sw * auxa;
              cmap2[jj] = cmap2[jj] + ucon;
              uconf[typeij - 1] = uconf[typeij - 1] + sw * ucon;
              uconf0[typeij - 1] = uconf0[typeij - 1] + sw * ucon0;
            } else {
              qforce = 0.0;
            }
            if (!lskip_ewald) {
              const double alphar = alphal / auxa;
              const double qt = 1.0 / (1.0 + QP * alphar);
              const double expcst = exp(-alphar * alphar);
              const double erfcst = ((((A5 * qt + A4) * qt + A3) * qt + A2) * qt + A1) * qt * expcst;
              const double furpar = Q[1] * chrgei * charge[j - 1];
              const double furpar0 = Q0[1] * chrgei * charge[j - 1];
              ucoul[typeij - 1] = ucoul[typeij - 1] + sw * furpar * erfcst * auxaa;
              ucoul0[typeij - 1] = ucoul0[typeij - 1] + sw * furpar0 * erfcst * auxa0;
              cmap2[jj] = cmap2[jj] + furpar * erfcst * auxaa;
              qforce = qforce + furpar * erfcst * auxbb * sw +
                       furpar * twrtpi * alphar * expcst * auxa * auxa * auxaa * sw;
            } else if (!ewald) {
              const double auxc = chrgei * charge[j - 1] * auxaa;
              const double ssvir = Q[1] * auxc;
              const double ssvir0 = Q0[1] * chrgei * charge[j - 1] * auxa0;
              cmap2[jj] = cmap2[jj] + ssvir;
              ucoul[typeij - 1] = ucoul[typeij - 1] + ssvir * sw;
              ucoul0[typeij - 1] = ucoul0[typeij - 1] + ssvir0 * sw;
              qforce = qforce + ssvir * sw * auxaa * auxaa;
            }
          accumulate:;
            const double emvir = qforce;
            fpx[i1 - 1] = fpx[i1 - 1] + qforce * xc;
            fpy[i1 - 1] = fpy[i1 - 1] + qforce * yc;
            fpz[i1 - 1] = fpz[i1 - 1] + qforce * zc;
            fpx[j - 1] = fpx[j - 1] - qforce * xc;
            fpy[j - 1] = fpy[j - 1] - qforce * yc;
            fpz[j - 1] = fpz[j - 1] - qforce * zc;
            st[0] = st[0] + emvir * xc * xxcm; st[1] = st[1] + emvir * xc * yycm; st[2] = st[2] + emvir * xc * zzcm;
            st[3] = st[3] + emvir * yc * xxcm; st[4] = st[4] + emvir * yc * yycm; st[5] = st[5] + emvir * yc * zzcm;
            st[6] = st[6] + emvir * zc * xxcm; st[7] = st[7] + emvir * zc * yycm; st[8] = st[8] + emvir * zc * zzcm;
          }
        }
        for (int t = 1; t <= la; ++t) maplg[mapnl[na + t]] = 1;
        na += la + 1;
      }
      /* S'(r) V by mass fraction, force_alchemy.f:559-622 */
      for (int jj = 0; jj < mapb; ++jj) {
        const double x3 = -dswrs[jj] * cmap2[jj] * xm2[jj];
        const double y3 = -dswrs[jj] * cmap2[jj] * ym2[jj];
        const double z3 = -dswrs[jj] * cmap2[jj] * zm2[jj];
        xm2[jj] = x3; ym2[jj] = y3; zm2[jj] = z3;  /* xmap3 (xmap2 is not read again) */
      }
      for (int i1 = ga; i1 <= gb; ++i1) {
        const double massi = gmass[i1 - 1];
        const int p1 = atomp[i1 - 1];
        const double xq = xpcm[p1 - 1], yq = ypcm[p1 - 1], zq = zpcm[p1 - 1];
        for (int jj = 0; jj < mapb; ++jj) {
          fpx[i1 - 1] = fpx[i1 - 1] + massi * xm2[jj];
          fpy[i1 - 1] = fpy[i1 - 1] + massi * ym2[jj];
          fpz[i1 - 1] = fpz[i1 - 1] + massi * zm2[jj];
          st[0] = st[0] + massi * xm2[jj] * xq; st[1] = st[1] + massi * xm2[jj] * yq;
          st[2] = st[2] + massi * xm2[jj] * zq; st[3] = st[3] + massi * ym2[jj] * xq;
          st[4] = st[4] + massi * ym2[jj] * yq; st[5] = st[5] + massi * ym2[jj] * zq;
          st[6] = st[6] + massi * zm2[jj] * xq; st[7] = st[7] + massi * zm2[jj] * yq;
          st[8] = st[8] + massi * zm2[jj] * zq;
        }
      }
      for (int jj = 0; jj < mapb; ++jj) {
        const int jg = index1[jj];
        for (int j1 = grppt[2 * (jg - 1)]; j1 <= grppt[2 * (jg - 1) + 1]; ++j1) {
          const double massj = gmass[j1 - 1];
          fpx[j1 - 1] = fpx[j1 - 1] - massj * xm2[jj];
          fpy[j1 - 1] = fpy[j1 - 1] - massj * ym2[jj];
          fpz[j1 - 1] = fpz[j1 - 1] - massj * zm2[jj];
          const int p2 = atomp[j1 - 1];
          const double xq = -(xpcm[p2 - 1] + xm1[jj]) * massj;
          const double yq = -(ypcm[p2 - 1] + ym1[jj]) * massj;
          const double zq = -(zpcm[p2 - 1] + zm1[jj]) * massj;
          st[0] = st[0] + xm2[jj] * xq; st[1] = st[1] + xm2[jj] * yq; st[2] = st[2] + xm2[jj] * zq;
          st[3] = st[3] + ym2[jj] * xq; st[4] = st[4] + ym2[jj] * yq; st[5] = st[5] + ym2[jj] * zq;
          st[6] = st[6] + zm2[jj] * xq; st[7] = st[7] + zm2[jj] * yq; st[8] = st[8] + zm2[jj] * zq;
        }
      }
    }
    if (iz == 1) n += m + 1;
  }
  *ncount1 = (int32_t)c1;
  free(index1); free(xm1); free(cmap2); free(maplg);
  if (status || iz == 0 || rinn0 >= rout) return status;
  e6[0] = uconf[0]; e6[1] = uconf[2]; e6[2] = uconf[1];
  e6[3] = ucoul[0]; e6[4] = ucoul[2]; e6[5] = ucoul[1];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < 3; ++k) stress[r + 3 * c] = stress[r + 3 * c] + st[3 * r + k] * co[c + 3 * k];
  for (int k = 0; k < 3; ++k) {
    const double auxa = ucoul[k] - ucoul0[k], auxb = uconf[k] - uconf0[k];
    *dwrk = *dwrk + auxa + auxb;
  }
  return 0;
}

int orc_force_alchemy(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
                      const double *xpg, const double *ypg, const double *zpg, const double *charge,
                      const int32_t *ss_alchemy, const double *rmin, const double *epsm, int nato,
                      const double *xpcm, const double *ypcm, const double *zpcm, const int32_t *atomp,
                      const double *co, double alphal, const int32_t *mapnl, int ngrp, const int32_t *grppt,
                      double *e6, double *fpx, double *fpy, double *fpz, double *stress, const int32_t *nnlpp,
                      int32_t *nnlpp1, long nnlpp1_cap, double rneigh, double rinn, double rout, double *rtolinn,
                      double rtolout, const double *gmass, int iz, int32_t *ncount1, const double *mass,
                      double massinfty, const double *lam, int intramol, int ljintra, int opls, double *dwrk,
                      int32_t lskip_ewald, int32_t ewald, double qsoft, double ljsoft, int shifted) {
  return force_alchemy_impl(ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, ss_alchemy, rmin, epsm, nato, xpcm,
                            ypcm, zpcm, atomp, co, alphal, mapnl, ngrp, grppt, e6, fpx, fpy, fpz, stress, nnlpp,
                            nnlpp1, nnlpp1_cap, rneigh, rinn, rout, rtolinn, rtolout, gmass, iz, ncount1, mass,
                            massinfty, lam, intramol, ljintra, opls, dwrk, lskip_ewald, ewald, qsoft, ljsoft,
                            shifted);
}

/* ferrf_alchemy.f:1-204 (the ERF_CORR spline branches are not restated):
 * ferrf on charge, plus en0 on charge0; dwrk = slt + slv - en0. */
int orc_ferrf_alchemy(const int32_t *ss_index, double alphal, const double *charge, const double *charge0,
                      double fudge, const double *x0, const double *y0, const double *z0, int iz,
                      const int32_t *list, int nlist, double *fs_slt, double *fs_slv, double *fpx, double *fpy,
                      double *fpz, int nbonds_added, double *dwrk) {
  const double twrtpi = 2.0 / sqrt(4.0 * atan(1.0));
  double fs[2] = {0.0, 0.0}, en0 = 0.0;
  for (int i = 0; i < nlist - nbonds_added; ++i) {
    const int ia = list[2 * i], ib = list[2 * i + 1];
    const int type = ss_index[ia - 1];
    /* ferrf_alchemy.f:144-145 (iz = 1) and 190-191, 232-233 (iz = 0) */
    const double furpar = iz == 1 ? fudge * charge[ia - 1] * charge[ib - 1] : charge[ia - 1] * charge[ib - 1];
    const double furpar0 = iz == 1 ? fudge * charge0[ia - 1] * charge0[ib - 1] : charge0[ia - 1] * charge0[ib - 1];
    const double xab = x0[ia - 1] - x0[ib - 1], yab = y0[ia - 1] - y0[ib - 1], zab = z0[ia - 1] - z0[ib - 1];
    const double rsq = xab * xab + yab * yab + zab * zab;
    const double rsp = sqrt(rsq);
    const double alphar = alphal * rsp;
    const double qt = 1.0 / (1.0 + QP * alphar);
    const double expcst = exp(-alphar * alphar);
    const double erfcst = ((((A5 * qt + A4) * qt + A3) * qt + A2) * qt + A1) * qt * expcst;
    const double erfst = 1.0 - erfcst;
    fs[type - 1] = fs[type - 1] - furpar * erfst / rsp;
    en0 = en0 - furpar0 * erfst / rsp;
    if (iz == 1) {
      const double qforce = -furpar * (erfst - twrtpi * alphar * expcst) / (rsp * rsq);
      fpx[ia - 1] = fpx[ia - 1] + qforce * xab;
      fpy[ia - 1] = fpy[ia - 1] + qforce * yab;
      fpz[ia - 1] = fpz[ia - 1] + qforce * zab;
      fpx[ib - 1] = fpx[ib - 1] - qforce * xab;
      fpy[ib - 1] = fpy[ib - 1] - qforce * yab;
      fpz[ib - 1] = fpz[ib - 1] - qforce * zab;
    }
  }
  *fs_slt = fs[0];
  *fs_slv = fs[1];
  *dwrk = fs[0] + fs[1] - en0;
  return 0;
}

/* cself_alchemy.f:1-66: solute atoms only */
void orc_cself_alchemy(const int32_t *ss_index, int nato, double alphal, double kcut, const double *charge,
                       const double *charge0, double *self_slt, double *dwrk) {
  double fs = 0.0, en0 = 0.0;
  const double pi = 4.0 * atan(1.0);
  for (int n = 0; n < nato; ++n)
    if (ss_index[n] == 1) {
      fs = fs + charge[n] * charge[n];
      en0 = en0 + charge0[n] * charge0[n];
    }
  const double alphar = kcut / (2.0 * alphal);
  const double qt = 1.0 / (1.0 + QP * alphar);
  const double expcst = exp(-alphar * alphar);
  const double erfcst = ((((A5 * qt + A4) * qt + A3) * qt + A2) * qt + A1) * qt * expcst;
  fs = -fs * alphal / sqrt(pi) * (1.0 - erfcst);
  en0 = -en0 * alphal / sqrt(pi) * (1.0 - erfcst);
  *self_slt = fs;
  *dwrk = fs - en0;
}

/* fnb14.f:1-219, serial build: the 1-4 pairs with LJ-FUDGE and QQ-FUDGE;
 * erfc when ewald and cut, plain Coulomb otherwise.  Energies by the
 * ss_index of the first atom. */
/* remint (fnb14_rem.f, null for fnb14.f): the flagged pairs' energy goes to
 * *seg and their force is scaled by f2 = rem_factor(2) (fnb14_rem.f:146-149,
 * :209-212) */
static void fnb14_impl(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
                       const double *charge, const double *ecc12, const double *ecc6, int cut, int ewald,
                       double alphal, const int32_t *int14, int int14p, const int32_t *type14, double fudge,
                       double lj_fudge, double *fpx, double *fpy, double *fpz, double *uconf_slt, double *uconf_slv,
                       double *ucoul_slt, double *ucoul_slv, const int32_t *remint, double f2, double *seg) {
  double uconf[2] = {0.0, 0.0}, ucoul[2] = {0.0, 0.0};
  *uconf_slt = *uconf_slv = *ucoul_slt = *ucoul_slv = 0.0;
  if (int14p == 0) return; /* (fnb14_rem.f:100: before seg_ene is zeroed, which is then left as it was) */
  if (seg) *seg = 0.0;
  const double twrtpi = 2.0 / sqrt(4.0 * atan(1.0));
  const int erfc_branch = ewald && cut;
  for (int i = 0; i < int14p; ++i) {
    const int i1 = int14[2 * i], i2 = int14[2 * i + 1];
    const int type = ss_index[i1 - 1];
    const double xa = xp0[i1 - 1] - xp0[i2 - 1], ya = yp0[i1 - 1] - yp0[i2 - 1], za = zp0[i1 - 1] - zp0[i2 - 1];
    const int lij = type14[i];
    const double chrgei = charge[i1 - 1], chrgej = charge[i2 - 1];
    const double rsq = xa * xa + ya * ya + za * za;
    const double rsp = sqrt(rsq);
    const double rsqi = 1.0 / rsq;
    const double r6 = rsqi * rsqi * rsqi, r12 = r6 * r6;
    const double ssvir = 12.0 * ecc12[lij - 1] * r12 - 6.0 * ecc6[lij - 1] * r6;
    double qforce;
    if (!erfc_branch) {
      /* fnb14.f:69-105 */
      qforce = lj_fudge * ssvir * rsqi;
      const double elj = lj_fudge * (ecc12[lij - 1] * r12 - ecc6[lij - 1] * r6);
      uconf[type - 1] = uconf[type - 1] + elj;
      const double sv = fudge * chrgei * chrgej / rsp;
      ucoul[type - 1] = ucoul[type - 1] + sv;
      qforce = qforce + sv * rsqi;
      if (remint && remint[i]) {
        *seg = *seg + elj + sv;
        qforce = qforce * f2;
      }
    } else {
      /* fnb14.f:107-152 */
      const double rspqi = rsqi / rsp;
      const double ucon = lj_fudge * (ecc12[lij - 1] * r12 - ecc6[lij - 1] * r6);
      uconf[type - 1] = uconf[type - 1] + ucon;
      const double alphar = alphal * rsp;
      const double qt = 1.0 / (1.0 + QP * alphar);
      const double expcst = exp(-alphar * alphar);
      const double erfcst = ((((A5 * qt + A4) * qt + A3) * qt + A2) * qt + A1) * qt * expcst;
      const double furpar = fudge * chrgei * chrgej;
      const double ucou = furpar * erfcst / rsp;
      ucoul[type - 1] = ucoul[type - 1] + ucou;
      qforce = lj_fudge * ssvir * rsqi + furpar * (erfcst + twrtpi * alphar * expcst) * rspqi;
      if (remint && remint[i]) {
        *seg = *seg + ucon + ucou;
        qforce = qforce * f2;
      }
    }
    fpx[i1 - 1] = fpx[i1 - 1] + qforce * xa;
    fpy[i1 - 1] = fpy[i1 - 1] + qforce * ya;
    fpz[i1 - 1] = fpz[i1 - 1] + qforce * za;
    fpx[i2 - 1] = fpx[i2 - 1] - qforce * xa;
    fpy[i2 - 1] = fpy[i2 - 1] - qforce * ya;
    fpz[i2 - 1] = fpz[i2 - 1] - qforce * za;
  }
  *uconf_slt = uconf[0];
  *uconf_slv = uconf[1];
  *ucoul_slt = ucoul[0];
  *ucoul_slv = ucoul[1];
}

void orc_fnb14(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
               const double *charge, int nato, const double *ecc12, const double *ecc6, int cut, int ewald,
               double alphal, const int32_t *int14, int int14p, const int32_t *type14, double fudge, double lj_fudge,
               double *fpx, double *fpy, double *fpz, double *uconf_slt, double *uconf_slv, double *ucoul_slt,
               double *ucoul_slv) {
  (void)nato;
  fnb14_impl(ss_index, xp0, yp0, zp0, charge, ecc12, ecc6, cut, ewald, alphal, int14, int14p, type14, fudge,
             lj_fudge, fpx, fpy, fpz, uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, NULL, 1.0, NULL);
}

/* fnb14_rem.f:1-234 (serial): fnb14 with remint14 */
void orc_fnb14_rem(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
                   const double *charge, int nato, const double *ecc12, const double *ecc6, int cut, int ewald,
                   double alphal, const int32_t *int14, int int14p, const int32_t *type14, double fudge,
                   double lj_fudge, double *fpx, double *fpy, double *fpz, double *uconf_slt, double *uconf_slv,
                   double *ucoul_slt, double *ucoul_slv, const int32_t *remint14, double f2, double *seg_ene) {
  (void)nato;
  fnb14_impl(ss_index, xp0, yp0, zp0, charge, ecc12, ecc6, cut, ewald, alphal, int14, int14p, type14, fudge,
             lj_fudge, fpx, fpy, fpz, uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, remint14, f2, seg_ene);
}

/* ------------------------------------------------------------------ */
/* fptors.f:1-345, serial build: proper torsions                       */
/* ------------------------------------------------------------------ */

static double powl_ltr(double x, int n) { /* x**n for integer n >= 1, products left to right */
  double r = x;
  for (int k = 1; k < n; ++k) r = r * x;
  return r;
}

static double fsign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

/* Returns 0, or 1 at a torsion whose bond angles are singular (the reference
 * then perturbs the four atoms at random, fptors.f:107-130, which is not
 * restated), or 2 at an unsupported multiplicity (the reference stops). */
/* rem (fptors_rem.f, 0 for fptors.f): the 60-degree-phase forms (12, 14)
 * without the sign factor and with dsoa = 0 where |sin| < 1e-8
 * (fptors_rem.f:262-301); the flagged torsions' energy goes to *seg and their
 * force is scaled by f2 (:310-313) */
static int fptors_impl(const int32_t *ss_index, const int32_t *ltor, int ltors, const double *xp0,
                       const double *yp0, const double *zp0, const double *pota, const double *potb,
                       double *utors_slt, double *utors_slv, double *fpx, double *fpy, double *fpz, int rem,
                       const int32_t *remltor, double f2, double *seg) {
  const double sin60 = (double)0.866025f, lim = (double)1.000001f;
  double utors[2] = {0.0, 0.0};
  int status = 0;
  if (seg) *seg = 0.0;
  for (int i = 0; i < ltors; ++i) {
    if (fabs(potb[i]) < 1.0e-12) continue;
    const int l1 = ltor[4 * i], l2 = ltor[4 * i + 1], l3 = ltor[4 * i + 2], l4 = ltor[4 * i + 3];
    const int type = ss_index[l1 - 1];
    const double x21 = xp0[l2 - 1] - xp0[l1 - 1], y21 = yp0[l2 - 1] - yp0[l1 - 1], z21 = zp0[l2 - 1] - zp0[l1 - 1];
    const double x32 = xp0[l3 - 1] - xp0[l2 - 1], y32 = yp0[l3 - 1] - yp0[l2 - 1], z32 = zp0[l3 - 1] - zp0[l2 - 1];
    const double x43 = xp0[l4 - 1] - xp0[l3 - 1], y43 = yp0[l4 - 1] - yp0[l3 - 1], z43 = zp0[l4 - 1] - zp0[l3 - 1];
    const double rsq21 = x21 * x21 + y21 * y21 + z21 * z21;
    const double rsq32 = x32 * x32 + y32 * y32 + z32 * z32;
    const double rsq43 = x43 * x43 + y43 * y43 + z43 * z43;
    const double rsp21 = sqrt(rsq21), rsp32 = sqrt(rsq32), rsp43 = sqrt(rsq43);
    const double cb1 = (x21 * x32 + y21 * y32 + z21 * z32) / (rsp21 * rsp32);
    const double cb2 = (x43 * x32 + y43 * y32 + z43 * z32) / (rsp43 * rsp32);
    const double cb3 = (x21 * x43 + y21 * y43 + z21 * z43) / (rsp21 * rsp43);
    const double sb1 = sqrt(fabs(1.0 - cb1 * cb1)), sb2 = sqrt(fabs(1.0 - cb2 * cb2)),
                 sb3 = sqrt(fabs(1.0 - cb3 * cb3));
    const double aux = sb1 * sb2;
    if (fabs(aux * sb3) < 1.0e-25) { status = 1; continue; }
    double coa = (cb1 * cb2 - cb3) / aux;
    const double pa = pota[i];
    const int tmp = (int)(pa >= 0.0 ? floor(pa + 0.5) : -floor(-pa + 0.5)); /* IDNINT */
    const int ntph = abs(tmp);
    const double sig = (double)(tmp / ntph);
    if (coa > 1.0 && coa < lim) coa = 1.0;
    if (coa < -1.0 && coa > -lim) coa = -1.0;
    double aux1 = cb1 * sb2, aux2 = sb1 * cb2;
    const double auxa = -(aux1 * coa + aux2) / aux;
    const double auxb = -(aux2 * coa + aux1) / aux;
    const double auxc = sb3 / aux;
    aux1 = rsp21 * rsp32 * sb1;
    aux2 = rsp32 * rsp43 * sb2;
    const double aux3 = rsp21 * rsp43 * sb3;
    /* derivatives of the three bond-angle cosines (fptors.f:152-183) */
    double d1[3][4], d2[3][4], d3[3][4]; /* [axis][atom] */
    const double r21[3] = {x21, y21, z21}, r32[3] = {x32, y32, z32}, r43[3] = {x43, y43, z43};
    for (int c = 0; c < 3; ++c) {
      d1[c][0] = r32[c] / aux1 - cb1 * r21[c] / (rsq21 * sb1);
      d1[c][2] = -r21[c] / aux1 + cb1 * r32[c] / (rsq32 * sb1);
      d1[c][1] = -d1[c][2] - d1[c][0];
      d2[c][1] = r43[c] / aux2 - cb2 * r32[c] / (rsq32 * sb2);
      d2[c][3] = -r32[c] / aux2 + cb2 * r43[c] / (rsq43 * sb2);
      d2[c][2] = -d2[c][1] - d2[c][3];
      d3[c][0] = r43[c] / aux3 - cb3 * r21[c] / (rsq21 * sb3);
      d3[c][2] = r21[c] / aux3 - cb3 * r43[c] / (rsq43 * sb3);
      d3[c][1] = -d3[c][0];
      d3[c][3] = -d3[c][2];
    }
    const double b = potb[i];
    double enrg, qforce;
#define P(n) powl_ltr(coa, n)
    switch (ntph) {
      case 1: enrg = sig * (fabs(b) - coa * b); qforce = sig * b; break;
      case 2: enrg = sig * (fabs(b) - (2.0 * P(2) - 1.0) * b); qforce = sig * (4.0 * b * coa); break;
      case 3: enrg = sig * (fabs(b) - (4.0 * P(3) - 3.0 * coa) * b); qforce = sig * ((12.0 * P(2) - 3.0) * b); break;
      case 4:
        enrg = sig * (fabs(b) - (8.0 * P(4) - 8.0 * P(2) + 1.0) * b);
        qforce = sig * ((32.0 * P(3) - 16.0 * coa) * b);
        break;
      case 5:
        enrg = sig * (fabs(b) - (16.0 * P(5) - 20.0 * P(3) + 5.0 * coa) * b);
        qforce = sig * ((80.0 * P(4) - 60.0 * P(2) + 5.0) * b);
        break;
      case 6:
        enrg = sig * (fabs(b) - b * (32.0 * P(6) - 48.0 * P(4) + 18.0 * P(2) - 1.0));
        qforce = sig * (b * (192.0 * P(5) - 192.0 * P(3) + 36.0 * coa));
        break;
      case 7:
        enrg = sig * (fabs(b) - b * (64.0 * P(7) - 112.0 * P(5) + 56.0 * P(3) - 7.0 * coa));
        qforce = sig * (b * (448.0 * P(6) - 560.0 * P(4) + 168.0 * P(2) - 7.0));
        break;
      case 8:
        enrg = sig * (fabs(b) - b * (128.0 * P(8) - 256.0 * P(6) + 160.0 * P(4) - 32.0 * P(2) + 1.0));
        qforce = sig * (b * (1024.0 * P(7) - 1536.0 * P(5) + 640.0 * P(3) - 64.0 * coa));
        break;
      case 12:
      case 14: {
        /* fptors.f:251-290: the sine from the dihedral's orientation */
        const double bb = acos(coa);
        const double xn = y21 * z32 - y32 * z21, yn = z21 * x32 - z32 * x21, zn = x21 * y32 - x32 * y21;
        const double aux4 = fsign(bb, xn * x43 + yn * y43 + zn * z43);
        const double soa = fsign(sqrt(1.0 - pow(coa, 2.0)), aux4);
        const double dsoa = (rem && fabs(soa) < 1.0e-8) ? 0.0 : -coa / soa;
        double auxn, dauxn;
        if (ntph == 14) {
          enrg = b * (1.0 + 0.5 * (8.0 * P(4) - 8.0 * P(2) + 1.0));
          auxn = b * sin60 * 4.0 * (soa * P(3) - coa * powl_ltr(soa, 3));
          dauxn = b * sin60 * 4.0 * (dsoa * (P(3) - 3.0 * coa * powl_ltr(soa, 2)) - powl_ltr(soa, 3) + 3.0 * soa * P(2));
          qforce = 0.5 * (32.0 * P(3) - 16.0 * coa) * b;
        } else {
          enrg = b * (1.0 + 0.5 * (2.0 * P(2) + 1.0));
          auxn = b * sin60 * 2.0 * soa * coa;
          dauxn = b * sin60 * 2.0 * (soa + dsoa * coa);
          qforce = 2.0 * coa * b;
        }
        if (rem) {
          enrg = enrg + auxn;
          qforce = -(qforce + dauxn);
        } else {
          enrg = sig * (enrg + auxn);
          qforce = sig * (-(qforce + dauxn));
        }
        break;
      }
      default: return 2;
    }
#undef P
    utors[type - 1] = utors[type - 1] + enrg;
    if (remltor && remltor[i]) {
      *seg = *seg + enrg;
      qforce = qforce * f2;
    }
    double u[3][4];
    for (int c = 0; c < 3; ++c) {
      u[c][0] = auxa * d1[c][0] + auxc * d3[c][0];
      u[c][1] = auxa * d1[c][1] + auxb * d2[c][1] + auxc * d3[c][1];
      u[c][2] = auxa * d1[c][2] + auxb * d2[c][2] + auxc * d3[c][2];
      u[c][3] = auxb * d2[c][3] + auxc * d3[c][3];
    }
    const int at[4] = {l1, l2, l3, l4};
    for (int k = 0; k < 4; ++k) fpx[at[k] - 1] = fpx[at[k] - 1] + qforce * u[0][k];
    for (int k = 0; k < 4; ++k) fpy[at[k] - 1] = fpy[at[k] - 1] + qforce * u[1][k];
    for (int k = 0; k < 4; ++k) fpz[at[k] - 1] = fpz[at[k] - 1] + qforce * u[2][k];
  }
  *utors_slt = utors[0];
  *utors_slv = utors[1];
  return status;
}

int orc_fptors(const int32_t *ss_index, const int32_t *ltor, int ltors, const double *xp0, const double *yp0,
               const double *zp0, const double *pota, const double *potb, double *utors_slt, double *utors_slv,
               double *fpx, double *fpy, double *fpz) {
  return fptors_impl(ss_index, ltor, ltors, xp0, yp0, zp0, pota, potb, utors_slt, utors_slv, fpx, fpy, fpz, 0, NULL,
                     1.0, NULL);
}

/* fptors_rem.f:1-352 (serial) */
int orc_fptors_rem(const int32_t *ss_index, const int32_t *ltor, int ltors, const double *xp0, const double *yp0,
                   const double *zp0, const double *pota, const double *potb, double *utors_slt, double *utors_slv,
                   double *fpx, double *fpy, double *fpz, const int32_t *remltor, double f2, double *seg_ene) {
  return fptors_impl(ss_index, ltor, ltors, xp0, yp0, zp0, pota, potb, utors_slt, utors_slv, fpx, fpy, fpz, 1,
                     remltor, f2, seg_ene);
}

/* fpbond.f:1-110 (serial): harmonic bonds lbnd(2, lbond), pota the rest
 * length and potb the force constant; forces accumulate, energies by the
 * ss_index of each bond's first atom. */
void orc_fpbond(const int32_t *ss_index, const int32_t *lbnd, int lbond, const double *xp0, const double *yp0,
                const double *zp0, const double *pota, const double *potb, double *ubond_slt, double *ubond_slv,
                double *fpx, double *fpy, double *fpz) {
  double ub[2] = {0.0, 0.0};
  for (int i = 0; i < lbond; ++i) {
    const int la = lbnd[2 * i] - 1, lb = lbnd[2 * i + 1] - 1;
    const int type = ss_index[la];
    const double x21 = xp0[lb] - xp0[la], y21 = yp0[lb] - yp0[la], z21 = zp0[lb] - zp0[la];
    const double rs21 = sqrt(x21 * x21 + y21 * y21 + z21 * z21);
    const double qforce = -2.0 * potb[i] * (pota[i] - rs21) / rs21;
    const double ux = qforce * x21, uy = qforce * y21, uz = qforce * z21;
    fpx[la] = fpx[la] + ux;
    fpy[la] = fpy[la] + uy;
    fpz[la] = fpz[la] + uz;
    fpx[lb] = fpx[lb] - ux;
    fpy[lb] = fpy[lb] - uy;
    fpz[lb] = fpz[lb] - uz;
    const double d = rs21 - pota[i];
    ub[type - 1] = ub[type - 1] + potb[i] * (d * d);
  }
  *ubond_slt = ub[0];
  *ubond_slv = ub[1];
}

/* fpbend.f:1-228 (serial): bends lbndg(3, lbend) with the vertex in the
 * middle; pota the rest angle (rad), potb its constant; the Urey-Bradley 1-3
 * term (potc rest distance, potd constant) when potd is set and pote is not,
 * or potf is; the stretch-stretch coupling (k12, r1, r2) = (potd, potc, pote)
 * or (pote, potf, potg) when those are set.  A rest angle within 0.01 of pi
 * uses 2 potb (cos + 1). */
void orc_fpbend(const int32_t *ss_index, const int32_t *lbndg, int lbend, const double *xp0, const double *yp0,
                const double *zp0, const double *pota, const double *potb, const double *potc, const double *potd,
                const double *pote, const double *potf, const double *potg, double *ubend_slt, double *ubend_slv,
                double *fpx, double *fpy, double *fpz) {
  double ub[2] = {0.0, 0.0};
  const double pi = 4.0 * atan(1.0);
  const double near_pi = (double)0.01f; /* default-real literal (fpbend.f) */
  for (int i = 0; i < lbend; ++i) {
    if (fabs(potb[i]) < 1.0e-12) continue;
    const int la = lbndg[3 * i] - 1, lb = lbndg[3 * i + 1] - 1, lc = lbndg[3 * i + 2] - 1;
    const int type = ss_index[la] - 1;
    const double x12 = xp0[la] - xp0[lb], y12 = yp0[la] - yp0[lb], z12 = zp0[la] - zp0[lb];
    const double x32 = xp0[lc] - xp0[lb], y32 = yp0[lc] - yp0[lb], z32 = zp0[lc] - zp0[lb];
    const double rs12 = x12 * x12 + y12 * y12 + z12 * z12;
    const double rs32 = x32 * x32 + y32 * y32 + z32 * z32;
    const double dcc2 = sqrt(rs12 * rs32);
    double cb = (x12 * x32 + y12 * y32 + z12 * z32) / dcc2;
    if (fabs(cb) > 1.0) cb = cb >= 0.0 ? 1.0 : -1.0;
    const double sb = sqrt(1.0 - cb * cb);
    const double bb = acos(cb);
    const int linear = fabs(pota[i] - pi) < near_pi;
    const double qforce = (linear || sb < 1.0e-6) ? -2.0 * potb[i] : 2.0 * potb[i] * (bb - pota[i]) / sb;
    const double ux1 = x32 / dcc2 - cb * x12 / rs12, ux3 = x12 / dcc2 - cb * x32 / rs32, ux2 = -ux1 - ux3;
    const double uy1 = y32 / dcc2 - cb * y12 / rs12, uy3 = y12 / dcc2 - cb * y32 / rs32, uy2 = -uy1 - uy3;
    const double uz1 = z32 / dcc2 - cb * z12 / rs12, uz3 = z12 / dcc2 - cb * z32 / rs32, uz2 = -uz1 - uz3;
    fpx[la] = fpx[la] + qforce * ux1;
    fpx[lc] = fpx[lc] + qforce * ux3;
    fpx[lb] = fpx[lb] + qforce * ux2;
    fpy[la] = fpy[la] + qforce * uy1;
    fpy[lc] = fpy[lc] + qforce * uy3;
    fpy[lb] = fpy[lb] + qforce * uy2;
    fpz[la] = fpz[la] + qforce * uz1;
    fpz[lc] = fpz[lc] + qforce * uz3;
    fpz[lb] = fpz[lb] + qforce * uz2;
    if (linear) {
      ub[type] = ub[type] + 2.0 * potb[i] * (cb + 1.0);
    } else {
      const double d = bb - pota[i];
      ub[type] = ub[type] + potb[i] * (d * d);
    }
    if (fabs(potd[i]) > 1.0e-5 && (fabs(pote[i]) < 1.0e-5 || fabs(potf[i]) > 1.0e-5)) {
      const double x31 = xp0[lc] - xp0[la], y31 = yp0[lc] - yp0[la], z31 = zp0[lc] - zp0[la];
      const double rsp31 = sqrt(x31 * x31 + y31 * y31 + z31 * z31);
      const double qf = -2.0 * potd[i] * (potc[i] - rsp31);
      const double vx = x31 / rsp31, vy = y31 / rsp31, vz = z31 / rsp31;
      fpx[la] = fpx[la] + qf * vx;
      fpy[la] = fpy[la] + qf * vy;
      fpz[la] = fpz[la] + qf * vz;
      fpx[lc] = fpx[lc] + qf * -vx;
      fpy[lc] = fpy[lc] + qf * -vy;
      fpz[lc] = fpz[lc] + qf * -vz;
      const double d = rsp31 - potc[i];
      ub[type] = ub[type] + potd[i] * (d * d);
    }
    int coupling = 0;
    double k12 = 0.0, r1 = 0.0, r2 = 0.0;
    if (fabs(potd[i]) > 1.0e-5 && fabs(pote[i]) > 1.0e-5 && fabs(potf[i]) < 1.0e-5) {
      coupling = 1;
      k12 = potd[i];
      r1 = potc[i];
      r2 = pote[i];
    }
    if (fabs(pote[i]) > 1.0e-5 && fabs(potf[i]) > 1.0e-5 && fabs(potg[i]) > 1.0e-5) {
      coupling = 1;
      k12 = pote[i];
      r1 = potf[i];
      r2 = potg[i];
    }
    if (coupling) {
      const double rsp12 = sqrt(rs12), rsp32 = sqrt(rs32);
      const double ax = x12 / rsp12, ay = y12 / rsp12, az = z12 / rsp12;
      const double cx = x32 / rsp32, cy = y32 / rsp32, cz = z32 / rsp32;
      const double qf = k12 * (rsp12 - r1), pf = k12 * (rsp32 - r2);
      fpx[la] = fpx[la] - pf * ax;
      fpy[la] = fpy[la] - pf * ay;
      fpz[la] = fpz[la] - pf * az;
      fpx[lb] = fpx[lb] + pf * ax + qf * cx;
      fpy[lb] = fpy[lb] + pf * ay + qf * cy;
      fpz[lb] = fpz[lb] + pf * az + qf * cz;
      fpx[lc] = fpx[lc] - qf * cx;
      fpy[lc] = fpy[lc] - qf * cy;
      fpz[lc] = fpz[lc] - qf * cz;
      ub[type] = ub[type] + k12 * (rsp12 - r1) * (rsp32 - r2);
    }
  }
  *ubend_slt = ub[0];
  *ubend_slv = ub[1];
}

/* fitors.f:1-286 (serial): improper torsions litr(4, litor).  potc > 0: the
 * harmonic form potb (phi - pota)^2 on the signed dihedral phi brought within
 * pi of pota; otherwise the cosine series of multiplicity |pota| (1, 2, 3, 4
 * or 6; another returns 2, where the reference stops).  The cosine-derivative
 * vectors are fptors'. */
/* remitor (fitors_rem.f, null for fitors.f): the flagged terms' energy goes
 * to *seg and their force is scaled by f2 (fitors_rem.f:206-209, :255-261) */
static int fitors_impl(const int32_t *ss_index, const int32_t *litr, int litor, const double *xp0,
                       const double *yp0, const double *zp0, const double *pota, const double *potb,
                       const double *potc, double *utors_slt, double *utors_slv, double *fpx, double *fpy,
                       double *fpz, const int32_t *remitor, double f2, double *seg) {
  const double lim = (double)1.000001f, half_pi4 = (double)1.57079f; /* default-real literals */
  double utors[2] = {0.0, 0.0};
  int status = 0;
  if (seg) *seg = 0.0;
  for (int i = 0; i < litor; ++i) {
    if (fabs(potb[i]) < 1.0e-12) continue;
    const int l[4] = {litr[4 * i] - 1, litr[4 * i + 1] - 1, litr[4 * i + 2] - 1, litr[4 * i + 3] - 1};
    const int type = ss_index[l[0]];
    const double x21 = xp0[l[1]] - xp0[l[0]], y21 = yp0[l[1]] - yp0[l[0]], z21 = zp0[l[1]] - zp0[l[0]];
    const double x32 = xp0[l[2]] - xp0[l[1]], y32 = yp0[l[2]] - yp0[l[1]], z32 = zp0[l[2]] - zp0[l[1]];
    const double x43 = xp0[l[3]] - xp0[l[2]], y43 = yp0[l[3]] - yp0[l[2]], z43 = zp0[l[3]] - zp0[l[2]];
    const double rsq21 = x21 * x21 + y21 * y21 + z21 * z21;
    const double rsq32 = x32 * x32 + y32 * y32 + z32 * z32;
    const double rsq43 = x43 * x43 + y43 * y43 + z43 * z43;
    const double rsp21 = sqrt(rsq21), rsp32 = sqrt(rsq32), rsp43 = sqrt(rsq43);
    const double cb1 = (x21 * x32 + y21 * y32 + z21 * z32) / (rsp21 * rsp32);
    const double cb2 = (x43 * x32 + y43 * y32 + z43 * z32) / (rsp43 * rsp32);
    const double cb3 = (x21 * x43 + y21 * y43 + z21 * z43) / (rsp21 * rsp43);
    const double sb1 = sqrt(1.0 - cb1 * cb1), sb2 = sqrt(1.0 - cb2 * cb2), sb3 = sqrt(1.0 - cb3 * cb3);
    const double aux = sb1 * sb2;
    double coa = (cb1 * cb2 - cb3) / aux;
    if (coa > 1.0 && coa < lim) coa = 1.0;
    if (coa < -1.0 && coa > -lim) coa = -1.0;
    double aux1 = cb1 * sb2, aux2 = sb1 * cb2;
    const double auxa = -(aux1 * coa + aux2) / aux;
    const double auxb = -(aux2 * coa + aux1) / aux;
    const double auxc = sb3 / aux;
    aux1 = rsp21 * rsp32 * sb1;
    aux2 = rsp32 * rsp43 * sb2;
    const double aux3 = rsp21 * rsp43 * sb3;
    double d1[3][4], d2[3][4], d3[3][4];
    const double r21[3] = {x21, y21, z21}, r32[3] = {x32, y32, z32}, r43[3] = {x43, y43, z43};
    for (int c = 0; c < 3; ++c) {
      d1[c][0] = r32[c] / aux1 - cb1 * r21[c] / (rsq21 * sb1);
      d1[c][2] = -r21[c] / aux1 + cb1 * r32[c] / (rsq32 * sb1);
      d1[c][1] = -d1[c][2] - d1[c][0];
      d2[c][1] = r43[c] / aux2 - cb2 * r32[c] / (rsq32 * sb2);
      d2[c][3] = -r32[c] / aux2 + cb2 * r43[c] / (rsq43 * sb2);
      d2[c][2] = -d2[c][1] - d2[c][3];
      d3[c][0] = r43[c] / aux3 - cb3 * r21[c] / (rsq21 * sb3);
      d3[c][2] = r21[c] / aux3 - cb3 * r43[c] / (rsq43 * sb3);
      d3[c][1] = -d3[c][0];
      d3[c][3] = -d3[c][2];
    }
    const double a = pota[i], b = potb[i];
    double qforce;
    if (potc[i] > 0.0) {
      double bb = acos(coa);
      const double xn = y21 * z32 - y32 * z21, yn = z21 * x32 - z32 * x21, zn = x21 * y32 - x32 * y21;
      const double aux4 = (xn * x43 + yn * y43 + zn * z43) >= 0.0 ? fabs(bb) : -fabs(bb);
      if (aux4 < 0.0 && fabs(aux4) >= half_pi4) {
        bb = 2.0 * acos(-1.0) + aux4;
        if (fabs(a - bb) > 3.14159) bb = -2 * 3.14159 + bb;
      } else if (aux4 < 0.0 && fabs(aux4) < half_pi4) {
        bb = aux4;
        if (fabs(a - bb) > 3.14159) bb = 2 * 3.14159 + bb;
      } else {
        bb = aux4;
        if (fabs(a - bb) > 3.14159) bb = 2 * 3.14159 + bb;
        if (fabs(a - bb) > 3.14159) bb = -4 * 3.14159 + bb;
      }
      if (fabs(a) < 1.0e-3)
        qforce = 2.0 * b;
      else if (fabs(a - 3.14159) < 1.0e-3)
        qforce = -2.0 * b;
      else
        qforce = (2.0 * b * (bb - a)) / sin(bb);
      const double da = bb - a;
      utors[type - 1] = utors[type - 1] + b * (da * da);
      if (remitor && remitor[i]) {
        *seg = *seg + b * (da * da);
        qforce = qforce * f2;
      }
    } else {
      const int tmp = (int)(a >= 0.0 ? floor(a + 0.5) : -floor(-a + 0.5)); /* IDNINT */
      const int ntph = abs(tmp);
      if (ntph == 0) { status = 2; continue; }
      const double sig = (double)(tmp / ntph);
      double enrg;
      switch (ntph) {
        case 1: enrg = sig * (fabs(b) - coa * b); qforce = sig * b; break;
        case 2: enrg = sig * (fabs(b) - (2.0 * powl_ltr(coa, 2) - 1.0) * b); qforce = sig * (4.0 * b * coa); break;
        case 3:
          enrg = sig * (fabs(b) - (4.0 * powl_ltr(coa, 3) - 3.0 * coa) * b);
          qforce = sig * ((12.0 * powl_ltr(coa, 2) - 3.0) * b);
          break;
        case 4:
          enrg = sig * (fabs(b) - (8.0 * powl_ltr(coa, 4) - 8.0 * powl_ltr(coa, 2) + 1.0) * b);
          qforce = sig * ((32.0 * powl_ltr(coa, 3) - 16.0 * coa) * b);
          break;
        case 6:
          enrg = sig * (fabs(b) - b * (32.0 * powl_ltr(coa, 6) - 48.0 * powl_ltr(coa, 4) + 18.0 * powl_ltr(coa, 2) - 1.0));
          qforce = sig * (b * (192.0 * powl_ltr(coa, 5) - 192.0 * powl_ltr(coa, 3) + 36.0 * coa));
          break;
        default: status = 2; continue;
      }
      utors[type - 1] = utors[type - 1] + enrg;
      if (remitor && remitor[i]) {
        *seg = *seg + enrg;
        qforce = qforce * f2;
      }
    }
    for (int c = 0; c < 3; ++c) {
      double *f = c == 0 ? fpx : (c == 1 ? fpy : fpz);
      const double u1 = auxa * d1[c][0] + auxc * d3[c][0];
      const double u2 = auxa * d1[c][1] + auxb * d2[c][1] + auxc * d3[c][1];
      const double u3 = auxa * d1[c][2] + auxb * d2[c][2] + auxc * d3[c][2];
      const double u4 = auxb * d2[c][3] + auxc * d3[c][3];
      f[l[0]] = f[l[0]] + qforce * u1;
      f[l[1]] = f[l[1]] + qforce * u2;
      f[l[2]] = f[l[2]] + qforce * u3;
      f[l[3]] = f[l[3]] + qforce * u4;
    }
  }
  *utors_slt = utors[0];
  *utors_slv = utors[1];
  return status;
}

int orc_fitors(const int32_t *ss_index, const int32_t *litr, int litor, const double *xp0, const double *yp0,
               const double *zp0, const double *pota, const double *potb, const double *potc, double *utors_slt,
               double *utors_slv, double *fpx, double *fpy, double *fpz) {
  return fitors_impl(ss_index, litr, litor, xp0, yp0, zp0, pota, potb, potc, utors_slt, utors_slv, fpx, fpy, fpz,
                     NULL, 1.0, NULL);
}

/* fitors_rem.f:1-301 (serial) */
int orc_fitors_rem(const int32_t *ss_index, const int32_t *litr, int litor, const double *xp0, const double *yp0,
                   const double *zp0, const double *pota, const double *potb, const double *potc, double *utors_slt,
                   double *utors_slv, double *fpx, double *fpy, double *fpz, const int32_t *remitor, double f2,
                   double *seg_ene) {
  return fitors_impl(ss_index, litr, litor, xp0, yp0, zp0, pota, potb, potc, utors_slt, utors_slv, fpx, fpy, fpz,
                     remitor, f2, seg_ene);
}

/* ------------------------------------------------------------------ */
/* Intra-alchemical reciprocal correction and alchemical 1-4 pairs      */
/* ------------------------------------------------------------------ */

/* ferrf_alchemy_add.f:1-204, serial build.  list(3, nlist): atoms ia, ib
 * (1-based) and facQ (1: both added or both removed, 0: one of each).
 * factl = (1 - |lambdaq_a|)(1 - |lambdaq_b|) - facQ scales the created
 * charges; with intramol only the facQ = 0 pairs are kept (:105, :175).
 * erf_corr selects the spline branch (:94-141).  dwrk = slt + slv - en0. */
int orc_ferrf_alchemy_add(const int32_t *ss_index, double alphal, const double *charge, const double *lambdaq,
                          const double *lambdaq0, const double *x0, const double *y0, const double *z0,
                          const int32_t *list, int nlist, double *fs_slt, double *fs_slv, double *fpx, double *fpy,
                          double *fpz, int32_t erf_corr, const double *erf_arr_corr, double delew, double rlew,
                          double *dwrk, int intramol) {
  const double twrtpi = 2.0 / sqrt(4.0 * atan(1.0));
  double fs[2] = {0.0, 0.0}, en0 = 0.0;
  for (int i = 0; i < nlist; ++i) {
    const int ia = list[3 * i], ib = list[3 * i + 1];
    const int type = ss_index[ia - 1];
    const double facq = (double)list[3 * i + 2];
    if (intramol && list[3 * i + 2] != 0) continue;
    const double factl = (1.0 - fabs(lambdaq[ia - 1])) * (1.0 - fabs(lambdaq[ib - 1])) - facq;
    const double factl0 = (1.0 - fabs(lambdaq0[ia - 1])) * (1.0 - fabs(lambdaq0[ib - 1])) - facq;
    const double qq = factl * charge[ia - 1] * charge[ib - 1];
    const double qq0 = factl0 * charge[ia - 1] * charge[ib - 1];
    const double xab = x0[ia - 1] - x0[ib - 1], yab = y0[ia - 1] - y0[ib - 1], zab = z0[ia - 1] - z0[ib - 1];
    const double rsq = xab * xab + yab * yab + zab * zab;
    const double rsp = sqrt(rsq);
    double qforce;
    if (erf_corr) {
      const int ic = (int)(1 + (rsp - rlew) / delew);
      const double h = rsp - (rlew + (ic - 1) * delew);
      const double *c = erf_arr_corr + 4 * (ic - 1);
      const double corr = c[0] + h * (c[1] + h * (c[2] + h * c[3] / 3) / 2);
      const double dcorr = (c[1] + h * (c[2] + h * c[3] / 2)) / rsp;
      const double alphar = alphal * rsp;
      const double qt = 1.0 / (1.0 + QP * alphar);
      const double expcst = exp(-alphar * alphar);
      const double erfcst = ((((A5 * qt + A4) * qt + A3) * qt + A2) * qt + A1) * qt * expcst;
      const double erfst = 1.0 - erfcst;
      qforce = -qq * (dcorr + (erfst - twrtpi * alphar * expcst) / (rsp * rsq));
      fs[type - 1] = fs[type - 1] - qq * (erfst / rsp - corr);
      en0 = en0 - qq0 * (erfst / rsp - corr);
    } else {
      const double alphar = alphal * rsp;
      const double qt = 1.0 / (1.0 + QP * alphar);
      const double expcst = exp(-alphar * alphar);
      const double erfcst = ((((A5 * qt + A4) * qt + A3) * qt + A2) * qt + A1) * qt * expcst;
      const double erfst = 1.0 - erfcst;
      qforce = -qq * (erfst - twrtpi * alphar * expcst) / (rsp * rsq);
      fs[type - 1] = fs[type - 1] - qq * erfst / rsp;
      en0 = en0 - qq0 * erfst / rsp;
    }
    fpx[ia - 1] = fpx[ia - 1] + qforce * xab;
    fpy[ia - 1] = fpy[ia - 1] + qforce * yab;
    fpz[ia - 1] = fpz[ia - 1] + qforce * zab;
    fpx[ib - 1] = fpx[ib - 1] - qforce * xab;
    fpy[ib - 1] = fpy[ib - 1] - qforce * yab;
    fpz[ib - 1] = fpz[ib - 1] - qforce * zab;
  }
  *fs_slt = fs[0];
  *fs_slv = fs[1];
  *dwrk = fs[0] + fs[1] - en0;
  return 0;
}

/* The scaling factors of one alchemical 1-4 pair (fnb14_alchemy.f:131-155 for
 * the LJ lambdas, :161-185 and :214-255 for the charge lambdas): normal pairs
 * and pairs of equal nonzero lambda keep the full potential, a pair with one
 * alchemical atom is scaled by 1 - |lambda| of that atom, and a pair of an
 * added and a removed atom is switched off. */
static void fnb14_alch_fact(double li, double lj, double li0, double lj0, double *f, double *f0) {
  if (li == 0.0 && lj != 0.0) {
    *f = 1.0 - fabs(lj);
    *f0 = 1.0 - fabs(lj0);
  } else if (li != 0.0 && lj == 0.0) {
    *f = 1.0 - fabs(li);
    *f0 = 1.0 - fabs(li0);
  } else if (li == 0.0 && lj == 0.0) {
    *f = 1.0;
    *f0 = 1.0;
  } else if (li != lj) {
    *f = 0.0;
    *f0 = 0.0;
  } else {
    *f = 1.0;
    *f0 = 1.0;
  }
}

/* fnb14_alchemy.f:1-361, serial build (dwrk zeroed on entry, :117).  The
 * reference's own arithmetic is kept where it departs from the plain
 * fnb14: without Ewald the LJ and Coulomb forces are unscaled and the LJ
 * lambda0 energy uses alch_fact (:203-210); with Ewald only the repulsive
 * LJ force term is scaled by alch_fact (:286), and the pairs of two
 * alchemical atoms add the erf term with alch_factq2 (:239-255, :305-311). */
void orc_fnb14_alchemy(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
                       const double *charge, int nato, const double *ecc12, const double *ecc6, int cut, int ewald,
                       double alphal, const int32_t *int14, int int14p, const int32_t *type14, double fudge,
                       double lj_fudge, double *fpx, double *fpy, double *fpz, double *uconf_slt, double *uconf_slv,
                       double *ucoul_slt, double *ucoul_slv, const double *lambda, const double *lambdaq,
                       const double *lambda0, const double *lambdaq0, double *dwrk) {
  (void)nato;
  double uconf[2] = {0.0, 0.0}, ucoul[2] = {0.0, 0.0}, uconf0[2] = {0.0, 0.0}, ucoul0[2] = {0.0, 0.0};
  *dwrk = 0.0;
  *uconf_slt = *uconf_slv = *ucoul_slt = *ucoul_slv = 0.0;
  if (int14p == 0) return;
  const double twrtpi = 2.0 / sqrt(4.0 * atan(1.0));
  for (int i = 0; i < int14p; ++i) {
    const int i1 = int14[2 * i], i2 = int14[2 * i + 1];
    const int a = i1 - 1, b = i2 - 1;
    double af, af0, aq, aq0;
    fnb14_alch_fact(lambda[a], lambda[b], lambda0[a], lambda0[b], &af, &af0);
    fnb14_alch_fact(lambdaq[a], lambdaq[b], lambdaq0[a], lambdaq0[b], &aq, &aq0);
    const int type = ss_index[a];
    const int lij = type14[i];
    const double chrgei = charge[a], chrgej = charge[b];
    const double xa = xp0[a] - xp0[b], ya = yp0[a] - yp0[b], za = zp0[a] - zp0[b];
    const double rsq = xa * xa + ya * ya + za * za;
    const double rsp = sqrt(rsq);
    const double rsqi = 1.0 / rsq;
    const double r6 = rsqi * rsqi * rsqi, r12 = r6 * r6;
    double qforce;
    if (!ewald || !cut) {
      const double ssvir = 12.0 * ecc12[lij - 1] * r12 - 6.0 * ecc6[lij - 1] * r6;
      qforce = lj_fudge * ssvir * rsqi;
      const double elj = lj_fudge * (ecc12[lij - 1] * r12 - ecc6[lij - 1] * r6);
      uconf[type - 1] = uconf[type - 1] + af * elj;
      uconf0[type - 1] = uconf0[type - 1] + af * elj;
      const double sv = fudge * chrgei * chrgej / rsp;
      ucoul[type - 1] = ucoul[type - 1] + aq * sv;
      ucoul0[type - 1] = ucoul0[type - 1] + aq0 * sv;
      qforce = qforce + sv * rsqi;
    } else {
      double aq2 = 0.0, aq20 = 0.0;
      const double lqa = lambdaq[a], lqb = lambdaq[b];
      if (lqa != 0.0 && lqb != 0.0) {
        const double p = (1.0 - fabs(lqa)) * (1.0 - fabs(lqb));
        const double p0 = (1.0 - fabs(lambdaq0[a])) * (1.0 - fabs(lambdaq0[b]));
        if (lqa != lqb) {
          aq2 = -p;
          aq20 = -p0;
        } else {
          aq2 = 1.0 - p;
          aq20 = 1.0 - p0;
        }
      }
      const double rspqi = rsqi / rsp;
      const double ssvir = af * 12.0 * ecc12[lij - 1] * r12 - 6.0 * ecc6[lij - 1] * r6;
      qforce = lj_fudge * ssvir * rsqi;
      double aux = lj_fudge * (ecc12[lij - 1] * r12 - ecc6[lij - 1] * r6);
      uconf[type - 1] = uconf[type - 1] + af * aux;
      uconf0[type - 1] = uconf0[type - 1] + af0 * aux;
      const double alphar = alphal * rsp;
      const double qt = 1.0 / (1.0 + QP * alphar);
      const double expcst = exp(-alphar * alphar);
      const double erfcst = ((((A5 * qt + A4) * qt + A3) * qt + A2) * qt + A1) * qt * expcst;
      const double furpar = aq * fudge * chrgei * chrgej;
      const double furpar0 = aq0 * fudge * chrgei * chrgej;
      ucoul[type - 1] = ucoul[type - 1] + furpar * erfcst / rsp;
      ucoul0[type - 1] = ucoul0[type - 1] + furpar0 * erfcst / rsp;
      const double erff = 1.0 - erfcst;
      const double furpar2 = aq2 * fudge * chrgei * chrgej;
      ucoul[type - 1] = ucoul[type - 1] + furpar2 * erff / rsp;
      const double furpar20 = aq20 * fudge * chrgei * chrgej;
      ucoul0[type - 1] = ucoul0[type - 1] + furpar20 * erff / rsp;
      aux = (erfcst + twrtpi * alphar * expcst) * rspqi;
      const double aux2 = (erff - twrtpi * alphar * expcst) * rspqi;
      qforce = qforce + furpar * aux + furpar2 * aux2;
    }
    fpx[a] = fpx[a] + qforce * xa;
    fpy[a] = fpy[a] + qforce * ya;
    fpz[a] = fpz[a] + qforce * za;
    fpx[b] = fpx[b] - qforce * xa;
    fpy[b] = fpy[b] - qforce * ya;
    fpz[b] = fpz[b] - qforce * za;
  }
  *uconf_slt = uconf[0];
  *uconf_slv = uconf[1];
  *ucoul_slt = ucoul[0];
  *ucoul_slv = ucoul[1];
  double dw = 0.0;
  for (int k = 0; k < 2; ++k) dw = dw + uconf[k] + ucoul[k] - (uconf0[k] + ucoul0[k]);
  *dwrk = dw;
}

/* ------------------------------------------------------------------ */
/* Constraints: SHAKE / RATTLE as ORAC applies them in its integrator. */
/* rattle_verlet.f (positions after a drift) and rattle_correc.f       */
/* (velocities after a kick).  Molecules of more than mim_lim          */
/* constraints iterate bond by bond; the others take the matrix-       */
/* inversion method (MIM), closed-form for up to 3 constraints (the    */
/* reference calls LAPACK dgesv beyond that; this restatement returns  */
/* 2 there).  cnst(2, ncnstr) is 1-based; cnst_protl holds, molecule   */
/* by molecule, the count of its constraints followed by their 1-based */
/* indices.                                                            */
/* ------------------------------------------------------------------ */

typedef struct {
  double b11, b12, b13, b21, b22, b23, b31, b32, b33, det;
} mim_inv;

/* rattle_verlet.f:366-397 (cnstp = 1, 2, 3 forms of the first solve) */
static void mim_first_solve(int cnstp, double mat[3][3], mim_inv *r, double *gam) {
  if (cnstp == 1) {
    gam[0] = gam[0] / mat[0][0];
  } else if (cnstp == 2) {
    r->b11 = mat[1][1];
    r->b12 = -mat[1][0];
    r->b21 = -mat[0][1];
    r->b22 = mat[0][0];
    r->det = r->b11 * r->b22 - r->b21 * r->b12;
    r->det = 1.0 / r->det;
    double xd = r->b11 * gam[0] + r->b21 * gam[1];
    double yd = r->b12 * gam[0] + r->b22 * gam[1];
    gam[0] = xd * r->det;
    gam[1] = yd * r->det;
  } else {
    double a11 = mat[0][0], a12 = mat[0][1], a13 = mat[0][2], a21 = mat[1][0], a22 = mat[1][1], a23 = mat[1][2],
           a31 = mat[2][0], a32 = mat[2][1], a33 = mat[2][2];
    r->b11 = a22 * a33 - a32 * a23;
    r->b12 = -a21 * a33 + a31 * a23;
    r->b13 = a21 * a32 - a22 * a31;
    r->b21 = -a12 * a33 + a32 * a13;
    r->b22 = a11 * a33 - a31 * a13;
    r->b23 = -a11 * a32 + a31 * a12;
    r->b31 = a12 * a23 - a22 * a13;
    r->b32 = -a11 * a23 + a21 * a13;
    r->b33 = a11 * a22 - a21 * a12;
    r->det = a11 * r->b11 + a21 * r->b21 + a31 * r->b31;
    r->det = 1.0 / r->det;
    double xd = r->b11 * gam[0] + r->b21 * gam[1] + r->b31 * gam[2];
    double yd = r->b12 * gam[0] + r->b22 * gam[1] + r->b32 * gam[2];
    double zd = r->b13 * gam[0] + r->b23 * gam[1] + r->b33 * gam[2];
    gam[0] = xd * r->det;
    gam[1] = yd * r->det;
    gam[2] = zd * r->det;
  }
}

/* rattle_verlet.f:430-444: later solves reuse the cofactors of the first */
static void mim_resolve(int cnstp, double mat[3][3], const mim_inv *r, double *gam) {
  if (cnstp == 1) {
    gam[0] = gam[0] / mat[0][0];
  } else if (cnstp == 2) {
    double xd = r->b11 * gam[0] + r->b21 * gam[1];
    double yd = r->b12 * gam[0] + r->b22 * gam[1];
    gam[0] = xd * r->det;
    gam[1] = yd * r->det;
  } else {
    double xd = r->b11 * gam[0] + r->b21 * gam[1] + r->b31 * gam[2];
    double yd = r->b12 * gam[0] + r->b22 * gam[1] + r->b32 * gam[2];
    double zd = r->b13 * gam[0] + r->b23 * gam[1] + r->b33 * gam[2];
    gam[0] = xd * r->det;
    gam[1] = yd * r->det;
    gam[2] = zd * r->det;
  }
}

/* direction of atom `at`'s move for constraint (lab, lbb) with bond vector d:
   rattle_verlet.f:286-303 */
static void mim_dir(int at, int lab, int lbb, const double *d, double *o) {
  if (at == lab) {
    o[0] = -d[0];
    o[1] = -d[1];
    o[2] = -d[2];
  } else if (at == lbb) {
    o[0] = d[0];
    o[1] = d[1];
    o[2] = d[2];
  } else {
    o[0] = o[1] = o[2] = 0.0;
  }
}

/* The MIM forms store the atom directions (matx) only for constraint pairs
   that share an atom (rattle_verlet.f:277-331) but read them back two per
   pair for every pair (:408-424, :514-541); a molecule with two constraints
   on disjoint atoms would read entries that were never written.  Such
   molecules are refused (return 3); triangles (rigid 3-site waters), pairs of
   bonds on a common atom and single constraints are what the path uses. */
static int mim_pairs_share(int cnstp, const int32_t *kl, const int32_t *cnst) {
  for (int ka = 0; ka < cnstp; ++ka)
    for (int kb = 0; kb < cnstp; ++kb) {
      int k1 = kl[ka] - 1, k2 = kl[kb] - 1;
      int la = cnst[2 * k1], lb = cnst[2 * k1 + 1], lab = cnst[2 * k2], lbb = cnst[2 * k2 + 1];
      if (la != lab && la != lbb && lb != lab && lb != lbb) return 0;
    }
  return 1;
}

int orc_rattle_verlet(double dt, double *xp1, double *yp1, double *zp1, const double *xp0, const double *yp0,
                      const double *zp0, double *vpx, double *vpy, double *vpz, int nato, const int32_t *cnst,
                      const double *dss, const double *coeff, int ncnstr, const double *mass, int nprot,
                      const int32_t *cnst_protl, int mim_lim, double *dnit) {
  const double tol = 1.0e-8, tol_mim = 1.0e-7; /* rattle_verlet.f:77 */
  if (ncnstr == 0) return 0;
  if (mim_lim > 20) return 1; /* :87-92 (nmat = 20) */
  double dti = dt != 0.0 ? 1.0 / dt : 0.0; /* :111-115 */
  double *mass0 = malloc(sizeof(double) * (size_t)nato);
  char *mask = malloc((size_t)nato);
  double *x0 = malloc(sizeof(double) * 3 * (size_t)ncnstr);
  for (int i = 0; i < nato; ++i) { /* :117-120 */
    mass0[i] = 1.0 / mass[i];
    mask[i] = 1;
  }
  for (int k = 0; k < ncnstr; ++k) { /* :127-136 old bond vectors */
    int la = cnst[2 * k] - 1, lb = cnst[2 * k + 1] - 1;
    x0[3 * k] = xp0[la] - xp0[lb];
    x0[3 * k + 1] = yp0[la] - yp0[lb];
    x0[3 * k + 2] = zp0[la] - zp0[lb];
  }
  int rc = 0, iter1 = 0, count = 0;
  for (int i = 0; i < nprot && rc == 0; ++i) {
    const int cnstp = cnst_protl[count];
    const int32_t *kl = cnst_protl + count + 1;
    if (cnstp > mim_lim) {
      /* bond-by-bond SHAKE, :157-220 */
      int iter = 0;
      for (;;) {
        int iox = 0;
        for (int ka = 0; ka < cnstp; ++ka) {
          int k = kl[ka] - 1, la = cnst[2 * k] - 1, lb = cnst[2 * k + 1] - 1;
          double dpx = xp1[la] - xp1[lb], dpy = yp1[la] - yp1[lb], dpz = zp1[la] - zp1[lb];
          double dpp = dpx * dpx + dpy * dpy + dpz * dpz;
          double dcnt = dpp - dss[k];
          if (fabs(dcnt) >= tol) {
            iox = 1;
            double xk = x0[3 * k], yk = x0[3 * k + 1], zk = x0[3 * k + 2];
            double dps = dpx * xk + dpy * yk + dpz * zk;
            double gg = dcnt / (dps * coeff[k]);
            double gg1 = -gg * mass0[la], gg2 = gg * mass0[lb];
            xp1[la] = xp1[la] + xk * gg1;
            xp1[lb] = xp1[lb] + xk * gg2;
            yp1[la] = yp1[la] + yk * gg1;
            yp1[lb] = yp1[lb] + yk * gg2;
            zp1[la] = zp1[la] + zk * gg1;
            zp1[lb] = zp1[lb] + zk * gg2;
            gg1 = gg1 * dti;
            gg2 = gg2 * dti;
            vpx[la] = vpx[la] + xk * gg1;
            vpx[lb] = vpx[lb] + xk * gg2;
            vpy[la] = vpy[la] + yk * gg1;
            vpy[lb] = vpy[lb] + yk * gg2;
            vpz[la] = vpz[la] + zk * gg1;
            vpz[lb] = vpz[lb] + zk * gg2;
          }
        }
        if (!iox) break;
        iter1 += 1;
        iter += 1;
        if (iter > 5000) { /* the reference stops the program here */
          rc = 1;
          break;
        }
      }
      /* The reference advances its cursor into cnst_protl only in the MIM
         branch (:562): after a molecule of more than mim_lim constraints the
         next molecule reads this one's entry again.  Kept. */
      continue;
    } else if (cnstp > 3) {
      rc = 2; /* dgesv / dgetrs in the reference (:398, :446) */
    } else if (!mim_pairs_share(cnstp, kl, cnst)) {
      rc = 3; /* see mim_pairs_share */
    } else {
      /* matrix-inversion method, :228-560 */
      int redo, nredo = 0;
      do {
        redo = 0;
        double gam[3] = {0, 0, 0}, gamo[3], dd[3], xc[3][3], mat[3][3], matx[18][3];
        int la = 0, lb = 0;
        for (int ka = 0; ka < cnstp; ++ka) { /* :238-251 */
          int k1 = kl[ka] - 1, a = cnst[2 * k1] - 1, b = cnst[2 * k1 + 1] - 1;
          xc[ka][0] = xp1[a] - xp1[b];
          xc[ka][1] = yp1[a] - yp1[b];
          xc[ka][2] = zp1[a] - zp1[b];
        }
        int count1 = 0;
        for (int ka = 0; ka < cnstp; ++ka) { /* :257-336 */
          int k1 = kl[ka] - 1;
          la = cnst[2 * k1] - 1;
          lb = cnst[2 * k1 + 1] - 1;
          dd[ka] = dss[k1] - (xc[ka][0] * xc[ka][0] + xc[ka][1] * xc[ka][1] + xc[ka][2] * xc[ka][2]);
          for (int kb = 0; kb < cnstp; ++kb) {
            int k2 = kl[kb] - 1, lab = cnst[2 * k2] - 1, lbb = cnst[2 * k2 + 1] - 1;
            if (la != lab && la != lbb && lb != lab && lb != lbb) {
              mat[ka][kb] = 0.0;
            } else {
              double dpa[3], dpb[3];
              mim_dir(la, lab, lbb, x0 + 3 * k2, dpa);
              mim_dir(lb, lab, lbb, x0 + 3 * k2, dpb);
              double dpx = -dpa[0], dpy = -dpa[1], dpz = -dpa[2];
              matx[count1][0] = dpa[0];
              matx[count1][1] = dpa[1];
              matx[count1][2] = dpa[2];
              count1++;
              double aux3 = (dpx * xc[ka][0] + dpy * xc[ka][1] + dpz * xc[ka][2]) * mass0[la];
              matx[count1][0] = dpb[0];
              matx[count1][1] = dpb[1];
              matx[count1][2] = dpb[2];
              count1++;
              aux3 = aux3 + (dpb[0] * xc[ka][0] + dpb[1] * xc[ka][1] + dpb[2] * xc[ka][2]) * mass0[lb];
              aux3 = aux3 * 4.0;
              mat[ka][kb] = -aux3;
            }
          }
        }
        /* Non-linear iteration, :342-508.  The quadratic term weights every
           constraint's displacement with the masses of the LAST constraint's
           atoms (la, lb as the loop above leaves them, :408-409); kept. */
        mim_inv inv = {0};
        int iox = 1, iter = 0;
        while (iox == 1) {
          for (int ka = 0; ka < cnstp; ++ka) gamo[ka] = gam[ka];
          if (iter == 0) {
            for (int ka = 0; ka < cnstp; ++ka) gam[ka] = dd[ka];
            mim_first_solve(cnstp, mat, &inv, gam);
          } else {
            double rhs[3];
            int c1 = 0;
            for (int ka = 0; ka < cnstp; ++ka) {
              double xd = 0.0, yd = 0.0, zd = 0.0;
              double gg1 = mass0[la] * 2.0, gg2 = mass0[lb] * 2.0;
              for (int kb = 0; kb < cnstp; ++kb) {
                double aux = gam[kb] * gg1;
                xd = xd + matx[c1][0] * aux;
                yd = yd + matx[c1][1] * aux;
                zd = zd + matx[c1][2] * aux;
                c1++;
                double aux1 = -matx[c1][0], aux2 = -matx[c1][1], aux3 = -matx[c1][2];
                aux = gam[kb] * gg2;
                xd = xd + aux1 * aux;
                yd = yd + aux2 * aux;
                zd = zd + aux3 * aux;
                c1++;
              }
              rhs[ka] = dd[ka] - (xd * xd + yd * yd + zd * zd);
            }
            for (int ka = 0; ka < cnstp; ++ka) gam[ka] = rhs[ka];
            mim_resolve(cnstp, mat, &inv, gam);
          }
          iter = iter + 1;
          if (iter > 120) redo = 1; /* :461-496: correct with what there is, then start over */
          iox = 0;
          for (int ka = 0; ka < cnstp; ++ka)
            if (gam[ka] != 0.0) {
              double aux = fabs((gam[ka] - gamo[ka]) / gam[ka]);
              if (aux > tol_mim) iox = 1;
            }
          if (redo) iox = 0;
        }
        iter1 += iter;
        /* corrected positions and velocities, :514-558 */
        count1 = 0;
        for (int ka = 0; ka < cnstp; ++ka) {
          int k1 = kl[ka] - 1, a = cnst[2 * k1] - 1, b = cnst[2 * k1 + 1] - 1;
          double xd = mass0[a] * 2.0, yd = mass0[b] * 2.0;
          for (int kb = 0; kb < cnstp; ++kb) {
            if (mask[a]) {
              double aux = gam[kb] * xd;
              xp1[a] = xp1[a] + matx[count1][0] * aux;
              yp1[a] = yp1[a] + matx[count1][1] * aux;
              zp1[a] = zp1[a] + matx[count1][2] * aux;
              aux = aux * dti;
              vpx[a] = vpx[a] + matx[count1][0] * aux;
              vpy[a] = vpy[a] + matx[count1][1] * aux;
              vpz[a] = vpz[a] + matx[count1][2] * aux;
            }
            count1++;
            if (mask[b]) {
              double aux = gam[kb] * yd;
              xp1[b] = xp1[b] + matx[count1][0] * aux;
              yp1[b] = yp1[b] + matx[count1][1] * aux;
              zp1[b] = zp1[b] + matx[count1][2] * aux;
              aux = aux * dti;
              vpx[b] = vpx[b] + matx[count1][0] * aux;
              vpy[b] = vpy[b] + matx[count1][1] * aux;
              vpz[b] = vpz[b] + matx[count1][2] * aux;
            }
            count1++;
          }
          mask[a] = 0;
          mask[b] = 0;
        }
        if (redo && ++nredo > 1000) { /* the reference starts over without bound */
          rc = 1;
          break;
        }
        if (redo) /* :546-557 */
          for (int ka = 0; ka < cnstp; ++ka) {
            int k1 = kl[ka] - 1;
            mask[cnst[2 * k1] - 1] = 1;
            mask[cnst[2 * k1 + 1] - 1] = 1;
          }
      } while (redo);
    }
    count = count + cnstp + 1;
  }
  if (dnit) *dnit = *dnit + (double)iter1 / (double)nprot;
  free(mass0);
  free(mask);
  free(x0);
  return rc;
}

int orc_rattle_correc(double dt, const double *xp0, const double *yp0, const double *zp0, double *vpx, double *vpy,
                      double *vpz, int nato, const int32_t *cnst, const double *dss, const double *coeff, int ncnstr,
                      const double *mass, int nprot, const int32_t *cnst_protl, int mim_lim, double *dnit) {
  (void)dt;
  const double tolh = 1.0e-6; /* rattle_correc.f:75,91 */
  double *mass0 = malloc(sizeof(double) * (size_t)(nato > 0 ? nato : 1));
  char *mask = malloc((size_t)(nato > 0 ? nato : 1));
  double *bv = malloc(sizeof(double) * 3 * (size_t)(ncnstr > 0 ? ncnstr : 1));
  double *coeffi = malloc(sizeof(double) * (size_t)(ncnstr > 0 ? ncnstr : 1));
  double *dssi = malloc(sizeof(double) * (size_t)(ncnstr > 0 ? ncnstr : 1));
  for (int i = 0; i < nato; ++i) { /* :92-95 */
    mass0[i] = 1.0 / mass[i];
    mask[i] = 1;
  }
  for (int k = 0; k < ncnstr; ++k) { /* :96-106 bond vectors at the current positions */
    int la = cnst[2 * k] - 1, lb = cnst[2 * k + 1] - 1;
    bv[3 * k] = xp0[la] - xp0[lb];
    bv[3 * k + 1] = yp0[la] - yp0[lb];
    bv[3 * k + 2] = zp0[la] - zp0[lb];
    coeffi[k] = 1.0 / coeff[k];
    dssi[k] = 1.0 / dss[k];
  }
  int rc = 0, iter1 = 0, count = 0;
  for (int i = 0; i < nprot && rc == 0; ++i) {
    const int cnstp = cnst_protl[count];
    const int32_t *kl = cnst_protl + count + 1;
    if (cnstp > mim_lim) {
      /* bond-by-bond RATTLE, :125-172 */
      int iter = 0;
      for (;;) {
        int iox = 0;
        for (int ka = 0; ka < cnstp; ++ka) {
          int k = kl[ka] - 1, la = cnst[2 * k] - 1, lb = cnst[2 * k + 1] - 1;
          double dpx = bv[3 * k], dpy = bv[3 * k + 1], dpz = bv[3 * k + 2];
          double vdpx = vpx[la] - vpx[lb], vdpy = vpy[la] - vpy[lb], vdpz = vpz[la] - vpz[lb];
          double dpp = dpx * vdpx + dpy * vdpy + dpz * vdpz;
          if (fabs(dpp) > tolh) {
            double gab = -dpp * dssi[k] * coeffi[k];
            iox = 1;
            double gg1 = mass0[la] * gab, gg2 = -mass0[lb] * gab;
            vpx[la] = vpx[la] + gg1 * dpx;
            vpx[lb] = vpx[lb] + gg2 * dpx;
            vpy[la] = vpy[la] + gg1 * dpy;
            vpy[lb] = vpy[lb] + gg2 * dpy;
            vpz[la] = vpz[la] + gg1 * dpz;
            vpz[lb] = vpz[lb] + gg2 * dpz;
          }
        }
        if (!iox) break;
        iter1 += 1;
        iter += 1;
        if (iter > 5000) {
          rc = 1;
          break;
        }
      }
    } else if (cnstp > 3) {
      rc = 2; /* dgesv in the reference (:339) */
    } else if (!mim_pairs_share(cnstp, kl, cnst)) {
      rc = 3;
    } else {
      /* matrix-inversion method, :176-385 (a linear system: no iteration) */
      double xc[3][3], dd[3], mat[3][3], matx[18][3], gam[3];
      for (int ka = 0; ka < cnstp; ++ka) { /* :179-190 relative velocities */
        int k1 = kl[ka] - 1, a = cnst[2 * k1] - 1, b = cnst[2 * k1 + 1] - 1;
        xc[ka][0] = vpx[a] - vpx[b];
        xc[ka][1] = vpy[a] - vpy[b];
        xc[ka][2] = vpz[a] - vpz[b];
      }
      int count1 = 0;
      for (int ka = 0; ka < cnstp; ++ka) { /* :196-272 */
        int k1 = kl[ka] - 1, la = cnst[2 * k1] - 1, lb = cnst[2 * k1 + 1] - 1;
        const double *r1 = bv + 3 * k1;
        dd[ka] = xc[ka][0] * r1[0] + xc[ka][1] * r1[1] + xc[ka][2] * r1[2];
        for (int kb = 0; kb < cnstp; ++kb) {
          int k2 = kl[kb] - 1, lab = cnst[2 * k2] - 1, lbb = cnst[2 * k2 + 1] - 1;
          if (la != lab && la != lbb && lb != lab && lb != lbb) {
            mat[ka][kb] = 0.0;
          } else {
            double dpa[3], dpb[3];
            mim_dir(la, lab, lbb, bv + 3 * k2, dpa);
            mim_dir(lb, lab, lbb, bv + 3 * k2, dpb);
            double dpx = -dpa[0], dpy = -dpa[1], dpz = -dpa[2];
            matx[count1][0] = dpa[0];
            matx[count1][1] = dpa[1];
            matx[count1][2] = dpa[2];
            count1++;
            double aux3 = (r1[0] * dpx + r1[1] * dpy + r1[2] * dpz) * mass0[la];
            matx[count1][0] = dpb[0];
            matx[count1][1] = dpb[1];
            matx[count1][2] = dpb[2];
            count1++;
            aux3 = aux3 + (r1[0] * dpb[0] + r1[1] * dpb[1] + r1[2] * dpb[2]) * mass0[lb];
            aux3 = aux3 * 2.0;
            mat[ka][kb] = -aux3;
          }
        }
      }
      for (int ka = 0; ka < cnstp; ++ka) gam[ka] = -dd[ka]; /* :278-281 */
      if (cnstp == 2) {
        /* rattle_correc.f:294-309 forms its 2x2 cofactors from named entries */
        double a11 = mat[0][0], a12 = mat[0][1], a21 = mat[1][0], a22 = mat[1][1];
        double b11 = a22, b12 = -a21, b21 = -a12, b22 = a11;
        double det = a11 * a22 - a21 * a12;
        det = 1.0 / det;
        double xd = b11 * gam[0] + b21 * gam[1];
        double yd = b12 * gam[0] + b22 * gam[1];
        gam[0] = xd * det;
        gam[1] = yd * det;
      } else {
        mim_inv inv;
        mim_first_solve(cnstp, mat, &inv, gam);
      }
      count1 = 0; /* :355-383 */
      for (int ka = 0; ka < cnstp; ++ka) {
        int k1 = kl[ka] - 1, a = cnst[2 * k1] - 1, b = cnst[2 * k1 + 1] - 1;
        double xd = mass0[a] * 2.0, yd = mass0[b] * 2.0;
        for (int kb = 0; kb < cnstp; ++kb) {
          if (mask[a]) {
            double aux = gam[kb];
            double aux1 = matx[count1][0] * aux, aux2 = matx[count1][1] * aux, aux3 = matx[count1][2] * aux;
            vpx[a] = vpx[a] + aux1 * xd;
            vpy[a] = vpy[a] + aux2 * xd;
            vpz[a] = vpz[a] + aux3 * xd;
          }
          count1++;
          if (mask[b]) {
            double aux = gam[kb];
            double aux1 = matx[count1][0] * aux, aux2 = matx[count1][1] * aux, aux3 = matx[count1][2] * aux;
            vpx[b] = vpx[b] + aux1 * yd;
            vpy[b] = vpy[b] + aux2 * yd;
            vpz[b] = vpz[b] + aux3 * yd;
          }
          count1++;
        }
        mask[a] = 0;
        mask[b] = 0;
      }
    }
    count = count + cnstp + 1;
  }
  if (dnit) *dnit = *dnit + (double)iter1 / (double)nprot;
  free(mass0);
  free(mask);
  free(bv);
  free(coeffi);
  free(dssi);
  return rc;
}
