/*
 * FP64 CPU oracle for the CHARMM nonbonded + bonded force path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker and the timed CPU baseline.  The HIP
 * product library never links or calls this file.
 *
 * Each routine restates the reference algorithm (file:line cited per
 * function, paths relative to the apoCHARMM source tree) in double precision.
 * Build: see oracle/Makefile (gcc -O3 -fopenmp -shared).
 */
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define ORC_TWO_SQRTPI 1.1283791670955126 /* 2/sqrt(pi) */
#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* Switching constants.  Restates src/CudaPMEDirectForce.cu:321-331          */
/* (k6, k12, dv6, dv12 for VDW_VFSW).                                         */
/* ------------------------------------------------------------------------ */
typedef struct {
  double roff2, ron2;
  double k6, k12, dv6, dv12, roffinv3, roffinv6;
  double kappa, kappa2;
} orc_consts;

static void orc_make_consts(double kappa, double roff, double ron,
                            orc_consts *c) {
  double ron3 = ron * ron * ron, roff3 = roff * roff * roff;
  double ron6 = ron3 * ron3, roff6 = roff3 * roff3;
  c->roff2 = roff * roff;
  c->ron2 = ron * ron;
  c->roffinv3 = 1.0 / roff3;
  c->roffinv6 = 1.0 / roff6;
  if (ron < roff) {
    c->k6 = roff3 / (roff3 - ron3);
    c->k12 = roff6 / (roff6 - ron6);
    c->dv6 = -1.0 / (ron3 * roff3);
    c->dv12 = -1.0 / (ron6 * roff6);
  } else {
    c->k6 = 1.0;
    c->k12 = 1.0;
    c->dv6 = -1.0 / roff6;
    c->dv12 = -1.0 / (roff6 * roff6);
  }
  c->kappa = kappa;
  c->kappa2 = kappa * kappa;
}

/* CHARMM force-switched LJ (VFSW).  Restates
 * src/CudaDirectForceKernels.cu:263-286 with c6s = 6*c6, c12s = 12*c12
 * (src/CudaPMEDirectForce.cu:534-540).  Returns r*dU/dr, energy in *e. */
static inline double orc_vfsw(double r2, double rinv, double c6, double c12,
                              const orc_consts *c, double *e) {
  double c6s = 6.0 * c6, c12s = 12.0 * c12;
  double rinv2 = rinv * rinv;
  double rinv3 = rinv * rinv2;
  double rinv6 = rinv3 * rinv3;
  int outer = r2 > c->ron2;
  double A6 = outer ? c->k6 : 1.0;
  double B6 = outer ? c->roffinv3 : 0.0;
  double A12 = outer ? c->k12 : 1.0;
  double B12 = outer ? c->roffinv6 : 0.0;
  double C6 = outer ? 0.0 : c->dv6;
  double C12 = outer ? 0.0 : c->dv12;
  double t6 = rinv3 - B6, t12 = rinv6 - B12;
  *e = (c12s / 12.0) * (A12 * t12 * t12 + C12) -
       (c6s / 6.0) * (A6 * t6 * t6 + C6);
  return c6s * A6 * t6 * rinv3 - c12s * A12 * t12 * rinv6;
}

/* Ewald real-space Coulomb.  Restates src/CudaDirectForceKernels.cu:476-490:
 * U = qq (erfc(kr) + e14fac - 1)/r, r dU/dr = -qq 2k/sqrt(pi) e^{-k^2 r^2} - U
 * (use_e14fac = 0 gives the plain erfc form). */
static inline double orc_ewald(double r2, double r, double rinv, double qq,
                               double e14fac, int use_e14fac,
                               const orc_consts *c, double *e) {
  double erfc_val = erfc(c->kappa * r);
  double exp_val = exp(-c->kappa2 * r2);
  double u = use_e14fac ? qq * (erfc_val + e14fac - 1.0) * rinv
                        : qq * erfc_val * rinv;
  *e = u;
  return -qq * ORC_TWO_SQRTPI * c->kappa * exp_val - u;
}

static inline double orc_lj_c6(const double *tab, int a, int b, double *c12) {
  int hi = a > b ? a : b, lo = a > b ? b : a;
  int p = hi * (hi + 1) / 2 + lo;
  *c12 = tab[2 * p + 1];
  return tab[2 * p];
}

static int orc_is_excluded(const int *ptr, const int *idx, int i, int j) {
  int lo = ptr[i], hi = ptr[i + 1] - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    if (idx[mid] == j)
      return 1;
    if (idx[mid] < j)
      lo = mid + 1;
    else
      hi = mid - 1;
  }
  return 0;
}

static inline double orc_wrap_frac(double x, double L) {
  double f = x / L;
  f -= floor(f);
  if (f >= 1.0)
    f = 0.0;
  return f;
}

/* One pair of the direct sum; shared by the cell-list and brute-force loops.
 * The pair test r^2 < roff^2 and the exclusion skip follow
 * include/CudaDirectForce_util.h:177-205. */
static inline void orc_direct_pair(int i, int j, const double *xyzq,
                                   const int *vdwtype, const double *vdwtab,
                                   const double *box, const orc_consts *c,
                                   double *g, double *en, double *vir) {
  double dx = xyzq[4 * i] - xyzq[4 * j];
  double dy = xyzq[4 * i + 1] - xyzq[4 * j + 1];
  double dz = xyzq[4 * i + 2] - xyzq[4 * j + 2];
  dx -= box[0] * nearbyint(dx / box[0]);
  dy -= box[1] * nearbyint(dy / box[1]);
  dz -= box[2] * nearbyint(dz / box[2]);
  double r2 = dx * dx + dy * dy + dz * dz;
  if (!(r2 < c->roff2))
    return;
  double r = sqrt(r2), rinv = 1.0 / r;
  double qq = ORC_CCELEC * xyzq[4 * i + 3] * xyzq[4 * j + 3];
  double eel, evdw, c12;
  double fel = orc_ewald(r2, r, rinv, qq, 0.0, 0, c, &eel);
  double c6 = orc_lj_c6(vdwtab, vdwtype[i], vdwtype[j], &c12);
  double fvdw = orc_vfsw(r2, rinv, c6, c12, c, &evdw);
  double fij = (fel + fvdw) * rinv * rinv;
  double gx = fij * dx, gy = fij * dy, gz = fij * dz;
  g[3 * i] += gx;
  g[3 * i + 1] += gy;
  g[3 * i + 2] += gz;
  g[3 * j] -= gx;
  g[3 * j + 1] -= gy;
  g[3 * j + 2] -= gz;
  en[0] += eel;
  en[1] += evdw;
  double d[3] = {dx, dy, dz}, gg[3] = {gx, gy, gz};
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      vir[3 * a + b] -= d[a] * gg[b];
}

/* Direct-space sum: semantics of the pair kernel
 * include/CudaDirectForce_util.h:16-333 (all non-excluded minimum-image
 * pairs inside roff, gradient +fij*d on i, -fij*d on j, virial -sum d (x) g)
 * evaluated by a plain cell list instead of tiles. */
int orc_direct(int n, const double *xyzq, const int *vdwtype, int ntypes,
               const double *vdwtab, const double *box, double kappa,
               double roff, double ron, const int *excl_ptr,
               const int *excl_idx, double *grad, double *energy,
               double *virial, int nthreads) {
  (void)ntypes;
  orc_consts c;
  orc_make_consts(kappa, roff, ron, &c);
  if (nthreads < 1)
    nthreads = 1;
  int nc[3];
  int brute = 0;
  for (int d = 0; d < 3; d++) {
    nc[d] = (int)floor(box[d] / roff);
    if (nc[d] < 3)
      brute = 1;
    if (roff > 0.5 * box[d])
      return -1; /* minimum image would be ambiguous */
  }
  double *tg = (double *)calloc((size_t)nthreads * 3 * n, sizeof(double));
  double *ten = (double *)calloc((size_t)nthreads * 2, sizeof(double));
  double *tvir = (double *)calloc((size_t)nthreads * 9, sizeof(double));
  if (!tg || !ten || !tvir)
    return -2;

  if (brute) {
#pragma omp parallel for schedule(static, 16) num_threads(nthreads)
    for (int i = 0; i < n; i++) {
      int t = 0;
#ifdef _OPENMP
      t = omp_get_thread_num();
#endif
      for (int j = i + 1; j < n; j++) {
        if (excl_ptr && orc_is_excluded(excl_ptr, excl_idx, i, j))
          continue;
        orc_direct_pair(i, j, xyzq, vdwtype, vdwtab, box, &c,
                        tg + (size_t)t * 3 * n, ten + 2 * t, tvir + 9 * t);
      }
    }
  } else {
    int ncell = nc[0] * nc[1] * nc[2];
    int *cnt = (int *)calloc(ncell + 1, sizeof(int));
    int *cellof = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    int *order = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
      int ci[3];
      for (int d = 0; d < 3; d++) {
        ci[d] = (int)(orc_wrap_frac(xyzq[4 * i + d], box[d]) * nc[d]);
        if (ci[d] >= nc[d])
          ci[d] = nc[d] - 1;
      }
      cellof[i] = ci[0] + nc[0] * (ci[1] + nc[1] * ci[2]);
      cnt[cellof[i] + 1]++;
    }
    for (int k = 0; k < ncell; k++)
      cnt[k + 1] += cnt[k];
    int *fill = (int *)calloc(ncell, sizeof(int));
    for (int i = 0; i < n; i++)
      order[cnt[cellof[i]] + fill[cellof[i]]++] = i;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (int cidx = 0; cidx < ncell; cidx++) {
      int t = 0;
#ifdef _OPENMP
      t = omp_get_thread_num();
#endif
      int cx = cidx % nc[0], cy = (cidx / nc[0]) % nc[1], cz = cidx / (nc[0] * nc[1]);
      for (int oz = -1; oz <= 1; oz++)
        for (int oy = -1; oy <= 1; oy++)
          for (int ox = -1; ox <= 1; ox++) {
            int jx = (cx + ox + nc[0]) % nc[0], jy = (cy + oy + nc[1]) % nc[1],
                jz = (cz + oz + nc[2]) % nc[2];
            int jc = jx + nc[0] * (jy + nc[1] * jz);
            for (int a = cnt[cidx]; a < cnt[cidx + 1]; a++) {
              int i = order[a];
              for (int b = cnt[jc]; b < cnt[jc + 1]; b++) {
                int j = order[b];
                if (j <= i)
                  continue;
                if (excl_ptr && orc_is_excluded(excl_ptr, excl_idx, i, j))
                  continue;
                orc_direct_pair(i, j, xyzq, vdwtype, vdwtab, box, &c,
                                tg + (size_t)t * 3 * n, ten + 2 * t,
                                tvir + 9 * t);
              }
            }
          }
    }
    free(cnt);
    free(cellof);
    free(order);
    free(fill);
  }
  /* fixed-order reduction over thread buffers */
  for (int t = 0; t < nthreads; t++) {
    const double *g = tg + (size_t)t * 3 * n;
    for (int k = 0; k < 3 * n; k++)
      grad[k] += g[k];
    energy[0] += ten[2 * t];
    energy[1] += ten[2 * t + 1];
    for (int k = 0; k < 9; k++)
      virial[k] += tvir[9 * t + k];
  }
  free(tg);
  free(ten);
  free(tvir);
  return 0;
}

/* 1-4 inclusion and 1-2/1-3 exclusion-correction pairs.  Restates
 * src/CudaDirectForceKernels.cu:640-870 (calc_ex14_force_device,
 * calc_in14_force_device): no cutoff test and no periodic image (ishift 13,
 * src/CudaPMEDirectForce.cu:964-993); the exclusion correction is the Ewald
 * term with e14fac = 0, i.e. qq (erfc - 1)/r. */
int orc_pairs14(int nin14, const int *in14, int nex14, const int *ex14,
                const double *xyzq, const int *vdwtype, int ntypes,
                const double *vdwtab14, double kappa, double roff, double ron,
                double e14fac, double *grad, double *energy, double *virial) {
  (void)ntypes;
  orc_consts c;
  orc_make_consts(kappa, roff, ron, &c);
  for (int pass = 0; pass < 2; pass++) {
    int np = pass == 0 ? nin14 : nex14;
    const int *lst = pass == 0 ? in14 : ex14;
    for (int p = 0; p < np; p++) {
      int i = lst[2 * p], j = lst[2 * p + 1];
      double dx = xyzq[4 * i] - xyzq[4 * j];
      double dy = xyzq[4 * i + 1] - xyzq[4 * j + 1];
      double dz = xyzq[4 * i + 2] - xyzq[4 * j + 2];
      double r2 = dx * dx + dy * dy + dz * dz;
      double r = sqrt(r2), rinv = 1.0 / r;
      double qq = ORC_CCELEC * xyzq[4 * i + 3] * xyzq[4 * j + 3];
      double fr, eel;
      if (pass == 0) {
        double c12, evdw;
        double c6 = orc_lj_c6(vdwtab14, vdwtype[i], vdwtype[j], &c12);
        double fvdw = orc_vfsw(r2, rinv, c6, c12, &c, &evdw);
        double fel = orc_ewald(r2, r, rinv, qq, e14fac, 1, &c, &eel);
        energy[0] += eel;
        energy[1] += evdw;
        fr = fvdw + fel;
      } else {
        fr = orc_ewald(r2, r, rinv, qq, 0.0, 1, &c, &eel);
        energy[2] += eel;
      }
      double fij = fr * rinv * rinv;
      double g[3] = {fij * dx, fij * dy, fij * dz}, d[3] = {dx, dy, dz};
      for (int a = 0; a < 3; a++) {
        grad[3 * i + a] += g[a];
        grad[3 * j + a] -= g[a];
      }
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
          virial[3 * a + b] -= d[a] * g[b];
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* PME reciprocal                                                             */
/* ------------------------------------------------------------------------ */

/* Cardinal B-spline values and derivatives at offset w.  Restates
 * fill_bspline_host, src/CudaPMERecip.cu:2680-2711 (same recursion as the
 * device calc_theta_dtheta, src/CudaPMERecip.cu:290-355). */
void orc_bspline(int order, double w, double *array, double *darray) {
  array[order - 1] = 0.0;
  array[1] = w;
  array[0] = 1.0 - w;
  for (int k = 3; k <= order - 1; k++) {
    double div = 1.0 / (double)(k - 1);
    array[k - 1] = div * w * array[k - 2];
    for (int j = 1; j <= k - 2; j++)
      array[k - j - 1] =
          div * ((w + j) * array[k - j - 2] + (k - j - w) * array[k - j - 1]);
    array[0] = div * (1.0 - w) * array[0];
  }
  darray[0] = -array[0];
  for (int j = 2; j <= order; j++)
    darray[j - 1] = array[j - 2] - array[j - 1];
  int k = order;
  double div = 1.0 / (double)(k - 1);
  array[k - 1] = div * w * array[k - 2];
  for (int j = 1; j <= k - 2; j++)
    array[k - j - 1] =
        div * ((w + j) * array[k - j - 2] + (k - j - w) * array[k - j - 1]);
  array[0] = div * (1.0 - w) * array[0];
}

/* 1/|b(m)|^2 Euler-spline moduli.  Restates dftmod + calc_prefac,
 * src/CudaPMERecip.cu:2656-2678 and 2719-2745. */
void orc_bsp_mod(int order, int nfft, double *bsp_mod) {
  double array[16], darray[16];
  double *bsp_arr = (double *)calloc(nfft + order + 2, sizeof(double));
  orc_bspline(order, 0.0, array, darray);
  for (int i = 2; i <= order + 1; i++)
    bsp_arr[i - 1] = array[i - 2];
  const double rsmall = 1.0e-10;
  double nfftr = (2.0 * ORC_PI) / (double)nfft;
  for (int k = 1; k <= nfft; k++) {
    double sum1 = 0.0, sum2 = 0.0, arg1 = (k - 1) * nfftr;
    for (int j = 1; j < nfft; j++) {
      double arg = arg1 * (j - 1);
      sum1 += bsp_arr[j - 1] * cos(arg);
      sum2 += bsp_arr[j - 1] * sin(arg);
    }
    bsp_mod[k - 1] = sum1 * sum1 + sum2 * sum2;
  }
  for (int k = 1; k <= nfft; k++)
    if (bsp_mod[k - 1] < rsmall)
      bsp_mod[k - 1] = 0.5 * (bsp_mod[k - 2] + bsp_mod[k]);
  for (int k = 1; k <= nfft; k++)
    bsp_mod[k - 1] = 1.0 / bsp_mod[k - 1];
  free(bsp_arr);
}

static int orc_smallest_factor(int n) {
  for (int p = 2; p * p <= n; p++)
    if (n % p == 0)
      return p;
  return n;
}

/* Recursive mixed-radix decimation-in-time DFT of length n (input strided,
 * output contiguous).  tw[e] = exp(sign*2*pi*i*e/N) for the top length N;
 * twstep = N/n. */
static void orc_fft_rec(int n, const cplx *x, int stride, cplx *y,
                        const cplx *tw, int twstep, int N) {
  if (n == 1) {
    y[0] = x[0];
    return;
  }
  int p = orc_smallest_factor(n);
  if (p == n) {
    for (int k = 0; k < n; k++) {
      cplx s = 0;
      for (int j = 0; j < n; j++)
        s += x[(size_t)j * stride] * tw[((size_t)j * k % n) * twstep];
      y[k] = s;
    }
    return;
  }
  int m = n / p;
  for (int r = 0; r < p; r++)
    orc_fft_rec(m, x + (size_t)r * stride, stride * p, y + (size_t)r * m, tw,
                twstep * p, N);
  cplx t[64], o[64];
  for (int k = 0; k < m; k++) {
    for (int r = 0; r < p; r++)
      t[r] = y[r * m + k] * tw[((size_t)r * k % n) * twstep];
    for (int q = 0; q < p; q++) {
      cplx s = 0;
      for (int r = 0; r < p; r++)
        s += t[r] * tw[((size_t)r * q % p) * (twstep * m)];
      o[q] = s;
    }
    for (int q = 0; q < p; q++)
      y[k + q * m] = o[q];
  }
}

int orc_fft3d(const int *K, double *data, int sign, int nthreads) {
  cplx *g = (cplx *)data;
  if (nthreads < 1)
    nthreads = 1;
  for (int d = 0; d < 3; d++)
    if (K[d] < 1)
      return -1;
  for (int d = 0; d < 3; d++) {
    int n = K[d];
    if (orc_smallest_factor(n) > 64 && orc_smallest_factor(n) != n)
      return -1;
    cplx *tw = (cplx *)malloc(sizeof(cplx) * n);
    for (int e = 0; e < n; e++)
      tw[e] = cexp(sign * 2.0 * ORC_PI * I * (double)e / (double)n);
    size_t stride = d == 0 ? 1 : (d == 1 ? (size_t)K[0] : (size_t)K[0] * K[1]);
    size_t nrows = (size_t)K[0] * K[1] * K[2] / n;
#pragma omp parallel num_threads(nthreads)
    {
      cplx *in = (cplx *)malloc(sizeof(cplx) * n);
      cplx *out = (cplx *)malloc(sizeof(cplx) * n);
#pragma omp for schedule(static)
      for (long row = 0; row < (long)nrows; row++) {
        size_t base;
        if (d == 0)
          base = (size_t)row * n;
        else if (d == 1)
          base = (size_t)(row % K[0]) + (size_t)(row / K[0]) * K[0] * K[1];
        else
          base = (size_t)row;
        for (int e = 0; e < n; e++)
          in[e] = g[base + e * stride];
        orc_fft_rec(n, in, 1, out, tw, 1, n);
        for (int e = 0; e < n; e++)
          g[base + e * stride] = out[e];
      }
      free(in);
      free(out);
    }
    free(tw);
  }
  return 0;
}

/* Fractional grid coordinate, restating include/cuda_pme_spread.h:64-69 and
 * include/cuda_pme_gather.h:60-68: w = x/L + 2, fr = K (w - (floor(w+0.5)-0.5)). */
static inline void orc_frac(double x, double L, int K, int *ix, double *w) {
  double ww = x / L + 2.0;
  double fr = (double)K * (ww - (floor(ww + 0.5) - 0.5));
  int fi = (int)fr;
  *ix = fi;
  *w = fr - (double)fi;
}

/* Smooth PME.  Spread restates include/cuda_pme_spread.h:60-141; the
 * k-space step restates scalar_sum_ortho_kernel src/CudaPMERecip.cu:673-929
 * (including its Hermitian-mirror energy/virial bookkeeping over the x half
 * grid); gather restates include/cuda_pme_gather.h:140-319.  The 3-D FFT is
 * a full complex DFT here (the reference uses R2C/C2R; identical results for
 * the real, Hermitian data involved). */
int orc_recip(int n, const double *xyzq, const double *box, const int *K,
              int order, double kappa, double *grad, double *energy,
              double *virial, int nthreads) {
  if (order < 2 || order > 12)
    return -1;
  size_t ng = (size_t)K[0] * K[1] * K[2];
  double *grid = (double *)calloc(2 * ng, sizeof(double));
  int *aix = (int *)malloc(sizeof(int) * 3 * (n > 0 ? n : 1));
  double *th = (double *)malloc(sizeof(double) * 3 * order * (n > 0 ? n : 1));
  double *dth = (double *)malloc(sizeof(double) * 3 * order * (n > 0 ? n : 1));
  if (!grid || !aix || !th || !dth)
    return -2;
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      double w;
      orc_frac(xyzq[4 * i + d], box[d], K[d], &aix[3 * i + d], &w);
      orc_bspline(order, w, th + (size_t)(3 * i + d) * order,
                  dth + (size_t)(3 * i + d) * order);
    }
  }
  /* spread (serial: fixed summation order) */
  for (int i = 0; i < n; i++) {
    double q = xyzq[4 * i + 3];
    const double *tx = th + (size_t)(3 * i) * order;
    const double *ty = tx + order, *tz = ty + order;
    for (int c = 0; c < order; c++) {
      int z = aix[3 * i + 2] + c;
      while (z >= K[2]) z -= K[2];
      for (int b = 0; b < order; b++) {
        int y = aix[3 * i + 1] + b;
        while (y >= K[1]) y -= K[1];
        for (int a = 0; a < order; a++) {
          int x = aix[3 * i] + a;
          while (x >= K[0]) x -= K[0];
          grid[2 * (x + (size_t)K[0] * (y + (size_t)K[1] * z))] +=
              q * tx[a] * ty[b] * tz[c];
        }
      }
    }
  }
  orc_fft3d(K, grid, -1, nthreads);

  double *pre[3];
  for (int d = 0; d < 3; d++) {
    pre[d] = (double *)malloc(sizeof(double) * K[d]);
    orc_bsp_mod(order, K[d], pre[d]);
  }
  double recip[3] = {1.0 / box[0], 1.0 / box[1], 1.0 / box[2]};
  double vol_inv = recip[0] * recip[1] * recip[2];
  double piv_inv = vol_inv / ORC_PI;
  double fac = ORC_PI * ORC_PI / (kappa * kappa);
  int nf[3];
  for (int d = 0; d < 3; d++)
    nf[d] = K[d] / 2 + (K[d] % 2);

#define ORC_ETERM(k0, k1, k2, mh0, mh1, mh2, msq)                             \
  do {                                                                       \
    int m0_ = (k0) >= nf[0] ? (k0) - K[0] : (k0);                            \
    int m1_ = (k1) >= nf[1] ? (k1) - K[1] : (k1);                            \
    int m2_ = (k2) >= nf[2] ? (k2) - K[2] : (k2);                            \
    mh0 = recip[0] * m0_;                                                    \
    mh1 = recip[1] * m1_;                                                    \
    mh2 = recip[2] * m2_;                                                    \
    msq = mh0 * mh0 + mh1 * mh1 + mh2 * mh2;                                 \
  } while (0)

  /* energy / virial over the x half grid with the mirror terms */
  double e = 0, v[6] = {0, 0, 0, 0, 0, 0};
  int size0 = K[0] / 2 + 1;
  for (int k2 = 0; k2 < K[2]; k2++)
    for (int k1 = 0; k1 < K[1]; k1++)
      for (int k0 = 0; k0 < size0; k0++) {
        if (k0 == 0 && k1 == 0 && k2 == 0)
          continue;
        size_t idx = k0 + (size_t)K[0] * (k1 + (size_t)K[1] * k2);
        double qre = grid[2 * idx], qim = grid[2 * idx + 1];
        double q2 = qre * qre + qim * qim;
        for (int mir = 0; mir < 2; mir++) {
          int a0 = k0, a1 = k1, a2 = k2;
          if (mir) {
            if (!(k0 >= 1 && k0 < K[0]))
              break;
            a0 = K[0] - k0;
            a1 = (K[1] - k1) % K[1];
            a2 = (K[2] - k2) % K[2];
          }
          double mh0, mh1, mh2, msq;
          ORC_ETERM(a0, a1, a2, mh0, mh1, mh2, msq);
          double eterm = exp(-fac * msq) * piv_inv * pre[0][a0] * pre[1][a1] *
                         pre[2][a2] / msq;
          double t1 = eterm * q2;
          double vterm = 2.0 * (fac + 1.0 / msq);
          e += t1;
          v[0] += t1 * (vterm * mh0 * mh0 - 1.0);
          v[1] += t1 * vterm * mh0 * mh1;
          v[2] += t1 * vterm * mh0 * mh2;
          v[3] += t1 * (vterm * mh1 * mh1 - 1.0);
          v[4] += t1 * vterm * mh1 * mh2;
          v[5] += t1 * (vterm * mh2 * mh2 - 1.0);
        }
      }
  double hc = 0.5 * ORC_CCELEC;
  energy[0] += e * hc;
  virial[0] -= hc * v[0];
  virial[1] -= hc * v[1];
  virial[2] -= hc * v[2];
  virial[3] -= hc * v[1];
  virial[4] -= hc * v[3];
  virial[5] -= hc * v[4];
  virial[6] -= hc * v[2];
  virial[7] -= hc * v[4];
  virial[8] -= hc * v[5];

  /* convolution over the full grid, then inverse transform */
  for (int k2 = 0; k2 < K[2]; k2++)
    for (int k1 = 0; k1 < K[1]; k1++)
      for (int k0 = 0; k0 < K[0]; k0++) {
        size_t idx = k0 + (size_t)K[0] * (k1 + (size_t)K[1] * k2);
        if (idx == 0) {
          grid[0] = grid[1] = 0.0;
          continue;
        }
        double mh0, mh1, mh2, msq;
        ORC_ETERM(k0, k1, k2, mh0, mh1, mh2, msq);
        double eterm = exp(-fac * msq) * piv_inv * pre[0][k0] * pre[1][k1] *
                       pre[2][k2] / msq;
        grid[2 * idx] *= eterm;
        grid[2 * idx + 1] *= eterm;
      }
#undef ORC_ETERM
  orc_fft3d(K, grid, +1, nthreads);

  /* gather */
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int i = 0; i < n; i++) {
    const double *tx = th + (size_t)(3 * i) * order;
    const double *ty = tx + order, *tz = ty + order;
    const double *dx = dth + (size_t)(3 * i) * order;
    const double *dy = dx + order, *dz = dy + order;
    double f1 = 0, f2 = 0, f3 = 0;
    for (int c = 0; c < order; c++) {
      int z = aix[3 * i + 2] + c;
      while (z >= K[2]) z -= K[2];
      for (int b = 0; b < order; b++) {
        int y = aix[3 * i + 1] + b;
        while (y >= K[1]) y -= K[1];
        for (int a = 0; a < order; a++) {
          int x = aix[3 * i] + a;
          while (x >= K[0]) x -= K[0];
          double phi = grid[2 * (x + (size_t)K[0] * (y + (size_t)K[1] * z))];
          f1 += dx[a] * ty[b] * tz[c] * phi;
          f2 += tx[a] * dy[b] * tz[c] * phi;
          f3 += tx[a] * ty[b] * dz[c] * phi;
        }
      }
    }
    double q = xyzq[4 * i + 3] * ORC_CCELEC;
    grad[3 * i] += q * recip[0] * f1 * K[0];
    grad[3 * i + 1] += q * recip[1] * f2 * K[1];
    grad[3 * i + 2] += q * recip[2] * f3 * K[2];
  }
  for (int d = 0; d < 3; d++)
    free(pre[d]);
  free(grid);
  free(aix);
  free(th);
  free(dth);
  return 0;
}

/* Restates include/cuda_pme_selfe.h:14-53. */
double orc_self_energy(int n, const double *xyzq, double kappa) {
  double s = 0.0;
  for (int i = 0; i < n; i++)
    s += xyzq[4 * i + 3] * xyzq[4 * i + 3];
  return -kappa * ORC_CCELEC / sqrt(ORC_PI) * s;
}

/* ------------------------------------------------------------------------ */
/* Bonded                                                                    */
/* ------------------------------------------------------------------------ */

static inline void orc_add(double *grad, int i, const double *g) {
  grad[3 * i] += g[0];
  grad[3 * i + 1] += g[1];
  grad[3 * i + 2] += g[2];
}

static inline void orc_vec(const double *xyzq, int i, int j, double *d) {
  for (int a = 0; a < 3; a++)
    d[a] = xyzq[4 * i + a] - xyzq[4 * j + a];
}

static void orc_virial_add(double *virial, const double *d, const double *g) {
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      virial[3 * a + b] -= d[a] * g[b];
}

/* Harmonic bond / Urey-Bradley, restates src/CudaBondedForce.cu:84-146. */
static double orc_bond_term(const double *xyzq, int i, int j, double b0,
                            double kb, double *grad, double *virial) {
  double d[3];
  orc_vec(xyzq, i, j, d);
  double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  double db = r - b0;
  double fij = 2.0 * kb * db / r;
  double g[3] = {fij * d[0], fij * d[1], fij * d[2]}, mg[3] = {-g[0], -g[1], -g[2]};
  orc_add(grad, i, g);
  orc_add(grad, j, mg);
  orc_virial_add(virial, d, g);
  return kb * db * db;
}

/* Dihedral angle phi and its Cartesian derivatives (F = ri-rj, G = rj-rk,
 * H = rl-rk, A = F x G, B = H x G, cos = A.B/|A||B|, sin = |G| H.A/|A||B|,
 * the convention of src/CudaBondedForce.cu:515-580).  Derivatives from the
 * Blondel-Karplus (J. Comput. Chem. 17, 1132 (1996)) expressions. */
static double orc_dihedral_angle(const double *xyzq, int i, int j, int k,
                                 int l, double *dphi /* 12 */,
                                 double *F, double *G, double *H) {
  orc_vec(xyzq, i, j, F);
  orc_vec(xyzq, j, k, G);
  orc_vec(xyzq, l, k, H);
  double A[3] = {F[1] * G[2] - F[2] * G[1], F[2] * G[0] - F[0] * G[2],
                 F[0] * G[1] - F[1] * G[0]};
  double B[3] = {H[1] * G[2] - H[2] * G[1], H[2] * G[0] - H[0] * G[2],
                 H[0] * G[1] - H[1] * G[0]};
  double ra2 = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
  double rb2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
  double rg = sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]);
  double rab = sqrt(ra2 * rb2);
  double ct = (A[0] * B[0] + A[1] * B[1] + A[2] * B[2]) / rab;
  double st = rg * (A[0] * H[0] + A[1] * H[1] + A[2] * H[2]) / rab;
  double phi = atan2(st, ct);
  double fg = F[0] * G[0] + F[1] * G[1] + F[2] * G[2];
  double hg = H[0] * G[0] + H[1] * G[1] + H[2] * G[2];
  /* dphi/dri = -|G|/A^2 A, dphi/drl = |G|/B^2 B,
   * dphi/drj = -dphi/dri + (F.G)/(|G| A^2) A - (H.G)/(|G| B^2) B,
   * dphi/drk = -dphi/drl - (F.G)/(|G| A^2) A + (H.G)/(|G| B^2) B */
  for (int a = 0; a < 3; a++) {
    double di = -rg / ra2 * A[a];
    double dl = rg / rb2 * B[a];
    double cross = fg / (rg * ra2) * A[a] - hg / (rg * rb2) * B[a];
    dphi[a] = di;
    dphi[3 + a] = -di + cross;
    dphi[6 + a] = -dl - cross;
    dphi[9 + a] = dl;
  }
  return phi;
}

int orc_bonded(const double *xyzq, int nb, const int *bonds,
               const double *bond_par, int nub, const int *ureyb,
               const double *ub_par, int na, const int *angles,
               const double *ang_par, int nd, const int *dihe,
               const double *dihe_par, int ni, const int *imp,
               const double *imp_par, double *grad, double *energy,
               double *virial) {
  for (int b = 0; b < nb; b++)
    energy[0] += orc_bond_term(xyzq, bonds[2 * b], bonds[2 * b + 1],
                               bond_par[2 * b], bond_par[2 * b + 1], grad,
                               virial);
  for (int b = 0; b < nub; b++)
    energy[2] += orc_bond_term(xyzq, ureyb[2 * b], ureyb[2 * b + 1],
                               ub_par[2 * b], ub_par[2 * b + 1], grad, virial);
  /* Angles: restates src/CudaBondedForce.cu:222-300, including the clamp of
   * cos(theta) to [-0.999, 0.999] used both in the energy and in the
   * derivative factors. */
  for (int a = 0; a < na; a++) {
    int i = angles[3 * a], j = angles[3 * a + 1], k = angles[3 * a + 2];
    double dij[3], dkj[3];
    orc_vec(xyzq, i, j, dij);
    orc_vec(xyzq, k, j, dkj);
    double rij = sqrt(dij[0] * dij[0] + dij[1] * dij[1] + dij[2] * dij[2]);
    double rkj = sqrt(dkj[0] * dkj[0] + dkj[1] * dkj[1] + dkj[2] * dkj[2]);
    double uij[3], ukj[3];
    for (int c = 0; c < 3; c++) {
      uij[c] = dij[c] / rij;
      ukj[c] = dkj[c] / rkj;
    }
    double cst = uij[0] * ukj[0] + uij[1] * ukj[1] + uij[2] * ukj[2];
    if (cst > 0.999) cst = 0.999;
    if (cst < -0.999) cst = -0.999;
    double th = acos(cst);
    double da = th - ang_par[2 * a];
    double kt = ang_par[2 * a + 1];
    energy[1] += kt * da * da;
    double dEdth = 2.0 * kt * da;
    double inv_s = 1.0 / sqrt(1.0 - cst * cst);
    double gi[3], gk[3], gj[3];
    for (int c = 0; c < 3; c++) {
      /* dtheta/dx_i = -(u_kj - cos u_ij)/(r_ij sin) */
      gi[c] = -dEdth * inv_s * (ukj[c] - cst * uij[c]) / rij;
      gk[c] = -dEdth * inv_s * (uij[c] - cst * ukj[c]) / rkj;
      gj[c] = -gi[c] - gk[c];
    }
    orc_add(grad, i, gi);
    orc_add(grad, j, gj);
    orc_add(grad, k, gk);
    orc_virial_add(virial, dij, gi);
    orc_virial_add(virial, dkj, gk);
  }
  /* Proper (periodic) and improper dihedrals: energies of
   * src/CudaBondedForce.cu:386-513 (k(1+cos(n phi - delta)) and harmonic
   * k (phi - phi0)^2 wrapped to (-pi, pi]). */
  for (int pass = 0; pass < 2; pass++) {
    int nn = pass == 0 ? nd : ni;
    const int *lst = pass == 0 ? dihe : imp;
    const double *par = pass == 0 ? dihe_par : imp_par;
    for (int t = 0; t < nn; t++) {
      int i = lst[4 * t], j = lst[4 * t + 1], k = lst[4 * t + 2],
          l = lst[4 * t + 3];
      double dphi[12], F[3], G[3], H[3];
      double phi = orc_dihedral_angle(xyzq, i, j, k, l, dphi, F, G, H);
      int nper = (int)par[3 * t];
      double kc = par[3 * t + 1], delta = par[3 * t + 2];
      double e, dEdphi;
      if (pass == 0 && nper == 0) {
        /* iper == 0: constant k, no force (src/CudaBondedForce.cu:419-421) */
        e = kc;
        dEdphi = 0.0;
      } else if (nper != 0 || pass == 0) {
        e = kc * (1.0 + cos(nper * phi - delta));
        dEdphi = -kc * nper * sin(nper * phi - delta);
      } else {
        double dp = phi - delta;
        while (dp > ORC_PI) dp -= 2.0 * ORC_PI;
        while (dp <= -ORC_PI) dp += 2.0 * ORC_PI;
        e = kc * dp * dp;
        dEdphi = 2.0 * kc * dp;
      }
      energy[pass == 0 ? 3 : 4] += e;
      double gi[3], gj[3], gk[3], gl[3];
      for (int c = 0; c < 3; c++) {
        gi[c] = dEdphi * dphi[c];
        gj[c] = dEdphi * dphi[3 + c];
        gk[c] = dEdphi * dphi[6 + c];
        gl[c] = dEdphi * dphi[9 + c];
      }
      orc_add(grad, i, gi);
      orc_add(grad, j, gj);
      orc_add(grad, k, gk);
      orc_add(grad, l, gl);
      /* W = -(F (x) g_i - G (x) g_k + (H - G) (x) g_l) */
      double mG[3] = {-G[0], -G[1], -G[2]};
      double HG[3] = {H[0] - G[0], H[1] - G[1], H[2] - G[2]};
      orc_virial_add(virial, F, gi);
      orc_virial_add(virial, mG, gk);
      orc_virial_add(virial, HG, gl);
    }
  }
  return 0;
}
