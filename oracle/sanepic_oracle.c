/*
 * sanepic_oracle.c -- FP64 CPU restatement of SANEPIC's CG GLS map-maker.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU product is
 * compared against, and the "port" CPU baseline that bench.py times beside
 * the GPU.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product (sanepic_amd/) never links or calls it.
 *
 * It follows the reference's operation order for every hot-path function
 * (citations are src/sanepic.cc line numbers of patanch/SANEPIC):
 *
 *   or_spn_row            SpN_all diagonal row            509-539
 *   or_pointingshift      boresight -> detector pointing  3412-3476
 *   or_ang2pix_ring       HEALPix RING pixel              2764-2843
 *   or_problem_pixelize   pointing loop + pixelisation    722-773, 848-951
 *   or_problem_load_tod   TOD NaN fix + r2c -> fdatas0    775-803
 *   or_problem_set_dipole orbital dipole out of fdatas0   700-720, 1055-1067
 *   or_deproject          P (map -> TOD)                  2845-2873
 *   or_apply_A            write_tfAS + do_PtNd            2390-2445, 2489-2637
 *   or_diag_corr          compute_diagPtNPCorr            2680-2762
 *   or_problem_solve      masks + diagonal PCG            1069-1761
 *   or_exact_blocks       exact 3x3 diagonal blocks of A  (no counterpart)
 *   or_problem_solve_pcg  block-Jacobi / A-DEF2 PCG       (no counterpart)
 *   or_chi2               chi2 of a map                   1514-1538
 *   or_problem_set_gains  relCalib of the recalibration   1041-1067, 2543-2545
 *   or_recal_sums         gain fit of one loop            2185-2228
 *   or_recal_chi2         MPI_compute_chi2                3206-3255
 *
 * FFT: the reference calls FFTW3 (version unpinned in the reference; its
 * Make.cm2 names fftw 3.3.10), which is absent from this image.  FFTW's
 * published r2c/c2r definitions are restated here: r2c is the unnormalised
 * forward DFT X[k] = sum_j x[j] e^{-2 pi i jk/n} for k <= n/2, and c2r is
 * the unnormalised inverse of the Hermitian extension that ignores the
 * imaginary parts of the DC and (even n) Nyquist bins, so c2r(r2c(x)) = n x.
 * The transform itself is this file's own Stockham mixed-radix FFT.
 *
 * Parity pinning (each checked in tests/): FFT against numpy.fft and a
 * brute-force DFT; ang2pix against an independent nested-scheme
 * implementation and pix2ang round trips; pointingshift against exact
 * zero-offset identities and an independent rotation-matrix construction;
 * P, N^-1 and A against dense brute-force matrices; the PCG against a dense
 * solve; the diagonal preconditioner against a brute-force pair sum.
 * Floating-point summation order differs from the reference wherever the
 * reference's own order is unspecified (qsort tie order at 2720, MPI reduce
 * order), so those results are compared to tolerances, not bits.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math), so the
 * pixel chain evaluates with the same IEEE operations and glibc libm calls as
 * the reference's GNU -O0 build (makefiles/Make.OR:1).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_D2PI 6.2831853071795864769252867665590057683943387987502

typedef struct {
  double re, im;
} cplx;

/* ======================================================================== */
/* FFT: Stockham autosort, radices 4, 2, 3, 5 and generic odd p              */
/* ======================================================================== */

typedef struct or_fft {
  long n;
  int nf;
  int fac[64];
  cplx *twf; /* exp(-2 pi i j / n) */
  cplx *twb; /* exp(+2 pi i j / n) */
} or_fft;

static void or_factorize(long n, int *fac, int *nf) {
  int k = 0;
  while (n % 4 == 0) {
    fac[k++] = 4;
    n /= 4;
  }
  while (n % 2 == 0) {
    fac[k++] = 2;
    n /= 2;
  }
  for (long p = 3; p * p <= n; p += 2)
    while (n % p == 0) {
      fac[k++] = (int)p;
      n /= p;
    }
  if (n > 1)
    fac[k++] = (int)n;
  *nf = k;
}

or_fft *or_fft_new(long n) {
  or_fft *f = (or_fft *)calloc(1, sizeof(or_fft));
  f->n = n;
  or_factorize(n, f->fac, &f->nf);
  f->twf = (cplx *)malloc(sizeof(cplx) * n);
  f->twb = (cplx *)malloc(sizeof(cplx) * n);
  for (long j = 0; j < n; j++) {
    /* angle reduced to [0, pi/4]-friendly form via long double for accuracy */
    long double a = (long double)2.0L * 3.141592653589793238462643383279503L *
                    (long double)j / (long double)n;
    double c = (double)cosl(a), s = (double)sinl(a);
    f->twf[j].re = c;
    f->twf[j].im = -s;
    f->twb[j].re = c;
    f->twb[j].im = s;
  }
  return f;
}

void or_fft_free(or_fft *f) {
  if (!f)
    return;
  free(f->twf);
  free(f->twb);
  free(f);
}

static inline cplx cmul(cplx a, cplx b) {
  cplx r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.re * b.im + a.im * b.re;
  return r;
}

/* One radix-p butterfly on a[0..p-1] in place; sign -1 forward, +1 inverse. */
static void or_dft_small(int p, cplx *a, const cplx *tw, long n, int sign) {
  if (p == 2) {
    cplx x0 = a[0], x1 = a[1];
    a[0].re = x0.re + x1.re;
    a[0].im = x0.im + x1.im;
    a[1].re = x0.re - x1.re;
    a[1].im = x0.im - x1.im;
    return;
  }
  if (p == 4) {
    cplx s02 = {a[0].re + a[2].re, a[0].im + a[2].im};
    cplx d02 = {a[0].re - a[2].re, a[0].im - a[2].im};
    cplx s13 = {a[1].re + a[3].re, a[1].im + a[3].im};
    cplx d13 = {a[1].re - a[3].re, a[1].im - a[3].im};
    /* forward: b1 = d02 - i d13, b3 = d02 + i d13 */
    double sg = (double)sign;
    a[0].re = s02.re + s13.re;
    a[0].im = s02.im + s13.im;
    a[2].re = s02.re - s13.re;
    a[2].im = s02.im - s13.im;
    a[1].re = d02.re - sg * d13.im;
    a[1].im = d02.im + sg * d13.re;
    a[3].re = d02.re + sg * d13.im;
    a[3].im = d02.im - sg * d13.re;
    return;
  }
  cplx out[64];
  cplx *o = out;
  cplx *heap = NULL;
  if (p > 64) {
    heap = (cplx *)malloc(sizeof(cplx) * p);
    o = heap;
  }
  long step = n / p;
  for (int s = 0; s < p; s++) {
    double re = 0.0, im = 0.0;
    for (int r = 0; r < p; r++) {
      cplx w = tw[((long)r * s % p) * step];
      re += a[r].re * w.re - a[r].im * w.im;
      im += a[r].re * w.im + a[r].im * w.re;
    }
    o[s].re = re;
    o[s].im = im;
  }
  memcpy(a, o, sizeof(cplx) * p);
  free(heap);
}

/* In-place unnormalised complex DFT of x[0..n-1]; work holds n entries. */
void or_fft_exec(const or_fft *f, cplx *x, cplx *work, int sign) {
  long n = f->n;
  if (n == 1)
    return;
  const cplx *tw = sign < 0 ? f->twf : f->twb;
  cplx *in = x, *out = work;
  long l = 1;
  cplx a[64];
  for (int st = 0; st < f->nf; st++) {
    int p = f->fac[st];
    long m = n / (l * p);
    cplx *aa = a;
    cplx *heap = NULL;
    if (p > 64) {
      heap = (cplx *)malloc(sizeof(cplx) * p);
      aa = heap;
    }
    for (long k = 0; k < l; k++) {
      for (long c = 0; c < m; c++) {
        for (int r = 0; r < p; r++) {
          cplx v = in[(k * p + r) * m + c];
          aa[r] = (r == 0 || k == 0) ? v : cmul(v, tw[(long)r * k * m]);
        }
        or_dft_small(p, aa, tw, n, sign);
        for (int s = 0; s < p; s++)
          out[(k + l * s) * m + c] = aa[s];
      }
    }
    free(heap);
    cplx *t = in;
    in = out;
    out = t;
    l *= p;
  }
  if (in != x)
    memcpy(x, in, sizeof(cplx) * n);
}

/* Real transforms with FFTW's r2c / c2r conventions. */
typedef struct or_rfft {
  long n;
  or_fft *c;  /* length n/2 (even n) or n (odd n) */
  cplx *wk;   /* exp(-2 pi i k / n), k <= n/2 (even n) */
} or_rfft;

or_rfft *or_rfft_new(long n) {
  or_rfft *r = (or_rfft *)calloc(1, sizeof(or_rfft));
  r->n = n;
  if (n % 2 == 0) {
    r->c = or_fft_new(n / 2);
    r->wk = (cplx *)malloc(sizeof(cplx) * (n / 2 + 1));
    for (long k = 0; k <= n / 2; k++) {
      long double a = (long double)2.0L * 3.141592653589793238462643383279503L *
                      (long double)k / (long double)n;
      r->wk[k].re = (double)cosl(a);
      r->wk[k].im = -(double)sinl(a);
    }
  } else {
    r->c = or_fft_new(n);
  }
  return r;
}

void or_rfft_free(or_rfft *r) {
  if (!r)
    return;
  or_fft_free(r->c);
  free(r->wk);
  free(r);
}

/* work: at least 2*n cplx entries. */
void or_rfft_r2c(const or_rfft *r, const double *x, cplx *X, cplx *work) {
  long n = r->n;
  if (n % 2 == 0) {
    long h = n / 2;
    cplx *z = work, *w2 = work + h;
    for (long j = 0; j < h; j++) {
      z[j].re = x[2 * j];
      z[j].im = x[2 * j + 1];
    }
    or_fft_exec(r->c, z, w2, -1);
    for (long k = 0; k <= h; k++) {
      cplx zk = z[k % h];
      cplx zc = z[(h - k) % h];
      zc.im = -zc.im;
      cplx e = {0.5 * (zk.re + zc.re), 0.5 * (zk.im + zc.im)};
      /* o = (zk - zc) / (2i) */
      cplx o = {0.5 * (zk.im - zc.im), -0.5 * (zk.re - zc.re)};
      cplx wo = cmul(r->wk[k], o);
      X[k].re = e.re + wo.re;
      X[k].im = e.im + wo.im;
    }
  } else {
    cplx *z = work, *w2 = work + n;
    for (long j = 0; j < n; j++) {
      z[j].re = x[j];
      z[j].im = 0.0;
    }
    or_fft_exec(r->c, z, w2, -1);
    for (long k = 0; k <= n / 2; k++)
      X[k] = z[k];
  }
}

void or_rfft_c2r(const or_rfft *r, const cplx *X, double *x, cplx *work) {
  long n = r->n;
  if (n % 2 == 0) {
    long h = n / 2;
    cplx *z = work, *w2 = work + h;
    for (long k = 0; k < h; k++) {
      cplx xk = X[k];
      cplx xc = X[h - k];
      if (k == 0) {
        xk.im = 0.0; /* FFTW ignores Im of DC and Nyquist */
        xc.im = 0.0;
      }
      xc.im = -xc.im;
      cplx s = {xk.re + xc.re, xk.im + xc.im};
      cplx d = {xk.re - xc.re, xk.im - xc.im};
      cplx winv = {r->wk[k].re, -r->wk[k].im}; /* exp(+2 pi i k/n) */
      cplx t = cmul(winv, d);                    /* then multiply by i */
      z[k].re = s.re - t.im;
      z[k].im = s.im + t.re;
    }
    or_fft_exec(r->c, z, w2, +1);
    for (long j = 0; j < h; j++) {
      x[2 * j] = z[j].re;
      x[2 * j + 1] = z[j].im;
    }
  } else {
    cplx *z = work, *w2 = work + n;
    for (long k = 0; k < n; k++) {
      if (k <= n / 2) {
        z[k] = X[k];
      } else {
        z[k].re = X[n - k].re;
        z[k].im = -X[n - k].im;
      }
    }
    z[0].im = 0.0;
    or_fft_exec(r->c, z, w2, +1);
    for (long j = 0; j < n; j++)
      x[j] = z[j].re;
  }
}

/* ctypes conveniences (plan per call). X is interleaved (re, im). */
void or_r2c(long n, const double *x, double *X) {
  or_rfft *r = or_rfft_new(n);
  cplx *w = (cplx *)malloc(sizeof(cplx) * 2 * n);
  or_rfft_r2c(r, x, (cplx *)X, w);
  free(w);
  or_rfft_free(r);
}

void or_c2r(long n, const double *X, double *x) {
  or_rfft *r = or_rfft_new(n);
  cplx *w = (cplx *)malloc(sizeof(cplx) * 2 * n);
  or_rfft_c2r(r, (const cplx *)X, x, w);
  free(w);
  or_rfft_free(r);
}

void or_c2c(long n, double *x, int sign) {
  or_fft *f = or_fft_new(n);
  cplx *w = (cplx *)malloc(sizeof(cplx) * n);
  or_fft_exec(f, (cplx *)x, w, sign);
  free(w);
  or_fft_free(f);
}

/* ======================================================================== */
/* Noise weights, pointing, pixelisation                                     */
/* ======================================================================== */

/* SpN_all diagonal row for one detector (sanepic.cc:533-537): bin ii takes
 * iSpf[long(double(ii)/nssp*ns)] / ns / ns, then bin 0 is replaced by
 * iSpf[1]*1e-6/ns/ns.  The index expression truncates exactly as written. */
void or_spn_row(const double *ispf, long nssp, long ns, double *row) {
  for (long ii = 0; ii < ns; ii++)
    row[ii] = ispf[(long)((double)ii / nssp * ns)] / (double)ns / (double)ns;
  row[0] = ispf[1] * 1e-6 / (double)ns / (double)ns;
}

/* Boresight -> detector pointing (sanepic.cc:3412-3476), same expression
 * order; x, y are the detector offsets in radians. */
void or_pointingshift(double x, double y, long nn, const double *thetap,
                      const double *phip, const double *psip, double *theta,
                      double *phi, double *psi) {
  double th_ = asin(sqrt(x * x + y * y));
  double ph_ = atan2(y, x);
  for (long i = 0; i < nn; i++) {
    theta[i] = acos(sin(thetap[i]) * sin(th_) * cos(ph_ + psip[i]) +
                    cos(thetap[i]) * cos(th_));
    phi[i] = atan2(-sin(th_) * sin(ph_ + psip[i]),
                   (-cos(thetap[i]) * sin(th_) * cos(ph_ + psip[i]) +
                    sin(thetap[i]) * cos(th_))) +
             phip[i];
  }
  double drp[3], dr[3], eph[3], R[3][3];
  drp[2] = -sin(th_) * cos(ph_);
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      R[a][b] = 0.0;
  R[1][1] = -1.0;
  eph[2] = 0.0;
  for (long i = 0; i < nn; i++) {
    drp[0] = cos(psip[i]) * cos(th_);
    drp[1] = sin(psip[i]) * cos(th_);
    R[0][0] = -cos(thetap[i]);
    R[0][2] = sin(thetap[i]);
    R[2][0] = sin(thetap[i]);
    R[2][2] = cos(thetap[i]);
    for (int k = 0; k < 3; k++)
      dr[k] = 0.0;
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        dr[k] += R[k][j] * drp[j];
    eph[0] = -sin(phi[i] - phip[i]);
    eph[1] = cos(phi[i] - phip[i]);
    double aaa = 0.0, rrr = 0.0;
    for (int k = 0; k < 3; k++) {
      rrr += dr[k] * dr[k];
      aaa += dr[k] * eph[k];
    }
    aaa = aaa / sqrt(rrr);
    if (aaa > 1.0)
      aaa = 1.0;
    if (aaa < -1.0)
      aaa = -1.0;
    double pt = acos(aaa);
    if (dr[2] < 0)
      pt = -pt;
    pt = (fmod(pt + OR_D2PI / 4.0 + 2.0 * OR_D2PI, OR_D2PI)) - OR_D2PI / 2.0;
    psi[i] = pt;
  }
}

/* HEALPix RING pixel (Gorski et al. 2005), arithmetic order of
 * sanepic.cc:2794-2842.  Returns -1 for out-of-range nside or theta (the
 * reference exits there, 2782-2792). */
long or_ang2pix_ring(long nside, double theta, double phi) {
  const double piover2 = 0.5 * M_PI, twopi = 2.0 * M_PI, z0 = 2.0 / 3.0;
  if (nside < 1 || nside > 8192)
    return -1;
  if (theta < 0. || theta > M_PI)
    return -1;
  int nl2 = 2 * (int)nside, nl4 = 4 * (int)nside;
  int ncap = nl2 * ((int)nside - 1);
  int npix = 12 * (int)nside * (int)nside;
  double z = cos(theta);
  double za = fabs(z);
  if (phi >= twopi)
    phi = phi - twopi;
  if (phi < 0.)
    phi = phi + twopi;
  double tt = phi / piover2;
  int jp, jm, ir, ip, ipix1;
  if (za <= z0) {
    jp = (int)floor(nside * (0.5 + tt - z * 0.75));
    jm = (int)floor(nside * (0.5 + tt + z * 0.75));
    ir = (int)nside + 1 + jp - jm;
    int kshift = ((ir & 1) == 0) ? 1 : 0;
    ip = (int)floor((jp + jm - (int)nside + kshift + 1) / 2) + 1;
    if (ip > nl4)
      ip = ip - nl4;
    ipix1 = ncap + nl4 * (ir - 1) + ip;
  } else {
    double tp = tt - floor(tt);
    double tmp = sqrt(3. * (1. - za));
    jp = (int)floor(nside * tp * tmp);
    jm = (int)floor(nside * (1. - tp) * tmp);
    ir = jp + jm + 1;
    ip = (int)floor(tt * ir) + 1;
    if (ip > 4 * ir)
      ip = ip - 4 * ir;
    ipix1 = 2 * ir * (ir - 1) + ip;
    if (z <= 0.)
      ipix1 = npix - 2 * ir * (ir + 1) + ip;
  }
  return (long)ipix1 - 1;
}

/* ======================================================================== */
/* Problem state                                                             */
/* ======================================================================== */

typedef struct or_problem {
  long nside, ndet, nseg, ns;
  int polar, hwp;
  double omega;
  int nthreads;
  double *offx, *offy, *psidet; /* radians, per detector */
  int galbp;                    /* 1: a 4th map component (-b, 551-575) */
  double *relfgal;              /* [ndet] its coefficient per detector */
  int nc;                       /* map components: 3, or 4 with galbp */
  double *spn;                  /* [ndet][ns] */
  long *datatopix;              /* [ndet][nseg][ns] full-sky RING */
  double *psip, *psipFP;        /* [ndet][nseg][ns] */
  long *indpix;                 /* 12 nside^2 */
  long npix;
  long n_bad_pixel;             /* samples that took the previous pixel */
  cplx *fraw;                   /* [ndet][nseg][ns/2+1] r2c of the NaN-filled TOD */
  cplx *fdatas0;                /* [ndet][nseg][ns/2+1] fraw / relCalib - fodipole */
  double *gain;                 /* [ndet][nseg] relCalib, NULL: all 1 */
  cplx *fodipole;               /* [ndet][nseg][ns/2+1] or NULL */
  double *b;                    /* PNdtot, 3 npix */
  double chi20;
  long *hits;                   /* 2 npix (doubled, as the reference) */
  double *Mp, *sumCross;        /* 3 npix */
  or_rfft *plan;
} or_problem;

#define IDX(P, idet, iseg, ii) \
  ((ii) + (P)->ns * (iseg) + (P)->ns * (P)->nseg * (idet))

static inline double or_gain(const or_problem *P, long idet, long iseg) {
  return P->gain ? P->gain[idet * P->nseg + iseg] : 1.0;
}

or_problem *or_problem_new(long nside, long ndet, long nseg, long ns,
                           int polar, int hwp, double omega, int nthreads) {
  or_problem *P = (or_problem *)calloc(1, sizeof(or_problem));
  P->nside = nside;
  P->ndet = ndet;
  P->nseg = nseg;
  P->ns = ns;
  P->polar = polar;
  P->hwp = hwp;
  P->omega = omega;
  P->nthreads = nthreads < 1 ? 1 : nthreads;
  P->offx = (double *)calloc(ndet, sizeof(double));
  P->offy = (double *)calloc(ndet, sizeof(double));
  P->psidet = (double *)calloc(ndet, sizeof(double));
  P->spn = (double *)calloc((size_t)ndet * ns, sizeof(double));
  P->plan = or_rfft_new(ns);
  P->nc = 3;
  P->relfgal = (double *)calloc(ndet, sizeof(double));
  return P;
}

/* The GalBP leakage component (-b 1): maps get a 4th block [I|Q|U|L] and
 * each detector sees L times its coefficient relfGal[idet] (deproject
 * 2869-2870, do_PtNd 2606-2607).  rel: [ndet], already normalised to zero
 * mean and unit variance over the detectors as at 561-570.  Call before
 * any operator use. */
void or_problem_set_leakage(or_problem *P, const double *rel) {
  P->galbp = 1;
  P->nc = 4;
  memcpy(P->relfgal, rel, sizeof(double) * P->ndet);
}

long or_problem_ncomp(const or_problem *P) { return P->nc; }

void or_problem_free(or_problem *P) {
  if (!P)
    return;
  free(P->offx);
  free(P->offy);
  free(P->psidet);
  free(P->relfgal);
  free(P->spn);
  free(P->datatopix);
  free(P->psip);
  free(P->psipFP);
  free(P->indpix);
  free(P->fdatas0);
  free(P->fraw);
  free(P->gain);
  free(P->fodipole);
  free(P->b);
  free(P->hits);
  free(P->Mp);
  free(P->sumCross);
  or_rfft_free(P->plan);
  free(P);
}

long or_problem_npix(const or_problem *P) { return P->npix; }
long or_problem_nbad(const or_problem *P) { return P->n_bad_pixel; }

/* Offsets as parsed from the bolometer file (read_bolo_offsets, 3173-3179:
 * offsets = {xel, lel, angl}) converted as at 458-463. */
void or_problem_set_detector(or_problem *P, long idet, double lel, double xel,
                             double angl, const double *ispf, long nssp) {
  double off0 = xel, off1 = lel, off2 = angl;
  P->offx[idet] = off1 / 180.0 * OR_D2PI / 2.0;
  P->offy[idet] = off0 / 180.0 * OR_D2PI / 2.0;
  P->psidet[idet] = off2 / 180.0 * OR_D2PI / 2.0;
  or_spn_row(ispf, nssp, P->ns, P->spn + idet * P->ns);
}

static void or_nan_fill(double *v, long n) {
  for (long i = 0; i < n; i++)
    if (!(v[i] > 0) && !(v[i] <= 0.0))
      if (i > 0)
        v[i] = v[i - 1];
}

/* Pointing per detector (722-773) and pixelisation (848-951).
 * ra, dec, psi: boresight [nseg][ns], shared by all detectors (read once per
 * segment at 680-692).  psi is NaN-filled in a private copy (722-727). */
int or_problem_pixelize(or_problem *P, const double *ra, const double *dec,
                        const double *psi) {
  long ns = P->ns, nseg = P->nseg, ndet = P->ndet;
  long ntot = ndet * nseg * ns;
  long npixsky = 12 * P->nside * P->nside;
  double *ra_d = (double *)malloc(sizeof(double) * ntot);
  double *dec_d = (double *)malloc(sizeof(double) * ntot);
  P->datatopix = (long *)malloc(sizeof(long) * ntot);
  P->psip = (double *)malloc(sizeof(double) * ntot);
  P->psipFP = (double *)malloc(sizeof(double) * ntot);
  /* Every sample is independent up to the previous-pixel substitution, so
   * the pointing and ang2pix run over segments in parallel and the
   * substitution follows in the reference's sample order. */
#pragma omp parallel num_threads(P->nthreads)
  {
    double *psis = (double *)malloc(sizeof(double) * ns);
#pragma omp for schedule(dynamic)
    for (long iseg = 0; iseg < nseg; iseg++) {
      memcpy(psis, psi + iseg * ns, sizeof(double) * ns);
      or_nan_fill(psis, ns);
      for (long idet = 0; idet < ndet; idet++) {
        for (long ii = 0; ii < ns; ii++) {
          double thetap_i = OR_D2PI / 4.0 - dec[iseg * ns + ii];
          double phip_i = ra[iseg * ns + ii];
          double psip_i = psis[ii];
          double th_o, ph_o, ps_o;
          or_pointingshift(P->offx[idet], P->offy[idet], 1, &thetap_i, &phip_i,
                           &psip_i, &th_o, &ph_o, &ps_o);
          long k = IDX(P, idet, iseg, ii);
          ra_d[k] = ph_o;
          dec_d[k] = OR_D2PI / 4.0 - th_o;
          P->psip[k] = ps_o + P->psidet[idet];
          P->psipFP[k] = P->psip[k];
          if (P->hwp)
            P->psip[k] = -(ps_o + P->psidet[idet]) +
                         2.0 * P->omega * (double)(ii + ns * iseg);
          double theta = OR_D2PI / 4.0 - dec_d[k];
          double phi = fmod(ra_d[k] + OR_D2PI, OR_D2PI);
          long v = -1;
          /* NaN angles make the reference's ang2pix undefined; they and any
           * out-of-range result take the previous pixel (862-872). */
          if (theta == theta && phi == phi)
            v = or_ang2pix_ring(P->nside, theta, phi);
          P->datatopix[k] = v;
        }
      }
    }
    free(psis);
  }
  long *pixon = (long *)calloc(npixsky, sizeof(long));
  long ipr = 0;
  P->n_bad_pixel = 0;
  for (long idet = 0; idet < ndet; idet++)
    for (long iseg = 0; iseg < nseg; iseg++)
      for (long ii = 0; ii < ns; ii++) {
        long k = IDX(P, idet, iseg, ii);
        long prev = ipr;
        ipr = P->datatopix[k];
        if (ipr < 0 || ipr > npixsky - 1) {
          ipr = prev;
          P->n_bad_pixel++;
        }
        pixon[ipr] += 1;
        P->datatopix[k] = ipr;
      }
  free(ra_d);
  free(dec_d);
  P->indpix = (long *)malloc(sizeof(long) * npixsky);
  long ll = 0;
  for (long i = 0; i < npixsky; i++) {
    P->indpix[i] = -1;
    if (pixon[i] != 0)
      P->indpix[i] = ll++;
  }
  P->npix = ll;
  free(pixon);
  return 0;
}

/* Direct setters used by tests that build P from known pixels. */
int or_problem_set_pixels(or_problem *P, const long *datatopix,
                          const double *psip) {
  long ntot = P->ndet * P->nseg * P->ns;
  long npixsky = 12 * P->nside * P->nside;
  P->datatopix = (long *)malloc(sizeof(long) * ntot);
  P->psip = (double *)malloc(sizeof(double) * ntot);
  P->psipFP = (double *)malloc(sizeof(double) * ntot);
  memcpy(P->datatopix, datatopix, sizeof(long) * ntot);
  memcpy(P->psip, psip, sizeof(double) * ntot);
  memcpy(P->psipFP, psip, sizeof(double) * ntot);
  long *pixon = (long *)calloc(npixsky, sizeof(long));
  for (long k = 0; k < ntot; k++)
    pixon[datatopix[k]]++;
  P->indpix = (long *)malloc(sizeof(long) * npixsky);
  long ll = 0;
  for (long i = 0; i < npixsky; i++) {
    P->indpix[i] = -1;
    if (pixon[i] != 0)
      P->indpix[i] = ll++;
  }
  P->npix = ll;
  free(pixon);
  return 0;
}

void or_problem_get_pixels(const or_problem *P, long *datatopix, double *psip,
                           double *psipFP, long *indpix) {
  long ntot = P->ndet * P->nseg * P->ns;
  if (datatopix)
    memcpy(datatopix, P->datatopix, sizeof(long) * ntot);
  if (psip)
    memcpy(psip, P->psip, sizeof(double) * ntot);
  if (psipFP)
    memcpy(psipFP, P->psipFP, sizeof(double) * ntot);
  if (indpix)
    memcpy(indpix, P->indpix, sizeof(long) * 12 * P->nside * P->nside);
}

/* TOD NaN fix (784-790) then r2c into fdatas0 (795-803). tod: [ndet][nseg][ns] */
int or_problem_load_tod(or_problem *P, const double *tod) {
  long ns = P->ns, nh = ns / 2 + 1;
  P->fdatas0 = (cplx *)malloc(sizeof(cplx) * nh * P->ndet * P->nseg);
  P->fraw = (cplx *)malloc(sizeof(cplx) * nh * P->ndet * P->nseg);
#pragma omp parallel num_threads(P->nthreads)
  {
    double *data = (double *)malloc(sizeof(double) * ns);
    cplx *work = (cplx *)malloc(sizeof(cplx) * 2 * ns);
#pragma omp for schedule(dynamic)
    for (long blk = 0; blk < P->ndet * P->nseg; blk++) {
      long idet = blk / P->nseg, iseg = blk % P->nseg;
      memcpy(data, tod + IDX(P, idet, iseg, 0), sizeof(double) * ns);
      or_nan_fill(data, ns);
      or_rfft_r2c(P->plan, data, P->fraw + (idet * P->nseg + iseg) * nh,
                  work);
      memcpy(P->fdatas0 + (idet * P->nseg + iseg) * nh,
             P->fraw + (idet * P->nseg + iseg) * nh, sizeof(cplx) * nh);
    }
    free(data);
    free(work);
  }
  return 0;
}

/* Orbital dipole (-o): fodipole = r2c(orbdata) per (det, seg) block, read
 * without NaN substitution (702-720), then fdatas0 -= fodipole (1055-1067).
 * odip: [ndet][nseg][ns]; call after or_problem_load_tod. */
int or_problem_set_dipole(or_problem *P, const double *odip) {
  long ns = P->ns, nh = ns / 2 + 1, nblk = P->ndet * P->nseg;
  if (!P->fdatas0)
    return -1;
  free(P->fodipole);
  P->fodipole = (cplx *)malloc(sizeof(cplx) * nh * nblk);
#pragma omp parallel num_threads(P->nthreads)
  {
    cplx *work = (cplx *)malloc(sizeof(cplx) * 2 * ns);
#pragma omp for schedule(dynamic)
    for (long blk = 0; blk < nblk; blk++) {
      long idet = blk / P->nseg, iseg = blk % P->nseg;
      cplx *F = P->fodipole + (idet * P->nseg + iseg) * nh;
      cplx *D = P->fdatas0 + (idet * P->nseg + iseg) * nh;
      const cplx *R = P->fraw + (idet * P->nseg + iseg) * nh;
      double g = P->gain ? P->gain[idet * P->nseg + iseg] : 1.0;
      or_rfft_r2c(P->plan, odip + IDX(P, idet, iseg, 0), F, work);
      for (long ii = 0; ii < nh; ii++) {
        D[ii].re = R[ii].re / g - F[ii].re;
        D[ii].im = R[ii].im / g - F[ii].im;
      }
    }
    free(work);
  }
  return 0;
}

/* P: Ps[ii] = S[m] + cos2psi S[npix+m] + sin2psi S[2npix+m]  (2845-2873,
 * polar branch with polard = 1, polara = 0, no GalBP). */
void or_deproject(const double *S, const long *indpix, const long *datatopix,
                  const double *psip, long ns, long npix, double *Ps,
                  int galbp, double rfgal) {
  for (long ii = 0; ii < ns; ii++) {
    long mpos = indpix[datatopix[ii]];
    if (mpos < 0)
      mpos = -mpos - 2;
    Ps[ii] = 0.0;
    Ps[ii] = S[mpos] + 1.0 * (S[npix + mpos] * cos(2.0 * (psip[ii] + 0.0)) +
                              S[2 * npix + mpos] * sin(2.0 * (psip[ii] + 0.0)));
    if (galbp)
      Ps[ii] += S[3 * npix + mpos] * rfgal;
  }
}

/* Scatter of one (det, seg) block (2584-2611).  If sumCross != NULL the
 * crosslink sums replace the map accumulation, exactly as the reference. */
static void or_scatter(const or_problem *P, const double *Nd, long idet,
                       long iseg, const double *psiarr, double *PNd,
                       double *sumCross) {
  long ns = P->ns, npix = P->npix;
  for (long ii = 0; ii < ns; ii++) {
    long ppos = IDX(P, idet, iseg, ii);
    long mpos = P->indpix[P->datatopix[ppos]];
    if (mpos >= 0) {
      if (sumCross != NULL) {
        sumCross[mpos] += cos(2.0 * (psiarr[ppos] + 0.0));
        sumCross[npix + mpos] += sin(2.0 * (psiarr[ppos] + 0.0));
      } else {
        PNd[mpos] += Nd[ii];
        PNd[npix + mpos] += 1.0 * cos(2.0 * (psiarr[ppos] + 0.0)) * Nd[ii];
        PNd[2 * npix + mpos] += 1.0 * sin(2.0 * (psiarr[ppos] + 0.0)) * Nd[ii];
        if (P->galbp)
          PNd[3 * npix + mpos] += P->relfgal[idet] * Nd[ii];
      }
    }
  }
}

/* compute_diagPtNPCorr (2680-2762) for one (det, seg) block: N_ = c2r(|Nk|),
 * then for every pair of samples in the same pixel with |dt| < ns/2 add
 * N_[|dt|].  Samples are grouped by pixel in time order (the reference's
 * qsort leaves the tie order unspecified; only the summation order differs). */
typedef struct {
  long pix, t;
} or_pt;

static int or_pt_cmp(const void *a, const void *b) {
  const or_pt *x = (const or_pt *)a, *y = (const or_pt *)b;
  if (x->pix != y->pix)
    return (x->pix > y->pix) - (x->pix < y->pix);
  return (x->t > y->t) - (x->t < y->t);
}

static void or_diag_block(const or_problem *P, long idet, long iseg,
                          double *dPtNP, double *N_, cplx *Nk_, cplx *work,
                          or_pt *pts) {
  long ns = P->ns, nh = ns / 2 + 1;
  const double *Nk = P->spn + idet * ns;
  for (long k = 0; k < nh; k++) {
    Nk_[k].re = fabs(Nk[k]);
    Nk_[k].im = 0.0;
  }
  or_rfft_c2r(P->plan, Nk_, N_, work);
  for (long ii = 0; ii < ns; ii++) {
    pts[ii].pix = P->indpix[P->datatopix[IDX(P, idet, iseg, ii)]];
    pts[ii].t = ii;
  }
  qsort(pts, ns, sizeof(or_pt), or_pt_cmp);
  long c0 = 0;
  while (c0 < ns) {
    long c1 = c0;
    while (c1 < ns && pts[c1].pix == pts[c0].pix)
      c1++;
    if (pts[c0].pix >= 0) {
      double acc = dPtNP[pts[c0].pix];
      for (long i = c0; i < c1; i++) {
        long i2 = pts[i].t;
        for (long k = c0; k < c1; k++) {
          long k2 = pts[k].t;
          if (labs(k2 - i2) < ns / 2)
            acc += N_[labs(i2 - k2)];
        }
      }
      dPtNP[pts[c0].pix] = acc;
    }
    c0 = c1;
  }
}

/* One full pass over all (det, seg) blocks of PtN^-1 applied to fdatas-like
 * spectra F (mode 0: F = r2c(P S); mode 1: F = fdatas0 (data)).
 * acc: 3 npix output (added to).  Per-thread private maps are summed in
 * thread order, so results are deterministic for a given thread count. */
static void or_pass(or_problem *P, const double *S, int from_data,
                    double *acc, const double *psiarr_scatter,
                    double *sumCross) {
  long ns = P->ns, nh = ns / 2 + 1, npix = P->npix, ntot3 = P->nc * npix;
  int nt = P->nthreads;
  double *priv = (double *)calloc((size_t)nt * ntot3, sizeof(double));
  long nblk = P->ndet * P->nseg;
#pragma omp parallel num_threads(nt)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double *mine = priv + (size_t)tid * ntot3;
    double *Ps = (double *)malloc(sizeof(double) * ns);
    double *Nd = (double *)malloc(sizeof(double) * ns);
    cplx *fdata = (cplx *)malloc(sizeof(cplx) * nh);
    cplx *work = (cplx *)malloc(sizeof(cplx) * 2 * ns);
    /* the reference loops segments outer, detectors inner (1429-1439) */
#pragma omp for schedule(static)
    for (long blk = 0; blk < nblk; blk++) {
      long iseg = blk / P->ndet, idet = blk % P->ndet;
      const double *Nk = P->spn + idet * ns;
      const double g = or_gain(P, idet, iseg);
      if (sumCross != NULL) {
        /* crosslink pass: the spectra are not used */
      } else if (from_data) {
        memcpy(fdata, P->fdatas0 + (idet * P->nseg + iseg) * nh,
               sizeof(cplx) * nh);
      } else {
        or_deproject(S, P->indpix, P->datatopix + IDX(P, idet, iseg, 0),
                     P->psip + IDX(P, idet, iseg, 0), ns, npix, Ps, P->galbp,
                     P->relfgal[idet]);
        or_rfft_r2c(P->plan, Ps, fdata, work);
      }
      if (sumCross == NULL) {
        for (long ii = 0; ii < nh; ii++) {
          double nk = Nk[ii] * g * g; /* SpN relCalib^2 (2543-2545) */
          fdata[ii].re = 0 + fdata[ii].re * nk;
          fdata[ii].im = 0 + fdata[ii].im * nk;
        }
        or_rfft_c2r(P->plan, fdata, Nd, work);
        or_scatter(P, Nd, idet, iseg, psiarr_scatter, mine, NULL);
      } else {
        or_scatter(P, NULL, idet, iseg, psiarr_scatter, NULL, mine);
      }
    }
    free(Ps);
    free(Nd);
    free(fdata);
    free(work);
  }
  for (long i = 0; i < ntot3; i++) {
    double s = acc[i];
    for (int t = 0; t < nt; t++)
      s += priv[(size_t)t * ntot3 + i];
    acc[i] = s;
  }
  free(priv);
}

/* y = P^T N^-1 P S (unmasked), S and y compressed block layout [I|Q|U]. */
void or_apply_A(or_problem *P, const double *S, double *y) {
  memset(y, 0, sizeof(double) * P->nc * P->npix);
  or_pass(P, S, 0, y, P->psip, NULL);
}

/* b = P^T N^-1 d (1069-1125). */
void or_rhs(or_problem *P, double *b) {
  memset(b, 0, sizeof(double) * P->nc * P->npix);
  or_pass(P, NULL, 1, b, P->psip, NULL);
}

/* Diagonal preconditioner accumulation for a single pass over all blocks
 * (I component only; the reference then sets Q = U = I/2). */
void or_diag_corr(or_problem *P, double *MpI) {
  long ns = P->ns, nh = ns / 2 + 1, npix = P->npix;
  int nt = P->nthreads;
  long nblk = P->ndet * P->nseg;
  double *priv = (double *)calloc((size_t)nt * npix, sizeof(double));
#pragma omp parallel num_threads(nt)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double *N_ = (double *)malloc(sizeof(double) * ns);
    cplx *Nk_ = (cplx *)malloc(sizeof(cplx) * nh);
    cplx *work = (cplx *)malloc(sizeof(cplx) * 2 * ns);
    or_pt *pts = (or_pt *)malloc(sizeof(or_pt) * ns);
#pragma omp for schedule(static)
    for (long blk = 0; blk < nblk; blk++) {
      long iseg = blk / P->ndet, idet = blk % P->ndet;
      or_diag_block(P, idet, iseg, priv + (size_t)tid * npix, N_, Nk_, work,
                    pts);
    }
    free(N_);
    free(Nk_);
    free(work);
    free(pts);
  }
  for (long i = 0; i < npix; i++) {
    double s = 0.0;
    for (int t = 0; t < nt; t++)
      s += priv[(size_t)t * npix + i];
    MpI[i] = s;
  }
  free(priv);
}

/* chi2 = sum over blocks of 2 |(fdatas0 - r2c(P S)) sqrt(SpN) relCalib|^2
 * over bins 0..ns/2 (1514-1538). */
double or_chi2(or_problem *P, const double *S) {
  long ns = P->ns, nh = ns / 2 + 1, npix = P->npix;
  long nblk = P->ndet * P->nseg;
  int nt = P->nthreads;
  double *part = (double *)calloc(nblk, sizeof(double));
#pragma omp parallel num_threads(nt)
  {
    double *Ps = (double *)malloc(sizeof(double) * ns);
    cplx *fdata = (cplx *)malloc(sizeof(cplx) * nh);
    cplx *work = (cplx *)malloc(sizeof(cplx) * 2 * ns);
#pragma omp for schedule(static)
    for (long blk = 0; blk < nblk; blk++) {
      long iseg = blk / P->ndet, idet = blk % P->ndet;
      const double *Nk = P->spn + idet * ns;
      const cplx *D = P->fdatas0 + (idet * P->nseg + iseg) * nh;
      const double g = or_gain(P, idet, iseg);
      or_deproject(S, P->indpix, P->datatopix + IDX(P, idet, iseg, 0),
                   P->psip + IDX(P, idet, iseg, 0), ns, npix, Ps, P->galbp,
                   P->relfgal[idet]);
      or_rfft_r2c(P->plan, Ps, fdata, work);
      double c = 0.0;
      for (long ii = 0; ii < nh; ii++) {
        double sq = sqrt(Nk[ii]) * g;
        double re = (D[ii].re - fdata[ii].re) * sq;
        double im = (D[ii].im - fdata[ii].im) * sq;
        c += (re * re + im * im) / 1.0 * 2.0;
      }
      part[blk] = c;
    }
    free(Ps);
    free(fdata);
    free(work);
  }
  /* serial sum in the reference's block order (segments outer) */
  double chi2 = 0.0;
  for (long blk = 0; blk < nblk; blk++)
    chi2 += part[blk];
  free(part);
  return chi2;
}

/* Relative calibration relCalib of every (det, seg) block, g: [ndet][nseg]
 * (NULL: all 1).  The data spectrum of the next solve is fraw / g - fodipole
 * (1044-1067, with the block index of 2128-2132: the reference's division at
 * 1050-1052 indexes relCalib[idet + iseg + iseg_min ndet] instead), and
 * N^-1 of the block is SpN g^2 in the operator and b (do_PtNd 2543-2545) and
 * in chi2 (1524-1530).  The preconditioner stays the one of unit gains, as
 * the reference builds Mp in its first loop only (1178-1199, 1352). */
int or_problem_set_gains(or_problem *P, const double *g) {
  long ns = P->ns, nh = ns / 2 + 1, nblk = P->ndet * P->nseg;
  if (!P->fraw)
    return -1;
  if (!g) {
    free(P->gain);
    P->gain = NULL;
  } else {
    if (!P->gain)
      P->gain = (double *)malloc(sizeof(double) * nblk);
    memcpy(P->gain, g, sizeof(double) * nblk);
  }
  for (long blk = 0; blk < nblk; blk++) {
    const double gg = P->gain ? P->gain[blk] : 1.0;
    const cplx *R = P->fraw + blk * nh;
    const cplx *F = P->fodipole ? P->fodipole + blk * nh : NULL;
    cplx *D = P->fdatas0 + blk * nh;
    for (long ii = 0; ii < nh; ii++) {
      D[ii].re = R[ii].re / gg;
      D[ii].im = R[ii].im / gg;
      if (F) {
        D[ii].re -= F[ii].re;
        D[ii].im -= F[ii].im;
      }
    }
  }
  return 0;
}

/* The gain fit after a solve (2185-2228): for every (det, seg) block, with
 * C = r2c(P S) + fodipole (the model with the dipole added back, 2116-2122)
 * and D = fraw (the data with the dipole and relCalib put back, 2116-2133),
 *   out[3 blk + 0] = sum_{ii <= ns/2} SpN (Re D Re C + Im D Im C)   (zeta)
 *   out[3 blk + 1] = sum_{ii <= ns/2} SpN |C|^2                     (gamma)
 *   out[3 blk + 2] = sum_{ii <= ns/2} SpN |D|^2
 * blk = idet nseg + iseg.  The gain that minimises the reference's chi2
 * (3206-3255) for that block is out[0] / out[1]; the reference's own ratio
 * sPNd / cPNPc carries a factor 2 in sPNd and 1/2 in cPNPc (2214, 2218-2220),
 * 4 times that minimiser, and scales the map by absrecal, which is read
 * from zeta_all / gamma_all (2282-2291) that only -R sets.  These sums are
 * the parts of that fit with both factors equal and absrecal = 1. */
void or_recal_sums(or_problem *P, const double *S, double *out) {
  long ns = P->ns, nh = ns / 2 + 1, npix = P->npix;
  long nblk = P->ndet * P->nseg;
#pragma omp parallel num_threads(P->nthreads)
  {
    double *Ps = (double *)malloc(sizeof(double) * ns);
    cplx *C = (cplx *)malloc(sizeof(cplx) * nh);
    cplx *work = (cplx *)malloc(sizeof(cplx) * 2 * ns);
#pragma omp for schedule(static)
    for (long blk = 0; blk < nblk; blk++) {
      long idet = blk / P->nseg, iseg = blk % P->nseg;
      const double *Nk = P->spn + idet * ns;
      const cplx *D = P->fraw + blk * nh;
      const cplx *F = P->fodipole ? P->fodipole + blk * nh : NULL;
      or_deproject(S, P->indpix, P->datatopix + IDX(P, idet, iseg, 0),
                   P->psip + IDX(P, idet, iseg, 0), ns, npix, Ps, P->galbp,
                   P->relfgal[idet]);
      or_rfft_r2c(P->plan, Ps, C, work);
      double zeta = 0.0, gamma = 0.0, dd = 0.0;
      for (long ii = 0; ii < nh; ii++) {
        double cr = C[ii].re, ci = C[ii].im;
        if (F) {
          cr += F[ii].re;
          ci += F[ii].im;
        }
        zeta += (D[ii].re * cr + D[ii].im * ci) * Nk[ii];
        gamma += (cr * cr + ci * ci) * Nk[ii];
        dd += (D[ii].re * D[ii].re + D[ii].im * D[ii].im) * Nk[ii];
      }
      out[3 * blk + 0] = zeta;
      out[3 * blk + 1] = gamma;
      out[3 * blk + 2] = dd;
    }
    free(Ps);
    free(C);
    free(work);
  }
}

/* MPI_compute_chi2 (3206-3255): sum over blocks of
 * 2 |(fraw - g (r2c(P S) + fodipole)) sqrt(SpN)|^2 over bins 0..ns/2, with
 * the gains g: [ndet][nseg] (NULL: 1).  Blocks are summed in the reference's
 * order (segments outer, 3219). */
double or_recal_chi2(or_problem *P, const double *S, const double *g) {
  long ns = P->ns, nh = ns / 2 + 1, npix = P->npix;
  long nblk = P->ndet * P->nseg;
  double *part = (double *)calloc(nblk, sizeof(double));
#pragma omp parallel num_threads(P->nthreads)
  {
    double *Ps = (double *)malloc(sizeof(double) * ns);
    cplx *C = (cplx *)malloc(sizeof(cplx) * nh);
    cplx *work = (cplx *)malloc(sizeof(cplx) * 2 * ns);
#pragma omp for schedule(static)
    for (long blk = 0; blk < nblk; blk++) {
      long iseg = blk / P->ndet, idet = blk % P->ndet;
      long b = idet * P->nseg + iseg;
      const double *Nk = P->spn + idet * ns;
      const cplx *D = P->fraw + b * nh;
      const cplx *F = P->fodipole ? P->fodipole + b * nh : NULL;
      const double gg = g ? g[b] : 1.0;
      or_deproject(S, P->indpix, P->datatopix + IDX(P, idet, iseg, 0),
                   P->psip + IDX(P, idet, iseg, 0), ns, npix, Ps, P->galbp,
                   P->relfgal[idet]);
      or_rfft_r2c(P->plan, Ps, C, work);
      double c = 0.0;
      for (long ii = 0; ii < nh; ii++) {
        double sq = sqrt(Nk[ii]);
        double re = (D[ii].re - gg * (C[ii].re + (F ? F[ii].re : 0.0))) * sq;
        double im = (D[ii].im - gg * (C[ii].im + (F ? F[ii].im : 0.0))) * sq;
        c += (re * re + im * im) / 1.0 * 2.0;
      }
      part[blk] = c;
    }
    free(Ps);
    free(C);
    free(work);
  }
  double chi2 = 0.0;
  for (long blk = 0; blk < nblk; blk++)
    chi2 += part[blk];
  free(part);
  return chi2;
}

/* Setup pass of the reference (1177-1314): sumCross via psipFP, Mp (twice,
 * from both do_PtNd calls), hits (twice), and A x0.  Writes the unmasked
 * products into the problem and Ax0. */
void or_setup_pass(or_problem *P, const double *x0, double *Ax0) {
  long npix = P->npix, ntot = P->ndet * P->nseg * P->ns, nc = P->nc;
  free(P->hits);
  free(P->Mp);
  free(P->sumCross);
  P->hits = (long *)calloc(2 * npix, sizeof(long));
  P->Mp = (double *)calloc(nc * npix, sizeof(double));
  P->sumCross = (double *)calloc(nc * npix, sizeof(double));
  /* crosslink sums (first do_PtNd call, 1258: sumCross != NULL) */
  or_pass(P, NULL, 0, P->sumCross, P->psipFP, P->sumCross);
  /* A x0 (second call, 1263) */
  double *Sx = (double *)calloc(nc * npix, sizeof(double));
  if (x0)
    memcpy(Sx, x0, sizeof(double) * nc * npix);
  memset(Ax0, 0, sizeof(double) * nc * npix);
  or_pass(P, Sx, 0, Ax0, P->psip, NULL);
  free(Sx);
  /* Mp: both calls accumulate the I pair sums; Q = U = I/2 afterwards */
  double *MpI = (double *)calloc(npix, sizeof(double));
  or_diag_corr(P, MpI);
  for (long i = 0; i < npix; i++) {
    P->Mp[i] = 2.0 * MpI[i];
    P->Mp[npix + i] = P->Mp[i] / 2.0;
    P->Mp[2 * npix + i] = P->Mp[i] / 2.0;
    /* the leakage entry scales I by the FIRST detector's coefficient
     * squared (2565-2566) */
    if (P->galbp)
      P->Mp[3 * npix + i] = P->Mp[i] * P->relfgal[0] * P->relfgal[0];
  }
  free(MpI);
  /* hits: both calls, polardet = 1 > 0.5 so pol hits count too (2613-2626) */
  for (long k = 0; k < ntot; k++) {
    long m = P->indpix[P->datatopix[k]];
    if (m >= 0) {
      P->hits[m] += 2;
      P->hits[npix + m] += 2;
    }
  }
}

void or_problem_get_setup(const or_problem *P, long *hits, double *Mp,
                          double *sumCross) {
  long npix = P->npix;
  if (hits)
    memcpy(hits, P->hits, sizeof(long) * 2 * npix);
  if (Mp)
    memcpy(Mp, P->Mp, sizeof(double) * P->nc * npix);
  if (sumCross)
    memcpy(sumCross, P->sumCross, sizeof(double) * P->nc * npix);
}

/* The reference's solve (1069-1761) with its masks, diagonal preconditioner,
 * residual refresh every `refresh` iterations and stop rule
 * while ((iter < maxiter && var_n/var00 > stop) || iter == 0).
 * Outputs: S (3 npix), Mptot (inverted, masked), masked sumCross, hits,
 * per-iteration crit, crit2, chi2.  Returns the iteration count, or -1 when
 * an I-component of Mp is zero (the reference exits there, 1344-1348). */
int or_problem_solve(or_problem *P, const double *x0, int maxiter, double stop,
                     int refresh, double *S_out, double *Mptot_out,
                     double *crit_out, double *crit2_out, double *chi2_out,
                     double *b_out) {
  long npix = P->npix, n3 = P->nc * npix;
  double *b = (double *)calloc(n3, sizeof(double));
  or_rhs(P, b);
  double *S = (double *)calloc(n3, sizeof(double));
  if (x0)
    memcpy(S, x0, sizeof(double) * n3);
  double *r = (double *)calloc(n3, sizeof(double));
  or_setup_pass(P, S, r); /* r = A x0 (unmasked) */
  double *Mpt = (double *)malloc(sizeof(double) * n3);
  memcpy(Mpt, P->Mp, sizeof(double) * n3);
  for (long ii = npix; ii < n3; ii++)
    if (ii < 3 * npix && P->hits[(ii % npix) + npix] < 4) {
      r[ii] = 0.0;
      S[ii] = 0.0;
      b[ii] = 0.0;
      Mpt[ii] = 0.0;
      P->sumCross[ii] = 0.0;
    }
  for (long ii = 0; ii < npix; ii++)
    if (Mpt[ii] == 0) {
      free(b);
      free(S);
      free(r);
      free(Mpt);
      return -1;
    }
  double *d = (double *)malloc(sizeof(double) * n3);
  double *q = (double *)malloc(sizeof(double) * n3);
  for (long ii = 0; ii < n3; ii++) {
    if (Mpt[ii] != 0.0)
      Mpt[ii] = 1.0 / Mpt[ii];
    r[ii] = b[ii] - r[ii];
    d[ii] = Mpt[ii] * r[ii];
  }
  double delta_n = 0.0, var_n = 0.0, var00 = 0.0;
  for (long ii = 0; ii < n3; ii++)
    delta_n += r[ii] * d[ii];
  for (long ii = 0; ii < n3; ii++)
    var_n += r[ii] * r[ii];
  for (long ii = 0; ii < n3; ii++)
    var00 += b[ii] * b[ii];
  double delta0 = delta_n, chi2 = 0.0;
  int iter = 0;
  while ((iter < maxiter && var_n / var00 > stop) || iter == 0) {
    or_apply_A(P, d, q);
    for (long ii = npix; ii < n3; ii++)
      if (P->hits[(ii % npix) + npix] < 4)
        q[ii] = 0.0;
    double rtq = 0.0;
    for (long ii = 0; ii < n3; ii++)
      rtq += q[ii] * d[ii];
    double alpha = delta_n / rtq;
    for (long ii = 0; ii < n3; ii++)
      S[ii] += alpha * d[ii];
    if ((iter % refresh) == 0) {
      chi2 = or_chi2(P, S);
      or_apply_A(P, S, r);
      for (long ii = npix; ii < n3; ii++)
        if (P->hits[(ii % npix) + npix] < 4)
          r[ii] = 0.0;
      for (long ii = 0; ii < n3; ii++)
        r[ii] = b[ii] - r[ii];
    } else {
      for (long ii = 0; ii < n3; ii++)
        r[ii] -= alpha * q[ii];
    }
    double delta_o = delta_n;
    delta_n = 0.0;
    for (long ii = 0; ii < n3; ii++)
      delta_n += r[ii] * Mpt[ii] * r[ii];
    var_n = 0.0;
    for (long ii = 0; ii < n3; ii++)
      var_n += r[ii] * r[ii];
    double beta = delta_n / delta_o;
    for (long ii = 0; ii < n3; ii++)
      d[ii] = Mpt[ii] * r[ii] + beta * d[ii];
    if (crit_out)
      crit_out[iter] = var_n / var00;
    if (crit2_out)
      crit2_out[iter] = delta_n / delta0;
    if (chi2_out)
      chi2_out[iter] = chi2;
    iter++;
  }
  if (S_out)
    memcpy(S_out, S, sizeof(double) * n3);
  if (Mptot_out)
    memcpy(Mptot_out, Mpt, sizeof(double) * n3);
  if (b_out)
    memcpy(b_out, b, sizeof(double) * n3);
  free(b);
  free(S);
  free(r);
  free(Mpt);
  free(d);
  free(q);
  return iter;
}

/* ======================================================================== */
/* Exact 3x3 diagonal blocks and the block-Jacobi / A-DEF2 PCG               */
/* ======================================================================== */

/* blocks[npix][6] = {B00, B01, B02, B11, B12, B22}: the diagonal 3x3 block
 * of A = P^T N^-1 P at each seen pixel, unmasked.  N^-1 of a (det, seg)
 * block is the circulant c2r(SpN r2c(.)) (2549-2581), i.e. the matrix
 * kern[(i - j) mod ns] with kern = c2r(SpN[0..ns/2]) (unnormalised); the
 * block is the brute-force sum over every ordered pair of samples (i, j) of
 * the pixel in one (det, seg) block of kern[(t_i - t_j) mod ns] w_i w_j^T,
 * w = (1, cos 2psi, sin 2psi) -- all lags, not the reference's |dt| < ns/2
 * truncation of compute_diagPtNPCorr (2680-2762), which builds only the I
 * entry.  O(sum of squared hits per block): test sizes only.  Pinned by
 * tests/test_oracle_blocks.py against columns of or_apply_A. */
void or_exact_blocks(or_problem *P, double *blocks) {
  long ns = P->ns, nh = ns / 2 + 1, npix = P->npix;
  int nt = P->nthreads;
  long nblk = P->ndet * P->nseg;
  double *kern = (double *)malloc(sizeof(double) * ns * P->ndet);
  {
    cplx *Nk_ = (cplx *)malloc(sizeof(cplx) * nh);
    cplx *work = (cplx *)malloc(sizeof(cplx) * 2 * ns);
    for (long idet = 0; idet < P->ndet; idet++) {
      for (long k = 0; k < nh; k++) {
        Nk_[k].re = P->spn[idet * ns + k];
        Nk_[k].im = 0.0;
      }
      or_rfft_c2r(P->plan, Nk_, kern + idet * ns, work);
    }
    free(Nk_);
    free(work);
  }
  double *priv = (double *)calloc((size_t)nt * 6 * npix, sizeof(double));
#pragma omp parallel num_threads(nt)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double *mine = priv + (size_t)tid * 6 * npix;
    or_pt *pts = (or_pt *)malloc(sizeof(or_pt) * ns);
#pragma omp for schedule(dynamic)
    for (long blk = 0; blk < nblk; blk++) {
      long idet = blk / P->nseg, iseg = blk % P->nseg;
      const double *kr = kern + idet * ns;
      for (long ii = 0; ii < ns; ii++) {
        pts[ii].pix = P->indpix[P->datatopix[IDX(P, idet, iseg, ii)]];
        pts[ii].t = ii;
      }
      qsort(pts, ns, sizeof(or_pt), or_pt_cmp);
      long c0 = 0;
      while (c0 < ns) {
        long c1 = c0;
        while (c1 < ns && pts[c1].pix == pts[c0].pix)
          c1++;
        long p = pts[c0].pix;
        if (p >= 0) {
          double acc[6] = {0, 0, 0, 0, 0, 0};
          for (long i = c0; i < c1; i++) {
            double ai = 2.0 * P->psip[IDX(P, idet, iseg, pts[i].t)];
            double wi[3] = {1.0, cos(ai), sin(ai)};
            for (long j = c0; j < c1; j++) {
              double aj = 2.0 * P->psip[IDX(P, idet, iseg, pts[j].t)];
              long lag = ((pts[i].t - pts[j].t) % ns + ns) % ns;
              double kv = kr[lag];
              double wj[3] = {1.0, cos(aj), sin(aj)};
              acc[0] += kv * wi[0] * wj[0];
              acc[1] += kv * wi[0] * wj[1];
              acc[2] += kv * wi[0] * wj[2];
              acc[3] += kv * wi[1] * wj[1];
              acc[4] += kv * wi[1] * wj[2];
              acc[5] += kv * wi[2] * wj[2];
            }
          }
          for (int k = 0; k < 6; k++)
            mine[6 * p + k] += acc[k];
        }
        c0 = c1;
      }
    }
    free(pts);
  }
  for (long i = 0; i < 6 * npix; i++) {
    double s = 0.0;
    for (int t = 0; t < nt; t++)
      s += priv[(size_t)t * 6 * npix + i];
    blocks[i] = s;
  }
  free(priv);
  free(kern);
}

/* Inverse of a symmetric 3x3 block by Gauss-Jordan elimination with partial
 * pivoting (m6 in, inv6 out, both {00, 01, 02, 11, 12, 22}). */
static void or_inv3(const double *m6, double *inv6) {
  double a[3][6] = {{m6[0], m6[1], m6[2], 1, 0, 0},
                    {m6[1], m6[3], m6[4], 0, 1, 0},
                    {m6[2], m6[4], m6[5], 0, 0, 1}};
  for (int c = 0; c < 3; c++) {
    int piv = c;
    for (int r = c + 1; r < 3; r++)
      if (fabs(a[r][c]) > fabs(a[piv][c]))
        piv = r;
    if (piv != c)
      for (int k = 0; k < 6; k++) {
        double t = a[c][k];
        a[c][k] = a[piv][k];
        a[piv][k] = t;
      }
    double d = a[c][c];
    for (int k = 0; k < 6; k++)
      a[c][k] /= d;
    for (int r = 0; r < 3; r++)
      if (r != c) {
        double f = a[r][c];
        for (int k = 0; k < 6; k++)
          a[r][k] -= f * a[c][k];
      }
  }
  inv6[0] = a[0][3];
  inv6[1] = 0.5 * (a[0][4] + a[1][3]);
  inv6[2] = 0.5 * (a[0][5] + a[2][3]);
  inv6[3] = a[1][4];
  inv6[4] = 0.5 * (a[1][5] + a[2][4]);
  inv6[5] = a[2][5];
}

/* Preconditioner of the GPU solver for one pixel (no counterpart in the
 * reference): a pixel whose Q/U are masked (pol hits < 4, doubled) keeps
 * 1/B00 on I only; a block that is not positive on its diagonal or whose
 * determinant is below 1e-10 B00 B11 B22 (one distinct angle) falls back to
 * the inverse diagonal; otherwise the inverse block. */
static void or_block_precond(const double *b6, int masked, double *m6) {
  for (int k = 0; k < 6; k++)
    m6[k] = 0.0;
  if (masked) {
    m6[0] = b6[0] != 0.0 ? 1.0 / b6[0] : 0.0;
    return;
  }
  double det = b6[0] * (b6[3] * b6[5] - b6[4] * b6[4]) -
               b6[1] * (b6[1] * b6[5] - b6[4] * b6[2]) +
               b6[2] * (b6[1] * b6[4] - b6[3] * b6[2]);
  double scale = b6[0] * b6[3] * b6[5];
  if (!(b6[0] > 0.0 && b6[3] > 0.0 && b6[5] > 0.0) || !(det > 1e-10 * scale)) {
    m6[0] = b6[0] > 0.0 ? 1.0 / b6[0] : 0.0;
    m6[3] = b6[3] > 0.0 ? 1.0 / b6[3] : 0.0;
    m6[5] = b6[5] > 0.0 ? 1.0 / b6[5] : 0.0;
    return;
  }
  or_inv3(b6, m6);
}

/* z = M r in block layout. */
static void or_apply_M(long npix, long n, int kind, const double *Mdiag,
                       const double *M6, const double *r, double *z) {
  if (kind == 0) {
    for (long i = 0; i < n; i++)
      z[i] = Mdiag[i] * r[i];
    return;
  }
  for (long p = 0; p < npix; p++) {
    const double *m = M6 + 6 * p;
    double r0 = r[p], r1 = r[npix + p], r2 = r[2 * npix + p];
    z[p] = m[0] * r0 + m[1] * r1 + m[2] * r2;
    z[npix + p] = m[1] * r0 + m[3] * r1 + m[4] * r2;
    z[2 * npix + p] = m[2] * r0 + m[4] * r1 + m[5] * r2;
  }
}

static void or_mask_qu(const or_problem *P, double *v) {
  long npix = P->npix;
  for (long p = 0; p < npix; p++)
    if (P->hits[npix + p] < 4) {
      v[npix + p] = 0.0;
      v[2 * npix + p] = 0.0;
    }
}

/* PCG of the GPU solver, for kind 0 (the reference's diagonal Mp,
 * 1351-1353) or kind 1 (per-pixel 3x3 block-Jacobi from or_exact_blocks),
 * optionally with the two-level A-DEF2 preconditioner of Tang, Nabben, Vuik
 * and Erlangga (J. Sci. Comput. 39, 2009) with the single coarse vector
 * Z = e_I (1 on every I component): E = Z^T A Z,
 *   z = M r + Z E^-1 (Z^T r - (A Z)^T M r),
 * and the starting guess corrected by Z E^-1 Z^T (b - A x0).  The loop, the
 * masks (1316-1330, 1455-1459, 1559-1563), the residual refresh every
 * `refresh` iterations and the stop rule are those of or_problem_solve
 * (1415-1761); var00_in > 0 replaces |b|^2 in crit.  Returns the iteration
 * count, or -1 when the reference's diagonal Mp has a zero I element
 * (1344-1348). */
int or_problem_solve_pcg(or_problem *P, const double *x0, int maxiter,
                         double stop, int refresh, int kind, int deflate,
                         double var00_in, double *S_out, double *crit_out,
                         double *crit2_out, double *b_out, double *M6_out) {
  long npix = P->npix, n3 = P->nc * npix;
  if (kind == 1 && P->galbp)
    return -2; /* no 3x3 blocks for the 4-component map */
  double *b = (double *)calloc(n3, sizeof(double));
  or_rhs(P, b);
  double *S = (double *)calloc(n3, sizeof(double));
  if (x0)
    memcpy(S, x0, sizeof(double) * n3);
  double *q = (double *)calloc(n3, sizeof(double));
  or_setup_pass(P, S, q); /* q = A x0 (unmasked), hits, Mp */
  or_mask_qu(P, q);
  or_mask_qu(P, S);
  or_mask_qu(P, b);
  for (long ii = 0; ii < npix; ii++)
    if (P->Mp[ii] == 0) {
      free(b);
      free(S);
      free(q);
      return -1;
    }
  double *Md = (double *)malloc(sizeof(double) * n3);
  double *M6 = NULL;
  for (long ii = 0; ii < n3; ii++)
    Md[ii] = P->Mp[ii];
  or_mask_qu(P, Md);
  for (long ii = 0; ii < n3; ii++)
    Md[ii] = Md[ii] != 0.0 ? 1.0 / Md[ii] : 0.0;
  if (kind == 1) {
    double *B6 = (double *)malloc(sizeof(double) * 6 * npix);
    or_exact_blocks(P, B6);
    M6 = (double *)malloc(sizeof(double) * 6 * npix);
    for (long p = 0; p < npix; p++)
      or_block_precond(B6 + 6 * p, P->hits[npix + p] < 4, M6 + 6 * p);
    free(B6);
    if (M6_out)
      memcpy(M6_out, M6, sizeof(double) * 6 * npix);
  }
  double *AZ = NULL, einv = 0.0;
  if (deflate) {
    double *Z = (double *)calloc(n3, sizeof(double));
    AZ = (double *)malloc(sizeof(double) * n3);
    for (long p = 0; p < npix; p++)
      Z[p] = 1.0;
    or_apply_A(P, Z, AZ);
    or_mask_qu(P, AZ);
    double e = 0.0, g = 0.0;
    for (long p = 0; p < npix; p++)
      e += AZ[p];
    einv = e != 0.0 ? 1.0 / e : 0.0;
    for (long p = 0; p < npix; p++)
      g += b[p] - q[p];
    double y0 = einv * g;
    for (long p = 0; p < npix; p++)
      S[p] += y0;
    for (long ii = 0; ii < n3; ii++)
      q[ii] += y0 * AZ[ii];
    free(Z);
  }
  double *r = (double *)malloc(sizeof(double) * n3);
  double *z = (double *)malloc(sizeof(double) * n3);
  double *d = (double *)malloc(sizeof(double) * n3);
  for (long ii = 0; ii < n3; ii++)
    r[ii] = b[ii] - q[ii];
  /* z = M r + c e_I, delta = r.z; returns c */
  double delta_n, var_n = 0.0, var00 = 0.0;
#define OR_PRECOND_STEP(DELTA)                                     \
  do {                                                             \
    or_apply_M(npix, n3, kind, Md, M6, r, z);                          \
    double rz = 0.0, zr_I = 0.0, azz = 0.0;                        \
    for (long ii = 0; ii < n3; ii++)                               \
      rz += r[ii] * z[ii];                                         \
    double c = 0.0;                                                \
    if (deflate) {                                                 \
      for (long p = 0; p < npix; p++)                              \
        zr_I += r[p];                                              \
      for (long ii = 0; ii < n3; ii++)                             \
        azz += AZ[ii] * z[ii];                                     \
      c = einv * (zr_I - azz);                                     \
      for (long p = 0; p < npix; p++)                              \
        z[p] += c;                                                 \
    }                                                              \
    DELTA = rz + c * zr_I;                                         \
  } while (0)
  OR_PRECOND_STEP(delta_n);
  for (long ii = 0; ii < n3; ii++) {
    d[ii] = z[ii];
    var_n += r[ii] * r[ii];
    var00 += b[ii] * b[ii];
  }
  if (var00_in > 0.0) /* the recalibration loop keeps rloop 1's |b|^2 (1377-1385) */
    var00 = var00_in;
  double delta0 = delta_n;
  int iter = 0;
  while ((iter < maxiter && var_n / var00 > stop) || iter == 0) {
    or_apply_A(P, d, q);
    or_mask_qu(P, q);
    double rtq = 0.0;
    for (long ii = 0; ii < n3; ii++)
      rtq += q[ii] * d[ii];
    double alpha = delta_n / rtq;
    for (long ii = 0; ii < n3; ii++)
      S[ii] += alpha * d[ii];
    if ((iter % refresh) == 0) {
      or_apply_A(P, S, r);
      or_mask_qu(P, r);
      for (long ii = 0; ii < n3; ii++)
        r[ii] = b[ii] - r[ii];
    } else {
      for (long ii = 0; ii < n3; ii++)
        r[ii] -= alpha * q[ii];
    }
    double delta_o = delta_n;
    OR_PRECOND_STEP(delta_n);
    var_n = 0.0;
    for (long ii = 0; ii < n3; ii++)
      var_n += r[ii] * r[ii];
    double beta = delta_n / delta_o;
    for (long ii = 0; ii < n3; ii++)
      d[ii] = z[ii] + beta * d[ii];
    if (crit_out)
      crit_out[iter] = var_n / var00;
    if (crit2_out)
      crit2_out[iter] = delta_n / delta0;
    iter++;
  }
#undef OR_PRECOND_STEP
  if (S_out)
    memcpy(S_out, S, sizeof(double) * n3);
  if (b_out)
    memcpy(b_out, b, sizeof(double) * n3);
  free(b);
  free(S);
  free(q);
  free(r);
  free(z);
  free(d);
  free(Md);
  free(M6);
  free(AZ);
  return iter;
}

/* Full-sky expansion of a compressed map (write_reduced_map, 2364-2388). */
void or_expand_map(long nside, const long *indpix, const double *mapr,
                   double *full) {
  long n = 12 * nside * nside;
  for (long i = 0; i < n; i++) {
    full[i] = 0.0;
    if (indpix[i] != -1) {
      if (indpix[i] < 0)
        full[i] = mapr[-indpix[i] - 2];
      else
        full[i] = mapr[indpix[i]];
    }
  }
}

int or_version(void) { return 1; }
