/*
 * sincdvr_oracle.c -- CPU restatement of the LBNL-AMO-MCTDHF sinc-DVR
 * orbital-operator hot path, in plain C (FP64, complex interleaved re/im).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * product library.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it; the product path never does.
 *
 * Every routine cites the reference routine it restates (paths relative to
 * the reference repository root, S/ = MCTDH.SRC, SD/ = MCTDH.SRC/SINCDVR).
 *
 * Parity pins (what checks each routine, see tests/test_oracle_*.py):
 *   orc_sinedvr        closed form + the reference's own sineDVR compiled
 *                      from SD/SineDVR.f90 (oracle/_ref) when present
 *   orc_ecs_axis       ECS-off identity (J=1, diag=0); finite differences
 *                      of F -> J -> J' -> J'' ; closed values at ramp ends
 *   orc_getinverse     residual of the defining linear system T3d Q = TR on
 *                      the small box; far field == Delta^2/r exactly;
 *                      brute-force TR at sample points; cube symmetry
 *   orc_fft1d          numpy.fft (pocketfft) and the reference's DFFTPACK
 *                      zfftf compiled into oracle/_ref when present
 *   orc_op_tinv        brute-force O(n^6) Toeplitz sum on tiny grids,
 *                      independence from the circulant edge M, linearity
 *   orc_twoe_matel     index bookkeeping by brute force on tiny grids;
 *                      hermnorm/cnorm symmetries
 *   orc_mult_ke        dense n^3 x n^3 brute force on tiny grids
 *   orc_mult_reducedpot, orc_reducedpot   numpy einsum / matmul
 */
#include <math.h>
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cplx;

static inline cplx cmul(cplx a, cplx b) { cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }
static inline cplx cadd(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cplx csub(cplx a, cplx b) { cplx r = {a.re - b.re, a.im - b.im}; return r; }
static inline cplx cconj(cplx a) { cplx r = {a.re, -a.im}; return r; }
static inline cplx cscale(cplx a, double s) { cplx r = {a.re * s, a.im * s}; return r; }
static inline cplx cdiv(cplx a, cplx b) {
  double d = b.re * b.re + b.im * b.im;
  cplx r = {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
  return r;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* a1: sineDVR  -- SD/SineDVR.f90:5-35                                       */
/* ke[i+(n-1)], fd[i+(n-1)] for i = 1-n .. n-1 ; points[0..n-1]               */
/* ------------------------------------------------------------------------ */
void orc_sinedvr(int n, double spacing, double *ke, double *fd, double *points) {
  const double mypi = 4.0 * atan(1.0);
  for (int i = 1; i <= n; i++) points[i - 1] = i * spacing;          /* :16-18 */
  const double mid = (points[0] + points[n - 1]) / 2.0;              /* :20 */
  for (int i = 0; i < n; i++) points[i] = points[i] - mid;
  for (int i = 1 - n; i <= n - 1; i++) {                              /* :22-30 */
    double k, f;
    if (i == 0) { k = mypi * mypi / 3.0; f = 0.0; }
    else {
      double sgn = (i % 2 == 0) ? 1.0 : -1.0;
      double ri = (double)i;
      k = 2.0 / (ri * ri) * sgn;
      f = 1.0 / ri * sgn;
    }
    ke[i + n - 1] = k / 2.0 / (spacing * spacing);                   /* :32 */
    fd[i + n - 1] = f / spacing;                                      /* :33 */
  }
}

/* ------------------------------------------------------------------------ */
/* a2: smooth exterior complex scaling  -- SD/init_sinc_new.f90:5-166       */
/* Evaluates F(x), J(x), J'(x), J''(x) for one coordinate.                   */
/* ------------------------------------------------------------------------ */
typedef struct {
  double theta, stretch, dist, smooth;
} orc_ecs;

static double fscaled(double x) {            /* :122-132 */
  const double pi = 4.0 * atan(1.0);
  double s = sin(pi * x);
  return (-1.0) / 12.0 / pi * s * s * s - 1.0 / 2.0 / pi * s + x / 2.0;
}
static double jscaled(double x) {            /* :134-144 */
  const double pi = 4.0 * atan(1.0);
  double c = cos(pi * x);
  return 1.0 / 4.0 * c * c * c - 3.0 / 4.0 * c + 0.5;
}
static double djscaled(double x) {           /* :146-156 */
  const double pi = 4.0 * atan(1.0);
  double s = sin(pi * x);
  return 0.75 * pi * s * s * s;
}
static double ddjscaled(double x) {          /* :158-166 */
  const double pi = 4.0 * atan(1.0);
  double s = sin(pi * x), c = cos(pi * x);
  return 9.0 / 4.0 * pi * pi * s * s * c;
}

void orc_ecs_point(double x, double theta, double stretch, double R0, double S,
                   double *F, double *J, double *dJ, double *ddJ) {
  cplx fac = {cos(theta) * stretch, sin(theta) * stretch};           /* :14 */
  cplx fm1 = {fac.re - 1.0, fac.im};
  cplx f, j, dj, ddj;
  if (x <= -(R0 + S)) {                                               /* :16-17,41-42 */
    f = cadd((cplx){x, 0}, cmul(fm1, (cplx){x + R0 + S / 2.0, 0}));
    j = fac; dj = (cplx){0, 0}; ddj = (cplx){0, 0};
  } else if (x >= R0 + S) {                                           /* :18-19 */
    f = cadd((cplx){x, 0}, cmul(fm1, (cplx){x - R0 - S / 2.0, 0}));
    j = fac; dj = (cplx){0, 0}; ddj = (cplx){0, 0};
  } else if (fabs(x) <= R0) {                                         /* :20-21 */
    f = (cplx){x, 0}; j = (cplx){1, 0}; dj = (cplx){0, 0}; ddj = (cplx){0, 0};
  } else if (x >= R0) {                                               /* :22-24,47-49,71-73,96-98 */
    double xx = (x - R0) / S;
    f = cadd((cplx){x, 0}, cscale(fm1, S * fscaled(xx)));
    j = cadd((cplx){1, 0}, cscale(fm1, jscaled(xx)));
    dj = cscale(fm1, djscaled(xx) / S);
    ddj = cscale(fm1, ddjscaled(xx) / (S * S));
  } else {                                                            /* :25-27,50-52,74-76,99-101 */
    double xx = (-x - R0) / S;
    f = csub((cplx){x, 0}, cscale(fm1, S * fscaled(xx)));
    j = cadd((cplx){1, 0}, cscale(fm1, jscaled(xx)));
    dj = cscale((cplx){-fm1.re, -fm1.im}, djscaled(xx) / S);
    ddj = cscale(fm1, ddjscaled(xx) / (S * S));
  }
  F[0] = f.re; F[1] = f.im; J[0] = j.re; J[1] = j.im;
  dJ[0] = dj.re; dJ[1] = dj.im; ddJ[0] = ddj.re; ddJ[1] = ddj.im;
}

/* Full-grid ECS arrays as the reference stores them:
 *   invjacobian(n^3,3) = 1/J_d(x_d), scalediag(n^3)
 * SD/init_sinc_new.f90:493-521 (dipoles(i,d) = grid coordinate along d,
 * SD/init_sinc_new.f90:619-645).  Arrays are complex interleaved.          */
void orc_ecs_arrays(int n, double spacing, double theta, double stretch, double R0, double S,
                    double *invjac /* 2*n^3*3 */, double *scalediag /* 2*n^3 */) {
  double *ke = malloc(sizeof(double) * (2 * n - 1)), *fd = malloc(sizeof(double) * (2 * n - 1));
  double *pts = malloc(sizeof(double) * n);
  orc_sinedvr(n, spacing, ke, fd, pts);
  cplx *ij = malloc(sizeof(cplx) * n), *dj = malloc(sizeof(cplx) * n), *ddj = malloc(sizeof(cplx) * n);
  for (int i = 0; i < n; i++) {
    double F[2], J[2], D[2], DD[2];
    orc_ecs_point(pts[i], theta, stretch, R0, S, F, J, D, DD);
    ij[i] = cdiv((cplx){1, 0}, (cplx){J[0], J[1]});                   /* :507 */
    dj[i] = (cplx){D[0], D[1]};
    ddj[i] = (cplx){DD[0], DD[1]};
  }
  const size_t tot = (size_t)n * n * n;
  cplx *IJ = (cplx *)invjac, *SD = (cplx *)scalediag;
  for (size_t p = 0; p < tot; p++) {
    int ix[3] = {(int)(p % n), (int)((p / n) % n), (int)(p / ((size_t)n * n))};
    cplx acc = {0, 0};
    for (int d = 0; d < 3; d++) {                                     /* :517-521 */
      cplx a = ij[ix[d]];
      IJ[p + tot * d] = a;
      cplx a2 = cmul(a, a), a3 = cmul(a2, a), a4 = cmul(a2, a2);
      cplx t1 = cscale(cmul(a4, cmul(dj[ix[d]], dj[ix[d]])), 3.0 / 8.0);
      cplx t2 = cscale(cmul(a3, ddj[ix[d]]), 1.0 / 4.0);
      acc = cadd(acc, csub(t1, t2));
    }
    SD[p] = acc;
  }
  free(ke); free(fd); free(pts); free(ij); free(dj); free(ddj);
}

/* ------------------------------------------------------------------------ */
/* Symmetric eigen-decomposition (cyclic Jacobi), used for the 1-D sinc     */
/* Laplacian on the small box.  Stands in for LAPACK DSYEV at              */
/* SD/tooth.f90:64 (the solve it feeds is unique; see orc_getinverse).      */
/* A is ns x ns column-major, overwritten; W gets eigenvectors (columns).   */
/* ------------------------------------------------------------------------ */
static void jacobi_eig(int ns, double *A, double *W, double *E) {
  for (int i = 0; i < ns * ns; i++) W[i] = 0;
  for (int i = 0; i < ns; i++) W[i + ns * i] = 1;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0, tot = 0;
    for (int j = 0; j < ns; j++)
      for (int i = 0; i < ns; i++) {
        double v = A[i + ns * j] * A[i + ns * j];
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-60 * tot || off == 0.0) break;
    for (int p = 0; p < ns - 1; p++)
      for (int q = p + 1; q < ns; q++) {
        double apq = A[p + ns * q];
        if (fabs(apq) < 1e-300) continue;
        double app = A[p + ns * p], aqq = A[q + ns * q];
        double tau = (aqq - app) / (2 * apq);
        double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1 + tau * tau));
        double c = 1 / sqrt(1 + t * t), s = t * c;
        for (int k = 0; k < ns; k++) {     /* columns p,q of A */
          double akp = A[k + ns * p], akq = A[k + ns * q];
          A[k + ns * p] = c * akp - s * akq;
          A[k + ns * q] = s * akp + c * akq;
        }
        for (int k = 0; k < ns; k++) {     /* rows p,q of A */
          double apk = A[p + ns * k], aqk = A[q + ns * k];
          A[p + ns * k] = c * apk - s * aqk;
          A[q + ns * k] = s * apk + c * aqk;
        }
        for (int k = 0; k < ns; k++) {
          double wkp = W[k + ns * p], wkq = W[k + ns * q];
          W[k + ns * p] = c * wkp - s * wkq;
          W[k + ns * q] = s * wkp + c * wkq;
        }
      }
  }
  for (int i = 0; i < ns; i++) E[i] = A[i + ns * i];
}

/* Y = (W (x) W (x) W) X on an nn^3 cube (column-major), SD/tooth.f90:191-231 */
static void tootheigmult(const double *X, double *Y, const double *W, int nn) {
  size_t n3 = (size_t)nn * nn * nn;
  double *X1 = calloc(n3, sizeof(double)), *X2 = calloc(n3, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int j = 0; j < nn; j++)                      /* X1(i,j,:) = W X(i,j,:)  :207-211 */
    for (int i = 0; i < nn; i++)
      for (int k = 0; k < nn; k++) {
        double s = 0;
        for (int l = 0; l < nn; l++) s += W[k + nn * l] * X[i + nn * (j + (size_t)nn * l)];
        X1[i + nn * (j + (size_t)nn * k)] = s;
      }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nn; k++)                      /* X2(i,:,k) = W X1(i,:,k) :213-217 */
    for (int i = 0; i < nn; i++)
      for (int j = 0; j < nn; j++) {
        double s = 0;
        for (int l = 0; l < nn; l++) s += W[j + nn * l] * X1[i + nn * (l + (size_t)nn * k)];
        X2[i + nn * (j + (size_t)nn * k)] = s;
      }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nn; k++)                      /* Y(:,j,k) = W X2(:,j,k)  :219-225 */
    for (int j = 0; j < nn; j++)
      for (int i = 0; i < nn; i++) {
        double s = 0;
        for (int l = 0; l < nn; l++) s += W[i + nn * l] * X2[l + nn * (j + (size_t)nn * k)];
        Y[i + nn * (j + (size_t)nn * k)] = s;
      }
  free(X1); free(X2);
}

/* ------------------------------------------------------------------------ */
/* a3: GETINVERSE / RVECTOR  -- SD/tooth.f90:12-187                          */
/* tinv[(i+N) + (2N+1)*((j+N) + (2N+1)*(k+N))], i,j,k in [-N,N]              */
/* Also returns the small-box right-hand side TR and correction Q when the   */
/* pointers are non-NULL (for the residual test of the solve).              */
/* The intended slice of SD/tooth.f90:165 (ALLTT(-NBIG-J:NBIG-J)) is used.   */
/* ------------------------------------------------------------------------ */
static inline double rvec(int i, int j, int k, double delta) {
  /* SD/tooth.f90:135-146 : R = DELTA^3/SQRT(PPOS(I)^2+PPOS(J)^2+PPOS(K)^2) */
  if (i == 0 && j == 0 && k == 0) return 0.0;
  double pi_ = i * delta, pj = j * delta, pk = k * delta;
  return (delta * delta * delta) / sqrt(pi_ * pi_ + pj * pj + pk * pk);
}

static inline double alltt(int m, double delta) {
  /* SD/tooth.f90:152-156 */
  const double pi = 4.0 * atan(1.0);
  if (m == 0) return pi * pi / (6.0 * delta * delta);
  double sgn = (m % 2 == 0) ? 1.0 : -1.0;
  return sgn * (1.0 / (delta * delta)) / ((double)m * (double)m);
}

int orc_getinverse(int N, int nbig, int nsmall, double delta, double *tinv,
                   double *tr_out, double *q_out) {
  if (nbig <= nsmall) return -1;                    /* :114-117 */
  if (N > nbig) return -2;
  const double pi = 4.0 * atan(1.0);
  const int NS = 2 * nsmall + 1;
  const size_t ns3 = (size_t)NS * NS * NS;
  double *TR = calloc(ns3, sizeof(double)), *Q = calloc(ns3, sizeof(double));
  double *QQ = calloc(ns3, sizeof(double)), *T = calloc((size_t)NS * NS, sizeof(double));
  double *W = calloc((size_t)NS * NS, sizeof(double)), *WT = calloc((size_t)NS * NS, sizeof(double));
  double *E = calloc(NS, sizeof(double));
  /* TR = -T_big R over all three axes, :158-181 */
#pragma omp parallel for schedule(static)
  for (int k = -nsmall; k <= nsmall; k++)
    for (int j = -nsmall; j <= nsmall; j++)
      for (int i = -nsmall; i <= nsmall; i++) {
        double s = 0;
        for (int p = -nbig; p <= nbig; p++)
          s += alltt(p - i, delta) * rvec(p, j, k, delta) + alltt(p - j, delta) * rvec(i, p, k, delta) +
               alltt(p - k, delta) * rvec(i, j, p, delta);
        TR[(i + nsmall) + NS * ((j + nsmall) + (size_t)NS * (k + nsmall))] = -s;
      }
  TR[nsmall + NS * (nsmall + (size_t)NS * nsmall)] += 2 * pi;          /* :44 */
  const double mycon = (pi / delta) * (pi / delta) / 6.0;             /* :38 */
  for (int j = -nsmall; j <= nsmall; j++) {                           /* :48-53 */
    for (int i = -nsmall; i <= nsmall; i++)
      if (i != j) {
        double sgn = ((i - j) % 2 == 0) ? 1.0 : -1.0;
        T[(i + nsmall) + NS * (j + nsmall)] = sgn / (delta * delta * (double)(i - j) * (double)(i - j));
      }
    T[(j + nsmall) + NS * (j + nsmall)] = mycon;
  }
  double *A = malloc(sizeof(double) * NS * NS);
  memcpy(A, T, sizeof(double) * NS * NS);
  jacobi_eig(NS, A, W, E);                                            /* :63-67 */
  free(A);
  for (int j = 0; j < NS; j++)
    for (int i = 0; i < NS; i++) WT[i + NS * j] = W[j + NS * i];      /* :76 */
  tootheigmult(TR, QQ, WT, NS);                                       /* :77 */
  for (int k = 0; k < NS; k++)                                        /* :69-73,78 */
    for (int j = 0; j < NS; j++)
      for (int i = 0; i < NS; i++) QQ[i + NS * (j + (size_t)NS * k)] /= (E[i] + E[j] + E[k]);
  tootheigmult(QQ, Q, W, NS);                                         /* :79 */
  /* TINV = R + Q on the small box, truncated to (-N:N)^3 :83-89 */
  const int L = 2 * N + 1;
#pragma omp parallel for schedule(static)
  for (int k = -N; k <= N; k++)
    for (int j = -N; j <= N; j++)
      for (int i = -N; i <= N; i++) {
        double v = rvec(i, j, k, delta);
        if (abs(i) <= nsmall && abs(j) <= nsmall && abs(k) <= nsmall)
          v += Q[(i + nsmall) + NS * ((j + nsmall) + (size_t)NS * (k + nsmall))];
        tinv[(i + N) + L * ((j + N) + (size_t)L * (k + N))] = v;
      }
  if (tr_out) memcpy(tr_out, TR, sizeof(double) * ns3);
  if (q_out) memcpy(q_out, Q, sizeof(double) * ns3);
  free(TR); free(Q); free(QQ); free(T); free(W); free(WT); free(E);
  return 0;
}

/* a4: threed_two(1-n:n-1)^3 = TINV/Delta^3 -- SD/sincproject.f90:155-156,182-187.
 * out has (2n-1)^3 entries, offsets -(n-1)..(n-1), first index fastest.      */
int orc_coulomb_kernel(int n, double spacing, int nbig, int nsmall, double *out) {
  int rc = orc_getinverse(n - 1, nbig, nsmall, spacing, out, NULL, NULL);
  if (rc) return rc;
  const size_t L = 2 * (size_t)n - 1;
  const double d3 = spacing * spacing * spacing;
  for (size_t p = 0; p < L * L * L; p++) out[p] = out[p] / d3;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* a8/a9: 1-D complex FFT, forward (e^{-2 pi i jk/M}), unnormalised, any M.  */
/* Same transform as DFFTPACK zfftf (S/DFFTPACK/zfftf.f) and FFTW FORWARD    */
/* (CP/ftcore.f90:84-92).  Recursive mixed radix, generic butterflies.       */
/* ------------------------------------------------------------------------ */
typedef struct {
  int M, nf, f[32];
  cplx *tw; /* tw[k] = exp(-2 pi i k / M) */
} fftplan;

static void plan_init(fftplan *pl, int M) {
  pl->M = M; pl->nf = 0;
  int r = M;
  int order[] = {4, 2, 3, 5, 7};
  for (int oi = 0; oi < 5; oi++)
    while (r % order[oi] == 0) { pl->f[pl->nf++] = order[oi]; r /= order[oi]; }
  for (int p = 11; r > 1; p += 2)
    while (r % p == 0) { pl->f[pl->nf++] = p; r /= p; }
  if (pl->nf == 0) pl->f[pl->nf++] = 1;
  pl->tw = malloc(sizeof(cplx) * M);
  const double twopi = 8.0 * atan(1.0);
  for (int k = 0; k < M; k++) {
    double a = twopi * (double)k / (double)M;
    pl->tw[k] = (cplx){cos(a), -sin(a)};
  }
}
static void plan_free(fftplan *pl) { free(pl->tw); }

static void fft_rec(const cplx *in, cplx *out, int N, size_t istride, const int *fac,
                    const fftplan *pl, int twstride, cplx *scratch) {
  const int p = fac[0];
  const int m = N / p;
  if (m == 1) {
    for (int j = 0; j < p; j++) out[j] = in[j * istride];
  } else {
    for (int j = 0; j < p; j++) fft_rec(in + j * istride, out + j * m, m, istride * p, fac + 1, pl, twstride * p, scratch);
  }
  /* butterflies: out[s*m+k] = sum_q out[q*m+k] w_N^{qk} w_p^{qs} */
  const int M = pl->M;
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) {
      size_t ti = ((size_t)q * k * twstride) % M;
      scratch[q] = cmul(out[q * m + k], pl->tw[ti]);
    }
    for (int s = 0; s < p; s++) {
      cplx acc = scratch[0];
      for (int q = 1; q < p; q++) {
        size_t ti = ((size_t)q * s * m * twstride) % M;
        acc = cadd(acc, cmul(scratch[q], pl->tw[ti]));
      }
      out[s * m + k] = acc;
    }
  }
}

/* Forward FFT of `howmany` contiguous lines of length M (in-place allowed). */
void orc_fft1d(int M, int howmany, const double *in, double *out) {
  fftplan pl;
  plan_init(&pl, M);
#pragma omp parallel
  {
    cplx *buf = malloc(sizeof(cplx) * M), *res = malloc(sizeof(cplx) * M), *scr = malloc(sizeof(cplx) * 64 * 1024);
#pragma omp for schedule(static)
    for (int h = 0; h < howmany; h++) {
      memcpy(buf, (const cplx *)in + (size_t)h * M, sizeof(cplx) * M);
      fft_rec(buf, res, M, 1, pl.f, &pl, 1, scr);
      memcpy((cplx *)out + (size_t)h * M, res, sizeof(cplx) * M);
    }
    free(buf); free(res); free(scr);
  }
  plan_free(&pl);
}

/* a7: myzfft3d = 3 x fftblock_withtranspose  -- SD/toeplitz_subs.f90:264-294.
 * In-place on a (d,d,d,howmany) cube: FFT along the first axis, then the
 * (123)->(231) transpose, three times.                                      */
static void zfft3d_inplace(cplx *a, int d, int howmany, const fftplan *pl, cplx *work) {
  const size_t d3 = (size_t)d * d * d;
  for (int pass = 0; pass < 3; pass++) {
    /* work(i,:,:,h) = FFT along i of a(:,j,k,h) ; stored as work(i,j,k) */
#pragma omp parallel
    {
      cplx *buf = malloc(sizeof(cplx) * d), *res = malloc(sizeof(cplx) * d), *scr = malloc(sizeof(cplx) * 4096);
#pragma omp for schedule(static)
      for (size_t line = 0; line < (size_t)d * d * howmany; line++) {
        memcpy(buf, a + line * d, sizeof(cplx) * d);
        fft_rec(buf, res, d, 1, pl->f, pl, 1, scr);
        memcpy(work + line * d, res, sizeof(cplx) * d);
      }
      free(buf); free(res); free(scr);
    }
    /* inout(:,:,i,:) = work1(i,:,:,:)   :278-280 */
#pragma omp parallel for schedule(static)
    for (int h = 0; h < howmany; h++)
      for (int k = 0; k < d; k++)
        for (int j = 0; j < d; j++)
          for (int i = 0; i < d; i++)
            a[h * d3 + j + (size_t)d * (k + (size_t)d * i)] = work[h * d3 + i + (size_t)d * (j + (size_t)d * k)];
  }
}

/* ------------------------------------------------------------------------ */
/* a5/a6: op_tinv_notscaled + circ3d_sub                                     */
/*   SD/coreproject.f90:366-537, SD/toeplitz_subs.f90:561-589                */
/* kernel: (2n-1)^3 Toeplitz kernel (offsets -(n-1)..(n-1)).                 */
/* den, red: (n^3, allsize) complex.  M = circulant edge (reference: 2n;     */
/* any M >= 2n-1 gives the same linear convolution).  The reference places   */
/* the kernel at offset +n (array threed_two(-n:n-1)) and reads the second   */
/* octant; for general M the kernel offset o sits at circulant index o mod M */
/* and the first octant is read, which is the same sum.                      */
/* ------------------------------------------------------------------------ */
int orc_op_tinv(int n, int M, const double *kernel, const double *den_, double *red_, int allsize,
                int circsize) {
  if (circsize <= 0 || allsize % circsize != 0) return -1;   /* :387-391 */
  if (M < 2 * n - 1) return -2;
  const cplx *den = (const cplx *)den_;
  cplx *red = (cplx *)red_;
  const size_t M3 = (size_t)M * M * M, n3 = (size_t)n * n * n;
  const int L = 2 * n - 1;
  const int ref_layout = (M == 2 * n);
  fftplan pl;
  plan_init(&pl, M);
  cplx *kmat = calloc(M3, sizeof(cplx));
  cplx *work = malloc(sizeof(cplx) * M3 * (size_t)circsize);
  cplx *vec = malloc(sizeof(cplx) * M3 * (size_t)circsize);
  for (int c0 = 0; c0 < allsize; c0 += circsize) {
    /* bigcirc: threed_two placed into the circulant (recomputed FFT per call, :574) */
    memset(kmat, 0, sizeof(cplx) * M3);
    for (int k = -(n - 1); k <= n - 1; k++)
      for (int j = -(n - 1); j <= n - 1; j++)
        for (int i = -(n - 1); i <= n - 1; i++) {
          double v = kernel[(i + n - 1) + L * ((j + n - 1) + (size_t)L * (k + n - 1))];
          int ci, cj, ck;
          if (ref_layout) { ci = i + n; cj = j + n; ck = k + n; }       /* threed_two(-n:n-1) */
          else { ci = (i + M) % M; cj = (j + M) % M; ck = (k + M) % M; }
          kmat[ci + (size_t)M * (cj + (size_t)M * ck)] = (cplx){v, 0};
        }
    zfft3d_inplace(kmat, M, 1, &pl, work);
    /* twoeden03huge(:,1,:,1,:,1,:) = twoeden03  (:449-450) */
    memset(vec, 0, sizeof(cplx) * M3 * circsize);
    for (int h = 0; h < circsize; h++)
      for (int z = 0; z < n; z++)
        for (int y = 0; y < n; y++)
          for (int x = 0; x < n; x++)
            vec[h * M3 + x + (size_t)M * (y + (size_t)M * z)] = den[(size_t)(c0 + h) * n3 + x + (size_t)n * (y + (size_t)n * z)];
    zfft3d_inplace(vec, M, circsize, &pl, work);                     /* :573 */
    const double inv = 1.0 / ((double)M * M * M);
#pragma omp parallel for schedule(static)
    for (int h = 0; h < circsize; h++)                                /* :576-584 */
      for (size_t p = 0; p < M3; p++) vec[h * M3 + p] = cconj(cscale(cmul(vec[h * M3 + p], kmat[p]), inv));
    zfft3d_inplace(vec, M, circsize, &pl, work);                     /* :585 */
    const int off = ref_layout ? n : 0;                               /* reducedhuge(:,2,:,2,:,2,:) :521 */
    for (int h = 0; h < circsize; h++)
      for (int z = 0; z < n; z++)
        for (int y = 0; y < n; y++)
          for (int x = 0; x < n; x++)
            red[(size_t)(c0 + h) * n3 + x + (size_t)n * (y + (size_t)n * z)] =
                cconj(vec[h * M3 + (x + off) + (size_t)M * ((y + off) + (size_t)M * (z + off))]);   /* :586 */
  }
  free(kmat); free(work); free(vec);
  plan_free(&pl);
  return 0;
}

/* Brute-force Toeplitz sum V(x) = sum_y K(x-y) rho(y), O(n^6): the closed  */
/* definition the circulant embedding must reproduce (tiny grids only).     */
void orc_toeplitz_direct(int n, const double *kernel, const double *den_, double *red_, int howmany) {
  const cplx *den = (const cplx *)den_;
  cplx *red = (cplx *)red_;
  const size_t n3 = (size_t)n * n * n;
  const int L = 2 * n - 1;
  for (int h = 0; h < howmany; h++)
#pragma omp parallel for schedule(static)
    for (size_t px = 0; px < n3; px++) {
      int x = px % n, y = (px / n) % n, z = px / ((size_t)n * n);
      cplx acc = {0, 0};
      for (int c = 0; c < n; c++)
        for (int b = 0; b < n; b++)
          for (int a = 0; a < n; a++) {
            double kv = kernel[(x - a + n - 1) + L * ((y - b + n - 1) + (size_t)L * (z - c + n - 1))];
            acc = cadd(acc, cscale(den[h * n3 + a + (size_t)n * (b + (size_t)n * c)], kv));
          }
      red[h * n3 + px] = acc;
    }
}

/* ------------------------------------------------------------------------ */
/* a10: call_twoe_matel00 / call_twoe_matelxxx00 (fft_batchdim=1,            */
/* fft_circbatchdim=1)  -- SD/coreproject.f90:945-1194                        */
/* phi1, phi2: (n^3, N).  twoematel: (N,N,N,low:high).                        */
/* twoereduced: (n^3, N, low:high).  cnorm != 0 drops the CONJUGATE           */
/* (S/Definitions.ALL:114 vs :122).                                           */
/* ------------------------------------------------------------------------ */
int orc_twoe_matel(int n, int M, const double *kernel, int N, int low, int high,
                   const double *phi1_, const double *phi2_, int cnorm, int notwo,
                   double *twoematel_, double *twoereduced_) {
  if (high > N || low < 1) return -1;                                 /* :994-996 */
  const cplx *phi1 = (const cplx *)phi1_, *phi2 = (const cplx *)phi2_;
  cplx *tm = (cplx *)twoematel_, *tr = (cplx *)twoereduced_;
  const size_t n3 = (size_t)n * n * n;
  const int nnn = high - low + 1;
  memset(tr, 0, sizeof(cplx) * n3 * N * nnn);                         /* :1019 */
  memset(tm, 0, sizeof(cplx) * (size_t)N * N * N * nnn);
  if (notwo) return 0;                                                /* :1021-1023 */
  cplx *den = malloc(sizeof(cplx) * n3 * N);
  for (int b2 = low; b2 <= high; b2++) {                              /* :1039 */
    for (int a2 = 1; a2 <= N; a2++)                                   /* :1057-1065 */
      for (size_t x = 0; x < n3; x++) {
        cplx v2 = phi2[(size_t)(b2 - 1) * n3 + x];
        if (!cnorm) v2 = cconj(v2);
        den[(size_t)(a2 - 1) * n3 + x] = cmul(phi1[(size_t)(a2 - 1) * n3 + x], v2);
      }
    int rc = orc_op_tinv(n, M, kernel, (double *)den, (double *)(tr + (size_t)(b2 - low) * N * n3), N, N);   /* :1078 */
    if (rc) { free(den); return rc; }
  }
  /* twoematel(1a,1b,:,low:high) = twoereduced^T . myden  (ZGEMV 'T', :1087-1100) */
#pragma omp parallel for schedule(dynamic) collapse(2)
  for (int b1 = 1; b1 <= N; b1++)
    for (int a1 = 1; a1 <= N; a1++) {
      for (int col = 0; col < N * nnn; col++) {
        cplx acc = {0, 0};
        const cplx *v = tr + (size_t)col * n3;
        for (size_t x = 0; x < n3; x++) {
          cplx p2 = phi2[(size_t)(b1 - 1) * n3 + x];
          if (!cnorm) p2 = cconj(p2);
          acc = cadd(acc, cmul(v[x], cmul(phi1[(size_t)(a1 - 1) * n3 + x], p2)));
        }
        /* twoemattemp(:,low:high) column col -> (2a, 2b) = (col%N+1, col/N+low) */
        int a2 = col % N, b2 = col / N;
        tm[(a1 - 1) + (size_t)N * ((b1 - 1) + (size_t)N * (a2 + (size_t)N * b2))] = acc;
      }
    }
  free(den);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* a11: mult_ke -> mult_general(option=1) -> mult_general_withbcast ->       */
/* mult_allpar -> mult_allone -> mult_all0                                    */
/*   SD/coreproject.f90:2104-2116, 1243-1390, 1634-1666, 2050-2097           */
/* ke: 1-D stencil ketot(1-n:n-1) (index i+n-1).  invjac/scalediag: NULL     */
/* when scalingflag == 0, else full-grid arrays (orc_ecs_arrays).            */
/* ------------------------------------------------------------------------ */
void orc_mult_ke(int n, const double *ke, const double *invjac_, const double *scalediag_,
                 const double *in_, double *out_, int howmany) {
  const cplx *in = (const cplx *)in_;
  cplx *out = (cplx *)out_;
  const cplx *ij = (const cplx *)invjac_, *sd = (const cplx *)scalediag_;
  const size_t n3 = (size_t)n * n * n;
  cplx *temp = malloc(sizeof(cplx) * n3 * howmany), *temp2 = malloc(sizeof(cplx) * n3 * howmany);
  memset(out, 0, sizeof(cplx) * n3 * howmany);
  for (int d = 0; d < 3; d++) {                                       /* :1340 */
    for (int h = 0; h < howmany; h++)
      for (size_t p = 0; p < n3; p++)
        temp[h * n3 + p] = ij ? cmul(in[h * n3 + p], ij[d * n3 + p]) : in[h * n3 + p];   /* :1344-1358 */
    /* mult_all0: out(nnn,q,mmm) = sum_p in(nnn,p,mmm) * ketot(p,q), ketot(p,q)=kevect(q-p) */
    const size_t nnn = d == 0 ? 1 : (d == 1 ? (size_t)n : (size_t)n * n);
    const size_t mmm = (d == 0 ? (size_t)n * n : (d == 1 ? (size_t)n : 1)) * howmany;
#pragma omp parallel for schedule(static)
    for (size_t mm = 0; mm < mmm; mm++)
      for (int q = 0; q < n; q++)
        for (size_t a = 0; a < nnn; a++) {
          cplx acc = {0, 0};
          for (int p = 0; p < n; p++) {
            double kv = ke[(q - p) + n - 1];
            cplx v = temp[a + nnn * (p + (size_t)n * mm)];
            acc.re += v.re * kv; acc.im += v.im * kv;
          }
          temp2[a + nnn * (q + (size_t)n * mm)] = acc;
        }
    for (int h = 0; h < howmany; h++)
      for (size_t p = 0; p < n3; p++) {
        cplx v = ij ? cmul(temp2[h * n3 + p], ij[d * n3 + p]) : temp2[h * n3 + p];          /* :1364-1378 */
        out[h * n3 + p] = cadd(out[h * n3 + p], v);
      }
  }
  if (sd)
    for (int h = 0; h < howmany; h++)
      for (size_t p = 0; p < n3; p++) out[h * n3 + p] = cadd(out[h * n3 + p], cmul(in[h * n3 + p], sd[p]));   /* :1382-1388 */
  free(temp); free(temp2);
}

/* maskfunction of init_project (SD/init_sinc_new.f90:580-599): 1, times
 * (1 + cos(pi ii/(mnp+1)))/2 at the ii-th point in from each end
 * (ii = 1..mnp; overlapping ends multiply twice). */
void orc_mask_function(int n, int mnp, double *m) {
  for (int i = 0; i < n; i++) m[i] = 1.0;
  for (int ii = 1; ii <= mnp; ii++) {
    const double rsum = (1.0 + cos(M_PI * ii / (mnp + 1))) / 2.0;
    m[(mnp + 1 - ii) - 1] *= rsum;
    m[(n - mnp + ii) - 1] *= rsum;
  }
}

/* mult_ke with maskflag (scalingflag = 0), SD/coreproject.f90:1355-1376 and
 * divide_mask / mult_mask (:1273-1319): for each axis d
 *   temp = ((in / m(x)) / m(y)) / m(z);  out += ((K_d temp * m(x)) * m(y)) * m(z)
 * with the same mask on the three axes (cubic grid). */
void orc_mult_ke_mask(int n, const double *ke, const double *mask, const double *in_, double *out_, int howmany) {
  const cplx *in = (const cplx *)in_;
  cplx *out = (cplx *)out_;
  const size_t n3 = (size_t)n * n * n;
  cplx *temp = malloc(sizeof(cplx) * n3 * howmany), *temp2 = malloc(sizeof(cplx) * n3 * howmany);
  memset(out, 0, sizeof(cplx) * n3 * howmany);
  for (int d = 0; d < 3; d++) {
    for (int h = 0; h < howmany; h++)
      for (size_t p = 0; p < n3; p++) {
        const int x = (int)(p % n), y = (int)((p / n) % n), z = (int)(p / ((size_t)n * n));
        cplx v = in[h * n3 + p];
        v.re = v.re / mask[x]; v.im = v.im / mask[x];
        v.re = v.re / mask[y]; v.im = v.im / mask[y];
        v.re = v.re / mask[z]; v.im = v.im / mask[z];
        temp[h * n3 + p] = v;
      }
    const size_t nnn = d == 0 ? 1 : (d == 1 ? (size_t)n : (size_t)n * n);
    const size_t mmm = (d == 0 ? (size_t)n * n : (d == 1 ? (size_t)n : 1)) * howmany;
#pragma omp parallel for schedule(static)
    for (size_t mm = 0; mm < mmm; mm++)
      for (int q = 0; q < n; q++)
        for (size_t a = 0; a < nnn; a++) {
          cplx acc = {0, 0};
          for (int p = 0; p < n; p++) {
            double kv = ke[(q - p) + n - 1];
            cplx v = temp[a + nnn * (p + (size_t)n * mm)];
            acc.re += v.re * kv; acc.im += v.im * kv;
          }
          temp2[a + nnn * (q + (size_t)n * mm)] = acc;
        }
    for (int h = 0; h < howmany; h++)
      for (size_t p = 0; p < n3; p++) {
        const int x = (int)(p % n), y = (int)((p / n) % n), z = (int)(p / ((size_t)n * n));
        cplx v = temp2[h * n3 + p];
        v = cscale(cscale(cscale(v, mask[x]), mask[y]), mask[z]);
        out[h * n3 + p] = cadd(out[h * n3 + p], v);
      }
  }
  free(temp); free(temp2);
}

/* Dense brute force of the same operator on tiny grids: builds the full
 * n^3 x n^3 matrix sum_d D_d^-1 (1 x .. K_d .. x 1) D_d^-1 + diag.          */
void orc_mult_ke_dense(int n, const double *ke, const double *invjac_, const double *scalediag_,
                       const double *in_, double *out_) {
  const cplx *in = (const cplx *)in_;
  cplx *out = (cplx *)out_;
  const cplx *ij = (const cplx *)invjac_, *sd = (const cplx *)scalediag_;
  const size_t n3 = (size_t)n * n * n;
  for (size_t r = 0; r < n3; r++) {
    int rx[3] = {(int)(r % n), (int)((r / n) % n), (int)(r / ((size_t)n * n))};
    cplx acc = {0, 0};
    for (size_t c = 0; c < n3; c++) {
      int cx[3] = {(int)(c % n), (int)((c / n) % n), (int)(c / ((size_t)n * n))};
      for (int d = 0; d < 3; d++) {
        int ok = 1;
        for (int e = 0; e < 3; e++) if (e != d && rx[e] != cx[e]) ok = 0;
        if (!ok) continue;
        cplx m = {ke[(rx[d] - cx[d]) + n - 1], 0};
        if (ij) m = cmul(m, cmul(ij[d * n3 + r], ij[d * n3 + c]));
        acc = cadd(acc, cmul(m, in[c]));
      }
    }
    if (sd) acc = cadd(acc, cmul(sd[r], in[r]));
    out[r] = acc;
  }
}

/* a12: mult_reducedpot -- SD/coreproject.f90:1200-1235
 * out(:,i) = sum_k reducedpot(:,k,i) * in(:,k), i = first..last             */
void orc_mult_reducedpot(int n3_, int N, int first, int last, const double *in_, double *out_,
                         const double *rp_) {
  const size_t n3 = (size_t)n3_;
  const cplx *in = (const cplx *)in_, *rp = (const cplx *)rp_;
  cplx *out = (cplx *)out_;
#pragma omp parallel for schedule(dynamic)
  for (int i = first; i <= last; i++) {
    for (size_t x = 0; x < n3; x++) {
      cplx acc = {0, 0};
      for (int k = 1; k <= N; k++)
        acc = cadd(acc, cmul(rp[x + n3 * ((k - 1) + (size_t)N * (i - first))], in[x + n3 * (k - 1)]));
      out[x + n3 * (i - first)] = acc;
    }
  }
}

/* a13: get_reducedpot0 -- S/mean.f90:503-588 (ZGEMM 'N','N' at :556-558)
 * reducedpot(n3, N^2) = twoereduced(n3, N^2) . twoden(N^2, N^2)              */
void orc_reducedpot(int n3_, int N, const double *tr_, const double *twoden_, double *out_) {
  const size_t n3 = (size_t)n3_;
  const int NN = N * N;
  const cplx *tr = (const cplx *)tr_, *g = (const cplx *)twoden_;
  cplx *out = (cplx *)out_;
#pragma omp parallel for schedule(static)
  for (size_t x = 0; x < n3; x++)
    for (int c = 0; c < NN; c++) {
      cplx acc = {0, 0};
      for (int k = 0; k < NN; k++) acc = cadd(acc, cmul(tr[x + n3 * k], g[k + (size_t)NN * c]));
      out[x + n3 * c] = acc;
    }
}

/* a4 (nuclear part): get_3dpoisson's potential  -- SD/sincproject.f90:192-250.
 * kernel: (2n-1)^3 threed_two values (offsets -(n-1)..(n-1)).  Each centre
 * sits at half-spacing offsets centershift (3 ints); grid point i (0-based)
 * sees offset o = i + (cs - n + 1)/2.  Inside the cube |o| <= n-1 the value
 * is the kernel (so the nucleus sits at -cs*spacing/2); outside it is
 * 1/(spacing*sqrt(o.o)) with the sqrt taken in
 * default REAL (realkind4 = 1: 4-byte, builds without -r8; 0: 8-byte).
 * pot = - sum_c Z_c value_c.  Returns -1 on the parity error of :199-204.   */
int orc_nuclear_pot(int n, double spacing, const double *kernel, int ncent, const int *cs, const double *Z,
                    int realkind4, double *pot) {
  const int L = 2 * n - 1;
  for (int c = 0; c < ncent; c++)
    for (int d = 0; d < 3; d++) {
      int pp = -n + cs[3 * c + d] + 1, qq = n + cs[3 * c + d] - 1;
      if ((pp % 2) != 0 || (qq % 2) != 0) return -1;
    }
  const size_t n3 = (size_t)n * n * n;
  for (size_t p = 0; p < n3; p++) pot[p] = 0.0;
  for (int c = 0; c < ncent; c++) {
    int off[3];
    for (int d = 0; d < 3; d++) off[d] = (-n + cs[3 * c + d] + 1) / 2;
    for (int k = 0; k < n; k++)
      for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
          int ox = i + off[0], oy = j + off[1], oz = k + off[2];
          double v;
          if (abs(ox) <= n - 1 && abs(oy) <= n - 1 && abs(oz) <= n - 1) {
            v = kernel[(ox + n - 1) + (size_t)L * ((oy + n - 1) + (size_t)L * (oz + n - 1))];
          } else {
            int s2 = ox * ox + oy * oy + oz * oz;
            double rval = realkind4 ? spacing * (double)sqrtf((float)s2) : spacing * sqrt((double)s2);
            v = 1.0 / rval;
          }
          size_t p = i + (size_t)n * (j + (size_t)n * k);
          pot[p] = pot[p] - Z[c] * v;                                  /* :245-247 */
        }
  }
  return 0;
}
