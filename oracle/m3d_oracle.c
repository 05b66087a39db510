/*
 * m3d_oracle.c -- FP64 CPU restatement of the MAT3d 3D FFT homogenization loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * product path (mat3d_amd/csrc).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * It restates, in plain C and double precision, the algorithm of
 *   src/3d/FFT_analysis3D.cpp      (class FFT_Analysis3D, behaviours "e","ep","ed")
 *   src/3dTherm/FFT_thermal3D.cpp  (class FFT_analysis3d_thermal)
 * following the reference's dataflow step by step: a stored 21-component
 * Green table (setGF), per-voxel 6x6 products (multipCx / update), two
 * forward and one inverse 3D FFT per accelerated iteration, the same stop
 * rule and the same homogenization assembly.  Every function cites the
 * reference lines it follows.
 *
 * FFT semantics.  FFTW is absent, so this file carries its own mixed-radix
 * Stockham FFT.  The rank-3 r2c is "r2c along z, then complex along y and x";
 * the c2r is "complex inverse along x and y, then c2r along z" where the
 * imaginary parts of the kz=0 and kz=Nz/2 (even Nz) bins are ignored.  That
 * equals numpy.fft.irfftn and the documented algorithm of FFTW's multi-
 * dimensional c2r.  Neither transform is normalised (FFTW convention).
 *
 * Parity pins (see tests/test_oracle.py): numpy.fft.rfftn/irfftn and a dense
 * DFT for the transforms; Green-operator identities (projector, Delta0 =
 * C0 - C0 Gamma0 C0); homogeneous and laminate closed forms; the reference's
 * own Hashin-Shtrikman bounds; the closed-form J2 uniaxial response.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------------ */
/* 1-D mixed-radix Stockham FFT                                              */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n;
    int nf;
    int fac[40];
    cplx *tw; /* tw[m] = exp(-2 pi i m / n), m in [0,n) */
} plan1d;

static void plan1d_init(plan1d *p, int n)
{
    p->n = n;
    p->nf = 0;
    int m = n;
    while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
    for (int f = 3; m > 1; f += 2)
        while (m % f == 0) { p->fac[p->nf++] = f; m /= f; }
    p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
    for (int k = 0; k < n; ++k) {
        /* angle reduced to [0, pi/4]-style symmetric evaluation for accuracy */
        double a = 2.0 * M_PI * (double)k / (double)n;
        p->tw[k] = cos(a) - I * sin(a);
    }
}

static void plan1d_free(plan1d *p) { free(p->tw); p->tw = NULL; }

/* sign = -1 forward, +1 backward.  a: data (length n), b: scratch. Result in a. */
static void fft1d(const plan1d *p, cplx *a, cplx *b, int sign)
{
    const int n = p->n;
    if (n == 1) return;
    cplx *src = a, *dst = b;
    int Ns = 1;
    cplx v[64], o[64];
    for (int f = 0; f < p->nf; ++f) {
        const int R = p->fac[f];
        const int nR = n / R;
        const int tstep = n / (Ns * R);
        for (int j = 0; j < nR; ++j) {
            const int k = j % Ns;
            for (int r = 0; r < R; ++r) {
                cplx x = src[j + r * nR];
                int ti = (k * r * tstep) % n;
                cplx w = p->tw[ti];
                if (sign > 0) w = conj(w);
                v[r] = x * w;
            }
            if (R == 2) {
                o[0] = v[0] + v[1];
                o[1] = v[0] - v[1];
            } else if (R == 4) {
                cplx s02 = v[0] + v[2], d02 = v[0] - v[2];
                cplx s13 = v[1] + v[3], d13 = v[1] - v[3];
                cplx jd13 = (sign < 0) ? (-I * d13) : (I * d13);
                o[0] = s02 + s13;
                o[2] = s02 - s13;
                o[1] = d02 + jd13;
                o[3] = d02 - jd13;
            } else {
                for (int q = 0; q < R; ++q) {
                    cplx acc = 0;
                    for (int r = 0; r < R; ++r) {
                        int ti = ((r * q) % R) * nR;
                        cplx w = p->tw[ti];
                        if (sign > 0) w = conj(w);
                        acc += v[r] * w;
                    }
                    o[q] = acc;
                }
            }
            const int base = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; ++r) dst[base + r * Ns] = o[r];
        }
        cplx *t = src; src = dst; dst = t;
        Ns *= R;
    }
    if (src != a) memcpy(a, src, sizeof(cplx) * (size_t)n);
}

/* ------------------------------------------------------------------------ */
/* Rank-3 r2c / c2r over C interleaved components (FFTW guru64 layout,       */
/* FFT_analysis3D.cpp:110-150): real in[((x*Ny+y)*Nz+z)*C+c],                */
/* complex out[((x*Ny+y)*Nzh+kz)*C+c], Nzh = Nz/2+1.                           */
/* ------------------------------------------------------------------------ */
typedef struct {
    int Nx, Ny, Nz, Nzh, C;
    plan1d px, py, pz;     /* full lengths */
    plan1d pzh;            /* Nz/2 for the even-length real trick */
    int nthreads;
} plan3d;

static void plan3d_init(plan3d *p, int Nx, int Ny, int Nz, int C, int nthreads)
{
    p->Nx = Nx; p->Ny = Ny; p->Nz = Nz; p->Nzh = Nz / 2 + 1; p->C = C;
    plan1d_init(&p->px, Nx);
    plan1d_init(&p->py, Ny);
    plan1d_init(&p->pz, Nz);
    plan1d_init(&p->pzh, (Nz % 2 == 0) ? Nz / 2 : 1);
    p->nthreads = nthreads > 0 ? nthreads : 1;
}

static void plan3d_free(plan3d *p)
{
    plan1d_free(&p->px); plan1d_free(&p->py); plan1d_free(&p->pz); plan1d_free(&p->pzh);
}

#define BLK 16

/* complex FFT along a strided axis: lines indexed by (outer o, inner i),
 * element e of line at base + o*ostride + e*estride + i, i contiguous. */
static void fft_axis(const plan1d *p, cplx *data, long nouter, long ostride, long estride,
                     long ninner, int sign, int nthreads)
{
    const int n = p->n;
    const long nblk = (ninner + BLK - 1) / BLK;
    const long total = nouter * nblk;
#pragma omp parallel num_threads(nthreads)
    {
        cplx *g = (cplx *)malloc(sizeof(cplx) * (size_t)n * BLK);
        cplx *a = (cplx *)malloc(sizeof(cplx) * (size_t)n);
        cplx *b = (cplx *)malloc(sizeof(cplx) * (size_t)n);
#pragma omp for schedule(static)
        for (long t = 0; t < total; ++t) {
            const long o = t / nblk, ib = (t % nblk) * BLK;
            const int nb = (int)((ninner - ib) < BLK ? (ninner - ib) : BLK);
            cplx *base = data + o * ostride + ib;
            for (int e = 0; e < n; ++e)
                for (int q = 0; q < nb; ++q) g[(long)q * n + e] = base[(long)e * estride + q];
            for (int q = 0; q < nb; ++q) {
                memcpy(a, g + (long)q * n, sizeof(cplx) * (size_t)n);
                fft1d(p, a, b, sign);
                memcpy(g + (long)q * n, a, sizeof(cplx) * (size_t)n);
            }
            for (int e = 0; e < n; ++e)
                for (int q = 0; q < nb; ++q) base[(long)e * estride + q] = g[(long)q * n + e];
        }
        free(g); free(a); free(b);
    }
}

/* forward r2c: unnormalised (FFTW_FORWARD convention) */
static void rfft3(const plan3d *p, const double *in, cplx *out)
{
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz, Nzh = p->Nzh, C = p->C;
    const long nlines = (long)Nx * Ny * C;
#pragma omp parallel num_threads(p->nthreads)
    {
        cplx *a = (cplx *)malloc(sizeof(cplx) * (size_t)Nz);
        cplx *b = (cplx *)malloc(sizeof(cplx) * (size_t)Nz);
#pragma omp for schedule(static)
        for (long l = 0; l < nlines; ++l) {
            const long xy = l / C;
            const int c = (int)(l % C);
            const double *src = in + xy * Nz * C + c;
            cplx *dst = out + xy * Nzh * C + c;
            if (Nz % 2 == 0 && Nz >= 2) {
                const int L = Nz / 2;
                for (int m = 0; m < L; ++m) a[m] = src[(2 * m) * C] + I * src[(2 * m + 1) * C];
                fft1d(&p->pzh, a, b, -1);
                for (int k = 0; k <= L; ++k) {
                    cplx zk = a[k % L];
                    cplx zc = conj(a[(L - k) % L]);
                    cplx Ek = 0.5 * (zk + zc);
                    cplx Ok = -0.5 * I * (zk - zc);
                    cplx w = (k < Nz) ? p->pz.tw[k] : 1.0;
                    dst[(long)k * C] = Ek + w * Ok;
                }
            } else {
                for (int m = 0; m < Nz; ++m) a[m] = src[(long)m * C];
                fft1d(&p->pz, a, b, -1);
                for (int k = 0; k < Nzh; ++k) dst[(long)k * C] = a[k];
            }
        }
        free(a); free(b);
    }
    /* y: lines over y at fixed (x), inner (kz,c) contiguous */
    fft_axis(&p->py, out, Nx, (long)Ny * Nzh * C, (long)Nzh * C, (long)Nzh * C, -1, p->nthreads);
    /* x: inner (y,kz,c) contiguous */
    fft_axis(&p->px, out, 1, 0, (long)Ny * Nzh * C, (long)Ny * Nzh * C, -1, p->nthreads);
}

/* backward c2r: unnormalised; spec is overwritten (as FFTW's c2r destroys its input) */
static void irfft3(const plan3d *p, cplx *spec, double *out)
{
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz, Nzh = p->Nzh, C = p->C;
    fft_axis(&p->px, spec, 1, 0, (long)Ny * Nzh * C, (long)Ny * Nzh * C, +1, p->nthreads);
    fft_axis(&p->py, spec, Nx, (long)Ny * Nzh * C, (long)Nzh * C, (long)Nzh * C, +1, p->nthreads);
    const long nlines = (long)Nx * Ny * C;
#pragma omp parallel num_threads(p->nthreads)
    {
        cplx *a = (cplx *)malloc(sizeof(cplx) * (size_t)Nz);
        cplx *b = (cplx *)malloc(sizeof(cplx) * (size_t)Nz);
#pragma omp for schedule(static)
        for (long l = 0; l < nlines; ++l) {
            const long xy = l / C;
            const int c = (int)(l % C);
            const cplx *src = spec + xy * Nzh * C + c;
            double *dst = out + xy * Nz * C + c;
            if (Nz % 2 == 0 && Nz >= 2) {
                const int L = Nz / 2;
                for (int k = 0; k < L; ++k) {
                    cplx Xk = src[(long)k * C];
                    cplx XL = src[(long)(L - k) * C];
                    if (k == 0) { Xk = creal(Xk); XL = creal(XL); } /* kz=0 and kz=Nz/2 bins: real part only */
                    cplx s = Xk + conj(XL);
                    cplx d = Xk - conj(XL);
                    cplx w = conj(p->pz.tw[k]); /* exp(+2 pi i k / Nz) */
                    a[k] = s + I * w * d;
                }
                fft1d(&p->pzh, a, b, +1);
                for (int m = 0; m < L; ++m) {
                    dst[(long)(2 * m) * C] = creal(a[m]);
                    dst[(long)(2 * m + 1) * C] = cimag(a[m]);
                }
            } else {
                /* odd Nz: Hermitian extension, kz=0 real part only */
                a[0] = creal(src[0]);
                for (int k = 1; k < Nzh; ++k) {
                    a[k] = src[(long)k * C];
                    a[Nz - k] = conj(src[(long)k * C]);
                }
                fft1d(&p->pz, a, b, +1);
                for (int m = 0; m < Nz; ++m) dst[(long)m * C] = creal(a[m]);
            }
        }
        free(a); free(b);
    }
}

/* exported transform entry points used by the oracle tests */
void m3o_rfft3(int Nx, int Ny, int Nz, int C, const double *in, double *out_ri, int nthreads)
{
    plan3d p;
    plan3d_init(&p, Nx, Ny, Nz, C, nthreads);
    rfft3(&p, in, (cplx *)out_ri);
    plan3d_free(&p);
}

void m3o_irfft3(int Nx, int Ny, int Nz, int C, const double *spec_ri, double *out, int nthreads)
{
    plan3d p;
    plan3d_init(&p, Nx, Ny, Nz, C, nthreads);
    long H = (long)Nx * Ny * (Nz / 2 + 1) * C;
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)H);
    memcpy(tmp, spec_ri, sizeof(cplx) * (size_t)H);
    irfft3(&p, tmp, out);
    free(tmp);
    plan3d_free(&p);
}

/* ------------------------------------------------------------------------ */
/* small dense helpers (Eigen semantics: column-major 6x6 / 3x3)              */
/* ------------------------------------------------------------------------ */
/* Gauss-Jordan inverse with partial pivoting; returns 0 if singular. */
static int mat_inv(const double *A, double *Ainv, int n)
{
    double M[72];
    double X[72];
    for (int i = 0; i < n * n; ++i) { M[i] = A[i]; X[i] = 0; }
    for (int i = 0; i < n; ++i) X[i + i * n] = 1.0;
    for (int col = 0; col < n; ++col) {
        int piv = col;
        double best = fabs(M[col + col * n]);
        for (int r = col + 1; r < n; ++r)
            if (fabs(M[r + col * n]) > best) { best = fabs(M[r + col * n]); piv = r; }
        if (best == 0.0) return 0;
        if (piv != col)
            for (int k = 0; k < n; ++k) {
                double t = M[col + k * n]; M[col + k * n] = M[piv + k * n]; M[piv + k * n] = t;
                t = X[col + k * n]; X[col + k * n] = X[piv + k * n]; X[piv + k * n] = t;
            }
        double d = M[col + col * n];
        for (int k = 0; k < n; ++k) { M[col + k * n] /= d; X[col + k * n] /= d; }
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            double f = M[r + col * n];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) { M[r + k * n] -= f * M[col + k * n]; X[r + k * n] -= f * X[col + k * n]; }
        }
    }
    for (int i = 0; i < n * n; ++i) Ainv[i] = X[i];
    return 1;
}

static void mat_mul(const double *A, const double *B, double *Cm, int n)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double s = 0;
            for (int k = 0; k < n; ++k) s += A[i + k * n] * B[k + j * n];
            Cm[i + j * n] = s;
        }
}

static inline void gemv6(const double *A, const double *x, double *y)
{
    for (int i = 0; i < 6; ++i) {
        double s = 0;
        for (int k = 0; k < 6; ++k) s += A[i + 6 * k] * x[k];
        y[i] = s;
    }
}

/* isotropic stiffness in the reference's Voigt convention, FFT_analysis3D.cpp:457-462 */
static void iso_C(double la, double mu, double *C6)
{
    memset(C6, 0, sizeof(double) * 36);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C6[i + 6 * j] = la + (i == j ? 2 * mu : 0.0);
    for (int i = 3; i < 6; ++i) C6[i + 6 * i] = 2 * mu;
}

/* max_min_Array, FFT_analysis3D.cpp:5-23: first index of max / min */
static void max_min(const double *a, int n, double out[4])
{
    out[0] = a[0]; out[1] = 0; out[2] = a[0]; out[3] = 0;
    for (int i = 0; i < n - 1; ++i) {
        if (a[i + 1] > out[0]) { out[0] = a[i + 1]; out[1] = i + 1; }
        if (a[i + 1] < out[2]) { out[2] = a[i + 1]; out[3] = i + 1; }
    }
}

/* mechanical frequency map, FFT_analysis3D.cpp:202-209 / 372-379 */
static inline double fmap_m(int k, int N) { return (k <= (N - 1) / 2) ? (double)k : -(double)(N - k); }
/* thermal gradient-base / check map, FFT_thermal3D.cpp:113-120 / 321-328 */
static inline double fmap_t(int k, int N) { return (k <= N / 2) ? (double)k : -(double)(N - k); }

/* ------------------------------------------------------------------------ */
/* FFT_Analysis3D restatement                                                */
/* ------------------------------------------------------------------------ */
enum { BEH_E = 0, BEH_EP = 1, BEH_ED = 2 };

#define HIST_MAX 4096

typedef struct {
    int Nx, Ny, Nz, Nzh;
    long N, H;
    int behav;        /* current mater_behav */
    int behav_ctor;   /* behaviour given to the constructor (picks the reference medium) */
    int nph;
    double *El, *nu, *mu, *la, *K, *Hh, *nn, *Sy, *P, *Y0, *vol;
    double *C, *S, *Cacc, *Sacc;  /* nph x 36 */
    double C0[36], S0[36], la0, mu0;
    int strainBase, stressBase;
    int basic;                    /* 1: basic scheme in the "e" loop (update with C_acc = I) */
    int32_t *rve;
    double *eps, *sig;            /* 6 x N, column-major (voxel-major AoS) */
    double *t1;
    cplx *t2, *t3, *t4;
    double *GF;                   /* 21 x H */
    double E[6], X0[6], dE[6];
    int itNo;
    double sum_field_old, diff_sum_field, avr_eqVar, aver_eqVarOld, diff_avr_eqVar;
    /* "ep" / "ed" state, FFT_analysis3D.cpp:56-98 */
    double *eps_old, *sig_old, *ep_eff, *sig_eff, *phi, *dep, *d, *d_old, *Y;
    double nInc, incNo;
    double *prop;                 /* nInc x 9, column-major */
    int prop_rows;
    double sig_iter_old[6], eps_iter_old[6];
    plan3d plan;
    int nthreads;
    int cap;                      /* iteration cap, FFT_analysis3D.cpp:787 */
    /* per-iteration log: itNo, eqVar, diff_sum_field, diff_avr_eqVar */
    int hist_n;
    double hist[HIST_MAX * 4];
    int last_iters;
    int status;                   /* 0 ok, 1 Newton NaN (reference: std::terminate) */
} m3o_el;

static void *xcalloc(size_t n, size_t s)
{
    void *p = calloc(n ? n : 1, s);
    if (!p) { fprintf(stderr, "m3d_oracle: out of memory\n"); abort(); }
    return p;
}

/* constructor, FFT_analysis3D.cpp:25-152 */
m3o_el *m3o_el_create(int Nx, int Ny, int Nz, int behav, int nthreads)
{
    m3o_el *s = (m3o_el *)xcalloc(1, sizeof(m3o_el));
    s->Nx = Nx; s->Ny = Ny; s->Nz = Nz; s->Nzh = Nz / 2 + 1;
    s->N = (long)Nx * Ny * Nz;
    s->H = (long)Nx * Ny * s->Nzh;
    s->behav = behav; s->behav_ctor = behav;
    s->nthreads = nthreads > 0 ? nthreads : 1;
    s->cap = 2000;
    s->rve = (int32_t *)xcalloc(s->N, sizeof(int32_t));
    s->eps = (double *)xcalloc(6 * s->N, sizeof(double));
    s->sig = (double *)xcalloc(6 * s->N, sizeof(double));
    s->t1 = (double *)xcalloc(6 * s->N, sizeof(double));
    s->t2 = (cplx *)xcalloc(6 * s->H, sizeof(cplx));
    s->t3 = (cplx *)xcalloc(6 * s->H, sizeof(cplx));
    s->GF = (double *)xcalloc(21 * s->H, sizeof(double));
    if (behav != BEH_E) s->t4 = (cplx *)xcalloc(6 * s->H, sizeof(cplx));
    if (behav == BEH_EP) {
        s->eps_old = (double *)xcalloc(6 * s->N, sizeof(double));
        s->sig_old = (double *)xcalloc(6 * s->N, sizeof(double));
        s->ep_eff = (double *)xcalloc(s->N, sizeof(double));
        for (long i = 0; i < s->N; ++i) s->ep_eff[i] = .000000001236547; /* :67-69 */
        s->sig_eff = (double *)xcalloc(s->N, sizeof(double));
        s->phi = (double *)xcalloc(s->N, sizeof(double));
        s->dep = (double *)xcalloc(s->N, sizeof(double));
    }
    if (behav == BEH_ED) {
        s->eps_old = (double *)xcalloc(6 * s->N, sizeof(double));
        s->d = (double *)xcalloc(s->N, sizeof(double));
        s->d_old = (double *)xcalloc(s->N, sizeof(double));
        s->Y = (double *)xcalloc(s->N, sizeof(double));
    }
    plan3d_init(&s->plan, Nx, Ny, Nz, 6, s->nthreads);
    return s;
}

void m3o_el_destroy(m3o_el *s)
{
    if (!s) return;
    free(s->El); free(s->nu); free(s->mu); free(s->la); free(s->K); free(s->Hh); free(s->nn);
    free(s->Sy); free(s->P); free(s->Y0); free(s->vol);
    free(s->C); free(s->S); free(s->Cacc); free(s->Sacc);
    free(s->rve); free(s->eps); free(s->sig); free(s->t1); free(s->t2); free(s->t3); free(s->t4);
    free(s->GF); free(s->eps_old); free(s->sig_old); free(s->ep_eff); free(s->sig_eff);
    free(s->phi); free(s->dep); free(s->d); free(s->d_old); free(s->Y); free(s->prop);
    plan3d_free(&s->plan);
    free(s);
}

void m3o_el_set_cap(m3o_el *s, int cap) { s->cap = cap; }

/* Threads of the FFTs and of the per-voxel loops, set apart: the reference
 * runs FFTW with 3 threads (FFT_analysis3D.cpp:101-102) and its voxel loops
 * serially, which (3, 1) reproduces for the CPU baseline. */
void m3o_el_set_threads(m3o_el *s, int fft_threads, int loop_threads)
{
    s->plan.nthreads = fft_threads > 0 ? fft_threads : 1;
    s->nthreads = loop_threads > 0 ? loop_threads : 1;
}

/* Basic Moulinec-Suquet scheme with the load held fixed (strain control): the
 * "e" loop with the update replaced by eps <- eps_c (sigma <- sigma_c in the
 * stress base), i.e. C_acc = I.  The reference runs its basic scheme only
 * with the updateE mixed control (FFT_analysis3D.cpp:796-853); this is the
 * plain fixed-E variant of the same iteration. */
void m3o_el_set_scheme(m3o_el *s, int basic) { s->basic = basic ? 1 : 0; }

/* setGF, FFT_analysis3D.cpp:155-233: 21 unique Gamma0 (strain base) or Delta0
 * (stress base) components per half-spectrum frequency. */
static const int QTAB[21][4] = {
    {0, 0, 0, 0}, {0, 0, 1, 1}, {0, 0, 2, 2}, {0, 0, 1, 2}, {0, 0, 0, 2}, {0, 0, 0, 1},
    {1, 1, 1, 1}, {1, 1, 2, 2}, {1, 1, 1, 2}, {1, 1, 0, 2}, {1, 1, 0, 1},
    {2, 2, 2, 2}, {2, 2, 1, 2}, {2, 2, 0, 2}, {2, 2, 0, 1},
    {1, 2, 1, 2}, {1, 2, 0, 2}, {1, 2, 0, 1},
    {2, 0, 0, 2}, {2, 0, 0, 1},
    {0, 1, 0, 1}};
/* symmetric 6x6 index map w(j,b), FFT_analysis3D.cpp:299-305 */
static const int WMAP[6][6] = {{0, 1, 2, 3, 4, 5},     {1, 6, 7, 8, 9, 10},    {2, 7, 11, 12, 13, 14},
                               {3, 8, 12, 15, 16, 17}, {4, 9, 13, 16, 18, 19}, {5, 10, 14, 17, 19, 20}};

static inline double dlt(int i, int j) { return i == j ? 1.0 : 0.0; }

/* one frequency's 21 components; zet = {z, y, x} (tensor index 0 = z axis, :211) */
static void green21(const double zet[3], double la0, double mu0, int strainBase, double *g)
{
    double c1 = 2 * mu0 * (3 * la0 + 2 * mu0) / (la0 + 2 * mu0);
    double c2 = 2 * mu0;
    double c3 = (c1 - c2) / 2.;
    double c4 = (c2) / 2.;
    double d1 = 1 / (4 * mu0);
    double d2 = (la0 + mu0) / (mu0 * (la0 + 2 * mu0));
    double n2 = zet[0] * zet[0] + zet[1] * zet[1] + zet[2] * zet[2];
    for (int s = 0; s < 21; ++s) {
        int i = QTAB[s][0], j = QTAB[s][1], k = QTAB[s][2], h = QTAB[s][3];
        if (strainBase)
            g[s] = (d1 / n2) * (dlt(k, i) * zet[h] * zet[j] + dlt(h, i) * zet[k] * zet[j] +
                                dlt(k, j) * zet[h] * zet[i] + dlt(h, j) * zet[k] * zet[i]) -
                   (d2 * zet[i] * zet[j] * zet[k] * zet[h]) / (n2 * n2);
        else
            g[s] = c3 * (dlt(i, j) - zet[i] * zet[j] / n2) * (dlt(k, h) - zet[k] * zet[h] / n2) +
                   c4 * (dlt(i, k) - zet[i] * zet[k] / n2) * (dlt(j, h) - zet[j] * zet[h] / n2) +
                   c4 * (dlt(i, h) - zet[i] * zet[h] / n2) * (dlt(j, k) - zet[j] * zet[k] / n2);
    }
}

static void el_setGF(m3o_el *s)
{
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz, Nzh = s->Nzh;
#pragma omp parallel for num_threads(s->nthreads) collapse(2)
    for (int x = 0; x < Nx; x++)
        for (int y = 0; y < Ny; y++)
            for (int z = 0; z < Nzh; z++) {
                double zet[3] = {fmap_m(z, Nz), fmap_m(y, Ny), fmap_m(x, Nx)};
                long cc = ((long)z + (long)y * Nzh + (long)x * Nzh * Ny) * 21;
                green21(zet, s->la0, s->mu0, s->strainBase, s->GF + cc);
            }
}

/* material, FFT_analysis3D.cpp:420-549.  mat is nph x 7 column-major
 * [E, nu, sigma_y, H, n, P, Y0] (example/3d/main.cpp:104-111). */
int m3o_el_material(m3o_el *s, const double *mat, const int32_t *rve, int nph)
{
    s->strainBase = 0; s->stressBase = 0;
    s->nph = nph;
#define ALLOC(p) do { free(s->p); s->p = (double *)xcalloc(nph, sizeof(double)); } while (0)
    ALLOC(El); ALLOC(nu); ALLOC(mu); ALLOC(la); ALLOC(K); ALLOC(Hh); ALLOC(nn); ALLOC(Sy); ALLOC(P); ALLOC(Y0); ALLOC(vol);
#undef ALLOC
    free(s->C); free(s->S); free(s->Cacc); free(s->Sacc);
    s->C = (double *)xcalloc(36 * nph, sizeof(double));
    s->S = (double *)xcalloc(36 * nph, sizeof(double));
    s->Cacc = (double *)xcalloc(36 * nph, sizeof(double));
    s->Sacc = (double *)xcalloc(36 * nph, sizeof(double));
    for (int i = 0; i < nph; ++i) {
        s->El[i] = mat[i + 0 * nph];
        s->nu[i] = mat[i + 1 * nph];
        s->Sy[i] = mat[i + 2 * nph];
        s->Hh[i] = mat[i + 3 * nph];
        s->nn[i] = mat[i + 4 * nph];
        s->P[i] = mat[i + 5 * nph];
        s->Y0[i] = mat[i + 6 * nph];
        s->mu[i] = s->El[i] / (1 + s->nu[i]) / 2.;
        s->la[i] = s->nu[i] * s->El[i] / (1 + s->nu[i]) / (1 - 2 * s->nu[i]);
        s->K[i] = s->la[i] + (2. / 3.) * s->mu[i];
        iso_C(s->la[i], s->mu[i], s->C + 36 * i);
        if (!mat_inv(s->C + 36 * i, s->S + 36 * i, 6)) return -1;
    }
    for (long i = 0; i < s->N; ++i) {
        if (rve[i] < 0 || rve[i] >= nph) return -2;
        s->rve[i] = rve[i];
    }
    long *cnt = (long *)xcalloc(nph, sizeof(long));
    for (long i = 0; i < s->N; ++i) cnt[s->rve[i]]++;
    for (int k = 0; k < nph; ++k) s->vol[k] = (double)cnt[k] / (double)s->N;
    free(cnt);

    double t[4];
    max_min(s->mu, nph, t);
    double mu_max = t[0], mu_min = t[2];
    int index_mu_max = (int)t[1];
    max_min(s->la, nph, t);
    double la_max = t[0], la_min = t[2];
    max_min(s->K, nph, t);
    double K_max = t[0], K_min = t[2];
    max_min(s->vol, nph, t);
    double vol_frac_max = t[0];

    if (s->behav_ctor == BEH_E) {
        double K0 = sqrt(K_min * K_max);                  /* :498-501 */
        s->mu0 = sqrt(mu_min * mu_max);
        s->la0 = K0 - s->mu0 * 2. / 3.;
        if (s->vol[index_mu_max] == vol_frac_max) s->strainBase = 1; /* :503-512 */
        else s->stressBase = 1;
    } else {
        s->strainBase = 1;                               /* :517-521 */
        s->mu0 = (mu_min + mu_max) / 2;
        s->la0 = (la_min + la_max) / 2;
    }
    iso_C(s->la0, s->mu0, s->C0);
    if (!mat_inv(s->C0, s->S0, 6)) return -1;
    for (int i = 0; i < nph; ++i) {                      /* :539-546 */
        double A[36], Ai[36];
        for (int q = 0; q < 36; ++q) A[q] = s->C[36 * i + q] + s->C0[q];
        if (!mat_inv(A, Ai, 6)) return -1;
        mat_mul(Ai, s->C0, s->Cacc + 36 * i, 6);
        for (int q = 0; q < 36; ++q) s->Cacc[36 * i + q] *= 2;
        for (int q = 0; q < 36; ++q) A[q] = s->S[36 * i + q] + s->S0[q];
        if (!mat_inv(A, Ai, 6)) return -1;
        mat_mul(Ai, s->S0, s->Sacc + 36 * i, 6);
        for (int q = 0; q < 36; ++q) s->Sacc[36 * i + q] *= 2;
    }
    el_setGF(s);
    return 0;
}

/* setLoad, FFT_analysis3D.cpp:551-564 */
void m3o_el_set_load(m3o_el *s, const double *E1, const double *X01)
{
    for (int i = 0; i < 6; ++i) { s->E[i] = E1[i]; s->X0[i] = X01[i]; }
    if (s->strainBase) {
        for (long i = 0; i < s->N; ++i) memcpy(s->eps + 6 * i, E1, sizeof(double) * 6);
    } else if (s->stressBase) {
        for (long i = 0; i < s->N; ++i) memcpy(s->sig + 6 * i, E1, sizeof(double) * 6);
    }
}

/* multipCx, FFT_analysis3D.cpp:236-247 */
static void el_multipCx(m3o_el *s)
{
#pragma omp parallel for num_threads(s->nthreads)
    for (long i = 0; i < s->N; ++i) {
        if (s->strainBase) gemv6(s->C + 36 * s->rve[i], s->eps + 6 * i, s->sig + 6 * i);
        else gemv6(s->S + 36 * s->rve[i], s->sig + 6 * i, s->eps + 6 * i);
    }
}

/* update, FFT_analysis3D.cpp:283-295 */
static void el_update(m3o_el *s)
{
#pragma omp parallel for num_threads(s->nthreads)
    for (long i = 0; i < s->N; ++i) {
        double dlt6[6], y[6];
        if (s->basic) {
            if (s->strainBase) for (int c = 0; c < 6; ++c) s->eps[6 * i + c] = s->sig[6 * i + c];
            else for (int c = 0; c < 6; ++c) s->sig[6 * i + c] = s->eps[6 * i + c];
            continue;
        }
        if (s->strainBase) {
            for (int c = 0; c < 6; ++c) dlt6[c] = s->sig[6 * i + c] - s->eps[6 * i + c];
            gemv6(s->Cacc + 36 * s->rve[i], dlt6, y);
            for (int c = 0; c < 6; ++c) s->eps[6 * i + c] += y[c];
        } else {
            for (int c = 0; c < 6; ++c) dlt6[c] = s->eps[6 * i + c] - s->sig[6 * i + c];
            gemv6(s->Sacc + 36 * s->rve[i], dlt6, y);
            for (int c = 0; c < 6; ++c) s->sig[6 * i + c] += y[c];
        }
    }
}

/* multiplyGF, FFT_analysis3D.cpp:297-348 */
static void el_multiplyGF(m3o_el *s)
{
    const long H = s->H;
    const int e_mode = (s->behav == BEH_E);
#pragma omp parallel for num_threads(s->nthreads)
    for (long i = 1; i < H; i++) {
        const long u = i * 6;
        const double *g = s->GF + i * 21;
        for (int j = 0; j < 6; j++) {
            cplx acc = g[WMAP[j][0]] * s->t2[u + 0] + g[WMAP[j][1]] * s->t2[u + 1] + g[WMAP[j][2]] * s->t2[u + 2] +
                       g[WMAP[j][3]] * s->t2[u + 3] * 2 + g[WMAP[j][4]] * s->t2[u + 4] * 2 +
                       g[WMAP[j][5]] * s->t2[u + 5] * 2;
            s->t3[u + j] = e_mode ? -acc : (s->t4[u + j] - acc);
        }
    }
    for (int i = 0; i < 6; ++i) s->t3[i] = s->E[i] * (double)s->N;
}

/* c2r + scale, FFTBCKWRD_sigma / FFTBCKWRD_epsilon, FFT_analysis3D.cpp:266-280 */
static void el_backward(m3o_el *s, double *dst)
{
    irfft3(&s->plan, s->t3, dst);
    const double inv = (double)s->N;
#pragma omp parallel for num_threads(s->nthreads)
    for (long i = 0; i < 6 * s->N; ++i) dst[i] = dst[i] / inv;
}

static double sum_all(const double *a, long n, int nthreads)
{
    double s = 0;
#pragma omp parallel for reduction(+ : s) num_threads(nthreads)
    for (long i = 0; i < n; ++i) s += a[i];
    return s;
}

/* convergenceCheck, FFT_analysis3D.cpp:351-416 */
static void el_convergenceCheck(m3o_el *s)
{
    rfft3(&s->plan, s->sig, s->t2);
    s->itNo += 1;
    const cplx *t2 = s->t2;
#define A2(q) (creal(t2[q]) * creal(t2[q]) + cimag(t2[q]) * cimag(t2[q]))
    double dd = sqrt(A2(0) + A2(1) + A2(2) + 2 * A2(3) + 2 * A2(4) + 2 * A2(5));
    double cc = 0;
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz, Nzh = s->Nzh;
#pragma omp parallel for reduction(+ : cc) num_threads(s->nthreads) collapse(2)
    for (int x = 0; x < Nx; x++)
        for (int y = 0; y < Ny; y++)
            for (int z = 1; z < Nzh; z++) {
                long i = ((long)z + (long)y * Nzh + (long)x * Nzh * Ny) * 6;
                double zt[3] = {fmap_m(z, Nz), fmap_m(y, Ny), fmap_m(x, Nx)};
                cc += sqrt(fabs(A2(i + 0) * zt[0] * zt[0] + A2(i + 1) * zt[1] * zt[1] + A2(i + 2) * zt[2] * zt[2] +
                                A2(i + 3) * ((zt[1] + zt[1]) + (zt[2] + zt[2])) +
                                A2(i + 4) * ((zt[0] + zt[0]) + (zt[2] + zt[2])) +
                                A2(i + 5) * ((zt[0] + zt[0]) + (zt[1] + zt[1]))));
            }
#undef A2
    s->avr_eqVar = cc / dd / (double)s->N;
    double Esum = 0;
    for (int i = 0; i < 6; ++i) Esum += s->E[i];
    if (s->strainBase) {
        double tot = sum_all(s->sig, 6 * s->N, s->nthreads);
        s->diff_sum_field = fabs((tot - s->sum_field_old) /
                                 ((double)s->N * (0.5 * s->C[0] + 0.5 * s->C[36 * (s->nph > 1 ? 1 : 0)]) * Esum));
    } else {
        double tot = sum_all(s->eps, 6 * s->N, s->nthreads);
        s->diff_sum_field = fabs((tot - s->sum_field_old) /
                                 ((double)s->N * (0.5 * s->S[0] + 0.5 * s->S[36 * (s->nph > 1 ? 1 : 0)]) * Esum));
    }
    s->diff_avr_eqVar = fabs(s->avr_eqVar - s->aver_eqVarOld);
    s->aver_eqVarOld = s->avr_eqVar;
    if (s->hist_n < HIST_MAX) {
        double *h = s->hist + 4 * s->hist_n++;
        h[0] = s->itNo; h[1] = s->avr_eqVar; h[2] = s->diff_sum_field; h[3] = s->diff_avr_eqVar;
    }
}

/* updateE with old_iter_fields_update and prod, FFT_analysis3D.cpp:566-594 */
static double prod6(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + 2 * a[3] * b[3] + 2 * a[4] * b[4] + 2 * a[5] * b[5];
}

static void el_updateE(m3o_el *s)
{
    for (int i = 0; i < 6; ++i) {
        double ss = 0, se = 0;
        for (long v = 0; v < s->N; ++v) { ss += s->sig[6 * v + i]; se += s->eps[6 * v + i]; }
        s->sig_iter_old[i] = ss / (double)s->N;
        s->eps_iter_old[i] = se / (double)s->N;
    }
    double x[6], y[6], S0s[6];
    gemv6(s->S0, s->sig_iter_old, S0s);
    for (int i = 0; i < 6; ++i) x[i] = S0s[i] - s->eps_iter_old[i];
    gemv6(s->S0, s->X0, y);
    double dEsum = 0;
    for (int i = 0; i < 6; ++i) dEsum += s->dE[i];
    double k = (dEsum * s->incNo + prod6(x, s->X0)) / prod6(y, s->X0);
    for (int i = 0; i < 6; ++i) s->E[i] = k * y[i] - S0s[i] + s->eps_iter_old[i];
}

/* sigma_eff_calculator, FFT_analysis3D.cpp:597-606; returns sigma_d in sd */
static double sig_eff_of(const double *sg, double *sd)
{
    double sm = (sg[0] + sg[1] + sg[2]) / 3.;
    for (int c = 0; c < 6; ++c) sd[c] = sg[c] - (c < 3 ? sm : 0.0);
    return sqrt(3. / 2.) * sqrt(sd[0] * sd[0] + sd[1] * sd[1] + sd[2] * sd[2] + 2 * sd[3] * sd[3] + 2 * sd[4] * sd[4] +
                               2 * sd[5] * sd[5]);
}

/* elastoPlastic, FFT_analysis3D.cpp:640-691 */
static void el_elastoPlastic(m3o_el *s)
{
    s->sum_field_old = sum_all(s->sig, 6 * s->N, s->nthreads);
    int bad = 0;
#pragma omp parallel for num_threads(s->nthreads) reduction(| : bad)
    for (long i = 0; i < s->N; ++i) {
        const int k = s->rve[i];
        const double *Ck = s->C + 36 * k;
        double de[6], y[6], sd[6];
        for (int c = 0; c < 6; ++c) de[c] = s->eps[6 * i + c] - s->eps_old[6 * i + c];
        gemv6(Ck, de, y);
        double *sg = s->sig + 6 * i;
        for (int c = 0; c < 6; ++c) sg[c] = s->sig_old[6 * i + c] + y[c];
        s->sig_eff[i] = sig_eff_of(sg, sd);
        s->phi[i] = s->sig_eff[i] - s->Sy[k] - s->Hh[k] * pow(s->ep_eff[i], s->nn[k]);
        if (s->phi[i] > 0) {
            double x = 0, x_o = 0, res;
            const double mu = s->mu[k], Hk = s->Hh[k], nk = s->nn[k], Syk = s->Sy[k], se = s->sig_eff[i],
                         ep = s->ep_eff[i];
            int guard = 0;
            while (1) {
                x = x_o - (se - 3 * mu * x_o - Syk - Hk * pow((ep + x_o), nk)) /
                              (-3 * mu - Hk * nk * pow((ep + x_o), nk - 1));
                x_o = x;
                res = (se - 3 * mu * x - Syk - Hk * pow((ep + x), nk));
                if (isnan(x)) { bad = 1; break; } /* reference: bare throw -> terminate (:672-673) */
                if (fabs(res) < 1e-5) break;
                if (++guard > 100000) { bad = 1; break; }
            }
            s->dep[i] = x;
            double cs[6];
            gemv6(Ck, sd, cs);
            double f = (1.5 * x / se);
            for (int c = 0; c < 6; ++c) sg[c] -= f * cs[c];
            s->sig_eff[i] = sig_eff_of(sg, sd);
        } else {
            s->dep[i] = 0;
        }
    }
    if (bad) s->status = 1;
}

/* elastoDamage, FFT_analysis3D.cpp:619-638 */
static void el_elastoDamage(m3o_el *s)
{
    s->sum_field_old = sum_all(s->sig, 6 * s->N, s->nthreads);
#pragma omp parallel for num_threads(s->nthreads)
    for (long i = 0; i < s->N; ++i) {
        const int k = s->rve[i];
        const double *Ck = s->C + 36 * k;
        const double *e = s->eps + 6 * i;
        double t[6];
        gemv6(Ck, e, t);
        t[3] *= 2; t[4] *= 2; t[5] *= 2;
        double q = 0;
        for (int c = 0; c < 6; ++c) q += t[c] * e[c];
        s->Y[i] = sqrt(q);
        s->d_old[i] = s->d[i];
        if (s->Y[i] > s->Y0[k]) {
            double dn = 1 - exp(-s->P[k] * (s->Y[i] - s->Y0[k]));
            s->d[i] = dn > s->d_old[i] ? dn : s->d_old[i];
        }
        double ce[6];
        gemv6(Ck, e, ce);
        for (int c = 0; c < 6; ++c) s->sig[6 * i + c] = (1 - s->d[i]) * ce[c];
    }
}

static void el_constitutive(m3o_el *s)
{
    if (s->behav == BEH_EP) el_elastoPlastic(s);
    else if (s->behav == BEH_ED) el_elastoDamage(s);
}

/* update_old_Variable, FFT_analysis3D.cpp:693-719 */
static void el_update_old(m3o_el *s)
{
    if (s->behav == BEH_EP) {
        for (long i = 0; i < s->N; ++i) {
            if (s->phi[i] > 0) s->ep_eff[i] += s->dep[i];
            memcpy(s->sig_old + 6 * i, s->sig + 6 * i, sizeof(double) * 6);
        }
    }
    for (long i = 0; i < 6 * s->N; ++i) {
        double tmp = s->eps_old[i];
        s->eps_old[i] = s->eps[i];
        s->eps[i] += (s->eps[i] - tmp);
    }
}

/* export_homogen_var, FFT_analysis3D.cpp:722-746 */
static void el_export(m3o_el *s, int row)
{
    int t = 0;
    for (int j = 0; j < 6; ++j)
        if (s->X0[j] != 0) { t = j; break; }
    double *P = s->prop;
    const int R = s->prop_rows;
    double se = 0;
    for (long v = 0; v < s->N; ++v) se += s->eps[6 * v + t];
    P[row + 0 * R] = se / (double)s->N;
    for (int j = 1; j < 7; ++j) {
        double ss = 0;
        for (long v = 0; v < s->N; ++v) ss += s->sig[6 * v + j - 1];
        P[row + j * R] = ss / (double)s->N;
    }
    if (s->behav == BEH_EP) {
        P[row + 7 * R] = sum_all(s->sig_eff, s->N, s->nthreads) / (double)s->N;
        P[row + 8 * R] = sum_all(s->ep_eff, s->N, s->nthreads) / (double)s->N;
    }
    if (s->behav == BEH_ED) {
        P[row + 7 * R] = sum_all(s->d, s->N, s->nthreads) / (double)s->N;
        P[row + 8 * R] = 0;
    }
}

/* runAnalysis, FFT_analysis3D.cpp:749-854.  Returns itNo of the last loop. */
int m3o_el_run(m3o_el *s, double tol_field, double tol_eq, int nInc1)
{
    s->status = 0;
    if (s->behav == BEH_E) {
        el_multipCx(s);
        while (1) {
            if (s->strainBase) {
                s->sum_field_old = sum_all(s->sig, 6 * s->N, s->nthreads);
#pragma omp parallel for num_threads(s->nthreads)
                for (long i = 0; i < s->N; ++i) {
                    double y[6];
                    gemv6(s->C0, s->eps + 6 * i, y);
                    for (int c = 0; c < 6; ++c) s->t1[6 * i + c] = s->sig[6 * i + c] - y[c];
                }
            } else {
                s->sum_field_old = sum_all(s->eps, 6 * s->N, s->nthreads);
#pragma omp parallel for num_threads(s->nthreads)
                for (long i = 0; i < s->N; ++i) {
                    double y[6];
                    gemv6(s->S0, s->sig + 6 * i, y);
                    for (int c = 0; c < 6; ++c) s->t1[6 * i + c] = s->eps[6 * i + c] - y[c];
                }
            }
            rfft3(&s->plan, s->t1, s->t2);                     /* FFT_sigma / FFT_epsilon */
            el_multiplyGF(s);
            if (s->strainBase) el_backward(s, s->sig);         /* sigma <- epsilon_c */
            else el_backward(s, s->eps);                       /* epsilon <- sigma_c */
            el_update(s);
            el_multipCx(s);
            el_convergenceCheck(s);
            if (s->itNo > 1 && fabs(s->diff_avr_eqVar) < tol_eq && fabs(s->diff_sum_field) < tol_field) break;
            if (s->itNo > s->cap) break;
        }
        s->last_iters = s->itNo;
        return s->itNo;
    }
    /* "ep" / "ed": incremental basic scheme, :796-853 */
    s->nInc = nInc1;
    for (int i = 0; i < 6; ++i) { s->dE[i] = s->E[i] / (s->nInc * 1.); s->E[i] = s->dE[i]; }
    {
        double Ec[6], Xc[6];
        memcpy(Ec, s->E, sizeof Ec); memcpy(Xc, s->X0, sizeof Xc);
        m3o_el_set_load(s, Ec, Xc);
    }
    free(s->prop);
    s->prop_rows = nInc1;
    s->prop = (double *)xcalloc((size_t)nInc1 * 9, sizeof(double));
    s->incNo = 0;
    int nx = 0;
    int total = 0;
    for (int ii = 0; ii < s->nInc; ++ii) {
        s->incNo += 1;
        s->itNo = 0;
        el_constitutive(s);
        while (1) {
            rfft3(&s->plan, s->sig, s->t2);   /* FFT_sigma */
            rfft3(&s->plan, s->eps, s->t4);   /* FFT_epsilon -> t4 */
            el_updateE(s);
            el_multiplyGF(s);
            el_backward(s, s->eps);           /* FFTBCKWRD_epsilon */
            el_constitutive(s);
            el_convergenceCheck(s);
            if (s->status) return -1;
            if (s->itNo > 1 && fabs(s->diff_avr_eqVar) < tol_eq && fabs(s->diff_sum_field) < tol_field) break;
            if (s->itNo > s->cap) break;
        }
        total += s->itNo;
        if (s->behav == BEH_ED && ii != 0) {
            /* softening exit, :829-847.  The reference compares row ii before
             * export_homogen_var(ii) writes it (reads an unwritten row); this
             * restatement evaluates the current means into row ii first. */
            el_export(s, ii);
            const double *P = s->prop;
            const int R = s->prop_rows;
            for (int j = 0; j < 6; ++j)
                if (s->X0[j] != 0 && P[ii + (j + 1) * R] < P[ii - 1 + (j + 1) * R]) nx += 1;
        }
        if (nx == 3) { s->nInc = ii; break; }
        el_export(s, ii);
        el_update_old(s);
        for (int i = 0; i < 6; ++i) s->E[i] += s->dE[i];
    }
    s->last_iters = total;
    return total;
}

/* homogenization, FFT_analysis3D.cpp:856-999 (report text omitted).
 * out: C[36], S[36], epHomo[36], sigHomo[36] column-major, hs[6] =
 * {E_min,K_min,G_min,E_max,K_max,G_max}; returns 1/S11+1/S22+1/S33. */
double m3o_el_homogenize(m3o_el *s, double tol_field, double tol_eq, double *Cout, double *Sout, double *ep,
                         double *sg, double *hs, int *iters)
{
    s->behav = BEH_E;  /* :858 */
    double epH[36] = {0}, sgH[36] = {0};
    for (int w = 0; w < 6; ++w) {
        s->itNo = 0;
        double E[6] = {0, 0, 0, 0, 0, 0};
        E[w] = s->strainBase ? 0.0005 : 1.e6;
        m3o_el_set_load(s, E, E);
        m3o_el_run(s, tol_field, tol_eq, 1);
        if (iters) iters[w] = s->itNo;
        for (int t = 0; t < 6; ++t) {
            double ss = 0, se = 0;
            for (long v = 0; v < s->N; ++v) { ss += s->sig[6 * v + t]; se += s->eps[6 * v + t]; }
            sgH[w + 6 * t] = ss / (double)s->N;  /* sigHomo(w,t) */
            epH[t + 6 * w] = se / (double)s->N;  /* epHomo(t,w)  */
        }
    }
    double Ch[36], Sh[36], inv[36];
    if (s->strainBase) {
        mat_inv(epH, inv, 6);
        mat_mul(inv, sgH, Ch, 6);
        mat_inv(Ch, Sh, 6);
    } else {
        mat_inv(sgH, inv, 6);
        mat_mul(inv, epH, Sh, 6);
        mat_inv(Sh, Ch, 6);
    }
    memcpy(Cout, Ch, sizeof Ch);
    memcpy(Sout, Sh, sizeof Sh);
    memcpy(ep, epH, sizeof epH);
    memcpy(sg, sgH, sizeof sgH);
    /* Hashin-Shtrikman bounds, :911-942 */
    double t[4];
    max_min(s->mu, s->nph, t);
    double Gn = t[0], G1 = t[2];
    max_min(s->K, s->nph, t);
    double Kn = t[0], K1 = t[2];
    double alfa1 = -3 / (3 * K1 + 4 * G1);
    double alfan = -3 / (3 * Kn + 4 * Gn);
    double beta1 = -(3 * (K1 + 2 * G1)) / ((3 * K1 + 4 * G1) * 5 * G1);
    double betan = -(3 * (Kn + 2 * Gn)) / ((3 * Kn + 4 * Gn) * 5 * Gn);
    double A1 = 0, An = 0, B1 = 0, Bn = 0;
    for (int i = 0; i < s->nph; ++i) {
        if (s->K[i] != K1) A1 += s->vol[i] / ((1. / (s->K[i] - K1)) - alfa1);
        if (s->K[i] != Kn) An += s->vol[i] / ((1. / (s->K[i] - Kn)) - alfan);
        if (s->mu[i] != G1) B1 += s->vol[i] / ((0.5 / (s->mu[i] - G1)) - beta1);
        if (s->mu[i] != Gn) Bn += s->vol[i] / ((0.5 / (s->mu[i] - Gn)) - betan);
    }
    double K_min = K1 + A1 / (1 + alfa1 * A1);
    double K_max = Kn + An / (1 + alfan * An);
    double G_min = G1 + 0.5 * B1 / (1 + beta1 * B1);
    double G_max = Gn + 0.5 * Bn / (1 + betan * Bn);
    hs[0] = 9 * K_min * G_min / (3 * K_min + G_min);
    hs[1] = K_min;
    hs[2] = G_min;
    hs[3] = 9 * K_max * G_max / (3 * K_max + G_max);
    hs[4] = K_max;
    hs[5] = G_max;
    return 1. / Sh[0] + 1. / Sh[7] + 1. / Sh[14];
}

/* accessors used by the tests */
void m3o_el_fields(const m3o_el *s, double *eps, double *sig)
{
    if (eps) memcpy(eps, s->eps, sizeof(double) * 6 * s->N);
    if (sig) memcpy(sig, s->sig, sizeof(double) * 6 * s->N);
}

/* info: {strainBase, la0, mu0, itNo, eqVar, diff_sum, diff_eq, status} */
void m3o_el_info(const m3o_el *s, double *out)
{
    out[0] = s->strainBase; out[1] = s->la0; out[2] = s->mu0; out[3] = s->itNo;
    out[4] = s->avr_eqVar; out[5] = s->diff_sum_field; out[6] = s->diff_avr_eqVar; out[7] = s->status;
}

void m3o_el_matrices(const m3o_el *s, double *C, double *S, double *Cacc, double *Sacc, double *C0, double *S0, double *vol)
{
    memcpy(C, s->C, sizeof(double) * 36 * s->nph);
    memcpy(S, s->S, sizeof(double) * 36 * s->nph);
    memcpy(Cacc, s->Cacc, sizeof(double) * 36 * s->nph);
    memcpy(Sacc, s->Sacc, sizeof(double) * 36 * s->nph);
    memcpy(C0, s->C0, sizeof(double) * 36);
    memcpy(S0, s->S0, sizeof(double) * 36);
    memcpy(vol, s->vol, sizeof(double) * s->nph);
}

int m3o_el_history(const m3o_el *s, double *out, int maxn)
{
    int n = s->hist_n < maxn ? s->hist_n : maxn;
    memcpy(out, s->hist, sizeof(double) * 4 * n);
    return n;
}

void m3o_el_clear_history(m3o_el *s) { s->hist_n = 0; }

void m3o_el_prop(const m3o_el *s, double *out, int *rows)
{
    *rows = s->prop_rows;
    if (s->prop && out) memcpy(out, s->prop, sizeof(double) * 9 * s->prop_rows);
}

/* J2 history per voxel: ep_eff, sig_eff */
void m3o_el_j2_state(const m3o_el *s, double *ep_eff, double *sig_eff)
{
    if (s->ep_eff && ep_eff) memcpy(ep_eff, s->ep_eff, sizeof(double) * s->N);
    if (s->sig_eff && sig_eff) memcpy(sig_eff, s->sig_eff, sizeof(double) * s->N);
}

void m3o_el_damage_state(const m3o_el *s, double *d)
{
    if (s->d && d) memcpy(d, s->d, sizeof(double) * s->N);
}

/* Green table at one frequency and the 6x6 it encodes, for identity tests */
void m3o_green21(const double *zet, double la0, double mu0, int strainBase, double *g21)
{
    green21(zet, la0, mu0, strainBase, g21);
}

/* apply multiplyGF to an arbitrary spectrum t2 (6 x H interleaved complex) */
void m3o_el_apply_green(m3o_el *s, const double *t2_ri, double *t3_ri)
{
    memcpy(s->t2, t2_ri, sizeof(cplx) * 6 * s->H);
    el_multiplyGF(s);
    memcpy(t3_ri, s->t3, sizeof(cplx) * 6 * s->H);
}

/* one accelerated iteration's building blocks, for golden vectors */
void m3o_el_eqvar(m3o_el *s, const double *sig, double *out3)
{
    /* out3 = {eqVar, cc, dd} of an arbitrary sigma field (convergenceCheck's first half) */
    rfft3(&s->plan, sig, s->t2);
    const cplx *t2 = s->t2;
#define A2(q) (creal(t2[q]) * creal(t2[q]) + cimag(t2[q]) * cimag(t2[q]))
    double dd = sqrt(A2(0) + A2(1) + A2(2) + 2 * A2(3) + 2 * A2(4) + 2 * A2(5));
    double cc = 0;
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz, Nzh = s->Nzh;
    for (int x = 0; x < Nx; x++)
        for (int y = 0; y < Ny; y++)
            for (int z = 1; z < Nzh; z++) {
                long i = ((long)z + (long)y * Nzh + (long)x * Nzh * Ny) * 6;
                double zt[3] = {fmap_m(z, Nz), fmap_m(y, Ny), fmap_m(x, Nx)};
                cc += sqrt(fabs(A2(i + 0) * zt[0] * zt[0] + A2(i + 1) * zt[1] * zt[1] + A2(i + 2) * zt[2] * zt[2] +
                                A2(i + 3) * ((zt[1] + zt[1]) + (zt[2] + zt[2])) +
                                A2(i + 4) * ((zt[0] + zt[0]) + (zt[2] + zt[2])) +
                                A2(i + 5) * ((zt[0] + zt[0]) + (zt[1] + zt[1]))));
            }
#undef A2
    out3[0] = cc / dd / (double)s->N;
    out3[1] = cc;
    out3[2] = dd;
}

/* ------------------------------------------------------------------------ */
/* FFT_analysis3d_thermal restatement                                         */
/* ------------------------------------------------------------------------ */
typedef struct {
    int Nx, Ny, Nz, Nzh;
    long N, H;
    int nph;
    double *k;          /* conductivity per phase id */
    double k0;
    int gradBase, fluxBase;
    int32_t *rve;
    double *tg, *q;     /* 3 x N: tempGradient, heatFlux */
    cplx *t2, *t3;
    double *GF;         /* 6 x H */
    double E[3];
    int itNo;
    double sum_field_old, diff_sum_field, avr_eqVar, aver_eqVarOld, diff_avr_eqVar;
    plan3d plan;
    int nthreads, cap, two_phase_api;
    int hist_n;
    double hist[HIST_MAX * 4];
} m3o_th;

/* constructor, FFT_thermal3D.cpp:14-91 */
m3o_th *m3o_th_create(int Nx, int Ny, int Nz, int nthreads)
{
    m3o_th *s = (m3o_th *)xcalloc(1, sizeof(m3o_th));
    s->Nx = Nx; s->Ny = Ny; s->Nz = Nz; s->Nzh = Nz / 2 + 1;
    s->N = (long)Nx * Ny * Nz;
    s->H = (long)Nx * Ny * s->Nzh;
    s->nthreads = nthreads > 0 ? nthreads : 1;
    s->cap = 500;
    s->rve = (int32_t *)xcalloc(s->N, sizeof(int32_t));
    s->tg = (double *)xcalloc(3 * s->N, sizeof(double));
    s->q = (double *)xcalloc(3 * s->N, sizeof(double));
    s->t2 = (cplx *)xcalloc(3 * s->H, sizeof(cplx));
    s->t3 = (cplx *)xcalloc(3 * s->H, sizeof(cplx));
    s->GF = (double *)xcalloc(6 * s->H, sizeof(double));
    plan3d_init(&s->plan, Nx, Ny, Nz, 3, s->nthreads);
    return s;
}

void m3o_th_destroy(m3o_th *s)
{
    if (!s) return;
    free(s->k); free(s->rve); free(s->tg); free(s->q); free(s->t2); free(s->t3); free(s->GF);
    plan3d_free(&s->plan);
    free(s);
}

void m3o_th_set_cap(m3o_th *s, int cap) { s->cap = cap; }

/* setGF, FFT_thermal3D.cpp:98-175 (isotropic K0 = k0 I) */
static void th_setGF(m3o_th *s)
{
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz, Nzh = s->Nzh;
    const double k0 = s->k0, R011 = 1.0 / k0;
#pragma omp parallel for num_threads(s->nthreads) collapse(2)
    for (int x = 0; x < Nx; x++)
        for (int y = 0; y < Ny; y++)
            for (int z = 0; z < Nzh; z++) {
                long cc = ((long)z + (long)Nzh * y + (long)Ny * Nzh * x) * 6;
                double *g = s->GF + cc;
                if (s->gradBase) {
                    double zet[3] = {fmap_t(z, Nz), fmap_t(y, Ny), fmap_t(x, Nx)};
                    double mm = zet[0] * (k0 * zet[0]) + zet[1] * (k0 * zet[1]) + zet[2] * (k0 * zet[2]);
                    g[0] = zet[0] * zet[0] / mm; g[1] = zet[0] * zet[1] / mm; g[2] = zet[0] * zet[2] / mm;
                    g[3] = zet[1] * zet[1] / mm; g[4] = zet[1] * zet[2] / mm; g[5] = zet[2] * zet[2] / mm;
                } else {
                    /* (N-1)/2 map; x uses -(Ny-x) as in :150-151 */
                    double x1 = (x <= (Nx - 1) / 2) ? (double)x : -(double)(Ny - x);
                    double zet[3] = {fmap_m(z, Nz), fmap_m(y, Ny), x1};
                    double mm = (zet[0] * zet[0] + zet[1] * zet[1] + zet[2] * zet[2]);
                    g[0] = (1 - zet[0] * zet[0] / mm) / R011; g[1] = (-zet[0] * zet[1] / mm) / R011;
                    g[2] = (-zet[0] * zet[2] / mm) / R011;    g[3] = (1 - zet[1] * zet[1] / mm) / R011;
                    g[4] = (-zet[1] * zet[2] / mm) / R011;    g[5] = (1 - zet[2] * zet[2] / mm) / R011;
                }
            }
}

/* material.  two_phase=1: the reference rule, FFT_thermal3D.cpp:373-422
 * (RVE != 0 -> km, RVE == 0 -> ki); k = {ki, km}.  two_phase=0: the N-phase
 * extension (conductivity table indexed by phase id, k0 = sqrt(kmin kmax),
 * gradient base iff the first max-k phase has the max volume fraction). */
int m3o_th_material(m3o_th *s, const double *k, int nph, const int32_t *rve, int two_phase)
{
    s->gradBase = 0; s->fluxBase = 0;
    s->two_phase_api = two_phase;
    s->nph = two_phase ? 2 : nph;
    free(s->k);
    s->k = (double *)xcalloc(s->nph, sizeof(double));
    for (int i = 0; i < s->nph; ++i) s->k[i] = k[i];
    for (long i = 0; i < s->N; ++i) {
        int p = rve[i];
        if (two_phase) p = (p != 0) ? 1 : 0;
        else if (p < 0 || p >= nph) return -2;
        s->rve[i] = p;
    }
    if (two_phase) {
        double ki = k[0], km = k[1];
        /* double vol_1 = RVE.sum()/xyz (:399) is an int/int division: 0 unless
         * every voxel is matrix.  Restated literally. */
        long cnt = 0;
        for (long i = 0; i < s->N; ++i) cnt += rve[i];
        double vol_1 = (double)(cnt / s->N);
        if ((vol_1 >= 0.5 && km > ki) || (vol_1 < 0.5 && km < ki)) s->gradBase = 1;
        else s->fluxBase = 1;
        s->k0 = sqrt(ki * km);
    } else {
        double *vol = (double *)xcalloc(nph, sizeof(double));
        for (long i = 0; i < s->N; ++i) vol[s->rve[i]] += 1.0;
        for (int p = 0; p < nph; ++p) vol[p] /= (double)s->N;
        double t[4], vmax;
        max_min(vol, nph, t);
        vmax = t[0];
        max_min(s->k, nph, t);
        int imax = (int)t[1];
        s->k0 = sqrt(t[0] * t[2]);
        if (vol[imax] == vmax) s->gradBase = 1;
        else s->fluxBase = 1;
        free(vol);
    }
    th_setGF(s);
    return 0;
}

/* setLoad, FFT_thermal3D.cpp:426-441 */
void m3o_th_set_load(m3o_th *s, const double *E1)
{
    for (int i = 0; i < 3; ++i) s->E[i] = E1[i];
    double *dst = s->gradBase ? s->tg : s->q;
    for (long i = 0; i < s->N; ++i) memcpy(dst + 3 * i, E1, sizeof(double) * 3);
}

static inline double th_kp(const m3o_th *s, long i) { return s->k[s->rve[i]]; }

/* multipKx, FFT_thermal3D.cpp:180-203 */
static void th_multipKx(m3o_th *s)
{
#pragma omp parallel for num_threads(s->nthreads)
    for (long i = 0; i < s->N; ++i) {
        double kp = th_kp(s, i);
        if (s->gradBase)
            for (int c = 0; c < 3; ++c) s->q[3 * i + c] = kp * s->tg[3 * i + c];
        else {
            double rp = 1.0 / kp; /* Rm = Km.inverse() */
            for (int c = 0; c < 3; ++c) s->tg[3 * i + c] = rp * s->q[3 * i + c];
        }
    }
}

/* multiplyGF, FFT_thermal3D.cpp:270-297: all k including DC, then DC <- E N */
static void th_multiplyGF(m3o_th *s)
{
    static const int w[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma omp parallel for num_threads(s->nthreads)
    for (long kk = 0; kk < s->H; ++kk) {
        const long u = kk * 3;
        const double *g = s->GF + kk * 6;
        for (int j = 0; j < 3; j++)
            s->t3[u + j] = g[w[j][0]] * s->t2[u + 0] + g[w[j][1]] * s->t2[u + 1] + g[w[j][2]] * s->t2[u + 2];
    }
    for (int j = 0; j < 3; ++j) s->t3[j] = s->E[j] * (double)s->N;
}

/* strainUpdate, FFT_thermal3D.cpp:240-268 */
static void th_strainUpdate(m3o_th *s)
{
#pragma omp parallel for num_threads(s->nthreads)
    for (long i = 0; i < s->N; ++i) {
        double kp = th_kp(s, i);
        if (s->gradBase) {
            double a = 2 * s->k0 / (kp + s->k0); /* 2 (K+K0)^-1 K0 */
            for (int c = 0; c < 3; ++c) s->tg[3 * i + c] += a * (s->q[3 * i + c] - s->tg[3 * i + c]);
        } else {
            double rp = 1.0 / kp, r0 = 1.0 / s->k0;
            double a = 2 * r0 / (rp + r0);       /* 2 (R+R0)^-1 R0 */
            for (int c = 0; c < 3; ++c) s->q[3 * i + c] += a * (s->tg[3 * i + c] - s->q[3 * i + c]);
        }
    }
}

/* convergenceCheck, FFT_thermal3D.cpp:303-365 */
static void th_convergenceCheck(m3o_th *s)
{
    rfft3(&s->plan, s->q, s->t2);
    s->itNo += 1;
    const cplx *t2 = s->t2;
#define A2(q) (creal(t2[q]) * creal(t2[q]) + cimag(t2[q]) * cimag(t2[q]))
    double dd = sqrt(A2(0) + A2(1) + A2(2));
    double cc = 0;
    const int Nx = s->Nx, Ny = s->Ny, Nz = s->Nz, Nzh = s->Nzh;
#pragma omp parallel for reduction(+ : cc) num_threads(s->nthreads) collapse(2)
    for (int x = 0; x < Nx; x++)
        for (int y = 0; y < Ny; y++)
            for (int z = 0; z < Nzh; z++) {
                long i = ((long)z + (long)y * Nzh + (long)x * Ny * Nzh) * 3;
                double zt[3] = {fmap_t(z, Nz), fmap_t(y, Ny), fmap_t(x, Nx)};
                cc += sqrt(A2(i + 0) * zt[0] * zt[0] + A2(i + 1) * zt[1] * zt[1] + A2(i + 2) * zt[2] * zt[2]);
            }
#undef A2
    s->avr_eqVar = cc / dd / (double)s->N;
    double Esum = s->E[0] + s->E[1] + s->E[2];
    if (s->gradBase) {
        double tot = sum_all(s->q, 3 * s->N, s->nthreads);
        s->diff_sum_field = fabs((tot - s->sum_field_old) / ((double)s->N * (0.5 * s->k[0] + 0.5 * s->k[1]) * Esum));
    } else {
        double tot = sum_all(s->tg, 3 * s->N, s->nthreads);
        s->diff_sum_field =
            fabs((tot - s->sum_field_old) / ((double)s->N * (0.5 / s->k[0] + 0.5 / s->k[1]) * Esum));
    }
    s->diff_avr_eqVar = fabs(s->avr_eqVar - s->aver_eqVarOld);
    s->aver_eqVarOld = s->avr_eqVar;
    if (s->hist_n < HIST_MAX) {
        double *h = s->hist + 4 * s->hist_n++;
        h[0] = s->itNo; h[1] = s->avr_eqVar; h[2] = s->diff_sum_field; h[3] = s->diff_avr_eqVar;
    }
}

/* runAnalysis, FFT_thermal3D.cpp:446-485 */
int m3o_th_run(m3o_th *s, double tol_field, double tol_eq)
{
    th_multipKx(s);
    while (1) {
        if (s->gradBase) {
            s->sum_field_old = sum_all(s->q, 3 * s->N, s->nthreads);
            for (long i = 0; i < 3 * s->N; ++i) s->q[i] = s->k0 * s->tg[i] - s->q[i];
            rfft3(&s->plan, s->q, s->t2);
        } else {
            s->sum_field_old = sum_all(s->tg, 3 * s->N, s->nthreads);
            for (long i = 0; i < 3 * s->N; ++i) s->tg[i] = (1.0 / s->k0) * s->q[i] - s->tg[i];
            rfft3(&s->plan, s->tg, s->t2);
        }
        th_multiplyGF(s);
        double *dst = s->gradBase ? s->q : s->tg;
        irfft3(&s->plan, s->t3, dst);
        for (long i = 0; i < 3 * s->N; ++i) dst[i] = dst[i] / (double)s->N;
        th_strainUpdate(s);
        th_multipKx(s);
        th_convergenceCheck(s);
        if (s->itNo > 1 && s->diff_avr_eqVar < tol_eq && s->diff_sum_field < tol_field) break;
        if (s->itNo > s->cap) break;
    }
    return s->itNo;
}

/* homogenization, FFT_thermal3D.cpp:489-622.  out K[9], R[9], G[9], Q[9]
 * column-major; est = {Mori-Tanaka, HS lower, HS upper} (2-phase only). */
void m3o_th_homogenize(m3o_th *s, double tol_field, double tol_eq, double *Kout, double *Rout, double *Gout,
                       double *Qout, double *est, int *iters)
{
    double G[9] = {0}, Q[9] = {0}, Kh[9], Rh[9], inv[9];
    for (int w = 0; w < 3; ++w) {
        s->itNo = 0;
        if (s->gradBase) s->sum_field_old = 0;
        double E[3] = {0, 0, 0};
        E[w] = 1;
        m3o_th_set_load(s, E);
        m3o_th_run(s, tol_field, tol_eq);
        if (iters) iters[w] = s->itNo;
        double e[3] = {0, 0, 0}, q[3] = {0, 0, 0};
        for (long t = 0; t < s->N; ++t)
            for (int c = 0; c < 3; ++c) { e[c] += s->tg[3 * t + c]; q[c] += s->q[3 * t + c]; }
        for (int c = 0; c < 3; ++c) {
            if (s->gradBase) {
                G[w + 3 * c] = e[c] / (double)s->N;  /* tempGradHomo(w,c) */
                Q[c + 3 * w] = q[c] / (double)s->N;  /* heatFluxHomo(c,w) */
            } else {
                Q[w + 3 * c] = q[c] / (double)s->N;  /* heatFluxHomo(w,c) */
                G[c + 3 * w] = e[c] / (double)s->N;  /* tempGradHomo(c,w) */
            }
        }
    }
    if (s->gradBase) {
        mat_inv(G, inv, 3);
        mat_mul(inv, Q, Kh, 3);
        mat_inv(Kh, Rh, 3);
    } else {
        mat_inv(Q, inv, 3);
        mat_mul(inv, G, Rh, 3);
        mat_inv(Rh, Kh, 3);
    }
    memcpy(Kout, Kh, sizeof Kh); memcpy(Rout, Rh, sizeof Rh);
    memcpy(Gout, G, sizeof G); memcpy(Qout, Q, sizeof Q);
    /* Mori-Tanaka and HS, :583-592 (phase 1 = matrix) */
    double km = s->k[1 < s->nph ? 1 : 0], ki = s->k[0];
    long cnt = 0;
    for (long i = 0; i < s->N; ++i) cnt += (s->rve[i] != 0);
    /* est[0..2]: literal reference values; fi_m = RVE.sum()/xyz is an int/int
     * division there (:586).  est[3..5]: the same formulas with the true
     * matrix volume fraction. */
    for (int lit = 0; lit < 2; ++lit) {
        double fi_m = lit == 0 ? (double)(cnt / s->N) : (double)cnt / (double)s->N, fi_i = 1 - fi_m;
        double *e = est + (lit == 0 ? 0 : 3);
        e[0] = km + (fi_i * (ki - km)) / (1 + fi_m * (ki - km) / (3 * km));
        e[1] = fi_i * ki + fi_m * km - fi_i * fi_m * (km - ki) * (km - ki) / (3 * km + fi_m * (ki - km));
        e[2] = fi_i * ki + fi_m * km - fi_i * fi_m * (km - ki) * (km - ki) / (3 * ki - fi_i * (ki - km));
    }
}

void m3o_th_fields(const m3o_th *s, double *tg, double *q)
{
    if (tg) memcpy(tg, s->tg, sizeof(double) * 3 * s->N);
    if (q) memcpy(q, s->q, sizeof(double) * 3 * s->N);
}

/* info: {gradBase, k0, itNo, eqVar, diff_sum, diff_eq} */
void m3o_th_info(const m3o_th *s, double *out)
{
    out[0] = s->gradBase; out[1] = s->k0; out[2] = s->itNo; out[3] = s->avr_eqVar;
    out[4] = s->diff_sum_field; out[5] = s->diff_avr_eqVar;
}

int m3o_th_history(const m3o_th *s, double *out, int maxn)
{
    int n = s->hist_n < maxn ? s->hist_n : maxn;
    memcpy(out, s->hist, sizeof(double) * 4 * n);
    return n;
}

void m3o_th_clear_history(m3o_th *s) { s->hist_n = 0; }

/* timing helper for the CPU baseline: run k accelerated iterations of load
 * case w without the stop rule (cap = k), returns the iterations done. */
int m3o_el_run_fixed(m3o_el *s, int w, int k)
{
    double E[6] = {0, 0, 0, 0, 0, 0};
    E[w] = s->strainBase ? 0.0005 : 1.e6;
    s->behav = BEH_E;
    s->itNo = 0;
    m3o_el_set_load(s, E, E);
    int old = s->cap;
    s->cap = k - 1;
    m3o_el_run(s, -1.0, -1.0, 1);
    s->cap = old;
    return s->itNo;
}

/* ------------------------------------------------------------------------ */
/* Phase-retrieval reconstruction: recPlan(double *TPCF, Nx, Ny, Nz) and     */
/* runRec(n, beta), src/rec/reconstruction.cpp:211-243 and :285-406.         */
/* Random phases: the reference draws them with Eigen's setRandom (std::rand */
/* through an unpinned Eigen version); here, as in the HIP path, a seeded    */
/* counter-based stream: value i of stream k is splitmix64(seed + (k << 40)  */
/* + i) scaled to [-1, 1).  Stream 0 is the start field, stream k the k-th   */
/* re-randomisation (every 150 iterations).                                  */
/* ------------------------------------------------------------------------ */
static inline uint64_t rec_mix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double m3o_rec_uniform(uint64_t seed, uint64_t stream, uint64_t i)
{
    return 2.0 * ((double)(rec_mix(seed + (stream << 40) + i) >> 11) * 0x1.0p-53) - 1.0;
}

uint64_t m3o_rec_mix(uint64_t z) { return rec_mix(z); }

typedef struct {
    int Nx, Ny, Nz;
    long xyz;
    int nthreads;
    plan1d px, py, pz;
    double *normC;
    cplx *mT, *MT, *mH, *MH;
    double *ang;
    double *err;
    int nerr, cap_err;
    const double *init; /* optional start field (tests), else stream 0 */
} m3o_rec;

static void rec_fft3(m3o_rec *r, cplx *d, int sign)
{
    const long Nx = r->Nx, Ny = r->Ny, Nz = r->Nz;
    fft_axis(&r->pz, d, Nx * Ny, Nz, 1, 1, sign, r->nthreads);
    fft_axis(&r->py, d, Nx, Ny * Nz, Nz, Nz, sign, r->nthreads);
    fft_axis(&r->px, d, 1, 0, Ny * Nz, Ny * Nz, sign, r->nthreads);
}

void m3o_c2c3(int Nx, int Ny, int Nz, double *ri, int sign, int nthreads)
{
    m3o_rec r;
    memset(&r, 0, sizeof r);
    r.Nx = Nx; r.Ny = Ny; r.Nz = Nz;
    r.nthreads = nthreads > 0 ? nthreads : 1;
    plan1d_init(&r.px, Nx); plan1d_init(&r.py, Ny); plan1d_init(&r.pz, Nz);
    rec_fft3(&r, (cplx *)ri, sign);
    plan1d_free(&r.px); plan1d_free(&r.py); plan1d_free(&r.pz);
}

/* recPlan(double *TPCF, ...), :211-243: normC = sqrt(|TPCF| xyz) */
m3o_rec *m3o_rec_create(const double *tpcf, int Nx, int Ny, int Nz, int nthreads)
{
    m3o_rec *r = (m3o_rec *)xcalloc(1, sizeof(m3o_rec));
    r->Nx = Nx; r->Ny = Ny; r->Nz = Nz;
    r->xyz = (long)Nx * Ny * Nz;
    r->nthreads = nthreads > 0 ? nthreads : 1;
    plan1d_init(&r->px, Nx); plan1d_init(&r->py, Ny); plan1d_init(&r->pz, Nz);
    r->normC = (double *)xcalloc((size_t)r->xyz, sizeof(double));
    r->mT = (cplx *)xcalloc((size_t)r->xyz, sizeof(cplx));
    r->MT = (cplx *)xcalloc((size_t)r->xyz, sizeof(cplx));
    r->mH = (cplx *)xcalloc((size_t)r->xyz, sizeof(cplx));
    r->MH = (cplx *)xcalloc((size_t)r->xyz, sizeof(cplx));
    r->ang = (double *)xcalloc((size_t)r->xyz, sizeof(double));
    for (long i = 0; i < r->xyz; ++i) r->normC[i] = sqrt(fabs(tpcf[i]) * (double)r->xyz);
    return r;
}

void m3o_rec_destroy(m3o_rec *r)
{
    if (!r) return;
    plan1d_free(&r->px); plan1d_free(&r->py); plan1d_free(&r->pz);
    free(r->normC); free(r->mT); free(r->MT); free(r->mH); free(r->MH); free(r->ang); free(r->err);
    free(r);
}

void m3o_rec_set_initial(m3o_rec *r, const double *m0) { r->init = m0; }

/* runRec(n, beta), :285-406.  Returns the iterations done (the length of
 * trackrealerrorA that was written); *status = 1 converged (error < 0.1),
 * 2 iteration number exceeded; *ctr_out = iterations of the final volume
 * fraction search. */
int m3o_rec_run(m3o_rec *r, int n, double beta, uint64_t seed, int *status, int *ctr_out)
{
    const long xyz = r->xyz;
    free(r->err);
    r->err = (double *)xcalloc((size_t)(n > 0 ? n : 1), sizeof(double));
    r->nerr = 0;
    /* ang = (setRandom + 1) * 0.5; mTilda = ang + 0i (:290-293) */
    for (long s = 0; s < xyz; ++s) {
        const double a = r->init ? r->init[s] : (m3o_rec_uniform(seed, 0, (uint64_t)s) + 1.0) * 0.5;
        r->mT[s] = a;
    }
    int j = 0, i = 0, st = 0;
    while (1) {
        j += 1;
        memcpy(r->MT, r->mT, sizeof(cplx) * (size_t)xyz);
        rec_fft3(r, r->MT, -1); /* p1 */
        if (i >= n) { st = 2; break; }
        /* trackrealerrorA(i) = funcNormC(MTilda, normC) (reconstruction.hpp funcNormC) */
        double e = 0;
        for (long k = 0; k < xyz; ++k) e += fabs(sqrt(creal(r->MT[k]) * creal(r->MT[k]) + cimag(r->MT[k]) * cimag(r->MT[k])) - r->normC[k]);
        e = e / (double)xyz;
        r->err[i] = e;
        r->nerr = i + 1;
        if (e < 0.1) { st = 1; break; }
        for (long s = 0; s < xyz; ++s) r->ang[s] = carg(r->MT[s]);
        if (remainder((double)(i + 1), 150.0) == 0) {
            /* ang.setRandom; mHat = ang; p3; ang2 = arg(MHat) .3; ang1 = arg(MTilda) .7 (:321-333) */
            const uint64_t k = (uint64_t)((i + 1) / 150);
            for (long s = 0; s < xyz; ++s) r->mH[s] = m3o_rec_uniform(seed, k, (uint64_t)s);
            memcpy(r->MH, r->mH, sizeof(cplx) * (size_t)xyz);
            rec_fft3(r, r->MH, -1);
            for (long s = 0; s < xyz; ++s) r->ang[s] = carg(r->MH[s]) * .3 + carg(r->MT[s]) * .7;
        }
        /* MHat = normC e^{i ang}; p4; / xyz (:335-346) */
        for (long s = 0; s < xyz; ++s) r->MH[s] = r->normC[s] * cos(r->ang[s]) + I * (r->normC[s] * sin(r->ang[s]));
        memcpy(r->mH, r->MH, sizeof(cplx) * (size_t)xyz);
        rec_fft3(r, r->mH, +1);
        for (long s = 0; s < xyz; ++s) r->mH[s] = creal(r->mH[s]) / (double)xyz + I * (cimag(r->mH[s]) / (double)xyz);
        if (j < 100) {
            /* :348-356 */
            for (long s = 0; s < xyz; ++s) {
                double re = creal(r->mH[s]);
                if (re < 0) { re = creal(r->mT[s]) - beta * re; r->mH[s] = re; }
                if (re > 1) { re = creal(r->mT[s]) - beta * (re - 1); r->mH[s] = re; }
                if (re > 1) { re = 1; r->mH[s] = re; }
            }
            memcpy(r->mT, r->mH, sizeof(cplx) * (size_t)xyz);
        }
        if (j >= 100) {
            /* :358-366 */
            for (long s = 0; s < xyz; ++s) {
                const double re = creal(r->mH[s]);
                if (re < 0) r->mH[s] = 0;
                if (re > 1) r->mH[s] = 1;
            }
            if (j > 200) j = 0;
            memcpy(r->mT, r->mH, sizeof(cplx) * (size_t)xyz);
        }
        i += 1;
    }
    /* volume fraction search (:374-402) */
    int ctr = 0;
    double x = 0, vol_dif = 0, thr = 0.5;
    while (1) {
        thr = .5 - x;
        double sum = 0;
        for (long s = 0; s < xyz; ++s) {
            r->ang[s] = (creal(r->mT[s]) < thr) ? 0.0 : 1.0;
            sum += r->ang[s];
        }
        vol_dif = (sum / (1. * xyz) - r->normC[0] / (1. * xyz));
        if (vol_dif > 0) x -= .001;
        else x += .001;
        if (fabs(vol_dif) < .0001 || ctr > 200) break;
        ctr += 1;
    }
    for (long s = 0; s < xyz; ++s) r->mT[s] = r->ang[s] + I * cimag(r->mT[s]);
    if (status) *status = st;
    if (ctr_out) *ctr_out = ctr;
    return r->nerr;
}

/* mTilda as interleaved (re, im) and trackrealerrorA */
int m3o_rec_result(const m3o_rec *r, double *mtilda_ri, double *err, int maxn)
{
    if (mtilda_ri) memcpy(mtilda_ri, r->mT, sizeof(cplx) * (size_t)r->xyz);
    const int n = r->nerr < maxn ? r->nerr : maxn;
    if (err && n > 0) memcpy(err, r->err, sizeof(double) * (size_t)n);
    return r->nerr;
}

/* ------------------------------------------------------------------------ */
/* FFT_Analysis2D (src/2d/FFT_analysis2D.cpp): basic scheme on complex 2D    */
/* fields [E11, E22, E12] (tensor shear), 25 fixed iterations.               */
/* ------------------------------------------------------------------------ */
typedef struct {
    int Nx, Ny;
    long xy;
    int nthreads;
    plan1d px, py;
    double Cm[9], Ci[9], C0[9], la0, mu0;
    double *rve;
    cplx *eps, *sig, *sigOld, *tau, *t;
    double *GF;
    cplx E[3];
    int itNo;
} m3o_2d;

m3o_2d *m3o_2d_create(int Nx, int Ny, int nthreads)
{
    m3o_2d *s = (m3o_2d *)xcalloc(1, sizeof(m3o_2d));
    s->Nx = Nx; s->Ny = Ny; s->xy = (long)Nx * Ny;
    s->nthreads = nthreads > 0 ? nthreads : 1;
    plan1d_init(&s->px, Nx); plan1d_init(&s->py, Ny);
    s->rve = (double *)xcalloc((size_t)s->xy, sizeof(double));
    s->eps = (cplx *)xcalloc((size_t)(3 * s->xy), sizeof(cplx));
    s->sig = (cplx *)xcalloc((size_t)(3 * s->xy), sizeof(cplx));
    s->sigOld = (cplx *)xcalloc((size_t)(3 * s->xy), sizeof(cplx));
    s->tau = (cplx *)xcalloc((size_t)(3 * s->xy), sizeof(cplx));
    s->t = (cplx *)xcalloc((size_t)(3 * s->xy), sizeof(cplx));
    s->GF = (double *)xcalloc((size_t)(6 * s->xy), sizeof(double));
    return s;
}

void m3o_2d_destroy(m3o_2d *s)
{
    if (!s) return;
    plan1d_free(&s->px); plan1d_free(&s->py);
    free(s->rve); free(s->eps); free(s->sig); free(s->sigOld); free(s->tau); free(s->t); free(s->GF);
    free(s);
}

/* 2D complex FFT of 3 interleaved components (p3 / p4, :29-50) */
static void fft2_3(m3o_2d *s, cplx *d, int sign)
{
    fft_axis(&s->py, d, s->Nx, (long)s->Ny * 3, 3, 3, sign, s->nthreads);
    fft_axis(&s->px, d, 1, 0, (long)s->Ny * 3, (long)s->Ny * 3, sign, s->nthreads);
}

/* material(properties, RVE1), :189-221 */
void m3o_2d_material(m3o_2d *s, const double *q, const double *rve)
{
    const double Em = q[0], vm = q[1], Ei = q[2], vi = q[3];
    const double mum = Em / (1 + vm) / 2, lam = vm * Em / (1 + vm) / (1 - 2 * vm);
    const double mui = Ei / (1 + vi) / 2, lai = vi * Ei / (1 + vi) / (1 - 2 * vi);
    const double cm[9] = {lam + 2 * mum, lam, 0, lam, lam + 2 * mum, 0, 0, 0, 2 * mum};
    const double ci[9] = {lai + 2 * mui, lai, 0, lai, lai + 2 * mui, 0, 0, 0, 2 * mui};
    const double alf = 1;
    for (int k = 0; k < 9; ++k) {
        s->Cm[k] = cm[k];
        s->Ci[k] = ci[k];
        s->C0[k] = alf * (cm[k] + ci[k]) / 2.;
    }
    s->la0 = alf * (lai + lam) / 2.;
    s->mu0 = alf * (mui + mum) / 2.;
    memcpy(s->rve, rve, sizeof(double) * (size_t)s->xy);
}

/* setLoad(E1), :223-230 */
void m3o_2d_set_load(m3o_2d *s, const double *E_ri)
{
    for (int c = 0; c < 3; ++c) s->E[c] = E_ri[2 * c] + I * E_ri[2 * c + 1];
    for (long i = 0; i < s->xy; ++i)
        for (int c = 0; c < 3; ++c) s->eps[3 * i + c] = s->E[c];
}

/* setGF, :53-93 (x > (Nx-1)/2 maps to -(Ny - x), as the reference) */
static void m3o_2d_setGF(m3o_2d *s)
{
    const double d1 = 1 / (4 * s->mu0);
    const double d2 = (s->la0 + s->mu0) / (s->mu0 * (s->la0 + 2 * s->mu0));
    static const int q[6][4] = {{0, 0, 0, 0}, {0, 0, 1, 1}, {0, 0, 0, 1}, {1, 1, 1, 1}, {1, 1, 0, 1}, {0, 1, 0, 1}};
    for (int x = 0; x < s->Nx; x++)
        for (int y = 0; y < s->Ny; y++) {
            const long cc = ((long)y + (long)x * s->Ny) * 6;
            const int x1 = (x <= (s->Nx - 1) / 2) ? x : -(s->Ny - x);
            const int y1 = (y <= (s->Ny - 1) / 2) ? y : -(s->Ny - y);
            const double zet[2] = {x1 * 1.0, y1 * 1.0};
            const double n2 = zet[0] * zet[0] + zet[1] * zet[1];
            for (int r = 0; r < 6; r++) {
                const int i = q[r][0], j = q[r][1], k = q[r][2], h = q[r][3];
                s->GF[cc + r] = (d1 / n2) * (dlt(k, i) * zet[h] * zet[j] + dlt(h, i) * zet[k] * zet[j] +
                                             dlt(k, j) * zet[h] * zet[i] + dlt(h, j) * zet[k] * zet[i]) -
                                (d2 * zet[i] * zet[j] * zet[k] * zet[h]) / (n2 * n2);
            }
        }
}

/* runAnalysis, :232-246: 25 x {multipCx, tau, FFT_tau, multiplyGF,
 * FFTBCKWRD_epsilon, convergenceCheck}.  convergenceCheck only prints (and
 * its cc reads past the 2-element frequency vector); its itNo is kept. */
void m3o_2d_run(m3o_2d *s)
{
    m3o_2d_setGF(s);
    const long xy = s->xy;
    static const int w[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    for (int it = 0; it < 25; ++it) {
        /* multipCx :96-106 */
        memcpy(s->sigOld, s->sig, sizeof(cplx) * (size_t)(3 * xy));
        for (long i = 0; i < xy; ++i) {
            const double *C = (s->rve[i] == 1) ? s->Cm : s->Ci;
            const cplx *e = s->eps + 3 * i;
            for (int r = 0; r < 3; ++r) s->sig[3 * i + r] = C[3 * r] * e[0] + C[3 * r + 1] * e[1] + C[3 * r + 2] * e[2];
        }
        /* tau = sigma - C0 epsilon (:240) */
        for (long i = 0; i < xy; ++i) {
            const cplx *e = s->eps + 3 * i;
            for (int r = 0; r < 3; ++r)
                s->tau[3 * i + r] = s->sig[3 * i + r] - (s->C0[3 * r] * e[0] + s->C0[3 * r + 1] * e[1] + s->C0[3 * r + 2] * e[2]);
        }
        /* FFT_tau :108-112 */
        memcpy(s->t, s->tau, sizeof(cplx) * (size_t)(3 * xy));
        fft2_3(s, s->t, -1);
        /* multiplyGF :114-141 (in place: every output reads its own frequency) */
        for (long i = 1; i < xy; i++) {
            const long u = i * 3, d = i * 6;
            const cplx a0 = s->t[u], a1 = s->t[u + 1], a2 = s->t[u + 2];
            for (int j = 0; j < 3; j++) {
                const double re = -(s->GF[d + w[j][0]] * creal(a0) + s->GF[d + w[j][1]] * creal(a1) +
                                    s->GF[d + w[j][2]] * creal(a2) * 2);
                const double im = -(s->GF[d + w[j][0]] * cimag(a0) + s->GF[d + w[j][1]] * cimag(a1) +
                                    s->GF[d + w[j][2]] * cimag(a2) * 2);
                s->t[u + j] = re + I * im;
            }
        }
        for (int c = 0; c < 3; ++c) s->t[c] = creal(s->E[c]) * xy + I * (cimag(s->E[c]) * xy);
        /* FFTBCKWRD_epsilon :143-148 */
        fft2_3(s, s->t, +1);
        for (long k = 0; k < 3 * xy; ++k) s->eps[k] = creal(s->t[k]) / xy + I * (cimag(s->t[k]) / xy);
        s->itNo += 1;
    }
}

/* homogenization, :248-272: three load cases E(w) = 0.005 from the current
 * fields (the reference does not reset epsilon between them), means of the
 * real parts, CHomo.col(k) = epHomo^-1 sigHomo.col(k).  Matrices column-major. */
void m3o_2d_homogenization(m3o_2d *s, double *CHomo, double *epHomo, double *sigHomo)
{
    double ep[9], sg[9];
    for (int w = 0; w < 3; ++w) {
        for (int c = 0; c < 3; ++c) s->E[c] = (c == w) ? 0.005 : 0.0;
        m3o_2d_run(s);
        double e[3] = {0, 0, 0}, g[3] = {0, 0, 0};
        for (long t = 0; t < s->xy; ++t)
            for (int c = 0; c < 3; ++c) {
                e[c] += creal(s->eps[3 * t + c]);
                g[c] += creal(s->sig[3 * t + c]);
            }
        for (int c = 0; c < 3; ++c) {
            ep[w + 3 * c] = e[c] / s->xy; /* epHomo(w, c) */
            sg[c + 3 * w] = g[c] / s->xy; /* sigHomo(c, w) */
        }
    }
    double inv[9];
    mat_inv(ep, inv, 3);
    double C[9];
    mat_mul(inv, sg, C, 3);
    if (CHomo) memcpy(CHomo, C, sizeof C);
    if (epHomo) memcpy(epHomo, ep, sizeof ep);
    if (sigHomo) memcpy(sigHomo, sg, sizeof sg);
}

void m3o_2d_fields(const m3o_2d *s, double *eps_ri, double *sig_ri)
{
    if (eps_ri) memcpy(eps_ri, s->eps, sizeof(cplx) * (size_t)(3 * s->xy));
    if (sig_ri) memcpy(sig_ri, s->sig, sizeof(cplx) * (size_t)(3 * s->xy));
}

/* ------------------------------------------------------------------------ */
/* percolation (src/percolation/percolation.cpp): cluster labelling of one   */
/* phase by the reference's sweeps, cluster volume fractions, and the count  */
/* of face contacts between two phases.  Grids are cubic (Nx = Ny = Nz), as  */
/* the reference's permuted-index sweeps require.                             */
/* ------------------------------------------------------------------------ */
typedef struct {
    int NX, NY, NZ, Nx, Ny, Nz;
    int noOfPhases, intrestPhase, phasePlus;
    int *C, *C1;
    double vol;
    int nclusters;
    int *all_clusters;       /* first-appearance order */
    double *vol_frac_sorted; /* descending */
} m3o_perc;

/* percolation(RVE, Nx1, Ny1, Nz1, noOfPhases, intrestPhase), :7-30 */
m3o_perc *m3o_perc_create(const int32_t *rve, int NX, int NY, int NZ, int nph, int phase)
{
    m3o_perc *p = (m3o_perc *)xcalloc(1, sizeof(m3o_perc));
    p->NX = NX; p->NY = NY; p->NZ = NZ;
    p->Nx = NX + 2; p->Ny = NY + 2; p->Nz = NZ + 2;
    p->noOfPhases = nph; p->intrestPhase = phase; p->phasePlus = nph + 1;
    const long n = (long)NX * NY * NZ;
    p->C1 = (int *)xcalloc((size_t)n, sizeof(int));
    p->C = (int *)xcalloc((size_t)p->Nx * p->Ny * p->Nz, sizeof(int));
    for (long i = 0; i < n; ++i) p->C1[i] = rve[i];
    return p;
}

void m3o_perc_destroy(m3o_perc *p)
{
    if (!p) return;
    free(p->C); free(p->C1); free(p->all_clusters); free(p->vol_frac_sorted);
    free(p);
}

#define PC(i, j, k) p->C[(long)(k) + (long)(j) * p->Nz + (long)(i) * p->Ny * p->Nz]

/* setMargin(cd), :36-59 */
static void perc_setMargin(m3o_perc *p, int cd)
{
    int lab = p->phasePlus;
    for (int i = 0; i < p->Nx; ++i)
        for (int j = 0; j < p->Ny; ++j)
            for (int k = 0; k < p->Nz; ++k) {
                if (i == p->Nx - 1 || j == p->Ny - 1 || k == p->Nz - 1 || i == 0 || j == 0 || k == 0)
                    PC(i, j, k) = -1;
                else {
                    const int v = p->C1[(long)(k - 1) + (long)(j - 1) * p->NZ + (long)(i - 1) * p->NY * p->NZ];
                    if (v == p->intrestPhase && cd == 0) {
                        PC(i, j, k) = lab;
                        lab += 1;
                        p->vol += 1;
                    } else
                        PC(i, j, k) = v;
                }
            }
}

/* neighborCheck(i, j, k), :75-114: the interest voxels of the cube
 * (i-1..i, j-1..j, k-1..k) all take their minimum label */
static void perc_neighborCheck(m3o_perc *p, int i, int j, int k)
{
    int ii[8], jj[8], kk[8], lab[8], cnt = 0;
    for (int t = 0; t < 2; ++t)
        for (int r = 0; r < 2; ++r)
            for (int w = 0; w < 2; ++w) {
                const int a = i + (t ? 0 : -1), b = j + (r ? 0 : -1), c = k + (w ? 0 : -1);
                if (PC(a, b, c) >= p->phasePlus) {
                    ii[cnt] = a; jj[cnt] = b; kk[cnt] = c; lab[cnt] = PC(a, b, c);
                    cnt += 1;
                }
            }
    double mn = 1e99;
    for (int q = 0; q < cnt; ++q) mn = fmin(mn, lab[q] * 1.);
    for (int q = 0; q < cnt; ++q) PC(ii[q], jj[q], kk[q]) = (int)mn;
}

static int perc_cond(m3o_perc *p, int t, int i, int j, int k)
{
    switch (t) {
    case 1: return PC(i, j, k) >= p->phasePlus;
    case 2: return PC(i, k, j) >= p->phasePlus;
    case 3: return PC(j, i, k) >= p->phasePlus;
    case 4: return PC(j, k, i) >= p->phasePlus;
    case 5: return PC(k, j, i) >= p->phasePlus;
    default: return PC(k, i, j) >= p->phasePlus;
    }
}

static int perc_cmp_desc(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return (x < y) - (x > y);
}

/* runPercolationAnalysis(total_iter), :117-380 (total_iter is unused there:
 * five rounds of the eight sweep orders, each over the six index
 * permutations of the trigger test); then the clusters in first-appearance
 * order, their volume fractions, and every interest voxel relabelled
 * phasePlus + rank of its cluster's fraction (equal fractions, equal rank) */
void m3o_perc_run_rounds(m3o_perc *p, int rounds);

void m3o_perc_run(m3o_perc *p) { m3o_perc_run_rounds(p, 5); }

/* the same with `rounds` rounds of the 8 x 6 sweeps instead of the
 * reference's fixed 5 (test use: how many rounds reach the fixed point) */
void m3o_perc_run_rounds(m3o_perc *p, int rounds)
{
    perc_setMargin(p, 0);
    const int lo = 1, hiX = p->Nx - 2, hiY = p->Ny - 2, hiZ = p->Nz - 2;
    for (int X = 0; X < rounds; ++X)
        for (int s = 0; s < 8; ++s) {
            const int di = (s & 4) ? -1 : 1, dj = (s & 2) ? -1 : 1, dk = (s & 1) ? -1 : 1;
            for (int t = 1; t < 7; ++t)
                for (int i = (di > 0 ? lo : hiX); di > 0 ? i <= hiX : i >= lo; i += di)
                    for (int j = (dj > 0 ? lo : hiY); dj > 0 ? j <= hiY : j >= lo; j += dj)
                        for (int k = (dk > 0 ? lo : hiZ); dk > 0 ? k <= hiZ : k >= lo; k += dk)
                            if (perc_cond(p, t, i, j, k)) perc_neighborCheck(p, i, j, k);
        }
    /* removeMargin, :62-70 */
    for (int i = 0; i < p->NX; ++i)
        for (int j = 0; j < p->NY; ++j)
            for (int k = 0; k < p->NZ; ++k)
                p->C1[(long)k + (long)j * p->NZ + (long)i * p->NY * p->NZ] = PC(i + 1, j + 1, k + 1);
    const long n = (long)p->NX * p->NY * p->NZ;
    /* clusters in first-appearance order (:318-336) and their fractions */
    int cap = 16, nc = 0;
    int *cl = (int *)malloc(sizeof(int) * (size_t)cap);
    for (long i = 0; i < n; ++i)
        if (p->C1[i] >= p->phasePlus) {
            int f = 0;
            for (int q = 0; q < nc; ++q)
                if (cl[q] == p->C1[i]) { f = 1; break; }
            if (!f) {
                if (nc == cap) { cap *= 2; cl = (int *)realloc(cl, sizeof(int) * (size_t)cap); }
                cl[nc++] = p->C1[i];
            }
        }
    double *fr = (double *)xcalloc((size_t)(nc > 0 ? nc : 1), sizeof(double));
    double *srt = (double *)xcalloc((size_t)(nc > 0 ? nc : 1), sizeof(double));
    for (int q = 0; q < nc; ++q) {
        double s = 0;
        for (long i = 0; i < n; ++i)
            if (p->C1[i] == cl[q]) s += 1;
        fr[q] = s / p->vol;
        srt[q] = fr[q];
    }
    qsort(srt, (size_t)nc, sizeof(double), perc_cmp_desc);
    /* relabel (:355-376): std::find of the cluster's fraction in the sorted list */
    for (long i = 0; i < n; ++i)
        if (p->C1[i] >= p->phasePlus) {
            int i1 = 0;
            while (cl[i1] != p->C1[i]) ++i1;
            int i2 = 0;
            while (srt[i2] != fr[i1]) ++i2;
            p->C1[i] = p->phasePlus + i2;
        }
    free(fr);
    p->nclusters = nc;
    p->all_clusters = cl;
    p->vol_frac_sorted = srt;
}

/* out: relabelled phase field; returns the cluster count */
int m3o_perc_result(const m3o_perc *p, int32_t *c1, int *all_clusters, double *vol_frac, double *vol_intrstd, int maxn)
{
    const long n = (long)p->NX * p->NY * p->NZ;
    if (c1) for (long i = 0; i < n; ++i) c1[i] = p->C1[i];
    for (int q = 0; q < p->nclusters && q < maxn; ++q) {
        if (all_clusters) all_clusters[q] = p->all_clusters[q];
        if (vol_frac) vol_frac[q] = p->vol_frac_sorted[q];
    }
    if (vol_intrstd) *vol_intrstd = p->vol;
    return p->nclusters;
}

/* SurfaceOfTwoPhase(RVE, phase1, phase2), :384-414: voxels of phase1 times
 * their -x, -y, -z face neighbours of phase2 (the margin is -1) */
int m3o_perc_surface(m3o_perc *p, const int32_t *rve, int phase1, int phase2)
{
    const long n = (long)p->NX * p->NY * p->NZ;
    for (long i = 0; i < n; ++i) p->C1[i] = rve[i];
    perc_setMargin(p, 1);
    int count = 0;
    for (int i = 0; i < p->Nx; ++i)
        for (int j = 0; j < p->Ny; ++j)
            for (int k = 0; k < p->Nz; ++k)
                if (i != p->Nx - 1 && j != p->Ny - 1 && k != p->Nz - 1 && i * j * k != 0 && PC(i, j, k) == phase1)
                    for (int t = 0; t < 2; ++t)
                        for (int r = 0; r < 2; ++r)
                            for (int w = 0; w < 2; ++w) {
                                const int a = i + (t ? 0 : -1), b = j + (r ? 0 : -1), c = k + (w ? 0 : -1);
                                const int diff = abs(a - i) + abs(b - j) + abs(c - k);
                                if (diff == 1 && PC(a, b, c) == phase2) count += 1;
                            }
    return count;
}
#undef PC

/* 1 when the last analysis stopped at a fixed point of its sweeps: every
 * cube that a trigger test would visit holds one label on its interest
 * voxels (then a further sweep would change nothing) */
int m3o_perc_fixed_point(const m3o_perc *p)
{
    m3o_perc *q = (m3o_perc *)p;
    for (int i = 1; i < q->Nx - 1; ++i)
        for (int j = 1; j < q->Ny - 1; ++j)
            for (int k = 1; k < q->Nz - 1; ++k) {
                int act = 0;
                for (int t = 1; t < 7 && !act; ++t) act = perc_cond(q, t, i, j, k);
                if (!act) continue;
                int lab = -2;
                for (int a = i - 1; a <= i; ++a)
                    for (int b = j - 1; b <= j; ++b)
                        for (int c = k - 1; c <= k; ++c) {
                            const int v = q->C[(long)c + (long)b * q->Nz + (long)a * q->Ny * q->Nz];
                            if (v < q->phasePlus) continue;
                            if (lab == -2) lab = v;
                            else if (v != lab) return 0;
                        }
            }
    return 1;
}
