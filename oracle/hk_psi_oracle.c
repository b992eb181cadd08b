/*
 * ORACLE -- test infrastructure only.
 *
 * CPU restatement, in plain C, of the cpw2000 H|psi> application
 * hk_psi_c16 and the routines it calls.  It is the checker for the HIP
 * implementation and the CPU baseline ("port") of the benchmark; nothing in
 * the product library links or calls it.
 *
 * Followed reference files (paths relative to the cpw2000 tree):
 *   Src/hk_psi/reference/hk_psi_c16.f90   driver, index map, local + kinetic
 *   Src/cpw_lib/hk_psi_nl_c16.f90         Kleinman-Bylander non-local term
 *   Src/fft/_cfft/cfft_wf_c16.f           pruned ("Goedecker") 3-D FFT
 *   Src/fft/_cfft/cfft_c16.f              full 3-D FFT
 *   Src/fft/_ncar/cfftmlt.f:46-50         sign convention of the 1-D transforms
 *   Src/ALFA/size_fft.f90                 mesh-size rule and work sizes
 *
 * The 1-D transforms are a mixed-radix Stockham FFT written for this file (the
 * reference uses NCAR FFTPACK scfftf/scfftb); they compute the same DFT:
 *   sign > 0 : X_k = sum_j x_j exp(-2 pi i jk/n)   (forward, SCFFTF)
 *   sign < 0 : X_k = sum_j x_j exp(+2 pi i jk/n)   (backward, SCFFTB)
 * Both are unnormalised; the 1/N of the r->G transform is applied where the
 * reference applies it.
 *
 * Complex numbers are interleaved (re, im) doubles, Fortran complex(8) layout.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "hk_psi_oracle.h"

#define ORC_MAXN 8192

/* ------------------------------------------------------------------ */
/* size_fft (Src/ALFA/size_fft.f90:53-76)                              */
/* ------------------------------------------------------------------ */
int orc_size_fft(const int *kmscr3, int *nsfft, int *mfft, int *mwrk)
{
    static const int inprim[11] = {2, 4, 6, 8, 10, 12, 16, 18, 20, 24, 32};
    static const int insec[6] = {36, 40, 48, 50, 54, 64};
    for (int n = 0; n < 3; n++) {
        int m = 2 * kmscr3[n] + 1;
        if (m < 32) {
            for (int j = 0; j < 11; j++) {
                nsfft[n] = inprim[j];
                if (m <= nsfft[n]) break;
            }
        } else {
            /* int(log(real(m))/log(real(2))) in default (single) precision */
            int k = (int)(logf((float)m) / logf(2.0f));
            for (int j = 0; j < 6; j++) {
                nsfft[n] = (1 << (k - 5)) * insec[j];
                if (m <= nsfft[n]) break;
            }
        }
    }
    *mfft = (nsfft[0] + 1) * nsfft[1] * nsfft[2];
    int a = nsfft[0] * nsfft[1], b = nsfft[0] * nsfft[2], c = nsfft[1] * nsfft[2];
    int mx = a > b ? a : b;
    mx = mx > c ? mx : c;
    *mwrk = 4 * mx;
    return 0;
}

/* ------------------------------------------------------------------ */
/* 1-D mixed-radix Stockham FFT                                        */
/* ------------------------------------------------------------------ */
typedef struct {
    int n;
    int nf;
    int fac[32];
    double *w; /* w[2k], w[2k+1] = cos, sin(2 pi k / n) */
} orc_plan;

static int factorize(int n, int *fac)
{
    int nf = 0;
    while (n % 4 == 0) { fac[nf++] = 4; n /= 4; }
    while (n % 2 == 0) { fac[nf++] = 2; n /= 2; }
    for (int p = 3; n > 1; p += 2) {
        while (n % p == 0) { fac[nf++] = p; n /= p; }
    }
    return nf;
}

static void plan_init(orc_plan *p, int n)
{
    p->n = n;
    p->nf = factorize(n, p->fac);
    p->w = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (int k = 0; k < n; k++) {
        double a = 2.0 * M_PI * (double)k / (double)n;
        p->w[2 * k] = cos(a);
        p->w[2 * k + 1] = sin(a);
    }
}

static void plan_free(orc_plan *p) { free(p->w); p->w = NULL; }

/* One length-n transform on contiguous interleaved data x; y is scratch of the
 * same size.  Result is left in x.  dir = -1 for exp(-i..), +1 for exp(+i..). */
static void fft_contig(const orc_plan *p, double *x, double *y, int dir)
{
    const int n = p->n;
    double *src = x, *dst = y;
    int L = 1;
    double vr[64], vi[64], tr[64], ti[64];
    for (int s = 0; s < p->nf; s++) {
        const int R = p->fac[s];
        const int m = n / R;
        const int tstep = n / (L * R);
        for (int j = 0; j < m; j++) {
            const int k = j % L;
            for (int r = 0; r < R; r++) {
                double a = src[2 * (j + r * m)], b = src[2 * (j + r * m) + 1];
                int e = (k * r * tstep) % n;
                double c = p->w[2 * e], sn = dir * p->w[2 * e + 1];
                vr[r] = a * c - b * sn;
                vi[r] = a * sn + b * c;
            }
            /* R-point DFT */
            if (R == 2) {
                tr[0] = vr[0] + vr[1]; ti[0] = vi[0] + vi[1];
                tr[1] = vr[0] - vr[1]; ti[1] = vi[0] - vi[1];
            } else if (R == 4) {
                double ar = vr[0] + vr[2], ai = vi[0] + vi[2];
                double br = vr[0] - vr[2], bi = vi[0] - vi[2];
                double cr = vr[1] + vr[3], ci = vi[1] + vi[3];
                double dr = vr[1] - vr[3], di = vi[1] - vi[3];
                /* (dir) * i * d */
                double er = -dir * di, ei = dir * dr;
                tr[0] = ar + cr; ti[0] = ai + ci;
                tr[2] = ar - cr; ti[2] = ai - ci;
                tr[1] = br + er; ti[1] = bi + ei;
                tr[3] = br - er; ti[3] = bi - ei;
            } else {
                for (int q = 0; q < R; q++) {
                    double sr = 0.0, si = 0.0;
                    for (int r = 0; r < R; r++) {
                        int e = ((q * r) % R) * (n / R);
                        double c = p->w[2 * e], sn = dir * p->w[2 * e + 1];
                        sr += vr[r] * c - vi[r] * sn;
                        si += vr[r] * sn + vi[r] * c;
                    }
                    tr[q] = sr; ti[q] = si;
                }
            }
            const int base = (j / L) * L * R + k;
            for (int r = 0; r < R; r++) {
                dst[2 * (base + r * L)] = tr[r];
                dst[2 * (base + r * L) + 1] = ti[r];
            }
        }
        double *t = src; src = dst; dst = t;
        L *= R;
    }
    if (src != x) memcpy(x, src, sizeof(double) * 2 * (size_t)n);
}

/* NUMBER strided transforms, as cfft_mlt_c16 (Src/fft/_ncar/cfftmlt.f:40-56):
 * line l starts at a + l*dist, element j at + j*stride (complex units).
 * isign > 0 -> forward exp(-i..), isign < 0 -> backward exp(+i..). */
static void fft_lines(const orc_plan *p, double *a, long stride, long dist, int number,
                      int isign, double *work)
{
    const int n = p->n;
    double *x = work, *y = work + 2 * n;
    const int dir = isign > 0 ? -1 : 1;
    for (int l = 0; l < number; l++) {
        double *base = a + 2 * (long)l * dist;
        for (int j = 0; j < n; j++) {
            x[2 * j] = base[2 * j * stride];
            x[2 * j + 1] = base[2 * j * stride + 1];
        }
        fft_contig(p, x, y, dir);
        for (int j = 0; j < n; j++) {
            base[2 * j * stride] = x[2 * j];
            base[2 * j * stride + 1] = x[2 * j + 1];
        }
    }
}

void orc_fft1d(int n, double *data, int isign)
{
    orc_plan p;
    plan_init(&p, n);
    double *work = (double *)malloc(sizeof(double) * 4 * (size_t)n);
    fft_lines(&p, data, 1, n, 1, isign, work);
    free(work);
    plan_free(&p);
}

/* helpers for the 3-D drivers: chd(re/im, id, n2, n3) column-major */
#define CHD(i, j, k) (chd + 2 * ((long)(i) + (long)id * ((long)(j) + (long)n2 * (long)(k))))

static void zero_range_x(double *chd, int id, int n2, int i0, int i1, int j, int k)
{
    for (int i = i0; i < i1; i++) { CHD(i, j, k)[0] = 0.0; CHD(i, j, k)[1] = 0.0; }
}

/* ------------------------------------------------------------------ */
/* full 3-D FFT (Src/fft/_cfft/cfft_c16.f:102-165)                     */
/* ------------------------------------------------------------------ */
static void cfft_full(double *chd, int id, int n1, int n2, int n3, int mode,
                      const orc_plan *p1, const orc_plan *p2, const orc_plan *p3)
{
    const double fac = 1.0 / ((double)n1 * (double)n2 * (double)n3);
    const int nmax = (n1 > n2 ? (n1 > n3 ? n1 : n3) : (n2 > n3 ? n2 : n3));
#pragma omp parallel
    {
        double *work = (double *)malloc(sizeof(double) * 4 * (size_t)nmax);
        /* dim 1, all planes; forward renormalises (cfft_c16.f:109-122) */
#pragma omp for schedule(static, 4)
        for (int k = 0; k < n3; k++) {
            fft_lines(p1, CHD(0, 0, k), 1, id, n2, mode, work);
            if (mode == 1)
                for (int j = 0; j < n2; j++)
                    for (int i = 0; i < n1; i++) { CHD(i, j, k)[0] *= fac; CHD(i, j, k)[1] *= fac; }
        }
        /* dim 2 (cfft_c16.f:143-149) */
#pragma omp for schedule(static, 4)
        for (int k = 0; k < n3; k++) fft_lines(p2, CHD(0, 0, k), id, 1, n1, mode, work);
        /* dim 3 (cfft_c16.f:158-163) */
#pragma omp for schedule(static, 4)
        for (int j = 0; j < n2; j++) fft_lines(p3, CHD(0, j, 0), (long)id * n2, 1, n1, mode, work);
        free(work);
    }
}

/* ------------------------------------------------------------------ */
/* pruned 3-D FFT (Src/fft/_cfft/cfft_wf_c16.f:117-408)                */
/* ------------------------------------------------------------------ */
static void cfft_pruned(double *chd, int id, int n1, int n2, int n3, int kd1, int kd2, int kd3,
                        int mode, const orc_plan *p1, const orc_plan *p2, const orc_plan *p3)
{
    const double fac = 1.0 / ((double)n1 * (double)n2 * (double)n3);
    const int nmax = (n1 > n2 ? (n1 > n3 ? n1 : n3) : (n2 > n3 ? n2 : n3));
    /* planes/lines in the box: index 0..kd and n-kd..n-1 */
    const int nk = 2 * kd3 + 1, nj = 2 * kd2 + 1;
    if (mode == 1) {
#pragma omp parallel
        {
            double *work = (double *)malloc(sizeof(double) * 4 * (size_t)nmax);
            /* dim 3 on all n1*n2 lines (cfft_wf_c16.f:129-139) */
#pragma omp for schedule(static, 4)
            for (int j = 0; j < n2; j++) fft_lines(p3, CHD(0, j, 0), (long)id * n2, 1, n1, 1, work);
            /* dim 2 on the 2*kd3+1 box planes (cfft_wf_c16.f:145-164) */
#pragma omp for schedule(static, 4)
            for (int t = 0; t < nk; t++) {
                int k = t <= kd3 ? t : n3 - kd3 + (t - kd3 - 1);
                fft_lines(p2, CHD(0, 0, k), id, 1, n1, 1, work);
            }
            /* dim 1 on box lines, times 1/N (cfft_wf_c16.f:170-238) */
#pragma omp for schedule(static, 4)
            for (int t = 0; t < nk; t++) {
                int k = t <= kd3 ? t : n3 - kd3 + (t - kd3 - 1);
                for (int s = 0; s < nj; s++) {
                    int j = s <= kd2 ? s : n2 - kd2 + (s - kd2 - 1);
                    fft_lines(p1, CHD(0, j, k), 1, id, 1, 1, work);
                    for (int i = 0; i < n1; i++) { CHD(i, j, k)[0] *= fac; CHD(i, j, k)[1] *= fac; }
                }
            }
            free(work);
        }
    } else {
#pragma omp parallel
        {
            double *work = (double *)malloc(sizeof(double) * 4 * (size_t)nmax);
            /* dim 1 on box lines after re-zeroing the x gap (cfft_wf_c16.f:250-336) */
#pragma omp for schedule(static, 4)
            for (int t = 0; t < nk; t++) {
                int k = t <= kd3 ? t : n3 - kd3 + (t - kd3 - 1);
                for (int s = 0; s < nj; s++) {
                    int j = s <= kd2 ? s : n2 - kd2 + (s - kd2 - 1);
                    zero_range_x(chd, id, n2, kd1 + 1, n1 - kd1, j, k);
                    fft_lines(p1, CHD(0, j, k), 1, id, 1, -1, work);
                }
            }
            /* dim 2 on box planes after zeroing the y gap (cfft_wf_c16.f:340-378) */
#pragma omp for schedule(static, 4)
            for (int t = 0; t < nk; t++) {
                int k = t <= kd3 ? t : n3 - kd3 + (t - kd3 - 1);
                for (int j = kd2 + 1; j < n2 - kd2; j++) zero_range_x(chd, id, n2, 0, n1, j, k);
                fft_lines(p2, CHD(0, 0, k), id, 1, n1, -1, work);
            }
            /* dim 3 on all lines after zeroing the z gap (cfft_wf_c16.f:384-403) */
#pragma omp for schedule(static, 4)
            for (int j = 0; j < n2; j++) {
                for (int k = kd3 + 1; k < n3 - kd3; k++) zero_range_x(chd, id, n2, 0, n1, j, k);
                fft_lines(p3, CHD(0, j, 0), (long)id * n2, 1, n1, -1, work);
            }
            free(work);
        }
    }
}

/* cfft_wf_c16 entry (Src/fft/_cfft/cfft_wf_c16.f:91-116): pruned when every
 * n_i > 2*kd_i+5, full transform otherwise. */
int orc_cfft_wf_c16(double *chd, int id, int n1, int n2, int n3, int kd1, int kd2, int kd3,
                    int mode)
{
    if (mode != 1 && mode != -1) return ORC_ERR_MODE;
    if (n1 > ORC_MAXN || n2 > ORC_MAXN || n3 > ORC_MAXN) return ORC_ERR_SIZE;
    if (n1 < 2 || n2 < 2 || n3 < 2) return ORC_ERR_SIZE;
    orc_plan p1, p2, p3;
    plan_init(&p1, n1);
    plan_init(&p2, n2);
    plan_init(&p3, n3);
    if (n1 > 2 * kd1 + 5 && n2 > 2 * kd2 + 5 && n3 > 2 * kd3 + 5)
        cfft_pruned(chd, id, n1, n2, n3, kd1, kd2, kd3, mode, &p1, &p2, &p3);
    else
        cfft_full(chd, id, n1, n2, n3, mode, &p1, &p2, &p3);
    plan_free(&p1);
    plan_free(&p2);
    plan_free(&p3);
    return 0;
}

/* ------------------------------------------------------------------ */
/* hk_psi_nl_c16 (Src/cpw_lib/hk_psi_nl_c16.f90:70-98)                 */
/* ------------------------------------------------------------------ */
void orc_hk_psi_nl_c16(int mtxd, int neig, const double *psi, double *hpsi, const double *anlga,
                       const double *xnlkb, int nanl, int ladd, int mxddim)
{
    if (nanl <= 0 || neig <= 0) {
        if (!ladd)
            for (int n = 0; n < neig; n++)
                memset(hpsi + 2 * (long)n * mxddim, 0, sizeof(double) * 2 * (size_t)mtxd);
        return;
    }
    double *dhd = (double *)calloc((size_t)2 * nanl * neig, sizeof(double));
    /* dhd = anlga^H psi  (zgemm('c','n'), :82-83), then Diag(xnlkb) (:87-91) */
#pragma omp parallel for collapse(2) schedule(static)
    for (int n = 0; n < neig; n++)
        for (int a = 0; a < nanl; a++) {
            const double *pa = anlga + 2 * (long)a * mxddim;
            const double *pp = psi + 2 * (long)n * mxddim;
            double sr = 0.0, si = 0.0;
            for (int i = 0; i < mtxd; i++) {
                double ar = pa[2 * i], ai = pa[2 * i + 1];
                double br = pp[2 * i], bi = pp[2 * i + 1];
                sr += ar * br + ai * bi;
                si += ar * bi - ai * br;
            }
            dhd[2 * ((long)n * nanl + a)] = xnlkb[a] * sr;
            dhd[2 * ((long)n * nanl + a) + 1] = xnlkb[a] * si;
        }
    /* hpsi = zu*hpsi + anlga dhd  (zgemm('n','n'), :95-96) */
#pragma omp parallel for schedule(static)
    for (int n = 0; n < neig; n++) {
        double *ph = hpsi + 2 * (long)n * mxddim;
        if (!ladd) memset(ph, 0, sizeof(double) * 2 * (size_t)mtxd);
        for (int a = 0; a < nanl; a++) {
            const double *pa = anlga + 2 * (long)a * mxddim;
            double dr = dhd[2 * ((long)n * nanl + a)], di = dhd[2 * ((long)n * nanl + a) + 1];
            for (int i = 0; i < mtxd; i++) {
                double ar = pa[2 * i], ai = pa[2 * i + 1];
                ph[2 * i] += ar * dr - ai * di;
                ph[2 * i + 1] += ar * di + ai * dr;
            }
        }
    }
    free(dhd);
}

/* ------------------------------------------------------------------ */
/* hk_psi_c16 (Src/hk_psi/reference/hk_psi_c16.f90:105-241)            */
/* ------------------------------------------------------------------ */
int orc_hk_psi_c16(int mtxd, int neig, const double *psi, double *hpsi, int ng, const int *kgv,
                   const double *ekpg, const int *isort, const double *vscr, const int *kmscr,
                   const double *anlga, const double *xnlkb, int nanl, int mxddim, int mxdbnd,
                   int mxdanl, int mxdgve, int mxdscr)
{
    (void)mxdbnd;
    (void)mxdanl;
    if (ng > mxdgve) return ORC_ERR_NG; /* :108 */
    int nsfft[3], mxdfft, mxdwrk;
    orc_size_fft(kmscr, nsfft, &mxdfft, &mxdwrk); /* :110 */
    if (mxdwrk > mxdscr) return ORC_ERR_MXDSCR;   /* :112-119 */
    if (mtxd > mxddim) return ORC_ERR_MTXD;
    const int n1 = kmscr[3], n2 = kmscr[4], n3 = kmscr[5], id = kmscr[6]; /* :121-125 */
    const long ntot = (long)id * n2 * n3;
    if (ntot > mxdfft) mxdfft = (int)ntot; /* :128-134 */

    /* ipoint (:143-158), 0-based here */
    long *ipoint = (long *)malloc(sizeof(long) * (size_t)(mtxd > 0 ? mtxd : 1));
    int kd1 = 0, kd2 = 0, kd3 = 0;
    for (int i = 0; i < mtxd; i++) {
        const int it = isort[i] - 1;
        int k1 = kgv[3 * it], k2 = kgv[3 * it + 1], k3 = kgv[3 * it + 2];
        if (abs(k1) > kd1) kd1 = abs(k1);
        if (k1 < 0) k1 += n1;
        if (abs(k2) > kd2) kd2 = abs(k2);
        if (k2 < 0) k2 += n2;
        if (abs(k3) > kd3) kd3 = abs(k3);
        if (k3 < 0) k3 += n3;
        ipoint[i] = ((long)k3 * n2 + k2) * id + k1;
    }
    /* no wrap-around (:162-172) */
    if (kd1 > (n1 - 1) / 2 || kd2 > (n2 - 1) / 2 || kd3 > (n3 - 1) / 2) {
        free(ipoint);
        return ORC_ERR_WRAP;
    }

    double *chd = (double *)malloc(sizeof(double) * 2 * (size_t)mxdfft);
    for (int n = 0; n < neig; n++) {
        const double *pp = psi + 2 * (long)n * mxddim;
        double *ph = hpsi + 2 * (long)n * mxddim;
        /* zero + scatter (:180-190) */
#pragma omp parallel for schedule(static)
        for (long i = 0; i < ntot; i++) { chd[2 * i] = 0.0; chd[2 * i + 1] = 0.0; }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < mtxd; i++) {
            chd[2 * ipoint[i]] = pp[2 * i];
            chd[2 * ipoint[i] + 1] = pp[2 * i + 1];
        }
        /* G -> r (:194) */
        orc_cfft_wf_c16(chd, id, n1, n2, n3, kd1, kd2, kd3, -1);
        /* V(r) psi(r) (:198-202) */
#pragma omp parallel for schedule(static)
        for (long i = 0; i < ntot; i++) { chd[2 * i] *= vscr[i]; chd[2 * i + 1] *= vscr[i]; }
        /* r -> G with 1/N (:206) */
        orc_cfft_wf_c16(chd, id, n1, n2, n3, kd1, kd2, kd3, 1);
        /* gather + kinetic (:210-222) */
#pragma omp parallel for schedule(static)
        for (int i = 0; i < mtxd; i++) {
            ph[2 * i] = chd[2 * ipoint[i]] + ekpg[i] * pp[2 * i];
            ph[2 * i + 1] = chd[2 * ipoint[i] + 1] + ekpg[i] * pp[2 * i + 1];
        }
    }
    free(chd);
    free(ipoint);
    /* non-local (:231-232) */
    orc_hk_psi_nl_c16(mtxd, neig, psi, hpsi, anlga, xnlkb, nanl, 1, mxddim);
    return 0;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Sets the OpenMP thread count of later calls (n < 1: all host cores). */
void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n < 1 ? omp_get_num_procs() : n);
#else
    (void)n;
#endif
}
