/*
 * dqmc_oracle.c -- CPU restatement of the reference DQMC hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the CPU
 * baseline: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it, and only as the checker.  The product (libdqmc_hip.so) never
 * links or calls it.
 *
 * It restates katherineding/dqmc-dev call for call, with the same BLAS/LAPACK
 * sequence (linked against MKL, LP64, sequential -- the BLAS the reference is
 * built with, build/Makefile:4-7):
 *   rand_uint / rand_doub / shuffle ..... src/rand.h:6-32
 *   B / B^-1 construction ............... src/dqmc.c:29-59
 *   mul_seq ............................. src/greens.c:8-41
 *   calc_eq_g (+ phase) ................. src/greens.c:69-208
 *   calc_o / bsofi / expand_g / calc_ue_g src/greens.c:247-494
 *   update_delayed ...................... src/updates.c:7-84
 *   measure_eqlt ........................ src/meas.c:66-407
 *   measure_uneqlt (default, energy vv/vn, bond_corr) src/meas.c:728-858, 860-955
 *   optional eq/uneqlt groups ........... meas_ext.inc (thermal, 2-bond, energy
 *                                         kv/kn, gen_suscept, nematic, chiral, local JQ)
 *   dqmc() sweep loop ................... src/dqmc.c:84-459
 * Compiled twice: without and with -DORACLE_CPLX (the reference's -DUSE_CPLX).
 * Parity pins: see tests/test_oracle_*.py.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "dqmc_hip.h"

#ifdef ORACLE_CPLX
typedef double _Complex num;
#define ABS(x) cabs(x)
#define RE(x) creal(x)
#define CONJ(x) conj(x)
#define SYM(name) oracle_c_##name
#else
typedef double num;
#define ABS(x) fabs(x)
#define RE(x) (x)
#define CONJ(x) (x)
#define SYM(name) oracle_r_##name
#endif

/* ------------------------------------------------------------------------ */
/* BLAS / LAPACK (Fortran ABI, LP64)                                         */
/* ------------------------------------------------------------------------ */
#ifdef ORACLE_CPLX
void zgemm_(const char *, const char *, const int *, const int *, const int *, const num *,
            const num *, const int *, const num *, const int *, const num *, num *, const int *);
void zgemv_(const char *, const int *, const int *, const num *, const num *, const int *,
            const num *, const int *, const num *, num *, const int *);
void ztrmm_(const char *, const char *, const char *, const char *, const int *, const int *,
            const num *, const num *, const int *, num *, const int *);
void zgetrf_(const int *, const int *, num *, const int *, int *, int *);
void zgetri_(const int *, num *, const int *, const int *, num *, const int *, int *);
void zgetrs_(const char *, const int *, const int *, const num *, const int *, const int *,
             num *, const int *, int *);
void zgeqp3_(const int *, const int *, num *, const int *, int *, num *, num *, const int *,
             double *, int *);
void zgeqrf_(const int *, const int *, num *, const int *, num *, num *, const int *, int *);
void zunmqr_(const char *, const char *, const int *, const int *, const int *, const num *,
             const int *, const num *, num *, const int *, num *, const int *, int *);
void ztrtri_(const char *, const char *, const int *, num *, const int *, int *);
#define XGEMM zgemm_
#define XGEMV zgemv_
#define XTRMM ztrmm_
#define XGETRF zgetrf_
#define XGETRI zgetri_
#define XGETRS zgetrs_
#define XGEQRF zgeqrf_
#define XTRTRI ztrtri_
#else
void dgemm_(const char *, const char *, const int *, const int *, const int *, const num *,
            const num *, const int *, const num *, const int *, const num *, num *, const int *);
void dgemv_(const char *, const int *, const int *, const num *, const num *, const int *,
            const num *, const int *, const num *, num *, const int *);
void dtrmm_(const char *, const char *, const char *, const char *, const int *, const int *,
            const num *, const num *, const int *, num *, const int *);
void dgetrf_(const int *, const int *, num *, const int *, int *, int *);
void dgetri_(const int *, num *, const int *, const int *, num *, const int *, int *);
void dgetrs_(const char *, const int *, const int *, const num *, const int *, const int *,
             num *, const int *, int *);
void dgeqp3_(const int *, const int *, num *, const int *, int *, num *, num *, const int *,
             int *);
void dgeqrf_(const int *, const int *, num *, const int *, num *, num *, const int *, int *);
void dormqr_(const char *, const char *, const int *, const int *, const int *, const num *,
             const int *, const num *, num *, const int *, num *, const int *, int *);
void dtrtri_(const char *, const char *, const int *, num *, const int *, int *);
#define XGEMM dgemm_
#define XGEMV dgemv_
#define XTRMM dtrmm_
#define XGETRF dgetrf_
#define XGETRI dgetri_
#define XGETRS dgetrs_
#define XGEQRF dgeqrf_
#define XTRTRI dtrtri_
#endif

static void gemm(const char *ta, const char *tb, int m, int n, int k, num alpha, const num *a,
                 int lda, const num *b, int ldb, num beta, num *c, int ldc)
{
    if (m == 0 || n == 0) return;
    if (k == 0) { /* BLAS semantics: C = beta*C */
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) c[i + ldc * j] *= beta;
        return;
    }
    XGEMM(ta, tb, &m, &n, &k, &alpha, a, &lda, b, &ldb, &beta, c, &ldc);
}

static void gemv(int m, int n, num alpha, const num *a, int lda, const num *x, int incx,
                 num beta, num *y, int incy)
{
    if (n == 0) return; /* y = 1*y */
    XGEMV("N", &m, &n, &alpha, a, &lda, x, &incx, &beta, y, &incy);
}

static void trmm(const char *side, const char *uplo, int m, int n, num alpha, const num *a,
                 int lda, num *b, int ldb)
{
    XTRMM(side, uplo, "N", "N", &m, &n, &alpha, a, &lda, b, &ldb);
}

static void geqp3(int m, int n, num *a, int lda, int *jpvt, num *tau, num *work, int lwork,
                  num *rwork)
{
    int info;
#ifdef ORACLE_CPLX
    zgeqp3_(&m, &n, a, &lda, jpvt, tau, work, &lwork, (double *)rwork, &info);
#else
    (void)rwork;
    dgeqp3_(&m, &n, a, &lda, jpvt, tau, work, &lwork, &info);
#endif
}

static void geqrf(int m, int n, num *a, int lda, num *tau, num *work, int lwork)
{
    int info;
    XGEQRF(&m, &n, a, &lda, tau, work, &lwork, &info);
}

/* trans 'C' = conjugate transpose (real: transpose), as src/linalg_mkl.h:166 */
static void unmqr(const char *side, const char *trans, int m, int n, int k, const num *a,
                  int lda, const num *tau, num *c, int ldc, num *work, int lwork)
{
    int info;
#ifdef ORACLE_CPLX
    zunmqr_(side, trans, &m, &n, &k, a, &lda, tau, c, &ldc, work, &lwork, &info);
#else
    dormqr_(side, trans[0] == 'C' ? "T" : trans, &m, &n, &k, a, &lda, tau, c, &ldc, work,
            &lwork, &info);
#endif
}

static void getrf(int m, int n, num *a, int lda, int *ipiv)
{
    int info;
    XGETRF(&m, &n, a, &lda, ipiv, &info);
}

static void getrs(int n, int nrhs, const num *a, int lda, const int *ipiv, num *b, int ldb)
{
    int info;
    XGETRS("N", &n, &nrhs, a, &lda, ipiv, b, &ldb, &info);
}

static void getri(int n, num *a, int lda, const int *ipiv, num *work, int lwork)
{
    int info;
    XGETRI(&n, a, &lda, ipiv, work, &lwork, &info);
}

static void trtri(int n, num *a, int lda)
{
    int info;
    XTRTRI("U", "N", &n, a, &lda, &info);
}

static void *zalloc(size_t bytes)
{
    void *p = NULL;
    if (posix_memalign(&p, 64, bytes ? bytes : 64)) return NULL;
    memset(p, 0, bytes ? bytes : 64);
    return p;
}

/* ------------------------------------------------------------------------ */
/* RNG (src/rand.h:6-32)                                                     */
/* ------------------------------------------------------------------------ */
static inline uint64_t rand_uint(uint64_t *s)
{
    const uint64_t s0 = s[s[16]];
    const int p = (int)((s[16] + 1) & 15);
    s[16] = (uint64_t)p;
    uint64_t s1 = s[p];
    s1 ^= s1 << 31;
    s[p] = s1 ^ s0 ^ (s1 >> 11) ^ (s0 >> 30);
    return s[p] * UINT64_C(1181783497276652981);
}

static inline double rand_doub(uint64_t *s)
{
    union { uint64_t i; double d; } u;
    u.i = (UINT64_C(0x3FF) << 52) | (rand_uint(s) >> 12);
    return u.d - 1.0;
}

static inline void shuffle(uint64_t *s, int n, int *a)
{
    for (int i = 0; i < n; i++) {
        const int j = (int)((rand_uint(s) >> 3) % (uint64_t)(i + 1));
        if (i != j) a[i] = a[j];
        a[j] = i;
    }
}

/* ------------------------------------------------------------------------ */
/* B_l and B_l^{-1} (src/dqmc.c:29-59)                                       */
/* spin 0 (up) couples to hs, spin 1 (down) to !hs.                          */
/* ------------------------------------------------------------------------ */
static void calc_B(int N, const num *expK, const double *exp_lambda, const int *hs_l, int spin,
                   num *B)
{
    for (int j = 0; j < N; j++) {
        const int s = spin ? !hs_l[j] : hs_l[j];
        const double el = exp_lambda[j + N * s];
        for (int i = 0; i < N; i++) B[i + N * j] = expK[i + N * j] * el;
    }
}

static void calc_iB(int N, const num *iexpK, const double *exp_lambda, const int *hs_l,
                    int spin, num *iB)
{
    for (int i = 0; i < N; i++) {
        const int s = spin ? hs_l[i] : !hs_l[i];
        const double el = exp_lambda[i + N * s];
        for (int j = 0; j < N; j++) iB[i + N * j] = el * iexpK[i + N * j];
    }
}

/* ------------------------------------------------------------------------ */
/* mul_seq (src/greens.c:8-41): A = alpha * B[maxp1-1] ... B[min]  (mod L)   */
/* ------------------------------------------------------------------------ */
static void mul_seq(int N, int L, int min, int maxp1, num alpha, const num *B, num *A, int ldA,
                    num *work)
{
    const int n_mul = (min == maxp1) ? L : (L + maxp1 - min) % L;
    const size_t NN = (size_t)N * N;
    if (n_mul == 1) {
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++) A[i + (size_t)ldA * j] = alpha * B[i + N * j + NN * min];
        return;
    }
    int l = min;
    if (n_mul % 2 == 0) {
        gemm("N", "N", N, N, N, alpha, B + NN * ((l + 1) % L), N, B + NN * l, N, 0.0, A, ldA);
        l = (l + 2) % L;
    } else {
        gemm("N", "N", N, N, N, alpha, B + NN * ((l + 1) % L), N, B + NN * l, N, 0.0, work, N);
        gemm("N", "N", N, N, N, 1.0, B + NN * ((l + 2) % L), N, work, N, 0.0, A, ldA);
        l = (l + 3) % L;
    }
    for (; l != maxp1; l = (l + 2) % L) {
        gemm("N", "N", N, N, N, 1.0, B + NN * l, N, A, ldA, 0.0, work, N);
        gemm("N", "N", N, N, N, 1.0, B + NN * ((l + 1) % L), N, work, N, 0.0, A, ldA);
    }
}

/* ------------------------------------------------------------------------ */
/* calc_eq_g (src/greens.c:69-208)                                           */
/* ------------------------------------------------------------------------ */
static num calc_eq_g(int l, int N, int L, int n_mul, const num *B, num *g, num *Q, num *T,
                     num *tau, num *d, num *v, int *pvt, num *work, int lwork)
{
    /* Algorithm 3 of Tomas et al., IPDPS 2012, with products of n_mul
     * consecutive matrices formed by GEMM. */
    int m = (l + 1 + (L - 1) % n_mul) % L;
    mul_seq(N, L, l, m, 1.0, B, Q, N, work);

    for (int i = 0; i < N; i++) pvt[i] = 0;
    geqp3(N, N, Q, N, pvt, tau, work, lwork, d); /* d doubles as rwork */

    for (int i = 0; i < N; i++) {
        d[i] = Q[i + i * N];
        if (d[i] == 0.0) d[i] = 1.0;
        v[i] = 1.0 / d[i];
    }
    for (int i = 0; i < N * N; i++) T[i] = 0.0;
    for (int j = 0; j < N; j++)
        for (int i = 0; i <= j; i++) T[i + (pvt[j] - 1) * N] = v[i] * Q[i + j * N];

    while (m != l) {
        const int next = (m + n_mul) % L;
        mul_seq(N, L, m, next, 1.0, B, g, N, work);
        m = next;

        unmqr("R", "N", N, N, N, Q, N, tau, g, N, work, lwork); /* g = g*Q */
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++) g[i + j * N] *= d[j];

        for (int j = 0; j < N; j++) { /* squared column norms */
            d[j] = 0.0;
            for (int i = 0; i < N; i++) d[j] += g[i + j * N] * CONJ(g[i + j * N]);
        }
        pvt[0] = 0; /* stable insertion sort, descending by norm */
        for (int i = 1; i < N; i++) {
            int j;
            for (j = i; j > 0 && RE(d[pvt[j - 1]]) < RE(d[i]); j--) pvt[j] = pvt[j - 1];
            pvt[j] = i;
        }
        for (int j = 0; j < N; j++) memcpy(Q + j * N, g + pvt[j] * N, N * sizeof(num));

        geqrf(N, N, Q, N, tau, work, lwork);

        for (int i = 0; i < N; i++) {
            d[i] = Q[i + i * N];
            if (d[i] == 0.0) d[i] = 1.0;
            v[i] = 1.0 / d[i];
        }
        for (int j = 0; j < N; j++)
            for (int i = 0; i <= j; i++) Q[i + j * N] *= v[i];
        for (int j = 0; j < N; j++) {
            for (int i = 0; i < N; i++) v[i] = T[pvt[i] + j * N];
            memcpy(T + j * N, v, N * sizeof(num));
        }
        trmm("L", "U", N, N, 1.0, Q, N, T, N);
    }

    /* g from Eq. 2.12 of Bai, Lee, Li, Xu, LAA 2011 */
    for (int i = 0; i < N * N; i++) g[i] = 0.0;
    for (int i = 0; i < N; i++) {
        if (ABS(d[i]) > 1.0) {
            v[i] = 1.0 / d[i];
            d[i] = 1.0;
        } else {
            v[i] = 1.0;
        }
        g[i + i * N] = v[i];
    }
    unmqr("R", "C", N, N, N, Q, N, tau, g, N, work, lwork);
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) T[i + j * N] *= d[i];
    for (int i = 0; i < N * N; i++) T[i] += g[i];
    getrf(N, N, T, N, pvt);
    getrs(N, N, T, N, pvt, g, N);

    num phase = 1.0;
    for (int i = 0; i < N; i++) {
        const num c = v[i] / T[i + N * i];
        phase *= c / ABS(c);
        double vv = 1.0;
        for (int j = i + 1; j < N; j++)
            vv += RE(Q[j + i * N] * CONJ(Q[j + i * N]));
        const num ref = 1.0 - tau[i] * vv;
        phase *= ABS(ref) / ref;
        if (pvt[i] != i + 1) phase *= -1.0;
    }
    return 1.0 / phase;
}

/* ------------------------------------------------------------------------ */
/* Unequal-time Green's functions (src/greens.c:247-494)                     */
/* ------------------------------------------------------------------------ */
static void calc_o(int N, int L, int n_mul, const num *B, num *G, num *work)
{
    const int E = 1 + (L - 1) / n_mul;
    const int NE = N * E;
    for (size_t i = 0; i < (size_t)NE * NE; i++) G[i] = 0.0;
    for (int e = 0; e < E - 1; e++)
        mul_seq(N, L, e * n_mul, (e + 1) * n_mul, -1.0, B, G + N * (e + 1) + (size_t)NE * N * e,
                NE, work);
    mul_seq(N, L, (E - 1) * n_mul, 0, 1.0, B, G + (size_t)NE * N * (E - 1), NE, work);
    for (int i = 0; i < NE; i++) G[i + (size_t)NE * i] += 1.0;
}

static void bsofi(int N, int L, num *G, num *tau, num *Q, num *work, int lwork)
{
    if (L == 1) {
        int *ipiv = (int *)tau;
        getrf(N, N, G, N, ipiv);
        getri(N, G, N, ipiv, work, lwork);
        return;
    }
    const int NL = N * L;
    const int N2 = 2 * N;
#define G_BLK(i, j) (G + N * (i) + (size_t)NL * N * (j))
    for (int l = 0; l < L - 2; l++) {
        geqrf(N2, N, G_BLK(l, l), NL, tau + N * l, work, lwork);
        unmqr("L", "C", N2, N, N, G_BLK(l, l), NL, tau + N * l, G_BLK(l, l + 1), NL, work, lwork);
        unmqr("L", "C", N2, N, N, G_BLK(l, l), NL, tau + N * l, G_BLK(l, L - 1), NL, work, lwork);
    }
    geqrf(N2, N2, G_BLK(L - 2, L - 2), NL, tau + N * (L - 2), work, lwork);

    if (L <= 2) {
        trtri(NL, G, NL);
    } else {
        trtri(3 * N, G_BLK(L - 3, L - 3), NL);
        if (L > 3) {
            trmm("R", "U", N * (L - 3), N, 1.0, G_BLK(L - 1, L - 1), NL, G_BLK(0, L - 1), NL);
            for (int l = L - 4; l >= 0; l--) {
                trtri(N, G_BLK(l, l), NL);
                trmm("L", "U", N, N, -1.0, G_BLK(l, l), NL, G_BLK(l, L - 1), NL);
                trmm("L", "U", N, N, -1.0, G_BLK(l, l), NL, G_BLK(l, l + 1), NL);
                gemm("N", "N", N, N * (L - l - 2), N, 1.0, G_BLK(l, l + 1), NL,
                     G_BLK(l + 1, l + 2), NL, 1.0, G_BLK(l, l + 2), NL);
                trmm("R", "U", N, N, 1.0, G_BLK(l + 1, l + 1), NL, G_BLK(l, l + 1), NL);
            }
        }
    }

    for (int i = 0; i < 4 * N * N; i++) Q[i] = 0.0;
    for (int j = 0; j < N2; j++)
        for (int i = j + 1; i < N2; i++) {
            Q[i + N2 * j] = G_BLK(L - 2, L - 2)[i + (size_t)NL * j];
            G_BLK(L - 2, L - 2)[i + (size_t)NL * j] = 0.0;
        }
    unmqr("R", "C", NL, N2, N2, Q, N2, tau + N * (L - 2), G_BLK(0, L - 2), NL, work, lwork);
    for (int l = L - 3; l >= 0; l--) {
        for (int j = 0; j < N; j++)
            for (int i = j + 1; i < N2; i++) {
                Q[i + N2 * j] = G_BLK(l, l)[i + (size_t)NL * j];
                G_BLK(l, l)[i + (size_t)NL * j] = 0.0;
            }
        unmqr("R", "C", NL, N2, N, Q, N2, tau + N * l, G_BLK(0, l), NL, work, lwork);
    }
#undef G_BLK
}

static void expand_g(int N, int L, int E, int n_matmul, const num *B, const num *iB,
                     const num *Gred, num *G0t, num *Gtt, num *Gt0)
{
    const size_t NN = (size_t)N * N;
    const int n_left = (n_matmul - 1) / 2;
    const int n_right = n_matmul / 2;
    const int n_up = n_left;
    const int n_down = n_right;
    const int rstop_last = ((E - 1) * n_matmul + L) / 2;
    const int lstop_first = (rstop_last + 1) % L;
    const int dstop_last = rstop_last;
    const int ustop_first = lstop_first;
    const size_t ldr = (size_t)N * E;

    for (int f = 0; f < E; f++) {
        const int t = f * n_matmul;
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++) G0t[i + N * j + NN * t] = Gred[i + ldr * (j + N * f)];
    }
    for (int f = 0; f < E; f++) {
        const int l = f * n_matmul;
        const int lstop = (f == 0) ? lstop_first : l - n_left;
        const int rstop = (f == E - 1) ? rstop_last : l + n_right;
        for (int m = l; m != lstop;) {
            const int next = (m - 1 + L) % L;
            const num alpha = (m == 0) ? -1.0 : 1.0;
            gemm("N", "N", N, N, N, alpha, G0t + NN * m, N, B + NN * next, N, 0.0, G0t + NN * next,
                 N);
            m = next;
        }
        for (int m = l; m != rstop;) {
            const int next = (m + 1) % L;
            const num alpha = (next == 0) ? -1.0 : 1.0;
            const num beta = (m == 0) ? -alpha : 0.0;
            if (m == 0)
                for (size_t k = 0; k < NN; k++) G0t[k + NN * next] = iB[k + NN * m];
            gemm("N", "N", N, N, N, alpha, G0t + NN * m, N, iB + NN * m, N, beta, G0t + NN * next,
                 N);
            m = next;
        }
    }

    for (int e = 0; e < E; e++) {
        const int k = e * n_matmul;
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++)
                Gtt[i + N * j + NN * k] = Gred[(i + N * e) + ldr * (j + N * e)];
    }
    for (int e = 0; e < E; e++) {
        const int k = e * n_matmul;
        const int ustop = (e == 0) ? ustop_first : k - n_up;
        const int dstop = (e == E - 1) ? dstop_last : k + n_down;
        for (int m = k; m != ustop;) {
            const int next = (m - 1 + L) % L;
            gemm("N", "N", N, N, N, 1.0, Gtt + NN * m, N, B + NN * next, N, 0.0, Gt0, N);
            gemm("N", "N", N, N, N, 1.0, iB + NN * next, N, Gt0, N, 0.0, Gtt + NN * next, N);
            m = next;
        }
        for (int m = k; m != dstop;) {
            const int next = (m + 1) % L;
            gemm("N", "N", N, N, N, 1.0, Gtt + NN * m, N, iB + NN * m, N, 0.0, Gt0, N);
            gemm("N", "N", N, N, N, 1.0, B + NN * m, N, Gt0, N, 0.0, Gtt + NN * next, N);
            m = next;
        }
    }

    for (int e = 0; e < E; e++) {
        const int t = e * n_matmul;
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++) Gt0[i + N * j + NN * t] = Gred[(i + N * e) + ldr * j];
    }
    for (int e = 0; e < E; e++) {
        const int k = e * n_matmul;
        const int ustop = (e == 0) ? ustop_first : k - n_up;
        const int dstop = (e == E - 1) ? dstop_last : k + n_down;
        for (int m = k; m != ustop;) {
            const int next = (m - 1 + L) % L;
            const num alpha = (m == 0) ? -1.0 : 1.0;
            const num beta = (m == 0) ? -alpha : 0.0;
            if (m == 0)
                for (size_t q = 0; q < NN; q++) Gt0[q + NN * next] = iB[q + NN * next];
            gemm("N", "N", N, N, N, alpha, iB + NN * next, N, Gt0 + NN * m, N, beta,
                 Gt0 + NN * next, N);
            m = next;
        }
        for (int m = k; m != dstop;) {
            const int next = (m + 1) % L;
            const num alpha = (next == 0) ? -1.0 : 1.0;
            gemm("N", "N", N, N, N, alpha, B + NN * m, N, Gt0 + NN * m, N, 0.0, Gt0 + NN * next,
                 N);
            if (next == 0)
                for (int i = 0; i < N; i++) Gt0[i + N * i + NN * next] += 1.0;
            m = next;
        }
    }
}

/* workspace size for bsofi/geqrf/unmqr on (N,E) */
static int lwork_for(int N, int E)
{
    int lw = N * N;
    const int b = 64 * (2 * N + 64) + 4 * N * E + 256;
    if (b > lw) lw = b;
    return lw;
}

static void calc_ue_g(int N, int L, int F, int n_mul, const num *B, const num *iB, const num *C,
                      num *G0t, num *Gtt, num *Gt0, num *Gred, num *tau, num *Q, num *work,
                      int lwork)
{
    const int E = 1 + (F - 1) / n_mul;
    calc_o(N, F, n_mul, C, Gred, work);
    bsofi(N, E, Gred, tau, Q, work, lwork);
    expand_g(N, L, E, (L / F) * n_mul, B, iB, Gred, G0t, Gtt, Gt0);
}

/* ------------------------------------------------------------------------ */
/* update_delayed (src/updates.c:7-84), without the nested OpenMP sections   */
/* ------------------------------------------------------------------------ */
static int64_t update_delayed(int N, int n_delay, const double *del, const int *site_order,
                              uint64_t *rng, int *hs, num *gu, num *gd, num *phase, num *au,
                              num *bu, num *du, num *ad, num *bd, num *dd)
{
    int64_t n_acc = 0;
    int k = 0;
    for (int j = 0; j < N; j++) du[j] = gu[j + N * j];
    for (int j = 0; j < N; j++) dd[j] = gd[j + N * j];
    for (int ii = 0; ii < N; ii++) {
        const int i = site_order[ii];
        const double delu = del[i + N * hs[i]];
        const double deld = del[i + N * !hs[i]];
        if (delu == 0.0 && deld == 0.0) continue;
        const num ru = 1.0 + (1.0 - du[i]) * delu;
        const num rd = 1.0 + (1.0 - dd[i]) * deld;
        const num prob = ru * rd;
        const double absprob = ABS(prob);
        if (rand_doub(rng) < absprob) {
            for (int j = 0; j < N; j++) au[j + N * k] = gu[j + N * i];
            for (int j = 0; j < N; j++) bu[j + N * k] = gu[i + N * j];
            gemv(N, k, 1.0, au, N, bu + i, N, 1.0, au + N * k, 1);
            gemv(N, k, 1.0, bu, N, au + i, N, 1.0, bu + N * k, 1);
            au[i + N * k] -= 1.0;
            for (int j = 0; j < N; j++) au[j + N * k] *= delu / ru;
            for (int j = 0; j < N; j++) du[j] += au[j + N * k] * bu[j + N * k];

            for (int j = 0; j < N; j++) ad[j + N * k] = gd[j + N * i];
            for (int j = 0; j < N; j++) bd[j + N * k] = gd[i + N * j];
            gemv(N, k, 1.0, ad, N, bd + i, N, 1.0, ad + N * k, 1);
            gemv(N, k, 1.0, bd, N, ad + i, N, 1.0, bd + N * k, 1);
            ad[i + N * k] -= 1.0;
            for (int j = 0; j < N; j++) ad[j + N * k] *= deld / rd;
            for (int j = 0; j < N; j++) dd[j] += ad[j + N * k] * bd[j + N * k];

            k++;
            n_acc++;
            hs[i] = !hs[i];
            *phase *= prob / absprob;
        }
        if (k == n_delay) {
            k = 0;
            gemm("N", "T", N, N, n_delay, 1.0, au, N, bu, N, 1.0, gu, N);
            for (int j = 0; j < N; j++) du[j] = gu[j + N * j];
            gemm("N", "T", N, N, n_delay, 1.0, ad, N, bd, N, 1.0, gd, N);
            for (int j = 0; j < N; j++) dd[j] = gd[j + N * j];
        }
    }
    gemm("N", "T", N, N, k, 1.0, au, N, bu, N, 1.0, gu, N);
    gemm("N", "T", N, N, k, 1.0, ad, N, bd, N, 1.0, gd, N);
    return n_acc;
}

/* ------------------------------------------------------------------------ */
/* Measurements                                                              */
/* ------------------------------------------------------------------------ */
#define ACC(m, k) ((num *)(m)->acc[k])

static inline num load_num(const double *p, long i)
{
#ifdef ORACLE_CPLX
    return p[2 * i] + I * p[2 * i + 1];
#else
    return p[i];
#endif
}

#include "meas_ext.inc"



/* measure_eqlt (src/meas.c:66-407) */
static void measure_eqlt(const dqmc_params *p, num phase, const num *gu, const num *gd,
                         dqmc_meas *m)
{
    const int N = p->N;
    const int num_b = p->num_b;
    m->n_sample++;
    {
        num s = load_num(m->sign, 0) + phase;
        m->sign[0] = RE(s);
#ifdef ORACLE_CPLX
        m->sign[1] = cimag(s);
#endif
    }
    const num *pu = (const num *)p->peierlsu, *pd = (const num *)p->peierlsd;
    for (int i = 0; i < N; i++) {
        const int r = p->map_i[i];
        const num pre = phase / p->degen_i;
        const num guii = gu[i + i * N], gdii = gd[i + i * N];
        ACC(m, DQ_EQ_density)[r] += pre * (2. - guii - gdii);
        ACC(m, DQ_EQ_density_u)[r] += pre * (1. - guii);
        ACC(m, DQ_EQ_density_d)[r] += pre * (1. - gdii);
        ACC(m, DQ_EQ_double_occ)[r] += pre * (1. - guii) * (1. - gdii);
    }
    for (int j = 0; j < N; j++) {
        for (int i = 0; i < N; i++) {
            const int delta = (i == j);
            const int r = p->map_ij[i + j * N];
            const num pre = phase / p->degen_ij;
            const num guii = gu[i + i * N], gdii = gd[i + i * N];
            const num guij = gu[i + j * N], gdij = gd[i + j * N];
            const num guji = gu[j + i * N], gdji = gd[j + i * N];
            const num gujj = gu[j + j * N], gdjj = gd[j + j * N];
#ifdef ORACLE_CPLX
            ACC(m, DQ_EQ_g00)[r] += 0.5 * pre * (guij * pu[j + i * N] + gdij * pd[j + i * N]);
            ACC(m, DQ_EQ_g00_u)[r] += pre * (guij * pu[j + i * N]);
            ACC(m, DQ_EQ_g00_d)[r] += pre * (gdij * pd[j + i * N]);
#else
            (void)pu; (void)pd;
            ACC(m, DQ_EQ_g00)[r] += 0.5 * pre * (guij + gdij);
            ACC(m, DQ_EQ_g00_u)[r] += pre * guij;
            ACC(m, DQ_EQ_g00_d)[r] += pre * gdij;
#endif
            const num x = delta * (guii + gdii) - (guji * guij + gdji * gdij);
            ACC(m, DQ_EQ_nn)[r] += pre * ((2. - guii - gdii) * (2. - gujj - gdjj) + x);
            ACC(m, DQ_EQ_xx)[r] += 0.25 * pre * (delta * (guii + gdii) - (guji * gdij + gdji * guij));
            ACC(m, DQ_EQ_zz)[r] += 0.25 * pre * ((gdii - guii) * (gdjj - gujj) + x);
            ACC(m, DQ_EQ_pair_sw)[r] += pre * guij * gdij;
            if (p->meas_energy_corr) {
                const num nuinuj = (1. - guii) * (1. - gujj) + (delta - guji) * guij;
                const num ndindj = (1. - gdii) * (1. - gdjj) + (delta - gdji) * gdij;
                ACC(m, DQ_EQ_vv)[r] += pre * nuinuj * ndindj;
                ACC(m, DQ_EQ_vn)[r] += pre * (nuinuj * (1. - gdii) + (1. - guii) * ndindj);
            }
        }
    }
    measure_eqlt_ext(p, phase, gu, gd, m);
    if (!p->meas_energy_corr) return;

    /* 1 bond - 1 site (src/meas.c:290-330) */
    for (int j = 0; j < N; j++) {
        for (int b = 0; b < num_b; b++) {
            const int i0 = p->bonds[b], i1 = p->bonds[b + num_b];
#ifdef ORACLE_CPLX
            const num pui0i1 = pu[i0 + N * i1], pui1i0 = pu[i1 + N * i0];
            const num pdi0i1 = pd[i0 + N * i1], pdi1i0 = pd[i1 + N * i0];
#else
            const num pui0i1 = 1, pui1i0 = 1, pdi0i1 = 1, pdi1i0 = 1;
#endif
            const int bs = p->map_bs[b + num_b * j];
            const num pre = phase / p->degen_bs;
            const int delta_i0i1 = 0;
            const int delta_i0j = (i0 == j), delta_i1j = (i1 == j);
            const num gui0j = gu[i0 + N * j], guji0 = gu[j + N * i0];
            const num gdi0j = gd[i0 + N * j], gdji0 = gd[j + N * i0];
            const num gui1j = gu[i1 + N * j], guji1 = gu[j + N * i1];
            const num gdi1j = gd[i1 + N * j], gdji1 = gd[j + N * i1];
            const num gui0i1 = gu[i0 + N * i1], gui1i0 = gu[i1 + N * i0];
            const num gdi0i1 = gd[i0 + N * i1], gdi1i0 = gd[i1 + N * i0];
            const num gujj = gu[j + N * j], gdjj = gd[j + N * j];
            const num ku = pui1i0 * (delta_i0i1 - gui0i1) + pui0i1 * (delta_i0i1 - gui1i0);
            const num kd = pdi1i0 * (delta_i0i1 - gdi0i1) + pdi0i1 * (delta_i0i1 - gdi1i0);
            const num xu = pui0i1 * (delta_i0j - guji0) * gui1j + pui1i0 * (delta_i1j - guji1) * gui0j;
            const num xd = pdi0i1 * (delta_i0j - gdji0) * gdi1j + pdi1i0 * (delta_i1j - gdji1) * gdi0j;
            ACC(m, DQ_EQ_kv)[bs] += pre * ((ku * (1. - gujj) + xu) * (1. - gdjj) +
                                           (kd * (1. - gdjj) + xd) * (1. - gujj));
            ACC(m, DQ_EQ_kn)[bs] += pre * ((ku + kd) * (2. - gujj - gdjj) + xu + xd);
        }
    }
    /* 1 bond - 1 bond (src/meas.c:333-406) */
    for (int c = 0; c < num_b; c++) {
        const int j0 = p->bonds[c], j1 = p->bonds[c + num_b];
#ifdef ORACLE_CPLX
        const num puj0j1 = pu[j0 + N * j1], puj1j0 = pu[j1 + N * j0];
        const num pdj0j1 = pd[j0 + N * j1], pdj1j0 = pd[j1 + N * j0];
#else
        const num puj0j1 = 1, puj1j0 = 1, pdj0j1 = 1, pdj1j0 = 1;
#endif
        for (int b = 0; b < num_b; b++) {
            const int i0 = p->bonds[b], i1 = p->bonds[b + num_b];
#ifdef ORACLE_CPLX
            const num pui0i1 = pu[i0 + N * i1], pui1i0 = pu[i1 + N * i0];
            const num pdi0i1 = pd[i0 + N * i1], pdi1i0 = pd[i1 + N * i0];
#else
            const num pui0i1 = 1, pui1i0 = 1, pdi0i1 = 1, pdi1i0 = 1;
#endif
            const int bb = p->map_bb[b + c * num_b];
            const num pre = phase / p->degen_bb;
            const int delta_i0j0 = (i0 == j0), delta_i1j0 = (i1 == j0);
            const int delta_i0j1 = (i0 == j1), delta_i1j1 = (i1 == j1);
            const num gui1i0 = gu[i1 + i0 * N], gui0i1 = gu[i0 + i1 * N];
            const num gui0j0 = gu[i0 + j0 * N], gui1j0 = gu[i1 + j0 * N];
            const num gui0j1 = gu[i0 + j1 * N], gui1j1 = gu[i1 + j1 * N];
            const num guj0i0 = gu[j0 + i0 * N], guj1i0 = gu[j1 + i0 * N];
            const num guj0i1 = gu[j0 + i1 * N], guj1i1 = gu[j1 + i1 * N];
            const num guj1j0 = gu[j1 + j0 * N], guj0j1 = gu[j0 + j1 * N];
            const num gdi1i0 = gd[i1 + i0 * N], gdi0i1 = gd[i0 + i1 * N];
            const num gdi0j0 = gd[i0 + j0 * N], gdi1j0 = gd[i1 + j0 * N];
            const num gdi0j1 = gd[i0 + j1 * N], gdi1j1 = gd[i1 + j1 * N];
            const num gdj0i0 = gd[j0 + i0 * N], gdj1i0 = gd[j1 + i0 * N];
            const num gdj0i1 = gd[j0 + i1 * N], gdj1i1 = gd[j1 + i1 * N];
            const num gdj1j0 = gd[j1 + j0 * N], gdj0j1 = gd[j0 + j1 * N];
            const num x = pui0i1 * puj0j1 * (delta_i0j1 - guj1i0) * gui1j0 +
                          pui1i0 * puj1j0 * (delta_i1j0 - guj0i1) * gui0j1 +
                          pdi0i1 * pdj0j1 * (delta_i0j1 - gdj1i0) * gdi1j0 +
                          pdi1i0 * pdj1j0 * (delta_i1j0 - gdj0i1) * gdi0j1;
            const num y = pui0i1 * puj1j0 * (delta_i0j0 - guj0i0) * gui1j1 +
                          pui1i0 * puj0j1 * (delta_i1j1 - guj1i1) * gui0j0 +
                          pdi0i1 * pdj1j0 * (delta_i0j0 - gdj0i0) * gdi1j1 +
                          pdi1i0 * pdj0j1 * (delta_i1j1 - gdj1i1) * gdi0j0;
            ACC(m, DQ_EQ_kk)[bb] +=
                pre * ((pui1i0 * gui0i1 + pui0i1 * gui1i0 + pdi1i0 * gdi0i1 + pdi0i1 * gdi1i0) *
                           (puj1j0 * guj0j1 + puj0j1 * guj1j0 + pdj1j0 * gdj0j1 + pdj0j1 * gdj1j0) +
                       x + y);
        }
    }
}

/* measure_uneqlt (src/meas.c:728-858 default + vv/vn; :860-955 bond_corr) */
static void measure_uneqlt(const dqmc_params *p, num phase, const num *Gu0t, const num *Gutt,
                           const num *Gut0, const num *Gd0t, const num *Gdtt, const num *Gdt0,
                           dqmc_meas *m)
{
    const int N = p->N, L = p->L, num_ij = p->num_ij;
    const int num_b = p->num_b, num_bb = p->num_bb;
    const size_t NN = (size_t)N * N;
    m->n_sample++;
    {
        num s = load_num(m->sign, 0) + phase;
        m->sign[0] = RE(s);
#ifdef ORACLE_CPLX
        m->sign[1] = cimag(s);
#endif
    }
    const num *Gu00 = Gutt, *Gd00 = Gdtt;
    const num *pu = (const num *)p->peierlsu, *pd = (const num *)p->peierlsd;
    for (int t = 0; t < L; t++) {
        const int delta_t = (t == 0);
        const num *Gu0t_t = Gu0t + NN * t, *Gutt_t = Gutt + NN * t, *Gut0_t = Gut0 + NN * t;
        const num *Gd0t_t = Gd0t + NN * t, *Gdtt_t = Gdtt + NN * t, *Gdt0_t = Gdt0 + NN * t;
        for (int j = 0; j < N; j++) {
            for (int i = 0; i < N; i++) {
                const int r = p->map_ij[i + j * N];
                const int delta_tij = delta_t * (i == j);
                const num pre = phase / p->degen_ij;
                const num guii = Gutt_t[i + N * i], guij = Gut0_t[i + N * j];
                const num guji = Gu0t_t[j + N * i], gujj = Gu00[j + N * j];
                const num gdii = Gdtt_t[i + N * i], gdij = Gdt0_t[i + N * j];
                const num gdji = Gd0t_t[j + N * i], gdjj = Gd00[j + N * j];
                const long o = r + (long)num_ij * t;
#ifdef ORACLE_CPLX
                ACC(m, DQ_UE_gt0)[o] += 0.5 * pre * (guij * pu[j + i * N] + gdij * pd[j + i * N]);
                ACC(m, DQ_UE_gt0_u)[o] += pre * (guij * pu[j + i * N]);
                ACC(m, DQ_UE_gt0_d)[o] += pre * (gdij * pd[j + i * N]);
#else
                (void)pu; (void)pd;
                ACC(m, DQ_UE_gt0)[o] += 0.5 * pre * (guij + gdij);
                ACC(m, DQ_UE_gt0_u)[o] += pre * guij;
                ACC(m, DQ_UE_gt0_d)[o] += pre * gdij;
#endif
                const num x = delta_tij * (guii + gdii) - (guji * guij + gdji * gdij);
                ACC(m, DQ_UE_nn)[o] += pre * ((2. - guii - gdii) * (2. - gujj - gdjj) + x);
                ACC(m, DQ_UE_xx)[o] +=
                    0.25 * pre * (delta_tij * (guii + gdii) - (guji * gdij + gdji * guij));
                ACC(m, DQ_UE_zz)[o] += 0.25 * pre * ((gdii - guii) * (gdjj - gujj) + x);
                ACC(m, DQ_UE_pair_sw)[o] += pre * guij * gdij;
                if (p->meas_energy_corr) {
                    const num nuinuj = (1. - guii) * (1. - gujj) + (delta_tij - guji) * guij;
                    const num ndindj = (1. - gdii) * (1. - gdjj) + (delta_tij - gdji) * gdij;
                    ACC(m, DQ_UE_vv)[o] += pre * nuinuj * ndindj;
                    ACC(m, DQ_UE_vn)[o] += pre * (nuinuj * (1. - gdii) + (1. - guii) * ndindj);
                }
            }
        }
    }
    measure_uneqlt_ext(p, phase, Gu0t, Gutt, Gut0, Gd0t, Gdtt, Gdt0, m);
    if (!p->meas_bond_corr) return;

    for (int t = 0; t < L; t++) {
        const int delta_t = (t == 0);
        const num *Gu0t_t = Gu0t + NN * t, *Gutt_t = Gutt + NN * t, *Gut0_t = Gut0 + NN * t;
        const num *Gd0t_t = Gd0t + NN * t, *Gdtt_t = Gdtt + NN * t, *Gdt0_t = Gdt0 + NN * t;
        for (int c = 0; c < num_b; c++) {
            const int j0 = p->bonds[c], j1 = p->bonds[c + num_b];
#ifdef ORACLE_CPLX
            const num puj0j1 = pu[j0 + N * j1], puj1j0 = pu[j1 + N * j0];
            const num pdj0j1 = pd[j0 + N * j1], pdj1j0 = pd[j1 + N * j0];
#else
            const num puj0j1 = 1, puj1j0 = 1, pdj0j1 = 1, pdj1j0 = 1;
#endif
            for (int b = 0; b < num_b; b++) {
                const int i0 = p->bonds[b], i1 = p->bonds[b + num_b];
#ifdef ORACLE_CPLX
                const num pui0i1 = pu[i0 + N * i1], pui1i0 = pu[i1 + N * i0];
                const num pdi0i1 = pd[i0 + N * i1], pdi1i0 = pd[i1 + N * i0];
#else
                const num pui0i1 = 1, pui1i0 = 1, pdi0i1 = 1, pdi1i0 = 1;
#endif
                const int bb = p->map_bb[b + c * num_b];
                const num pre = phase / p->degen_bb;
                const num gui1i0 = Gutt_t[i1 + i0 * N], gui0i1 = Gutt_t[i0 + i1 * N];
                const num gui0j0 = Gut0_t[i0 + j0 * N], gui1j0 = Gut0_t[i1 + j0 * N];
                const num gui0j1 = Gut0_t[i0 + j1 * N], gui1j1 = Gut0_t[i1 + j1 * N];
                const num guj0i0 = Gu0t_t[j0 + i0 * N], guj1i0 = Gu0t_t[j1 + i0 * N];
                const num guj0i1 = Gu0t_t[j0 + i1 * N], guj1i1 = Gu0t_t[j1 + i1 * N];
                const num guj1j0 = Gu00[j1 + j0 * N], guj0j1 = Gu00[j0 + j1 * N];
                const num gdi1i0 = Gdtt_t[i1 + i0 * N], gdi0i1 = Gdtt_t[i0 + i1 * N];
                const num gdi0j0 = Gdt0_t[i0 + j0 * N], gdi1j0 = Gdt0_t[i1 + j0 * N];
                const num gdi0j1 = Gdt0_t[i0 + j1 * N], gdi1j1 = Gdt0_t[i1 + j1 * N];
                const num gdj0i0 = Gd0t_t[j0 + i0 * N], gdj1i0 = Gd0t_t[j1 + i0 * N];
                const num gdj0i1 = Gd0t_t[j0 + i1 * N], gdj1i1 = Gd0t_t[j1 + i1 * N];
                const num gdj1j0 = Gd00[j1 + j0 * N], gdj0j1 = Gd00[j0 + j1 * N];
                const int delta_i0j0 = delta_t * (i0 == j0), delta_i1j0 = delta_t * (i1 == j0);
                const int delta_i0j1 = delta_t * (i0 == j1), delta_i1j1 = delta_t * (i1 == j1);
                const num arr[8] = {
                    (delta_i0j1 - gdj1i0) * gdi1j0, (delta_i0j0 - gdj0i0) * gdi1j1,
                    (delta_i1j1 - gdj1i1) * gdi0j0, (delta_i1j0 - gdj0i1) * gdi0j1,
                    (delta_i0j1 - guj1i0) * gui1j0, (delta_i0j0 - guj0i0) * gui1j1,
                    (delta_i1j1 - guj1i1) * gui0j0, (delta_i1j0 - guj0i1) * gui0j1,
                };
                const num arrp[8] = {pdi0i1 * pdj0j1 * arr[0], pdi0i1 * pdj1j0 * arr[1],
                                     pdi1i0 * pdj0j1 * arr[2], pdi1i0 * pdj1j0 * arr[3],
                                     pui0i1 * puj0j1 * arr[4], pui0i1 * puj1j0 * arr[5],
                                     pui1i0 * puj0j1 * arr[6], pui1i0 * puj1j0 * arr[7]};
                const long o = bb + (long)num_bb * t;
                ACC(m, DQ_UE_pair_bb)[o] +=
                    0.5 * pre * (gui0j0 * gdi1j1 + gui1j0 * gdi0j1 + gui0j1 * gdi1j0 + gui1j1 * gdi0j0);
                const num x = arrp[0] + arrp[3] + arrp[4] + arrp[7];
                const num y = arrp[1] + arrp[2] + arrp[5] + arrp[6];
                const num tt[8] = {
                    pui1i0 * gui0i1, pui0i1 * gui1i0, pdi1i0 * gdi0i1, pdi0i1 * gdi1i0,
                    puj1j0 * guj0j1, puj0j1 * guj1j0, pdj1j0 * gdj0j1, pdj0j1 * gdj1j0,
                };
                ACC(m, DQ_UE_jj)[o] +=
                    pre * ((tt[0] - tt[1] + tt[2] - tt[3]) * (tt[4] - tt[5] + tt[6] - tt[7]) + x - y);
                ACC(m, DQ_UE_jsjs)[o] +=
                    pre * ((tt[0] - tt[1] - tt[2] + tt[3]) * (tt[4] - tt[5] - tt[6] + tt[7]) + x - y);
                ACC(m, DQ_UE_kk)[o] +=
                    pre * ((tt[0] + tt[1] + tt[2] + tt[3]) * (tt[4] + tt[5] + tt[6] + tt[7]) + x + y);
                ACC(m, DQ_UE_ksks)[o] +=
                    pre * ((tt[0] + tt[1] - tt[2] - tt[3]) * (tt[4] + tt[5] - tt[6] - tt[7]) + x + y);
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Per-chain work arrays                                                     */
/* ------------------------------------------------------------------------ */
typedef struct {
    int N, L, F, E, lwork, ue;
    num *B[2], *iB[2], *C[2], *g[2], *NN1[2], *NN2[2], *N1[2], *N2[2], *N3[2];
    int *pvt[2];
    num *G0t[2], *Gtt[2], *Gt0[2], *Gred[2], *tau[2], *Q[2], *work[2];
    int *site_order;
} ws_t;

static void ws_free(ws_t *w)
{
    for (int s = 0; s < 2; s++) {
        free(w->B[s]); free(w->iB[s]); free(w->C[s]); free(w->g[s]); free(w->NN1[s]);
        free(w->NN2[s]); free(w->N1[s]); free(w->N2[s]); free(w->N3[s]); free(w->pvt[s]);
        free(w->G0t[s]); free(w->Gtt[s]); free(w->Gt0[s]); free(w->Gred[s]); free(w->tau[s]);
        free(w->Q[s]); free(w->work[s]);
    }
    free(w->site_order);
}

static int ws_alloc(ws_t *w, const dqmc_params *p, int want_ue)
{
    memset(w, 0, sizeof(*w));
    const int N = p->N, L = p->L, F = p->F;
    const size_t NN = (size_t)N * N;
    w->N = N; w->L = L; w->F = F;
    w->E = 1 + (F - 1) / 2;
    w->ue = want_ue;
    w->lwork = lwork_for(N, w->E);
    for (int s = 0; s < 2; s++) {
        w->B[s] = zalloc(NN * L * sizeof(num));
        w->iB[s] = zalloc(NN * L * sizeof(num));
        w->C[s] = zalloc(NN * F * sizeof(num));
        w->g[s] = zalloc(NN * sizeof(num));
        w->NN1[s] = zalloc(NN * sizeof(num));
        w->NN2[s] = zalloc(NN * sizeof(num));
        w->N1[s] = zalloc(N * sizeof(num));
        w->N2[s] = zalloc(N * sizeof(num));
        w->N3[s] = zalloc(N * sizeof(num));
        w->pvt[s] = zalloc(N * sizeof(int) * 2);
        w->work[s] = zalloc((size_t)w->lwork * sizeof(num));
        if (!w->B[s] || !w->iB[s] || !w->C[s] || !w->work[s]) return -1;
        if (want_ue) {
            const size_t NE = (size_t)N * w->E;
            w->G0t[s] = zalloc(NN * L * sizeof(num));
            w->Gtt[s] = zalloc(NN * L * sizeof(num));
            w->Gt0[s] = zalloc(NN * L * sizeof(num));
            w->Gred[s] = zalloc(NE * NE * sizeof(num));
            w->tau[s] = zalloc(NE * sizeof(num) + 64);
            w->Q[s] = zalloc(4 * NN * sizeof(num));
            if (!w->G0t[s] || !w->Gtt[s] || !w->Gt0[s] || !w->Gred[s]) return -1;
        }
    }
    w->site_order = zalloc(N * sizeof(int));
    return 0;
}

static const num *expK_of(const dqmc_params *p, int s) { return (const num *)(s ? p->exp_Kd : p->exp_Ku); }
static const num *iexpK_of(const dqmc_params *p, int s) { return (const num *)(s ? p->inv_exp_Kd : p->inv_exp_Ku); }

/* build B_l, B_l^-1 (all l) and C_f (all f) for one spin from hs */
static void build_BC(const dqmc_params *p, const int *hs, ws_t *w, int s)
{
    const int N = p->N, L = p->L, n_matmul = p->n_matmul;
    const size_t NN = (size_t)N * N;
    for (int l = 0; l < L; l++) {
        calc_iB(N, iexpK_of(p, s), p->exp_lambda, hs + N * l, s, w->iB[s] + NN * l);
        calc_B(N, expK_of(p, s), p->exp_lambda, hs + N * l, s, w->B[s] + NN * l);
    }
    for (int f = 0; f < p->F; f++)
        mul_seq(N, L, f * n_matmul, ((f + 1) * n_matmul) % L, 1.0, w->B[s], w->C[s] + NN * f, N,
                w->NN1[s]);
}

static num eq_g_spin(const dqmc_params *p, ws_t *w, int s, int f)
{
    return calc_eq_g(f, p->N, p->F, 2, w->C[s], w->g[s], w->NN1[s], w->NN2[s], w->N1[s],
                     w->N2[s], w->N3[s], w->pvt[s], w->work[s], w->lwork);
}

static void half_wrap(int N, const num *G, const num *ehalf, const num *iehalf, num *tmp, num *out)
{
    gemm("N", "N", N, N, N, 1.0, G, N, ehalf, N, 0.0, tmp, N);
    gemm("N", "N", N, N, N, 1.0, iehalf, N, tmp, N, 0.0, out, N);
}

static void store_phase(num ph, double *out)
{
    out[0] = RE(ph);
#ifdef ORACLE_CPLX
    out[1] = cimag(ph);
#else
    out[1] = 0.0;
#endif
}

/* ------------------------------------------------------------------------ */
/* Public entry points (ctypes)                                              */
/* ------------------------------------------------------------------------ */

/* dqmc() of src/dqmc.c:84-459 for up to max_sweeps sweeps (no checkpoints,
 * no signals).  rng/sweep/hs are updated in place.  Returns the number of
 * sweeps run, negative on allocation failure.  n_acc (may be NULL) receives
 * the number of accepted flips. */
int SYM(dqmc)(const dqmc_params *p, uint64_t *rng, int *sweep, int *hs, dqmc_meas *meq,
              dqmc_meas *mue, int max_sweeps, int64_t *n_acc, double *phase_out)
{
    const int N = p->N, L = p->L, n_matmul = p->n_matmul, F = p->F;
    const size_t NN = (size_t)N * N;
    ws_t w;
    if (ws_alloc(&w, p, p->period_uneqlt > 0)) { ws_free(&w); return -1; }
    int64_t acc = 0;
    num phase;
    {
        num ph[2];
#pragma omp parallel for num_threads(2)
        for (int s = 0; s < 2; s++) {
            build_BC(p, hs, &w, s);
            ph[s] = eq_g_spin(p, &w, s, 0);
        }
        phase = ph[0] * ph[1];
    }
    int done = 0;
    for (; *sweep < p->n_sweep && done < max_sweeps; (*sweep)++, done++) {
        for (int l = 0; l < L; l++) {
            shuffle(rng, N, w.site_order);
            acc += update_delayed(N, p->n_delay, p->del, w.site_order, rng, hs + N * l, w.g[0],
                                  w.g[1], &phase, w.NN1[0], w.NN2[0], w.N1[0], w.NN1[1],
                                  w.NN2[1], w.N1[1]);
            const int f = l / n_matmul;
            const int recalc = ((l + 1) % n_matmul == 0);
            num ph[2];
#pragma omp parallel for num_threads(2)
            for (int s = 0; s < 2; s++) {
                num *Bl = w.B[s] + NN * l, *iBl = w.iB[s] + NN * l;
                calc_B(N, expK_of(p, s), p->exp_lambda, hs + N * l, s, Bl);
                if (!recalc || p->period_uneqlt > 0)
                    calc_iB(N, iexpK_of(p, s), p->exp_lambda, hs + N * l, s, iBl);
                if (recalc) {
                    mul_seq(N, L, f * n_matmul, ((f + 1) * n_matmul) % L, 1.0, w.B[s],
                            w.C[s] + NN * f, N, w.NN1[s]);
                    ph[s] = eq_g_spin(p, &w, s, (f + 1) % F);
                } else {
                    gemm("N", "N", N, N, N, 1.0, w.g[s], N, iBl, N, 0.0, w.NN1[s], N);
                    gemm("N", "N", N, N, N, 1.0, Bl, N, w.NN1[s], N, 0.0, w.g[s], N);
                }
            }
            if (recalc) phase = ph[0] * ph[1];

            if (*sweep >= p->n_sweep_warm && p->period_eqlt > 0 && (l + 1) % p->period_eqlt == 0) {
#pragma omp parallel for num_threads(2)
                for (int s = 0; s < 2; s++)
                    half_wrap(N, w.g[s], (const num *)(s ? p->exp_halfKd : p->exp_halfKu),
                              (const num *)(s ? p->inv_exp_halfKd : p->inv_exp_halfKu), w.NN1[s],
                              w.NN2[s]);
                measure_eqlt(p, phase, w.NN2[0], w.NN2[1], meq);
            }
        }
        if (*sweep >= p->n_sweep_warm && p->period_uneqlt > 0 && *sweep % p->period_uneqlt == 0) {
#pragma omp parallel for num_threads(2)
            for (int s = 0; s < 2; s++) {
                calc_ue_g(N, L, F, 2, w.B[s], w.iB[s], w.C[s], w.G0t[s], w.Gtt[s], w.Gt0[s],
                          w.Gred[s], w.tau[s], w.Q[s], w.work[s], w.lwork);
                const num *eh = (const num *)(s ? p->exp_halfKd : p->exp_halfKu);
                const num *ieh = (const num *)(s ? p->inv_exp_halfKd : p->inv_exp_halfKu);
                num *arrs[3] = {w.G0t[s], w.Gtt[s], w.Gt0[s]};
                for (int a = 0; a < 3; a++)
                    for (int l = 0; l < L; l++) {
                        gemm("N", "N", N, N, N, 1.0, arrs[a] + NN * l, N, eh, N, 0.0, w.NN1[s], N);
                        gemm("N", "N", N, N, N, 1.0, ieh, N, w.NN1[s], N, 0.0, arrs[a] + NN * l, N);
                    }
            }
            measure_uneqlt(p, phase, w.G0t[0], w.Gtt[0], w.Gt0[0], w.G0t[1], w.Gtt[1], w.Gt0[1],
                           mue);
        }
    }
    if (n_acc) *n_acc = acc;
    if (phase_out) store_phase(phase, phase_out);
    ws_free(&w);
    return done;
}

/* G of spin s from scratch at group boundary f for field hs: calc_eq_g(f, ...)
 * after building B and C as dqmc() does on entry.  Returns 0. */
int SYM(calc_eq_g)(const dqmc_params *p, const int *hs, int spin, int f, double *G, double *phase)
{
    ws_t w;
    if (ws_alloc(&w, p, 0)) { ws_free(&w); return -1; }
    build_BC(p, hs, &w, spin);
    num ph = eq_g_spin(p, &w, spin, f);
    memcpy(G, w.g[spin], (size_t)p->N * p->N * sizeof(num));
    store_phase(ph, phase);
    ws_free(&w);
    return 0;
}

/* G(0,t), G(t,t), G(t,0) for all t (calc_ue_g), before the half-wrap. */
int SYM(calc_ue_g)(const dqmc_params *p, const int *hs, int spin, double *G0t, double *Gtt,
                   double *Gt0)
{
    ws_t w;
    if (ws_alloc(&w, p, 1)) { ws_free(&w); return -1; }
    build_BC(p, hs, &w, spin);
    const size_t bytes = (size_t)p->N * p->N * p->L * sizeof(num);
    calc_ue_g(p->N, p->L, p->F, 2, w.B[spin], w.iB[spin], w.C[spin], w.G0t[spin], w.Gtt[spin],
              w.Gt0[spin], w.Gred[spin], w.tau[spin], w.Q[spin], w.work[spin], w.lwork);
    memcpy(G0t, w.G0t[spin], bytes);
    memcpy(Gtt, w.Gtt[spin], bytes);
    memcpy(Gt0, w.Gt0[spin], bytes);
    ws_free(&w);
    return 0;
}

/* One equal-time measurement on given (already half-wrapped) G's. */
void SYM(measure_eqlt)(const dqmc_params *p, const double *phase, const double *gu,
                       const double *gd, dqmc_meas *m)
{
    measure_eqlt(p, load_num(phase, 0), (const num *)gu, (const num *)gd, m);
}

/* One unequal-time measurement on given (already half-wrapped) arrays. */
void SYM(measure_uneqlt)(const dqmc_params *p, const double *phase, const double *Gu0t,
                         const double *Gutt, const double *Gut0, const double *Gd0t,
                         const double *Gdtt, const double *Gdt0, dqmc_meas *m)
{
    measure_uneqlt(p, load_num(phase, 0), (const num *)Gu0t, (const num *)Gutt,
                   (const num *)Gut0, (const num *)Gd0t, (const num *)Gdtt, (const num *)Gdt0, m);
}

/* rand_uint stream, for RNG pinning: writes n draws, advances rng. */
void SYM(rand_uint)(uint64_t *rng, int n, uint64_t *out)
{
    for (int i = 0; i < n; i++) out[i] = rand_uint(rng);
}

void SYM(shuffle)(uint64_t *rng, int n, int *a) { shuffle(rng, n, a); }

/* CHECK_G_ACC of src/dqmc.c:283-286: G at slice boundary l from the B's with a
 * QR after every single slice (calc_eq_g(l, N, L, 1, B, ...)) -- the most
 * stable variant of the stabilized product, used to referee accuracy. */
int SYM(calc_eq_g_fine)(const dqmc_params *p, const int *hs, int spin, int l, double *G,
                        double *phase)
{
    ws_t w;
    if (ws_alloc(&w, p, 0)) { ws_free(&w); return -1; }
    build_BC(p, hs, &w, spin);
    num ph = calc_eq_g(l, p->N, p->L, 1, w.B[spin], w.g[spin], w.NN1[spin], w.NN2[spin],
                       w.N1[spin], w.N2[spin], w.N3[spin], w.pvt[spin], w.work[spin], w.lwork);
    memcpy(G, w.g[spin], (size_t)p->N * p->N * sizeof(num));
    store_phase(ph, phase);
    ws_free(&w);
    return 0;
}
