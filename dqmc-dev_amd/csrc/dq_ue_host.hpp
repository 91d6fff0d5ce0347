// dq_ue_host.hpp -- unequal-time Green's functions on the GPU for every
// (chain, spin): calc_ue_g = calc_o -> bsofi -> expand_g (src/greens.c:247-494),
// the half-wrap of all 3L matrices (src/dqmc.c:415-451) and the contraction
// (src/meas.c:728-801).  Included by engine.hip.
//
// Every step is one batched launch over all (chain, spin) matrices.  The
// block-structured orthogonal factorization (bsofi) keeps the reference's
// block order: per block column a Householder QR of the 2N x N panel (with the
// explicit Q kept for the final right-multiplication by Q^H), the 2N x 2N QR
// of the last panel, the structured inversion of R (triangular solves for the
// diagonal blocks, GEMMs for the rest) and the right-multiplications by the
// Q_l^H -- each Q application is a GEMM on FP64 MFMA.
#pragma once
#include "dq_ue.hpp"

static long ue_stride(const dqmc_engine* e)
{
    const long N = e->N, E = 1 + (e->F - 1) / 2, NE = N * E;
    long s = NE * NE;
    if (2 * N * NE > s) s = 2 * N * NE;
    if (9 * N * N > s) s = 9 * N * N;
    return s;
}

static int ue_alloc(dqmc_engine* e)
{
    const long N = e->N, E = 1 + (e->F - 1) / 2, NE = N * E, B = 2L * e->nc;
    const long sU2 = ue_stride(e);
    const size_t es = e->esz;
    DALLOC(e->Gred, B * sU2 * es);
    DALLOC(e->Wa, B * sU2 * es);
    DALLOC(e->Wb, B * sU2 * es);
    DALLOC(e->Qb, B * (E > 1 ? E - 1 : 1) * 4 * N * N * es);
    DALLOC(e->taub, B * (NE + N) * es);
    DALLOC(e->T2, B * (long)e->L * N * N * es);
    DALLOC(e->Bthi, B * (long)e->L * N * N * es);
    return 0;
}

template <typename T>
static int launch_trsm_upper(dqmc_engine* e, int n, int ncols, const T* U, long sU, int ldu, T* X,
                             long sX, int ldx, int batch)
{
    dim3 grid((ncols + 4 * TS_NC - 1) / (4 * TS_NC), batch);
#define TRS(R) k_trsm_upper<T, R><<<grid, 256, 0, e->st>>>(n, ncols, U, sU, ldu, X, sX, ldx)
    if (n <= 64) TRS(1);
    else if (n <= 128) TRS(2);
    else if (n <= 256) TRS(4);
    else if (n <= 512) TRS(8);
    else if (n <= 1024) TRS(16);
    else return set_err(DQMC_ERR_UNSUPPORTED, "triangular block of order %d > 1024", n);
#undef TRS
    LCHK();
    return 0;
}

// X (n x n, ld) <- X^-1 in the upper triangle (xtrtri), lower part untouched
template <typename T>
static int trtri_upper(dqmc_engine* e, int n, T* X, long sX, int ldx, T* W, long sW, int B)
{
    k_set_identity<T><<<ew_grid(n, n, B), ew_threads(n), 0, e->st>>>(n, W, sW, n, 0, B);
    LCHK();
    RC(trsm_upper_blocked<T>(e, n, n, X, ldx, sX, W, n, sW, B));
    k_store_upper<T><<<ew_grid(n, n, B), ew_threads(n), 0, e->st>>>(n, W, sW, n, X, sX, ldx, B);
    LCHK();
    return 0;
}

// The slice products of the expansion, in the plain frame (B_l = expK
// diag(e_l), B_l^-1 = diag(e'_l) invexpK, scales fused into the GEMM) or in
// the half-wrapped frame (B~_l = H B_l H^-1 and its inverse, H = e^{-dtK/2}
// side factor `iehalf`, both precomputed for all l by ue_half_factors).
// C = alpha X op + beta C (right), C = alpha op X + beta C (left), C = B_l^-1.
template <typename T>
struct UeSteps {
    dqmc_engine* e;
    bool half;
    int N, L, B;
    long NN, LNN, LN;
    T one, zero;
    int right(T* C, const T* X, int l, bool inv, T alpha, T beta) const
    {
        if (half)
            return gemm<T>(e, OP_N, OP_N, N, N, N, alpha, X, N, LNN,
                           dp<T>(inv ? e->Bthi : e->T2) + l * NN, N, LNN, beta, C, N, LNN, B);
        if (inv)
            return gemm<T>(e, OP_N, OP_N, N, N, N, alpha, X, N, LNN, dp<T>(e->iexpK), N, NN, beta, C,
                           N, LNN, B, nullptr, 0, e->iscale + (long)N * l, LN);
        return gemm<T>(e, OP_N, OP_N, N, N, N, alpha, X, N, LNN, dp<T>(e->expK), N, NN, beta, C, N,
                       LNN, B, nullptr, 0, nullptr, 0, e->escale + (long)N * l, LN);
    }
    // X, C with the stride of T1 (one matrix per b) when `tmp`
    int left(T* C, long sC, const T* X, long sX, int l, bool inv, T alpha, T beta) const
    {
        if (half)
            return gemm<T>(e, OP_N, OP_N, N, N, N, alpha, dp<T>(inv ? e->Bthi : e->T2) + l * NN, N,
                           LNN, X, N, sX, beta, C, N, sC, B);
        if (inv)
            return gemm<T>(e, OP_N, OP_N, N, N, N, alpha, dp<T>(e->iexpK), N, NN, X, N, sX, beta, C,
                           N, sC, B, e->iscale + (long)N * l, LN);
        return gemm<T>(e, OP_N, OP_N, N, N, N, alpha, dp<T>(e->expK), N, NN, X, N, sX, beta, C, N,
                       sC, B, nullptr, 0, e->escale + (long)N * l, LN);
    }
    int right_tmp(T* C, const T* X, int l, bool inv) const  // C (stride NN) = X op
    {
        if (half)
            return gemm<T>(e, OP_N, OP_N, N, N, N, one, X, N, LNN,
                           dp<T>(inv ? e->Bthi : e->T2) + l * NN, N, LNN, zero, C, N, NN, B);
        if (inv)
            return gemm<T>(e, OP_N, OP_N, N, N, N, one, X, N, LNN, dp<T>(e->iexpK), N, NN, zero, C,
                           N, NN, B, nullptr, 0, e->iscale + (long)N * l, LN);
        return gemm<T>(e, OP_N, OP_N, N, N, N, one, X, N, LNN, dp<T>(e->expK), N, NN, zero, C, N,
                       NN, B, nullptr, 0, nullptr, 0, e->escale + (long)N * l, LN);
    }
    int inverse(T* C, int l) const  // C (stride LNN) = B_l^-1
    {
        if (half)
            k_copy_mn<T><<<ew_grid(N, N, B), ew_threads(N), 0, e->st>>>(N, N, dp<T>(e->Bthi) + l * NN, LNN, N,
                                                           one, C, LNN, N, B);
        else
            k_rowscale<T><<<ew_grid(N, N, B), ew_threads(N), 0, e->st>>>(N, dp<T>(e->iexpK), NN, 1,
                                                          e->iscale + (long)N * l, LN, C, LNN, B);
        LCHK();
        return 0;
    }
};

// B~_l = H expK diag(e_l) H^-1 into T2 and B~_l^-1 = H diag(e'_l) invexpK H^-1
// into Bthi for all (b, l), H = iehalf, H^-1 = ehalf: two batched GEMMs over
// all slices after the per-b products H expK (Y2) and invexpK H^-1 (Z).
template <typename T>
static int ue_half_factors(dqmc_engine* e)
{
    const int N = e->N, L = e->L, B = 2 * e->nc;
    const long NN = (long)N * N;
    const T one = from_re<T>(1.0), zero = from_re<T>(0.0);
    T* M1 = dp<T>(e->Y2);
    T* M2 = dp<T>(e->Z);
    RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, dp<T>(e->iehalf), N, NN, dp<T>(e->expK), N, NN, zero,
               M1, N, NN, B));
    RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, dp<T>(e->iexpK), N, NN, dp<T>(e->ehalf), N, NN, zero,
               M2, N, NN, B));
    e->gemm_adiv = L;
    e->gemm_bdiv = L;
    RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, M1, N, NN, dp<T>(e->ehalf), N, NN, zero, dp<T>(e->T2),
               N, NN, B * L, nullptr, 0, e->escale, N));
    e->gemm_adiv = L;
    e->gemm_bdiv = L;
    RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, dp<T>(e->iehalf), N, NN, M2, N, NN, zero,
               dp<T>(e->Bthi), N, NN, B * L, nullptr, 0, e->iscale, N));
    if (e->ragged) {
        // slices past a chain's own L are identity slices: B~_l = B~_l^-1 = I.
        // They sit after the chain's last slice, so G(t,0), G(0,t), G(t,t) of
        // its own slices t < L_c are those of the chain alone.
        k_ue_identity_slices<T><<<nblk((long)B * L * NN), 256, 0, e->st>>>(
            N, L, dp<T>(e->T2), dp<T>(e->Bthi), e->d_Lc, B);
        LCHK();
    }
    return 0;
}

// X[b] (one slice of an [b][L][NN] array) <- H X H^-1 through T1
template <typename T>
static int half_wrap_slice(dqmc_engine* e, T* X)
{
    const int N = e->N, B = 2 * e->nc;
    const long NN = (long)N * N, LNN = (long)e->L * NN;
    const T one = from_re<T>(1.0), zero = from_re<T>(0.0);
    T* T1 = dp<T>(e->T1);
    RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, X, N, LNN, dp<T>(e->ehalf), N, NN, zero, T1, N, NN, B));
    RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, dp<T>(e->iehalf), N, NN, T1, N, NN, zero, X, N, LNN, B));
    return 0;
}

// expand_g (src/greens.c:339-469) with nm = (L/F)*2 slices per block: the
// block-boundary values from bsofi (Gred) are spread over all slices by
// products with B_l and B_l^-1.  half: the boundary values are half-wrapped
// first and the products are the half-wrapped ones, so that G0t, Gtt, Gt0 come
// out half-wrapped (src/dqmc.c:415-451) without a pass over all slices; the
// identity terms are unchanged by the half-wrap.
template <typename T>
static int ue_expand(dqmc_engine* e, T* Gred, long sU2, bool half)
{
    const int N = e->N, L = e->L, F = e->F, B = 2 * e->nc;
    const int E = 1 + (F - 1) / 2, NE = N * E, NL = NE;
    const long NN = (long)N * N, LNN = (long)L * NN, LN = (long)L * N;
    const T one = from_re<T>(1.0), mone = from_re<T>(-1.0), zero = from_re<T>(0.0);
    auto BLK = [&](int i, int j) { return Gred + (long)N * i + (long)NL * N * j; };
    T* G0t = dp<T>(e->G0t);
    T* Gtt = dp<T>(e->Gtt);
    T* Gt0 = dp<T>(e->Gt0);
    T* tmp = dp<T>(e->T1);
    const UeSteps<T> st{e, half, N, L, B, NN, LNN, LN, one, zero};
    if (half) {
        ProfSec ps(e, PS_half_wrap);
        RC(ue_half_factors<T>(e));
    }
    ProfSec ps(e, PS_expand_g);
    const int nm = (L / F) * 2;
    const int n_left = (nm - 1) / 2, n_right = nm / 2, n_up = n_left, n_down = n_right;
    const int rstop_last = ((E - 1) * nm + L) / 2;
    const int lstop_first = (rstop_last + 1) % L;
    const int dstop_last = rstop_last, ustop_first = lstop_first;
    auto cpy = [&](const T* src, T* dst) -> int {
        k_copy_mn<T><<<ew_grid(N, N, B), ew_threads(N), 0, e->st>>>(N, N, src, sU2, NE, one, dst, LNN, N, B);
        LCHK();
        if (half) RC(half_wrap_slice<T>(e, dst));
        return 0;
    };
    for (int f = 0; f < E; f++) RC(cpy(BLK(0, f), G0t + (long)f * nm * NN));
    for (int f = 0; f < E; f++) {
        const int l = f * nm;
        const int lstop = (f == 0) ? lstop_first : l - n_left;
        const int rstop = (f == E - 1) ? rstop_last : l + n_right;
        for (int m = l; m != lstop;) {  // G(0, next) = alpha G(0, m) B_next
            const int next = (m - 1 + L) % L;
            const T alpha = (m == 0) ? mone : one;
            RC(st.right(G0t + next * NN, G0t + m * NN, next, false, alpha, zero));
            m = next;
        }
        for (int m = l; m != rstop;) {  // G(0, next) = alpha G(0, m) B_m^-1 (+ beta B_m^-1)
            const int next = (m + 1) % L;
            const T alpha = (next == 0) ? mone : one;
            const T beta = (m == 0) ? from_re<T>(-re(alpha)) : zero;
            if (m == 0) RC(st.inverse(G0t + next * NN, m));
            RC(st.right(G0t + next * NN, G0t + m * NN, m, true, alpha, beta));
            m = next;
        }
    }
    for (int q = 0; q < E; q++) RC(cpy(BLK(q, q), Gtt + (long)q * nm * NN));
    for (int q = 0; q < E; q++) {
        const int k = q * nm;
        const int ustop = (q == 0) ? ustop_first : k - n_up;
        const int dstop = (q == E - 1) ? dstop_last : k + n_down;
        for (int m = k; m != ustop;) {  // G(next, next) = B_next^-1 G(m, m) B_next
            const int next = (m - 1 + L) % L;
            RC(st.right_tmp(tmp, Gtt + m * NN, next, false));
            RC(st.left(Gtt + next * NN, LNN, tmp, NN, next, true, one, zero));
            m = next;
        }
        for (int m = k; m != dstop;) {  // G(next, next) = B_m G(m, m) B_m^-1
            const int next = (m + 1) % L;
            RC(st.right_tmp(tmp, Gtt + m * NN, m, true));
            RC(st.left(Gtt + next * NN, LNN, tmp, NN, m, false, one, zero));
            m = next;
        }
    }
    for (int q = 0; q < E; q++) RC(cpy(BLK(q, 0), Gt0 + (long)q * nm * NN));
    for (int q = 0; q < E; q++) {
        const int k = q * nm;
        const int ustop = (q == 0) ? ustop_first : k - n_up;
        const int dstop = (q == E - 1) ? dstop_last : k + n_down;
        for (int m = k; m != ustop;) {  // G(next, 0) = alpha B_next^-1 G(m, 0) (+ beta B_next^-1)
            const int next = (m - 1 + L) % L;
            const T alpha = (m == 0) ? mone : one;
            const T beta = (m == 0) ? from_re<T>(-re(alpha)) : zero;
            if (m == 0) RC(st.inverse(Gt0 + next * NN, next));
            RC(st.left(Gt0 + next * NN, LNN, Gt0 + m * NN, LNN, next, true, alpha, beta));
            m = next;
        }
        for (int m = k; m != dstop;) {  // G(next, 0) = alpha B_m G(m, 0) (+ I at next = 0)
            const int next = (m + 1) % L;
            const T alpha = (next == 0) ? mone : one;
            RC(st.left(Gt0 + next * NN, LNN, Gt0 + m * NN, LNN, m, false, alpha, zero));
            if (next == 0) {
                k_set_identity<T><<<ew_grid(N, N, B), ew_threads(N), 0, e->st>>>(N, Gt0 + next * NN, LNN, N, 1, B);
                LCHK();
            }
            m = next;
        }
    }
    return 0;
}

// G0t, Gtt, Gt0 <- calc_ue_g for all b; half: half-wrapped (as measured)
template <typename T>
static int ue_calc(dqmc_engine* e, bool half)
{
    const int N = e->N, F = e->F, B = 2 * e->nc;
    const int E = 1 + (F - 1) / 2, NE = N * E, NL = NE, N2 = 2 * N;
    const long NN = (long)N * N;
    const long sU2 = ue_stride(e);
    T* Gred = dp<T>(e->Gred);
    T* Wa = dp<T>(e->Wa);
    T* Wb = dp<T>(e->Wb);
    T* tauB = dp<T>(e->taub);
    T* Qb = dp<T>(e->Qb);
    const long sQb = (long)(E > 1 ? E - 1 : 1) * 4 * NN;
    const T one = from_re<T>(1.0), mone = from_re<T>(-1.0), zero = from_re<T>(0.0);
    auto BLK = [&](int i, int j) { return Gred + (long)N * i + (long)NL * N * j; };

    // calc_o (src/greens.c:247-264)
    ProfSec ps_o(e, PS_calc_o);
    {
        const long nd = (long)B * sU2 * (long)(sizeof(T) / 8);
        k_zero<<<nblk(nd), 256, 0, e->st>>>(nd, (double*)Gred);
        LCHK();
    }
    for (int q = 0; q < E - 1; q++)
        RC(mul_seq_C<T>(e, q * 2, (q + 1) * 2, mone, BLK(q + 1, q), NE, sU2));
    RC(mul_seq_C<T>(e, (E - 1) * 2, 0, one, BLK(0, E - 1), NE, sU2));
    k_set_identity<T><<<ew_grid(NE, NE, B), ew_threads(NE), 0, e->st>>>(NE, Gred, sU2, NE, 1, B);
    LCHK();
    ps_o.stop();

    // bsofi (src/greens.c:266-337)
    {
    ProfSec ps_b(e, PS_bsofi);
    if (E == 1) {
        k_set_identity<T><<<ew_grid(N, N, B), ew_threads(N), 0, e->st>>>(N, Wa, sU2, N, 0, B);
        LCHK();
        k_getrf<T><<<B, 1024, 0, e->st>>>(N, Gred, Wa, sU2, e->ipiv, N);
        LCHK();
        const int wpb = 4;
        dim3 grid((N + wpb * LU_NC - 1) / (wpb * LU_NC), B);
#define GRS(R)                                                                                 \
    k_getrs_phase<T, R><<<grid, 256, 0, e->st>>>(N, Gred, Wa, sU2, e->ipiv, N, nullptr, 0,       \
                                                 nullptr, 0, nullptr, 0, nullptr)
        if (N <= 64) GRS(1);
        else if (N <= 128) GRS(2);
        else if (N <= 256) GRS(4);
        else if (N <= 512) GRS(8);
        else GRS(16);
#undef GRS
        LCHK();
        k_copy_mn<T><<<ew_grid(N, N, B), ew_threads(N), 0, e->st>>>(N, N, Wa, sU2, N, one, Gred, sU2, NE, B);
        LCHK();
    } else {
        for (int l = 0; l < E - 2; l++) {
            T* Ql = Qb + (long)l * 4 * NN;
            RC(qr_blocked<T>(e, N2, N, BLK(l, l), NL, sU2, tauB + (long)N * l, NE, Ql, N2, sQb,
                             nullptr, B));
            const int cols[2] = {l + 1, E - 1};
            for (int t = 0; t < 2; t++) {
                T* P = BLK(l, cols[t]);
                RC(gemm<T>(e, OP_C, OP_N, N2, N, N2, one, Ql, N2, sQb, P, NL, sU2, zero, Wa, N2,
                           sU2, B));
                k_copy_mn<T><<<ew_grid(N2, N, B), ew_threads(N2), 0, e->st>>>(N2, N, Wa, sU2, N2, one, P,
                                                                       sU2, NL, B);
                LCHK();
            }
        }
        T* Qlast = Qb + (long)(E - 2) * 4 * NN;
        RC(qr_blocked<T>(e, N2, N2, BLK(E - 2, E - 2), NL, sU2, tauB + (long)N * (E - 2), NE,
                         Qlast, N2, sQb, nullptr, B));
        // invert R
        if (E <= 2) {
            RC(trtri_upper<T>(e, NL, Gred, sU2, NL, Wa, sU2, B));
        } else {
            RC(trtri_upper<T>(e, 3 * N, BLK(E - 3, E - 3), sU2, NL, Wa, sU2, B));
            if (E > 3) {
                const int m3 = N * (E - 3);
                k_copy_upper<T><<<ew_grid(N, N, B), ew_threads(N), 0, e->st>>>(N, BLK(E - 1, E - 1), sU2, NL, Wb,
                                                                sU2, N, B);
                LCHK();
                RC(gemm<T>(e, OP_N, OP_N, m3, N, N, one, BLK(0, E - 1), NL, sU2, Wb, N, sU2, zero,
                           Wa, m3, sU2, B));
                k_copy_mn<T><<<ew_grid(m3, N, B), ew_threads(m3), 0, e->st>>>(m3, N, Wa, sU2, m3, one,
                                                                       BLK(0, E - 1), sU2, NL, B);
                LCHK();
                for (int l = E - 4; l >= 0; l--) {
                    RC(trtri_upper<T>(e, N, BLK(l, l), sU2, NL, Wa, sU2, B));
                    k_copy_upper<T><<<ew_grid(N, N, B), ew_threads(N), 0, e->st>>>(N, BLK(l, l), sU2, NL, Wb,
                                                                    sU2, N, B);
                    LCHK();
                    const int cols[2] = {E - 1, l + 1};
                    for (int t = 0; t < 2; t++) {
                        RC(gemm<T>(e, OP_N, OP_N, N, N, N, mone, Wb, N, sU2, BLK(l, cols[t]), NL,
                                   sU2, zero, Wa, N, sU2, B));
                        k_copy_mn<T><<<ew_grid(N, N, B), ew_threads(N), 0, e->st>>>(N, N, Wa, sU2, N, one,
                                                                     BLK(l, cols[t]), sU2, NL, B);
                        LCHK();
                    }
                    RC(gemm<T>(e, OP_N, OP_N, N, N * (E - l - 2), N, one, BLK(l, l + 1), NL, sU2,
                               BLK(l + 1, l + 2), NL, sU2, one, BLK(l, l + 2), NL, sU2, B));
                    k_copy_upper<T><<<ew_grid(N, N, B), ew_threads(N), 0, e->st>>>(N, BLK(l + 1, l + 1), sU2, NL,
                                                                    Wb, sU2, N, B);
                    LCHK();
                    RC(gemm<T>(e, OP_N, OP_N, N, N, N, one, BLK(l, l + 1), NL, sU2, Wb, N, sU2,
                               zero, Wa, N, sU2, B));
                    k_copy_mn<T><<<ew_grid(N, N, B), ew_threads(N), 0, e->st>>>(N, N, Wa, sU2, N, one,
                                                                 BLK(l, l + 1), sU2, NL, B);
                    LCHK();
                }
            }
        }
        // multiply by Q^H from the right
        k_zero_lower<T><<<ew_grid(N2, N2, B), ew_threads(N2), 0, e->st>>>(N2, N2, BLK(E - 2, E - 2), sU2, NL, B);
        LCHK();
        RC(gemm<T>(e, OP_N, OP_C, NL, N2, N2, one, BLK(0, E - 2), NL, sU2, Qlast, N2, sQb, zero,
                   Wa, NL, sU2, B));
        k_copy_mn<T><<<ew_grid(NL, N2, B), ew_threads(NL), 0, e->st>>>(NL, N2, Wa, sU2, NL, one,
                                                                BLK(0, E - 2), sU2, NL, B);
        LCHK();
        for (int l = E - 3; l >= 0; l--) {
            T* Ql = Qb + (long)l * 4 * NN;
            k_zero_lower<T><<<ew_grid(N2, N, B), ew_threads(N2), 0, e->st>>>(N2, N, BLK(l, l), sU2, NL, B);
            LCHK();
            RC(gemm<T>(e, OP_N, OP_C, NL, N2, N2, one, BLK(0, l), NL, sU2, Ql, N2, sQb, zero, Wa,
                       NL, sU2, B));
            k_copy_mn<T><<<ew_grid(NL, N2, B), ew_threads(NL), 0, e->st>>>(NL, N2, Wa, sU2, NL, one,
                                                                    BLK(0, l), sU2, NL, B);
            LCHK();
        }
    }
    }

    return ue_expand<T>(e, Gred, sU2, half);
}

template <typename T>
static int ue_measure(dqmc_engine* e)
{
    RC(ue_calc<T>(e, true));
    ProfSec ps(e, PS_meas_uneq);
    MeasArgs<T> ma = meas_args<T>(e, 1);
    ma.G0t = dp<T>(e->G0t);
    ma.Gtt = dp<T>(e->Gtt);
    ma.Gt0 = dp<T>(e->Gt0);
    ma.sU = (long)e->L * e->N * e->N;
    const int L = e->L, nc = e->nc;
    if (e->ij_percol) {
        const int thr = (e->num_ij + 63) / 64 * 64;
        k_meas_ue_strip<T, 8><<<dim3(L, nc), thr, 0, e->st>>>(ma);
    } else {
        k_meas_ue<T><<<dim3((e->num_ij + 255) / 256, L, nc), 256, 0, e->st>>>(ma);
    }
    LCHK();
    // the pair kernels read the per-bond tables of the maps they walk
    if (((e->bond || e->thermal || e->pbbonly) && e->num_bb > 0 && !e->btab) ||
        (e->energy && e->num_bs > 0 && !e->btab))
        return set_err(DQMC_ERR_UNSUPPORTED, "bond-pair measurements without bonds");
    if ((e->thermal || e->twobond) && (e->num_bb2 > 0 || e->num_b2b > 0 || e->num_b2b2 > 0) &&
        (!e->btab || !e->b2tab))
        return set_err(DQMC_ERR_UNSUPPORTED, "two-hop bond measurements need bonds and two-hop bonds");
    ProfSec ps_sub(e, PS_meas_uneq_sub);
    if (e->bond || e->thermal || e->pbbonly || e->twobond || e->energy) {
        // G(0,t)^T for the bond-pair kernels, into T2 (the expansion's
        // half-wrapped B_l, free until the next measurement rebuilds it)
        const int tn = (e->N + 31) / 32;
        const long count = 2L * nc * L;
        k_transpose_batch<T><<<(unsigned)(count * tn * tn), 256, 0, e->st>>>(
            e->N, tn, dp<T>(e->G0t), dp<T>(e->T2), (long)e->N * e->N);
        LCHK();
        ma.G0tT = dp<T>(e->T2);
        // per-(chain, slice, bond) hop terms, densities and phases
        if (e->btab) {
            ma.btab = dp<T>(e->btab);
            ma.bph = dp<T>(e->bph);
            k_bond_tab<T><<<dim3((e->num_b + 255) / 256, L, nc), 256, 0, e->st>>>(
                ma, dp<T>(e->btab), dp<T>(e->bph));
            LCHK();
        }
        if (e->b2tab) {
            ma.b2tab = dp<T>(e->b2tab);
            k_bond2_tab<T><<<dim3((e->num_b2 + 255) / 256, L, nc), 256, 0, e->st>>>(
                ma, dp<T>(e->b2tab));
            LCHK();
        }
    }
    if ((e->bond || e->thermal || e->pbbonly) && e->num_b > 0 && e->num_bb > 0) {
        k_meas_ue_pairs<T, PK_BB><<<dim3((e->num_bb + 255) / 256, L, nc), 256, 0, e->st>>>(ma);
        LCHK();
    }
    if (e->thermal || e->twobond) {
        k_meas_ue_pairs<T, PK_B2B><<<dim3((e->num_bb2 + 255) / 256, L, nc), 256, 0, e->st>>>(ma);
        LCHK();
        k_meas_ue_pairs<T, PK_BB2><<<dim3((e->num_b2b + 255) / 256, L, nc), 256, 0, e->st>>>(ma);
        LCHK();
    }
    if (e->twobond) {
        k_meas_ue_pairs<T, PK_B2B2><<<dim3((e->num_b2b2 + 255) / 256, L, nc), 256, 0, e->st>>>(ma);
        LCHK();
    }
    if (e->energy) {
        k_meas_ue_kv<T><<<dim3((e->num_bs + 255) / 256, L, nc), 256, 0, e->st>>>(ma);
        LCHK();
    }
    if (e->gensus) {
        const long nij = e->num_ij;
        k_gensus_ab<T><<<dim3((nij + 255) / 256, L, nc), 256, 0, e->st>>>(ma, 1);
        LCHK();
        k_gensus_outer<T><<<dim3((nij * nij + 255) / 256, L, nc), 256, 0, e->st>>>(ma, 1);
        LCHK();
    }
    if constexpr (!is_cplx<T>::value) {
        if (e->nematic) {
            k_meas_ue_nematic<<<dim3((e->num_bb + 255) / 256, L, nc), 256, 0, e->st>>>(ma);
            LCHK();
        }
    }
    return 0;
}

template <typename T>
static int ue_hook(dqmc_engine* e, int c, int spin, double* G0t, double* Gtt, double* Gt0)
{
    const int N = e->N;
    k_scales<<<nblk((long)e->nc * e->L * N), 256, 0, e->st>>>(N, e->L, e->nc, e->hs, e->exp_lambda,
                                                               e->escale, e->iscale);
    LCHK();
    RC(build_C<T>(e, 0, e->F));
    RC(ue_calc<T>(e, false));
    HIPCHK(hipStreamSynchronize(e->st));
    const size_t bytes = (size_t)e->L * N * N * sizeof(T);
    const int b = 2 * c + spin;
    HIPCHK(hipMemcpy(G0t, (char*)e->G0t + b * bytes, bytes, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(Gtt, (char*)e->Gtt + b * bytes, bytes, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(Gt0, (char*)e->Gt0 + b * bytes, bytes, hipMemcpyDeviceToHost));
    e->stale = true;
    return 0;
}
