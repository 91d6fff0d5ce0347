/* dq_bucket.c -- padding of files of uneven size into one engine (see
 * dq_bucket.h).  Matrices are column-major (element (r, c) at r + N c), as the
 * .h5 datasets the reference reads. */
#include "dq_bucket.h"

#include <stdlib.h>
#include <string.h>

static const int BUCKETS[] = {64, 121, 196, 289, 400, 512};

int dq_bucket_of(const dqmc_params *p)
{
    if (p->meas_bond_corr || p->meas_thermal || p->meas_2bond_corr || p->meas_energy_corr ||
        p->meas_nematic_corr || p->meas_pair_bb_only || p->meas_chiral || p->meas_local_JQ ||
        p->meas_gen_suscept)
        return 0;
    for (size_t k = 0; k < sizeof BUCKETS / sizeof BUCKETS[0]; k++)
        if (p->N <= BUCKETS[k]) return BUCKETS[k];
    return 0;
}

int dq_bucket_compatible(const dqmc_params *a, const dqmc_params *b)
{
    const int ka = dq_bucket_of(a);
    return ka != 0 && ka == dq_bucket_of(b) && a->n_matmul == b->n_matmul &&
           a->period_eqlt == b->period_eqlt && a->period_uneqlt == b->period_uneqlt &&
           a->is_complex == b->is_complex &&
           a->L % a->n_matmul == 0 && b->L % b->n_matmul == 0;
}

static void *own(dq_pad *d, size_t bytes)
{
    if (d->n_own >= (int)(sizeof d->own / sizeof d->own[0])) return NULL;
    void *q = calloc(1, bytes ? bytes : 1);
    if (q) d->own[d->n_own++] = q;
    return q;
}

/* N x N `num` matrix (es doubles per element) into N_pad x N_pad, the rest
 * zero with `diag` on the padding diagonal */
static double *pad_mat(dq_pad *d, const double *a, int N, int N_pad, int es, double diag)
{
    double *o = own(d, (size_t)N_pad * N_pad * es * sizeof(double));
    if (!o) return NULL;
    for (int c = 0; c < N; c++)
        memcpy(o + (size_t)N_pad * c * es, a + (size_t)N * c * es, (size_t)N * es * sizeof(double));
    for (int r = N; r < N_pad; r++) o[((size_t)r + (size_t)N_pad * r) * es] = diag;
    return o;
}

int dq_pad_build(const dq_sim *s, int N_pad, int L_pad, int num_i_pad, int num_ij_pad,
                 dq_pad *d)
{
    memset(d, 0, sizeof *d);
    const dqmc_params *p = &s->p;
    const int N = p->N, es = p->is_complex ? 2 : 1;
    if (N > N_pad || p->L > L_pad || L_pad % p->n_matmul || p->num_i > num_i_pad ||
        p->num_ij > num_ij_pad)
        return -1;
    d->N0 = N;
    d->L0 = p->L;
    d->p = *p;
    dqmc_params *q = &d->p;
    q->N = N_pad;
    q->L = L_pad;
    q->F = L_pad / p->n_matmul;
    q->num_i = num_i_pad;
    q->num_ij = num_ij_pad;
    q->num_plaq_accum = q->num_b_accum = q->num_b2_accum = 0;
    q->num_plaq = q->num_b = q->num_b2 = q->num_bs = q->num_bb = 0;
    q->num_bb2 = q->num_b2b = q->num_b2b2 = 0;
    const double *src[8] = {p->exp_Ku, p->exp_Kd, p->inv_exp_Ku, p->inv_exp_Kd,
                            p->exp_halfKu, p->exp_halfKd, p->inv_exp_halfKu, p->inv_exp_halfKd};
    const double *dst[8];
    for (int k = 0; k < 8; k++)
        if (!(dst[k] = pad_mat(d, src[k], N, N_pad, es, 1.0))) return -2;
    q->exp_Ku = dst[0]; q->exp_Kd = dst[1]; q->inv_exp_Ku = dst[2]; q->inv_exp_Kd = dst[3];
    q->exp_halfKu = dst[4]; q->exp_halfKd = dst[5];
    q->inv_exp_halfKu = dst[6]; q->inv_exp_halfKd = dst[7];
    if (p->peierlsu) {  /* phases of padding pairs: 1 */
        double *pu = own(d, (size_t)N_pad * N_pad * es * sizeof(double));
        double *pd = own(d, (size_t)N_pad * N_pad * es * sizeof(double));
        if (!pu || !pd) return -2;
        for (int c = 0; c < N_pad; c++)
            for (int r = 0; r < N_pad; r++) {
                const size_t o = ((size_t)r + (size_t)N_pad * c) * es;
                const int in = r < N && c < N;
                const size_t i = ((size_t)r + (size_t)N * c) * es;
                pu[o] = in ? p->peierlsu[i] : 1.0;
                pd[o] = in ? p->peierlsd[i] : 1.0;
                if (es == 2) {
                    pu[o + 1] = in ? p->peierlsu[i + 1] : 0.0;
                    pd[o + 1] = in ? p->peierlsd[i + 1] : 0.0;
                }
            }
        q->peierlsu = pu;
        q->peierlsd = pd;
    }
    double *el = own(d, 2 * (size_t)N_pad * sizeof(double));
    double *dl = own(d, 2 * (size_t)N_pad * sizeof(double));
    int32_t *mi = own(d, (size_t)N_pad * sizeof(int32_t));
    int32_t *mij = own(d, (size_t)N_pad * N_pad * sizeof(int32_t));
    d->hs = own(d, (size_t)L_pad * N_pad * sizeof(int32_t));
    if (!el || !dl || !mi || !mij || !d->hs) return -2;
    for (int r = 0; r < 2; r++)
        for (int i = 0; i < N_pad; i++) {
            el[r * N_pad + i] = i < N ? p->exp_lambda[r * N + i] : 1.0;
            dl[r * N_pad + i] = i < N ? p->del[r * N + i] : 0.0;
        }
    for (int i = 0; i < N_pad; i++) mi[i] = i < N ? p->map_i[i] : -1;
    for (int c = 0; c < N_pad; c++)
        for (int r = 0; r < N_pad; r++)
            mij[r + (size_t)N_pad * c] = (r < N && c < N) ? p->map_ij[r + (size_t)N * c] : -1;
    q->exp_lambda = el;
    q->del = dl;
    q->map_i = mi;
    q->map_ij = mij;
    /* accumulators of the padded bin (and, unequal-time, slice) counts */
    d->meq = s->meq;
    d->mue = s->mue;
    for (int k = 0; k < DQMC_MAX_ACC; k++) {
        const long len = k < DQ_EQ_COUNT ? dqmc_acc_len(q, 0, k) : 0;
        d->meq.acc[k] = NULL;
        if (len > 0 && s->meq.acc[k]) {
            d->meq.acc[k] = own(d, (size_t)len * es * sizeof(double));
            if (!d->meq.acc[k]) return -2;
        }
        const long lue = (p->period_uneqlt > 0 && k < DQ_UE_COUNT) ? dqmc_acc_len(q, 1, k) : 0;
        d->mue.acc[k] = NULL;
        if (lue > 0 && s->mue.acc[k]) {
            d->mue.acc[k] = own(d, (size_t)lue * es * sizeof(double));
            if (!d->mue.acc[k]) return -2;
        }
    }
    return 0;
}

/* unequal-time accumulator k: [L][n] slices of the file <-> [L_pad][n_pad] of
 * the padded engine (the file's slices and bins first, zeros elsewhere) */
static void ue_copy(const dq_pad *d, const dqmc_params *p0, int k, double *pad, double *file,
                    int to_pad)
{
    const int es = d->p.is_complex ? 2 : 1;
    const long lp = dqmc_acc_len(&d->p, 1, k), l0 = dqmc_acc_len(p0, 1, k);
    const long np = lp / d->p.L, n0 = l0 / d->L0;
    if (to_pad) memset(pad, 0, (size_t)lp * es * sizeof(double));
    for (int t = 0; t < d->L0; t++) {
        double *a = pad + (size_t)np * t * es, *b = file + (size_t)n0 * t * es;
        if (to_pad)
            memcpy(a, b, (size_t)n0 * es * sizeof(double));
        else
            memcpy(b, a, (size_t)n0 * es * sizeof(double));
    }
}

void dq_pad_in(const dq_sim *s, dq_pad *d)
{
    const int N_pad = d->p.N, L_pad = d->p.L, es = d->p.is_complex ? 2 : 1;
    for (int l = 0; l < L_pad; l++)
        for (int i = 0; i < N_pad; i++)
            d->hs[(size_t)N_pad * l + i] = (l < d->L0 && i < d->N0) ? s->hs[(size_t)d->N0 * l + i] : 0;
    d->meq.n_sample = s->meq.n_sample;
    d->meq.sign[0] = s->meq.sign[0];
    d->meq.sign[1] = s->meq.sign[1];
    for (int k = 0; k < DQ_EQ_COUNT; k++) {
        if (!d->meq.acc[k]) continue;
        const long lp = dqmc_acc_len(&d->p, 0, k), l0 = dqmc_acc_len(&s->p, 0, k);
        memset(d->meq.acc[k], 0, (size_t)lp * es * sizeof(double));
        memcpy(d->meq.acc[k], s->meq.acc[k], (size_t)l0 * es * sizeof(double));
    }
    d->mue.n_sample = s->mue.n_sample;
    d->mue.sign[0] = s->mue.sign[0];
    d->mue.sign[1] = s->mue.sign[1];
    for (int k = 0; k < DQ_UE_COUNT; k++)
        if (d->mue.acc[k]) ue_copy(d, &s->p, k, d->mue.acc[k], s->mue.acc[k], 1);
}

void dq_pad_out(const dq_pad *d, dq_sim *s)
{
    const int N_pad = d->p.N, es = d->p.is_complex ? 2 : 1;
    for (int l = 0; l < d->L0; l++)
        for (int i = 0; i < d->N0; i++)
            s->hs[(size_t)d->N0 * l + i] = d->hs[(size_t)N_pad * l + i];
    s->meq.n_sample = d->meq.n_sample;
    s->meq.sign[0] = d->meq.sign[0];
    s->meq.sign[1] = d->meq.sign[1];
    for (int k = 0; k < DQ_EQ_COUNT; k++) {
        if (!d->meq.acc[k] || !s->meq.acc[k]) continue;
        const long l0 = dqmc_acc_len(&s->p, 0, k);
        memcpy(s->meq.acc[k], d->meq.acc[k], (size_t)l0 * es * sizeof(double));
    }
    if (s->p.period_uneqlt <= 0) return;
    s->mue.n_sample = d->mue.n_sample;
    s->mue.sign[0] = d->mue.sign[0];
    s->mue.sign[1] = d->mue.sign[1];
    for (int k = 0; k < DQ_UE_COUNT; k++)
        if (d->mue.acc[k] && s->mue.acc[k]) ue_copy(d, &s->p, k, d->mue.acc[k], s->mue.acc[k], 0);
}

void dq_pad_free(dq_pad *d)
{
    for (int k = 0; k < d->n_own; k++) free(d->own[k]);
    d->n_own = 0;
}
