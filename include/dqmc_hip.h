/*
 * dqmc_hip.h -- C ABI of the MI355X DQMC engine (libdqmc_hip.so).
 *
 * The engine replaces the hot path of katherineding/dqmc-dev behind its own
 * seams:
 *   - the sweep engine `static int dqmc(struct sim_data *)`  (src/dqmc.c:84-511)
 *   - the kernels it calls: mul_seq / calc_eq_g / calc_ue_g  (src/greens.c:8-494),
 *     update_delayed (src/updates.c:7-84), measure_eqlt / measure_uneqlt
 *     (src/meas.c:66-1383), the RNG (src/rand.h:6-32)
 *   - the in-memory data it works on: struct params / state / meas_eqlt /
 *     meas_uneqlt (src/data.h:7-105), flattened here into plain pointers.
 * The reference has no plugin API; its FFI-equivalent seam is the C call
 * `dqmc(sim)` made by dqmc_wrapper (src/dqmc.c:621).  dqmc_engine_run() is the
 * drop-in for that call, dqmc_engine_{set,get}_* carry struct sim_data across
 * the boundary, and the host-side CLI (dqmc_1 / dqmc_stack) keeps the file
 * seam (the .h5 layout of src/data.c:345-789).
 *
 * Conventions
 *   - every matrix is column-major, element (i,j) at i + N*j, exactly the raw
 *     buffer the reference reads from the .h5 file (a numpy C-order (N,N) array
 *     A therefore appears as A^T, as in the reference);
 *   - `num` arrays are double (real build) or interleaved (re,im) double pairs
 *     (complex build, the reference's USE_CPLX); the choice is per engine at
 *     run time via dqmc_params.is_complex;
 *   - callers own every host buffer; the engine owns device memory;
 *   - return codes: 0 ok, DQMC_INTERRUPTED when the stop flag was seen,
 *     negative on error (dqmc_last_error() gives the message).
 *   - one HIP stream per engine; an engine handle is not thread-safe.
 */
#ifndef DQMC_HIP_H
#define DQMC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DQMC_OK 0
#define DQMC_INTERRUPTED 5 /* same value as src/dqmc.c:73 */
#define DQMC_ERR_ARG (-10)
#define DQMC_ERR_HIP (-11)
#define DQMC_ERR_NOMEM (-12)
#define DQMC_ERR_UNSUPPORTED (-13)

/* struct params (src/data.h:7-39), flattened.  Pointers refer to host memory
 * laid out exactly as the .h5 datasets of the same name. */
typedef struct dqmc_params {
    int N, L, Nx, Ny, F;
    int n_matmul, n_delay;
    int n_sweep_warm, n_sweep_meas, n_sweep;
    int period_eqlt, period_uneqlt;
    int checkpoint_every;
    int meas_bond_corr, meas_thermal, meas_2bond_corr, meas_energy_corr;
    int meas_nematic_corr, meas_pair_bb_only, meas_chiral, meas_local_JQ;
    int meas_gen_suscept;
    int num_i, num_ij, num_plaq_accum, num_b_accum, num_b2_accum;
    int num_plaq, num_b, num_b2, num_bs, num_bb, num_bb2, num_b2b, num_b2b2;
    int degen_i, degen_ij, degen_plaq, degen_b, degen_b2, degen_bs, degen_bb;
    int degen_bb2, degen_b2b, degen_b2b2;
    int is_complex;
    const int32_t *map_i, *map_ij, *bonds, *bond2s, *map_bs, *map_bb;
    const int32_t *map_b2b, *map_bb2, *map_b2b2, *plaqs, *map_plaq, *map_b, *map_b2;
    const double *peierlsu, *peierlsd, *pp_u, *pp_d, *ppr_u, *ppr_d; /* num */
    const double *exp_Ku, *exp_Kd, *inv_exp_Ku, *inv_exp_Kd;         /* num N*N */
    const double *exp_halfKu, *exp_halfKd, *inv_exp_halfKu, *inv_exp_halfKd;
    const double *exp_lambda, *del; /* double 2*N */
} dqmc_params;

/* Equal-time accumulators (struct meas_eqlt, src/data.h:46-64). */
enum dqmc_eq_acc {
    DQ_EQ_density, DQ_EQ_density_u, DQ_EQ_density_d, DQ_EQ_double_occ,
    DQ_EQ_g00, DQ_EQ_g00_u, DQ_EQ_g00_d, DQ_EQ_nn, DQ_EQ_xx, DQ_EQ_zz, DQ_EQ_pair_sw,
    DQ_EQ_uuuu, DQ_EQ_dddd, DQ_EQ_uudd, DQ_EQ_dduu, DQ_EQ_chi,
    DQ_EQ_kk, DQ_EQ_kv, DQ_EQ_kn, DQ_EQ_vv, DQ_EQ_vn,
    DQ_EQ_j, DQ_EQ_j2, DQ_EQ_jn,
    DQ_EQ_COUNT
};

/* Unequal-time accumulators (struct meas_uneqlt, src/data.h:66-91). */
enum dqmc_ue_acc {
    DQ_UE_gt0, DQ_UE_gt0_u, DQ_UE_gt0_d, DQ_UE_nn, DQ_UE_xx, DQ_UE_zz, DQ_UE_pair_sw,
    DQ_UE_uuuu, DQ_UE_dddd, DQ_UE_uudd, DQ_UE_dduu,
    DQ_UE_pair_bb, DQ_UE_jj, DQ_UE_jsjs, DQ_UE_kk, DQ_UE_ksks,
    DQ_UE_pair_b2b2, DQ_UE_j2jn, DQ_UE_jnj2, DQ_UE_jnjn, DQ_UE_j2j, DQ_UE_jj2,
    DQ_UE_j2j2, DQ_UE_js2js2, DQ_UE_k2k2, DQ_UE_ks2ks2, DQ_UE_jnj, DQ_UE_jjn,
    DQ_UE_kv, DQ_UE_kn, DQ_UE_vv, DQ_UE_vn, DQ_UE_nem_nnnn, DQ_UE_nem_ssss,
    DQ_UE_COUNT
};

#define DQMC_MAX_ACC 40

/* One accumulator group.  acc[k] points to a `num` array of the size the
 * reference allocates for it (src/data.c:441-517), or NULL when the group is
 * switched off by the meas_* toggles. */
typedef struct dqmc_meas {
    int n_sample;
    double sign[2]; /* num: (re, im); im is 0 in real builds */
    double *acc[DQMC_MAX_ACC];
} dqmc_meas;

/* Length (in num elements) of accumulator `k` of the equal-time (ue=0) or
 * unequal-time (ue=1) group, or 0 when the parameters switch it off. */
long dqmc_acc_len(const dqmc_params *p, int ue, int k);

/* ---------------------------------------------------------------------- */
/* Engine: a batch of independent Markov chains of one shape (N, L, F,
 * n_matmul, dtype and measurement set equal; all physical parameters may
 * differ per chain).  Chains of uneven size share an engine by padding (see
 * chain_L / chain_N below); otherwise other shapes go into other engines.    */
typedef struct dqmc_engine dqmc_engine;

/* Tuning knobs picked by the autotuner; zero fields mean "engine default". */
typedef struct dqmc_tuning {
    int delay_rank; /* rank of the delayed Metropolis update; 0 = the params'
                       n_delay.  Rounded to a compiled variant: powers of two,
                       real 8..16 (N <= 64), 8..32 (N <= 256), 8..16 (N <= 512);
                       complex 4..8, 8..16, 4; any rank <= N above N = 512 */
    int flags;      /* bit 0: time the N x N x N GEMMs with HIP events (the
                       sweeps then run launch by launch); bit 1: do not capture
                       the sweeps into HIP graphs (needed when other host
                       threads make HIP calls while this engine runs: HIP's
                       stream capture fails then); bit 2: the wrap drift check
                       (dqmc_engine_drift); bit 3: the per-section profile
                       (dqmc_engine_profile; the sweeps then run launch by
                       launch) */
    int gemm_tile;  /* C tile of the MFMA GEMM: 0 = time the shapes at creation
                       and keep the fastest, 1 = 64x64, 2 = 128x64 with K
                       steps of 16, 3 = 128x64 with K steps of 8, 4 = 128x64
                       on the 16x16x4 form with 128-bit LDS reads for the real
                       products whose sizes are multiples of 128 x 64 x 16
                       (others take 64x64) */
    /* Chains of uneven size in one engine (both NULL: every chain has the
     * params' N and L).  chain_L[c] <= L and chain_N[c] <= N are chain c's own
     * slice and site counts; its params are padded to the common N and L by
     * the caller: sites >= chain_N[c] carry the identity (exp_K blocks = I,
     * exp_lambda = 1, del = 0) and map to measurement bins the caller drops,
     * and slices >= chain_L[c] are identity slices (no update, no wrap, no
     * measurement; the chain's last group of n_matmul slices may be partial).
     * Its random stream advances exactly as for a chain of its own size,
     * and it measures at its own params' period_eqlt (which, unlike the
     * other shape fields, may differ between the chains of a padded engine).
     * Unequal-time measurements run on padded chains (period_uneqlt shared
     * by the chains): the identity slices follow a chain's last slice, so
     * its blocks G(t,0), G(0,t), G(t,t) for t < chain_L[c] are its own, and
     * the accumulator slices t >= chain_L[c] hold nothing the caller keeps.
     * The adjoint and the calc_ue_g hook need equal sizes.
     * No counterpart in the reference, which runs one chain per process. */
    const int *chain_L;
    const int *chain_N;
} dqmc_tuning;

/* Create an engine for n_chains chains on HIP device `device`.  params[c]
 * describes chain c (its arrays are copied to the device). */
int dqmc_engine_create(dqmc_engine **out, const dqmc_params *const *params, int n_chains,
                       int device, const dqmc_tuning *tune);
void dqmc_engine_destroy(dqmc_engine *e);

/* struct state (src/data.h:41-45): rng[17], sweep, hs[L*N] (0/1).  Setting the
 * state marks G stale: the next run recomputes it from scratch as dqmc()
 * does on entry (src/dqmc.c:188-220). */
int dqmc_engine_set_state(dqmc_engine *e, int chain, const uint64_t *rng, int sweep,
                          const int32_t *hs);
int dqmc_engine_get_state(dqmc_engine *e, int chain, uint64_t *rng, int *sweep, int32_t *hs);

/* Accumulators: copy host -> device (set) or device -> host (get). ue may be
 * NULL when period_uneqlt == 0. */
int dqmc_engine_set_meas(dqmc_engine *e, int chain, const dqmc_meas *eq, const dqmc_meas *ue);
int dqmc_engine_get_meas(dqmc_engine *e, int chain, dqmc_meas *eq, dqmc_meas *ue);

/* Advance every chain by up to n_sweeps sweeps (a chain stops at its
 * n_sweep).  The loop body is that of src/dqmc.c:222-459 without the
 * checkpoint, which the host performs between calls.  *stop_flag (may be NULL)
 * is polled before each sweep as sig_check_state is (src/dqmc.c:224-230);
 * when it is non-zero the call returns DQMC_INTERRUPTED.  The call queues a
 * sweep only once the sweep two before it has finished on the device, so a
 * flag that a signal handler or a timer sets while the call runs ends it
 * within two sweeps. */
int dqmc_engine_run(dqmc_engine *e, int n_sweeps, const volatile int *stop_flag);

/* Oracle hooks (read-only):
 *  - the engine's current equal-time Green's function of one spin and the
 *    current phase of the weight;
 *  - G computed from scratch for the current field at slice boundary l by the
 *    stabilized product (calc_eq_g((l/n_matmul)%F ...), src/greens.c:69-208)
 *    and the phase of det(I + B..B) it returns;
 *  - the unequal-time blocks G(0,t), G(t,t), G(t,0) for all t
 *    (calc_ue_g, src/greens.c:471-494), before the half-wrap. */
int dqmc_engine_green(dqmc_engine *e, int chain, int spin, double *G, double *phase);
int dqmc_engine_calc_eq_g(dqmc_engine *e, int chain, int spin, int f, double *G,
                          double *phase);
int dqmc_engine_calc_ue_g(dqmc_engine *e, int chain, int spin, double *G0t, double *Gtt,
                          double *Gt0);

/* Adjoint of log|det M_s| through the stabilized B-chain.  No counterpart in
 * the reference, which computes only the phase (its determinant code is
 * commented out, src/greens.c:180-191).  For the current HS field of every
 * chain (no updates, state unchanged), with B_l = E e^{V_l}, E = e^{-dt K}
 * the engine's exp_K and G_l the Green's function with B_l at the right end:
 *   logdet[2c+s] = log|det(I + B_{L-1}..B_0)|                 (double)
 *   W[2c+s]      = sum_l e^{V_l} (I - G_l) e^{-V_l}, N x N column-major (num);
 *                  d logdet = Re tr(W E^-1 dE)
 *   trG[2c+s]    = sum_l tr G_l (num);  d logdet / d mu = dt (N L - trG)
 *   sG[2c+s]     = sum_l sum_i sigma_s s_li (1 - G_l,ii) (num), s = 2 hs - 1,
 *                  sigma = +1 up / -1 down;  d logdet / d lambda = sG
 * G is advanced by wraps and recomputed from the stored group products every
 * n_matmul slices.  Any pointer may be NULL.  The next run recomputes G.
 * "num" is one double per value for real engines and a (re, im) pair for
 * complex ones (W: N*N of them per (chain, spin)). */
int dqmc_engine_adjoint(dqmc_engine *e, double *logdet, double *W, double *trG, double *sG);

/* Adjoint of a G-based observable.  No counterpart in the reference.  For the
 * current HS field of every chain and X[s] (N x N column-major, num; X[0] for
 * spin up, X[1] for spin down, shared by the chains), with G the equal-time
 * Green's function at slice 0 and Z_l = G(t_l,0) X G^(0,t_l) built from the
 * unequal-time Green's functions of calc_ue_g (G^(0,0) = G - I):
 *   obs[2c+s] = tr(X_s G)                                              (num)
 *   W[2c+s]   = sum_l e^{V_l} Z_l e^{-V_l}  (num, N x N);  d obs = Re tr(W E^-1 dE)
 *   trZ[2c+s] = sum_l tr Z_l  (num);  d obs / d mu = dt trZ
 *   sZ[2c+s]  = sum_l sum_i sigma_s s_li (Z_l)_ii  (num);  d obs / d lambda = sZ
 * The unequal-time buffers are allocated on first use when the files measure
 * no unequal-time quantities.  Any output pointer may be NULL.  The next run
 * recomputes G. */
int dqmc_engine_adjoint_green(dqmc_engine *e, const double *X, double *obs, double *W,
                              double *trZ, double *sZ);

/* Accepted flips so far, summed over chains (for the flop model). */
int64_t dqmc_engine_accepts(dqmc_engine *e);

/* Oracle hook: accepted flips of chain c per slice, counts[l] for l < L,
 * summed over every sweep since the engine was created (the count of
 * acceptances of update_delayed at slice l, src/updates.c:26-50).  In one
 * sweep it equals the number of sites whose field changed at slice l. */
int dqmc_engine_slice_accepts(dqmc_engine *e, int chain, int64_t *counts);

/* Time per section of the reference's profile table (src/prof.h:8-22, printed
 * by src/dqmc.c:661-663), accumulated over every run of an engine created with
 * dqmc_tuning.flags bit 3: names[k], ms[k] and calls[k] for the k-th section
 * in the reference's order (wall, updates, calcb, multb, recalc, wrap,
 * half_wrap, meas_eq, calc_o, bsofi, expand_g, meas_uneq, meas_uneq_sub,
 * checkpoint).  Device sections are timed by HIP event pairs on the engine
 * stream (the sweeps then run launch by launch); wall is the host time of the
 * dqmc_engine_run calls; meas_uneq_sub lies inside meas_uneq, as in the
 * reference; checkpoint is the host's (0 here).  Any pointer may be NULL.
 * Returns the number of sections (14), or a negative code. */
int dqmc_engine_profile(dqmc_engine *e, const char **names, double *ms, int64_t *calls, int max);

/* Average duration (ms) of the dominant GEMM kernel in the last run, measured
 * with HIP events on the engine stream, and its launch count. */
int dqmc_engine_gemm_stats(dqmc_engine *e, double *avg_ms, int64_t *launches,
                           double *flops_per_launch);

/* Kernel-level test hook (no counterpart in the reference): one GEMM through
 * the engine's FP64 MFMA kernel on host buffers,
 *   C = alpha diag(r) op(A) diag(u) op(B) diag(w) + beta C
 * op: 0 = N, 1 = T, 2 = conjugate transpose; alpha/beta are num; r/u/w may
 * be NULL; tile as dqmc_tuning.gemm_tile (0 = 64x64).  Used by the parity
 * tests of the GEMM kernel. */
int dqmc_test_gemm(int is_complex, int opa, int opb, int M, int N, int K, const double *alpha,
                   const double *A, int lda, const double *B, int ldb, const double *beta,
                   double *C, int ldc, const double *r, const double *u, const double *w,
                   int tile);

/* Measurement hook (no counterpart in the reference): the sustained FP64
 * matrix-core rate of this device in TFLOP/s, from a register-only loop of
 * independent MFMA chains on every CU (form 0: the 16x16x4 instruction,
 * form 1: the 4x4x4 one, which the GEMM tiles use).  Roofline context for
 * the published peak. */
int dqmc_probe_mfma(int form, double *tflops);

/* Replace chain c's numeric parameters (the exponentials, exp_lambda, del,
 * the Peierls phases and those of the two-hop bonds) in an existing engine;
 * chain -1: every chain (one
 * upload, copied on the device).  N, L, n_matmul and the element type must be
 * the engine's, and the maps are kept.  No counterpart
 * in the reference, whose parameters are fixed per file; it serves callers
 * that evaluate the same HS fields for many parameter values (the adjoint's
 * parameter scans) without rebuilding the engine.  G is recomputed before
 * the next sweep.  Not for padded (chain_L / chain_N) engines. */
int dqmc_engine_set_params(dqmc_engine *e, int chain, const dqmc_params *params);

/* Wrap drift (CHECK_G_WRP, src/dqmc.c:277-287; a compile-time check there):
 * with dqmc_tuning.flags bit 2 the engine also wraps G at every
 * recalculation slice and compares it with the recomputed G.  max_abs[c] <-
 * the largest |G_wrapped - G_recomputed| of chain c over both spins since the
 * previous call (which resets it).  Used to pick the stabilization interval. */
int dqmc_engine_drift(dqmc_engine *e, double *max_abs);

/* The GEMM tile shape the engine uses (dqmc_tuning.gemm_tile values 1..4). */
int dqmc_engine_gemm_tile(dqmc_engine *e);

/* The delay rank the engine's Metropolis kernel runs at (n_delay of
 * src/updates.c:7, after rounding to a compiled variant). */
int dqmc_engine_delay_rank(dqmc_engine *e);

/* Stabilized equal-time Green's functions from caller-held group products,
 * allocating no memory.  It has no counterpart in the reference.  It is for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  It computes what
 * dqmc_engine_calc_eq_g computes (calc_eq_g, src/greens.c:69-208) for the
 * 2*n_chains matrices b = 2*chain + spin at once:
 *   C      device, [2 n_chains][F][N*N] num: C_f = B_{(f+1)n-1} .. B_{fn},
 *          column-major (the group products of mul_seq, src/greens.c:8-41)
 *   G      device, [2 n_chains][N*N] num  <- (I + C_{f-1} .. C_f)^-1 (cyclic)
 *   phase  device, [2 n_chains] num       <- phase of det(I + C_{f-1} .. C_f)
 *   stream a hipStream_t (NULL: the null stream); the call only queues work
 *          on it and returns.
 * The work space is one static device array of DQMC_NOALLOC_SCRATCH_BYTES on
 * each GPU, held until the process ends and cleared at the start of every
 * call.  Every call on a GPU shares that GPU's array: calls that overlap on
 * one GPU race.  A request whose work space exceeds the array fails with
 * DQMC_ERR_ARG (about 11 N*N num per matrix: real N = 256 fits 5 chains,
 * complex N = 144 fits 7, real N = 64 fits 59). */
#define DQMC_NOALLOC_SCRATCH_BYTES (64L << 20)
int dqmc_calc_eq_g_noalloc(int is_complex, int N, int F, int n_chains, int f, const double *C,
                           double *G, double *phase, void *stream);

const char *dqmc_last_error(void);
const char *dqmc_version(void);

#ifdef __cplusplus
}
#endif
#endif
