/*
 * dq_driver.c -- dqmc_wrapper (src/dqmc.c:513-671) on the GPU engine.
 *
 * Per file, the log carries the same milestones as the reference's wrapper:
 * header, memory requirement, dry-run exit, open/read, checkpoint policy,
 * progress, interruption, save and wall time.  The sweep loop is
 * src/dqmc.c:222-459 with the engine doing the sweeps:
 *   - before every sweep the stop flag is polled (the engine does it between
 *     sweeps); a signal or the time limit ends the run with INTERRUPTED, and
 *     nothing is saved (the file keeps its last checkpoint);
 *   - a chain whose sweep is a multiple of checkpoint_every is saved before
 *     that sweep runs (src/dqmc.c:233-244);
 *   - after the last sweep the file is saved unless in benchmark mode.
 * Files of one shape share an engine (one chain each); files of uneven size
 * that dq_bucket.h admits share one too, padded to the group's largest N and
 * L.  Each group runs in its own host thread with its own engine and stream.
 */
#define _GNU_SOURCE
#include "dq_driver.h"

#include <limits.h>
#include <pthread.h>
#include <signal.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "dq_bucket.h"
#include "dq_h5.h"
#include "dqmc_hip.h"

#define MAX_CHAINS_PER_ENGINE 256

static volatile sig_atomic_t g_stop = 0; /* signal number, or -1 at the time limit */

static void on_stop(int s) { g_stop = s; }

void dq_sig_init(void)
{
    static int done = 0;
    if (done) return;
    done = 1;
    struct sigaction sa;
    memset(&sa, 0, sizeof sa);
    sa.sa_handler = on_stop;
    sigaction(SIGUSR1, &sa, NULL);
    sigaction(SIGINT, &sa, NULL);
    sigaction(SIGTERM, &sa, NULL);
    sigaction(SIGHUP, &sa, NULL);
}

int64_t dq_time_ns(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (int64_t)t.tv_sec * 1000000000LL + t.tv_nsec;
}

#define N_PROF 14

typedef struct job {
    const char *file;
    FILE *log;
    int code;
    int loaded, failed, interrupted;
    int last_ckpt;
    dq_sim sim;
    /* the engine's profile table (src/prof.h:8-22) and this file's checkpoints */
    int prof_chains;
    const char *prof_name[N_PROF];
    double prof_ms[N_PROF];
    int64_t prof_calls[N_PROF];
    int64_t ckpt_ns, ckpt_calls;
} job;

static bool g_profile = false;

void dq_set_profile(bool on) { g_profile = on; }



typedef struct group {
    job **j;
    int n;
    int bucket; /* size bucket (dq_bucket_of), 0 for a group of one shape */
    int alone;  /* the only group: its sweeps may be captured into HIP graphs */
} group;

static pthread_mutex_t g_log_mu = PTHREAD_MUTEX_INITIALIZER;

static void logf_(job *j, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static void logf_(job *j, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    pthread_mutex_lock(&g_log_mu);
    vfprintf(j->log, fmt, ap);
    fflush(j->log);
    pthread_mutex_unlock(&g_log_mu);
    va_end(ap);
}

/* the table of src/prof.c:12-50 (profile_print): sections sorted by time,
 * share of the wall time, total, time per call and calls; the device
 * sections are the engine's, shared by the chains that ran in it */
static void print_profile(job *j, double wall_s)
{
    if (!j->prof_chains) return;
    int idx[N_PROF];
    for (int i = 0; i < N_PROF; i++) {
        int k = i;
        while (k > 0 && j->prof_ms[idx[k - 1]] < j->prof_ms[i]) {
            idx[k] = idx[k - 1];
            k--;
        }
        idx[k] = i;
    }
    logf_(j, "gpu_engine_(%d_chain%s)|_%% of all_|___total (s)_|___us per call_|___# calls\n",
          j->prof_chains, j->prof_chains == 1 ? "" : "s");
    for (int q = 0; q < N_PROF; q++) {
        const int i = idx[q];
        if (j->prof_calls[i] == 0) continue;
        logf_(j, "%16s |%9.3f |%12.3f |%14.3f |%10lld\n", j->prof_name[i],
              100.0 * j->prof_ms[i] * 1e-3 / wall_s, j->prof_ms[i] * 1e-3,
              1e3 * j->prof_ms[i] / (double)j->prof_calls[i], (long long)j->prof_calls[i]);
    }
    logf_(j, "---------------------------------------------------------------------\n");
}

static void now_str(char *buf, size_t n)
{
    time_t t = time(NULL);
    struct tm tmv;
    localtime_r(&t, &tmv);
    strftime(buf, n, "%a %b %d %H:%M:%S %Y", &tmv);
}

static int same_shape(const dqmc_params *a, const dqmc_params *b)
{
    return a->N == b->N && a->L == b->L && a->F == b->F && a->n_matmul == b->n_matmul &&
           a->is_complex == b->is_complex && a->period_eqlt == b->period_eqlt &&
           a->period_uneqlt == b->period_uneqlt && a->num_i == b->num_i &&
           a->num_ij == b->num_ij && a->meas_energy_corr == b->meas_energy_corr &&
           a->meas_bond_corr == b->meas_bond_corr && a->meas_thermal == b->meas_thermal &&
           a->meas_2bond_corr == b->meas_2bond_corr && a->meas_chiral == b->meas_chiral &&
           a->meas_gen_suscept == b->meas_gen_suscept && a->meas_local_JQ == b->meas_local_JQ &&
           a->meas_nematic_corr == b->meas_nematic_corr &&
           a->meas_pair_bb_only == b->meas_pair_bb_only && a->num_b == b->num_b &&
           a->num_bb == b->num_bb && a->num_bs == b->num_bs && a->num_b2 == b->num_b2 &&
           a->num_b2b2 == b->num_b2b2 && a->num_b2b == b->num_b2b && a->num_bb2 == b->num_bb2;
}

/* chain c's state and accumulators into its file's buffers; a padded chain
 * goes through its padded buffers */
static int fetch(dqmc_engine *e, int c, job *j, dq_pad *pad)
{
    if (pad) {
        int rc = dqmc_engine_get_state(e, c, j->sim.rng, &j->sim.sweep, pad->hs);
        if (!rc)
            rc = dqmc_engine_get_meas(e, c, &pad->meq,
                                      j->sim.p.period_uneqlt > 0 ? &pad->mue : NULL);
        if (!rc) dq_pad_out(pad, &j->sim);
        return rc;
    }
    int rc = dqmc_engine_get_state(e, c, j->sim.rng, &j->sim.sweep, j->sim.hs);
    if (!rc)
        rc = dqmc_engine_get_meas(e, c, &j->sim.meq,
                                  j->sim.p.period_uneqlt > 0 ? &j->sim.mue : NULL);
    return rc;
}

static void fail_group(group *g, const char *what)
{
    for (int c = 0; c < g->n; c++) {
        logf_(g->j[c], "%s: %s\n", what, dqmc_last_error());
        g->j[c]->code = WRAP_FAIL;
    }
}

static void *run_group(void *arg)
{
    group *g = (group *)arg;
    const int n = g->n;
    /* uneven sizes: pad every file to the group's largest N, L and bin counts */
    int uneven = 0, N_pad = 0, L_pad = 0, ni_pad = 0, nij_pad = 0;
    for (int c = 0; c < n; c++) {
        const dqmc_params *p = &g->j[c]->sim.p;
        uneven |= !same_shape(p, &g->j[0]->sim.p);
        if (p->N > N_pad) N_pad = p->N;
        if (p->L > L_pad) L_pad = p->L;
        if (p->num_i > ni_pad) ni_pad = p->num_i;
        if (p->num_ij > nij_pad) nij_pad = p->num_ij;
    }
    dq_pad *pads = uneven ? calloc(n, sizeof *pads) : NULL;
    int *Lc = uneven ? calloc(n, sizeof *Lc) : NULL, *Nc = uneven ? calloc(n, sizeof *Nc) : NULL;
    const dqmc_params **P = calloc(n, sizeof *P);
    dqmc_engine *e = NULL;
    /* HIP stream capture fails while other host threads make HIP calls, so
     * graphs are used only when this group is the only one */
    dqmc_tuning tune = {0, (g->alone ? 0 : 2) | (g_profile ? 8 : 0), 0, NULL, NULL};
    int rc = 0;
    for (int c = 0; c < n && !rc; c++) {
        job *j = g->j[c];
        P[c] = &j->sim.p;
        if (!uneven) continue;
        rc = dq_pad_build(&j->sim, N_pad, L_pad, ni_pad, nij_pad, &pads[c]);
        if (rc) {
            logf_(j, "padding to N=%d, L=%d failed: %d\n", N_pad, L_pad, rc);
            break;
        }
        P[c] = &pads[c].p;
        Lc[c] = j->sim.p.L;
        Nc[c] = j->sim.p.N;
    }
    if (uneven) {
        tune.chain_L = Lc;
        tune.chain_N = Nc;
    }
    if (!rc) {
        rc = dqmc_engine_create(&e, P, n, 0, &tune);
        if (rc) fail_group(g, "engine creation failed");
    } else {
        for (int c = 0; c < n; c++) g->j[c]->code = WRAP_FAIL;
    }
    free(P);
    free(Lc);
    free(Nc);
    if (rc) goto done;
    for (int c = 0; c < n && !rc; c++) {
        job *j = g->j[c];
        if (uneven) {
            dq_pad_in(&j->sim, &pads[c]);
            rc = dqmc_engine_set_state(e, c, j->sim.rng, j->sim.sweep, pads[c].hs);
            if (!rc)
                rc = dqmc_engine_set_meas(e, c, &pads[c].meq,
                                          j->sim.p.period_uneqlt > 0 ? &pads[c].mue : NULL);
            continue;
        }
        rc = dqmc_engine_set_state(e, c, j->sim.rng, j->sim.sweep, j->sim.hs);
        if (!rc)
            rc = dqmc_engine_set_meas(e, c, &j->sim.meq,
                                      j->sim.p.period_uneqlt > 0 ? &j->sim.mue : NULL);
    }
    if (rc) {
        fail_group(g, "engine upload failed");
        dqmc_engine_destroy(e);
        goto done;
    }
    char ts[64];
    now_str(ts, sizeof ts);
    for (int c = 0; c < n; c++)
        logf_(g->j[c], "[%s] starting dqmc (MI355X engine, %d chain%s in this batch)\n", ts, n,
              n == 1 ? "" : "s");
    if (uneven)
        for (int c = 0; c < n; c++)
            logf_(g->j[c], "files of uneven size in this batch: padded from N=%d, L=%d to N=%d, L=%d\n",
                  g->j[c]->sim.p.N, g->j[c]->sim.p.L, N_pad, L_pad);

    int interrupted = 0, failed = 0;
    for (;;) {
        int any = 0, kmin = INT_MAX;
        if (g_stop) {
            interrupted = 1;
            break;
        }
        for (int c = 0; c < n; c++) {
            job *j = g->j[c];
            const int sw = j->sim.sweep, ns = j->sim.p.n_sweep, ce = j->sim.p.checkpoint_every;
            if (j->failed || sw >= ns) continue;
            any = 1;
            if (ce > 0 && sw % ce == 0 && j->last_ckpt != sw) {
                j->last_ckpt = sw;
                const int64_t t0 = dq_time_ns();
                const int bad = fetch(e, c, j, uneven ? &pads[c] : NULL) != 0 ||
                                dq_sim_save(&j->sim) < 0;
                j->ckpt_ns += dq_time_ns() - t0;
                j->ckpt_calls++;
                if (bad) {
                    logf_(j, "dqmc() checkpoint failed\n");
                    j->failed = 1;
                    j->code = WRAP_FAIL;
                    continue;
                }
            }
            int k = ns - sw;
            if (ce > 0 && ce - sw % ce < k) k = ce - sw % ce;
            if (k < kmin) kmin = k;
        }
        if (!any || kmin == INT_MAX) break;
        rc = dqmc_engine_run(e, kmin, (const volatile int *)&g_stop);
        for (int c = 0; c < n; c++) dqmc_engine_get_state(e, c, NULL, &g->j[c]->sim.sweep, NULL);
        if (rc == DQMC_INTERRUPTED) {
            interrupted = 1;
            break;
        }
        if (rc < 0) {
            failed = 1;
            break;
        }
    }
    const char *why = g_stop == -1 ? "time limit reached" : "signal received";
    for (int c = 0; c < n; c++) {
        job *j = g->j[c];
        if (j->failed) continue;
        if (failed) {
            logf_(j, "engine run failed: %s\n", dqmc_last_error());
            j->code = WRAP_FAIL;
            continue;
        }
        if (fetch(e, c, j, uneven ? &pads[c] : NULL) != 0) {
            logf_(j, "engine download failed: %s\n", dqmc_last_error());
            j->code = WRAP_FAIL;
            continue;
        }
        if (j->sim.sweep < j->sim.p.n_sweep && interrupted) {
            logf_(j, "%d/%d sweeps completed (%s)\n", j->sim.sweep, j->sim.p.n_sweep,
                  j->sim.sweep >= j->sim.p.n_sweep_warm ? "measuring" : "warming up");
            logf_(j, "stop_flag = %d (%s), stopping\n", (int)g_stop, why);
            logf_(j, "dqmc() interrupted\n");
            j->interrupted = 1;
            j->code = WRAP_INCOMPLETE;
        }
    }
    now_str(ts, sizeof ts);
    for (int c = 0; c < n; c++) {
        job *j = g->j[c];
        if (j->failed || j->interrupted || j->code == WRAP_FAIL) continue;
        logf_(j, "[%s] %d/%d sweeps completed\n", ts, j->sim.sweep, j->sim.p.n_sweep);
    }
    if (g_profile) {
        const char *names[N_PROF];
        double ms[N_PROF];
        int64_t calls[N_PROF];
        if (dqmc_engine_profile(e, names, ms, calls, N_PROF) == N_PROF) {
            for (int c = 0; c < n; c++) {
                job *j = g->j[c];
                j->prof_chains = n;
                for (int i = 0; i < N_PROF; i++) {
                    j->prof_name[i] = names[i];
                    j->prof_ms[i] = ms[i];
                    j->prof_calls[i] = calls[i];
                    if (!strcmp(names[i], "checkpoint")) {  /* the host's own */
                        j->prof_ms[i] = (double)j->ckpt_ns * 1e-6;
                        j->prof_calls[i] = j->ckpt_calls;
                    }
                }
            }
        }
    }
    dqmc_engine_destroy(e);
done:
    if (pads) {
        for (int c = 0; c < n; c++) dq_pad_free(&pads[c]);
        free(pads);
    }
    return NULL;
}

void dqmc_wrapper_batch(int n, const char *const *files, const char *const *logs,
                        int64_t max_time_ns, bool dry, bool bench, int *codes)
{
    const int64_t wall_start = dq_time_ns();
    job *jobs = calloc(n, sizeof *jobs);
    for (int i = 0; i < n; i++) {
        job *j = &jobs[i];
        j->file = files[i];
        j->code = 0;
        j->last_ckpt = -1;
        j->log = (logs && logs[i]) ? fopen(logs[i], "a") : stdout;
        if (!j->log) {
            fprintf(stderr, "fopen() failed to open: %s\n", logs[i]);
            j->log = NULL;
            j->code = WRAP_FAIL;
            continue;
        }
        logf_(j, "executable %s\n", dqmc_version());
        char commit[256];
        if (dq_sim_commit(j->file, commit, sizeof commit) == 0)
            logf_(j, "hdf5 generation script commit id %s\n", commit);
        logf_(j, "integer size on this system: %zu bytes\n", sizeof(int));
        const int64_t mem = dq_sim_memory_req(j->file);
        if (mem < 0) logf_(j, "minimum RAM check failed: %lld\n", (long long)mem);
        else logf_(j, "minimum RAM requirement: %.2f MB\n", (double)mem * 1e-6);
        if (dry) {
            logf_(j, "toggled dry run for mem check only\n");
            logf_(j, "Not performing: read sim file, alloc mem, run core DQMC logic, or save data\n");
            continue;
        }
        logf_(j, "opening %s\n", j->file);
        struct stat info;
        if (stat(j->file, &info) == 0)
            logf_(j, "size of this hdf5 file: %.2f MB\n", (double)info.st_size * 1e-6);
        const int rs = dq_sim_read(j->file, &j->sim);
        if (rs) {
            logf_(j, "sim_data_read_alloc() failed: %d\n", rs);
            j->code = WRAP_FAIL;
            dq_sim_free(&j->sim);
            continue;
        }
        j->loaded = 1;
        if (j->sim.p.checkpoint_every)
            logf_(j, "checkpointing every %d sweeps\n", j->sim.p.checkpoint_every);
        else
            logf_(j, "warning: running with no checkpoints\n");
        logf_(j, "%d/%d sweeps completed\n", j->sim.sweep, j->sim.p.n_sweep);
        if (j->sim.sweep > j->sim.p.n_sweep) {
            logf_(j, "sweep > n_sweep in %s\n", j->file);
            j->code = WRAP_FAIL;
        } else if (j->sim.sweep == j->sim.p.n_sweep) {
            logf_(j, "already finished so don't do anything\n");
        }
    }

    if (!dry) {
        dq_sig_init();
        /* group the runnable files by shape, or by size bucket where the
         * files allow padding */
        group *gs = calloc(n > 0 ? n : 1, sizeof *gs);
        int ng = 0;
        for (int i = 0; i < n; i++) {
            job *j = &jobs[i];
            if (!j->loaded || j->code != 0 || j->sim.sweep >= j->sim.p.n_sweep) continue;
            const dqmc_params *p = &j->sim.p;
            const int bk = dq_bucket_compatible(p, p) ? dq_bucket_of(p) : 0;
            int gi = -1;
            for (int q = 0; q < ng && gi < 0; q++) {
                if (gs[q].n >= MAX_CHAINS_PER_ENGINE || (gs[q].bucket != 0) != (bk != 0)) continue;
                const dqmc_params *p0 = &gs[q].j[0]->sim.p;
                if (bk ? dq_bucket_compatible(p0, p) : same_shape(p0, p)) gi = q;
            }
            if (gi < 0) {
                gi = ng++;
                gs[gi].j = calloc(n, sizeof(job *));
                gs[gi].bucket = bk;
            }
            gs[gi].j[gs[gi].n++] = j;
        }
        pthread_t *th = calloc(ng > 0 ? ng : 1, sizeof *th);
        int *joined = calloc(ng > 0 ? ng : 1, sizeof *joined);
        const int64_t deadline = max_time_ns > 0 ? wall_start + max_time_ns : 0;
        for (int q = 0; q < ng; q++) gs[q].alone = (ng == 1);
        for (int q = 0; q < ng; q++) pthread_create(&th[q], NULL, run_group, &gs[q]);
        if (deadline) {
            /* the time limit is a stop flag raised here; the engine polls the
             * flag before each sweep (sig_check_state, src/sig.c:70-75) */
            int left = ng;
            while (left > 0) {
                for (int q = 0; q < ng; q++)
                    if (!joined[q] && pthread_tryjoin_np(th[q], NULL) == 0) {
                        joined[q] = 1;
                        left--;
                    }
                if (left == 0) break;
                if (dq_time_ns() >= deadline && !g_stop) g_stop = -1;
                struct timespec ts = {0, 20 * 1000 * 1000};
                nanosleep(&ts, NULL);
            }
        } else {
            for (int q = 0; q < ng; q++) pthread_join(th[q], NULL);
        }
        free(joined);
        for (int q = 0; q < ng; q++) free(gs[q].j);
        free(gs);
        free(th);

        for (int i = 0; i < n; i++) {
            job *j = &jobs[i];
            if (!j->loaded || j->code != 0 || j->interrupted) continue;
            if (bench) {
                logf_(j, "benchmark mode enabled; not saving data to disk\n");
                continue;
            }
            logf_(j, "saving data to disk\n");
            const int ss = dq_sim_save(&j->sim);
            if (ss < 0) {
                logf_(j, "sim_data_save() failed: %d\n", ss);
                j->code = WRAP_FAIL;
            } else {
                logf_(j, "sim_data_save() succeeded\n");
            }
        }
    }

    const double wall = (double)(dq_time_ns() - wall_start) * 1e-9;
    for (int i = 0; i < n; i++) {
        job *j = &jobs[i];
        codes[i] = j->code;
        if (j->loaded) dq_sim_free(&j->sim);
        if (!j->log) continue;
        logf_(j, "wall time: %.3f\n", wall);
        print_profile(j, wall);
        if (j->log != stdout) fclose(j->log);
        else fflush(stdout);
    }
    free(jobs);
}

int dqmc_wrapper(const char *sim_file, const char *log_file, int64_t max_time_ns, bool dry,
                 bool bench)
{
    int code = 0;
    const char *logs[1] = {log_file};
    dqmc_wrapper_batch(1, &sim_file, logs, max_time_ns, dry, bench, &code);
    return code;
}
