/*
 * dq_driver.h -- the library seam of the reference: dqmc_wrapper
 * (src/dqmc.h:8-10, src/dqmc.c:513-671), on the GPU engine, for one file or
 * for a batch of files that run together.
 */
#ifndef DQ_DRIVER_H
#define DQ_DRIVER_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WRAP_FAIL (-1)
#define WRAP_INCOMPLETE 1

/* Runs sim_file to completion (or until a signal / the time limit), logging
 * to log_file (appended; NULL = stdout).  max_time_ns <= 0: no limit.
 * Returns 0 (done, or nothing to do), WRAP_INCOMPLETE (interrupted; state as
 * of the last checkpoint is on disk) or WRAP_FAIL. */
int dqmc_wrapper(const char *sim_file, const char *log_file, int64_t max_time_ns, bool dry,
                 bool bench);

/* The same for n files at once.  Files of one shape (N, L, n_matmul, element
 * type, measurement set) share one engine; engines of different shapes run
 * concurrently on their own streams.  codes[i] receives file i's return code;
 * logs[i] may be NULL (stdout). */
void dqmc_wrapper_batch(int n, const char *const *files, const char *const *logs,
                        int64_t max_time_ns, bool dry, bool bench, int *codes);

/* Print the reference's profile table (src/prof.h:8-22, src/prof.c:12-50:
 * time per section, the engine's, and the host's checkpoints) into every log
 * at the end of the run; the engines then time each section with HIP events
 * and launch kernel by kernel.  The reference selects this at compile time
 * (PROFILE_ENABLE). */
void dq_set_profile(bool on);

/* install the stop-signal handlers (SIGUSR1, SIGINT, SIGTERM, SIGHUP), as
 * sig_init does (src/sig.c:43-59) */
void dq_sig_init(void);

/* monotonic wall clock in ns */
int64_t dq_time_ns(void);

#ifdef __cplusplus
}
#endif
#endif
