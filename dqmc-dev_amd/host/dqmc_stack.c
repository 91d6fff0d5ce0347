/*
 * dqmc_stack -- work through a stack file of simulation paths on the MI355X
 * engine (src/main_stack.c:309-479).
 *
 *   dqmc_stack [-t TIME] [-n|--dry-run] [-b|--benchmark] [-B|--batch N] [-p|--profile] stack_file
 *
 * The stack file holds one .h5 path per line and is used LIFO: paths are
 * popped from the end and a run that did not finish is pushed back.  Access
 * is serialised between processes by a lock made by hard-linking the stack
 * file to "<stack_file>~" (src/main_stack.c:47-118); a lock whose mtime has
 * not changed for 10 polls is taken to be left by a dead process and removed.
 * Each file logs to "<file>.log".  Where the reference pops one file per
 * iteration, this driver pops up to N (default 64) and runs them as one batch
 * on the GPU (files of one shape share an engine).  After a batch with an
 * incomplete file the driver stops, as the reference does.  The exit status
 * is EXIT_FAILURE only if the stack file cannot be accessed.
 */
#define _GNU_SOURCE
#include <argp.h>
#include <fcntl.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

#include "dq_driver.h"
#include "dqmc_hip.h"

#define MAX_LEN 512

const char *argp_program_version = "dqmc-dev_amd (MI355X HIP engine)";

static char hostname[65];
static int pid;

#define my_printf(...)                          \
    do {                                        \
        printf("%16s %6d: ", hostname, pid);    \
        printf(__VA_ARGS__);                    \
        fflush(stdout);                         \
    } while (0)

static void sleep_rand(double lo, double hi)
{
    const double t = lo + (hi - lo) * (double)rand() / RAND_MAX;
    const struct timespec ts = {(time_t)t, (long)(1e9 * (t - (long)t))};
    nanosleep(&ts, NULL);
}

static char *lock_name(const char *file)
{
    const size_t n = strlen(file);
    char *l = calloc(n + 2, 1);
    memcpy(l, file, n);
    l[n] = '~';
    return l;
}

/* 0 on success; with retry, waits until the lock is obtained */
static int lock_file(const char *file, int retry)
{
    char *lf = lock_name(file);
    struct timespec seen = {0, 0};
    int same = 0;
    for (;;) {
        if (link(file, lf) == 0) {
            free(lf);
            return 0;
        }
        if (!retry) {
            free(lf);
            return 1;
        }
        struct stat st;
        if (stat(lf, &st) != 0) {
            same = 0;
        } else if (st.st_mtim.tv_sec == seen.tv_sec && st.st_mtim.tv_nsec == seen.tv_nsec) {
            same++;
        } else {
            same = 0;
            seen = st.st_mtim;
        }
        if (same >= 10) {
            remove(lf);
            my_printf("warning: zombie lock released\n");
            sleep_rand(1.0, 2.0);
        }
        sleep_rand(1.0, 3.0);
    }
}

static void unlock_file(const char *file)
{
    char *lf = lock_name(file);
    if (remove(lf) != 0) my_printf("warning: lock release failed (already removed?)\n");
    free(lf);
}

/* remove the last non-empty line of the stack into line: 0 popped, 1 empty,
 * -1 I/O error (stack unchanged) */
static int pop_stack(const char *file, char *line)
{
    line[0] = 0;
    lock_file(file, 1);
    int ret = -1;
    FILE *f = fopen(file, "r+");
    if (!f) {
        unlock_file(file);
        my_printf("error: open() failed in pop_stack\n");
        return -1;
    }
    char *buf = NULL;
    long n = 0;
    if (fseek(f, 0, SEEK_END) == 0 && (n = ftell(f)) >= 0 && fseek(f, 0, SEEK_SET) == 0) {
        buf = malloc(n + 1);
        if (buf && (long)fread(buf, 1, n, f) == n) {
            buf[n] = 0;
            long end = n - 1;
            while (end >= 0 && (buf[end] == '\n' || buf[end] == 0)) end--;
            if (end < 0) {
                ret = 1;
            } else {
                long start = end;
                while (start > 0 && buf[start - 1] != '\n') start--;
                const long len = end - start + 1;
                if (len > MAX_LEN) {
                    my_printf("error: last line length > %d\n", MAX_LEN);
                } else {
                    memcpy(line, buf + start, len);
                    line[len] = 0;
                    ret = ftruncate(fileno(f), start) == 0 ? 0 : -1;
                    if (ret) my_printf("error: ftruncate failed in pop_stack()\n");
                }
            }
        }
    }
    if (ret == -1 && !buf) my_printf("error: read failed in pop_stack()\n");
    free(buf);
    if (fclose(f) != 0) my_printf("error: close() failed in pop_stack()\n");
    unlock_file(file);
    return ret;
}

/* append line; if the first lock attempt fails, the line is first saved to
 * "<stack>_<host>_<pid>" so that it survives if this process is killed while
 * waiting for the lock (src/main_stack.c:229-298) */
static void push_stack(const char *file, const char *line)
{
    char *backup = NULL;
    if (lock_file(file, 0) != 0) {
        const size_t n = strlen(file) + 96;
        backup = calloc(n, 1);
        snprintf(backup, n, "%s_%.20s_%d", file, hostname, pid);
        FILE *b = fopen(backup, "a");
        if (b) {
            fprintf(b, "%s\n", line);
            fclose(b);
        }
        lock_file(file, 1);
    }
    int ok = 0;
    FILE *f = fopen(file, "a");
    if (f) {
        ok = fprintf(f, "%s\n", line) > 0;
        ok = (fclose(f) == 0) && ok;
    }
    unlock_file(file);
    if (!ok) my_printf("error: push_stack() failed for %s\n", line);
    if (backup) {
        if (ok) remove(backup);
        free(backup);
    }
}

static struct argp_option options[] = {
    {0, 't', "TIME", 0, "Max allowed wall time in integer seconds", 0},
    {"dry-run", 'n', 0, 0, "Run through stack, report mem requirement only", 0},
    {"benchmark", 'b', 0, 0, "Don't save sim_data to disk", 0},
    {"profile", 'p', 0, 0, "Print the per-section profile table into the log", 0},
    {"batch", 'B', "N", 0, "Run up to N files together on the GPU (default 64)", 0},
    {0}};

struct arguments {
    long time, batch;
    bool dry_run, bench;
    char *stack;
};

static error_t parse_opt(int key, char *arg, struct argp_state *state)
{
    struct arguments *a = state->input;
    switch (key) {
    case 't': a->time = atol(arg); break;
    case 'B': a->batch = atol(arg); break;
    case 'n': a->dry_run = true; break;
    case 'b': a->bench = true; break;
    case 'p': dq_set_profile(true); break;
    case ARGP_KEY_ARG:
        if (state->arg_num >= 1) argp_usage(state);
        a->stack = arg;
        break;
    case ARGP_KEY_END:
        if (state->arg_num < 1) argp_usage(state);
        break;
    default: return ARGP_ERR_UNKNOWN;
    }
    return 0;
}

static struct argp argp = {options, parse_opt, "stack_file",
                           "Complex-capable DQMC, stack version (MI355X engine)", 0, 0, 0};

int main(int argc, char **argv)
{
    struct arguments a = {0, 64, false, false, NULL};
    argp_parse(&argp, argc, argv, 0, 0, &a);
    if (a.batch < 1) a.batch = 1;
    gethostname(hostname, 64);
    pid = getpid();
    srand((unsigned)pid);
    const int64_t t_start = dq_time_ns();
    const int64_t t_stop = t_start + (int64_t)a.time * 1000000000LL;
    if (access(a.stack, R_OK | W_OK) != 0) {
        my_printf("[ERROR] %s does not exist or you don't have access; idling\n", a.stack);
        return EXIT_FAILURE;
    }
    char **files = calloc(a.batch, sizeof *files);
    char **logs = calloc(a.batch, sizeof *logs);
    int *codes = calloc(a.batch, sizeof *codes);
    for (;;) {
        int64_t t_remain = 0;
        if (a.time > 0) {
            t_remain = t_stop - dq_time_ns();
            if (t_remain <= 0) {
                my_printf("dqmc_stack reached wall time limit; idling\n");
                break;
            }
        }
        int nf = 0, err = 0;
        char line[MAX_LEN + 1];
        while (nf < a.batch) {
            const int ps = pop_stack(a.stack, line);
            if (ps == -1) {
                my_printf("[ERROR] pop_stack() failed; idling\n");
                err = 1;
                break;
            }
            if (ps == 1) break;
            files[nf] = strdup(line);
            logs[nf] = malloc(strlen(line) + 5);
            sprintf(logs[nf], "%s.log", line);
            my_printf("starting:   %s\n", files[nf]);
            my_printf("logging to: %s\n", logs[nf]);
            nf++;
        }
        if (nf == 0) {
            if (!err) my_printf("pop_stack() returned 1, %s empty; idling\n", a.stack);
            break;
        }
        dqmc_wrapper_batch(nf, (const char *const *)files, (const char *const *)logs, t_remain,
                           a.dry_run, a.bench, codes);
        int incomplete = 0;
        for (int i = 0; i < nf; i++) {
            if (codes[i] > 0) {
                push_stack(a.stack, files[i]);
                my_printf("dqmc_wrapper() incomplete, pushed: %s\n", files[i]);
                incomplete = 1;
            } else if (codes[i] == 0) {
                my_printf("completed: %s\n", files[i]);
            } else {
                my_printf("dqmc_wrapper() failed: %s\n", files[i]);
            }
            free(files[i]);
            free(logs[i]);
        }
        if (incomplete || err) break;
    }
    free(files);
    free(logs);
    free(codes);
    return EXIT_SUCCESS;
}
