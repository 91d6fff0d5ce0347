/*
 * dqmc_1 -- run one simulation file on the MI355X engine.
 *
 * Command line and exit status of src/main_1.c:20-113:
 *   dqmc_1 [-t TIME] [-n|--dry-run] [-l|--log LOGFILE] [-b|--benchmark] [-p|--profile] sim_file.h5
 * The exit status is dqmc_wrapper's code (0 done, 1 incomplete, -1 -> 255 failed).
 * TIME is taken as seconds (the reference passes the seconds value where its
 * wrapper expects clock ticks, src/main_1.c:103-104 vs src/sig.c:75).
 */
#include <argp.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>

#include "dq_driver.h"
#include "dqmc_hip.h"

const char *argp_program_version = "dqmc-dev_amd (MI355X HIP engine)";

static struct argp_option options[] = {
    {0, 't', "TIME", 0, "Max allowed wall time in integer seconds", 0},
    {"dry-run", 'n', 0, 0, "Report the memory requirement only", 0},
    {"log", 'l', "LOGFILE", 0, "Redirect stdout to LOGFILE", 0},
    {"benchmark", 'b', 0, 0, "Don't save sim_data to disk", 0},
    {"profile", 'p', 0, 0, "Print the per-section profile table into the log", 0},
    {0}};

struct arguments {
    long time;
    char *log_file;
    bool dry_run, bench;
    char *file;
};

static error_t parse_opt(int key, char *arg, struct argp_state *state)
{
    struct arguments *a = state->input;
    switch (key) {
    case 't': a->time = atol(arg); break;
    case 'l': a->log_file = arg; break;
    case 'n': a->dry_run = true; break;
    case 'b': a->bench = true; break;
    case 'p': dq_set_profile(true); break;
    case ARGP_KEY_ARG:
        if (state->arg_num >= 1) argp_usage(state);
        a->file = arg;
        break;
    case ARGP_KEY_END:
        if (state->arg_num < 1) argp_usage(state);
        break;
    default: return ARGP_ERR_UNKNOWN;
    }
    return 0;
}

static struct argp argp = {options, parse_opt, "sim_file.h5",
                           "Complex-capable DQMC, single file version (MI355X engine)", 0, 0, 0};

int main(int argc, char **argv)
{
    struct arguments a = {0, NULL, false, false, NULL};
    argp_parse(&argp, argc, argv, 0, 0, &a);
    return dqmc_wrapper(a.file, a.log_file, (int64_t)a.time * 1000000000LL, a.dry_run, a.bench);
}
