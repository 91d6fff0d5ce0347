/*
 * dq_h5.h -- HDF5 side of the drop-in: the simulation-file layout written by
 * util/gen_1band_unified_hub.py:929-1160 and read/written by
 * src/data.c:345-789, plus a small generic dataset API that the Python
 * generator and readers use through ctypes.
 *
 * Error codes follow src/data.c:10-13: -1 open, -2 corrupt (partial write),
 * -3 read/write, -4 close.
 */
#ifndef DQ_H5_H
#define DQ_H5_H

#include <stdint.h>
#include <stdio.h>

#include "dqmc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define DQH5_OPEN_FAIL (-1)
#define DQH5_CORRUPT_FAIL (-2)
#define DQH5_RW_FAIL (-3)
#define DQH5_CLOSE_FAIL (-4)

/* element kinds of the generic API */
enum { DQH5_I32 = 1, DQH5_I64 = 2, DQH5_U64 = 3, DQH5_F64 = 4, DQH5_C128 = 5, DQH5_STR = 6 };

/* generic API (returns >= 0 on success) */
int64_t dqh5_create(const char *path);           /* truncating create; file id */
int64_t dqh5_open(const char *path, int rw);     /* file id */
int dqh5_close(int64_t fid);
/* write a dataset, creating missing groups; STR: data is a NUL-terminated
 * string written as a variable-length UTF-8 scalar (h5py's str) */
int dqh5_write(int64_t fid, const char *path, int kind, int ndim, const int64_t *dims,
               const void *data);
/* overwrite an existing dataset's contents (same element count) */
int dqh5_overwrite(int64_t fid, const char *path, int kind, const void *data);
/* external link path -> target_file:target_obj (h5py.ExternalLink) */
int dqh5_external_link(int64_t fid, const char *path, const char *target_file,
                       const char *target_obj);
/* kind (class of the stored type), rank and dims (<= 8) of a dataset; for
 * STR the string length without NUL is returned in dims[0] */
int dqh5_info(int64_t fid, const char *path, int *kind, int *ndim, int64_t *dims);
int dqh5_read(int64_t fid, const char *path, int kind, void *data, int64_t cap_bytes);
/* member names of a group, '\n'-separated, NUL-terminated */
int dqh5_list(int64_t fid, const char *group, char *names, int64_t cap);

/* one simulation file in memory (struct sim_data, src/data.h:98-104) */
typedef struct dq_sim {
    char *file;
    dqmc_params p;
    uint64_t rng[17];
    int sweep;
    int32_t *hs;
    dqmc_meas meq, mue;
    /* owned storage behind the pointers above */
    void *own[128];
    int n_own;
} dq_sim;

/* sim_data_read_alloc (src/data.c:345-658): refuses a file whose last
 * checkpoint was a partial write; the element type (real / complex) is taken
 * from the file. */
int dq_sim_read(const char *file, dq_sim *s);
/* sim_data_save (src/data.c:668-789), with the partial_write protocol */
int dq_sim_save(const dq_sim *s);
void dq_sim_free(dq_sim *s);
/* get_memory_req (src/data.c:121-216) as a 64-bit byte count (the reference
 * returns int, which overflows above 2 GiB); < 0 on error */
int64_t dq_sim_memory_req(const char *file);
/* /metadata/commit of the file (consistency_check, src/data.c:69-99); writes
 * at most cap bytes; 0 on success */
int dq_sim_commit(const char *file, char *buf, int cap);

#ifdef __cplusplus
}
#endif
#endif
