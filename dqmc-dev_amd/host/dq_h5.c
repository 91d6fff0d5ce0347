/*
 * dq_h5.c -- HDF5 I/O of simulation files (see dq_h5.h).
 *
 * The dataset list, element types and the partial_write protocol are those of
 * src/data.c:345-789; the files themselves are the ones
 * util/gen_1band_unified_hub.py:929-1160 writes (h5py conventions: Python
 * ints as int64 scalars, numpy int32 arrays, complex128 as the compound
 * {"r" f64 @0, "i" f64 @8}, str as variable-length UTF-8).
 */
#include "dq_h5.h"

#include <hdf5.h>
#include <stdlib.h>
#include <string.h>

#define ERR(code, ...)                    \
    do {                                  \
        fprintf(stderr, __VA_ARGS__);     \
        return (code);                    \
    } while (0)

static hid_t complex_mem(void)
{
    static hid_t t = -1;
    if (t < 0) {
        t = H5Tcreate(H5T_COMPOUND, 16);
        H5Tinsert(t, "r", 0, H5T_NATIVE_DOUBLE);
        H5Tinsert(t, "i", 8, H5T_NATIVE_DOUBLE);
    }
    return t;
}

static hid_t complex_file(void)
{
    static hid_t t = -1;
    if (t < 0) {
        t = H5Tcreate(H5T_COMPOUND, 16);
        H5Tinsert(t, "r", 0, H5T_IEEE_F64LE);
        H5Tinsert(t, "i", 8, H5T_IEEE_F64LE);
    }
    return t;
}

static hid_t mem_type(int kind)
{
    switch (kind) {
    case DQH5_I32: return H5T_NATIVE_INT32;
    case DQH5_I64: return H5T_NATIVE_INT64;
    case DQH5_U64: return H5T_NATIVE_UINT64;
    case DQH5_F64: return H5T_NATIVE_DOUBLE;
    case DQH5_C128: return complex_mem();
    }
    return -1;
}

static hid_t file_type(int kind)
{
    switch (kind) {
    case DQH5_I32: return H5T_STD_I32LE;
    case DQH5_I64: return H5T_STD_I64LE;
    case DQH5_U64: return H5T_STD_U64LE;
    case DQH5_F64: return H5T_IEEE_F64LE;
    case DQH5_C128: return complex_file();
    }
    return -1;
}

static size_t kind_size(int kind)
{
    switch (kind) {
    case DQH5_I32: return 4;
    case DQH5_C128: return 16;
    default: return 8;
    }
}

static hid_t vlen_str(void)
{
    hid_t t = H5Tcopy(H5T_C_S1);
    H5Tset_size(t, H5T_VARIABLE);
    H5Tset_strpad(t, H5T_STR_NULLTERM);
    H5Tset_cset(t, H5T_CSET_UTF8);
    return t;
}

/* ---------------------------------------------------------------- generic */

int64_t dqh5_create(const char *path)
{
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    return f < 0 ? DQH5_OPEN_FAIL : (int64_t)f;
}

int64_t dqh5_open(const char *path, int rw)
{
    hid_t f = H5Fopen(path, rw ? H5F_ACC_RDWR : H5F_ACC_RDONLY, H5P_DEFAULT);
    return f < 0 ? DQH5_OPEN_FAIL : (int64_t)f;
}

int dqh5_close(int64_t fid) { return H5Fclose((hid_t)fid) < 0 ? DQH5_CLOSE_FAIL : 0; }

/* create every missing group on the way to path's last component */
static int ensure_groups(hid_t fid, const char *path)
{
    char buf[512];
    const size_t n = strlen(path);
    if (n >= sizeof buf) return DQH5_RW_FAIL;
    memcpy(buf, path, n + 1);
    for (size_t i = 1; i < n; i++) {
        if (buf[i] != '/') continue;
        buf[i] = 0;
        htri_t ex = H5Lexists(fid, buf, H5P_DEFAULT);
        if (ex < 0) return DQH5_RW_FAIL;
        if (ex == 0) {
            hid_t g = H5Gcreate2(fid, buf, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
            if (g < 0) return DQH5_RW_FAIL;
            H5Gclose(g);
        }
        buf[i] = '/';
    }
    return 0;
}

int dqh5_write(int64_t fid_, const char *path, int kind, int ndim, const int64_t *dims,
               const void *data)
{
    const hid_t fid = (hid_t)fid_;
    if (ensure_groups(fid, path)) ERR(DQH5_RW_FAIL, "dqh5_write: cannot create groups for %s\n", path);
    hid_t space, ftype, mtype;
    int own_type = 0;
    if (kind == DQH5_STR) {
        space = H5Screate(H5S_SCALAR);
        ftype = mtype = vlen_str();
        own_type = 1;
    } else {
        if (ndim < 0 || ndim > 8) return DQH5_RW_FAIL;
        hsize_t hd[8];
        for (int i = 0; i < ndim; i++) hd[i] = (hsize_t)dims[i];
        space = ndim == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(ndim, hd, NULL);
        ftype = file_type(kind);
        mtype = mem_type(kind);
        if (ftype < 0) return DQH5_RW_FAIL;
    }
    hid_t d = H5Dcreate2(fid, path, ftype, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    herr_t st = -1;
    if (d >= 0) {
        if (kind == DQH5_STR) {
            const char *s = (const char *)data;
            st = H5Dwrite(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, &s);
        } else {
            st = H5Dwrite(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
        }
        H5Dclose(d);
    }
    H5Sclose(space);
    if (own_type) H5Tclose(ftype);
    if (d < 0 || st < 0) ERR(DQH5_RW_FAIL, "dqh5_write failed for %s\n", path);
    return 0;
}

int dqh5_overwrite(int64_t fid, const char *path, int kind, const void *data)
{
    hid_t d = H5Dopen2((hid_t)fid, path, H5P_DEFAULT);
    if (d < 0) ERR(DQH5_RW_FAIL, "H5Dopen2() failed for %s\n", path);
    herr_t st = H5Dwrite(d, mem_type(kind), H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Dclose(d);
    if (st < 0) ERR(DQH5_RW_FAIL, "H5Dwrite() failed for %s\n", path);
    return 0;
}

int dqh5_external_link(int64_t fid, const char *path, const char *target_file,
                       const char *target_obj)
{
    if (ensure_groups((hid_t)fid, path)) return DQH5_RW_FAIL;
    if (H5Lcreate_external(target_file, target_obj, (hid_t)fid, path, H5P_DEFAULT,
                           H5P_DEFAULT) < 0)
        ERR(DQH5_RW_FAIL, "H5Lcreate_external failed for %s\n", path);
    return 0;
}

static int class_kind(hid_t t)
{
    switch (H5Tget_class(t)) {
    case H5T_INTEGER:
        if (H5Tget_size(t) == 4) return DQH5_I32;
        return H5Tget_sign(t) == H5T_SGN_NONE ? DQH5_U64 : DQH5_I64;
    case H5T_FLOAT: return DQH5_F64;
    case H5T_COMPOUND: return DQH5_C128;
    case H5T_STRING: return DQH5_STR;
    default: return -1;
    }
}

int dqh5_info(int64_t fid, const char *path, int *kind, int *ndim, int64_t *dims)
{
    hid_t d = H5Dopen2((hid_t)fid, path, H5P_DEFAULT);
    if (d < 0) return DQH5_RW_FAIL;
    hid_t t = H5Dget_type(d), s = H5Dget_space(d);
    const int k = class_kind(t);
    int nd = H5Sget_simple_extent_ndims(s);
    hsize_t hd[8] = {0};
    if (nd > 8) nd = -1;
    if (nd > 0) H5Sget_simple_extent_dims(s, hd, NULL);
    if (kind) *kind = k;
    if (ndim) *ndim = nd;
    if (dims && nd >= 0)
        for (int i = 0; i < nd; i++) dims[i] = (int64_t)hd[i];
    if (k == DQH5_STR && dims) {
        char *str = NULL;
        hid_t mt = vlen_str();
        if (H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, &str) >= 0 && str) {
            dims[0] = (int64_t)strlen(str);
            H5free_memory(str);
        }
        H5Tclose(mt);
    }
    H5Tclose(t);
    H5Sclose(s);
    H5Dclose(d);
    return (k < 0 || nd < 0) ? DQH5_RW_FAIL : 0;
}

int dqh5_read(int64_t fid, const char *path, int kind, void *data, int64_t cap_bytes)
{
    hid_t d = H5Dopen2((hid_t)fid, path, H5P_DEFAULT);
    if (d < 0) ERR(DQH5_RW_FAIL, "H5Dopen2() failed for %s\n", path);
    int rc = 0;
    if (kind == DQH5_STR) {
        char *str = NULL;
        hid_t mt = vlen_str();
        if (H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, &str) < 0 || !str) {
            rc = DQH5_RW_FAIL;
        } else {
            const size_t n = strlen(str);
            if ((int64_t)n + 1 > cap_bytes) rc = DQH5_RW_FAIL;
            else memcpy(data, str, n + 1);
            H5free_memory(str);
        }
        H5Tclose(mt);
    } else {
        hid_t s = H5Dget_space(d);
        const hssize_t n = H5Sget_simple_extent_npoints(s);
        H5Sclose(s);
        if (n < 0 || (int64_t)n * (int64_t)kind_size(kind) > cap_bytes) rc = DQH5_RW_FAIL;
        else if (H5Dread(d, mem_type(kind), H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
            rc = DQH5_RW_FAIL;
    }
    H5Dclose(d);
    if (rc) fprintf(stderr, "dqh5_read failed for %s\n", path);
    return rc;
}

struct list_ctx {
    char *buf;
    int64_t cap, used;
    int overflow;
};

static herr_t list_cb(hid_t g, const char *name, const H5L_info_t *info, void *op)
{
    (void)g; (void)info;
    struct list_ctx *c = (struct list_ctx *)op;
    const int64_t n = (int64_t)strlen(name);
    if (c->used + n + 2 > c->cap) {
        c->overflow = 1;
        return 1;
    }
    memcpy(c->buf + c->used, name, n);
    c->used += n;
    c->buf[c->used++] = '\n';
    c->buf[c->used] = 0;
    return 0;
}

int dqh5_list(int64_t fid, const char *group, char *names, int64_t cap)
{
    if (cap < 1) return DQH5_RW_FAIL;
    names[0] = 0;
    hid_t g = H5Gopen2((hid_t)fid, group, H5P_DEFAULT);
    if (g < 0) return DQH5_RW_FAIL;
    struct list_ctx c = {names, cap, 0, 0};
    hsize_t idx = 0;
    H5Literate(g, H5_INDEX_NAME, H5_ITER_INC, &idx, list_cb, &c);
    H5Gclose(g);
    return c.overflow ? DQH5_RW_FAIL : 0;
}

/* ------------------------------------------------------------ simulations */

/* accumulator names in the order of enum dqmc_eq_acc / dqmc_ue_acc */
static const char *const eq_names[DQ_EQ_COUNT] = {
    "density", "density_u", "density_d", "double_occ", "g00", "g00_u", "g00_d", "nn",
    "xx", "zz", "pair_sw", "uuuu", "dddd", "uudd", "dduu", "chi",
    "kk", "kv", "kn", "vv", "vn", "j", "j2", "jn"};
static const char *const ue_names[DQ_UE_COUNT] = {
    "gt0", "gt0_u", "gt0_d", "nn", "xx", "zz", "pair_sw", "uuuu", "dddd", "uudd", "dduu",
    "pair_bb", "jj", "jsjs", "kk", "ksks", "pair_b2b2", "j2jn", "jnj2", "jnjn", "j2j", "jj2",
    "j2j2", "js2js2", "k2k2", "ks2ks2", "jnj", "jjn", "kv", "kn", "vv", "vn",
    "nem_nnnn", "nem_ssss"};

/* accumulator lengths as allocated by src/data.c:441-517 (0 = not present) */
static long eq_len(const dqmc_params *p, int k)
{
    const long ni = p->num_i, nij = p->num_ij;
    switch (k) {
    case DQ_EQ_density: case DQ_EQ_density_u: case DQ_EQ_density_d: case DQ_EQ_double_occ:
        return ni;
    case DQ_EQ_g00: case DQ_EQ_g00_u: case DQ_EQ_g00_d: case DQ_EQ_nn: case DQ_EQ_xx:
    case DQ_EQ_zz: case DQ_EQ_pair_sw:
        return nij;
    case DQ_EQ_uuuu: case DQ_EQ_dddd: case DQ_EQ_uudd: case DQ_EQ_dduu:
        return p->meas_gen_suscept ? nij * nij : 0;
    case DQ_EQ_chi: return p->meas_chiral ? p->num_plaq_accum : 0;
    case DQ_EQ_kk: return p->meas_energy_corr ? p->num_bb : 0;
    case DQ_EQ_kv: case DQ_EQ_kn: return p->meas_energy_corr ? p->num_bs : 0;
    case DQ_EQ_vv: case DQ_EQ_vn: return p->meas_energy_corr ? nij : 0;
    case DQ_EQ_j: case DQ_EQ_jn: return p->meas_local_JQ ? p->num_b_accum : 0;
    case DQ_EQ_j2: return p->meas_local_JQ ? p->num_b2_accum : 0;
    }
    return 0;
}

static long ue_len(const dqmc_params *p, int k)
{
    if (p->period_uneqlt <= 0) return 0;
    const long L = p->L, nij = p->num_ij;
    switch (k) {
    case DQ_UE_gt0: case DQ_UE_gt0_u: case DQ_UE_gt0_d: case DQ_UE_nn: case DQ_UE_xx:
    case DQ_UE_zz: case DQ_UE_pair_sw:
        return nij * L;
    case DQ_UE_uuuu: case DQ_UE_dddd: case DQ_UE_uudd: case DQ_UE_dduu:
        return p->meas_gen_suscept ? nij * nij * L : 0;
    case DQ_UE_pair_bb:
        return (p->meas_bond_corr || p->meas_pair_bb_only) ? (long)p->num_bb * L : 0;
    case DQ_UE_jj: case DQ_UE_jsjs: case DQ_UE_kk: case DQ_UE_ksks:
        return p->meas_bond_corr ? (long)p->num_bb * L : 0;
    case DQ_UE_jnjn: case DQ_UE_jnj: case DQ_UE_jjn:
        return p->meas_thermal ? (long)p->num_bb * L : 0;
    case DQ_UE_j2jn: return p->meas_thermal ? (long)p->num_b2b * L : 0;
    case DQ_UE_jnj2: return p->meas_thermal ? (long)p->num_bb2 * L : 0;
    case DQ_UE_pair_b2b2: case DQ_UE_j2j2: case DQ_UE_js2js2: case DQ_UE_k2k2:
    case DQ_UE_ks2ks2:
        return p->meas_2bond_corr ? (long)p->num_b2b2 * L : 0;
    case DQ_UE_j2j: return p->meas_2bond_corr ? (long)p->num_b2b * L : 0;
    case DQ_UE_jj2: return p->meas_2bond_corr ? (long)p->num_bb2 * L : 0;
    case DQ_UE_kv: case DQ_UE_kn: return p->meas_energy_corr ? (long)p->num_bs * L : 0;
    case DQ_UE_vv: case DQ_UE_vn: return p->meas_energy_corr ? nij * L : 0;
    case DQ_UE_nem_nnnn: case DQ_UE_nem_ssss:
        return p->meas_nematic_corr ? (long)p->num_bb * L : 0;
    }
    return 0;
}

static void *own(dq_sim *s, size_t bytes)
{
    if (s->n_own >= (int)(sizeof s->own / sizeof s->own[0])) return NULL;
    void *p = calloc(bytes ? bytes : 1, 1);
    if (p) s->own[s->n_own++] = p;
    return p;
}

static int rd(hid_t f, const char *path, hid_t mt, void *buf, size_t n_expect)
{
    hid_t d = H5Dopen2(f, path, H5P_DEFAULT);
    if (d < 0) ERR(DQH5_RW_FAIL, "H5Dopen2() failed for %s\n", path);
    hid_t sp = H5Dget_space(d);
    const hssize_t n = H5Sget_simple_extent_npoints(sp);
    H5Sclose(sp);
    herr_t st = -1;
    if (n == (hssize_t)n_expect) st = H5Dread(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
    H5Dclose(d);
    if (n != (hssize_t)n_expect)
        ERR(DQH5_RW_FAIL, "%s has %lld elements, expected %zu\n", path, (long long)n, n_expect);
    if (st < 0) ERR(DQH5_RW_FAIL, "H5Dread() failed for %s\n", path);
    return 0;
}

#define RD(path, mt, buf, n) \
    do { if ((rc = rd(f, (path), (mt), (buf), (n))) != 0) goto done; } while (0)
#define RDI(path, dst) RD(path, H5T_NATIVE_INT, &(dst), 1)

int dq_sim_read(const char *file, dq_sim *s)
{
    memset(s, 0, sizeof *s);
    s->file = strdup(file);
    const hid_t f = H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) ERR(DQH5_OPEN_FAIL, "H5Fopen() failed for %s\n", file);
    int rc = 0, partial = 0;
    dqmc_params *p = &s->p;
    RDI("/state/partial_write", partial);
    if (partial) {
        fprintf(stderr, "Last checkpoint in %s contained partial write\n", file);
        rc = DQH5_CORRUPT_FAIL;
        goto done;
    }
    {
        /* element type from the file: the complex build stores compound num */
        hid_t d = H5Dopen2(f, "/params/exp_Ku", H5P_DEFAULT);
        if (d < 0) { rc = DQH5_RW_FAIL; fprintf(stderr, "no /params/exp_Ku in %s\n", file); goto done; }
        hid_t t = H5Dget_type(d);
        p->is_complex = (H5Tget_class(t) == H5T_COMPOUND);
        H5Tclose(t);
        H5Dclose(d);
    }
    RDI("/params/N", p->N); RDI("/params/L", p->L);
    RDI("/params/Nx", p->Nx); RDI("/params/Ny", p->Ny);
    RDI("/params/num_i", p->num_i); RDI("/params/num_ij", p->num_ij);
    RDI("/params/num_plaq_accum", p->num_plaq_accum);
    RDI("/params/num_b_accum", p->num_b_accum); RDI("/params/num_b2_accum", p->num_b2_accum);
    RDI("/params/num_plaq", p->num_plaq); RDI("/params/num_b", p->num_b);
    RDI("/params/num_b2", p->num_b2); RDI("/params/num_bs", p->num_bs);
    RDI("/params/num_bb", p->num_bb); RDI("/params/num_b2b", p->num_b2b);
    RDI("/params/num_bb2", p->num_bb2); RDI("/params/num_b2b2", p->num_b2b2);
    RDI("/params/degen_i", p->degen_i); RDI("/params/degen_ij", p->degen_ij);
    RDI("/params/degen_plaq", p->degen_plaq); RDI("/params/degen_b", p->degen_b);
    RDI("/params/degen_b2", p->degen_b2); RDI("/params/degen_bs", p->degen_bs);
    RDI("/params/degen_bb", p->degen_bb); RDI("/params/degen_b2b", p->degen_b2b);
    RDI("/params/degen_bb2", p->degen_bb2); RDI("/params/degen_b2b2", p->degen_b2b2);
    RDI("/params/period_uneqlt", p->period_uneqlt);
    RDI("/params/meas_bond_corr", p->meas_bond_corr);
    RDI("/params/meas_thermal", p->meas_thermal);
    RDI("/params/meas_2bond_corr", p->meas_2bond_corr);
    RDI("/params/meas_local_JQ", p->meas_local_JQ);
    RDI("/params/meas_gen_suscept", p->meas_gen_suscept);
    RDI("/params/meas_energy_corr", p->meas_energy_corr);
    RDI("/params/meas_nematic_corr", p->meas_nematic_corr);
    RDI("/params/meas_pair_bb_only", p->meas_pair_bb_only);
    RDI("/params/meas_chiral", p->meas_chiral);
    RDI("/params/checkpoint_every", p->checkpoint_every);
    RDI("/params/n_matmul", p->n_matmul); RDI("/params/n_delay", p->n_delay);
    RDI("/params/n_sweep_warm", p->n_sweep_warm); RDI("/params/n_sweep_meas", p->n_sweep_meas);
    RDI("/params/period_eqlt", p->period_eqlt);
    RDI("/params/F", p->F); RDI("/params/n_sweep", p->n_sweep);
    {
        const long N = p->N, L = p->L, nb = p->num_b, nb2 = p->num_b2, npl = p->num_plaq;
        if (N <= 0 || L <= 0 || nb < 0 || nb2 < 0 || npl < 0) {
            fprintf(stderr, "bad sizes in %s\n", file);
            rc = DQH5_RW_FAIL;
            goto done;
        }
        const size_t es = p->is_complex ? 16 : 8;
        const hid_t nt = p->is_complex ? complex_mem() : H5T_NATIVE_DOUBLE;
        int32_t *map_i = own(s, N * 4), *map_ij = own(s, N * N * 4);
        int32_t *bonds = own(s, nb * 2 * 4), *bond2s = own(s, nb2 * 2 * 4);
        int32_t *plaqs = own(s, npl * 3 * 4), *map_plaq = own(s, npl * 4);
        int32_t *map_b = own(s, nb * 4), *map_b2 = own(s, nb2 * 4);
        int32_t *map_bs = own(s, nb * N * 4), *map_bb = own(s, nb * nb * 4);
        int32_t *map_b2b2 = own(s, nb2 * nb2 * 4), *map_b2b = own(s, nb2 * nb * 4);
        int32_t *map_bb2 = own(s, nb * nb2 * 4);
        double *peu = own(s, N * N * es), *ped = own(s, N * N * es);
        double *ppu = own(s, nb2 * es), *ppd = own(s, nb2 * es);
        double *ppru = own(s, nb2 * es), *pprd = own(s, nb2 * es);
        double *eK[8];
        for (int i = 0; i < 8; i++) eK[i] = own(s, N * N * es);
        double *el = own(s, 2 * N * 8), *de = own(s, 2 * N * 8);
        s->hs = own(s, N * L * 4);
        if (!s->hs || !de) { rc = DQH5_RW_FAIL; fprintf(stderr, "out of memory\n"); goto done; }
        RD("/params/map_i", H5T_NATIVE_INT32, map_i, N);
        RD("/params/map_ij", H5T_NATIVE_INT32, map_ij, N * N);
        RD("/params/map_plaq", H5T_NATIVE_INT32, map_plaq, npl);
        RD("/params/plaqs", H5T_NATIVE_INT32, plaqs, npl * 3);
        RD("/params/bonds", H5T_NATIVE_INT32, bonds, nb * 2);
        RD("/params/bond2s", H5T_NATIVE_INT32, bond2s, nb2 * 2);
        RD("/params/map_bs", H5T_NATIVE_INT32, map_bs, nb * N);
        RD("/params/map_bb", H5T_NATIVE_INT32, map_bb, nb * nb);
        RD("/params/map_b", H5T_NATIVE_INT32, map_b, nb);
        RD("/params/map_b2", H5T_NATIVE_INT32, map_b2, nb2);
        RD("/params/map_bb2", H5T_NATIVE_INT32, map_bb2, nb * nb2);
        RD("/params/map_b2b", H5T_NATIVE_INT32, map_b2b, nb2 * nb);
        RD("/params/map_b2b2", H5T_NATIVE_INT32, map_b2b2, nb2 * nb2);
        RD("/params/peierlsu", nt, peu, N * N);
        RD("/params/peierlsd", nt, ped, N * N);
        RD("/params/pp_u", nt, ppu, nb2);
        RD("/params/pp_d", nt, ppd, nb2);
        RD("/params/ppr_u", nt, ppru, nb2);
        RD("/params/ppr_d", nt, pprd, nb2);
        static const char *const kn[8] = {"/params/exp_Ku", "/params/exp_Kd",
                                          "/params/inv_exp_Ku", "/params/inv_exp_Kd",
                                          "/params/exp_halfKu", "/params/exp_halfKd",
                                          "/params/inv_exp_halfKu", "/params/inv_exp_halfKd"};
        for (int i = 0; i < 8; i++) RD(kn[i], nt, eK[i], N * N);
        RD("/params/exp_lambda", H5T_NATIVE_DOUBLE, el, 2 * N);
        RD("/params/del", H5T_NATIVE_DOUBLE, de, 2 * N);
        RD("/state/rng", H5T_NATIVE_UINT64, s->rng, 17);
        RDI("/state/sweep", s->sweep);
        RD("/state/hs", H5T_NATIVE_INT32, s->hs, N * L);
        p->map_i = map_i; p->map_ij = map_ij; p->bonds = bonds; p->bond2s = bond2s;
        p->plaqs = plaqs; p->map_plaq = map_plaq; p->map_b = map_b; p->map_b2 = map_b2;
        p->map_bs = map_bs; p->map_bb = map_bb; p->map_b2b2 = map_b2b2; p->map_b2b = map_b2b;
        p->map_bb2 = map_bb2;
        p->peierlsu = peu; p->peierlsd = ped; p->pp_u = ppu; p->pp_d = ppd;
        p->ppr_u = ppru; p->ppr_d = pprd;
        p->exp_Ku = eK[0]; p->exp_Kd = eK[1]; p->inv_exp_Ku = eK[2]; p->inv_exp_Kd = eK[3];
        p->exp_halfKu = eK[4]; p->exp_halfKd = eK[5]; p->inv_exp_halfKu = eK[6];
        p->inv_exp_halfKd = eK[7];
        p->exp_lambda = el; p->del = de;

        char path[128];
        RDI("/meas_eqlt/n_sample", s->meq.n_sample);
        RD("/meas_eqlt/sign", nt, s->meq.sign, 1);
        for (int k = 0; k < DQ_EQ_COUNT; k++) {
            const long n = eq_len(p, k);
            if (n <= 0) continue;
            s->meq.acc[k] = own(s, n * es);
            if (!s->meq.acc[k]) { rc = DQH5_RW_FAIL; goto done; }
            snprintf(path, sizeof path, "/meas_eqlt/%s", eq_names[k]);
            RD(path, nt, s->meq.acc[k], n);
        }
        if (p->period_uneqlt > 0) {
            RDI("/meas_uneqlt/n_sample", s->mue.n_sample);
            RD("/meas_uneqlt/sign", nt, s->mue.sign, 1);
            for (int k = 0; k < DQ_UE_COUNT; k++) {
                const long n = ue_len(p, k);
                if (n <= 0) continue;
                s->mue.acc[k] = own(s, n * es);
                if (!s->mue.acc[k]) { rc = DQH5_RW_FAIL; goto done; }
                snprintf(path, sizeof path, "/meas_uneqlt/%s", ue_names[k]);
                RD(path, nt, s->mue.acc[k], n);
            }
        }
    }
done:
    if (H5Fclose(f) < 0 && rc == 0) rc = DQH5_CLOSE_FAIL;
    return rc;
}

static int wr(hid_t f, const char *path, hid_t mt, const void *buf)
{
    hid_t d = H5Dopen2(f, path, H5P_DEFAULT);
    if (d < 0) ERR(DQH5_RW_FAIL, "H5Dopen2() failed for %s\n", path);
    herr_t st = H5Dwrite(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
    herr_t sc = H5Dclose(d);
    if (st < 0) ERR(DQH5_RW_FAIL, "H5Dwrite() failed for %s\n", path);
    if (sc < 0) ERR(DQH5_RW_FAIL, "H5Dclose() failed for %s\n", path);
    return 0;
}

#define WR(path, mt, buf) \
    do { if ((rc = wr(f, (path), (mt), (buf))) != 0) goto done; } while (0)

int dq_sim_save(const dq_sim *s)
{
    const hid_t f = H5Fopen(s->file, H5F_ACC_RDWR, H5P_DEFAULT);
    if (f < 0) ERR(DQH5_OPEN_FAIL, "H5Fopen() failed for %s\n", s->file);
    const dqmc_params *p = &s->p;
    const hid_t nt = p->is_complex ? complex_mem() : H5T_NATIVE_DOUBLE;
    int rc = 0, partial = 0;
    if ((rc = rd(f, "/state/partial_write", H5T_NATIVE_INT, &partial, 1)) != 0) goto done;
    if (partial) {
        fprintf(stderr, "Last checkpoint in %s contained partial write\n", s->file);
        rc = DQH5_CORRUPT_FAIL;
        goto done;
    }
    partial = 1;
    WR("/state/partial_write", H5T_NATIVE_INT, &partial);
    WR("/state/rng", H5T_NATIVE_UINT64, s->rng);
    WR("/state/sweep", H5T_NATIVE_INT, &s->sweep);
    WR("/state/hs", H5T_NATIVE_INT32, s->hs);
    {
        char path[128];
        WR("/meas_eqlt/n_sample", H5T_NATIVE_INT, &s->meq.n_sample);
        WR("/meas_eqlt/sign", nt, s->meq.sign);
        for (int k = 0; k < DQ_EQ_COUNT; k++) {
            if (eq_len(p, k) <= 0) continue;
            snprintf(path, sizeof path, "/meas_eqlt/%s", eq_names[k]);
            WR(path, nt, s->meq.acc[k]);
        }
        if (p->period_uneqlt > 0) {
            WR("/meas_uneqlt/n_sample", H5T_NATIVE_INT, &s->mue.n_sample);
            WR("/meas_uneqlt/sign", nt, s->mue.sign);
            for (int k = 0; k < DQ_UE_COUNT; k++) {
                if (ue_len(p, k) <= 0) continue;
                snprintf(path, sizeof path, "/meas_uneqlt/%s", ue_names[k]);
                WR(path, nt, s->mue.acc[k]);
            }
        }
    }
    partial = 0;
    WR("/state/partial_write", H5T_NATIVE_INT, &partial);
done:
    if (H5Fclose(f) < 0 && rc == 0) ERR(DQH5_CLOSE_FAIL, "H5Fclose() failed for %s\n", s->file);
    return rc;
}

void dq_sim_free(dq_sim *s)
{
    for (int i = 0; i < s->n_own; i++) free(s->own[i]);
    free(s->file);
    memset(s, 0, sizeof *s);
}

int64_t dq_sim_memory_req(const char *file)
{
    const hid_t f = H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) ERR(DQH5_OPEN_FAIL, "H5Fopen() failed for %s\n", file);
    dqmc_params q;
    memset(&q, 0, sizeof q);
    dqmc_params *p = &q;
    int rc = 0;
    RDI("/params/N", p->N); RDI("/params/L", p->L);
    RDI("/params/num_i", p->num_i); RDI("/params/num_ij", p->num_ij);
    RDI("/params/num_plaq_accum", p->num_plaq_accum);
    RDI("/params/num_b_accum", p->num_b_accum); RDI("/params/num_b2_accum", p->num_b2_accum);
    RDI("/params/num_plaq", p->num_plaq); RDI("/params/num_b", p->num_b);
    RDI("/params/num_b2", p->num_b2); RDI("/params/num_bs", p->num_bs);
    RDI("/params/num_bb", p->num_bb); RDI("/params/num_b2b", p->num_b2b);
    RDI("/params/num_bb2", p->num_bb2); RDI("/params/num_b2b2", p->num_b2b2);
    RDI("/params/period_uneqlt", p->period_uneqlt);
    RDI("/params/meas_bond_corr", p->meas_bond_corr);
    RDI("/params/meas_thermal", p->meas_thermal);
    RDI("/params/meas_2bond_corr", p->meas_2bond_corr);
    RDI("/params/meas_energy_corr", p->meas_energy_corr);
    RDI("/params/meas_local_JQ", p->meas_local_JQ);
    RDI("/params/meas_gen_suscept", p->meas_gen_suscept);
    RDI("/params/meas_nematic_corr", p->meas_nematic_corr);
    RDI("/params/meas_pair_bb_only", p->meas_pair_bb_only);
    RDI("/params/meas_chiral", p->meas_chiral);
    RDI("/params/F", p->F);
    {
        hid_t d = H5Dopen2(f, "/params/exp_Ku", H5P_DEFAULT);
        if (d >= 0) {
            hid_t t = H5Dget_type(d);
            p->is_complex = (H5Tget_class(t) == H5T_COMPOUND);
            H5Tclose(t);
            H5Dclose(d);
        }
    }
done:
    H5Fclose(f);
    if (rc) return rc;
    const int64_t es = p->is_complex ? 16 : 8, N = p->N, L = p->L;
    const int64_t nb = p->num_b, nb2 = p->num_b2, npl = p->num_plaq;
    /* host-side sim_data, as src/data.c:152-215 counts it */
    int64_t b = 4 * (N + N * N + 2 * nb + 2 * nb2 + 4 * npl + nb + nb2 + nb * N + nb * nb +
                     nb2 * nb2 + 2 * nb2 * nb + N * L) +
                es * (2 * N * N + 4 * nb2 + 8 * N * N) + 8 * 4 * N;
    for (int k = 0; k < DQ_EQ_COUNT; k++) b += es * eq_len(p, k);
    for (int k = 0; k < DQ_UE_COUNT; k++) b += es * ue_len(p, k);
    return b;
}

int dq_sim_commit(const char *file, char *buf, int cap)
{
    const hid_t f = H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0 || cap < 1) return DQH5_OPEN_FAIL;
    buf[0] = 0;
    int rc = dqh5_read((int64_t)f, "/metadata/commit", DQH5_STR, buf, cap);
    H5Fclose(f);
    return rc;
}
