"""ctypes mirror of include/dqmc_hip.h plus helpers that turn a simulation dict
(the /params, /state, /meas_eqlt, /meas_uneqlt groups of the .h5 layout, as
numpy arrays) into the C structs the engine takes.

The struct layouts follow struct params / meas_eqlt / meas_uneqlt of the
reference (src/data.h:7-105); the accumulator order is that of the
dqmc_eq_acc / dqmc_ue_acc enums."""
import ctypes as C

import numpy as np

EQ_ACC = ["density", "density_u", "density_d", "double_occ", "g00", "g00_u", "g00_d", "nn",
          "xx", "zz", "pair_sw", "uuuu", "dddd", "uudd", "dduu", "chi", "kk", "kv", "kn", "vv",
          "vn", "j", "j2", "jn"]
UE_ACC = ["gt0", "gt0_u", "gt0_d", "nn", "xx", "zz", "pair_sw", "uuuu", "dddd", "uudd", "dduu",
          "pair_bb", "jj", "jsjs", "kk", "ksks", "pair_b2b2", "j2jn", "jnj2", "jnjn", "j2j", "jj2",
          "j2j2", "js2js2", "k2k2", "ks2ks2", "jnj", "jjn", "kv", "kn", "vv", "vn", "nem_nnnn",
          "nem_ssss"]
MAX_ACC = 40

_INT_FIELDS = [
    "N", "L", "Nx", "Ny", "F", "n_matmul", "n_delay", "n_sweep_warm", "n_sweep_meas", "n_sweep",
    "period_eqlt", "period_uneqlt", "checkpoint_every", "meas_bond_corr", "meas_thermal",
    "meas_2bond_corr", "meas_energy_corr", "meas_nematic_corr", "meas_pair_bb_only",
    "meas_chiral", "meas_local_JQ", "meas_gen_suscept", "num_i", "num_ij", "num_plaq_accum",
    "num_b_accum", "num_b2_accum", "num_plaq", "num_b", "num_b2", "num_bs", "num_bb", "num_bb2",
    "num_b2b", "num_b2b2", "degen_i", "degen_ij", "degen_plaq", "degen_b", "degen_b2",
    "degen_bs", "degen_bb", "degen_bb2", "degen_b2b", "degen_b2b2", "is_complex"]
_MAP_FIELDS = ["map_i", "map_ij", "bonds", "bond2s", "map_bs", "map_bb", "map_b2b", "map_bb2",
               "map_b2b2", "plaqs", "map_plaq", "map_b", "map_b2"]
_NUM_FIELDS = ["peierlsu", "peierlsd", "pp_u", "pp_d", "ppr_u", "ppr_d", "exp_Ku", "exp_Kd",
               "inv_exp_Ku", "inv_exp_Kd", "exp_halfKu", "exp_halfKd", "inv_exp_halfKu",
               "inv_exp_halfKd"]
_DBL_FIELDS = ["exp_lambda", "del"]


class DqmcParams(C.Structure):
    _fields_ = ([(f, C.c_int) for f in _INT_FIELDS] +
                [(f, C.POINTER(C.c_int32)) for f in _MAP_FIELDS] +
                [(f, C.POINTER(C.c_double)) for f in _NUM_FIELDS] +
                [(f, C.POINTER(C.c_double)) for f in _DBL_FIELDS])


class DqmcMeas(C.Structure):
    _fields_ = [("n_sample", C.c_int), ("sign", C.c_double * 2),
                ("acc", C.POINTER(C.c_double) * MAX_ACC)]


def is_complex(sim):
    return np.iscomplexobj(sim["params"]["exp_Ku"])


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def make_params(p_dict):
    """DqmcParams from a /params dict; returns (struct, keepalive list).

    Arrays are passed as the raw C-order buffers, i.e. a numpy (N, N) array A is
    seen by the engine as the column-major matrix A^T -- the reading the
    reference's data.c makes of the same .h5 dataset."""
    keep = []
    cplx = np.iscomplexobj(p_dict["exp_Ku"])
    dnum = np.complex128 if cplx else np.float64
    s = DqmcParams()
    for f in _INT_FIELDS:
        if f == "is_complex":
            s.is_complex = int(cplx)
        else:
            s.__setattr__(f, int(np.asarray(p_dict[f]).reshape(())) if f in p_dict else 0)
    for f in _MAP_FIELDS:
        a = np.ascontiguousarray(p_dict[f], dtype=np.int32)
        keep.append(a)
        setattr(s, f, _ptr(a, C.c_int32))
    for f in _NUM_FIELDS:
        a = np.ascontiguousarray(p_dict[f], dtype=dnum)
        keep.append(a)
        setattr(s, f, _ptr(a, C.c_double))
    for f in _DBL_FIELDS:
        a = np.ascontiguousarray(p_dict[f], dtype=np.float64)
        keep.append(a)
        setattr(s, f, _ptr(a, C.c_double))
    return s, keep


def make_meas(group, names, cplx):
    """DqmcMeas pointing into the arrays of one accumulator group (dict of
    numpy arrays).  The arrays are converted in place to the engine dtype and
    are updated by the engine/oracle; returns (struct, arrays dict)."""
    m = DqmcMeas()
    dnum = np.complex128 if cplx else np.float64
    arrays = {}
    if not group:
        return m, arrays
    m.n_sample = int(group["n_sample"])
    sg = complex(np.asarray(group["sign"]).reshape(()))
    m.sign[0] = sg.real
    m.sign[1] = sg.imag
    for k, name in enumerate(names):
        if name in group:
            a = np.ascontiguousarray(group[name], dtype=dnum).copy()
            arrays[name] = a
            m.acc[k] = _ptr(a, C.c_double)
    return m, arrays


def meas_to_dict(m, arrays, cplx):
    out = dict(arrays)
    out["n_sample"] = np.array(m.n_sample, dtype=np.int32)
    out["sign"] = np.array(complex(m.sign[0], m.sign[1]) if cplx else m.sign[0])
    return out
