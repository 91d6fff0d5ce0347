"""Helpers shared by the tests (not part of the package)."""
import numpy as np

from dqmc_amd import capi

_NUM_PARAMS = ("peierlsu", "peierlsd", "pp_u", "pp_d", "ppr_u", "ppr_d", "exp_Ku", "exp_Kd",
               "inv_exp_Ku", "inv_exp_Kd", "exp_halfKu", "exp_halfKd", "inv_exp_halfKu",
               "inv_exp_halfKd", "Ku", "Kd")


def as_complex(sim):
    """The same simulation as a complex (USE_CPLX) file: every `num` dataset
    complex128 with zero imaginary parts, the field and the random state
    unchanged."""
    p = dict(sim["params"])
    for k in _NUM_PARAMS:
        p[k] = np.asarray(p[k]).astype(np.complex128)
    out = {"params": p, "state": dict(sim["state"]), "metadata": dict(sim.get("metadata", {}))}
    for g in ("meas_eqlt", "meas_uneqlt"):
        out[g] = {k: (v.astype(np.complex128) if k != "n_sample" else v.copy())
                  for k, v in sim[g].items()}
    assert capi.is_complex(out)
    return out


# Accumulators that differ between the alpha = 1/2 and alpha = 1 gauges of
# the generator: the pair fields pick up e^{2i theta}; and g00 / gt0 weight
# G[i,j] with the straight-line Peierls phase of every site pair, which for
# pairs whose path wraps both torus directions differs by pi between the two
# gauges (the phase construction of util/tight_binding.py, reproduced by the
# generator) -- the bond phases, and so K and the trajectory, are covariant.
GAUGE_VARIANT = {"pair_sw", "pair_bb", "pair_b2b2", "g00", "g00_u", "g00_d", "gt0", "gt0_u",
                 "gt0_d"}
