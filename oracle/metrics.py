"""Parity metrics -- TEST INFRASTRUCTURE.

Per (species k, point i): e_raw = |dw| / |w_ref| and the gross-normalised
e_gross = |dw| / (|w_ref| + G_ki), G_ki = sum_j |nu_kj| (|q_f,j| + |q_b,j|): the size of
the terms whose sum is w_ki.  A species close to steady state has |w| far below G, so
any change of summation order moves e_raw by orders of magnitude while e_gross stays at
rounding level; the gate is on e_gross, e_raw is reported.
"""
import numpy as np

from oracle.getrates_np import gross_scale


def parity_stats(mech, w, ref, qf, qb):
    G = gross_scale(mech, qf, qb)
    d = np.abs(np.asarray(w) - ref)
    nz = ref != 0.0
    with np.errstate(divide="ignore", invalid="ignore"):
        e_gross = d / (np.abs(ref) + G)
        e_raw = np.where(nz, d / np.abs(ref), 0.0)
    e_gross = np.where(np.isfinite(e_gross), e_gross, np.where(d == 0.0, 0.0, np.inf))
    return {
        "max_gross": float(np.max(e_gross)) if e_gross.size else 0.0,
        # raw relative error over entries with a non-zero reference value
        "max_raw": float(np.max(e_raw[nz])) if nz.any() else 0.0,
        "frac_raw_le_1e-12": float(np.mean(e_raw[nz] <= 1e-12)) if nz.any() else 1.0,
        # entries whose reference is exactly zero (inert species) but the result is not
        "n_zero_ref_nonzero": int(np.sum(~nz & (d != 0.0))),
        "finite": bool(np.isfinite(np.asarray(w)).all()),
    }
