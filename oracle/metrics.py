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
    with np.errstate(divide="ignore", invalid="ignore"):
        e_gross = d / (np.abs(ref) + G)
        e_raw = np.where(ref != 0.0, d / np.abs(ref), np.where(d == 0.0, 0.0, np.inf))
    e_gross = np.where(np.isfinite(e_gross), e_gross, np.where(d == 0.0, 0.0, np.inf))
    return {
        "max_gross": float(np.max(e_gross)) if e_gross.size else 0.0,
        "max_raw": float(np.max(e_raw)) if e_raw.size else 0.0,
        "frac_raw_le_1e-12": float(np.mean(e_raw <= 1e-12)) if e_raw.size else 1.0,
        "finite": bool(np.isfinite(np.asarray(w)).all()),
    }
