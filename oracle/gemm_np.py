"""The GEMM formulation of getrates evaluated on the host in NumPy -- TEST INFRASTRUCTURE.

Checks the plan that chemgen_amd.gemm builds for the GPU path (dense stoichiometric and
efficiency matrices, per-reaction records into the gpu_coef blob) independently of the GPU:
the same products and rate formulas, evaluated here in float64 NumPy, must reproduce the
oracle's wdot (oracle/getrates_np.py).  The formulation is the reference's parallel level 4
(chemgen/templates/reactions_hip/hip_v4_cpp.j2:1-19, :9-16 method 1: kf (2^sum nu_r log2 C -
2^(sum nu_p log2 C - dsmh - dnu log2 pfac))).  Only tests import this module.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from chemgen_amd.gemm import GemmPlan, gemm_plan
from chemgen_amd.mechanism import Mechanism


def gemm_formulation_np(mech: Mechanism, P: np.ndarray, T: np.ndarray, Y: np.ndarray,
                        plan: Optional[GemmPlan] = None) -> np.ndarray:
    """The GEMM formulation on the host in NumPy (dense products of the same plan arrays, the
    rates from the same blob records): a check of the plan's matrices and records against the
    oracle without a GPU.  Y [Ns, ng] -> wdot [Ns, ng]."""
    pl = plan or gemm_plan(mech)
    cf = pl.coef
    ns = mech.n_species
    lg2T, invT = np.log2(T), 1.0 / T
    csum = (Y * cf[:ns, None]).sum(axis=0)
    scale = P / (8.314510e7 * csum * T)
    C = Y * cf[:ns, None] * scale
    S = np.empty_like(C)
    for k in range(ns):
        lo = T < pl.tmid[k]
        for rng, sel in ((0, lo), (7, ~lo)):
            a = cf[pl.off_smh + 14 * k + rng: pl.off_smh + 14 * k + rng + 7]
            S[k, sel] = (a[6] * invT + a[1] * lg2T + a[2] * T + a[3] * T ** 2 + a[4] * T ** 3
                         + a[5] * T ** 4 + a[0])[sel]
    with np.errstate(divide="ignore"):
        L2C = np.where(C > 0, np.log2(np.where(C > 0, C, 1.0)), -1100.0)
    SGr, SGp = pl.nu_r @ L2C, pl.nu_p @ L2C
    DS = (pl.nu_p - pl.nu_r) @ S
    M = pl.eff @ C
    lg2pfac = np.log2(1013250.0 / 8.314510e7) - lg2T
    Q = np.zeros((mech.n_reactions, len(T)))
    for j in range(mech.n_reactions):
        kind, ro, tb, rev, neg, p0, npt, _ = (int(x) for x in pl.rec[j])
        c = cf[ro:]
        if kind == 7:
            lgP = np.log2(P)

            def point(q):
                o, n = int(pl.plog[q, 0]), int(pl.plog[q, 1])
                cc = cf[o + 3:]
                if n == 1:
                    return cc[2] * lg2T + cc[1] - cc[3] * invT
                return np.log2(sum(cc[4 * r] * np.exp2(cc[4 * r + 2] * lg2T + cc[4 * r + 1] - cc[4 * r + 3] * invT)
                                   for r in range(n)))
            lo = np.full(len(T), -1)
            for q in range(npt):
                lo = np.where(P >= cf[int(pl.plog[p0 + q, 0]) + 1], q, lo)
            vals = [point(p0 + q) for q in range(npt)]
            lk = np.choose(np.maximum(lo, 0), vals)
            for q in range(npt - 1):
                a, b = int(pl.plog[p0 + q, 0]), int(pl.plog[p0 + q + 1, 0])
                interp = vals[q] + (vals[q + 1] - vals[q]) * (lgP - cf[a]) * cf[b + 2]
                lk = np.where(lo == q, interp, lk)
            kf = np.exp2(lk)
        else:
            lki = c[1] * lg2T + c[0] - c[2] * invT
            kf = np.exp2(lki)
            if kind >= 2:
                lk0 = c[4] * lg2T + c[3] - c[5] * invT
                pr = M[tb] * np.exp2(lk0 - lki)
                F = np.ones_like(T)
                if kind in (3, 4):
                    fcent = c[6] * np.exp2(-T * c[7]) + c[8] * np.exp2(-T * c[9])
                    if kind == 4:
                        fcent = fcent + np.exp2(-c[10] * invT)
                    lf, lp = np.log10(fcent), np.log10(pr)
                    cc, nn = -0.4 - 0.67 * lf, 0.75 - 1.27 * lf
                    f1 = (lp + cc) / (nn - 0.14 * (lp + cc))
                    F = np.exp2(np.log2(fcent) / (1.0 + f1 * f1))
                elif kind >= 5:
                    lp = np.log10(pr)
                    X = 1.0 / (1.0 + lp * lp)
                    F = c[9] * np.exp2(X * np.log2(c[6] * np.exp2(-c[7] * invT) + np.exp2(-T * c[8])))
                    if kind == 6:
                        F = F * np.exp2(c[10] * lg2T)
                kf = kf * pr / (1.0 + pr) * F
        if neg:
            kf = -kf
        bwd = np.exp2(SGp[j] - DS[j] - pl.dnu[j] * lg2pfac) if rev else 0.0
        q = kf * (np.exp2(SGr[j]) - bwd)
        if kind == 1:
            q = q * M[tb]
        Q[j] = q
    return (pl.nu_p - pl.nu_r).T @ Q
