"""CPU FP64 oracle for getrates with quasi-steady-state (QSS) species -- TEST INFRASTRUCTURE.

The reference only carries the scaffolding of QSS species (chemgen/chemistry.py:30-52: a
``qssa_species`` list, a reduced species list and index, ``is_qssa``; constants.py:485-489:
NSP_RED / NSP_SK / NSP_QSSA) and never computes with them, so no reference output exists to
match.  This restates the definition the generated kernels implement (linear QSS
approximation):

* the transported species R are the species not in Q, in mechanism order; Y holds R only;
  C_k = (Y_k/W_k) P / (R0 T sum_{j in R} Y_j/W_j) for k in R (ytoc.j2:3-17 over R);
* ctot = sum_{k in R} C_k and M = ctot + sum_{k in R} (eff_k - 1) C_k: QSS species, whose
  concentrations are small by assumption, take no part in either;
* rate constants as in the plain path (oracle/getrates_np.py); for each QSS species i the
  net production is zero, sum_j nu_ij (q_f,j - q_b,j) = 0, which is linear in the QSS
  concentrations when every QSS species enters with coefficient 1 and no reaction side holds
  two QSS species; that N_Q x N_Q system is solved per point (LAPACK, partial pivoting);
* wdot_k for k in R from the rates with those QSS concentrations.

Pinned by tests/test_qssa.py: a one-QSS-species mechanism against its closed form, the QSS
balance residual, element conservation of the transported rates (exact only when the QSS
balance holds), and equality with the plain oracle for an empty QSS set.
"""
from __future__ import annotations

import numpy as np

from oracle.getrates_np import (PATM, R0, _arrh, _falloff_kf, _kb, _plog_kf, _smh, r15)


def qss_getrates(mech, qss, P, T, Y, return_parts=False):
    """wdot[N_R, ng] for Y[N_R, ng] (transported species in mechanism order); with
    ``return_parts`` also (C_Q[N_Q, ng], |q_f|[Nr, ng], |q_b|[Nr, ng], residual[N_Q, ng])."""
    P = np.asarray(P, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    names = [s.name for s in mech.species]
    qset = list(qss)
    R = [k for k, n in enumerate(names) if n not in qset]
    Q = [names.index(n) for n in qset]
    qpos = {k: i for i, k in enumerate(Q)}
    ng = T.shape[0]
    assert Y.shape[0] == len(R)
    idx = {n: i for i, n in enumerate(names)}
    C = np.zeros((len(names), ng))
    for r, k in enumerate(R):
        C[k] = Y[r] / r15(mech.species[k].mw)
    csum = np.zeros(ng)
    for k in R:
        csum = csum + C[k]
    for k in R:
        C[k] = C[k] * P / (R0 * csum * T)
    ctot = np.zeros(ng)
    for k in R:
        ctot = ctot + C[k]
    pfac = PATM / (R0 * T)
    lnT = np.log(T)
    EG = np.empty((len(names), ng))
    for k, s in enumerate(mech.species):
        EG[k] = np.exp(_smh(s.thermo, T, lnT))
    # F_j, B_j: forward / reverse rates without the QSS factor; a_j, b_j: QSS species on
    # the reactant / product side (or None)
    F, B, qa, qb = [], [], [], []
    for r in mech.reactions:
        M = None
        if r.rtype != "standard" and r.rtype != "plog":
            M = ctot
            for sp, e in r.third_body.items():
                if sp in qset:
                    continue
                M = M + r15(abs(e)) * C[idx[sp]] if e > 0.0 else M - r15(abs(e)) * C[idx[sp]]
        if r.rtype in ("standard", "third_body"):
            kf = _arrh(*r.arrh, T)
        elif r.rtype == "plog":
            kf = _plog_kf(r, T, P)
        else:
            kf = _falloff_kf(r, T, M)
        kb = _kb(r, kf, EG, idx, pfac)
        f, b = kf, kb
        a_sp = b_sp = None
        for sp, nu in r.reactants.items():
            if sp in qset:
                a_sp = qpos[idx[sp]]
            else:
                f = f * (C[idx[sp]] if nu == 1.0 else C[idx[sp]] ** float(nu))
        for sp, nu in r.products.items():
            if sp in qset:
                b_sp = qpos[idx[sp]]
            else:
                b = b * (C[idx[sp]] if nu == 1.0 else C[idx[sp]] ** float(nu))
        if r.rtype == "third_body":
            f, b = M * f, M * b
        F.append(f)
        B.append(b)
        qa.append(a_sp)
        qb.append(b_sp)
    nq = len(Q)
    x = np.zeros((nq, ng))
    if nq:
        A = np.zeros((ng, nq, nq))
        c = np.zeros((ng, nq))
        for j in range(mech.n_reactions):
            a, b = qa[j], qb[j]
            if a is None and b is None:
                continue
            if a is not None and b == a:
                continue                                  # the same QSS species on both sides
            if a is not None:                             # -(q_f - q_b) into species a
                A[:, a, a] -= F[j]
                if b is not None:
                    A[:, a, b] += B[j]
                else:
                    c[:, a] += B[j]
            if b is not None:                             # +(q_f - q_b) into species b
                A[:, b, b] -= B[j]
                if a is not None:
                    A[:, b, a] += F[j]
                else:
                    c[:, b] += F[j]
        x = np.linalg.solve(A, -c[..., None])[..., 0].T
    wdot = np.zeros((len(R), ng))
    rpos = {k: i for i, k in enumerate(R)}
    qf_abs = np.zeros((mech.n_reactions, ng))
    qb_abs = np.zeros_like(qf_abs)
    resid = np.zeros((nq, ng))
    for j, r in enumerate(mech.reactions):
        qf = F[j] * (x[qa[j]] if qa[j] is not None else 1.0)
        qbk = B[j] * (x[qb[j]] if qb[j] is not None else 1.0)
        rr = qf - qbk
        for sp, nu in r.reactants.items():
            k = idx[sp]
            if k in rpos:
                wdot[rpos[k]] = wdot[rpos[k]] - (rr if nu == 1.0 else rr * float(nu))
            else:
                resid[qpos[k]] -= rr
        for sp, nu in r.products.items():
            k = idx[sp]
            if k in rpos:
                wdot[rpos[k]] = wdot[rpos[k]] + (rr if nu == 1.0 else rr * float(nu))
            else:
                resid[qpos[k]] += rr
        qf_abs[j] = np.abs(qf)
        qb_abs[j] = np.abs(qbk)
    if return_parts:
        return wdot, x, qf_abs, qb_abs, resid
    return wdot


def qss_gross_scale(mech, qss, qf_abs, qb_abs):
    """G[k, i] over the transported species: sum_j (nu'_kj + nu''_kj)(|q_f,j| + |q_b,j|)."""
    names = [s.name for s in mech.species]
    R = [n for n in names if n not in set(qss)]
    rpos = {n: i for i, n in enumerate(R)}
    G = np.zeros((len(R),) + qf_abs.shape[1:])
    for j, r in enumerate(mech.reactions):
        g = qf_abs[j] + qb_abs[j]
        for st in (r.reactants, r.products):
            for sp, n in st.items():
                if sp in rpos:
                    G[rpos[sp]] += abs(n) * g
    return G
