"""getrates in the GEMM formulation (the reference's parallel level 4) on MI355X.

The reference's level 4 (chemgen/templates/reactions_hip/hip_v4_cpp.j2:1-19, driver
:422-597) computes the stoichiometric sums of every reaction as rocBLAS DGEMMs of the
per-point logarithms with dense [nr x Ns] coefficient matrices, the net rates element-wise,
and the production rates as a DGEMM of the rates with the [Ns x nr] matrix.  ``GemmPlan``
builds the dense matrices and per-reaction records of a ``Mechanism`` (rates read from the
same gpu_coef blob as the generated kernels); ``GemmGetRates`` runs them through
``cg_gemm_create`` / ``cg_getrates_gemm`` (chemgen_amd/csrc/cg_gemm.hip: two HIP kernels and
four FP64 DGEMMs per slab of points).  It is a second execution strategy beside the generated
kernels, measured against them (profiles/gemm_formulation.md); the generated kernels stay the
default.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List

import numpy as np

from .codegen import _table_kind, build_coef, negative_a, r15
from .mechanism import Mechanism
from .runtime import ChemgenError, _check_dev, _ld, _stream_ptr, check, lib

REC_WIDTH = 8


@dataclass
class GemmPlan:
    """Host arrays of the GEMM formulation (the argument list of cg_gemm_create)."""
    coef: np.ndarray       # gpu_coef blob
    off_smh: int
    tmid: np.ndarray       # [Ns]
    nu_r: np.ndarray       # [Nr, Ns]
    nu_p: np.ndarray       # [Nr, Ns]
    eff: np.ndarray        # [Ntb, Ns]: collision efficiencies, 1 for species not listed
    rec: np.ndarray        # [Nr, 8] int32
    dnu: np.ndarray        # [Nr]
    plog: np.ndarray       # [n_plog, 2] int32: blob offset of a pressure point, its Arrhenius terms


def gemm_plan(mech: Mechanism) -> GemmPlan:
    """Dense stoichiometric and efficiency matrices and one record per reaction (kind, blob
    offset, efficiency row, reversible, negative prefactor, pressure points) of ``mech``."""
    coef, off_smh, roffs = build_coef(mech)
    idx = {s.name: i for i, s in enumerate(mech.species)}
    ns, nr = mech.n_species, mech.n_reactions
    nu_r = np.zeros((nr, ns))
    nu_p = np.zeros((nr, ns))
    rec = np.zeros((nr, REC_WIDTH), dtype=np.int32)
    dnu = np.zeros(nr)
    eff_rows: List[np.ndarray] = []
    plog: List[List[int]] = []
    for j, r in enumerate(mech.reactions):
        for s, n in r.reactants.items():
            nu_r[j, idx[s]] += n
        for s, n in r.products.items():
            nu_p[j, idx[s]] += n
        kind = _table_kind(r)
        tb = -1
        if r.rtype in ("third_body", "lindemann", "troe", "sri"):
            row = np.ones(ns)
            for s, e in r.third_body.items():    # stored as eff - 1 (chemgen/parser.py semantics)
                row[idx[s]] = 1.0 + r15(e)
            tb = len(eff_rows)
            eff_rows.append(row)
        p0 = np.int32(len(plog))
        npt = 0
        if r.rtype == "plog":
            o = roffs[j]
            for p in sorted(set(c[0] for c in r.plog)):
                n_rates = sum(1 for c in r.plog if c[0] == p)
                plog.append([o, n_rates])
                o += 3 + 4 * n_rates
                npt += 1
        rec[j] = (kind, roffs[j], tb, int(r.reversible), int(negative_a(r)), p0 if npt else 0, npt, 0)
        dnu[j] = r.coeff_sum
    eff = np.array(eff_rows) if eff_rows else np.zeros((0, ns))
    tmid = np.array([s.thermo.t_mid for s in mech.species], dtype=np.float64)
    return GemmPlan(coef, off_smh, tmid, nu_r, nu_p, eff, rec, dnu,
                    np.array(plog, dtype=np.int32).reshape(-1, 2))


class GemmGetRates:
    """A GEMM-formulation plan on the current device: getrates_(P, T, Y, wdot) with the same
    arguments and layout as ``DeviceMechanism.getrates_``.  ``slab``: grid points per pass (the
    plan's workspace holds (3 Ns + 4 Nr + Ntb + 4) x slab doubles)."""

    def __init__(self, mech: Mechanism, slab: int = 1 << 16):
        L = lib()
        c_p, c_i32, c_i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        L.cg_gemm_create.argtypes = [c_p, c_i64, c_i32, c_p, c_i32, c_i32, c_i32, c_p, c_p, c_p, c_p, c_p, c_p,
                                     c_i32, c_i64, ctypes.POINTER(c_p)]
        L.cg_getrates_gemm.argtypes = [c_p, c_p, c_p, c_i64, c_i64, c_p, c_p, c_p]
        L.cg_gemm_free.argtypes = [c_p]
        for f in ("cg_gemm_create", "cg_getrates_gemm", "cg_gemm_free"):
            getattr(L, f).restype = ctypes.c_int
        self.mech = mech
        self.n_species = mech.n_species
        self.plan = pl = gemm_plan(mech)
        arrs = [np.ascontiguousarray(a) for a in (pl.coef, pl.tmid, pl.nu_r, pl.nu_p, pl.eff, pl.rec, pl.dnu,
                                                   pl.plog)]
        coef, tmid, nu_r, nu_p, eff, rec, dnu, plog = arrs
        self._h = ctypes.c_void_p()
        check(L.cg_gemm_create(coef.ctypes.data, len(coef), pl.off_smh, tmid.ctypes.data, mech.n_species,
                               mech.n_reactions, eff.shape[0], nu_r.ctypes.data, nu_p.ctypes.data,
                               eff.ctypes.data if eff.size else None, rec.ctypes.data, dnu.ctypes.data,
                               plog.ctypes.data if plog.size else None, plog.shape[0], slab,
                               ctypes.byref(self._h)), "cg_gemm_create")

    def getrates_(self, P, T, Y, wdot, stream=None) -> None:
        """In-place, on device tensors: P[ng], T[ng], Y[Ns, ld], wdot[Ns, ld] (float64)."""
        if not self._h:
            raise ChemgenError("GemmGetRates is closed")
        ng = int(P.shape[0])
        _check_dev(P, T, Y, wdot, self.n_species, ng)
        ld = _ld(Y)
        if _ld(wdot) != ld:
            raise ChemgenError("Y and wdot must share the leading dimension")
        check(lib().cg_getrates_gemm(self._h, P.data_ptr(), T.data_ptr(), ng, ld, Y.data_ptr(), wdot.data_ptr(),
                                     _stream_ptr(stream)), "cg_getrates_gemm")

    def getrates(self, P, T, Y, stream=None):
        import torch
        W = torch.empty_like(Y)
        self.getrates_(P, T, Y, W, stream)
        return W

    def close(self) -> None:
        if self._h:
            lib().cg_gemm_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
