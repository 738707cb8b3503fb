"""Python call surface of the HIP target: ``wdot = getrates(P, T, Y)``.

The reference exposes its Fortran target to Python through f2py as
``wdot = getrates(P_cgs, T, Y)`` (tests/test_getrates.py:53-54,81) and its Python target as
``getrates(T, Y, P, wdot)`` filling ``wdot`` in place (chemgen/templates/python_header.j2:25).
:class:`GetRates` keeps both call forms, with the same units (P in dyn/cm^2, T in K, Y mass
fractions, wdot in mol/cm^3/s), and extends them to many grid points per call:

* scalar P and T with ``Y[Ns]`` -> ``wdot[Ns]`` (one point, the f2py form);
* ``P[ng]``, ``T[ng]`` with ``Y[ng, Ns]`` (``layout="point"``, the default for NumPy) or
  ``Y[Ns, ng]`` (``layout="species"``, the Fortran ``Y(ng, Ns)`` memory order and the
  default for device tensors) -> ``wdot`` in the same layout;
* NumPy arrays (copied to the GPU and back) or float64 device tensors (used in place).

Every evaluation runs the generated HIP kernel through libchemgen_amd.so; there is no
host fallback.
"""
from __future__ import annotations

import json
import os
from typing import List, Optional

import numpy as np

from .codegen import Variant, build_coef, structure_hash
from .mechanism import Mechanism, load_mechanism
from .runtime import ChemgenError, DeviceMechanism


class GetRates:
    """A mechanism loaded on the current GPU, callable as ``getrates(P, T, Y)``."""

    def __init__(self, mech: Optional[Mechanism] = None, variant: Optional[Variant] = None,
                 code_object: Optional[str] = None, coef: Optional[np.ndarray] = None,
                 build: bool = True):
        if mech is None:
            raise ChemgenError("a Mechanism is required")
        if variant is None:
            from .autotune import best_variant
            variant = best_variant(mech)
        self.mech = mech
        q = set(getattr(mech, "qss", ()))
        self.species: List[str] = [n for n in mech.species_names if n not in q]
        self._dm = DeviceMechanism(mech, variant, build=build, coef=coef, code_object=code_object)

    @property
    def n_species(self) -> int:
        """Species in Y and wdot (the transported ones when the mechanism has QSS species)."""
        return getattr(self.mech, "n_transported", self.mech.n_species)

    def __call__(self, P, T, Y, layout: Optional[str] = None):
        import torch
        if isinstance(Y, torch.Tensor) and Y.is_cuda:
            return self._device_call(P, T, Y, layout or "species")
        return self._host_call(P, T, Y, layout or "point")

    def getrates_inplace(self, T, Y, P, wdot) -> None:
        """The reference Python target's form: fills ``wdot`` (python_header.j2:25)."""
        out = self(P, T, Y)
        wdot[...] = out

    def _device_call(self, P, T, Y, layout):
        import torch
        ns = self.n_species
        if Y.dim() == 1:
            if Y.shape[0] != ns:
                raise ChemgenError(f"Y has {Y.shape[0]} entries, the mechanism {ns} species")
            Ys = Y.reshape(ns, 1).contiguous()
            ng, squeeze = 1, True
        elif layout == "species":
            Ys, ng, squeeze = Y, Y.shape[1], False
        elif layout == "point":
            Ys, ng, squeeze = Y.t().contiguous(), Y.shape[0], False
        else:
            raise ChemgenError(f"unknown layout {layout}")
        dev = Y.device
        Pd = _as_points(P, ng, dev)
        Td = _as_points(T, ng, dev)
        W = torch.empty_like(Ys)
        self._dm.getrates_(Pd, Td, Ys, W)
        if squeeze:
            return W.reshape(ns)
        return W if layout == "species" else W.t().contiguous()

    def _host_call(self, P, T, Y, layout):
        import torch
        if not torch.cuda.is_available():
            raise ChemgenError("getrates needs a GPU (no host fallback)")
        Yn = np.asarray(Y, dtype=np.float64)
        dev = torch.device("cuda", torch.cuda.current_device())
        Yd = torch.from_numpy(np.ascontiguousarray(Yn)).to(dev)
        out = self._device_call(P, T, Yd, layout)
        torch.cuda.synchronize()
        return out.cpu().numpy()


def _as_points(x, ng, dev):
    import torch
    if isinstance(x, torch.Tensor):
        t = x.to(device=dev, dtype=torch.float64).reshape(-1)
    else:
        t = torch.as_tensor(np.asarray(x, dtype=np.float64).reshape(-1), device=dev)
    if t.numel() == 1 and ng != 1:
        t = t.expand(ng).contiguous()
    if t.numel() != ng:
        raise ChemgenError(f"P/T have {t.numel()} entries for {ng} points")
    return t.contiguous()


def getrates(P, T, Y, mech: Mechanism):
    """One-shot convenience: load ``mech`` (cached per mechanism structure and coefficient
    values, so a mechanism with perturbed rates never reuses another's blob) and evaluate."""
    key = cache_key(mech)
    g = _CACHE.get(key)
    if g is None:
        g = _CACHE[key] = GetRates(mech)
    return g(P, T, Y)


_CACHE = {}


def cache_key(mech: Mechanism):
    """(name, structure hash, hash of the gpu_coef blob): one loaded mechanism per key."""
    from .codegen import coef_hash
    return (mech.name, structure_hash(mech), coef_hash(build_coef(mech)[0]))


# ---- generated-module support (generate_chemistry.py --mode getrates --language hip) ----

def save_generated(stem: str, mech: Mechanism, variant: Variant, code_object: str) -> None:
    """Write the artefacts a generated Python module loads: <stem>.json (metadata),
    <stem>.coef.npy (gpu_coef blob) and the mechanism as Chemkin text (<stem>.inp)."""
    from .mechanism import write_chemkin
    coef, _, _ = build_coef(mech)
    np.save(stem + ".coef.npy", coef)
    write_chemkin(mech, stem + ".inp")
    meta = {"mechanism": mech.name, "species": mech.species_names,
            "mw": [s.mw for s in mech.species], "structure_hash": f"{structure_hash(mech):016x}",
            "variant": __import__("dataclasses").asdict(variant),
            "qss": list(getattr(mech, "qss", ())),
            "code_object": os.path.basename(code_object)}
    with open(stem + ".json", "w") as fh:
        json.dump(meta, fh, indent=1)


def load_generated(stem: str) -> GetRates:
    """Load the artefacts written by :func:`save_generated` next to a generated module."""
    with open(stem + ".json") as fh:
        meta = json.load(fh)
    mech = load_mechanism(stem + ".inp", name=meta["mechanism"])
    if meta.get("qss"):
        from .qssa import with_qss
        mech = with_qss(mech, meta["qss"], name=meta["mechanism"])
    if f"{structure_hash(mech):016x}" != meta["structure_hash"]:
        raise ChemgenError(f"{stem}.inp does not match the generated code object")
    coef = np.load(stem + ".coef.npy")
    co = os.path.join(os.path.dirname(os.path.abspath(stem)), meta["code_object"])
    return GetRates(mech, Variant(**meta["variant"]), code_object=co, coef=coef, build=False)
