"""ctypes layer over libchemgen_amd.so (include/chemgen_amd.h) and the code-object cache.

PyTorch is used only for device buffers and streams: tensors are handed to the C ABI as
``data_ptr()`` device pointers together with ``torch.cuda.current_stream().cuda_stream``.
There is no host or PyTorch fallback: if the library or a code object is missing the
call raises.
"""
from __future__ import annotations

import ctypes
import os
import threading
from typing import Optional

import numpy as np

from .codegen import Variant, build_coef, compile_code_object, emit_hip, structure_hash, variant_flags
from .mechanism import Mechanism

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "_lib", "libchemgen_amd.so")
KERNEL_DIR = os.path.join(PKG_DIR, "_kernels")

_lib = None
_lib_lock = threading.Lock()


class ChemgenError(RuntimeError):
    pass


def lib():
    """Load libchemgen_amd.so (raises if it was not built)."""
    global _lib
    with _lib_lock:
        if _lib is None:
            # PyTorch-ROCm ships its own libamdhip64.so.7; load it first so that this library
            # binds to the same HIP runtime (one runtime per process: torch's streams and
            # allocations are then valid handles here)
            import torch  # noqa: F401
            if not os.path.exists(LIB_PATH):
                raise ChemgenError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
            L = ctypes.CDLL(LIB_PATH)
            c_i64, c_i32, c_u64, c_p = ctypes.c_int64, ctypes.c_int32, ctypes.c_uint64, ctypes.c_void_p
            L.cg_last_error.restype = ctypes.c_char_p
            L.cg_mech_load.argtypes = [ctypes.c_char_p, c_p, c_i64, c_u64, ctypes.POINTER(c_p)]
            L.cg_mech_set_coef.argtypes = [c_p, c_p, c_i64]
            L.cg_mech_info.argtypes = [c_p, ctypes.POINTER(c_i32), ctypes.POINTER(c_i32),
                                       ctypes.POINTER(c_i32), ctypes.POINTER(c_u64)]
            L.cg_mech_free.argtypes = [c_p]
            L.cg_getrates.argtypes = [c_p, c_p, c_p, c_i64, c_i64, c_p, c_p, c_p]
            L.cg_getrates_batched.argtypes = [c_p, c_p, c_p, c_p, c_p, c_i64, c_p, c_i32, c_i32, c_p]
            L.cg_batched_prepare.argtypes = [c_p, c_p, c_i32]
            L.cg_getrates_batched_prepared.argtypes = [c_p, c_p, c_p, c_p, c_p, c_i64, c_i32, c_p]
            L.cg_batch_create.argtypes = [c_p, c_p, c_i32, c_p, ctypes.POINTER(c_p)]
            L.cg_batch_info.argtypes = [c_p, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)]
            L.cg_getrates_batch.argtypes = [c_p, c_p, c_p, c_p, c_p, c_i64, c_i32, c_p]
            L.cg_batch_free.argtypes = [c_p]
            L.cg_batch_tiles.argtypes = [c_p, c_i32, c_i32, c_p, c_p, c_i64]
            L.cg_batch_tiles.restype = c_i64
            L.cg_getrates_batched_noalloc.argtypes = [c_p, c_p, c_p, c_p, c_p, c_i64, c_p, c_i32, c_i32, c_p]
            L.cg_getrates_tiles.argtypes = [c_p, c_p, c_p, c_p, c_p, c_i64, c_p, c_p, c_i64, c_i64, c_i32, c_p]
            for f in ("cg_mech_load", "cg_mech_set_coef", "cg_mech_info", "cg_mech_free", "cg_getrates",
                      "cg_getrates_batched", "cg_batched_prepare", "cg_getrates_batched_prepared",
                      "cg_batch_create", "cg_batch_info", "cg_getrates_batch", "cg_batch_free",
                      "cg_getrates_batched_noalloc", "cg_getrates_tiles"):
                getattr(L, f).restype = ctypes.c_int
            _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().cg_last_error().decode(errors="replace")
        raise ChemgenError(f"{what} failed ({rc}): {msg}")


def code_object_path(mech: Mechanism, variant: Variant) -> str:
    h = structure_hash(mech)
    if variant.baked:   # the coefficient values are part of the code: name them too
        from .codegen import build_coef, coef_hash
        h ^= coef_hash(build_coef(mech)[0])
    return os.path.join(KERNEL_DIR, f"{mech.name}_{h:016x}_{variant.tag}.hsaco")


def ensure_code_object(mech: Mechanism, variant: Variant = Variant(), build: bool = True,
                       force: bool = False) -> str:
    """Path of the compiled code object for (mechanism structure, variant); compiles it
    with hipcc if it is missing and ``build`` is true."""
    path = code_object_path(mech, variant)
    if os.path.exists(path) and not force:
        return path
    if not build:
        raise ChemgenError(f"code object {path} was not built")
    os.makedirs(KERNEL_DIR, exist_ok=True)
    src, _, _ = emit_hip(mech, variant)
    src_path = path[:-len(".hsaco")] + ".hip"
    with open(src_path, "w") as fh:
        fh.write(src)
    compile_code_object(src_path, path, variant_flags(variant))
    return path


def _stream_ptr(stream) -> int:
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return int(stream.cuda_stream)


class DeviceMechanism:
    """One mechanism loaded on the current GPU (code object + gpu_coef blob)."""

    def __init__(self, mech: Mechanism, variant: Optional[Variant] = None, build: bool = True,
                 coef: Optional[np.ndarray] = None, code_object: Optional[str] = None):
        from .autotune import best_variant
        self.mech = mech
        variant = variant or best_variant(mech)
        self.variant = variant
        self.path = code_object or ensure_code_object(mech, variant, build=build)
        c, _, _ = build_coef(mech)
        if coef is not None:
            c = np.asarray(coef, dtype=np.float64)
        self.coef = np.ascontiguousarray(c)
        self._h = ctypes.c_void_p()
        check(lib().cg_mech_load(self.path.encode(), self.coef.ctypes.data, len(self.coef),
                                 structure_hash(mech), ctypes.byref(self._h)), "cg_mech_load")
        ns, nr, blk, hh = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_uint64()
        check(lib().cg_mech_info(self._h, ctypes.byref(ns), ctypes.byref(nr), ctypes.byref(blk),
                                 ctypes.byref(hh)), "cg_mech_info")
        self.n_species, self.n_reactions, self.block = ns.value, nr.value, blk.value

    @property
    def handle(self):
        return self._h

    def set_coef(self, coef: np.ndarray) -> None:
        c = np.ascontiguousarray(np.asarray(coef, dtype=np.float64))
        check(lib().cg_mech_set_coef(self._h, c.ctypes.data, len(c)), "cg_mech_set_coef")
        self.coef = c

    def getrates_(self, P, T, Y, wdot, stream=None) -> None:
        """In-place, on device tensors: P[ng], T[ng], Y[Ns, ld], wdot[Ns, ld] (float64)."""
        ng = int(P.shape[0])
        _check_dev(P, T, Y, wdot, self.n_species, ng)
        ld = _ld(Y)
        if _ld(wdot) != ld:
            raise ChemgenError("Y and wdot must share the leading dimension")
        check(lib().cg_getrates(self._h, P.data_ptr(), T.data_ptr(), ng, ld, Y.data_ptr(),
                                wdot.data_ptr(), _stream_ptr(stream)), "cg_getrates")

    def getrates(self, P, T, Y, stream=None):
        """wdot = getrates(P, T, Y) on device tensors; returns a new [Ns, ng] tensor."""
        import torch
        wdot = torch.empty_like(Y)
        self.getrates_(P, T, Y, wdot, stream)
        return wdot

    def getrates_batched_(self, P, T, Y, wdot, offsets, grid: int = 0, stream=None) -> None:
        """Ragged batch: patch b = points offsets[b]:offsets[b+1] of the packed arrays."""
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        ng = int(off[-1])
        _check_dev(P, T, Y, wdot, self.n_species, ng)
        ld = _ld(Y)
        if _ld(wdot) != ld:
            raise ChemgenError("Y and wdot must share the leading dimension")
        check(lib().cg_getrates_batched(self._h, P.data_ptr(), T.data_ptr(), Y.data_ptr(),
                                        wdot.data_ptr(), ld, off.ctypes.data, len(off) - 1, grid,
                                        _stream_ptr(stream)), "cg_getrates_batched")

    def batched_prepare(self, offsets) -> None:
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        check(lib().cg_batched_prepare(self._h, off.ctypes.data, len(off) - 1), "cg_batched_prepare")
        self._prepared_points = int(off[-1])

    def getrates_batched_prepared_(self, P, T, Y, wdot, grid: int = 0, stream=None) -> None:
        ng = getattr(self, "_prepared_points", None)
        if ng is None:
            raise ChemgenError("getrates_batched_prepared_: no batch prepared")
        _check_dev(P, T, Y, wdot, self.n_species, ng)
        ld = _ld(Y)
        if _ld(wdot) != ld:
            raise ChemgenError("Y and wdot must share the leading dimension")
        check(lib().cg_getrates_batched_prepared(self._h, P.data_ptr(), T.data_ptr(), Y.data_ptr(),
                                                 wdot.data_ptr(), ld, grid, _stream_ptr(stream)),
              "cg_getrates_batched_prepared")

    def getrates_tiles_(self, P, T, Y, wdot, tile_start, tile_len, n_points: int, grid: int = -1,
                        stream=None) -> None:
        """One launch over a device tile list (int64 starts, int32 lengths; zero-length tiles are
        skipped) of points in [0, n_points)."""
        _check_dev(P, T, Y, wdot, self.n_species, n_points)
        ld = _ld(Y)
        if _ld(wdot) != ld:
            raise ChemgenError("Y and wdot must share the leading dimension")
        check(lib().cg_getrates_tiles(self._h, P.data_ptr(), T.data_ptr(), Y.data_ptr(), wdot.data_ptr(), ld,
                                      tile_start.data_ptr(), tile_len.data_ptr(), int(tile_start.numel()),
                                      n_points, grid, _stream_ptr(stream)), "cg_getrates_tiles")

    def batch(self, offsets, stream=None) -> "Batch":
        """A ragged-batch object (cg_batch_create): its tile list lives until ``close()``, so
        graphs captured over its launches stay valid that long."""
        return Batch(self, offsets, stream)

    def close(self) -> None:
        if self._h:
            lib().cg_mech_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """A ragged batch of patches on one mechanism (owns its device tile list)."""

    def __init__(self, dm: DeviceMechanism, offsets, stream=None):
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        self.dm = dm
        self.n_points = int(off[-1])
        self._b = ctypes.c_void_p()
        check(lib().cg_batch_create(dm.handle, off.ctypes.data, len(off) - 1, _stream_ptr(stream),
                                    ctypes.byref(self._b)), "cg_batch_create")
        nt, npt = ctypes.c_int64(), ctypes.c_int64()
        check(lib().cg_batch_info(self._b, ctypes.byref(nt), ctypes.byref(npt)), "cg_batch_info")
        self.n_tiles = nt.value

    def getrates_(self, P, T, Y, wdot, grid: int = 0, stream=None) -> None:
        _check_dev(P, T, Y, wdot, self.dm.n_species, self.n_points)
        ld = _ld(Y)
        if _ld(wdot) != ld:
            raise ChemgenError("Y and wdot must share the leading dimension")
        check(lib().cg_getrates_batch(self._b, P.data_ptr(), T.data_ptr(), Y.data_ptr(), wdot.data_ptr(),
                                      ld, grid, _stream_ptr(stream)), "cg_getrates_batch")

    def close(self) -> None:
        if self._b:
            lib().cg_batch_free(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_tiles(offsets, tile: int):
    """(tile starts, tile lengths) of a ragged batch for ``tile``-point workgroups, as the
    library builds them (cg_batch_tiles; host only)."""
    off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
    n = lib().cg_batch_tiles(off.ctypes.data, len(off) - 1, tile, None, None, 0)
    if n < 0:
        check(int(n), "cg_batch_tiles")
    st = np.zeros(max(n, 1), dtype=np.int64)
    ln = np.zeros(max(n, 1), dtype=np.int32)
    n2 = lib().cg_batch_tiles(off.ctypes.data, len(off) - 1, tile, st.ctypes.data, ln.ctypes.data, n)
    assert n2 == n
    return st[:n], ln[:n]


def _check_dev(P, T, Y, wdot, ns, ng):
    import torch
    for name, t in (("P", P), ("T", T), ("Y", Y), ("wdot", wdot)):
        if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64:
            raise ChemgenError(f"{name} must be a float64 device tensor")
    if P.dim() != 1 or T.dim() != 1 or P.shape[0] < ng or T.shape[0] < ng:
        raise ChemgenError("P and T must be 1-D with at least ng points")
    if not (P.is_contiguous() and T.is_contiguous()):
        raise ChemgenError("P and T must be contiguous")
    for name, t in (("Y", Y), ("wdot", wdot)):
        if t.dim() != 2 or t.shape[0] != ns or t.shape[1] < ng or (t.shape[1] > 1 and t.stride(1) != 1):
            raise ChemgenError(f"{name} must be species-major [{ns}, >=ng] with unit point stride")
        if ns > 1 and t.stride(0) < t.shape[1]:
            raise ChemgenError(f"{name}: species rows overlap (stride {t.stride(0)} < {t.shape[1]})")


def _ld(t) -> int:
    """Leading dimension (distance between species rows) of a [Ns, n] tensor."""
    return int(t.stride(0)) if t.shape[0] > 1 else int(t.shape[1])
