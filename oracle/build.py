"""Build the C oracle (oracle/cgen.py output) for the bundled mechanisms -- TEST INFRASTRUCTURE.

Each mechanism becomes oracle/_build/oracle_<name>.so, compiled with
``gcc -O3 -march=x86-64-v3 -fopenmp`` (AVX2/FMA, no fast-math, so libm's exp/pow/log keep
their accuracy): the checker.  x86-64-v3 rather than -march=native: the library is built
here and also runs on the GPU host.

The CPU baseline is timed with the same C source built for speed as well: ``-ffast-math``
lets gcc vectorise the strip loops' exp/pow/log calls with glibc's vector math library
(libmvec, AVX2: oracle_<name>_vec.so; AVX-512 with -march=x86-64-v4: oracle_<name>_vec512.so),
as a vectorising Fortran compiler does for the reference's getrates_i.  Its results stay
within 1e-13 (gross-normalised) of the strict build; the benchmark checks that on a sample.
"""
import ctypes
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
MECHS = ("h2_burke", "h2_burke2012", "gri30", "synth_gri53", "synth_nhept88", "synth_jetsurf348")
CFLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-fno-math-errno", "-shared", "-fPIC"]
# timing builds of the CPU baseline (vectorised exp/pow/log through libmvec)
FLAVORS = {"": CFLAGS,
           "vec": ["-O3", "-march=x86-64-v3", "-ffast-math", "-fopenmp", "-shared", "-fPIC"],
           "vec512": ["-O3", "-march=x86-64-v4", "-mprefer-vector-width=512", "-ffast-math", "-fopenmp",
                      "-shared", "-fPIC"]}
TIMED = ("h2_burke", "h2_burke2012", "synth_gri53", "h2_burke_nofalloff")


def so_path(name: str, flavor: str = "") -> str:
    return os.path.join(BUILD, f"oracle_{name}{'_' + flavor if flavor else ''}.so")


def build_one(mech, name: str = None, flavors=("",)) -> str:
    from oracle.cgen import emit_c
    name = name or mech.name
    os.makedirs(BUILD, exist_ok=True)
    src = os.path.join(BUILD, f"oracle_{name}.c")
    with open(src, "w") as fh:
        fh.write(emit_c(mech))
    for fl in flavors:
        out = so_path(name, fl)
        libs = ["-lmvec", "-lm"] if fl else ["-lm"]
        res = subprocess.run(["gcc", *FLAVORS[fl], src, "-o", out + ".tmp", *libs], capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"gcc failed for {src}:\n{res.stderr[-3000:]}")
        os.replace(out + ".tmp", out)
    return so_path(name)


def h2_burke_nofalloff():
    """H2 Burke without its third-body and Troe reactions: what the reference's CPU mini-app
    computes when built with --rtypes-together and without -DTROE/-DTHIRD_BODY
    (chemgen/chemistry.py:383-397, mini_app/CMakeLists.txt:24-34), for the comparison with
    its published per-core figure."""
    import dataclasses
    from chemgen_amd.kernels import load_bundled
    m = load_bundled("h2_burke")
    return dataclasses.replace(m, name="h2_burke_nofalloff",
                               reactions=[r for r in m.reactions if r.rtype == "standard"])


def build_all(names=MECHS, jobs=6):
    sys.path.insert(0, ROOT)
    from chemgen_amd.kernels import load_bundled
    work = [(n, load_bundled(n)) for n in names] + [("h2_burke_nofalloff", h2_burke_nofalloff())]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda nm: build_one(nm[1], nm[0], ("",) + (("vec", "vec512") if nm[0] in TIMED else ())),
                    work))


class COracle:
    """ctypes handle on one built C oracle."""

    def __init__(self, name: str, flavor: str = ""):
        path = so_path(name, flavor)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path}: run oracle.build.build_all() first")
        self.lib = ctypes.CDLL(path)
        self.lib.oracle_getrates.argtypes = [ctypes.c_int64, ctypes.c_int64] + [ctypes.c_void_p] * 4
        self.lib.oracle_getrates.restype = None
        self.ns = self.lib.oracle_n_species()
        self.threads = self.lib.oracle_threads()

    def set_threads(self, n: int):
        """OpenMP threads of the following calls (process-wide)."""
        self.lib.oracle_set_threads(int(n))
        self.threads = self.lib.oracle_threads()

    def getrates(self, P, T, Y):
        import numpy as np
        P = np.ascontiguousarray(P, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        assert Y.shape[0] == self.ns
        ng = P.shape[0]
        W = np.empty_like(Y)
        self.lib.oracle_getrates(ng, Y.shape[1], P.ctypes.data, T.ctypes.data, Y.ctypes.data, W.ctypes.data)
        return W


if __name__ == "__main__":
    build_all()
