"""HIP (gfx950) emitter for getrates and the gpu_coef coefficient blob.

The reference renders its HIP kernels from chemgen/templates/reactions_hip/*.j2 and keeps
the mechanism numbers in ``__constant__`` arrays copied by ``copyConstantsToDevice``
(chemgen/constants.py:119-240, chemgen/write_constants.py:408-493).  Here the generator
compiles the mechanism *structure* (which species each reaction touches, its type, its
zero coefficients) into straight-line code, one lane per grid point, and keeps the
*numbers* in a flat FP64 blob (gpu_coef) that the kernel reads through uniform loads:

  [invmw[Ns] | smh[Ns][14] | per-reaction records]

smh record per species: low range (a6-a0, a0, a1/2, a2/6, a3/12, a4/20, -a5), then the
high range -- the transform of chemgen/write_constants.py:146-189 -- evaluated as
``c0 + c1 lnT + T(c2 + T(c3 + T(c4 + T c5))) + c6/T``.

Reaction records (kernel order = mechanism order):
  standard/third-body: lnA, beta, Ea/Rc, then (eff-1) per listed third-body species
  falloff: lnA_inf, beta_inf, Ea_inf/Rc, lnA_0, beta_0, Ea_0/Rc, then
           Troe (1-a, 1/T3, a, 1/T1[, T2]) | SRI (a, b, 1/c, d, e) | Lindemann (none),
           then (eff-1) per listed species
  plog: per pressure point (ln p, n_rates, lnA, beta, Ea/Rc ...)

Numbers are taken after the reference's 16-significant-digit printing (``%.15e``,
chemgen/templates/reactions/standard.j2:5, chemgen/write_constants.py:14-70).
"""
from __future__ import annotations

import hashlib
import math
import os
import subprocess
from dataclasses import dataclass
from typing import Dict, List, Tuple

import numpy as np

from .mechanism import Mechanism, Reaction
from .writer_utils import dlit, power_expr, stoich_product, tree_sum

R0 = 8.314510e+07
RC = 1.9872155832
PATM = 1013250.0
LN10 = math.log(10.0)

TEMPLATE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "templates")


def r15(x: float) -> float:
    return float("%.15e" % x)


@dataclass(frozen=True)
class Variant:
    """Launch/compile shape of one generated kernel.

    block: points (lanes) per workgroup.
    waves_per_eu: occupancy the register allocator targets (0 = compiler default);
      1 wave/SIMD lets a lane keep ~512 registers (VGPR+AGPR), 2 -> 256.
    """
    block: int = 64
    waves_per_eu: int = 1

    @property
    def tag(self) -> str:
        return f"b{self.block}w{self.waves_per_eu}"


class _Coef:
    def __init__(self):
        self.vals: List[float] = []

    def add(self, *xs: float) -> int:
        off = len(self.vals)
        self.vals.extend(float(x) for x in xs)
        return off


def _lk(off: int, beta: float, ea: float) -> str:
    """ln k = lnA + beta lnT - (Ea/Rc)/T with zero terms compiled out."""
    e = f"cf[{off}]"
    if beta != 0.0:
        e = f"__builtin_fma(cf[{off + 1}], lnT, {e})"
    if ea != 0.0:
        e = f"__builtin_fma(-cf[{off + 2}], invT, {e})"
    return e


def _arr_vals(A: float, beta: float, Ea: float) -> Tuple[float, float, float]:
    return (math.log(r15(A)), float(beta), r15(Ea) / RC)


def structure_key(mech: Mechanism) -> str:
    """Everything the compiled code depends on (not the coefficient values)."""
    idx = {s.name: i for i, s in enumerate(mech.species)}
    parts = [f"ns={mech.n_species}", "tmid=" + ",".join(repr(s.thermo.t_mid) for s in mech.species)]
    for r in mech.reactions:
        parts.append(repr((
            r.rtype, r.reversible,
            tuple((idx[s], n) for s, n in r.reactants.items()),
            tuple((idx[s], n) for s, n in r.products.items()),
            r.arrh[1] != 0.0, r.arrh[2] != 0.0,
            None if r.low is None else (r.low[1] != 0.0, r.low[2] != 0.0),
            tuple(idx[s] for s in r.third_body),
            None if r.troe is None else len(r.troe),
            None if r.sri is None else len(r.sri),
            None if r.plog is None else tuple(len([c for c in r.plog if c[0] == p])
                                              for p in sorted(set(c[0] for c in r.plog))),
        )))
    return "\n".join(parts)


def structure_hash(mech: Mechanism) -> int:
    return int(hashlib.sha256(structure_key(mech).encode()).hexdigest()[:16], 16)


def build_coef(mech: Mechanism):
    """Return (coef array, per-reaction offsets) in the layout documented above."""
    cb = _Coef()
    off_invmw = cb.add(*[1.0 / r15(s.mw) for s in mech.species])
    assert off_invmw == 0
    off_smh = len(cb.vals)
    for s in mech.species:
        th = s.thermo
        for a in (th.low, th.high):
            cb.add(r15(a[6] - a[0]), a[0] if a[0] == 0 else math.copysign(r15(abs(a[0])), a[0]),
                   math.copysign(r15(abs(a[1] / 2.0)), a[1]), math.copysign(r15(abs(a[2] / 6.0)), a[2]),
                   math.copysign(r15(abs(a[3] / 12.0)), a[3]), math.copysign(r15(abs(a[4] / 20.0)), a[4]),
                   -math.copysign(r15(abs(a[5])), a[5]))
    roffs: List[int] = []
    for r in mech.reactions:
        if r.rtype in ("standard", "third_body"):
            off = cb.add(*_arr_vals(*r.arrh))
        elif r.rtype in ("troe", "lindemann", "sri"):
            off = cb.add(*_arr_vals(*r.arrh))
            cb.add(*_arr_vals(*r.low))
            if r.rtype == "troe":
                a, T3, T1 = r.troe[0], r.troe[1], r.troe[2]
                cb.add(r15(1.0 - a), (1.0 / r15(T3)) if T3 != 0.0 else float("inf"), r15(a),
                       (1.0 / r15(T1)) if T1 != 0.0 else float("inf"))
                if len(r.troe) == 4:
                    cb.add(r15(r.troe[3]))
            elif r.rtype == "sri":
                sp = tuple(r.sri) + ((1.0, 0.0) if len(r.sri) == 3 else ())
                cb.add(r15(sp[0]), r15(sp[1]), 1.0 / r15(sp[2]), r15(sp[3]), float(sp[4]))
        elif r.rtype == "plog":
            pts = sorted(set(c[0] for c in r.plog))
            off = len(cb.vals)
            for p in pts:
                rates = [c for c in r.plog if c[0] == p]
                cb.add(r15(math.log(p)), r15(p))
                for (_, A, b, Ea) in rates:
                    cb.add(r15(A), float(b), r15(-Ea))
        else:
            raise ValueError(r.rtype)
        if r.third_body:
            cb.add(*[r15(e) for e in r.third_body.values()])
        roffs.append(off)
    return np.array(cb.vals, dtype=np.float64), off_smh, roffs


def _reaction_code(j: int, r: Reaction, off: int, idx: Dict[str, int]) -> List[str]:
    L = [f"  {{  // R{j}: {_eq(r)}"]
    cr = stoich_product(power_expr(f"c{idx[s]}", n) for s, n in r.reactants.items())
    cp = stoich_product(power_expr(f"c{idx[s]}", n) for s, n in r.products.items())
    # Delta(smh) + dnu*ln(pfac): kb = kf * exp(-(sum_p nu smh - sum_r nu smh) - dnu ln pfac)
    dg_terms = []
    for s, n in r.products.items():
        dg_terms.append(f"+ {'' if n == 1.0 else dlit(n) + ' * '}s{idx[s]}")
    for s, n in r.reactants.items():
        dg_terms.append(f"- {'' if n == 1.0 else dlit(n) + ' * '}s{idx[s]}")
    dnu = r.coeff_sum
    dg = " ".join(dg_terms)
    if dnu != 0.0:
        dg += f" + {dlit(dnu)} * lnpfac"
    if dg.startswith("+ "):
        dg = dg[2:]
    tb_off = None
    if r.rtype in ("standard", "third_body"):
        L.append(f"    const double lk = {_lk(off, r.arrh[1], r.arrh[2])};")
        L.append("    const double kf = cg_exp(lk);")
        if r.reversible:
            L.append(f"    const double kr = cg_exp(lk - ({dg}));")
        tb_off = off + 3
    elif r.rtype in ("troe", "lindemann", "sri"):
        nx = {"troe": 4 + (1 if r.troe and len(r.troe) == 4 else 0), "lindemann": 0, "sri": 5}[r.rtype]
        tb_off = off + 6 + nx
    elif r.rtype == "plog":
        raise NotImplementedError("PLOG reactions are not supported by the HIP target yet")
    M = None
    if r.rtype != "standard":
        # M = ctot + sum (eff-1) C, in the order the efficiencies are listed
        M = "ctot"
        for q, s in enumerate(r.third_body):
            M = f"__builtin_fma(cf[{tb_off + q}], c{idx[s]}, {M})"
        L.append(f"    const double M = {M};")
    if r.rtype in ("troe", "lindemann", "sri"):
        L.append(f"    const double lki = {_lk(off, r.arrh[1], r.arrh[2])};")
        L.append(f"    const double lk0 = {_lk(off + 3, r.low[1], r.low[2])};")
        L.append("    const double kinf = cg_exp(lki);")
        L.append("    const double pr = M * cg_exp(lk0 - lki);")
        o = off + 6
        if r.rtype == "troe":
            fc = f"cf[{o}] * cg_exp(-t * cf[{o + 1}]) + cf[{o + 2}] * cg_exp(-t * cf[{o + 3}])"
            if len(r.troe) == 4:
                fc += f" + cg_exp(-cf[{o + 4}] * invT)"
            L.append(f"    const double fcent = {fc};")
            L.append(f"    const double lf = cg_log(fcent) * {dlit(1.0 / LN10)};")
            L.append(f"    const double lp = cg_log(pr) * {dlit(1.0 / LN10)};")
            L.append("    const double cc = -0.4 - 0.67 * lf;")
            L.append("    const double nn = 0.75 - 1.27 * lf;")
            L.append("    const double f1 = (lp + cc) / (nn - 0.14 * (lp + cc));")
            L.append(f"    const double F = cg_exp({dlit(LN10)} * lf / (1.0 + f1 * f1));")
            L.append("    const double kf = kinf * (pr / (1.0 + pr)) * F;")
        elif r.rtype == "sri":
            L.append(f"    const double lp = cg_log(pr) * {dlit(1.0 / LN10)};")
            L.append("    const double X = 1.0 / (1.0 + lp * lp);")
            L.append(f"    const double base = cf[{o}] * cg_exp(-cf[{o + 1}] * invT) + cg_exp(-t * cf[{o + 2}]);")
            L.append(f"    double F = cf[{o + 3}] * cg_exp(X * cg_log(base));")
            if len(r.sri) == 5 and r.sri[4] != 0.0:
                L.append(f"    F *= cg_exp(cf[{o + 4}] * lnT);")
            L.append("    const double kf = kinf * (pr / (1.0 + pr)) * F;")
        else:
            L.append("    const double kf = kinf * (pr / (1.0 + pr));")
        if r.reversible:
            L.append(f"    const double kr = kf * cg_exp(-({dg}));")
    if r.reversible:
        q = f"kf * {cr} - kr * {cp}"
    else:
        q = f"kf * {cr}"
    if r.rtype == "third_body":
        q = f"M * ({q})"
    L.append(f"    const double q = {q};")
    for s, n in r.reactants.items():
        L.append(f"    w{idx[s]} -= {'q' if n == 1.0 else dlit(n) + ' * q'};")
    for s, n in r.products.items():
        L.append(f"    w{idx[s]} += {'q' if n == 1.0 else dlit(n) + ' * q'};")
    L.append("  }")
    return L


def _eq(r: Reaction) -> str:
    def side(d):
        return " + ".join((s if n == 1.0 else f"{n:g} {s}") for s, n in d.items())
    return f"{side(r.reactants)} {'<=>' if r.reversible else '=>'} {side(r.products)} [{r.rtype}]"


def emit_hip(mech: Mechanism, variant: Variant = Variant()):
    """Render the HIP source.  Returns (source, coef, meta dict)."""
    import jinja2
    coef, off_smh, roffs = build_coef(mech)
    idx = {s.name: i for i, s in enumerate(mech.species)}
    ns = mech.n_species
    lines: List[str] = []
    for j, r in enumerate(mech.reactions):
        lines.extend(_reaction_code(j, r, roffs[j], idx))
    tmids = []
    group_of = []
    for s in mech.species:
        if s.thermo.t_mid not in tmids:
            tmids.append(s.thermo.t_mid)
        group_of.append(tmids.index(s.thermo.t_mid))
    need_ctot = any(r.rtype != "standard" for r in mech.reactions)
    env = jinja2.Environment(loader=jinja2.FileSystemLoader(TEMPLATE_DIR), keep_trailing_newline=True)
    tpl = env.get_template("getrates_hip.j2")
    h = structure_hash(mech)
    src = tpl.render(
        name=mech.name, ns=ns, nr=mech.n_reactions, ncoef=len(coef), block=variant.block,
        waves_per_eu=variant.waves_per_eu, hash_hex=f"0x{h:016x}",
        off_invmw=0, off_smh=off_smh,
        csum_expr=tree_sum([f"c{k}" for k in range(ns)]),
        ctot_expr=tree_sum([f"c{k}" for k in range(ns)]),
        need_ctot=need_ctot, R0=dlit(R0), ln_patm_r0=dlit(math.log(PATM / R0)),
        tmid_groups=[{"tmid": dlit(t)} for t in tmids], tmid_group_of=group_of,
        reactions_code="\n".join(lines),
    )
    meta = dict(ns=ns, nr=mech.n_reactions, ncoef=len(coef), block=variant.block,
                waves_per_eu=variant.waves_per_eu, structure_hash=h)
    return src, coef, meta


HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def compile_code_object(src_path: str, out_path: str, extra: List[str] = ()) -> None:
    """hipcc --genco: a gfx950 code object loadable with hipModuleLoad."""
    cmd = [HIPCC, "--genco", *HIP_FLAGS, *extra, "-o", out_path + ".tmp", src_path]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({' '.join(cmd)}):\n{res.stderr[-4000:]}")
    os.replace(out_path + ".tmp", out_path)
