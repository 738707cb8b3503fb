"""HIP (gfx950) emitter for getrates and the gpu_coef coefficient blob.

The reference renders its HIP kernels from chemgen/templates/reactions_hip/*.j2 and keeps
the mechanism numbers in ``__constant__`` arrays copied by ``copyConstantsToDevice``
(chemgen/constants.py:119-240, chemgen/write_constants.py:408-493).  Here the generator
compiles the mechanism *structure* (which species each reaction touches, its type, its
zero coefficients) into straight-line code, one lane per grid point, and keeps the
*numbers* in a flat FP64 blob (gpu_coef) that the kernel reads through uniform loads:

  [invmw[Ns] | smh[Ns][14] | per-reaction records]

The kernels work in base 2 (one exp2 per Arrhenius rate, no ln 2 reduction step), so the
blob holds base-2 forms; L = log2(e):

smh record per species: low range L*(a6-a0), a0, L*a1/2, L*a2/6, L*a3/12, L*a4/20, -L*a5,
then the high range -- the transform of chemgen/write_constants.py:146-189 scaled by L --
evaluated as ``c0 + c1 log2T + c2 T + c3 T^2 + c4 T^3 + c5 T^4 + c6/T = smh log2(e)``.

Reaction records (kernel order = mechanism order):
  standard/third-body: log2 A, beta, L Ea/Rc, then (eff-1) per listed third-body species
  falloff: the same three for the high- and the low-pressure limit, then
           Troe (1-a, L/T3, a, L/T1[, L T2]) | SRI (a, L b, L/c, d, e) | Lindemann (none),
           then (eff-1) per listed species
  plog: per pressure point (log2 p, p, 1/(log2 p - log2 p_prev) [0 for the first], then per
        rate: sign(A), log2|A|, beta, L Ea/Rc), points in increasing pressure

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
LOG2E = 1.0 / math.log(2.0)
LOG2_10 = math.log2(10.0)
LOG10_2 = math.log10(2.0)

LDS_PER_CU = 160 * 1024

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
    exp_impl: str = "lean"   # "lean": own exp/log (range reduction + polynomials); "ocml": libm
    kr_mode: str = "gh"      # "gh": q = kf prod(g_r) (prod(h_r) - pfac^-dnu prod(h_p)) with
                             #   g = exp(smh), h = C/g per species (no per-reaction division);
                             # "eg": kb from g products and one division; "exp": one exp per kb
    poly: str = "horner"     # exp2 polynomial: "horner" (degree 10), "estrin" (shorter chains), "d9" (degree 9, 1.5e-14)
    g_lds: bool = False      # keep the per-species EG_k in LDS instead of registers
    group: int = 4           # with LDS state: reactions between re-reads from LDS
    split: int = 1           # wavefronts per 64-point tile sharing the species state in LDS
                             # (each runs 1/split of the reactions); 1 = lane-per-point kernel
    table: int = 0           # > 0: table-driven kernel, `table` points per tile and `split`
                             # wavefronts, lanes = (point, reaction slot); for large mechanisms
    baked: bool = False      # compile the coefficient values into the code (no scalar loads that
                             # share the LGKM counter with LDS reads); the blob must then match
    live: bool = False       # lane kernel: reactions ordered to keep few species live, each
                             # species' state computed at its first use and its wdot stored after
                             # its last (register pressure of the lane-per-point kernel)
    pair: bool = False       # straight-line kernels: the exp2 of two consecutive Arrhenius rates
                             # (and of two species' EG) evaluated interleaved, so that each
                             # wavefront has two independent FMA chains in flight
    fdiv: bool = False       # divisions inside log2, h_k = C_k / EG_k and the falloff blending by
                             # a reciprocal with one Newton step (cg_div_fast, a few ulp)
    tp: int = 64             # reaction-split tiles of tp < 64 points: lanes tp..63 repeat lanes
                             # 0..63-tp (same values, same LDS writes), so that the species state
                             # of one more tile fits the LDS of a CU
    bcast: bool = False      # table kernel of 16- or 8-point tiles: per-reaction data (records,
                             # coefficients, efficiencies, PLOG pressures of the compiled class
                             # loops; gather stream words) loaded once per lane group and
                             # broadcast across its lanes (DPP), not loaded by every lane

    @property
    def tag(self) -> str:
        if self.table:
            return (f"t{self.table}s{self.split}w{self.waves_per_eu}{self.exp_impl[0]}{self.poly[0]}"
                    + ("d" if self.fdiv else "") + ("b" if self.bcast else ""))
        c = (("k" if self.baked else "") + ("o" if self.live else "") + ("p" if self.pair else "")
             + ("d" if self.fdiv else ""))
        if self.split > 1:
            return (f"s{self.split}w{self.waves_per_eu}{self.exp_impl[0]}{self.kr_mode}"
                    f"{self.poly[0]}g{self.group}{c}" + (f"n{self.tp}" if self.tp != 64 else ""))
        return (f"b{self.block}w{self.waves_per_eu}{self.exp_impl[0]}{self.kr_mode}"
                f"{self.poly[0]}{'L%d' % self.group if self.g_lds else ''}{c}")

    @property
    def threads(self) -> int:
        return 64 * self.split if (self.split > 1 or self.table) else self.block

    @property
    def ppb(self) -> int:
        """Grid points per workgroup."""
        if self.table:
            return self.table
        return self.tp if self.split > 1 else self.block


class _Coef:
    def __init__(self):
        self.vals: List[float] = []

    def add(self, *xs: float) -> int:
        off = len(self.vals)
        self.vals.extend(float(x) for x in xs)
        return off


def _lk(off: int, beta: float, ea: float) -> str:
    """log2 k = log2 A + beta log2 T - (Ea/(Rc ln2))/T with zero terms compiled out; written
    so that every operation takes at most one coefficient (one scalar operand per VALU op)."""
    terms = None
    if ea != 0.0:
        terms = f"(-cf[{off + 2}] * invT)"
    if beta != 0.0:
        terms = f"(cf[{off + 1}] * lg2T)" if terms is None else f"__builtin_fma(cf[{off + 1}], lg2T, {terms})"
    return f"cf[{off}]" if terms is None else f"({terms} + cf[{off}])"


def _arr_vals(A: float, beta: float, Ea: float, signed: bool = False) -> Tuple[float, float, float]:
    """(log2 |A|, beta, Ea / (Rc ln 2)); A = 0 (a placeholder rate, kf = 0) is stored as
    log2 A = -1100, which exp2 takes to 0.  A < 0 (one of a pair of duplicate elementary
    reactions whose rates are summed) is allowed where ``signed``: the sign is part of the
    mechanism structure and compiled into the kernels (``negative_a``)."""
    if A < 0.0 and not signed:
        raise ValueError(f"Arrhenius prefactor {A} < 0 in a falloff reaction")
    return (math.log2(r15(abs(A))) if A != 0.0 else -1100.0, float(beta), r15(Ea) / RC * LOG2E)


def negative_a(r: Reaction) -> bool:
    """An elementary or third-body reaction with a negative prefactor: kf = -2^(log2|A| + ...)."""
    return r.rtype in ("standard", "third_body") and r.arrh[0] < 0.0


def structure_key(mech: Mechanism) -> str:
    """Everything the compiled code depends on (not the coefficient values)."""
    idx = {s.name: i for i, s in enumerate(mech.species)}
    parts = [f"ns={mech.n_species}", "tmid=" + ",".join(repr(s.thermo.t_mid) for s in mech.species)]
    if getattr(mech, "qss", ()):
        parts.append("qss=" + ",".join(str(idx[s]) for s in mech.qss))
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
        ) + (("negative A",) if negative_a(r) else ())))
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
            L = LOG2E
            cb.add(L * r15(a[6] - a[0]), a[0] if a[0] == 0 else math.copysign(r15(abs(a[0])), a[0]),
                   L * math.copysign(r15(abs(a[1] / 2.0)), a[1]), L * math.copysign(r15(abs(a[2] / 6.0)), a[2]),
                   L * math.copysign(r15(abs(a[3] / 12.0)), a[3]), L * math.copysign(r15(abs(a[4] / 20.0)), a[4]),
                   -L * math.copysign(r15(abs(a[5])), a[5]))
    roffs: List[int] = []
    for r in mech.reactions:
        if r.rtype in ("standard", "third_body"):
            off = cb.add(*_arr_vals(*r.arrh, signed=True))
        elif r.rtype in ("troe", "lindemann", "sri"):
            off = cb.add(*_arr_vals(*r.arrh))
            cb.add(*_arr_vals(*r.low))
            if r.rtype == "troe":
                a, T3, T1 = r.troe[0], r.troe[1], r.troe[2]
                cb.add(r15(1.0 - a), (LOG2E / r15(T3)) if T3 != 0.0 else float("inf"), r15(a),
                       (LOG2E / r15(T1)) if T1 != 0.0 else float("inf"))
                if len(r.troe) == 4:
                    cb.add(LOG2E * r15(r.troe[3]))
            elif r.rtype == "sri":
                sp = tuple(r.sri) + ((1.0, 0.0) if len(r.sri) == 3 else ())
                cb.add(r15(sp[0]), LOG2E * r15(sp[1]), LOG2E / r15(sp[2]), r15(sp[3]), float(sp[4]))
        elif r.rtype == "plog":
            pts = sorted(set(c[0] for c in r.plog))
            off = len(cb.vals)
            prev = None
            for p in pts:
                rates = [c for c in r.plog if c[0] == p]
                lp = math.log2(r15(p))
                cb.add(lp, r15(p), 0.0 if prev is None else 1.0 / (lp - prev))
                prev = lp
                for (_, A, b, Ea) in rates:
                    A = r15(A)
                    cb.add(1.0 if A >= 0 else -1.0, math.log2(abs(A)) if A != 0 else -1100.0, float(b),
                           r15(Ea) / RC * LOG2E)
        else:
            raise ValueError(r.rtype)
        if r.third_body:
            cb.add(*[r15(e) for e in r.third_body.values()])
        roffs.append(off)
    return np.array(cb.vals, dtype=np.float64), off_smh, roffs


def _pfac_pow(dnu: float) -> str:
    """pfac^(-dnu) as a product (kb = kf K_c^-1, K_c = exp(dsmh) pfac^dnu)."""
    if dnu == 0.0:
        return ""
    if float(dnu).is_integer() and abs(dnu) <= 3:
        base = "pinv" if dnu > 0 else "pfac"
        return " * " + " * ".join([base] * int(abs(dnu)))
    return f" * cg_exp2({dlit(-dnu)} * lg2pfac)"


def _qss_terms(r: Reaction, idx: Dict[str, int], qpos: Dict[str, int], stage: int) -> List[str]:
    """After kf (and M): F = kf prod_{transported reactants} C, B = kb prod_{transported
    products} C (times M for three-body), with C = g h and kb = kf prod(g_r) / prod(g_p)
    pfac^-dnu, so B = kf prod(g_r) pfac^-dnu prod_{transported products} h / g_b.  Stage 1
    adds the reaction's terms to the QSS balance rows qA<i>_<k>, qc<i>; stage 2 forms
    q = F x_a - B x_b and updates wdot of the transported species."""
    L: List[str] = []
    a = next((s for s in r.reactants if s in qpos), None)
    b = next((s for s in r.products if s in qpos), None)
    fr = [power_expr(f"(g{idx[s]} * h{idx[s]})", n) for s, n in r.reactants.items() if s not in qpos]
    F = "kf" + "".join(f" * {t}" for t in fr)
    if r.rtype == "third_body":
        F = f"M * ({F})"
    L.append(f"    const double qF = {F};")
    rev = r.reversible
    if rev:
        gr = stoich_product(power_expr(f"g{idx[s]}", n) for s, n in r.reactants.items())
        hp = [power_expr(f"h{idx[s]}", n) for s, n in r.products.items() if s not in qpos]
        pf = _pf_factor(r.coeff_sum)
        B = f"(kf * {gr})" + (f" * {pf}" if pf else "") + "".join(f" * {t}" for t in hp)
        if b is not None:
            B += f" * ig{idx[b]}"
        if r.rtype == "third_body":
            B = f"M * ({B})"
        L.append(f"    const double qB = {B};")
    if stage == 1:
        if a is not None and a == b:
            return L
        if a is not None:
            ia = qpos[a]
            L.append(f"    qA{ia}_{ia} -= qF;")
            if rev:
                L.append(f"    qA{ia}_{qpos[b]} += qB;" if b is not None else f"    qc{ia} += qB;")
        if b is not None:
            ib = qpos[b]
            if rev:
                L.append(f"    qA{ib}_{ib} -= qB;")
            L.append(f"    qA{ib}_{qpos[a]} += qF;" if a is not None else f"    qc{ib} += qF;")
        return L
    fq = f"qF * x{qpos[a]}" if a is not None else "qF"
    if rev:
        bq = f"qB * x{qpos[b]}" if b is not None else "qB"
        L.append(f"    const double q = {fq} - {bq};")
    else:
        L.append(f"    const double q = {fq};")
    for s_, n in r.reactants.items():
        if s_ not in qpos:
            L.append(f"    w{idx[s_]} -= {'q' if n == 1.0 else dlit(n) + ' * q'};")
    for s_, n in r.products.items():
        if s_ not in qpos:
            L.append(f"    w{idx[s_]} += {'q' if n == 1.0 else dlit(n) + ' * q'};")
    return L


def _kr_eg(r: Reaction, idx: Dict[str, int], kf: str) -> str:
    num = stoich_product(power_expr(f"g{idx[s]}", n) for s, n in r.reactants.items())
    den = stoich_product(power_expr(f"g{idx[s]}", n) for s, n in r.products.items())
    return f"cg_div_fast({kf} * {num}, {den}){_pfac_pow(r.coeff_sum)}"


def _pf_factor(dnu: float) -> str:
    """pfac^(-dnu) as one expression ("" for dnu = 0)."""
    if dnu == 0.0:
        return ""
    if float(dnu).is_integer() and abs(dnu) <= 3:
        base = "pinv" if dnu > 0 else "pfac"
        return "(" + " * ".join([base] * int(abs(dnu))) + ")"
    return f"cg_exp2({dlit(-dnu)} * lg2pfac)"


def _plog_code(r: Reaction, off: int) -> List[str]:
    """kf of a PLOG reaction (chemgen/templates/reactions/plog.j2:3-38): log k at each pressure
    point (the sum of its rates), linear in log P between the two points around P, the end
    point's value outside the range.  Written without branches: the point values are all
    evaluated and the interval is picked by selects in increasing pressure."""
    pts = sorted(set(c[0] for c in r.plog))
    L = ["    const double lgP = cg_log2(p);"]
    o = off
    rec = []
    for i, pt in enumerate(pts):
        rates = [c for c in r.plog if c[0] == pt]
        po = o
        o += 3
        if len(rates) == 1:
            expr = f"__builtin_fma(cf[{o + 2}], lg2T, cf[{o + 1}] - cf[{o + 3}] * invT)"
            L.append(f"    const double lkp{i} = {expr};")
            o += 4
        else:
            terms = []
            for _ in rates:
                terms.append(f"cf[{o}] * cg_exp2(__builtin_fma(cf[{o + 2}], lg2T, cf[{o + 1}] - cf[{o + 3}] * invT))")
                o += 4
            L.append(f"    const double lkp{i} = cg_log2({' + '.join(terms)});")
        rec.append(po)
    L.append("    double lk = lkp0;")
    for i in range(1, len(pts)):
        a, b = rec[i - 1], rec[i]
        L.append(f"    lk = p >= cf[{a + 1}] ? __builtin_fma(lkp{i} - lkp{i - 1}, (lgP - cf[{a}]) * cf[{b + 2}], lkp{i - 1}) : lk;")
    L.append(f"    lk = p >= cf[{rec[-1] + 1}] ? lkp{len(pts) - 1} : lk;")
    L.append("    const double kf = cg_exp2(lk);")
    return L


def _reaction_code(j: int, r: Reaction, off: int, idx: Dict[str, int], kr_mode: str = "exp",
                   skip_eff=frozenset(), qss=None, kf_var: str = None, fdiv: bool = False) -> List[str]:
    """Straight-line code of one reaction: kf (and M), q and the wdot updates.  ``skip_eff``:
    species left out of the third-body sum.  ``qss`` = (QSS species -> unknown index, stage):
    stage 1 emits the reaction's terms of the QSS balance, stage 2 its rate with the solved
    QSS concentrations x<i> (see chemgen_amd/qssa.py).  ``kf_var``: an elementary or
    third-body reaction's 2^(log2 k) was computed before (``_paired_code``) into this variable.
    ``fdiv``: the falloff divisions use the one-Newton-step reciprocal."""
    div = "cg_div_fast" if fdiv else "cg_div"
    L = [f"  {{  // R{j}: {_eq(r)}" + (f" [QSS stage {qss[1]}]" if qss else "")]
    if kr_mode == "gh":
        cvar = lambda k: f"(g{k} * h{k})"
    else:
        cvar = lambda k: f"c{k}"
    cr = stoich_product(power_expr(cvar(idx[s]), n) for s, n in r.reactants.items())
    cp = stoich_product(power_expr(cvar(idx[s]), n) for s, n in r.products.items())
    # Delta(smh) + dnu*ln(pfac): kb = kf * exp(-(sum_p nu smh - sum_r nu smh) - dnu ln pfac)
    dg_terms = []
    for s, n in r.products.items():
        dg_terms.append(f"+ {'' if n == 1.0 else dlit(n) + ' * '}s{idx[s]}")
    for s, n in r.reactants.items():
        dg_terms.append(f"- {'' if n == 1.0 else dlit(n) + ' * '}s{idx[s]}")
    dnu = r.coeff_sum
    dg = " ".join(dg_terms)
    if dnu != 0.0:
        dg += f" + {dlit(dnu)} * lg2pfac"
    if dg.startswith("+ "):
        dg = dg[2:]
    tb_off = None
    if r.rtype in ("standard", "third_body"):
        sg = "-" if negative_a(r) else ""
        if kf_var is not None:
            L.append(f"    const double kf = {sg}{kf_var};")
        else:
            L.append(f"    const double lk = {_lk(off, r.arrh[1], r.arrh[2])};")
            L.append(f"    const double kf = {sg}cg_exp2_fin(lk);")
        if r.reversible and kr_mode != "gh":
            if kr_mode == "eg":
                L.append(f"    const double kr = {_kr_eg(r, idx, 'kf')};")
            else:
                L.append(f"    const double kr = {sg}cg_exp2(lk - ({dg}));")
        tb_off = off + 3
    elif r.rtype in ("troe", "lindemann", "sri"):
        nx = {"troe": 4 + (1 if r.troe and len(r.troe) == 4 else 0), "lindemann": 0, "sri": 5}[r.rtype]
        tb_off = off + 6 + nx
    elif r.rtype == "plog":
        L.extend(_plog_code(r, off))
    if r.rtype == "plog" and r.reversible and kr_mode != "gh":
        if kr_mode == "eg":
            L.append(f"    const double kr = {_kr_eg(r, idx, 'kf')};")
        else:
            L.append(f"    const double kr = kf * cg_exp2(-({dg}));")
    M = None
    if r.rtype not in ("standard", "plog"):
        # M = ctot + sum (eff-1) C, in the order the efficiencies are listed
        M = "ctot"
        for q, s in enumerate(r.third_body):
            if s in skip_eff:
                continue
            M = f"__builtin_fma(cf[{tb_off + q}], {cvar(idx[s])}, {M})"
        L.append(f"    const double M = {M};")
    if r.rtype in ("troe", "lindemann", "sri"):
        L.append(f"    const double lki = {_lk(off, r.arrh[1], r.arrh[2])};")
        L.append(f"    const double lk0 = {_lk(off + 3, r.low[1], r.low[2])};")
        if r.rtype == "lindemann":
            L.append("    const double kinf = cg_exp2_fin(lki);")
        L.append("    const double pr = M * cg_exp2_fin(lk0 - lki);")
        o = off + 6
        if r.rtype == "troe":
            fc = f"cf[{o}] * cg_exp2(-t * cf[{o + 1}]) + cf[{o + 2}] * cg_exp2(-t * cf[{o + 3}])"
            if len(r.troe) == 4:
                fc += f" + cg_exp2(-cf[{o + 4}] * invT)"
            L.append(f"    const double fcent = {fc};")
            L.append("    const double l2f = cg_log2(fcent);")
            L.append(f"    const double lf = l2f * {dlit(LOG10_2)};            // log10 Fcent")
            L.append(f"    const double lp = cg_log2(pr) * {dlit(LOG10_2)};     // log10 Pr")
            L.append("    const double cc = -0.4 - 0.67 * lf;")
            L.append("    const double nn = 0.75 - 1.27 * lf;")
            # log10 F = log10 Fcent / (1 + (x/d)^2) = log10 Fcent d^2 / (d^2 + x^2): one division
            L.append("    const double x = lp + cc, d = nn - 0.14 * x;")
            # k_inf F = 2^(log2 k_inf + log2 F): one exp2 for both
            L.append(f"    const double l2F = {div}(l2f * (d * d), __builtin_fma(x, x, d * d));")
            L.append(f"    const double kf = cg_exp2(lki + l2F) * {div}(pr, 1.0 + pr);")
        elif r.rtype == "sri":
            L.append(f"    const double lp = cg_log2(pr) * {dlit(LOG10_2)};")
            L.append(f"    const double X = {div}(1.0, 1.0 + lp * lp);")
            L.append(f"    const double base = cf[{o}] * cg_exp2(-cf[{o + 1}] * invT) + cg_exp2(-t * cf[{o + 2}]);")
            # k_inf F = d 2^(log2 k_inf + X log2 base [+ e log2 T]): one exp2 for all three
            l2F = "X * cg_log2(base)"
            if len(r.sri) == 5:   # compiled for e = 0 too: e is a coefficient, not structure
                l2F = f"__builtin_fma(cf[{o + 4}], lg2T, {l2F})"
            L.append(f"    const double kf = cf[{o + 3}] * cg_exp2(lki + {l2F}) * {div}(pr, 1.0 + pr);")
        else:
            L.append(f"    const double kf = kinf * {div}(pr, 1.0 + pr);")
        if r.reversible and kr_mode != "gh":
            if kr_mode == "eg":
                L.append(f"    const double kr = {_kr_eg(r, idx, 'kf')};")
            else:
                L.append(f"    const double kr = kf * cg_exp2(-({dg}));")
    if qss is not None:
        L.extend(_qss_terms(r, idx, qss[0], qss[1]))
        L.append("  }")
        return L
    if kr_mode == "gh":
        # q = kf prod(C_r) - kb prod(C_p) with C = g h and kb = kf prod(g_r)/prod(g_p) pfac^-dnu:
        # q = kf prod(g_r) (prod(h_r) - pfac^-dnu prod(h_p))
        gr = stoich_product(power_expr(f"g{idx[s]}", n) for s, n in r.reactants.items())
        hr = stoich_product(power_expr(f"h{idx[s]}", n) for s, n in r.reactants.items())
        hp = stoich_product(power_expr(f"h{idx[s]}", n) for s, n in r.products.items())
        if r.reversible:
            pf = _pf_factor(r.coeff_sum)
            diff = f"__builtin_fma(-{pf}, {hp}, {hr})" if pf else f"({hr} - {hp})"
            q = f"(kf * {gr}) * {diff}"
        else:
            q = f"(kf * {gr}) * ({hr})"
    elif r.reversible:
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


def _pairable(r: Reaction, kr_mode: str) -> bool:
    """An Arrhenius rate whose only transcendental is its 2^(log2 k)."""
    return r.rtype in ("standard", "third_body") and kr_mode == "gh"


def _reactions_code(js: List[int], mech: Mechanism, roffs: List[int], idx: Dict[str, int], kr_mode: str,
                    pair: bool, indent: str = "", fdiv: bool = False) -> List[List[str]]:
    """Code blocks of the reactions ``js`` in order.  With ``pair``, two consecutive Arrhenius
    rates share one block whose two exp2 evaluations are interleaved (cg_exp2_fin2), then each
    reaction's body follows as before; returns one block per reaction or pair."""
    out: List[List[str]] = []
    q = 0
    while q < len(js):
        j = js[q]
        r = mech.reactions[j]
        if pair and q + 1 < len(js) and _pairable(r, kr_mode) and _pairable(mech.reactions[js[q + 1]], kr_mode):
            j2 = js[q + 1]
            r2 = mech.reactions[j2]
            L = [f"  {{  // R{j} and R{j2}: their rates' exp2 interleaved",
                 "    double kfp0, kfp1;",
                 f"    cg_exp2_fin2({_lk(roffs[j], r.arrh[1], r.arrh[2])}, "
                 f"{_lk(roffs[j2], r2.arrh[1], r2.arrh[2])}, kfp0, kfp1);"]
            L += _reaction_code(j, r, roffs[j], idx, kr_mode, kf_var="kfp0", fdiv=fdiv)
            L += _reaction_code(j2, r2, roffs[j2], idx, kr_mode, kf_var="kfp1", fdiv=fdiv)
            L.append("  }")
            out.append([indent + ln for ln in L])
            q += 2
        else:
            out.append([indent + ln for ln in _reaction_code(j, r, roffs[j], idx, kr_mode, fdiv=fdiv)])
            q += 1
    return out


def _eq(r: Reaction) -> str:
    def side(d):
        return " + ".join((s if n == 1.0 else f"{n:g} {s}") for s, n in d.items())
    return f"{side(r.reactants)} {'<=>' if r.reversible else '=>'} {side(r.products)} [{r.rtype}]"


def reaction_cost(r: Reaction) -> float:
    """Rough VALU instruction count of one generated reaction (for load balancing)."""
    n = len(r.reactants) + len(r.products)
    c = 30.0 + 3 * n + (12.0 if r.reversible else 0.0)
    if r.rtype != "standard":
        c += 2 * len(r.third_body) + 2
    c += {"troe": 120.0, "sri": 110.0, "lindemann": 35.0}.get(r.rtype, 0.0)
    return c


def partition_reactions(mech: Mechanism, nparts: int, tol: float = 1.02, iters: int = 40000,
                        seed: int = 0) -> List[List[int]]:
    """Split reactions over wavefronts.  Every wavefront keeps one partial-wdot register pair
    per species its reactions touch, so the split minimises the largest touched set under a
    cost balance (each part at most ``tol`` x the mean cost): a greedy start (costliest
    reaction first, into the part it adds the fewest new species to), then seeded random
    single-reaction moves, kept when the touched-set sizes sorted in decreasing order do not
    get worse (lexicographically)."""
    import random
    if nparts <= 1:
        return [list(range(mech.n_reactions))]
    idx = {s.name: i for i, s in enumerate(mech.species)}
    S = [frozenset({idx[s] for s in r.reactants} | {idx[s] for s in r.products}) for r in mech.reactions]
    cost = [reaction_cost(r) for r in mech.reactions]
    cap = sum(cost) / nparts * tol
    load = [0.0] * nparts
    where = [0] * mech.n_reactions
    cnt: List[Dict[int, int]] = [dict() for _ in range(nparts)]

    def put(j, v, sign):
        load[v] += sign * cost[j]
        for k in S[j]:
            c = cnt[v].get(k, 0) + sign
            if c:
                cnt[v][k] = c
            else:
                del cnt[v][k]

    for j in sorted(range(mech.n_reactions), key=lambda j: -cost[j]):
        fits = [v for v in range(nparts) if load[v] + cost[j] <= cap]
        v = (min(fits, key=lambda u: (len(S[j] - cnt[u].keys()), load[u])) if fits
             else min(range(nparts), key=lambda u: load[u]))
        where[j] = v
        put(j, v, 1)

    def score():
        return sorted((len(c) for c in cnt), reverse=True)

    rng = random.Random(seed)
    best = score()
    for _ in range(iters):
        j = rng.randrange(mech.n_reactions)
        a, b = where[j], rng.randrange(nparts)
        if a == b or load[b] + cost[j] > cap:
            continue
        put(j, a, -1)
        put(j, b, 1)
        sc = score()
        if sc <= best:
            best, where[j] = sc, b
        else:
            put(j, b, -1)
            put(j, a, 1)
    parts: List[List[int]] = [[] for _ in range(nparts)]
    for j in range(mech.n_reactions):
        parts[where[j]].append(j)
    return parts


def coef_hash(coef: np.ndarray) -> int:
    """FNV-1a (64 bit) over the bytes of the FP64 blob, as cg_mech_load computes it."""
    h = 0xCBF29CE484222325
    for b in np.ascontiguousarray(coef, dtype=np.float64).tobytes():
        h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def _baked_args(coef: np.ndarray, variant: "Variant") -> dict:
    if not variant.baked:
        return {"baked": False}
    vals = ",\n".join(", ".join(repr(float(x)) if math.isfinite(x) else ("__builtin_inf()" if x > 0 else "-__builtin_inf()")
                                for x in coef[i:i + 4]) for i in range(0, len(coef), 4))
    return {"baked": True, "baked_values": vals, "coef_hash_hex": f"0x{coef_hash(coef):016x}"}


def emit_hip(mech: Mechanism, variant: Variant = Variant()):
    """Render the HIP source.  Returns (source, coef, meta dict)."""
    if variant.table:
        return _emit_table(mech, variant)
    if variant.split > 1:
        return _emit_split(mech, variant)
    if getattr(mech, "qss", ()):
        return _emit_qss(mech, variant)
    if variant.live:
        return _emit_live(mech, variant)
    import jinja2
    coef, off_smh, roffs = build_coef(mech)
    idx = {s.name: i for i, s in enumerate(mech.species)}
    ns = mech.n_species
    lines: List[str] = []
    for j, r in enumerate(mech.reactions):
        if variant.g_lds and j > 0 and j % max(1, variant.group) == 0:
            lines.append("  CG_FENCE();")
        lines.extend(_reaction_code(j, r, roffs[j], idx, variant.kr_mode, fdiv=variant.fdiv))
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
        need_ctot=need_ctot, R0=dlit(R0), lg2_patm_r0=dlit(math.log2(PATM / R0)),
        patm_r0=dlit(PATM / R0), r0_patm=dlit(R0 / PATM),
        exp_impl=variant.exp_impl, kr_mode=variant.kr_mode, poly=variant.poly,
        g_lds=variant.g_lds and variant.kr_mode in ("eg", "gh"),
        tmid_groups=[{"tmid": dlit(t)} for t in tmids], tmid_group_of=group_of,
        reactions_code="\n".join(lines), threads=variant.threads, ppb=variant.ppb, fdiv=variant.fdiv,
        **_baked_args(coef, variant),
    )
    meta = dict(ns=ns, nr=mech.n_reactions, ncoef=len(coef), block=variant.block,
                waves_per_eu=variant.waves_per_eu, structure_hash=h)
    return src, coef, meta


def live_order(mech: Mechanism) -> Tuple[List[int], int]:
    """Reaction order for the lane kernel that keeps the set of live species small: greedily take
    the reaction that brings in the fewest species not yet live minus the species it retires
    (its last use).  A species (its EG, h and wdot registers) is live from its first to its last
    reaction.  Returns (order, the largest live set)."""
    idx = {s.name: i for i, s in enumerate(mech.species)}
    S = [frozenset(idx[x] for x in list(r.reactants) + list(r.products) + list(r.third_body))
         for r in mech.reactions]
    remaining = [0] * mech.n_species
    for sp in S:
        for k in sp:
            remaining[k] += 1
    live: set = set()
    left = set(range(mech.n_reactions))
    order: List[int] = []
    peak = 0
    while left:
        def score(j):
            return (len(S[j] - live) - sum(1 for k in S[j] if remaining[k] == 1), j)
        j = min(left, key=score)
        left.remove(j)
        order.append(j)
        live |= S[j]
        peak = max(peak, len(live))
        for k in S[j]:
            remaining[k] -= 1
            if remaining[k] == 0:
                live.discard(k)
    return order, peak


def _emit_live(mech: Mechanism, variant: Variant):
    import jinja2
    coef, off_smh, roffs = build_coef(mech)
    idx = {s.name: i for i, s in enumerate(mech.species)}
    ns = mech.n_species
    order, peak = live_order(mech)
    S = [set(idx[x] for x in list(r.reactants) + list(r.products) + list(r.third_body)) for r in mech.reactions]
    last = {}
    for pos, j in enumerate(order):
        for k in S[j]:
            last[k] = pos
    tmids, group_of = [], []
    for sp in mech.species:
        if sp.thermo.t_mid not in tmids:
            tmids.append(sp.thermo.t_mid)
        group_of.append(tmids.index(sp.thermo.t_mid))
    lines: List[str] = []
    for k in range(ns):
        if k not in last:
            lines.append(f"  W[{k} * ld + i] = 0.0;  // species {mech.species[k].name} takes part in no reaction")
    started: set = set()
    for pos, j in enumerate(order):
        for k in sorted(S[j] - started):
            started.add(k)
            lines += [f"  // species {k} ({mech.species[k].name}): state at its first use",
                      f"  double g{k}, h{k};",
                      "  {",
                      f"    const double ck = Y[{k} * ld + i] * cf[{k}] * scale;",
                      f"    const auto* a = cf + {off_smh + 14 * k};",
                      "    double hl = a[6] * invT, hh = a[13] * invT;",
                      "    hl = __builtin_fma(a[1], lg2T, hl); hh = __builtin_fma(a[8], lg2T, hh);",
                      "    hl = __builtin_fma(a[2], t, hl);   hh = __builtin_fma(a[9], t, hh);",
                      "    hl = __builtin_fma(a[3], t2, hl);  hh = __builtin_fma(a[10], t2, hh);",
                      "    hl = __builtin_fma(a[4], t3, hl);  hh = __builtin_fma(a[11], t3, hh);",
                      "    hl = __builtin_fma(a[5], t4, hl);  hh = __builtin_fma(a[12], t4, hh);",
                      "    hl = hl + a[0];                    hh = hh + a[7];",
                      f"    g{k} = cg_exp2_fin(lo{group_of[k]} ? hl : hh);",
                      f"    h{k} = {'cg_div_fast' if variant.fdiv else 'cg_div'}(ck, g{k});",
                      "  }",
                      f"  double w{k} = 0.0;"]
        lines.extend(_reaction_code(j, mech.reactions[j], roffs[j], idx, "gh", fdiv=variant.fdiv))
        for k in sorted(S[j]):
            if last[k] == pos:
                lines.append(f"  W[{k} * ld + i] = w{k};")
    env = jinja2.Environment(loader=jinja2.FileSystemLoader(TEMPLATE_DIR), keep_trailing_newline=True)
    tpl = env.get_template("getrates_live_hip.j2")
    h = structure_hash(mech)
    src = tpl.render(
        name=mech.name, ns=ns, nr=mech.n_reactions, ncoef=len(coef), block=variant.block,
        waves_per_eu=variant.waves_per_eu, hash_hex=f"0x{h:016x}", peak=peak,
        csum_expr=tree_sum([f"Y[{k} * ld + i] * cf[{k}]" for k in range(ns)]),
        R0=dlit(R0), lg2_patm_r0=dlit(math.log2(PATM / R0)),
        patm_r0=dlit(PATM / R0), r0_patm=dlit(R0 / PATM), exp_impl=variant.exp_impl, poly=variant.poly,
        tmid_groups=[{"tmid": dlit(t)} for t in tmids], program_code="\n".join(lines),
        threads=variant.threads, ppb=variant.ppb, fdiv=variant.fdiv, **_baked_args(coef, variant),
    )
    meta = dict(ns=ns, nr=mech.n_reactions, ncoef=len(coef), block=variant.block,
                waves_per_eu=variant.waves_per_eu, structure_hash=h, peak_live=peak)
    return src, coef, meta


def _emit_qss(mech, variant: Variant):
    """Lane-per-point kernel of a mechanism with QSS species (chemgen_amd/qssa.py): Y and
    wdot hold the transported species only; the QSS concentrations come from the linear QSS
    balance, assembled in registers (stage 1 over the reactions that involve a QSS species)
    and solved by Gaussian elimination without pivoting before any wdot accumulator is live;
    then every reaction runs in mechanism order, those with QSS species as stage 2."""
    import jinja2
    if variant.split != 1 or variant.table or variant.g_lds or variant.live:
        raise ValueError("QSS species: lane-per-point kernels only")
    coef, off_smh, roffs = build_coef(mech)
    idx = {s.name: i for i, s in enumerate(mech.species)}
    qset = set(mech.qss)
    qpos = {s: i for i, s in enumerate(mech.qss)}
    nq = len(mech.qss)
    R = [k for k, s in enumerate(mech.species) if s.name not in qset]
    row = {k: r for r, k in enumerate(R)}
    tmids, group_of = [], []
    for sp in mech.species:
        if sp.thermo.t_mid not in tmids:
            tmids.append(sp.thermo.t_mid)
        group_of.append(tmids.index(sp.thermo.t_mid))
    involved = [j for j, r in enumerate(mech.reactions)
                if any(s in qset for s in list(r.reactants) + list(r.products))]
    # every QSS species must be consumed somewhere, or its balance row has no diagonal
    consumed = set()
    for j in involved:
        r = mech.reactions[j]
        a = next((s for s in r.reactants if s in qset), None)
        b = next((s for s in r.products if s in qset), None)
        if a is not None and a != b:
            consumed.add(a)
        if b is not None and a != b and r.reversible:
            consumed.add(b)
    if consumed != qset:
        raise ValueError(f"QSS species never consumed: {sorted(qset - consumed)}")
    L: List[str] = []
    for k, sp in enumerate(mech.species):
        L += [f"  // species {k} ({sp.name}){' [QSS]' if sp.name in qset else ''}",
              f"  double g{k};", "  {",
              f"    const auto* a = cf + {off_smh + 14 * k};",
              "    double hl = a[6] * invT, hh = a[13] * invT;",
              "    hl = __builtin_fma(a[1], lg2T, hl); hh = __builtin_fma(a[8], lg2T, hh);",
              "    hl = __builtin_fma(a[2], t, hl);   hh = __builtin_fma(a[9], t, hh);",
              "    hl = __builtin_fma(a[3], t2, hl);  hh = __builtin_fma(a[10], t2, hh);",
              "    hl = __builtin_fma(a[4], t3, hl);  hh = __builtin_fma(a[11], t3, hh);",
              "    hl = __builtin_fma(a[5], t4, hl);  hh = __builtin_fma(a[12], t4, hh);",
              "    hl = hl + a[0];                    hh = hh + a[7];",
              f"    g{k} = cg_exp2_fin(lo{group_of[k]} ? hl : hh);", "  }"]
        if sp.name in qset:
            L.append(f"  const double ig{k} = cg_div(1.0, g{k});")
        else:
            L.append(f"  const double h{k} = cg_div(Y[{row[k]} * ld + i] * cf[{k}] * scale, g{k});")
    # stage 1 (the QSS balance) and the solve come first, before any wdot accumulator is live
    for i in range(nq):
        L.append("  double " + ", ".join(f"qA{i}_{k} = 0.0" for k in range(nq)) + f", qc{i} = 0.0;")
    for j in involved:
        L.extend(_reaction_code(j, mech.reactions[j], roffs[j], idx, "gh", qset, (qpos, 1)))
    # A x = -c by Gaussian elimination (no pivoting), fill-in tracked structurally
    nz = set()
    for j in involved:
        r = mech.reactions[j]
        a = next((s for s in r.reactants if s in qset), None)
        b = next((s for s in r.products if s in qset), None)
        if a is not None and a == b:
            continue
        if a is not None:
            nz.add((qpos[a], qpos[a]))
            if b is not None and r.reversible:
                nz.add((qpos[a], qpos[b]))
        if b is not None:
            if r.reversible:
                nz.add((qpos[b], qpos[b]))
            if a is not None:
                nz.add((qpos[b], qpos[a]))
    L.append("  // QSS balance A x = -c: forward elimination")
    for p in range(nq):
        L.append(f"  const double qp{p} = cg_div(1.0, qA{p}_{p});")
        for i in range(p + 1, nq):
            if (i, p) not in nz:
                continue
            L.append(f"  {{ const double f = qA{i}_{p} * qp{p};")
            for k in range(p + 1, nq):
                if (p, k) in nz:
                    L.append(f"    qA{i}_{k} = __builtin_fma(-f, qA{p}_{k}, qA{i}_{k});")
                    nz.add((i, k))
            L.append(f"    qc{i} = __builtin_fma(-f, qc{p}, qc{i}); }}")
    L.append("  // back substitution: the QSS concentrations")
    for p in reversed(range(nq)):
        expr = f"-qc{p}"
        for k in range(p + 1, nq):
            if (p, k) in nz:
                expr = f"__builtin_fma(-qA{p}_{k}, x{k}, {expr})"
        L.append(f"  const double x{p} = ({expr}) * qp{p};")
    # stage 2 recomputes the rates of the QSS reactions from fresh (opaque) copies of the
    # point's inputs, so the compiler does not carry every stage-1 rate across the solve
    names = ("t", "lg2T", "invT", "ctot", "pfac", "pinv", "lg2pfac")
    L.append("  " + " ".join(f"double w{k} = 0.0;" for k in R))
    L.append("  {")
    L.append("    double " + ", ".join(f"o_{n} = {n}" for n in names) + ";")
    L.append('    __asm__ volatile("" : ' + ", ".join(f'"+v"(o_{n})' for n in names) + ");")
    L.append("    const double " + ", ".join(f"{n} = o_{n}" for n in names) + ";")
    L.append("    " + " ".join(f"(void){n};" for n in names))
    for j, r in enumerate(mech.reactions):
        if j in involved:
            L.extend("  " + ln for ln in _reaction_code(j, r, roffs[j], idx, "gh", qset, (qpos, 2)))
        else:
            L.extend("  " + ln for ln in _reaction_code(j, r, roffs[j], idx, "gh", qset))
    L.append("  }")
    for k in R:
        L.append(f"  W[{row[k]} * ld + i] = w{k};")
    env = jinja2.Environment(loader=jinja2.FileSystemLoader(TEMPLATE_DIR), keep_trailing_newline=True)
    tpl = env.get_template("getrates_live_hip.j2")
    h = structure_hash(mech)
    note = (f"// FP64 getrates for gfx950 with {nq} quasi-steady-state species ({', '.join(mech.qss)}):\n"
            f"// one lane per grid point; Y and wdot hold the {len(R)} transported species (rows in\n"
            f"// mechanism order without the QSS species); the QSS concentrations solve the linear QSS\n"
            f"// balance of the {len(involved)} reactions that involve them.")
    src = tpl.render(
        name=mech.name, ns=len(R), nr=mech.n_reactions, ncoef=len(coef), block=variant.block,
        waves_per_eu=variant.waves_per_eu, hash_hex=f"0x{h:016x}", peak=0, note=note,
        csum_expr=tree_sum([f"Y[{row[k]} * ld + i] * cf[{k}]" for k in R]),
        R0=dlit(R0), lg2_patm_r0=dlit(math.log2(PATM / R0)),
        patm_r0=dlit(PATM / R0), r0_patm=dlit(R0 / PATM), exp_impl=variant.exp_impl, poly=variant.poly,
        tmid_groups=[{"tmid": dlit(t)} for t in tmids], program_code="\n".join(L),
        threads=variant.threads, ppb=variant.ppb, **_baked_args(coef, variant),
    )
    meta = dict(ns=len(R), nr=mech.n_reactions, ncoef=len(coef), block=variant.block,
                waves_per_eu=variant.waves_per_eu, structure_hash=h, n_qss=nq, qss_reactions=len(involved))
    return src, coef, meta


def _emit_split(mech: Mechanism, variant: Variant):
    import jinja2
    coef, off_smh, roffs = build_coef(mech)
    idx = {s.name: i for i, s in enumerate(mech.species)}
    ns = mech.n_species
    W = variant.split
    parts = partition_reactions(mech, W)
    # the per-wave partial sums get LDS rows of their own unless that LDS would keep the
    # workgroups per CU below the occupancy the registers are sized for
    tp = variant.tp
    if not 32 <= tp <= 64:
        raise ValueError("split kernel: 32 to 64 points per tile")
    lds_full = 8 * (2 * ns * tp + W * tp)
    compact_lds = (W <= ns and (LDS_PER_CU // lds_full) * W < 4 * variant.waves_per_eu
                   and (LDS_PER_CU // (8 * 2 * ns * tp)) > LDS_PER_CU // lds_full)
    codes, touched = [], []
    for v in range(W):
        lines: List[str] = []
        tv = set()
        gsz = max(1, variant.group)
        for g0 in range(0, len(parts[v]), gsz):
            if g0 > 0:
                lines.append("    CG_FENCE();")
            for blk in _reactions_code(parts[v][g0:g0 + gsz], mech, roffs, idx, variant.kr_mode,
                                       variant.pair, "  ", fdiv=variant.fdiv):
                lines.extend(blk)
        for j in parts[v]:
            r = mech.reactions[j]
            tv |= {idx[s] for s in r.reactants} | {idx[s] for s in r.products}
        codes.append("\n".join(lines))
        touched.append(sorted(tv))
    bounds = [round(ns * v / W) for v in range(W + 1)]
    species_of = [list(range(bounds[v], bounds[v + 1])) for v in range(W)]
    tmids, group_of = [], []
    for s in mech.species:
        if s.thermo.t_mid not in tmids:
            tmids.append(s.thermo.t_mid)
        group_of.append(tmids.index(s.thermo.t_mid))
    env = jinja2.Environment(loader=jinja2.FileSystemLoader(TEMPLATE_DIR), keep_trailing_newline=True)
    tpl = env.get_template("getrates_split_hip.j2")
    h = structure_hash(mech)
    src = tpl.render(
        name=mech.name, ns=ns, nr=mech.n_reactions, ncoef=len(coef), W=W, threads=variant.threads,
        ppb=variant.ppb, waves_per_eu=variant.waves_per_eu, hash_hex=f"0x{h:016x}", off_invmw=0,
        off_smh=off_smh, R0=dlit(R0), lg2_patm_r0=dlit(math.log2(PATM / R0)), patm_r0=dlit(PATM / R0),
        r0_patm=dlit(R0 / PATM), exp_impl=variant.exp_impl, kr_mode=variant.kr_mode, poly=variant.poly,
        tmid_groups=[{"tmid": dlit(t)} for t in tmids], tmid_group_of=group_of, species_of=species_of,
        zero_of=species_of, touched=touched, reactions_code=codes, compact_lds=compact_lds,
        pair=variant.pair, fdiv=variant.fdiv, tp=tp,
        **_baked_args(coef, variant),
    )
    meta = dict(ns=ns, nr=mech.n_reactions, ncoef=len(coef), block=variant.threads, ppb=variant.ppb,
                waves_per_eu=variant.waves_per_eu, structure_hash=h,
                touched_per_wave=[len(t) for t in touched], compact_lds=compact_lds)
    return src, coef, meta


_TABLE_KIND = {"standard": 0, "third_body": 1, "lindemann": 2, "troe": 3, "sri": 5, "plog": 7}
TABLE_LONG = 1 << 12      # class bit of a reaction written as an extended record
TABLE_NEG = 1 << 13       # class bit of a reaction with a negative prefactor (kf < 0)


def _table_long(r: Reaction) -> bool:
    """Whether a reaction needs the extended record: a side with more than three species
    slots, or a non-integer stoichiometric coefficient."""
    for side in (r.reactants, r.products):
        if any(not float(nu).is_integer() for nu in side.values()):
            return True
        if sum(int(nu) for nu in side.values()) > 3:
            return True
    return False


def _table_slots(r: Reaction, idx: Dict[str, int]) -> Tuple[List[int], List[int]]:
    out = []
    for side in (r.reactants, r.products):
        sl: List[int] = []
        for sp, nu in side.items():
            if not float(nu).is_integer():
                raise ValueError(f"table kernel: stoichiometric coefficient {nu} needs an extended record")
            sl += [idx[sp]] * int(nu)
        if not 1 <= len(sl) <= 3:
            raise ValueError("table kernel: a side of more than 3 species slots needs an extended record")
        out.append(sl)
    return out[0], out[1]


def _table_kind(r: Reaction) -> int:
    kind = _TABLE_KIND[r.rtype]
    if r.rtype == "troe" and len(r.troe) == 4:
        kind = 4
    if r.rtype == "sri" and len(r.sri) == 5:   # e read from the blob, also when it is 0
        kind = 6
    return kind


def table_class(mech: Mechanism, j: int) -> Tuple[int, int, int, int, int, int]:
    """The shape the table kernel compiles a loop for: (family, reversible, reactant slots,
    product slots, dnu if reversible else 0, negative prefactor); family 0 elementary,
    1 third-body, 2 Lindemann, 3 Troe (with or without T2), 5 SRI (with or without e), 7 PLOG.
    A reaction that needs an extended record has slot counts 0; neither it nor one with a
    negative prefactor ever gets a loop of its own."""
    r = mech.reactions[j]
    fam = _TABLE_KIND[r.rtype]
    neg = int(negative_a(r))
    if _table_long(r):
        return (fam, int(r.reversible), 0, 0, 0, neg)
    idx = {s.name: i for i, s in enumerate(mech.species)}
    rs, ps = _table_slots(r, idx)
    return (fam, int(r.reversible), len(rs), len(ps), (len(ps) - len(rs)) if r.reversible else 0, neg)


def table_order(mech: Mechanism, min_class: int = 0):
    """Processing order of the table kernel and its class segments.  Reactions of a class
    (``table_class``) with at least ``min_class`` members are contiguous and get a loop of
    their own, compiled for that shape; the rest follow in one generic segment, sorted by
    (extended record, kind, reversibility, slot counts, efficiency count) so that the lane
    groups of one wavefront mostly take the same branches.  Returns (order, [(start, end,
    class)...], start of the generic segment)."""
    by: Dict[Tuple, List[int]] = {}
    for j in range(mech.n_reactions):
        by.setdefault(table_class(mech, j), []).append(j)
    order: List[int] = []
    segs = []
    compiled = lambda key: key[2] > 0 and not key[5] and len(by[key]) >= max(1, min_class)
    for key in sorted(by):
        if compiled(key):
            segs.append((len(order), len(order) + len(by[key]), key))
            order.extend(by[key])
    rest = sorted((j for key in by if not compiled(key) for j in by[key]),
                  key=lambda j: (_table_long(mech.reactions[j]), negative_a(mech.reactions[j]),
                                 _table_kind(mech.reactions[j]),
                                 mech.reactions[j].reversible, table_class(mech, j)[2:4],
                                 len(mech.reactions[j].third_body), j))
    gen_start = len(order)
    order.extend(rest)
    return order, segs, gen_start


@dataclass
class TableLayout:
    """What the table kernel reads besides the gpu_coef blob (see ``table_layout``)."""
    rows: List[Tuple[int, int, int, int]]
    effsp: List[int]
    order: List[int]
    segs: List[Tuple[int, int, Tuple]]
    gen_start: int
    plog: List[Tuple[int, int]]   # per PLOG pressure point: (blob offset of the point, rates)
    xsl: List[int]                # extended records: [n reactant, n product entries, rows...]
    xnu: List[float]              # parallel to xsl: [dnu, 0, nu of each entry...]


def table_layout(mech: Mechanism, npt: int = 16, min_class: int = 0) -> TableLayout:
    """The table kernel's reaction table for ``npt``-point tiles.  Each side of a reaction is
    written as up to three species *slots* (a species with coefficient nu fills nu slots),
    absent slots point at an extra LDS row that holds EG = h = 1; slots are stored as LDS byte
    offsets of the rows (species index x npt x 8; below 2^16 whenever the tile fits the LDS).
      row 0: x = byte offset of the reaction's record in the blob, y = r0 | r1 << 16,
             z = r2 | p0 << 16, w = p1 | p2 << 16
      row 1: x = kind | reversible << 3 | reactant slots << 4 | product slots << 6
                 | efficiencies << 8 | extended << 12 | negative prefactor << 13
                 | (dnu + 8) << 24,
             y = offset of the collision-efficiency species in the efficiency list (PLOG: of
             its pressure points in the PLOG list), z = blob offset of their (eff - 1) values
             (PLOG: the number of pressure points), w = offset of an extended record in
             xsl / xnu (0 otherwise)
    kind: 0 elementary, 1 third-body, 2 Lindemann, 3 Troe, 4 Troe with T2, 5 SRI, 6 SRI with
    e, 7 PLOG.  A reaction with a side of more than three slots or a non-integer coefficient
    is an extended record: its row-0 slots are unused (the padding row), and xsl[w:] holds the
    number of reactant and of product entries and then each entry's LDS row offset in doubles
    (species index x npt), xnu[w:]
    its dnu (sum of product minus reactant coefficients), 0 and each entry's coefficient.
    Raises ValueError for what the format does not hold (more than 15 efficiencies, an LDS
    row offset beyond 16 bits)."""
    coef, off_smh, roffs = build_coef(mech)
    idx = {s.name: i for i, s in enumerate(mech.species)}
    pad = mech.n_species
    if pad * npt * 8 >= 1 << 16:
        raise ValueError("table kernel: LDS row byte offsets exceed 16 bits")
    for r in mech.reactions:
        if r.rtype not in _TABLE_KIND:
            raise ValueError(f"table kernel: reaction type {r.rtype} not supported")
        if len(r.third_body) > 15:
            raise ValueError("table kernel: more than 15 collision efficiencies")
    rows: List[Tuple[int, int, int, int]] = []
    effsp: List[int] = []
    plog: List[Tuple[int, int]] = []
    xsl: List[int] = [0, 0]           # offset 0: no extended record
    xnu: List[float] = [0.0, 0.0]
    order, segs, gen_start = table_order(mech, min_class)
    for j in order:
        r = mech.reactions[j]
        kind = _table_kind(r)
        lng = _table_long(r)
        if lng:
            nrs = nps = dnu = 0
            rr = pp = [8 * pad * npt] * 3
            xo = len(xsl)
            xsl += [len(r.reactants), len(r.products)]
            xsl += [idx[sp] * npt for sp in r.reactants] + [idx[sp] * npt for sp in r.products]
            xnu += [float(r.coeff_sum), 0.0]
            xnu += [float(nu) for nu in r.reactants.values()] + [float(nu) for nu in r.products.values()]
        else:
            rsl, psl = _table_slots(r, idx)
            nrs, nps = len(rsl), len(psl)
            dnu = nps - nrs
            rr = [8 * k * npt for k in rsl] + [8 * pad * npt] * (3 - nrs)
            pp = [8 * k * npt for k in psl] + [8 * pad * npt] * (3 - nps)
            xo = 0
        x = kind | (int(r.reversible) << 3) | (nrs << 4) | (nps << 6) | (len(r.third_body) << 8) \
            | (TABLE_LONG if lng else 0) | (TABLE_NEG if negative_a(r) else 0) | ((dnu + 8) << 24)
        rows.append((8 * roffs[j], rr[0] | (rr[1] << 16), rr[2] | (pp[0] << 16), pp[1] | (pp[2] << 16)))
        if r.rtype == "plog":
            pts = sorted(set(c[0] for c in r.plog))
            y, o = len(plog), roffs[j]
            for pt in pts:
                n = len([c for c in r.plog if c[0] == pt])
                plog.append((o, n))
                o += 3 + 4 * n
            rows.append((x, y, len(pts), xo))
        else:
            nx = {"standard": 0, "third_body": 0, "lindemann": 3,
                  "troe": 3 + (5 if len(r.troe or ()) == 4 else 4), "sri": 3 + 5}[r.rtype]
            rows.append((x, len(effsp), roffs[j] + 3 + nx, xo))
            effsp.extend(idx[sp] for sp in r.third_body)
    return TableLayout(rows, effsp, order, segs, gen_start, plog, xsl, xnu)


def table_chunks(spans: List[Tuple[int, int]], rc: int, slots: int) -> List[Tuple[int, int, List[Tuple[int, int, int]]]]:
    """Split the table's reaction ranges (``spans``: the compiled classes then the generic
    range, contiguous in table order) into chunks of at most ``rc`` reactions, whose rates of
    progress fit the table kernel's LDS buffer at once.  Returns [(c0, c1, pieces)], pieces =
    [(a, b, span index)] covering [c0, c1).  Chunks are filled to rc (ceil(total / rc) of
    them): a range cut at a chunk end leaves a part-empty round there, which costs less than
    the chunk more that cuts after whole rounds would need (a gather and two barriers)."""
    total = spans[-1][1] if spans else 0
    if rc < min(slots, total):
        raise ValueError(f"table kernel: {rc} reactions per chunk is less than one round of {slots} slots")
    chunks: List[Tuple[int, int, List[Tuple[int, int, int]]]] = []
    c0, pieces = 0, []
    for si, (a, b) in enumerate(spans):
        while a < b:
            room = c0 + rc - a
            if room == 0:
                chunks.append((c0, a, pieces))
                c0, pieces, room = a, [], rc
            take = min(b - a, room)
            pieces.append((a, a + take, si))
            a += take
    if pieces:
        chunks.append((c0, spans[-1][1] if spans else 0, pieces))
    return chunks


def table_incidence(mech: Mechanism, order: List[int], chunks, npt: int):
    """Per chunk and species, the reactions of the chunk that change the species: CSR offsets
    goff[c (ns + 1) + k], and per entry the row of the rate of progress in the chunk's LDS
    buffer (in doubles: (j - c0) npt) and the net coefficient nu'' - nu' (entries in table
    order; species on both sides with equal coefficients get none)."""
    ns = mech.n_species
    idx = {s.name: i for i, s in enumerate(mech.species)}
    goff: List[int] = []
    grow: List[int] = []
    gnu: List[float] = []
    for (c0, c1, _) in chunks:
        lists: List[List[Tuple[int, float]]] = [[] for _ in range(ns)]
        for j in range(c0, c1):
            r = mech.reactions[order[j]]
            net: Dict[int, float] = {}
            for sp, nu in r.reactants.items():
                net[idx[sp]] = net.get(idx[sp], 0.0) - float(nu)
            for sp, nu in r.products.items():
                net[idx[sp]] = net.get(idx[sp], 0.0) + float(nu)
            for k, nu in net.items():
                if nu != 0.0:
                    lists[k].append(((j - c0) * npt, nu))
        for k in range(ns):
            goff.append(len(grow))
            for row, nu in lists[k]:
                grow.append(row)
                gnu.append(nu)
        goff.append(len(grow))
    return goff, grow, gnu


GATHER_FIRST = 1 << 30     # writer flag: the species' first chunk stores wdot, later chunks add


def _gather_order(lens: List[int]) -> List[Tuple[int, int]]:
    """Order of a lane group's stream words in a chunk: slot by slot (each slot's sum is
    stored as soon as it is complete, which frees its registers)."""
    return [(s_, u) for s_ in range(len(lens)) for u in range(lens[s_])]


def _gather_coef_word(nu: float):
    """The bits 18-31 of a gather entry for coefficient ``nu``: the high word of the double,
    when its low word and the high word's bits 0-17 are zero (the double is then exactly
    {0, word}); None otherwise."""
    import struct
    lo, hi = struct.unpack("<II", struct.pack("<d", float(nu)))
    return hi if lo == 0 and hi & 0x3ffff == 0 else None


def _gather_rows(lens_entries: Dict[int, list], W: int, G: int, thr: float):
    """Rows of one chunk's gather: a species with more than ``thr`` entries (and more than G)
    is split over the G lane groups of one wavefront (a combined row); the others are sorted
    by entry count and taken G at a time.  Rows sorted longest first; slot s of the chunk is
    rows [W s, W s + W), padded to the length of its first (longest) row."""
    rows, single = [], []
    for k, v in lens_entries.items():
        if len(v) > thr and len(v) > G:
            part = -(-len(v) // G)
            rows.append((part, True, [(k if g == 0 else -1, v[g * part:(g + 1) * part]) for g in range(G)]))
        else:
            single.append((k, v))
    single.sort(key=lambda kv: (-len(kv[1]), kv[0]))
    for i in range(0, len(single), G):
        grp = single[i:i + G] + [(-1, [])] * (G - len(single[i:i + G]))
        rows.append((max(len(v) for _, v in grp), False, grp))
    rows.sort(key=lambda r: -r[0])
    lens = [rows[i][0] for i in range(0, len(rows), W)]
    return rows, lens


def table_gather_streams(mech: Mechanism, chunks, goff: List[int], grow: List[int], gnu: List[float],
                         npt: int, W: int, rc: int) -> Dict:
    """The table kernel's gather, scheduled on the host.  A wavefront's 64 lanes form G = 64 /
    npt lane groups (one lane per point of the tile).  Chunk c has slots s = 0 .. S_c - 1 of
    fixed lengths L_s (steps), the same in every wavefront, so the kernel's code for a chunk is
    straight-line with static register indices.  In slot s of wavefront w each lane group sums
    the entries nu q of one species (the W x G species of a slot have similar entry counts), or
    all G groups of w sum parts of one long species and combine them with a butterfly over the
    groups (bit s of the wavefront's combine mask).  Shorter rows are padded with zero-row
    entries.  Every species is written by one group of the first chunk with entries for it
    (untouched species: a one-entry zero term in chunk 0), which stores its wdot; later chunks
    add to it.  Of the split thresholds 1 and 1/2 times the mean entries per group, the one
    with fewer steps is used.

    Entry (one 32-bit word): bits 0-17 the LDS byte offset of a rate-of-progress row of the
    chunk ((j - c0) npt 8; the zero row rc npt 8 pads), bits 18-31 the high word of the net
    coefficient nu'' - nu' as a double, whose other bits are zero for every integer up to 8
    and the halves (``_gather_coef_word``); a mechanism with another coefficient writes 0
    there and the parallel double array ``fnu`` holds every entry's coefficient.

    Returns per chunk c (``chunks``): lens (L_s), steps (sum, padded to a multiple of 4),
    str_base, spk_base, comb_base; lane group (w, g) reads words
    stream[str_base + (w G + g) steps + t], writers spk[spk_base + (w G + g) S_c + s]
    (species | GATHER_FIRST, or -1) and the wavefront's combine mask comb[comb_base + w]."""
    ns = mech.n_species
    G = 64 // npt
    frac = not all(_gather_coef_word(v) is not None for v in gnu)
    zero_row = rc * npt * 8
    if zero_row >= 1 << 18:
        raise ValueError("table kernel: rate-of-progress rows beyond the gather's 18-bit offsets")
    nch = len(chunks)
    touched = [any(goff[c * (ns + 1) + k + 1] > goff[c * (ns + 1) + k] for c in range(nch)) for k in range(ns)]
    first = {}
    for c in range(nch):
        for k in range(ns):
            if goff[c * (ns + 1) + k + 1] > goff[c * (ns + 1) + k] and k not in first:
                first[k] = c
    for k in range(ns):
        first.setdefault(k, 0)
    out, stream, spk, comb, fnu = [], [], [], [], []
    for c in range(nch):
        ents = {}
        for k in range(ns):
            a, b = goff[c * (ns + 1) + k], goff[c * (ns + 1) + k + 1]
            if b > a:
                ents[k] = [(grow[e] * 8, gnu[e]) for e in range(a, b)]
            elif c == 0 and not touched[k]:
                ents[k] = [(zero_row, 0.0)]
        mean = sum(len(v) for v in ents.values()) / (W * G)
        best = None
        for f in (1.0, 0.5):
            rows, lens = _gather_rows(ents, W, G, max(1.0, mean * f))
            if best is None or sum(lens) < sum(best[1]):
                best = (rows, lens)
        rows, lens = best
        nsl = len(lens)
        steps = -(-sum(lens) // 16) * 16
        rec = dict(lens=lens, steps=steps, str_base=len(stream), spk_base=len(spk), comb_base=len(comb),
                   adds=[any(r[2][g][0] >= 0 and first[r[2][g][0]] != c for r in rows[W * s_:W * s_ + W]
                             for g in range(G)) for s_ in range(nsl)])
        for w in range(W):
            mask = 0
            lane_rows = [rows[W * s_ + w] if W * s_ + w < len(rows) else (0, False, [(-1, [])] * G)
                         for s_ in range(nsl)]
            for s_, r in enumerate(lane_rows):
                if r[1]:
                    mask |= 1 << s_
            comb.append(mask)
            for g in range(G):
                words, fv = [], []
                for (s_, t) in _gather_order(lens):
                    v = lane_rows[s_][2][g][1]
                    off, nu = v[t] if t < len(v) else (zero_row, 0.0)
                    cw = _gather_coef_word(nu)
                    words.append(off | (cw if cw is not None else 0))
                    fv.append(float(nu))
                for s_, r in enumerate(lane_rows):
                    k = r[2][g][0]
                    spk.append(-1 if k < 0 else (k | (GATHER_FIRST if first[k] == c else 0)))
                words += [zero_row] * (steps - len(words))
                fv += [0.0] * (steps - len(fv))
                stream.extend(words)
                fnu.extend(fv)
        out.append(rec)
    if nch and any(m >= 1 << 31 for m in comb):
        raise ValueError("table kernel: more than 31 gather slots in a chunk")
    return dict(chunks=out, stream=stream, spk=spk, comb=comb, fnu=fnu if frac else None, G=G)


def _gather_code(ci: int, rec: Dict, frac: bool, bcast: int = 0) -> str:
    """Straight-line gather of chunk ``ci`` for the table kernel (see ``table_gather_streams``):
    writer words, the first block of the lane group's entry stream, the writers' old wdot
    (addresses clamped into the arrays, so every load is unconditional), then the slot sums
    with static register indices, each slot's store right after its last step.  ``bcast``
    (the lane group's size, 16 or 8): each lane loads one word of a block of ``bcast`` steps
    and step t takes lane t mod ``bcast``'s word by a cross-lane broadcast (DPP), instead of
    every lane of the group loading every word (``bcast`` times fewer bytes through the
    vector memory pipeline, which the redundant loads saturate)."""
    if bcast:
        return _gather_code_bcast(ci, rec, frac, bcast)
    lens, steps = rec["lens"], rec["steps"]
    nsl = len(lens)
    L = [f"  {{  // gather of chunk {ci}: {nsl} slot(s) of {', '.join(map(str, lens))} steps",
         f"    const int lg_ = cg_gw * CG_GG + cg_gg;",
         f"    const int* sk = cg_gspk + {rec['spk_base']} + lg_ * {nsl};",
         f"    const int cm = cg_gcomb[{rec['comb_base']} + cg_gw];"]
    L += [f"    const int k{s_} = sk[{s_}];" for s_ in range(nsl)]
    L.append(f"    const uint4* st = (const uint4*)(cg_gstr + {rec['str_base']} + lg_ * {steps});")
    if frac:
        L.append(f"    const double* sf = cg_gfnu + {rec['str_base']} + lg_ * {steps};")
    # the stream in blocks of 4 uint4 (16 steps), each loaded one block ahead of its use.
    # Memory counters complete in order: block 0 is issued before the writers' old values, so
    # that the first steps do not wait for those (scheduling barriers keep the order, and keep
    # the scheduler from hoisting every load to the top, which spills)
    order = _gather_order(lens)
    nb = -(-len(order) // 4)
    blk = 4
    def load_block(b0):
        return [f"    const uint4 v{b} = st[{b}];" for b in range(b0, min(nb, b0 + blk))]
    L += load_block(0)
    L.append("    __builtin_amdgcn_sched_barrier(0);")
    # one row pointer per slot for the old value and the store (row 0 for no writer: the load
    # stays in bounds and the store is skipped); k ld as a 32 x 32 -> 64 product (ld < 2^32)
    L += [f"    double* const w{s_} = cg_gwc + (uint64_t)(unsigned)(k{s_} >= 0 ? (k{s_} & 0xffffff) : 0) * cg_gld;"
          for s_ in range(nsl)]
    # the old values, for the slots with a writer that adds (none in a chunk where every
    # writer stores: the first chunk)
    adds = rec["adds"]
    L += [f"    const double p{s_} = *w{s_};" if adds[s_] else f"    const double p{s_} = 0.0;" for s_ in range(nsl)]
    L += [f"    double a{s_} = 0.0;" for s_ in range(nsl)]
    last = {}
    for t, (s_, u) in enumerate(order):
        last[s_] = t
    for b0 in range(0, nb, blk):
        L.append("    __builtin_amdgcn_sched_barrier(0);")
        L += load_block(b0 + blk)
        for t in range(4 * b0, min(len(order), 4 * (b0 + blk))):
            s_ = order[t][0]
            wd = f"v{t // 4}.{'xyzw'[t % 4]}"
            nu = f"sf[{t}]" if frac else f"CG_GNU({wd})"
            L.append(f"    a{s_} = __builtin_fma({nu}, CG_GQ({wd}), a{s_});")
            if last[s_] == t:
                L.append(f"    if (cm & {1 << s_}) {{")
                L.append(f"      for (int o = CG_NPT; o < 64; o <<= 1) a{s_} += __shfl_xor(a{s_}, o);")
                L.append("    }")
                L.append(f"    if (k{s_} >= 0 && valid) *w{s_} = (k{s_} & (1 << 30)) ? a{s_} : p{s_} + a{s_};")
    L.append("  }")
    return "\n".join(L) + "\n"


def _gather_code_bcast(ci: int, rec: Dict, frac: bool, B: int) -> str:
    """``_gather_code`` with the stream words broadcast across the lane group (B lanes)."""
    lens, steps = rec["lens"], rec["steps"]
    nsl = len(lens)
    L = [f"  {{  // gather of chunk {ci}: {nsl} slot(s) of {', '.join(map(str, lens))} steps",
         f"    const int lg_ = cg_gw * CG_GG + cg_gg;",
         f"    const int* sk = cg_gspk + {rec['spk_base']} + lg_ * {nsl};",
         f"    const int cm = cg_gcomb[{rec['comb_base']} + cg_gw];"]
    L += [f"    const int k{s_} = sk[{s_}];" for s_ in range(nsl)]
    # lane pt of the group holds word B b + pt of block b (steps is a multiple of 16 >= B nb)
    L.append(f"    const unsigned* st = cg_gstr + {rec['str_base']} + lg_ * {steps} + pt;")
    if frac:
        L.append(f"    const double* sf = cg_gfnu + {rec['str_base']} + lg_ * {steps};")
    order = _gather_order(lens)
    nb = -(-len(order) // B)
    per = 16 // B                     # blocks per 16 steps, loaded one such group ahead
    def load_block(b0):
        return [f"    const unsigned u{b} = st[{B * b}];" for b in range(b0, min(nb, b0 + per))]
    L += load_block(0)
    L.append("    __builtin_amdgcn_sched_barrier(0);")
    L += [f"    double* const w{s_} = cg_gwc + (uint64_t)(unsigned)(k{s_} >= 0 ? (k{s_} & 0xffffff) : 0) * cg_gld;"
          for s_ in range(nsl)]
    adds = rec["adds"]
    L += [f"    const double p{s_} = *w{s_};" if adds[s_] else f"    const double p{s_} = 0.0;" for s_ in range(nsl)]
    L += [f"    double a{s_} = 0.0;" for s_ in range(nsl)]
    last = {}
    for t, (s_, u) in enumerate(order):
        last[s_] = t
    for b0 in range(0, nb, per):
        L.append("    __builtin_amdgcn_sched_barrier(0);")
        L += load_block(b0 + per)
        for t in range(B * b0, min(len(order), B * (b0 + per))):
            s_ = order[t][0]
            L.append(f"    const unsigned x{t} = CG_BCW(u{t // B}, {t % B});")
            nu = f"sf[{t}]" if frac else f"CG_GNU(x{t})"
            L.append(f"    a{s_} = __builtin_fma({nu}, CG_GQ(x{t}), a{s_});")
            if last[s_] == t:
                L.append(f"    if (cm & {1 << s_}) {{")
                L.append(f"      for (int o = CG_NPT; o < 64; o <<= 1) a{s_} += __shfl_xor(a{s_}, o);")
                L.append("    }")
                L.append(f"    if (k{s_} >= 0 && valid) *w{s_} = (k{s_} & (1 << 30)) ? a{s_} : p{s_} + a{s_};")
    L.append("  }")
    return "\n".join(L) + "\n"


def table_gather_plan(mech: Mechanism, npt: int, W: int) -> Dict:
    """How the table kernel of ``npt``-point tiles and ``W`` wavefronts runs its reactions:
    the table layout (a class gets a compiled loop when it fills at least half the reaction
    slots of a tile; JetSurf-sized: 2.6% faster than a full round of slots, every class
    compiled 5% slower), the chunk size rc (the rates of progress of rc reactions fill the LDS
    left after EG and h [ns + 1][npt] and the per-wavefront partial sums [W][npt]: every table
    shape runs one workgroup per CU, W wavefronts = 4 SIMDs x its waves per SIMD), the chunks
    and their species incidence lists, and the LDS bytes of a tile."""
    ns, nr = mech.n_species, mech.n_reactions
    slots = W * 64 // npt
    fixed = 2 * (ns + 1) * npt + W * npt + npt       # ... and the gather's zero row
    rc = min(nr, (LDS_PER_CU // 8 - fixed) // npt)
    if rc < min(nr, slots):
        raise ValueError(f"table kernel: {npt}-point tiles of {ns} species leave LDS for {rc} rates of progress")
    lay = table_layout(mech, npt, min_class=max(1, slots // 2))
    spans = [(a, b) for (a, b, _) in lay.segs] + ([(lay.gen_start, nr)] if lay.gen_start < nr else [])
    chunks = table_chunks(spans, rc, slots) if nr else []
    goff, grow, gnu = table_incidence(mech, lay.order, chunks, npt)
    streams = table_gather_streams(mech, chunks, goff, grow, gnu, npt, W, rc) if chunks else None
    return dict(layout=lay, rc=rc, slots=slots, spans=spans, chunks=chunks, goff=goff, grow=grow, gnu=gnu,
                streams=streams, lds=8 * (fixed + rc * npt))


def table_records(mech: Mechanism, npt: int = 16, min_class: int = 0):
    """(reaction table rows [2 nr][4], efficiency species, order, class segments, start of the
    generic segment) of ``table_layout``."""
    t = table_layout(mech, npt, min_class)
    return t.rows, t.effsp, t.order, t.segs, t.gen_start


def _emit_table(mech: Mechanism, variant: Variant):
    import jinja2
    coef, off_smh, roffs = build_coef(mech)
    npt, W = variant.table, variant.split
    if 64 % npt:
        raise ValueError("table kernel: points per tile must divide 64")
    slots = W * 64 // npt
    nr = mech.n_reactions
    plan = table_gather_plan(mech, npt, W)
    lay, rc, lds = plan["layout"], plan["rc"], plan["lds"]
    rows, effsp, segs, gen_start = lay.rows, lay.effsp, lay.segs, lay.gen_start

    def most(a: int, b: int):
        """(most collision efficiencies, most PLOG pressure points) of table positions [a, b)."""
        rs = [mech.reactions[lay.order[t]] for t in range(a, b)]
        ne = max([len(r.third_body) for r in rs if r.rtype in ("third_body", "lindemann", "troe", "sri")],
                 default=0)
        npt_ = max([len(set(c[0] for c in r.plog)) for r in rs if r.rtype == "plog"], default=1)
        return ne, npt_

    classes = []
    for (a, b, (fam, rev, nrs, nps, dnu, _neg)) in segs:
        pf = None
        if rev:
            pf = {0: None, 1: "pinv", -1: "pfac"}.get(dnu, f"cg_exp2({dlit(-float(dnu))} * lg2pfac)")
        nemax, npmax = most(a, b)
        classes.append(dict(start=a, end=b, fam=fam, rev=rev, nrs=nrs, nps=nps, dnu=dnu, pf=pf,
                            nemax=nemax, npmax=npmax,
                            name=f"family {fam}, {'reversible' if rev else 'irreversible'}, "
                                 f"{nrs} -> {nps} slots, dnu {dnu}"))
    chunk_list, gs = plan["chunks"], plan["streams"]
    gfrac = bool(gs and gs["fnu"])
    bc = npt if (variant.bcast and npt in (8, 16)) else 0
    chunks = [dict(c0=c0, c1=c1, gather=_gather_code(ci, gs["chunks"][ci], gfrac, bc),
                   pieces=[dict(a=a, b=b, cl=classes[si] if si < len(classes) else None)
                                         for (a, b, si) in pieces])
              for ci, (c0, c1, pieces) in enumerate(chunk_list)]
    env = jinja2.Environment(loader=jinja2.FileSystemLoader(TEMPLATE_DIR), keep_trailing_newline=True)
    tpl = env.get_template("getrates_table_hip.j2")
    h = structure_hash(mech)
    rtab = ",\n".join(f"  {{{a}, {b}, {c}, {d}}}" for (a, b, c, d) in rows)
    src = tpl.render(
        name=mech.name, ns=mech.n_species, nr=mech.n_reactions, ncoef=len(coef), npt=npt, W=W,
        threads=variant.threads, ppb=variant.ppb, slots=slots, waves_per_eu=variant.waves_per_eu,
        hash_hex=f"0x{h:016x}", off_invmw=0, off_smh=off_smh, R0=dlit(R0),
        lg2_patm_r0=dlit(math.log2(PATM / R0)), patm_r0=dlit(PATM / R0), r0_patm=dlit(R0 / PATM),
        log10_2=dlit(LOG10_2), log2_10=dlit(LOG2_10), exp_impl=variant.exp_impl, poly=variant.poly,
        fdiv=variant.fdiv, bcast=bc, rtab=rtab, effsp=", ".join(str(k) for k in effsp + [mech.n_species]), n_effsp=len(effsp) + 1,
        classes=classes, gen_start=gen_start, chunks=chunks, rc=rc,
        gG=gs["G"] if gs else 1,
        gspk=", ".join(str(v) for v in gs["spk"]) if gs and gs["spk"] else "-1",
        n_gspk=max(1, len(gs["spk"])) if gs else 1,
        gcomb=", ".join(str(v) for v in gs["comb"]) if gs and gs["comb"] else "0",
        n_gcomb=max(1, len(gs["comb"])) if gs else 1,
        gstr=", ".join(f"{v}u" for v in gs["stream"]) if gs and gs["stream"] else "0u",
        n_gstr=max(4, len(gs["stream"])) if gs else 4,
        gfrac=gfrac, gfnu=", ".join(dlit(v) for v in gs["fnu"]) if gfrac else "0.0",
        has_plog=bool(lay.plog), plog=", ".join(f"{{{o}, {n}}}" for o, n in lay.plog), n_plog=len(lay.plog),
        has_long=len(lay.xsl) > 2, xsl=", ".join(str(v) for v in lay.xsl), n_xsl=len(lay.xsl),
        has_neg=any(negative_a(r) for r in mech.reactions),
        xnu=", ".join(dlit(v) for v in lay.xnu),
        tmid=", ".join(dlit(s.thermo.t_mid) for s in mech.species),
    )
    meta = dict(ns=mech.n_species, nr=mech.n_reactions, ncoef=len(coef), block=variant.threads,
                ppb=variant.ppb, waves_per_eu=variant.waves_per_eu, structure_hash=h, slots=slots,
                compiled_classes=len(classes), generic_reactions=mech.n_reactions - gen_start,
                lds_bytes=lds, chunks=len(chunks), chunk_reactions=rc)
    return src, coef, meta


HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def variant_flags(variant: "Variant") -> List[str]:
    """Extra compiler flags of a variant.  Straight-line kernels turn off machine-level
    loop-invariant code motion: in the persistent ragged-batch kernel it hoists the
    materialised FP64 constants of the whole tile body out of the tile loop, which pushes
    the kernel to the SGPR limit (40 spills, twice the LGKM waits of the loop-free body);
    the kernels without a loop are unchanged by it.  The table kernel's reaction loops
    keep it."""
    return [] if variant.table else ["-mllvm", "-disable-machine-licm"]


def compile_code_object(src_path: str, out_path: str, extra: List[str] = ()) -> None:
    """hipcc --genco: a gfx950 code object loadable with hipModuleLoad."""
    cmd = [HIPCC, "--genco", *HIP_FLAGS, *extra, "-o", out_path + ".tmp", src_path]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({' '.join(cmd)}):\n{res.stderr[-4000:]}")
    os.replace(out_path + ".tmp", out_path)
