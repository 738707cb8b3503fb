"""C/OpenMP restatement of the reference's vector Fortran getrates -- TEST INFRASTRUCTURE.

Emits, per mechanism, C code with the structure of the reference's ``getrates_i``
(chemgen/templates/ftn_header.j2:1-36, ytoc_ftn_vec.j2, exp_g_ftn_vec.j2,
reactions_ftn_vec/*.j2): array statements over a strip of up to 32 points (the mini-app's
MAXVL, chemgen/templates/chemistry_mini_app.j2:77-82), reactions unrolled in mechanism
order, coefficients printed with ``%.15e`` as the templates print them, and the formulas
of chemgen/templates/reactions_ftn_vec/{standard,third_body,troe}.j2.  An
``omp parallel for`` over strips spreads it over the host cores.

It is the checker the GPU kernel is compared with (besides the NumPy oracle) and the
CPU baseline that is timed beside it; it is never called by the product path.
Lindemann/SRI falloff extend the Troe template with the textbook broadening factors,
exactly as the NumPy oracle does.
"""
from __future__ import annotations

import math
from typing import List

R0 = 8.314510e+07
RC = 1.9872155832
PATM = 1013250.0


def e15(x: float) -> str:
    return "%.15e" % x


def _pw(var: str, nu: float) -> str:
    return var if nu == 1.0 else f"pow({var}, {e15(nu)})"


def _kb(r, idx) -> str:
    s = "kf[i]"
    for sp, nu in r.reactants.items():
        s += f" * {_pw(f'EG[{idx[sp]}][i]', nu)}"
    for sp, nu in r.products.items():
        s += f" / {_pw(f'EG[{idx[sp]}][i]', nu)}"
    if r.coeff_sum != 0.0:
        s += f" / pow(pfac[i], {e15(r.coeff_sum)})"
    return s


def _arrh(A, b, Ea) -> str:
    s = e15(A)
    if b != 0.0:
        s += f" * pow(T[i], {repr(float(b))})"
    return s + f" * exp({e15(-1.0 * Ea)} / (RC * T[i]))"


def _m(r, idx) -> str:
    s = "ctot[i]"
    for sp, e in r.third_body.items():
        s += (" + " if e > 0.0 else " - ") + f"{e15(abs(e))} * C[{idx[sp]}][i]"
    return s


def _rr(r, idx) -> str:
    f = "kf[i]" + "".join(f" * {_pw(f'C[{idx[s]}][i]', n)}" for s, n in r.reactants.items())
    b = "kb[i]" + "".join(f" * {_pw(f'C[{idx[s]}][i]', n)}" for s, n in r.products.items())
    return f"{f} - {b}"


def emit_c(mech, vl: int = 32) -> str:
    idx = {s.name: i for i, s in enumerate(mech.species)}
    ns = mech.n_species
    L: List[str] = []
    a = L.append
    a("#include <math.h>")
    a("#include <stdint.h>")
    a("#include <omp.h>")
    a(f"#define NS {ns}")
    a(f"#define VL {vl}")
    a(f"static const double R0 = {repr(R0)};")
    a(f"static const double RC = {repr(RC)};")
    a(f"static const double PATM = {repr(PATM)};")
    a("static void strip(int n, const double* Pg, const double* Tg, const double* Y, double* W,")
    a("                  int64_t ld, int64_t i0) {")
    a("  double C[NS][VL], EG[NS][VL], WL[NS][VL];")
    a("  double P[VL], T[VL], ctot[VL], pfac[VL], kf[VL], kb[VL], rr[VL], M[VL];")
    a("  double k0[VL], kinf[VL], Pr[VL], Fcent[VL], C1[VL], N1[VL], f1[VL], F[VL], X[VL];")
    a("  int i, k;")
    a("  for (i = 0; i < n; ++i) { P[i] = Pg[i0 + i]; T[i] = Tg[i0 + i]; }")
    a("  for (k = 0; k < NS; ++k) for (i = 0; i < n; ++i) WL[k][i] = 0.0;")
    a("  for (i = 0; i < n; ++i) pfac[i] = PATM / (R0 * T[i]);")
    # ytoc (ytoc_ftn_vec.j2)
    for k, s in enumerate(mech.species):
        a(f"  for (i = 0; i < n; ++i) C[{k}][i] = Y[(int64_t){k} * ld + i0 + i] / {e15(s.mw)};")
    a("  for (i = 0; i < n; ++i) ctot[i] = 0.0;")
    a("  for (k = 0; k < NS; ++k) for (i = 0; i < n; ++i) ctot[i] = ctot[i] + C[k][i];")
    a("  for (k = 0; k < NS; ++k) for (i = 0; i < n; ++i) C[k][i] = C[k][i] * P[i] / (R0 * ctot[i] * T[i]);")
    if any(r.rtype != "standard" for r in mech.reactions):
        a("  for (i = 0; i < n; ++i) ctot[i] = 0.0;")
        a("  for (k = 0; k < NS; ++k) for (i = 0; i < n; ++i) ctot[i] = ctot[i] + C[k][i];")
    # EG (exp_g_ftn_vec.j2)
    for k, s in enumerate(mech.species):
        th = s.thermo

        def poly(c):
            sg = lambda v: "+" if v >= 0 else "-"
            return (f"{e15(c[6] - c[0])} {sg(c[0])} {e15(abs(c[0]))} * log(T[i])"
                    f" {sg(c[1])} {e15(abs(c[1] / 2.0))} * T[i]"
                    f" {sg(c[2])} {e15(abs(c[2] / 6.0))} * (T[i] * T[i])"
                    f" {sg(c[3])} {e15(abs(c[3] / 12.0))} * pow(T[i], 3.0)"
                    f" {sg(c[4])} {e15(abs(c[4] / 20.0))} * pow(T[i], 4.0)"
                    f" {'-' if c[5] >= 0 else '+'} {e15(abs(c[5]))} / T[i]")
        a(f"  for (i = 0; i < n; ++i) {{ double smh = (T[i] < {repr(float(th.t_mid))}) ? "
          f"({poly(th.low)}) : ({poly(th.high)}); EG[{k}][i] = exp(smh); }}")
    for j, r in enumerate(mech.reactions):
        a(f"  /* R{j} {r.rtype} */")
        a("  for (i = 0; i < n; ++i) {")
        if r.rtype in ("standard", "third_body"):
            a(f"    kf[i] = {_arrh(*r.arrh)};")
            a(f"    kb[i] = {_kb(r, idx) if r.reversible else '0.0'};")
            if r.rtype == "third_body":
                a(f"    M[i] = {_m(r, idx)};")
                a(f"    rr[i] = M[i] * ({_rr(r, idx)});")
            else:
                a(f"    rr[i] = {_rr(r, idx)};")
        elif r.rtype in ("troe", "lindemann", "sri"):
            a(f"    M[i] = {_m(r, idx)};")
            a(f"    k0[i] = {_arrh(*r.low)};")
            a(f"    kinf[i] = {_arrh(*r.arrh)};")
            a("    Pr[i] = (k0[i] * M[i]) / kinf[i];")
            if r.rtype == "troe":
                aa, T3, T1 = r.troe[0], r.troe[1], r.troe[2]
                t3 = f"{e15(1.0 - aa)} * exp(-T[i] / {e15(T3)})" if T3 != 0.0 else "0.0"
                t1 = f"{e15(aa)} * exp(-T[i] / {e15(T1)})" if T1 != 0.0 else "0.0"
                fc = f"{t3} + {t1}"
                if len(r.troe) == 4:
                    fc += f" + exp({e15(-r.troe[3])} / T[i])"
                a(f"    Fcent[i] = {fc};")
                a("    C1[i] = -0.4 - 0.67 * log10(Fcent[i]);")
                a("    N1[i] = 0.75 - 1.27 * log10(Fcent[i]);")
                a("    f1[i] = (log10(Pr[i]) + C1[i]) / (N1[i] - 0.14 * (log10(Pr[i]) + C1[i]));")
                a("    F[i] = pow(10.0, log10(Fcent[i]) / (1.0 + f1[i] * f1[i]));")
            elif r.rtype == "sri":
                sp = tuple(r.sri) + ((1.0, 0.0) if len(r.sri) == 3 else ())
                a("    X[i] = 1.0 / (1.0 + log10(Pr[i]) * log10(Pr[i]));")
                f = (f"{e15(sp[3])} * pow({e15(sp[0])} * exp({e15(-sp[1])} / T[i]) + "
                     f"exp(-T[i] / {e15(sp[2])}), X[i])")
                if sp[4] != 0.0:
                    f += f" * pow(T[i], {repr(float(sp[4]))})"
                a(f"    F[i] = {f};")
            else:
                a("    F[i] = 1.0;")
            a("    kf[i] = kinf[i] * (Pr[i] / (1.0 + Pr[i])) * F[i];")
            a(f"    kb[i] = {_kb(r, idx) if r.reversible else '0.0'};")
            a(f"    rr[i] = {_rr(r, idx)};")
        else:
            raise NotImplementedError(f"C oracle: reaction type {r.rtype}")
        for sp, nu in r.reactants.items():
            a(f"    WL[{idx[sp]}][i] = WL[{idx[sp]}][i] - rr[i]{'' if nu == 1.0 else ' * ' + e15(nu)};")
        for sp, nu in r.products.items():
            a(f"    WL[{idx[sp]}][i] = WL[{idx[sp]}][i] + rr[i]{'' if nu == 1.0 else ' * ' + e15(nu)};")
        a("  }")
    a("  for (k = 0; k < NS; ++k) for (i = 0; i < n; ++i) W[(int64_t)k * ld + i0 + i] = WL[k][i];")
    a("}")
    a("")
    a("int oracle_n_species(void) { return NS; }")
    a("int oracle_threads(void) { return omp_get_max_threads(); }")
    a("void oracle_set_threads(int n) { omp_set_num_threads(n); }")
    a("void oracle_getrates(int64_t ng, int64_t ld, const double* P, const double* T, const double* Y,")
    a("                     double* W) {")
    a("  #pragma omp parallel for schedule(static)")
    a("  for (int64_t s = 0; s < ng; s += VL) {")
    a("    int n = (int)((ng - s) < VL ? (ng - s) : VL);")
    a("    strip(n, P, T, Y, W, ld, s);")
    a("  }")
    a("}")
    return "\n".join(L) + "\n"
