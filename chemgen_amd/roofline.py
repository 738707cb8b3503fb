"""Algorithmic work per grid point and the MI355X peaks it is priced against.

Work model (flop-equivalents per point, transcendentals weighted by W_T):
  setup               10 arithmetic + 1 transcendental (ln T)
  Y -> C              3 Ns + 2
  NASA-7 smh          13 Ns arithmetic + Ns transcendentals (the reference GPU code takes
                      log C per species, chemgen/templates/reactions_hip/device_functions_h.j2:701-719)
  per reaction, n = species entries on both sides:
                      4 + 2n (ln kf and sum nu ln C), + 2n + 2 if reversible (delta smh and
                      dnu ln pfac), + 1 (net rate), + 2n (scatter);
                      transcendentals: 2 if reversible else 1
  third-body/falloff  + 2 n_eff + 1 (the collision efficiency sum)
  Troe                + 25 arithmetic, + 5 transcendentals
  SRI                 + 20 arithmetic, + 5 transcendentals
  Lindemann           + 8 arithmetic, + 1 transcendental
Bytes per point: 8 (2 + 2 Ns) (T, P, Y in; wdot out).

W_T = 20 flop-equivalents per FP64 exp/log is the fixed weight this repository prices
transcendentals at; the FP64 vector peak of one MI355X is 78.6 TFLOP/s and its HBM3E
peak 8 TB/s.
"""
from __future__ import annotations

W_T = 20.0
PEAK_FP64_VECTOR = 78.6e12   # FLOP/s
PEAK_HBM = 8.0e12            # B/s
PEAK_BF16_DENSE = 2.5e15     # FLOP/s


def work_per_point(mech) -> dict:
    ns = mech.n_species
    arith = 10.0 + 3 * ns + 2 + 13 * ns
    transc = 1.0 + ns
    for r in mech.reactions:
        n = len(r.reactants) + len(r.products)
        arith += 4 + 2 * n + 1 + 2 * n
        if r.reversible:
            arith += 2 * n + 2
            transc += 2
        else:
            transc += 1
        if r.rtype != "standard":
            arith += 2 * len(r.third_body) + 1
        if r.rtype == "troe":
            arith += 25
            transc += 5
        elif r.rtype == "sri":
            arith += 20
            transc += 5
        elif r.rtype == "lindemann":
            arith += 8
            transc += 1
    flops = arith + W_T * transc
    bytes_ = 8.0 * (2 + 2 * ns)
    return {"flops": flops, "arith": arith, "transc": transc, "bytes": bytes_,
            "intensity": flops / bytes_}


def binding_roof(mech) -> str:
    w = work_per_point(mech)
    return "fp64" if w["intensity"] > PEAK_FP64_VECTOR / PEAK_HBM else "hbm"
