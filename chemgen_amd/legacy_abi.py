"""The reference's C ABI over a generated kernel: a per-mechanism compatibility library.

The reference's Fortran module ``chemgen_m`` binds three C entry points of its generated HIP
code (chemgen/templates/chemgen_mod.j2:26-104, called from :108-166):

* ``copyConstantsToDevice(...)`` -- the mechanism's coefficient arrays (host pointers), the
  argument list depending on which reaction types exist (chemgen/constants.py:119-240);
* ``getrates_hip_wrapper_vN(P, T, ng, Y, wdot, logEG, logC)``, N = 1..4 (hip_v1_h.j2:7-13,
  hip_v2_h.j2:7-13, hip_v3_h.j2:7-13, hip_v4_h.j2:26-32), device pointers, ``Y(ng, Ns)``;
* ``allocate_intermediate_DeviceMemory(ng)`` for level 4 (constants.py:435-470).

:func:`emit_legacy_source` writes C++ that exports exactly these symbols for one mechanism:
``copyConstantsToDevice`` converts the reference's arrays (ln A, (beta, -Ea/Rc) pairs, the
Troe centre coefficients, the dense 1 + (eff - 1) table, the NASA-7 transform, molecular
weights) into this build's gpu_coef blob and loads the generated code object for the same
mechanism; every ``getrates_hip_wrapper_vN`` runs ``cg_getrates`` on the null stream and
synchronises, as the reference's wrappers do (hip_v1_cpp.j2:363,371); the scratch pointers
are ignored (no scratch is needed) and any ``ng`` works.  Errors are printed to stderr and the
call returns, the reference's behaviour (constants.py:8-24).  The conversion itself is also
exported as ``cg_legacy_convert`` (host only), so it can be checked without a GPU.

Only the reaction types the reference's HIP path holds are accepted: standard, falloff-Troe
and three-body (chemgen/parser.py:44-49; its PLOG HIP code is disabled,
device_functions_h.j2:554-586).
"""
from __future__ import annotations

import math
import os
import subprocess
from typing import List

from .codegen import HIP_FLAGS, HIPCC, build_coef, structure_hash
from .mechanism import Mechanism

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib")
INCLUDE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")

_TYPES = ("standard", "troe", "third_body")


def _params(present) -> List[str]:
    p = []
    if "standard" in present:
        p += ["const double* A_h", "const double* B_h", "const int* sk_map_h", "const double* sk_coef_h",
              "const int* map_r_h", "const double* coef_r_h", "const int* map_p_h", "const double* coef_p_h"]
    if "troe" in present:
        p += ["const double* A_0_troe_h", "const double* B_0_troe_h", "const double* A_inf_troe_h",
              "const double* B_inf_troe_h", "const int* sk_map_troe_h", "const double* sk_coef_troe_h",
              "const int* map_r_troe_h", "const double* coef_r_troe_h", "const int* map_p_troe_h",
              "const double* coef_p_troe_h", "const double* eff_fac_troe_h", "const double* fcent_coef_troe_h"]
    if "third_body" in present:
        p += ["const double* A_third_h", "const double* B_third_h", "const int* sk_map_third_h",
              "const double* sk_coef_third_h", "const int* map_r_third_h", "const double* coef_r_third_h",
              "const int* map_p_third_h", "const double* coef_p_third_h", "const double* eff_fac_third_h"]
    p += ["const double* mw_h", "const double* smh_coef_h", "const double* T_mid_h"]
    return p


def _params4(present) -> List[str]:
    """copyConstantsToDevice of the rocBLAS level (chemgen/constants.py:327-420): dense
    coefficient matrices instead of species maps."""
    p = []
    if "standard" in present:
        p += ["const double* A_h", "const double* B_h", "const double* sk_coef_h", "const double* coef_r_h",
              "const double* coef_p_h", "const double* wdot_coef_h"]
    if "troe" in present:
        p += ["const double* A_0_troe_h", "const double* B_0_troe_h", "const double* A_inf_troe_h",
              "const double* B_inf_troe_h", "const double* sk_coef_troe_h", "const double* coef_r_troe_h",
              "const double* coef_p_troe_h", "const double* eff_fac_troe_h", "const double* fcent_coef_troe_h",
              "const double* wdot_coef_troe_h"]
    if "third_body" in present:
        p += ["const double* A_third_h", "const double* B_third_h", "const double* sk_coef_third_h",
              "const double* coef_r_third_h", "const double* coef_p_third_h", "const double* eff_fac_third_h",
              "const double* wdot_coef_third_h"]
    p += ["const double* mw_h", "const double* smh_coef_h", "const double* T_mid_h"]
    return p


def emit_legacy_source(mech: Mechanism, code_object: str, level: int = 1) -> str:
    """C++ of the compatibility library for the coefficient layout of parallel level
    ``level``: levels 1-3 pass species maps with MAX_SP entries per side (rounded up to a power
    of two at level 3), level 4 the dense matrices of the rocBLAS path with another argument
    list.  Dummy reactions that pad a type to a multiple of the reactions per block follow its
    real reactions in every array, so they need no handling here."""
    from .legacy_coef import max_species
    if level not in (1, 2, 3, 4):
        raise ValueError("parallel level must be 1..4")
    for r in mech.reactions:
        if r.rtype not in _TYPES:
            raise ValueError(f"the reference's HIP ABI has no {r.rtype} reactions (load with compat='reference')")
    coef, off_smh, roffs = build_coef(mech)
    idx = {s.name: i for i, s in enumerate(mech.species)}
    present = [t for t in _TYPES if any(r.rtype == t for r in mech.reactions)]
    counters = {t: 0 for t in _TYPES}
    max_sp = max_species(mech, level)
    ns_ = mech.n_species
    rows, effsp, checks = [], [], []
    suffix = {0: "", 1: "_troe", 2: "_third"}
    chk_lines = []
    for j, r in enumerate(mech.reactions):
        t = _TYPES.index(r.rtype)
        i = counters[r.rtype]
        counters[r.rtype] += 1
        t2 = int(r.rtype == "troe" and len(r.troe) == 4)
        rows.append(f"  {{{t}, {i}, {roffs[j]}, {t2}, {len(r.third_body)}, {len(effsp)}}}")
        effsp += [idx[s] for s in r.third_body]
        sfx = suffix[t]
        if level < 4:
            for q, s_ in enumerate(r.reactants):
                chk_lines.append(f"  if (map_r{sfx}_h[{i * max_sp + q}] != {idx[s_]}) return 2;")
            for q, s_ in enumerate(r.products):
                chk_lines.append(f"  if (map_p{sfx}_h[{i * max_sp + q}] != {idx[s_]}) return 2;")
        else:
            # the dense rows of this reaction: its coefficients where its species are, and row
            # sums that leave no room for another species
            for side, arr in ((r.reactants, "coef_r"), (r.products, "coef_p")):
                for s_, nu in side.items():
                    chk_lines.append(f"  if ({arr}{sfx}_h[{i * ns_ + idx[s_]}] != {float(nu)!r}) return 2;")
                chk_lines.append(f"  if (rowsum({arr}{sfx}_h + {i * ns_}) != {float(sum(side.values()))!r}) return 2;")
    params = _params(present) if level < 4 else _params4(present)
    args = ", ".join(p.split()[-1] for p in params)
    h = structure_hash(mech)
    ns, nr = mech.n_species, mech.n_reactions
    tmid = ", ".join(repr(float(s.thermo.t_mid)) for s in mech.species)
    conv_std = '''
      const double* A = r.type == 0 ? A_h : A_third_h;
      const double* B = r.type == 0 ? B_h : B_third_h;
      cf[r.off] = A[r.idx] * LOG2E;
      cf[r.off + 1] = B[2 * r.idx];
      cf[r.off + 2] = -B[2 * r.idx + 1] * LOG2E;
      for (int q = 0; q < r.ne; ++q)
        cf[r.off + 3 + q] = (r.type == 0 ? 0.0 : eff_fac_third_h[(long)r.idx * NS + kEffSp[r.eoff + q]]) - 1.0;'''
    conv_troe = '''
      cf[r.off] = A_inf_troe_h[r.idx] * LOG2E;
      cf[r.off + 1] = B_inf_troe_h[2 * r.idx];
      cf[r.off + 2] = -B_inf_troe_h[2 * r.idx + 1] * LOG2E;
      cf[r.off + 3] = A_0_troe_h[r.idx] * LOG2E;
      cf[r.off + 4] = B_0_troe_h[2 * r.idx];
      cf[r.off + 5] = -B_0_troe_h[2 * r.idx + 1] * LOG2E;
      const double* fc = fcent_coef_troe_h + 6 * r.idx;  // 1-a, T3, a, T1, 1|0, T2
      if ((fc[4] != 0.0) != (r.t2 != 0)) return 3;
      cf[r.off + 6] = fc[0];
      cf[r.off + 7] = fc[1] != 0.0 ? LOG2E / fc[1] : INFINITY;
      cf[r.off + 8] = fc[2];
      cf[r.off + 9] = fc[3] != 0.0 ? LOG2E / fc[3] : INFINITY;
      if (r.t2) cf[r.off + 10] = LOG2E * fc[5];
      for (int q = 0; q < r.ne; ++q)
        cf[r.off + 10 + r.t2 + q] = eff_fac_troe_h[(long)r.idx * NS + kEffSp[r.eoff + q]] - 1.0;'''
    body_std = conv_std if ("standard" in present or "third_body" in present) else ""
    if "standard" not in present:
        body_std = body_std.replace("r.type == 0 ? A_h : A_third_h", "A_third_h").replace(
            "r.type == 0 ? B_h : B_third_h", "B_third_h")
    if "third_body" not in present:
        body_std = body_std.replace("r.type == 0 ? A_h : A_third_h", "A_h").replace(
            "r.type == 0 ? B_h : B_third_h", "B_h").replace(
            "(r.type == 0 ? 0.0 : eff_fac_third_h[(long)r.idx * NS + kEffSp[r.eoff + q]])", "0.0")
    src = f'''// Generated by chemgen_amd: the reference's C ABI (copyConstantsToDevice,
// getrates_hip_wrapper_v1..v4, allocate_intermediate_DeviceMemory) for mechanism
// '{mech.name}' ({ns} species, {nr} reactions) over this build's generated getrates kernel,
// for the coefficient arrays of parallel level {level}{" (species maps with MAX_SP = %d)" % max_sp if level < 4 else " (dense rocBLAS matrices)"}.
#include "chemgen_amd.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {{
constexpr int NS = {ns};
constexpr int NR = {nr};
constexpr long NCOEF = {len(coef)};
constexpr unsigned long long HASH = 0x{h:016x}ull;
constexpr double LOG2E = {repr(1.0 / math.log(2.0))};
constexpr int OFF_SMH = {off_smh};
// per reaction (mechanism order): reference type (0 standard, 1 troe, 2 three-body), index in
// that type's arrays, offset in the gpu_coef blob, Troe T2 present, number and offset of
// its listed collision partners
struct Rec {{ int type, idx, off, t2, ne, eoff; }};
const Rec kRec[NR] = {{
{chr(10).join(r + "," for r in rows)}
}};
const int kEffSp[{max(1, len(effsp))}] = {{ {", ".join(str(k) for k in effsp) or "0"} }};
const double kTmid[NS] = {{ {tmid} }};
cg_mech_t g_h = nullptr;

const char* code_object() {{
  const char* e = std::getenv("CHEMGEN_AMD_CODE_OBJECT");
  return e ? e : "{code_object}";
}}
}}  // namespace

// The reference's arrays -> this build's gpu_coef blob (cf, NCOEF doubles).  Returns 0, or 1
// for a NULL argument, 2 if the species maps do not match this mechanism, 3 for a Troe form
// (with / without T2) or T_mid that differs from it.
static double rowsum(const double* a) {{
  double s = 0.0;
  for (int k = 0; k < NS; ++k) s += std::fabs(a[k]);
  return s;
}}

extern "C" int cg_legacy_convert({", ".join(params)}, double* cf) {{
  if (!cf || !mw_h || !smh_coef_h || !T_mid_h) return 1;
{chr(10).join(chk_lines)}
  for (int k = 0; k < NS; ++k) {{
    if (std::fabs(T_mid_h[k] - kTmid[k]) > 0.005 * kTmid[k]) return 3;
    cf[k] = 1.0 / mw_h[k];
    for (int hl = 0; hl < 2; ++hl) {{
      const double* a = smh_coef_h + 14 * k + 7 * hl;  // a6-a0, a0, a1/2, a2/6, a3/12, a4/20, -a5
      double* c = cf + OFF_SMH + 14 * k + 7 * hl;
      c[0] = LOG2E * a[0];
      c[1] = a[1];
      for (int q = 2; q < 7; ++q) c[q] = LOG2E * a[q];
    }}
  }}
  for (int j = 0; j < NR; ++j) {{
    const Rec& r = kRec[j];
    if (r.type != 1) {{{body_std}
    }} else {{{conv_troe if "troe" in present else ""}
    }}
  }}
  return 0;
}}

extern "C" void copyConstantsToDevice({", ".join(params)}) {{
  std::vector<double> cf(NCOEF);
  const int rc = cg_legacy_convert({args}, cf.data());
  if (rc != 0) {{
    std::fprintf(stderr, "copyConstantsToDevice: the arrays do not belong to mechanism '{mech.name}' (code %d)\\n", rc);
    return;
  }}
  if (g_h) cg_mech_free(g_h);
  g_h = nullptr;
  if (cg_mech_load(code_object(), cf.data(), NCOEF, HASH, &g_h) != CG_OK)
    std::fprintf(stderr, "copyConstantsToDevice: %s\\n", cg_last_error());
}}

static void run(double* P, double* Temp, int ng, double* Ysp, double* WDOT_RR) {{
  if (!g_h) {{
    std::fprintf(stderr, "getrates_hip_wrapper: copyConstantsToDevice has not loaded the mechanism\\n");
    return;
  }}
  if (cg_getrates(g_h, P, Temp, ng, ng, Ysp, WDOT_RR, nullptr) != CG_OK)
    std::fprintf(stderr, "getrates_hip_wrapper: %s\\n", cg_last_error());
  (void)hipDeviceSynchronize();
}}

extern "C" void getrates_hip_wrapper_v1(double* P, double* Temp, int ng, double* Ysp, double* WDOT_RR,
                                        double* logEG_d, double* logC_d) {{
  (void)logEG_d; (void)logC_d;
  run(P, Temp, ng, Ysp, WDOT_RR);
}}
extern "C" void getrates_hip_wrapper_v2(double* P, double* Temp, int ng, double* Ysp, double* WDOT_RR,
                                        double* logEG_d, double* logC_d) {{
  (void)logEG_d; (void)logC_d;
  run(P, Temp, ng, Ysp, WDOT_RR);
}}
extern "C" void getrates_hip_wrapper_v3(double* P, double* Temp, int ng, double* Ysp, double* WDOT_RR,
                                        double* logEG_d, double* logC_d) {{
  (void)logEG_d; (void)logC_d;
  run(P, Temp, ng, Ysp, WDOT_RR);
}}
extern "C" void getrates_hip_wrapper_v4(double* P, double* Temp, int ng, double* Ysp, double* WDOT_RR,
                                        double* logEG_d, double* logC_d) {{
  (void)logEG_d; (void)logC_d;
  run(P, Temp, ng, Ysp, WDOT_RR);
}}
// level 4 needs caller-visible scratch in the reference; this build needs none
extern "C" void allocate_intermediate_DeviceMemory(int ng) {{ (void)ng; }}
'''
    return src


def build_legacy_library(mech: Mechanism, code_object: str, out_path: str, level: int = 1) -> str:
    """Write and compile the compatibility library for parallel level ``level`` (links
    libchemgen_amd.so)."""
    src = emit_legacy_source(mech, code_object, level)
    src_path = os.path.splitext(out_path)[0] + ".cpp"
    with open(src_path, "w") as fh:
        fh.write(src)
    # libchemgen_amd.so is found by its soname: the copy a process has already loaded (the
    # Python package loads it first), else through the rpath, relative to this library
    rel = os.path.relpath(LIB_DIR, os.path.dirname(os.path.abspath(out_path)))
    cmd = [HIPCC, *HIP_FLAGS, "-shared", "-fPIC", "-I", INCLUDE_DIR, src_path, "-L", LIB_DIR, "-lchemgen_amd",
           f"-Wl,-rpath,$ORIGIN/{rel}", "-o", out_path + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"legacy ABI library build failed:\n{res.stderr[-3000:]}")
    os.replace(out_path + ".tmp", out_path)
    return out_path
