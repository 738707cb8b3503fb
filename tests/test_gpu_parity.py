"""GPU parity: generated FP64 HIP getrates vs the CPU oracle, through the C ABI."""
import numpy as np
import pytest

from chemgen_amd.kernels import load_bundled
from chemgen_amd.synth import make_states
from oracle.getrates_np import getrates as oracle_getrates
from oracle.metrics import parity_stats

pytestmark = pytest.mark.gpu

TOL_GROSS = 1e-12   # max |dw| / (|w_ref| + sum_j (nu'+nu'')(|qf|+|qb|)) per species and point


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    return torch


def _run(torch, mech, P, T, Y, dm=None):
    from chemgen_amd.runtime import DeviceMechanism
    dm = dm or DeviceMechanism(mech, build=False)
    d = torch.device("cuda:0")
    Pd, Td, Yd = (torch.from_numpy(np.ascontiguousarray(x)).to(d) for x in (P, T, Y))
    w = dm.getrates(Pd, Td, Yd)
    torch.cuda.synchronize()
    return w.cpu().numpy()


@pytest.mark.parametrize("name,ng,trange", [
    ("h2_burke", 4096, (300.0, 2500.0)),
    ("h2_burke2012", 4096, (300.0, 2500.0)),
    ("gri30", 2048, (300.0, 2500.0)),
    ("synth_gri53", 2048, (300.0, 2500.0)),
    ("synth_nhept88", 1024, (600.0, 2800.0)),
    ("synth_jetsurf348", 512, (600.0, 2800.0)),
    ("synth_plog53", 2048, (300.0, 2500.0)),
    ("synth_ext53", 2048, (300.0, 2500.0)),
    ("fuller2021", 256, (600.0, 2800.0)),
    ("hashemi2016", 2048, (300.0, 2500.0)),
    ("ic8_llnl", 128, (600.0, 2800.0)),
])
def test_parity_vs_oracle(torch_cuda, name, ng, trange):
    m = load_bundled(name)
    print(f"[parity] {name} ng={ng}", flush=True)
    P, T, Y = make_states(m.n_species, ng, seed=0, T_range=trange)
    w = _run(torch_cuda, m, P, T, Y)
    ref, qf, qb = oracle_getrates(m, P, T, Y, return_parts=True)
    st = parity_stats(m, w, ref, qf, qb)
    print(name, st)
    assert st["finite"]
    assert st["max_gross"] <= TOL_GROSS, st


@pytest.mark.parametrize("ng", [1, 63, 65, 1000])
def test_ragged_tail_sizes(torch_cuda, ng):
    """Any ng works (the reference drops ng % VECLEN tails, hip_v1_cpp.j2:352)."""
    m = load_bundled("h2_burke")
    P, T, Y = make_states(m.n_species, ng, seed=ng)
    w = _run(torch_cuda, m, P, T, Y)
    ref, qf, qb = oracle_getrates(m, P, T, Y, return_parts=True)
    assert parity_stats(m, w, ref, qf, qb)["max_gross"] <= TOL_GROSS


@pytest.mark.parametrize("ng", [1, 55, 56, 57, 113])
def test_short_split_tile_tails(torch_cuda, ng):
    """The n-heptane-sized mechanism's tuned kernel has 56-point tiles (lanes 56..63 repeat
    lanes 0..7): tails of every length around the tile, against the oracle."""
    from chemgen_amd.autotune import best_variant
    m = load_bundled("synth_nhept88")
    assert best_variant(m).ppb == 56
    print(f"[tp56] ng={ng}", flush=True)
    P, T, Y = make_states(m.n_species, ng, seed=ng, T_range=(600.0, 2800.0))
    w = _run(torch_cuda, m, P, T, Y)
    ref, qf, qb = oracle_getrates(m, P, T, Y, return_parts=True)
    st = parity_stats(m, w, ref, qf, qb)
    assert st["finite"] and st["max_gross"] <= TOL_GROSS, st


@pytest.mark.parametrize("name", ["gri30", "synth_gri53", "synth_jetsurf348", "synth_plog53", "synth_ext53",
                                  "fuller2021", "hashemi2016", "ic8_llnl", "synth_eff53"])
def test_table_kernels_vs_oracle(torch_cuda, name):
    """Every table-driven kernel built for the mechanism (all tile lengths and wavefront counts
    that fit the LDS) against the oracle; gri30 brings Troe with and without T2 and Lindemann
    reactions, the JetSurf-sized mechanism SRI, synth_plog53 PLOG reactions (in compiled class
    loops and the generic loop), synth_ext53 extended records (sides of 4-5 species slots,
    coefficients 0.5, with PLOG, third-body and falloff rates), Fuller 2021 (835 species) all of
    them in 1- to 8-point tiles, Hashemi 2016 duplicate reactions with negative prefactors,
    synth_eff53 compiled third-body and Troe rounds whose 9-15 efficiencies the broadcast
    kernel of 8-point tiles loads in two parts; 300 points leave a partial tile for every tile
    length."""
    from chemgen_amd.kernels import default_variants
    from chemgen_amd.runtime import DeviceMechanism
    m = load_bundled(name)
    P, T, Y = make_states(m.n_species, 300, seed=4, T_range=(300.0, 2500.0))
    ref, qf, qb = oracle_getrates(m, P, T, Y, return_parts=True)
    tables = [v for v in default_variants(m) if v.table]
    assert tables
    for v in tables:
        print(f"[table] {name} {v.tag}", flush=True)
        w = _run(torch_cuda, m, P, T, Y, DeviceMechanism(m, v, build=False))
        st = parity_stats(m, w, ref, qf, qb)
        assert st["finite"] and st["max_gross"] <= TOL_GROSS, (v.tag, st)


def test_baked_coefficients_vs_oracle_and_guard(torch_cuda):
    """A kernel with the coefficients compiled in: parity with the oracle, and a blob that
    differs from the compiled one is refused at load and at cg_mech_set_coef."""
    from chemgen_amd.autotune import candidate_variants
    from chemgen_amd.codegen import build_coef
    from chemgen_amd.runtime import ChemgenError, DeviceMechanism
    m = load_bundled("synth_gri53")
    v = next(v for v in candidate_variants(m) if v.baked and v.split == 4)
    P, T, Y = make_states(m.n_species, 2048, seed=6)
    dm = DeviceMechanism(m, v, build=False)
    w = _run(torch_cuda, m, P, T, Y, dm)
    ref, qf, qb = oracle_getrates(m, P, T, Y, return_parts=True)
    st = parity_stats(m, w, ref, qf, qb)
    assert st["finite"] and st["max_gross"] <= TOL_GROSS, st
    coef = build_coef(m)[0].copy()
    coef[-1] *= 1.5
    with pytest.raises(ChemgenError):
        dm.set_coef(coef)
    with pytest.raises(ChemgenError):
        DeviceMechanism(m, v, build=False, coef=coef)


@pytest.mark.parametrize("name", ["h2_burke", "gri30", "synth_gri53", "synth_nhept88", "synth_plog53"])
def test_live_order_kernels_vs_oracle(torch_cuda, name):
    """Lane kernels with reactions in live-range order (species state made at first use, wdot
    stored after last use), with the coefficients read and, where built, compiled in."""
    from chemgen_amd.codegen import Variant
    from chemgen_amd.kernels import default_variants
    from chemgen_amd.runtime import DeviceMechanism
    m = load_bundled(name)
    P, T, Y = make_states(m.n_species, 1000, seed=8, T_range=(300.0, 2500.0))
    ref, qf, qb = oracle_getrates(m, P, T, Y, return_parts=True)
    lives = [v for v in default_variants(m) if v.live]
    assert Variant(64, 1, "lean", "gh", "horner", False, 4, 1, 0, False, True) in lives
    for v in lives:
        print(f"[live] {name} {v.tag}", flush=True)
        w = _run(torch_cuda, m, P, T, Y, DeviceMechanism(m, v, build=False))
        st = parity_stats(m, w, ref, qf, qb)
        assert st["finite"] and st["max_gross"] <= TOL_GROSS, (v.tag, st)


def _edge_states(m):
    """Crafted states: T at 300 K, at and next to each NASA-7 range switch, at 3000 and
    3500 K (beyond the fitted range); P from 0.01 to 100 atm; compositions that are a single
    species, sparse (most mass fractions exactly 0), and uniform."""
    rng = np.random.default_rng(11)
    ns = m.n_species
    tmids = sorted({s.thermo.t_mid for s in m.species})
    temps = [300.0, 3000.0, 3500.0] + [x for t in tmids for x in (t, np.nextafter(t, 0.0), np.nextafter(t, 1e9))]
    pres = [0.01 * 1013250.0, 1013250.0, 100.0 * 1013250.0]
    comps = []
    for k in (0, ns // 2, ns - 1):
        y = np.zeros(ns); y[k] = 1.0; comps.append(y)
    for _ in range(3):
        y = rng.dirichlet(0.3 * np.ones(ns)); y[y < np.quantile(y, 0.8)] = 0.0; comps.append(y / y.sum())
    comps.append(np.full(ns, 1.0 / ns))
    P, T, Y = [], [], []
    for t in temps:
        for p in pres:
            for y in comps:
                P.append(p); T.append(t); Y.append(y)
    return np.array(P), np.array(T), np.ascontiguousarray(np.array(Y).T)


@pytest.mark.parametrize("name", ["h2_burke", "gri30", "synth_gri53", "synth_nhept88", "synth_jetsurf348",
                                  "synth_plog53", "synth_ext53", "fuller2021", "hashemi2016"])
def test_edge_states_vs_oracle(torch_cuda, name):
    """The tuned kernel of each mechanism on the crafted states of _edge_states.  Where the
    oracle itself overflows (Fuller 2021 at 3000 K and above: exp(smh) of species whose NASA-7
    fits end below that) the point is left out."""
    m = load_bundled(name)
    P, T, Y = _edge_states(m)
    print(f"[edge] {name} ng={P.shape[0]}", flush=True)
    w = _run(torch_cuda, m, P, T, Y)
    with np.errstate(over="ignore", invalid="ignore"):
        ref, qf, qb = oracle_getrates(m, P, T, Y, return_parts=True)
    keep = np.isfinite(ref).all(axis=0) & np.isfinite(qf).all(axis=0) & np.isfinite(qb).all(axis=0)
    assert keep.mean() > 0.95 and (keep.all() or name == "fuller2021")
    st = parity_stats(m, w[:, keep], ref[:, keep], qf[:, keep], qb[:, keep])
    assert st["finite"] and st["max_gross"] <= TOL_GROSS, st


def test_zero_points(torch_cuda):
    """ng = 0 launches nothing and is not an error."""
    from chemgen_amd.runtime import DeviceMechanism
    m = load_bundled("synth_gri53")
    dm = DeviceMechanism(m, build=False)
    d = torch_cuda.device("cuda:0")
    P = torch_cuda.empty(0, dtype=torch_cuda.float64, device=d)
    Y = torch_cuda.empty(m.n_species, 0, dtype=torch_cuda.float64, device=d)
    W = torch_cuda.empty_like(Y)
    dm.getrates_(P, P.clone(), Y, W)
    torch_cuda.cuda.synchronize()
