"""The generated module's ``wdot = getrates(P, T, Y)`` and the ragged-batch path, on the GPU."""
import importlib.util
import os

import numpy as np
import pytest

from chemgen_amd.kernels import load_bundled
from chemgen_amd.synth import make_states
from oracle.getrates_np import getrates as oracle_getrates
from oracle.metrics import parity_stats

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, "chemgen_amd", "_kernels", "gen_h2_burke", "getrates_hip.py")
TOL_GROSS = 1e-12


def _module():
    spec = importlib.util.spec_from_file_location("getrates_hip", GEN)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_generated_module_single_state():
    """The reference test's state (tests/test_getrates.py:63-92, conftest.py defaults):
    H2/air at equivalence ratio 0.5, T = 1200 K, P = 101325 Pa; compared with the oracle
    (the reference compared with Cantera, which is not available here)."""
    mod = _module()
    m = load_bundled("h2_burke")
    X = np.zeros(m.n_species)
    names = m.species_names
    X[names.index("H2")], X[names.index("O2")], X[names.index("N2")] = 1.0, 1.0, 3.76
    W = m.mw_array()
    Y = X * W / np.sum(X * W)
    P, T = 101325.0 * 10.0, 1200.0
    print("[api] single state", flush=True)
    w = mod.getrates(P, T, Y)
    assert w.shape == (m.n_species,)
    ref, qf, qb = oracle_getrates(m, np.array([P]), np.array([T]), Y[:, None], return_parts=True)
    assert parity_stats(m, w[:, None], ref, qf, qb)["max_gross"] <= TOL_GROSS


def test_generated_module_point_and_species_layouts():
    mod = _module()
    m = load_bundled("h2_burke")
    P, T, Y = make_states(m.n_species, 333, seed=4)
    ref, qf, qb = oracle_getrates(m, P, T, Y, return_parts=True)
    w_point = mod.getrates(P, T, Y.T)               # Y[ng, Ns]
    assert w_point.shape == (333, m.n_species)
    assert parity_stats(m, w_point.T, ref, qf, qb)["max_gross"] <= TOL_GROSS
    import torch
    d = torch.device("cuda:0")
    w_dev = mod._impl(torch.from_numpy(P).to(d), torch.from_numpy(T).to(d), torch.from_numpy(Y).to(d))
    torch.cuda.synchronize()
    assert parity_stats(m, w_dev.cpu().numpy(), ref, qf, qb)["max_gross"] <= TOL_GROSS


@pytest.mark.parametrize("grid", [7, -1])
@pytest.mark.parametrize("name", ["h2_burke2012", "synth_gri53"])
def test_ragged_batch_matches_oracle(name, grid):
    """Ragged patches (sizes 1..300, some empty) in one launch: 7 persistent workgroups
    walking the tile list, or one workgroup per tile (the loop-free kernel)."""
    import torch
    from chemgen_amd.autotune import best_variant
    from chemgen_amd.runtime import DeviceMechanism
    m = load_bundled(name)
    rng = np.random.default_rng(3)
    sizes = rng.integers(0, 300, 17)
    sizes[[2, 9]] = 0
    sizes[5] = 1
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    ng = int(offs[-1])
    P, T, Y = make_states(m.n_species, ng, seed=6)
    print(f"[batch] {name} ng={ng} patches={len(sizes)} grid={grid}", flush=True)
    dm = DeviceMechanism(m, best_variant(m), build=False)
    d = torch.device("cuda:0")
    Pd, Td, Yd = (torch.from_numpy(x).to(d) for x in (P, T, Y))
    W = torch.full_like(Yd, float("nan"))
    dm.getrates_batched_(Pd, Td, Yd, W, offs, grid=grid)
    torch.cuda.synchronize()
    ref, qf, qb = oracle_getrates(m, P, T, Y, return_parts=True)
    assert parity_stats(m, W.cpu().numpy(), ref, qf, qb)["max_gross"] <= TOL_GROSS


@pytest.mark.parametrize("name,tags", [
    ("synth_gri53", ("b64w1lghhd", "b64w1lghhkod", "s4w2lghhg8d", "s4w3lghhg2d", "s8w4lghhg4d")),
    ("gri30", ("t16s8w2lhdb", "b64w2lghhL4d")),
    ("synth_ext53", ("t16s4w1lh", "t4s8w2lh")),
    ("fuller2021", ("t8s16w4lhdb", "t2s8w2lhd", "t1s4w1lhd")),
    ("synth_nhept88", ("s8w4lghhg2kdn56",)),
])
def test_ragged_batch_kernel_families(name, tags):
    """Both batched kernels (persistent walk, one workgroup per tile) of each kernel family
    -- lane, live-range lane with compiled-in coefficients, reaction-split (also with the
    compact LDS layout, and with 56-point tiles whose last 8 lanes repeat the first 8), table
    (also with PLOG and extended records, and with the 1- to 8-point tiles of an 835-species
    mechanism), EG in LDS -- against the oracle on ragged patches."""
    import torch
    from chemgen_amd.autotune import candidate_variants
    from chemgen_amd.kernels import default_variants
    from chemgen_amd.runtime import DeviceMechanism
    m = load_bundled(name)
    sizes = np.array([130, 0, 64, 1, 250, 63, 65])
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    ng = int(offs[-1])
    P, T, Y = make_states(m.n_species, ng, seed=9)
    ref, qf, qb = oracle_getrates(m, P, T, Y, return_parts=True)
    d = torch.device("cuda:0")
    Pd, Td, Yd = (torch.from_numpy(x).to(d) for x in (P, T, Y))
    by_tag = {v.tag: v for v in candidate_variants(m) + default_variants(m)}
    for tag in tags:
        dm = DeviceMechanism(m, by_tag[tag], build=False)
        for grid in (3, -1):
            print(f"[batch-family] {name} {tag} grid={grid}", flush=True)
            W = torch.full_like(Yd, float("nan"))
            dm.getrates_batched_(Pd, Td, Yd, W, offs, grid=grid)
            torch.cuda.synchronize()
            st = parity_stats(m, W.cpu().numpy(), ref, qf, qb)
            assert st["finite"] and st["max_gross"] <= TOL_GROSS, (tag, grid, st)
        dm.close()


def test_batched_graph_replay():
    """The prepared batched launch captured in a hipGraph and replayed; every replay is
    checked against the eager result, with eager work in between."""
    import torch
    from chemgen_amd.autotune import best_variant
    from chemgen_amd.runtime import DeviceMechanism
    m = load_bundled("synth_gri53")
    sizes = np.array([1000, 37, 4096, 5, 700])
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    ng = int(offs[-1])
    P, T, Y = make_states(m.n_species, ng, seed=8)
    dm = DeviceMechanism(m, best_variant(m), build=False)
    d = torch.device("cuda:0")
    Pd, Td, Yd = (torch.from_numpy(x).to(d) for x in (P, T, Y))
    eager = torch.empty_like(Yd)
    dm.getrates_batched_(Pd, Td, Yd, eager, offs)
    torch.cuda.synchronize()
    W = torch.zeros_like(Yd)
    dm.batched_prepare(offs)
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            dm.getrates_batched_prepared_(Pd, Td, Yd, W, stream=s)
    torch.cuda.synchronize()
    for rep in range(3):
        W.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(W, eager), f"replay {rep}"
        _ = (Yd * 2.0).sum().item()  # eager work between replays


def _capture(fn, s):
    import torch
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            fn()
    return g


def test_batch_graphs_survive_reprepare():
    """Graphs over different batches stay valid after a re-prepare with more tiles: two batch
    objects and two handle-prepared batches (the second prepare replaces the first with about
    three times the tiles) are captured, then all four graphs are replayed interleaved, three
    rounds, each replay bit-identical to the eager one-shot result of its batch."""
    import torch
    from chemgen_amd.autotune import best_variant
    from chemgen_amd.runtime import DeviceMechanism
    m = load_bundled("synth_gri53")
    dm = DeviceMechanism(m, best_variant(m), build=False)
    d = torch.device("cuda:0")
    layouts = [np.array([300, 0, 77, 1200]), np.array([64, 64, 5, 2000, 13]),
               np.array([900, 31, 0, 650]), np.array([2500, 1, 800, 1700, 3, 990])]
    offs = [np.concatenate([[0], np.cumsum(sz)]).astype(np.int64) for sz in layouts]
    ng = max(int(o[-1]) for o in offs)
    P, T, Y = make_states(m.n_species, ng, seed=12)
    Pd, Td, Yd = (torch.from_numpy(x).to(d) for x in (P, T, Y))
    eager = []
    for o in offs:
        e = torch.full_like(Yd, -1.0)
        dm.getrates_batched_(Pd, Td, Yd, e, o)
        eager.append(e)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    outs = [torch.full_like(Yd, -1.0) for _ in offs]
    b0, b1 = dm.batch(offs[0], stream=s), dm.batch(offs[1], stream=s)
    print(f"[batch-graphs] tiles {b0.n_tiles} {b1.n_tiles}", flush=True)
    graphs = [_capture(lambda: b0.getrates_(Pd, Td, Yd, outs[0], stream=s), s),
              _capture(lambda: b1.getrates_(Pd, Td, Yd, outs[1], stream=s), s)]
    dm.batched_prepare(offs[2])
    graphs.append(_capture(lambda: dm.getrates_batched_prepared_(Pd, Td, Yd, outs[2], stream=s), s))
    dm.batched_prepare(offs[3])          # more tiles than offs[2]: the earlier list must survive
    graphs.append(_capture(lambda: dm.getrates_batched_prepared_(Pd, Td, Yd, outs[3], stream=s), s))
    torch.cuda.synchronize()
    for rnd in range(3):
        for k in (2, 0, 3, 1):
            outs[k].fill_(-1.0)
            graphs[k].replay()
            torch.cuda.synchronize()
            assert torch.equal(outs[k], eager[k]), f"round {rnd} graph {k}"
            _ = (Yd * 2.0).sum().item()  # eager work between replays
    b0.close()
    b1.close()


def test_oneshot_batch_refused_under_capture():
    """The one-shot form builds a tile list per call; under capture it refuses (a batch
    object is the capturable form)."""
    import torch
    from chemgen_amd.runtime import ChemgenError, DeviceMechanism
    m = load_bundled("h2_burke")
    dm = DeviceMechanism(m, build=False)
    offs = np.array([0, 100, 300], dtype=np.int64)
    P, T, Y = make_states(m.n_species, 300, seed=2)
    d = torch.device("cuda:0")
    Pd, Td, Yd = (torch.from_numpy(x).to(d) for x in (P, T, Y))
    W = torch.empty_like(Yd)
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    err = None
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            _ = Yd * 1.0          # one captured node, so the capture itself is not empty
            try:
                dm.getrates_batched_(Pd, Td, Yd, W, offs, stream=s)
            except ChemgenError as e:
                err = e
    torch.cuda.synchronize()
    assert err is not None and "capturing" in str(err)


def test_fp64_probe_values():
    """The FMA probe behind the bench's measured FP64 ceiling computes what it claims: eight
    chains x <- a x + b per lane, x0 = lane/1000 + c for chain c."""
    import ctypes
    import torch
    from chemgen_amd.runtime import lib
    L = lib()
    L.cg_probe_fp64_fma.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]
    L.cg_probe_fp64_fma.restype = ctypes.c_int
    blocks, iters = 3, 50
    out = torch.zeros(blocks * 256, dtype=torch.float64, device="cuda:0")
    assert L.cg_probe_fp64_fma(out.data_ptr(), blocks, iters, None) == 0
    torch.cuda.synchronize()
    a, b = 0.999999, 1e-7
    lane = np.tile(np.arange(256), blocks) * 1e-3
    x = np.stack([lane + c for c in range(8)])
    for _ in range(iters):
        x = a * x + b
    ref = ((x[0] + x[1]) + (x[2] + x[3])) + ((x[4] + x[5]) + (x[6] + x[7]))
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-13)
    assert L.cg_probe_fp64_fma(out.data_ptr(), 0, iters, None) != 0


def test_batched_launch_checks_leading_dimension():
    """A prepared batch refuses arrays whose leading dimension is shorter than the batch's
    last point (the tiles would run past the species rows)."""
    import torch
    from chemgen_amd.runtime import ChemgenError, DeviceMechanism
    m = load_bundled("h2_burke")
    dm = DeviceMechanism(m, build=False)
    offs = np.array([0, 100, 300], dtype=np.int64)
    d = torch.device("cuda:0")
    P = torch.ones(300, dtype=torch.float64, device=d) * 1013250.0
    T = torch.ones(300, dtype=torch.float64, device=d) * 1200.0
    Y = torch.full((m.n_species, 200), 1.0 / m.n_species, dtype=torch.float64, device=d)
    W = torch.empty_like(Y)
    dm.batched_prepare(offs)
    with pytest.raises(ChemgenError):
        dm.getrates_batched_prepared_(P, T, Y, W)
    with pytest.raises(ChemgenError):
        dm.getrates_batched_(P, T, Y, W, offs)
