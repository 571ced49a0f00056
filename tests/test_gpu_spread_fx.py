"""GPU: the FP32-bulk spread in 32-bit fixed point (k_spread_fx.hip).

Each window product is rounded to an integer at a scale set per call from max |q| and a
per-configuration bound on the summed window weights; the integer mesh is converted back
to float by the first FFT pass (or by cr_spread's readback).  Checks: the oracle's FP64
spread to FP32 accuracy (rel <= 1e-6) for generic and compile-time tile shapes and odd P,
bit-identical repeats (integer sums do not depend on the order of the atomics), scale
invariance, exact zeros, and NaN propagation.
"""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import GpuCase, n1k_system, rel, small_system

pytestmark = pytest.mark.gpu

N1K = dict(xi=0.929, errortol=1e-6, rc=4.0, mesh=(64, 64, 64), P=8)
SMALL = dict(xi=0.8, errortol=1e-5, rc=3.5, mesh=(30, 30, 30), P=7)


def _report(name, value):
    print(f"[spread_fx] {name}: {value:.3e}", flush=True)
    sys.stdout.flush()


def _maxrel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module", params=["n1k", "small"])
def fx_case(request, require_gpu):
    if request.param == "n1k":
        s, prm = n1k_system(), N1K
    else:
        s, prm = small_system(), SMALL
    print(f"[case] fx {request.param} N={s.N}", flush=True)
    g = GpuCase(s, precision="fp32", **prm)
    o = O.Oracle(s, prm["xi"], prm["errortol"], rc=prm["rc"], mesh=prm["mesh"], P=prm["P"])
    return request.param, s, g, o


def test_fx_spread_matches_oracle(fx_case):
    name, s, g, o = fx_case
    q = np.random.default_rng(21).normal(size=s.N)
    ref = o.spread(q)
    got = g.spread(q)
    v, m = rel(got, ref), _maxrel(got, ref)
    _report(f"{name} rel", v)
    _report(f"{name} max", m)
    assert v < 1e-6 and m < 1e-6


@pytest.mark.parametrize("tile", [(16, 16, 16), (8, 8, 8), (16, 8, 16), (32, 16, 8)])
def test_fx_tile_shapes(tile, require_gpu):
    """Compile-time tile shapes (16^3, 8^3, 16x8x16 at P = 8) and the run-time-shape kernel."""
    s = n1k_system()
    g = GpuCase(s, precision="fp32", **N1K)
    g.cr.set_tiles(tile, 160 * 1024)
    o = O.Oracle(s, N1K["xi"], N1K["errortol"], rc=N1K["rc"], mesh=N1K["mesh"], P=N1K["P"])
    q = np.random.default_rng(22).normal(size=s.N)
    ref = o.spread(q)
    got = g.spread(q)
    v = rel(got, ref)
    _report(f"tile {tile}", v)
    assert v < 1e-6 and _maxrel(got, ref) < 1e-6


@pytest.mark.parametrize("mesh,tile", [((128, 64, 64), (8, 8, 8)), ((128, 128, 64), (8, 8, 8)),
                                       ((128, 128, 128), (8, 8, 8)), ((256, 128, 128), (16, 16, 16))])
def test_fx_persistent_runs(mesh, tile, require_gpu):
    """Meshes of 1024-4096 tiles: the persistent spread takes runs of 2, 4 and 8 tiles along
    z and carries each tile's z halo in LDS into the next tile (the last halo of a run goes
    to the first tile of the next run, wrapping at the mesh edge); the potential contraction
    of the same tiling reads the mesh back."""
    s = n1k_system()
    prm = dict(N1K, mesh=mesh)
    g = GpuCase(s, precision="fp32", **prm)
    g.cr.set_tiles(tile, 160 * 1024)
    o = O.Oracle(s, prm["xi"], prm["errortol"], rc=prm["rc"], mesh=mesh, P=prm["P"])
    q = np.random.default_rng(23).normal(size=s.N)
    ref = o.spread(q)
    got = g.spread(q)
    v = rel(got, ref)
    _report(f"runs {mesh} {tile}", v)
    assert v < 1e-6 and _maxrel(got, ref) < 1e-6
    assert np.array_equal(got, g.spread(q))
    _, phi = g.contract_after_recip(ref)
    vc = rel(phi, o.contract(o.recip_grid(ref)))
    _report(f"runs {mesh} {tile} contract", vc)
    assert vc < 1e-6


def test_fx_spread_is_deterministic(fx_case):
    _, s, g, o = fx_case
    q = np.random.default_rng(23).normal(size=s.N)
    a = g.spread(q)
    b = g.spread(q)
    assert np.array_equal(a.view(np.int64), b.view(np.int64))


def test_fx_spread_scale_invariant(fx_case):
    """The fixed-point scale follows max |q|: tiny and huge charges keep full precision."""
    _, s, g, o = fx_case
    q = np.random.default_rng(24).normal(size=s.N)
    ref = o.spread(q)
    for f in (1e-25, 1e25):
        v = rel(g.spread(q * f) / f, ref)
        _report(f"scale {f:g}", v)
        assert v < 1e-6


def test_fx_spread_zero_and_nan(fx_case):
    _, s, g, o = fx_case
    assert not np.any(g.spread(np.zeros(s.N)))
    q = np.random.default_rng(25).normal(size=s.N)
    q[s.N // 2] = np.nan
    assert np.any(np.isnan(g.spread(q)))
    # the context recovers for finite charges
    q[s.N // 2] = 0.0
    assert rel(g.spread(q), o.spread(q)) < 1e-6


def test_fx_matvec_and_solve(fx_case):
    """FP32-bulk operator against the oracle's FP64 operator, and the solve against FP64."""
    name, s, g, o = fx_case
    x = np.random.default_rng(26).normal(size=g.n)
    v = rel(g.matvec(x), o.matvec(x))
    _report(f"{name} fp32 matvec", v)
    assert v < 1e-5


@pytest.mark.parametrize("tile", [None, (8, 8, 8), (16, 8, 16), (16, 16, 16), (8, 16, 32)])
def test_fx_spread_p9_batch_box(tile, require_gpu):
    """P = 9 on a 32^3 mesh (a box of the batched workload): compile-time and run-time
    tile shapes against the oracle."""
    from conductorrigid_amd.system import make_batch

    s, p = make_batch(1, seed=11, n_range=(400, 3000), phi_range=(0.1, 0.3))[0]
    prm = dict(xi=p["xi"], errortol=p["errortol"], rc=p["rc"], mesh=tuple(p["mesh"]), P=p["P"])
    print(f"[case] fx p9 N={s.N} {prm} tile={tile}", flush=True)
    g = GpuCase(s, precision="fp32", **prm)
    if tile is not None:
        if any(m % t for m, t in zip(prm["mesh"], tile)):
            pytest.skip("tile does not divide this mesh")
        g.cr.set_tiles(tile, 160 * 1024)
    o = O.Oracle(s, prm["xi"], prm["errortol"], rc=prm["rc"], mesh=prm["mesh"], P=prm["P"])
    q = np.random.default_rng(27).normal(size=s.N)
    ref = o.spread(q)
    got = g.spread(q)
    v = rel(got, ref)
    _report(f"p9 tile {tile}", v)
    x = np.random.default_rng(28).normal(size=g.n)
    vm = rel(g.matvec(x), o.matvec(x))
    _report(f"p9 tile {tile} matvec", vm)
    assert v < 1e-6 and vm < 1e-5


@pytest.mark.parametrize("tile", [None, (16, 16, 16), (8, 8, 8), (16, 8, 16), (32, 16, 8)])
def test_fx_contraction_matches_oracle(tile, require_gpu):
    """FP32-bulk potential contraction (k_contract_fx): compile-time and run-time tile
    shapes against the oracle's FP64 contraction of the same mesh (rel <= 1e-6)."""
    s = n1k_system()
    g = GpuCase(s, precision="fp32", **N1K)
    if tile is not None:
        g.cr.set_tiles(tile, 160 * 1024)
    o = O.Oracle(s, N1K["xi"], N1K["errortol"], rc=N1K["rc"], mesh=N1K["mesh"], P=N1K["P"])
    grid = o.spread(np.random.default_rng(29).normal(size=s.N))
    og = o.recip_grid(grid)
    gg, phi = g.contract_after_recip(grid)
    v = rel(phi, o.contract(og))
    _report(f"contract tile {tile}", v)
    assert v < 1e-6


def test_fx_contraction_odd_p(fx_case):
    name, s, g, o = fx_case
    grid = o.spread(np.random.default_rng(30).normal(size=s.N))
    og = o.recip_grid(grid)
    v = rel(g.contract_after_recip(grid)[1], o.contract(og))
    _report(f"{name} contract", v)
    assert v < 1e-6


@pytest.mark.parametrize("tile", [None, (16, 16, 16), (8, 8, 8), (16, 8, 16), (32, 16, 8)])
def test_fx_force_contraction_matches_oracle(tile, require_gpu):
    """FP32-bulk reciprocal force (k_contract_fx in force mode): compile-time and run-time
    tile shapes against the oracle's FP64 contract_force of the same mesh."""
    s = n1k_system()
    g = GpuCase(s, precision="fp32", **N1K)
    if tile is not None:
        g.cr.set_tiles(tile, 160 * 1024)
    o = O.Oracle(s, N1K["xi"], N1K["errortol"], rc=N1K["rc"], mesh=N1K["mesh"], P=N1K["P"])
    q = np.random.default_rng(31).normal(size=s.N)
    grid = o.spread(q)
    og = o.recip_grid(grid)
    F = g.contract_force_after_recip(grid, q)
    v = rel(F[:, :3], o.contract_force(q, og)[:, :3])
    _report(f"force contract tile {tile}", v)
    assert v < 5e-6
