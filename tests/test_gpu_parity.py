"""GPU parity: the HIP library (through its C ABI) against the CPU FP64 oracle.

Bar: FP64 ("plain method") stages agree with the oracle to 1e-12 relative, the
full solve at a tight GMRES tolerance to 1e-10 relative (charges, body potentials,
forces); neighbour lists are compared as exact sets.  The FP32-bulk mode is held
to an accuracy gate of 1e-5 relative force error against the FP64 run.
"""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import GpuCase, n1k_system, rel, small_system

pytestmark = pytest.mark.gpu

N1K = dict(xi=0.929, errortol=1e-6, rc=4.0, mesh=(64, 64, 64), P=8)
SMALL = dict(xi=0.8, errortol=1e-5, rc=3.5, mesh=(30, 30, 30), P=7)   # radix 2/3/5 mesh, odd P


def _report(name, value):
    print(f"[parity] {name}: {value:.3e}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", params=["n1k", "small"])
def case(request, require_gpu):
    if request.param == "n1k":
        s, prm = n1k_system(), N1K
    else:
        s, prm = small_system(), SMALL
    print(f"[case] {request.param} N={s.N} bodies={s.N_isolated}", flush=True)
    g = GpuCase(s, gmres_tol=1e-12, restart=30, max_iter=600, **prm)
    head, nn, nl = g.cr.dsys.nlist_host()
    o = O.Oracle(s, prm["xi"], prm["errortol"], rc=prm["rc"], mesh=prm["mesh"], P=prm["P"],
                 nlist=(head, nn, nl))
    return request.param, s, g, o


def test_resolved_params_match_oracle(case):
    _, s, g, o = case
    p = g.cr.params
    assert p["mesh"] == list(o.p.N)
    assert p["P"] == o.p.P and p["N_table"] == o.p.N_table
    for k in ("rc", "selfterm", "prefac", "xiterm", "quadW"):
        assert abs(p[k] - getattr(o.p, k)) <= 1e-15 * abs(getattr(o.p, k)), k


def test_nlist_matches_oracle(case):
    _, s, g, o = case
    head, nn, nl = g.cr.dsys.nlist_host()
    ho, no, lo = O.build_nlist(s.pos, s.L, g.cr.rcut + g.cr.skin)
    assert np.array_equal(nn, no)
    for i in range(s.N):
        a = set(nl[head[i]:head[i] + nn[i]].tolist())
        b = set(lo[ho[i]:ho[i] + no[i]].tolist())
        assert a == b, i


def test_spread(case):
    _, s, g, o = case
    q = np.random.default_rng(3).normal(size=s.N)
    v = rel(g.spread(q), o.spread(q))
    _report("spread", v)
    assert v < 1e-12


def test_fft_c2c_matches_numpy(case):
    _, s, g, o = case
    rng = np.random.default_rng(4)
    z = rng.normal(size=g.mesh) + 1j * rng.normal(size=g.mesh)
    f = g.fft(z, -1)
    v1 = rel(f, np.fft.fftn(z))
    b = g.fft(z, +1)
    v2 = rel(b, np.fft.ifftn(z) * z.size)
    _report("fft fwd", v1)
    _report("fft inv", v2)
    assert v1 < 1e-13 and v2 < 1e-13


def test_recip_and_contract(case):
    _, s, g, o = case
    q = np.random.default_rng(5).normal(size=s.N)
    grid = o.spread(q)
    gg, phi = g.contract_after_recip(grid)
    og = o.recip_grid(grid)
    v = rel(gg, og)
    _report("recip grid", v)
    assert v < 1e-12
    v = rel(phi, o.contract(og))
    _report("contract", v)
    assert v < 1e-12
    F = g.contract_force_after_recip(grid, q)
    v = rel(F[:, :3], o.contract_force(q, og)[:, :3])
    _report("contract_force", v)
    assert v < 1e-12


def test_real_space(case):
    _, s, g, o = case
    rng = np.random.default_rng(6)
    x = rng.normal(size=g.n)
    phi0 = rng.normal(size=s.N) * 0.01
    y = g.real_space(x, phi0)
    # oracle: matvec pieces
    import ctypes as C
    yo = phi0.copy()
    tab, _ = o.tables()
    tabf = np.ascontiguousarray(tab.ravel())
    iso = np.ascontiguousarray(s.iso_membership(), dtype=np.uint32)
    O.lib().cro_real_space(C.byref(o.p), C.byref(o.sys), O._d(tabf), O._u(iso), O._d(np.ascontiguousarray(x[:s.N])),
                           O._d(np.ascontiguousarray(x[s.N:])), O._d(yo))
    v = rel(y, yo)
    _report("real_space", v)
    assert v < 1e-12
    F0 = np.zeros((s.N, 4))
    F0[:, :3] = rng.normal(size=(s.N, 3))
    F = g.real_space_force(x[:s.N], F0)
    _, gtab = o.tables()
    Fo = F0.copy()
    O.lib().cro_real_space_force(C.byref(o.p), C.byref(o.sys), O._d(np.ascontiguousarray(gtab.ravel())),
                                 O._d(np.ascontiguousarray(x[:s.N])), O._d(np.array([0.0, 0.0, 1.0])), O._d(Fo))
    v = rel(F[:, :3], Fo[:, :3])
    _report("real_space_force", v)
    assert v < 1e-12


def test_objective_precondition(case):
    _, s, g, o = case
    b = g.objective()
    assert rel(b, o.objective([0.0, 0.0, 1.0])) < 1e-14
    x = np.random.default_rng(8).normal(size=g.n)
    assert rel(g.precondition(x), o.precondition(x)) < 1e-15


def test_matvec(case):
    _, s, g, o = case
    x = np.random.default_rng(9).normal(size=g.n)
    v = rel(g.matvec(x), o.matvec(x))
    _report("matvec", v)
    assert v < 1e-12


def test_solve_tight_matches_oracle(case):
    name, s, g, o = case
    q, Phi, F, st = g.solve("zero")
    qo, Po, Fo, so = o.compute_forces([0.0, 0.0, 1.0], restart=30, maxiter=600, tol=1e-12)
    print(f"[solve] {name} gpu iters={st['iterations']} conv={st['converged']} oracle iters={so['iters']}", flush=True)
    assert st["converged"] and so["converged"]
    for nm, a, b in (("charges", q, qo), ("potentials", Phi, Po), ("forces", F[:, :3], Fo[:, :3])):
        v = rel(a, b)
        _report(f"solve {nm}", v)
        assert v < 1e-10, nm
    # constraints hold in FP64
    sums = np.bincount(s.iso_membership(), weights=q, minlength=s.N_isolated)
    assert np.max(np.abs(sums - s.body_charge)) < 1e-9


@pytest.mark.parametrize("mode", ["fp32"])
def test_fp32_bulk_accuracy(mode, require_gpu):
    s = n1k_system()
    g64 = GpuCase(s, gmres_tol=1e-10, restart=30, max_iter=400, **N1K)
    q64, P64, F64, st64 = g64.solve("zero")
    g32 = GpuCase(s, gmres_tol=1e-6, restart=30, max_iter=400, precision="fp32", **N1K)
    q32, P32, F32, st32 = g32.solve("zero")
    print(f"[fp32] iters fp64={st64['iterations']} fp32={st32['iterations']}", flush=True)
    vq, vF = rel(q32, q64), rel(F32[:, :3], F64[:, :3])
    _report("fp32 charges vs fp64", vq)
    _report("fp32 forces vs fp64", vF)
    assert vF < 1e-5 and vq < 1e-5


@pytest.mark.parametrize("tile", [(8, 8, 8), (16, 8, 8), (8, 16, 32)])
def test_tiling_does_not_change_the_answer(case, tile):
    """Every admissible tile shape (and the autotuner's choice) gives the oracle's
    spread and contraction and the same tight solve."""
    name, s, g, o = case
    mesh = g.mesh
    if any(m % t for m, t in zip(mesh, tile)):
        pytest.skip("tile does not divide this mesh")
    g.cr.set_tiles(tile, 160 * 1024)
    q = np.random.default_rng(5).normal(size=s.N)
    assert rel(g.spread(q), o.spread(q)) < 1e-12
    og = o.recip_grid(o.spread(q))
    v = rel(g.contract_after_recip(o.spread(q))[1], o.contract(og))
    _report(f"contract tile {tile}", v)
    assert v < 1e-12
    tuned = g.cr.autotune()
    assert tuned["candidates"] >= 1 and all(m % t == 0 for m, t in zip(mesh, tuned["tile"]))
    assert rel(g.spread(q), o.spread(q)) < 1e-12
    qs, Phi, F, st = g.solve("zero")
    qo, Po, Fo, so = o.compute_forces([0.0, 0.0, 1.0], restart=30, maxiter=600, tol=1e-12)
    assert rel(qs, qo) < 1e-10 and rel(F[:, :3], Fo[:, :3]) < 1e-10


# Register four-step FFT lengths n = A * B (32 = 8x4, 128 = 16x8, 256 = 16x16, 512 = 32x16,
# 1024 = 32x32 beside the square 64): every axis pass (z R2C / C2R / C2C, strided y, fused
# x + scale) against numpy's FFT and against the oracle's k-space step.  256 is the
# headline mesh's length on every axis.
@pytest.mark.parametrize("mesh", [(32, 128, 64), (128, 32, 512), (512, 64, 32), (256, 32, 64), (64, 256, 32),
                                  (32, 64, 256), (1024, 32, 32), (32, 32, 1024)])
def test_register_fft_lengths(mesh, require_gpu):
    s = n1k_system()
    prm = dict(xi=0.929, errortol=1e-6, rc=4.0, mesh=mesh, P=8)
    g = GpuCase(s, **prm)
    rng = np.random.default_rng(9)
    z = rng.normal(size=mesh) + 1j * rng.normal(size=mesh)
    v1 = rel(g.fft(z, -1), np.fft.fftn(z))
    v2 = rel(g.fft(z, +1), np.fft.ifftn(z) * z.size)
    o = O.Oracle(s, prm["xi"], prm["errortol"], rc=prm["rc"], mesh=mesh, P=prm["P"])
    grid = rng.normal(size=mesh)
    og = o.recip_grid(grid)
    v3 = rel(g.recip(grid), og)
    g32 = GpuCase(s, precision="fp32", **prm)
    v4 = rel(g32.recip(grid), og)
    _report(f"register fft {mesh} fwd", v1)
    _report(f"register fft {mesh} inv", v2)
    _report(f"register recip {mesh} fp64", v3)
    _report(f"register recip {mesh} fp32", v4)
    assert v1 < 1e-13 and v2 < 1e-13 and v3 < 1e-12 and v4 < 1e-5


@pytest.mark.parametrize("stretch", [(1.0, 1.0, 1.0), (1.0, 1.05, 1.02)])
def test_register_recip_box_shapes(stretch, require_gpu):
    """k-space step of the register x pass on a cubic box (S(k) from the radial m^2 table)
    and on a stretched box (S(k) evaluated per mode), both against the oracle."""
    import copy

    s = copy.deepcopy(n1k_system())
    s.L = np.asarray(s.L, dtype=np.float64) * np.asarray(stretch)
    mesh = (64, 64, 64)
    prm = dict(xi=0.929, errortol=1e-6, rc=4.0, mesh=mesh, P=8)
    rng = np.random.default_rng(11)
    grid = rng.normal(size=mesh)
    o = O.Oracle(s, prm["xi"], prm["errortol"], rc=prm["rc"], mesh=mesh, P=prm["P"])
    og = o.recip_grid(grid)
    v64 = rel(GpuCase(s, **prm).recip(grid), og)
    v32 = rel(GpuCase(s, precision="fp32", **prm).recip(grid), og)
    _report(f"recip box {stretch} fp64", v64)
    _report(f"recip box {stretch} fp32", v32)
    assert v64 < 1e-12 and v32 < 1e-5
