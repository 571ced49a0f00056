"""Accuracy of the GPU force compute against the FP64 direct-Ewald oracle
(tests/golden/direct_ewald_n1000.npz: a seeded 1000-bead system at the N=1M
workload's method parameters, xi = 0.6, rc = 4, h = 0.89, P = 8), for the plain FP64
method and the FP32-bulk mode, and of FP32-bulk against FP64."""
import numpy as np
import pytest

from direct_fixture import load
from helpers import GpuCase, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fixture():
    return load()


def _solve(s, f, precision):
    c = GpuCase(s, float(f["xi"]), 1e-6, rc=float(f["rc"]), mesh=(27, 27, 27), P=8, precision=precision,
                restart=30, max_iter=3000, gmres_tol=1e-11, preconditioner="block")
    q, Phi, F, st = c.solve()
    assert st["converged"]
    return q, Phi, F[:, :3]


def test_error_vs_direct_ewald(fixture):
    s, f = fixture
    q64, P64, F64 = _solve(s, f, "fp64")
    q32, P32, F32 = _solve(s, f, "fp32")
    e64 = (rel(q64, f["q"]), rel(P64, f["Phi"]), rel(F64, f["F"]))
    e32 = (rel(q32, f["q"]), rel(P32, f["Phi"]), rel(F32, f["F"]))
    d = (rel(q32, q64), rel(P32, P64), rel(F32, F64))
    print(f"\n[direct-ewald N=1000] fp64 vs direct: q {e64[0]:.3e} Phi {e64[1]:.3e} F {e64[2]:.3e}")
    print(f"[direct-ewald N=1000] fp32 vs direct: q {e32[0]:.3e} Phi {e32[1]:.3e} F {e32[2]:.3e}")
    print(f"[direct-ewald N=1000] fp32 vs fp64:   q {d[0]:.3e} Phi {d[1]:.3e} F {d[2]:.3e}")
    # discretisation error of the method at these parameters (the same bound pins the
    # CPU oracle's spectral Ewald against the fixture)
    assert e64[0] < 2e-4 and e64[1] < 2e-4 and e64[2] < 1e-3
    # FP32 bulk adds nothing visible on top of it ...
    assert abs(e32[2] - e64[2]) < 0.02 * e64[2] and abs(e32[0] - e64[0]) < 0.02 * e64[0]
    # ... and stays within 1e-5 of the FP64 method
    assert d[0] < 1e-5 and d[2] < 1e-5


@pytest.mark.parametrize("precision,mesh", [("fp64", (60, 64, 27)), ("fp64", (64, 64, 64)), ("fp32", (64, 64, 64)),
                                            ("fp32", (48, 50, 45))])
def test_fft_cross_check_rocfft(fixture, precision, mesh):
    """The hand-written mixed-radix FFT against rocFFT (through torch.fft on the GPU):
    unnormalised forward e^{-ikx} and inverse e^{+ikx}."""
    import torch
    s, f = fixture
    c = GpuCase(s, float(f["xi"]), 1e-6, rc=float(f["rc"]), mesh=mesh, P=8, precision=precision)
    rng = np.random.default_rng(sum(mesh))
    z = rng.normal(size=mesh) + 1j * rng.normal(size=mesh)
    zt = torch.from_numpy(z).to("cuda")
    ref_f = torch.fft.fftn(zt).cpu().numpy()
    ref_i = (torch.fft.ifftn(zt) * float(np.prod(mesh))).cpu().numpy()
    tol = 1e-12 if precision == "fp64" else 5e-6
    assert rel(c.fft(z, -1), ref_f) < tol
    assert rel(c.fft(z, +1), ref_i) < tol
