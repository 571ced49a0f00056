"""OutputData's text format (cppmodule/ConductorRigid.cc:560-622) without a GPU: the
writer behind ConductorRigid.OutputData against the layout the reference writes --
"Bead Position" / "Bead Charge" / "Bead Force" sections, rows in particle-tag order
(global index rtag[tag]), numbers as an ostream with setprecision(10) prints them (the
C library's "%.10g", used here as the independent formatter), vector components joined
and followed by two spaces, and the file name prefix.%010d.txt."""
import ctypes as C
import os

import numpy as np

from conductorrigid_amd.compute import output_filename, write_output
from helpers import small_system

_libc = C.CDLL(None)


def _g10(v: float) -> str:
    buf = C.create_string_buffer(64)
    _libc.snprintf(buf, 64, b"%.10g", C.c_double(float(v)))
    return buf.value.decode()


def test_filename():
    assert output_filename("run/cr", 42) == "run/cr.0000000042.txt"
    assert output_filename("x", 0) == "x.0000000000.txt"


def test_sections_order_and_digits(tmp_path):
    s = small_system(seed=9, shuffle=True)
    assert not np.array_equal(s.tag, np.arange(s.N))  # permuted tags exercise the ordering
    rng = np.random.default_rng(1)
    q = rng.normal(size=s.N) * np.array([1e-7, 1.0, 123.456789012345, -3.0] * (s.N // 4 + 1))[:s.N]
    q[0] = -0.0
    F = np.zeros((s.N, 4))
    F[:, :3] = rng.normal(size=(s.N, 3)) * 1e3
    fn = os.path.join(tmp_path, output_filename("out", 7))
    write_output(fn, s.tag, s.pos, q, F)
    lines = open(fn).read().split("\n")
    assert lines[-1] == ""
    lines = lines[:-1]
    N = s.N
    assert len(lines) == 3 * N + 3
    assert lines[0] == "Bead Position" and lines[N + 1] == "Bead Charge" and lines[2 * N + 2] == "Bead Force"
    rtag = {int(t): i for i, t in enumerate(s.tag)}
    for t in range(N):
        i = rtag[t]
        p = s.pos[i]
        assert lines[1 + t] == f"{_g10(p[0])}  {_g10(p[1])}  {_g10(p[2])}  "
        assert lines[N + 2 + t] == _g10(q[i])
        assert lines[2 * N + 3 + t] == f"{_g10(F[i, 0])}  {_g10(F[i, 1])}  {_g10(F[i, 2])}  "
