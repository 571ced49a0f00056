"""GPU: the allocation-free neighbour-list build (cr_build_nlist_noalloc) against
cr_build_nlist through the C ABI.

Integer work, so the bar is bit-exact: the same totals, head_list, n_neigh and neighbour
arrays, on two calls in a row, the second with a smaller system (the static scratch keeps
the first call's contents, which the second must not see).  Inputs beyond the scratch
are rejected before any device work.  Device memory does not change across a call."""
import ctypes as C

import numpy as np
import pytest

from helpers import n1k_system, small_system

pytestmark = pytest.mark.gpu


def _buffers(torch, s):
    pos = torch.from_numpy(np.ascontiguousarray(s.pos, dtype=np.float64)).cuda()
    head = torch.full((s.N,), -1, dtype=torch.int32, device="cuda")
    nn = torch.full((s.N,), -1, dtype=torch.int32, device="cuda")
    nl = torch.full((s.N * 64,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    return pos, head, nn, nl


def _call(lib, ctx, torch, s, rlist, name, bufs):
    from conductorrigid_amd import _lib

    pos, head, nn, nl = bufs
    total = C.c_long(0)
    L = (C.c_double * 3)(*[float(v) for v in s.L])
    st = torch.cuda.current_stream().cuda_stream
    _lib.check(getattr(lib, name)(ctx, s.N, pos.data_ptr(), L, float(rlist), head.data_ptr(), nn.data_ptr(),
                                  nl.data_ptr(), nl.numel(), C.byref(total), st))
    torch.cuda.synchronize()
    return int(total.value)


def _build(lib, ctx, torch, s, rlist, name):
    bufs = _buffers(torch, s)
    total = _call(lib, ctx, torch, s, rlist, name, bufs)
    return total, bufs[1].cpu().numpy(), bufs[2].cpu().numpy(), bufs[3].cpu().numpy()


def test_noalloc_nlist_matches_allocating_build(require_gpu):
    import torch

    from conductorrigid_amd import _lib
    from conductorrigid_amd.compute import ConductorRigid

    lib = _lib.load()
    big, small = n1k_system(), small_system(seed=41)
    assert small.N < big.N
    cr = ConductorRigid(system=big, xi=0.929, errortol=1e-6, rc=4.0, mesh=(64, 64, 64), P=8, field=(0.0, 0.0, 1.0))
    for s in (big, small):  # two calls in a row, the second smaller
        rlist = 4.4
        t0, h0, n0, l0 = _build(lib, cr.ctx, torch, s, rlist, "cr_build_nlist")
        t1, h1, n1, l1 = _build(lib, cr.ctx, torch, s, rlist, "cr_build_nlist_noalloc")
        print(f"[noalloc] N={s.N} pairs {t0} / {t1}", flush=True)
        assert t0 == t1 and t0 > 0
        assert np.array_equal(h0, h1) and np.array_equal(n0, n1)
        assert np.array_equal(l0[:t0], l1[:t1])
        assert np.all(l1[t1:] == -1)  # nothing written past the list
    # no device memory is taken across a call (outputs allocated beforehand)
    bufs = _buffers(torch, big)
    free0 = torch.cuda.mem_get_info()[0]
    _call(lib, cr.ctx, torch, big, 4.4, "cr_build_nlist_noalloc", bufs)
    free1 = torch.cuda.mem_get_info()[0]
    print(f"[noalloc] device free memory {free0} -> {free1}", flush=True)
    assert free1 == free0


def test_noalloc_nlist_rejects_inputs_beyond_scratch(require_gpu):
    """The size check runs on the host before any device work: a bead count whose cell
    lists exceed the static scratch returns CR_ERR_INVALID with a message."""
    import torch

    from conductorrigid_amd import _lib
    from conductorrigid_amd.compute import ConductorRigid

    lib = _lib.load()
    s = small_system(seed=42)
    cr = ConductorRigid(system=s, xi=0.8, errortol=1e-5, rc=3.5, mesh=(30, 30, 30), P=7, field=(0.0, 0.0, 1.0))
    pos = torch.zeros(4, dtype=torch.float64, device="cuda")
    buf = torch.zeros(4, dtype=torch.int32, device="cuda")
    total = C.c_long(0)
    L = (C.c_double * 3)(400.0, 400.0, 400.0)
    code = lib.cr_build_nlist_noalloc(cr.ctx, 4_000_000, pos.data_ptr(), L, 4.4, buf.data_ptr(), buf.data_ptr(),
                                      buf.data_ptr(), 0, C.byref(total), torch.cuda.current_stream().cuda_stream)
    msg = lib.cr_last_error().decode()
    print(f"[noalloc] oversize: code {code}: {msg}", flush=True)
    assert code == _lib.CR_ERR_INVALID and "scratch" in msg
