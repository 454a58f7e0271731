"""The allocation-free forward (adt_eikonal3d_fwd_noalloc_f64) against the
allocating one (adt_eikonal3d_fwd_f64), through the C ABI: travel times bit
for bit and the same diagnostics, on two calls in a row (the second on a
smaller event, so stale contents of the shared workspace would show), and
the size limit of the fixed workspace."""
import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path[:0] = [ROOT, os.path.join(ROOT, "adtomo.jl_amd")]

from oracle import oracle  # noqa: E402
from adtomo_hip import synth  # noqa: E402

pytestmark = pytest.mark.gpu


def _event(shape, seed):
    f = 1.0 / synth.layered_grf_model(shape, 0.5, seed=seed)
    u0 = synth.point_source_u0(shape, synth.interior_source(shape, seed=seed + 1))
    return f, u0


def _run(fn, f, u0, shape, tol):
    import torch
    from adtomo_hip import _lib as L
    fd = torch.from_numpy(f).cuda()
    u0d = torch.from_numpy(u0).cuda()
    u = torch.full_like(fd, -7.0)
    info = L.AdtInfo()
    stream = torch.cuda.current_stream().cuda_stream
    if fn == "alloc":
        rc = L.lib().adt_eikonal3d_fwd_f64(u0d.data_ptr(), fd.data_ptr(), 0.5, *shape, tol, 0,
                                           u.data_ptr(), ctypes.byref(info), 0, stream)
    else:
        rc = L.lib().adt_eikonal3d_fwd_noalloc_f64(u0d.data_ptr(), fd.data_ptr(), 0.5, *shape, tol,
                                                   0, u.data_ptr(), ctypes.byref(info), stream)
    assert rc == 0, L.lib().adt_last_error()
    torch.cuda.synchronize()
    return u.cpu().numpy(), info.as_dict()


def test_noalloc_forward_matches_allocating_forward():
    for k, (shape, seed) in enumerate((((40, 36, 30), 61), ((20, 18, 12), 62))):
        print(f"call {k}: event {shape}", flush=True)
        f, u0 = _event(shape, seed)
        u_a, info_a = _run("alloc", f, u0, shape, 1e-6)
        u_n, info_n = _run("noalloc", f, u0, shape, 1e-6)
        assert np.array_equal(u_n.view(np.uint64), u_a.view(np.uint64)), shape
        assert info_n["iterations"] == info_a["iterations"]
        assert info_n["converged"] == info_a["converged"]
        assert info_n["err_log"] == info_a["err_log"]
        u_ref, it, _ = oracle.eikonal3d(u0, f, 0.5, 1e-6)
        assert np.array_equal(u_n.view(np.uint64), u_ref.view(np.uint64)) and it == info_n["iterations"]


def test_noalloc_forward_rejects_events_beyond_the_workspace():
    from adtomo_hip import _lib as L
    lib = L.lib()
    buf = np.zeros(8)
    p = buf.ctypes.data
    # 129 x 128 x 128 = 2^21 + 16384 nodes: past the 2^21-node snapshot
    rc = lib.adt_eikonal3d_fwd_noalloc_f64(p, p, 0.5, 129, 128, 128, 1e-3, 0, p, None, None)
    assert rc == -1 and b"limit" in lib.adt_last_error()
    assert lib.adt_eikonal3d_fwd_noalloc_f64(None, p, 0.5, 4, 4, 4, 1e-3, 0, p, None, None) == -1
