"""ddh_solve_batch_noalloc (the solve on the library's own scratch, no allocation) against
ddh_solve_batch (the caller's workspace), both through the C-ABI: the same results on two calls
in a row, the second with a smaller batch, and a batch beyond the scratch refused."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(torch, lib, batch, V, opts, noalloc):
    from jaxdd_amd import _lib
    from jaxdd_amd.batch import _sc_now

    dev = batch["_pt"].device
    n, NP, nV = batch["n"], batch["NP"], len(V)
    sc = _sc_now(batch)
    Vd = torch.as_tensor(np.ascontiguousarray(V, dtype=np.float64), device=dev)
    J = torch.full((n, nV), np.nan, dtype=torch.float64, device=dev)
    ok = torch.zeros((n, nV), dtype=torch.uint8, device=dev)
    its = torch.zeros((n, nV), dtype=torch.int32, device=dev)
    status = torch.full((n,), -7, dtype=torch.int32, device=dev)
    b = _lib.Batch(n, NP, batch["nseg"], 0, batch["_pt"].data_ptr(), batch["_sid"].data_ptr(),
                   batch["_seg"].data_ptr(), sc.data_ptr())
    stream = torch.cuda.current_stream(dev).cuda_stream
    if noalloc:
        rc = lib.ddh_solve_batch_noalloc(ctypes.byref(b), Vd.data_ptr(), nV, ctypes.byref(opts),
                                         J.data_ptr(), ok.data_ptr(), its.data_ptr(),
                                         status.data_ptr(), stream)
        _lib.check(rc, "ddh_solve_batch_noalloc")
    else:
        nbytes = lib.ddh_workspace_bytes(n, NP, nV, opts.precision)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        rc = lib.ddh_solve_batch(ctypes.byref(b), Vd.data_ptr(), nV, ctypes.byref(opts),
                                 J.data_ptr(), ok.data_ptr(), its.data_ptr(), status.data_ptr(),
                                 ws.data_ptr(), nbytes, stream)
        _lib.check(rc, "ddh_solve_batch")
    torch.cuda.synchronize()
    return tuple(t.cpu().numpy() for t in (J, ok, its, status))


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_noalloc_matches_the_workspace_entry(torch_cuda, precision):
    from jaxdd_amd import _lib
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import prepare

    torch = torch_cuda
    lib = _lib.load()
    V = S.VOLTAGES[::3].copy()
    opts = _lib.Opts(1e-10, 60, 60, 0, 0, _lib.PRECISIONS[precision], 0, 2)
    big = prepare(S.devices("c3", 40, variant="free"), "cuda")
    small = prepare(S.devices("c2", 12), "cuda")
    print(f"noalloc {precision}: 40 c3 devices then 12 c2 devices, {len(V)} voltages", flush=True)
    for batch in (big, small):  # the second call smaller, on the scratch the first one left
        ref = _call(torch, lib, batch, V, opts, noalloc=False)
        got = _call(torch, lib, batch, V, opts, noalloc=True)
        assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2])
        assert np.array_equal(got[3], ref[3])
        assert np.all(np.isfinite(got[0]))
        assert np.array_equal(got[0], ref[0]), float(np.max(np.abs(got[0] - ref[0])))


def test_noalloc_refuses_a_batch_beyond_its_scratch(torch_cuda):
    from jaxdd_amd import _lib
    from jaxdd_amd import synthetic as S
    from jaxdd_amd.batch import from_packed, merge, pack

    torch = torch_cuda
    lib = _lib.load()
    limit = lib.ddh_noalloc_scratch_bytes()
    part = pack(S.devices("c2", 2))
    n = 2
    while lib.ddh_workspace_bytes(n, part["pt"].shape[0] - 2, 1, 1) <= 2 * limit:
        n *= 2
    batch = from_packed(merge([part] * (n // 2)), "cuda")
    opts = _lib.Opts(1e-10, 60, 60, 0, 0, _lib.DDH_MIXED, 0, 2)
    with pytest.raises(RuntimeError, match="scratch"):
        _call(torch, lib, batch, np.array([0.0]), opts, noalloc=True)
