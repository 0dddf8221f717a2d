"""The request server's queueing, batching and error paths on a host-only
context (no GPU): every request completes with the device error, requests
for the same geometry that wait together are merged into one batch."""
import numpy as np
import pytest

from cpanel_amd import abi, serve
from cpanel_amd.case import Context, CPanelError


def test_requests_complete_with_device_error_and_merge():
    ctx = Context(-1)
    S = serve.Server(ctx, slots=1)
    stream, geoms = serve.make_stream(10, 1e6, seed=2, n_lo=200, n_hi=400, reuse=0.6, max_cond=8)
    res, wall = serve.run_stream(S, stream, geoms)
    assert all(r["status"] == abi.CP_NO_DEVICE for r in res)
    st = S.stats()
    assert st["submitted"] == 10 and st["completed"] == 10
    assert st["conditions"] == sum(len(r["conds"]) for r in stream)
    # failed geometries are not kept
    assert st["resident_geometries"] == 0
    S.close()


def test_stream_is_seeded():
    a, ga = serve.make_stream(20, 5.0, seed=7, n_lo=300, n_hi=600)
    b, gb = serve.make_stream(20, 5.0, seed=7, n_lo=300, n_hi=600)
    assert [r["t"] for r in a] == [r["t"] for r in b]
    assert [r["key"] for r in a] == [r["key"] for r in b]
    assert all(np.array_equal(x[0]["xyz"], y[0]["xyz"]) for x, y in zip(ga, gb))
    nb = [int(np.sum(g[0]["ids"] < 1000)) for g in ga]
    assert min(nb) >= 150 and max(nb) <= 1200


def test_server_bad_args():
    ctx = Context(-1)
    L = abi.lib()
    import ctypes as C
    h = C.c_void_p()
    assert L.cp_server_create(ctx.h, 0, 0, C.byref(h)) == abi.CP_BAD_ARG
    S = serve.Server(ctx, slots=2)
    m = serve.request_geometry(np.random.default_rng(0), 200, 300)[0]
    with pytest.raises(CPanelError):
        S.submit(1, m, [(1.0, 0.0, 0.0)] * 65, serve.solve_opts())
    S.close()
