"""mpf_ctx_stream_probe: the streaming reference kernels compute what they claim (copy and
y = x + 0.5 y) on a ragged size, and report a positive rate."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("pattern,loads", [(0, 1), (0, 4), (1, 2)])
def test_stream_probe_results(gpu_ctx, pattern, loads):
    import torch
    from mpflow_amd import capi
    from mpflow_amd.capi import lib
    n = 4 * 100003
    x = torch.rand(n, device="cuda")
    y0 = torch.rand(n, device="cuda")
    y = y0.clone()
    g = C.c_double()
    reps = 3
    print("case: stream probe pattern %d loads %d" % (pattern, loads), flush=True)
    capi.check(lib().mpf_ctx_stream_probe(gpu_ctx.h, pattern, loads, C.c_void_p(x.data_ptr()),
                                          C.c_void_p(y.data_ptr()), n, 37, reps, C.byref(g)))
    assert g.value > 0.0
    want = x.clone()
    if pattern == 1:   # reps + 1 launches of y = x + 0.5 y
        want = y0.clone()
        for _ in range(reps + 1):
            want = x + 0.5 * want
    np.testing.assert_allclose(y.cpu().numpy(), want.cpu().numpy(), rtol=1e-6, atol=1e-6)
