"""hps_multi_dot_noalloc (the entry point that allocates nothing) against
hps_multi_dot, through the C ABI: two calls in a row on one GPU, the second
with a smaller input, each bitwise equal to the existing entry point's
results (same kernels and reduction order); sizes beyond the scratch array
are refused before any GPU work."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from hpsint_amd import CartesianMesh, SinteringOperatorGeneric  # noqa: E402
from hpsint_amd import _ffi  # noqa: E402


def P(t):
    return ctypes.c_void_p(t.data_ptr())


def test_multi_dot_noalloc_matches_multi_dot():
    L = _ffi.lib()
    op = SinteringOperatorGeneric(CartesianMesh((4, 4, 4), (0.0,) * 3, (4.0,) * 3), 1)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(21)
    for n, k in ((1_000_003, 27), (70_001, 5)):
        V = torch.from_numpy(rng.standard_normal(k * n)).cuda()
        w = torch.from_numpy(rng.standard_normal(n)).cuda()
        # stale values in the output that the call must overwrite
        h_dev = torch.full((k,), np.nan, dtype=torch.float64, device="cuda")
        _ffi.check(L.hps_multi_dot_noalloc(stream, n, k, P(V), P(w), P(h_dev)))
        ref = (ctypes.c_double * k)()
        _ffi.check(L.hps_multi_dot(op.handle, n, k, P(V), P(w), ref))
        torch.cuda.synchronize()
        got = h_dev.cpu().numpy()
        want = np.array(ref[:])
        assert np.array_equal(got, want), (n, k, got - want)
        exact = (V.view(k, n).cpu().numpy() @ w.cpu().numpy())
        assert np.allclose(got, exact, rtol=1e-12, atol=1e-9)
    # k beyond the scratch array's 32 rows: refused, nothing launched
    h_dev = torch.zeros(33, dtype=torch.float64, device="cuda")
    assert L.hps_multi_dot_noalloc(stream, 10, 33, P(h_dev), P(h_dev), P(h_dev)) == \
        _ffi.HPS_ERR_INVALID_ARGUMENT
    assert b"k must be 1..32" in L.hps_last_error()
