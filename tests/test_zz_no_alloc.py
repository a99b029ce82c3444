"""GPU test of the no-allocation entry point bex_mlp_predict_mfma_noalloc: through the C ABI,
on two calls in a row (the second with fewer rows and another network), its output equals
bex_mlp_predict_mfma's bit for bit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _params(seed):
    from bubbleensemble_implicitextension_amd.warm_start import init_params

    rng = np.random.default_rng(seed)
    p = init_params(seed).astype(np.float64) + rng.normal(0.0, 0.05, init_params(seed).size)
    return np.ascontiguousarray(p, dtype=np.float32)


def test_noalloc_predict_matches_allocating_entry_point():
    import torch
    from bubbleensemble_implicitextension_amd import _native
    from bubbleensemble_implicitextension_amd.warm_start import _ptr, _stream

    lib = _native.load()
    for n, seed in ((1000, 31), (77, 32)):
        print(f"case: predict_mfma_noalloc, n={n}, network seed {seed}", flush=True)
        rng = np.random.default_rng(seed)
        rows = np.zeros((n, 10), dtype=np.float32)
        rows[:, :5] = rng.normal(0.0, 1.0, (n, 5))
        params = _params(seed)
        pp = params.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        r = torch.from_numpy(rows).cuda()
        want = torch.zeros(n, dtype=torch.float32, device="cuda")
        got = torch.full((n,), float("nan"), dtype=torch.float32, device="cuda")
        _native.check(lib.bex_mlp_predict_mfma(_ptr(r), 10, n, pp, _ptr(want), _stream(torch)))
        _native.check(lib.bex_mlp_predict_mfma_noalloc(_ptr(r), 10, n, pp, _ptr(got),
                                                       _stream(torch)))
        torch.cuda.synchronize()
        a, b = want.cpu().numpy(), got.cpu().numpy()
        assert np.all(np.isfinite(a))
        np.testing.assert_array_equal(a.view(np.int32), b.view(np.int32))
    # Argument checks happen before any device work.
    assert lib.bex_mlp_predict_mfma_noalloc(None, 10, 5, pp, None, None) == -1
    assert lib.bex_mlp_predict_mfma_noalloc(_ptr(r), 4, 5, pp, _ptr(got), _stream(torch)) == -1
