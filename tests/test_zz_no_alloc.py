"""ccnn_stage1_noalloc: stage 1 of one level on device buffers with its intermediate data
in the library's fixed device array, against ccnn_stage1 (which allocates per call) on
the same levels. Bit-exact, two calls in a row with the second on a smaller level (the
array keeps the first call's contents), and the size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_stage1_noalloc_matches_stage1():
    import torch
    from compactcnncascade_amd import api
    L = api.load()
    L.ccnn_stage1_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    rng = np.random.default_rng(31)
    for lh, lw in [(729, 1296), (211, 333)]:
        lvl = (rng.random((lh, lw)) * 255).astype(np.float32)
        exp = api.stage1(lvl)
        d_lvl = torch.from_numpy(lvl).to("cuda:0")
        d_out = torch.full(exp.shape, float("nan"), dtype=torch.float32, device="cuda:0")
        torch.cuda.synchronize()
        print("noalloc level", lw, lh, flush=True)
        assert L.ccnn_stage1_noalloc(d_lvl.data_ptr(), lw, lh, d_out.data_ptr()) == 0
        assert np.array_equal(bits(d_out.cpu().numpy()), bits(exp)), (lw, lh)
    # a level whose L1-sum planes exceed the fixed array is refused
    big = torch.zeros((3100, 5600), dtype=torch.float32, device="cuda:0")
    out = torch.zeros((800, 1400), dtype=torch.float32, device="cuda:0")
    assert L.ccnn_stage1_noalloc(big.data_ptr(), 5600, 3100, out.data_ptr()) == -1
    assert L.ccnn_stage1_noalloc(0, 100, 100, out.data_ptr()) == -1
