"""The no-allocation C^H entry point equals the workspace entry point (GPU, through the C ABI).

pgj_effective_tensor_noalloc keeps its partial sums in the library's own device
array; two calls in a row, the second on a smaller grid, must each give the
bits of pgj_effective_tensor with a caller workspace (same kernels, same
summation order), and a grid beyond the array's limit is refused.
"""
import ctypes

import numpy as np
import pytest
import torch

from pgj import _lib

pytestmark = pytest.mark.gpu


def _inputs(shape, seed):
    rng = np.random.default_rng(seed)
    N = int(np.prod(shape))
    a = torch.tensor(1e-3 + rng.random(shape), dtype=torch.float64, device='cuda')
    u6 = torch.tensor(rng.standard_normal((6, 3 * N)), dtype=torch.float64, device='cuda')
    return a, u6


def _both(shape, seed):
    g = _lib.make_grid(shape, tuple(1.0 / v for v in shape[::-1]), 15 / 26, 5 / 13)
    a, u6 = _inputs(shape, seed)
    lib = _lib.lib()
    ws = torch.empty(lib.pgj_effective_tensor_ws_bytes(g) // 8 + 1, dtype=torch.float64, device='cuda')
    ref = torch.full((36,), np.nan, dtype=torch.float64, device='cuda')
    out = torch.full((36,), np.nan, dtype=torch.float64, device='cuda')
    _lib.call('pgj_effective_tensor', g, _lib.ptr(a), _lib.ptr(u6), _lib.ptr(ws), _lib.ptr(ref), _lib.stream())
    _lib.call('pgj_effective_tensor_noalloc', g, _lib.ptr(a), _lib.ptr(u6), _lib.ptr(out), _lib.stream())
    torch.cuda.synchronize()
    return ref.cpu().numpy(), out.cpu().numpy()


def test_noalloc_matches_workspace_entry_point_twice():
    print('no-alloc C^H: 32^3 then 12x20x24', flush=True)
    for shape, seed in (((32, 32, 32), 1), ((12, 20, 24), 2)):
        ref, out = _both(shape, seed)
        assert np.all(np.isfinite(ref))
        np.testing.assert_array_equal(out, ref)
        assert np.abs(ref.reshape(6, 6) - ref.reshape(6, 6).T).max() == 0


def test_noalloc_refuses_grids_beyond_its_scratch():
    lib = _lib.lib()
    limit = lib.pgj_effective_tensor_noalloc_max_elements()
    assert limit == 1 << 22
    a, u6 = _inputs((8, 8, 8), 3)
    out = torch.zeros(36, dtype=torch.float64, device='cuda')
    big = _lib.make_grid((256, 128, 129), (1 / 129, 1 / 128, 1 / 256), 15 / 26, 5 / 13)
    with pytest.raises(ValueError, match='limit'):
        _lib.check(lib.pgj_effective_tensor_noalloc(big, _lib.ptr(a), _lib.ptr(u6), _lib.ptr(out), _lib.stream()))
    torch.cuda.synchronize()
    assert torch.all(out == 0)
