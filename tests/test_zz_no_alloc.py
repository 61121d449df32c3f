"""spme_block_gram_noalloc: the block Gram product C = X Y^T with its partial
sums in the library's static device array instead of a caller workspace.

Checked through the C ABI against spme_block_gram (workspace form) on two
calls in a row, the second with a smaller input, FP32 and FP64:
  - where the workspace form's partial sums fit the array, the partition is
    the same and the results are bitwise equal;
  - the tall form at sizes whose partial sums do not fit uses fewer blocks
    along n; both forms accumulate in FP64, so they agree within 1e-13 of
    the largest |C| entry (about 500 ulps at these lengths);
  - sizes beyond the limits are refused with a message, and a refused call
    leaves the next one working.
"""
import numpy as np
import pytest

from spme_amd import lib as L

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _bind():
    return L.load()


def _gram_ws(h, prec, X, Y, n, torch):
    sx, sy = X.shape[0], Y.shape[0]
    need = h.spme_block_gram_worksize(n, sx, sy)
    work = torch.empty(max(1, need), dtype=torch.float64, device="cuda")
    C = torch.empty(sx * sy, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    L.check(h.spme_block_gram(prec, n, sx, sy, X.data_ptr(), X.stride(0), Y.data_ptr(),
                              Y.stride(0), C.data_ptr(), work.data_ptr(), need, s),
            "spme_block_gram")
    torch.cuda.synchronize()
    return C.cpu().numpy().reshape(sx, sy)


def _gram_na(h, prec, X, Y, n, torch):
    sx, sy = X.shape[0], Y.shape[0]
    C = torch.empty(sx * sy, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    L.check(h.spme_block_gram_noalloc(prec, n, sx, sy, X.data_ptr(), X.stride(0),
                                      Y.data_ptr(), Y.stride(0), C.data_ptr(), s),
            "spme_block_gram_noalloc")
    torch.cuda.synchronize()
    return C.cpu().numpy().reshape(sx, sy)


def _blocks(torch, prec, sx, sy, n, seed):
    dt = torch.float32 if prec == L.SPME_FP32 else torch.float64
    g = torch.Generator(device="cpu").manual_seed(seed)
    X = torch.randn(sx, n + 7, generator=g, dtype=torch.float64).to(dt).cuda()
    Y = torch.randn(sy, n + 3, generator=g, dtype=torch.float64).to(dt).cuda()
    return X, Y


@pytest.mark.parametrize("prec", [L.SPME_FP64, L.SPME_FP32])
def test_square_form_bitwise_two_calls(prec):
    torch = _torch()
    h = _bind()
    print(f"[no_alloc] square form prec={prec}", flush=True)
    for (sx, sy, n, seed) in [(16, 16, 300_000, 1), (3, 5, 4_321, 2)]:
        X, Y = _blocks(torch, prec, sx, sy, n, seed)
        a = _gram_ws(h, prec, X, Y, n, torch)
        b = _gram_na(h, prec, X, Y, n, torch)
        assert np.array_equal(a, b), (sx, sy, n)
        ref = X[:, :n].double().cpu().numpy() @ Y[:, :n].double().cpu().numpy().T
        assert np.abs(b - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("prec", [L.SPME_FP64, L.SPME_FP32])
def test_tall_form_two_calls(prec):
    torch = _torch()
    h = _bind()
    lim = h.spme_block_gram_noalloc_max()
    print(f"[no_alloc] tall form prec={prec}, scratch {lim} doubles", flush=True)
    # first: partial sums beyond the array (own partition); second, smaller:
    # they fit (the workspace form's partition, bitwise)
    for (sx, sy, n, seed, fits) in [(64, 16, 200_000, 3, False), (40, 8, 3_000, 4, True)]:
        assert (h.spme_block_gram_worksize(n, sx, sy) <= lim) == fits
        X, Y = _blocks(torch, prec, sx, sy, n, seed)
        a = _gram_ws(h, prec, X, Y, n, torch)
        b = _gram_na(h, prec, X, Y, n, torch)
        if fits:
            assert np.array_equal(a, b), (sx, sy, n)
        else:
            assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max(), (sx, sy, n)
        ref = X[:, :n].double().cpu().numpy() @ Y[:, :n].double().cpu().numpy().T
        assert np.abs(b - ref).max() <= 1e-12 * np.abs(ref).max()


def test_largest_block_and_errors():
    torch = _torch()
    h = _bind()
    X, Y = _blocks(torch, L.SPME_FP64, 4096, 16, 2_000, 5)
    a = _gram_ws(h, L.SPME_FP64, X, Y, 2_000, torch)
    b = _gram_na(h, L.SPME_FP64, X, Y, 2_000, torch)
    assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max()
    s = torch.cuda.current_stream().cuda_stream
    C = torch.empty(17 * 16, dtype=torch.float64, device="cuda")
    rc = h.spme_block_gram_noalloc(L.SPME_FP64, 100, 4, 17, X.data_ptr(), X.stride(0),
                                   Y.data_ptr(), Y.stride(0), C.data_ptr(), s)
    assert rc != 0 and "block widths" in L.last_error()
    rc = h.spme_block_gram_noalloc(L.SPME_FP64, 100, 4, 4, None, 100, Y.data_ptr(),
                                   Y.stride(0), C.data_ptr(), s)
    assert rc != 0 and "null" in L.last_error()
    # the array still serves the next call
    X2, Y2 = _blocks(torch, L.SPME_FP64, 2, 2, 1_000, 6)
    assert np.array_equal(_gram_ws(h, L.SPME_FP64, X2, Y2, 1_000, torch),
                          _gram_na(h, L.SPME_FP64, X2, Y2, 1_000, torch))
