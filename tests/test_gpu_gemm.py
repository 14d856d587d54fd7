"""GPU parity: the overlap product C = A B^T behind the Lehmann sums
(the ``np.dot`` of cmpy/exactdiag.py:112), in every mode (VALU kernel,
FP64 matrix-core kernel, rocBLAS, and the size-based default) against numpy
on ragged shapes.

Tolerance: |C - C_ref| <= 1e-13 * sum_k |A_ik| |B_jk| (only the order of the
FP64 sums differs).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (7, 5, 3), (16, 16, 4), (17, 15, 5), (64, 64, 16), (65, 130, 17), (300, 400, 300),
          (1225, 35, 1225), (35, 1225, 1225), (1, 900, 2000), (513, 2, 64), (40, 30, 0)]


def _gemm(a, b):
    import torch

    from cmpy_amd import _lib as L

    at = torch.from_numpy(a).cuda()
    bt = torch.from_numpy(b).cuda()
    c = torch.full((a.shape[0], b.shape[0]), np.nan, dtype=torch.float64, device="cuda")
    L.call("ed_gemm_nt", L.dptr(at), L.dptr(bt), L.dptr(c), a.shape[0], b.shape[0], a.shape[1], L.stream_ptr())
    return c.cpu().numpy()


@pytest.mark.parametrize("mode", ["0", "1", "2", ""])
@pytest.mark.parametrize("shape", SHAPES)
def test_gemm_nt_matches_numpy(gpu, monkeypatch, mode, shape):
    monkeypatch.setenv("CMPY_AMD_GEMM", mode)
    m, n, k = shape
    rng = np.random.default_rng(m * 7919 + n * 31 + k)
    a = rng.standard_normal((m, k))
    b = rng.standard_normal((n, k))
    c = _gemm(a, b)
    ref = a @ b.T
    bound = 1e-13 * (np.abs(a) @ np.abs(b).T) + 1e-300
    assert np.all(np.isfinite(c))
    assert np.all(np.abs(c - ref) <= bound), float(np.max(np.abs(c - ref) / bound))


def test_gemm_nt_kernels_agree_on_eigenvector_overlaps(gpu, monkeypatch):
    """Orthogonal factors: V V^T = 1 for the eigenvectors of a symmetric
    matrix, in every mode."""
    rng = np.random.default_rng(5)
    h = rng.standard_normal((257, 257))
    _, v = np.linalg.eigh(h + h.T)
    vt = np.ascontiguousarray(v.T)
    for mode in ("0", "1", "2"):
        monkeypatch.setenv("CMPY_AMD_GEMM", mode)
        np.testing.assert_allclose(_gemm(vt, vt), np.eye(257), rtol=0, atol=1e-12)


def test_gemm_nt_rejects_unknown_mode(gpu, monkeypatch):
    from cmpy_amd import _lib as L

    monkeypatch.setenv("CMPY_AMD_GEMM", "3")
    with pytest.raises(L.EDInvalidArgument):
        _gemm(np.ones((2, 2)), np.ones((2, 2)))
