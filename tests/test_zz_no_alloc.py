"""ed_lehmann_pairs_noalloc (fixed device scratch, no allocation) against
ed_lehmann_pairs through the C ABI: two calls in a row, the second smaller,
each bit-identical to the allocating entry point (same kernels, same
summation order); a size beyond the fixed scratch is refused."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _case(rng, n, m, nz):
    ov = torch.from_numpy(rng.standard_normal((m, n))).cuda()
    en = torch.from_numpy(np.sort(rng.uniform(-3, 3, n))).cuda()
    em = torch.from_numpy(np.sort(rng.uniform(-3, 3, m))).cuda()
    z = np.linspace(-5, 5, nz) + 0.1j
    zd = torch.from_numpy(np.ascontiguousarray(np.stack([z.real, z.imag], axis=1).ravel())).cuda()
    return ov, en, em, zd


def _run(L, name, ov, en, em, zd, nz, beta, emin, g0):
    g = g0.clone()
    L.call(name, L.dptr(ov), L.dptr(en), en.numel(), L.dptr(em), em.numel(), beta, emin, L.dptr(zd), nz,
           L.dptr(g), L.stream_ptr())
    torch.cuda.synchronize()
    return g.cpu().numpy()


def test_noalloc_pair_sum_matches_allocating_entry_point(gpu):
    from cmpy_amd import _lib as L

    rng = np.random.default_rng(12)
    for (n, m, nz, beta) in ((700, 900, 300, 7.0), (40, 25, 64, float("inf"))):
        ov, en, em, zd = _case(rng, n, m, nz)
        g0 = torch.from_numpy(rng.standard_normal(2 * nz)).cuda()  # both accumulate into g
        emin = float(min(en.min(), em.min()))
        ref = _run(L, "ed_lehmann_pairs", ov, en, em, zd, nz, beta, emin, g0)
        got = _run(L, "ed_lehmann_pairs_noalloc", ov, en, em, zd, nz, beta, emin, g0)
        assert np.array_equal(got, ref), (n, m, nz, np.abs(got - ref).max())


def test_noalloc_pair_sum_refuses_oversize(gpu):
    from cmpy_amd import _lib as L

    rng = np.random.default_rng(13)
    # 2 * ceil(1100 * 1100 / 8192) * 2000 + 2 = 592,002 > 524,288 doubles
    n = m = 1100
    nz = 2000
    ov, en, em, zd = _case(rng, n, m, nz)
    g = torch.zeros(2 * nz, dtype=torch.float64, device="cuda")
    with pytest.raises(L.EDError, match="scratch"):
        L.call("ed_lehmann_pairs_noalloc", L.dptr(ov), L.dptr(en), n, L.dptr(em), m, 1.0, 0.0, L.dptr(zd), nz,
               L.dptr(g), L.stream_ptr())
    # the allocating entry point takes the same size
    L.call("ed_lehmann_pairs", L.dptr(ov), L.dptr(en), n, L.dptr(em), m, 1.0, 0.0, L.dptr(zd), nz, L.dptr(g),
           L.stream_ptr())
    torch.cuda.synchronize()
    assert np.isfinite(g.cpu().numpy()).all()
