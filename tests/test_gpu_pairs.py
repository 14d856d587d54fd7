"""GPU parity: the finite-T pair sum of _accumulate_sum
(cmpy/exactdiag.py:110-129) through the C ABI, against a numpy brute-force
sum, with and without Boltzmann screening.

Screening skips a pair only when both of its Boltzmann factors are below
1e-20; the skipped terms of one call then add up to at most
(N + M) 1e-20 / Im z for normalised overlaps.  Tolerances: 1e-11 absolute
against the brute-force sum (|G| up to 21 here; unscreened and screened),
1e-12 absolute between screened and unscreened G of whole models, 1e-13
relative at beta = inf, where nothing is screened.
"""

import numpy as np
import pytest
from numpy.testing import assert_allclose

pytestmark = pytest.mark.gpu


def _pairs(ov, en, em, beta, emin, z, ov2=None):
    import torch

    from cmpy_amd import _lib as L

    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    ovt, ent, emt = d(ov), d(en), d(em)
    zt = d(np.stack([z.real, z.imag], axis=1).reshape(-1))
    g = torch.zeros(2 * len(z), dtype=torch.float64, device="cuda")
    if ov2 is None:
        L.call("ed_lehmann_pairs", L.dptr(ovt), L.dptr(ent), len(en), L.dptr(emt), len(em), beta, emin,
               L.dptr(zt), len(z), L.dptr(g), L.stream_ptr())
    else:
        ov2t = d(ov2)
        L.call("ed_lehmann_pairs_ij", L.dptr(ovt), L.dptr(ov2t), L.dptr(ent), len(en), L.dptr(emt), len(em), beta,
               emin, L.dptr(zt), len(z), L.dptr(g), L.stream_ptr())
    gh = g.cpu().numpy().reshape(-1, 2)
    return gh[:, 0] + 1j * gh[:, 1]


def _brute(ov, en, em, beta, emin, z, ov2=None):
    w = ov * (ov if ov2 is None else ov2)
    if np.isfinite(beta):
        w = w * (np.exp(-beta * (en - emin))[None, :] + np.exp(-beta * (em - emin))[:, None])
    else:
        w = 2.0 * w  # both Boltzmann factors are 1 (cmpy/exactdiag.py:117-119)
    poles = em[:, None] - en[None, :]
    return np.array([np.sum(w / (zk - poles)) for zk in z])


def _case(seed, n, m, shuffle):
    rng = np.random.default_rng(seed)
    en = np.sort(rng.uniform(0.0, 12.0, n))
    em = np.sort(rng.uniform(0.5, 14.0, m))
    if shuffle:
        # significant states at the end of the list: the kept prefix must reach them
        en = en[::-1].copy()
        em = rng.permutation(em)
    ov = rng.standard_normal((m, n)) / np.sqrt(m)
    ov2 = rng.standard_normal((m, n)) / np.sqrt(m)
    z = np.linspace(-8, 8, 97) + 0.05j
    return en, em, ov, ov2, z


@pytest.mark.parametrize("screen", ["1", "0"])
@pytest.mark.parametrize("shuffle", [False, True])
@pytest.mark.parametrize("beta", [0.5, 5.0, 40.0])
def test_pairs_match_brute_force(gpu, monkeypatch, screen, shuffle, beta):
    monkeypatch.setenv("CMPY_AMD_PAIR_SCREEN", screen)
    en, em, ov, ov2, z = _case(3, 300, 211, shuffle)
    emin = min(en.min(), em.min())
    assert_allclose(_pairs(ov, en, em, beta, emin, z), _brute(ov, en, em, beta, emin, z), rtol=0, atol=1e-11)
    assert_allclose(_pairs(ov, en, em, beta, emin, z, ov2), _brute(ov, en, em, beta, emin, z, ov2), rtol=0,
                    atol=1e-11)


def test_pairs_unit_weights_at_infinite_beta(gpu):
    en, em, ov, _, z = _case(4, 120, 90, True)
    ref = _brute(ov, en, em, np.inf, 0.0, z)
    assert_allclose(_pairs(ov, en, em, np.inf, 0.0, z), ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def test_screening_leaves_gf_lehmann_unchanged(gpu, monkeypatch):
    """Finite-T G and G_ij of small models, screened and unscreened, at a
    beta where most pairs are screened out."""
    from cmpy_amd import HubbardModel, SingleImpurityAndersonModel, gf_lehmann, greens_lehmann
    from oracle import ed_oracle as O

    z = np.linspace(-6, 6, 301) + 0.05j
    siam = SingleImpurityAndersonModel(4.0, -2.0, [-1.5, -0.4, 0.3, 1.2], [0.6, 0.5, 0.7, 0.4], mu=0.0)
    ring = HubbardModel(6, O.chain_bonds(6, periodic=True), inter=4.0, eps=0.0, hop=1.0, mu=2.0)
    out = {}
    for screen in ("0", "1"):
        monkeypatch.setenv("CMPY_AMD_PAIR_SCREEN", screen)
        out[screen] = (gf_lehmann(siam, z, 30.0).gf, greens_lehmann(ring, z, 12.0, 0, 2).gf)
    for a, b in zip(out["0"], out["1"]):
        assert_allclose(b, a, rtol=0, atol=1e-12)
