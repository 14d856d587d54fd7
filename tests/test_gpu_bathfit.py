"""Adjoint backward of the batched Anderson solve (bath-fit gradients).

Checked against the oracle's dense FP64 reference gradient (small baths) and
against central finite differences (h = 1e-5) of the batched forward at the
bath-fit workload's size (7 bath sites, D = 4,900), over the workload's
whole 256-model draw.  Tolerance: 1e-6 of the largest gradient entry of each
model.
"""
import numpy as np
import pytest

from oracle import ed_oracle as O

pytestmark = pytest.mark.gpu

GRAD_RTOL = 1e-6


def _loss_grads(b, wn, target):
    """Loss, dL/dRe G and dL/dIm G of sum |G - target|^2 at the batch's last solve."""
    res = b.solve(wn)
    diff = res.g - target
    return np.sum(np.abs(diff) ** 2, axis=1), 2 * diff.real, 2 * diff.imag, res


def _check(got, ref, what):
    for i in range(ref.shape[0]):
        scale = np.abs(ref[i]).max()
        assert np.abs(got[i] - ref[i]).max() <= GRAD_RTOL * scale, (what, i, got[i], ref[i])


@pytest.mark.parametrize("width_mode", ["reference", "fermion"])
def test_backward_matches_dense_reference(gpu, width_mode):
    from cmpy_amd.batch import AndersonBatch, matsubara

    rng = np.random.default_rng(17)
    M = 6
    nb = rng.integers(3, 6, size=M)
    u = np.full(M, 3.0)
    eb = [rng.uniform(-3, 3, n) for n in nb]
    v = [rng.uniform(0.1, 1.2, n) for n in nb]
    wn = matsubara(100.0, 48)
    b = AndersonBatch(u=u, eps_imp=-u / 2, eps_bath=eb, v=v, mu=0.0, keep_basis=True, tol=1e-12,
                      cf_tol=1e-14, sign_mode=width_mode)
    targets = []
    for i in range(M):
        n = (nb[i] + 1) // 2
        w = 0 if width_mode == "reference" else nb[i] + 1
        targets.append(O.siam_gf_t0(3.0, -1.5, rng.uniform(-3, 3, nb[i]), rng.uniform(0.1, 1.2, nb[i]), 0.0,
                                    n, n, wn, w))
    target = np.array(targets)
    loss, gr, gi, res = _loss_grads(b, wn, target)
    ge, gv, cg, _ = b.backward(gr, gi)
    ref_e = np.zeros_like(ge)
    ref_v = np.zeros_like(gv)
    for i in range(M):
        n = (nb[i] + 1) // 2
        w = 0 if width_mode == "reference" else nb[i] + 1
        L, de, dv = O.bathfit_grad_dense(3.0, -1.5, eb[i], v[i], 0.0, n, n, wn, target[i], w)
        assert loss[i] == pytest.approx(L, rel=1e-9)
        ref_e[i, :nb[i]] = de
        ref_v[i, :nb[i]] = dv
    _check(np.concatenate([ge, gv], axis=1), np.concatenate([ref_e, ref_v], axis=1), "grad")
    assert np.all(cg > 0)


def test_backward_matches_finite_differences_at_fit_size(gpu):
    """Seven bath sites, sector (4, 4), beta = 100 (the bath-fit workload's
    model size), four models; every parameter differenced in all models at once."""
    from cmpy_amd.batch import AndersonBatch, matsubara

    rng = np.random.default_rng(23)
    M, nb = 4, 7
    u = np.full(M, 3.0)
    eb = rng.uniform(-3, 3, (M, nb))
    v = rng.uniform(0.1, 1.2, (M, nb))
    wn = matsubara(100.0, 128)
    tb = AndersonBatch(u=u, eps_imp=-u / 2, eps_bath=list(rng.uniform(-3, 3, (M, nb))),
                       v=list(rng.uniform(0.1, 1.2, (M, nb))), mu=0.0, tol=1e-12, cf_tol=1e-14)
    target = tb.solve(wn).g
    b = AndersonBatch(u=u, eps_imp=-u / 2, eps_bath=list(eb), v=list(v), mu=0.0, keep_basis=True, tol=1e-12,
                      cf_tol=1e-14)
    loss, gr, gi, res = _loss_grads(b, wn, target)
    ge, gv, cg, _ = b.backward(gr, gi)
    h = 1e-5
    fd = np.zeros((M, 2 * nb))
    for kind in range(2):
        for k in range(nb):
            lp = []
            for sgn in (1, -1):
                e2, v2 = eb.copy(), v.copy()
                (e2 if kind == 0 else v2)[:, k] += sgn * h
                b.set_bath(e2, v2)
                lp.append(_loss_grads(b, wn, target)[0])
            fd[:, kind * nb + k] = (lp[0] - lp[1]) / (2 * h)
    _check(np.concatenate([ge, gv], axis=1), fd, "fd")


def _fd_grads(b, wn, target, eb, v, h=1e-5):
    """Central differences of every model's loss in every bath parameter,
    all models perturbed at once (the losses are independent)."""
    M, nb = eb.shape
    fd = np.zeros((M, 2 * nb))
    for kind in range(2):
        for k in range(nb):
            lp = []
            for sgn in (1, -1):
                e2, v2 = eb.copy(), v.copy()
                (e2 if kind == 0 else v2)[:, k] += sgn * h
                b.set_bath(e2, v2)
                lp.append(_loss_grads(b, wn, target)[0])
            fd[:, kind * nb + k] = (lp[0] - lp[1]) / (2 * h)
    b.set_bath(eb, v)
    return fd


def test_backward_long_continued_fractions(gpu, monkeypatch):
    """Two models of the bath-fit workload's draw (seed 101, indices 130 and
    242) whose particle fractions need 590 and 660 steps at beta = 100: past
    the 512 columns one wave group of the backward's Krylov product covers,
    and past the step cap these models used to stop at unconverged (then the
    fixed-basis adjoint differs from the derivative of the truncated forward,
    by 1.8e-6 relative for model 242).  Gradients against central differences."""
    from cmpy_amd.batch import AndersonBatch, matsubara

    ra = np.random.default_rng(101)
    ea, va = ra.uniform(-3, 3, (256, 7)), ra.uniform(0.1, 1.2, (256, 7))
    rb = np.random.default_rng(202)
    et, vt = rb.uniform(-3, 3, (256, 7)), rb.uniform(0.1, 1.2, (256, 7))
    sel = [130, 242]
    M = len(sel)
    eb, v = ea[sel], va[sel]
    u = np.full(M, 3.0)
    wn = matsubara(100.0, 512)
    tb = AndersonBatch(u=u, eps_imp=-u / 2, eps_bath=list(et[sel]), v=list(vt[sel]), mu=0.0)
    target = tb.solve(wn).g
    b = AndersonBatch(u=u, eps_imp=-u / 2, eps_bath=list(eb), v=list(v), mu=0.0, keep_basis=True, tol=1e-12)
    loss, gr, gi, res = _loss_grads(b, wn, target)
    assert res.cf_steps.max() > 512, res.cf_steps  # two wave groups of the Krylov product
    assert not res.cf_capped.any(), res.cf_steps
    ge, gv, cg, _ = b.backward(gr, gi)
    # these parts' K products run device-wide by default (>= 96 steps); the
    # per-model workgroup path gives the same gradient
    monkeypatch.setenv("CMPY_AMD_BWD_SPLIT_M", "100000")
    ge1, gv1, _, _ = b.backward(gr, gi)
    monkeypatch.delenv("CMPY_AMD_BWD_SPLIT_M")
    ref = np.abs(np.concatenate([ge1, gv1], axis=1)).max()
    assert np.abs(np.concatenate([ge - ge1, gv - gv1], axis=1)).max() <= 1e-10 * ref
    _check(np.concatenate([ge, gv], axis=1), _fd_grads(b, wn, target, eb, v), "fd")


def test_backward_whole_bathfit_draw(gpu):
    """The bath-fit workload itself: 256 models, 7 bath sites, sector (4, 4),
    U = 3, 512 Matsubara frequencies at beta = 100, theta from seed 101 and the
    target from seed 202.  Every continued fraction converges before the
    cap, and every model's adjoint gradient is within 1e-6 (of its largest
    entry) of central differences with h = 1e-5 (28 batched solves)."""
    from cmpy_amd.batch import AndersonBatch, matsubara

    M, nb = 256, 7
    ra, rb = np.random.default_rng(101), np.random.default_rng(202)
    u = np.full(M, 3.0)
    ea, va = ra.uniform(-3, 3, (M, nb)), ra.uniform(0.1, 1.2, (M, nb))
    et, vt = rb.uniform(-3, 3, (M, nb)), rb.uniform(0.1, 1.2, (M, nb))
    wn = matsubara(100.0, 512)
    target = AndersonBatch(u=u, eps_imp=-u / 2, eps_bath=list(et), v=list(vt), mu=0.0).solve(wn).g
    b = AndersonBatch(u=u, eps_imp=-u / 2, eps_bath=list(ea), v=list(va), mu=0.0, keep_basis=True, tol=1e-12)
    loss, gr, gi, res = _loss_grads(b, wn, target)
    assert not res.cf_capped.any()
    assert res.residual.max() <= 1e-10
    ge, gv, cg, _ = b.backward(gr, gi)
    _check(np.concatenate([ge, gv], axis=1), _fd_grads(b, wn, target, ea, va), "fd")


def test_capped_fraction_is_flagged(gpu):
    """A fraction stopped by the step cap is reported (cf_capped) and warned
    about; with enough steps the same model converges."""
    from cmpy_amd.batch import AndersonBatch, matsubara

    ra = np.random.default_rng(101)
    ea, va = ra.uniform(-3, 3, (256, 7)), ra.uniform(0.1, 1.2, (256, 7))
    wn = matsubara(100.0, 64)
    b = AndersonBatch(u=3.0, eps_imp=-1.5, eps_bath=[ea[242]], v=[va[242]], mu=0.0, cf_steps=256)
    with pytest.warns(RuntimeWarning, match="step cap"):
        r = b.solve(wn)
    assert r.cf_capped[0, 0] and r.cf_steps[0, 0] == 256
    r2 = AndersonBatch(u=3.0, eps_imp=-1.5, eps_bath=[ea[242]], v=[va[242]], mu=0.0).solve(wn)
    assert not r2.cf_capped.any() and 512 < r2.cf_steps[0, 0] < 1024


def test_autograd_function(gpu):
    """anderson_gf inside a torch loss gives the batch backward's gradients."""
    import torch

    from cmpy_amd.batch import AndersonBatch, anderson_gf, matsubara

    rng = np.random.default_rng(31)
    M, nb = 3, 4
    u = np.full(M, 2.0)
    eb = rng.uniform(-2, 2, (M, nb))
    v = rng.uniform(0.2, 1.0, (M, nb))
    wn = matsubara(50.0, 32)
    b = AndersonBatch(u=u, eps_imp=-u / 2, eps_bath=list(eb), v=list(v), mu=0.0, keep_basis=True, tol=1e-12,
                      cf_tol=1e-14)
    target = torch.from_numpy(b.solve(wn).g * 0.9)
    te = torch.tensor(eb, requires_grad=True)
    tv = torch.tensor(v, requires_grad=True)
    g = anderson_gf(te, tv, b, wn)
    loss = (g - target).abs().pow(2).sum()
    loss.backward()
    diff = g.detach().numpy() - target.numpy()
    ge, gv, _, _ = b.backward(2 * diff.real, 2 * diff.imag)
    np.testing.assert_allclose(te.grad.numpy(), ge, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(tv.grad.numpy(), gv, rtol=1e-12, atol=1e-14)
    for i in range(M):
        n = (nb + 1) // 2
        L, de, dv = O.bathfit_grad_dense(2.0, -1.0, eb[i], v[i], 0.0, n, n, wn, target.numpy()[i])
        assert np.abs(te.grad.numpy()[i] - de).max() <= GRAD_RTOL * np.abs(np.r_[de, dv]).max()


def test_fit_bath_reduces_the_loss(gpu):
    """L-BFGS on the adjoint gradients fits G(i w_n) of a known bath from a
    perturbed start.  Matsubara-axis bath fits are ill-conditioned (G at
    beta = 100 barely constrains some parameter combinations), so the test
    asks for the loss, not for the parameters: with each model's loss
    normalised by its starting value, every model's misfit drops by three
    orders of magnitude (unnormalised, the worst model sets the steps)."""
    from cmpy_amd.batch import AndersonBatch, fit_bath, matsubara

    rng = np.random.default_rng(7)
    M, nb = 8, 5
    u = np.full(M, 3.0)
    eb = np.sort(rng.uniform(-2.5, 2.5, (M, nb)), axis=1)
    v = rng.uniform(0.3, 1.0, (M, nb))
    wn = matsubara(100.0, 256)
    b = AndersonBatch(u=u, eps_imp=-u / 2, eps_bath=list(eb), v=list(v), mu=0.0, keep_basis=True, tol=1e-12,
                      cf_tol=1e-14)
    target = b.solve(wn).g
    e0 = eb + 0.03 * rng.standard_normal(eb.shape)
    v0 = v + 0.03 * rng.standard_normal(v.shape)
    g0 = AndersonBatch(u=u, eps_imp=-u / 2, eps_bath=list(e0), v=list(v0), mu=0.0).solve(wn).g
    loss0 = (np.abs(g0 - target) ** 2).sum(axis=1)
    fit = fit_bath(b, wn, target, e0, v0, max_iter=100)
    assert fit.history[-1] < 1e-3 * fit.history[0]
    assert np.all(fit.loss < 1e-3 * loss0)
    with pytest.raises(ValueError):
        fit_bath(AndersonBatch(u=u, eps_imp=-u / 2, eps_bath=list(eb), v=list(v), mu=0.0), wn, target)


def test_backward_errors(gpu):
    from cmpy_amd import _lib as L
    from cmpy_amd.batch import AndersonBatch, matsubara

    b = AndersonBatch(u=2.0, eps_imp=-1.0, eps_bath=[np.zeros(3)], v=[np.ones(3)])
    with pytest.raises(ValueError, match="keep_basis"):
        b.backward(np.zeros((1, 4)), np.zeros((1, 4)))
    bk = AndersonBatch(u=2.0, eps_imp=-1.0, eps_bath=[np.zeros(3)], v=[np.ones(3)], keep_basis=True)
    with pytest.raises(L.EDInvalidArgument, match="solve"):
        bk.backward(np.zeros((1, 4)), np.zeros((1, 4)))
    bk.solve(matsubara(10.0, 4))
    bk.set_bath([np.ones(3)], [np.ones(3)])
    with pytest.raises(L.EDInvalidArgument, match="solve"):
        bk.backward(np.zeros((1, 4)), np.zeros((1, 4)))
