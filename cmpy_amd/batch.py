"""Batched Anderson impurity solves (ground state + G(i w_n)) in one launch.

No reference counterpart: cmpy solves one ``SingleImpurityAndersonModel`` per
call (cmpy/models/anderson.py:163-164).  ``AndersonBatch`` builds every model
exactly as that class does (cmpy/models/anderson.py:24-76,147-158):
site 0 is the impurity with level ``eps_imp - mu`` and interaction ``u``,
sites 1..nb are the bath levels ``eps_bath`` (not shifted), bonds (0, k) carry
``v_k``; ``mu`` defaults to ``u / 2`` (anderson.py:61).  For each model it
returns the ground-state energy of the (n_up, n_dn) sector and the impurity
Green's function on the given Matsubara frequencies,

    G(i w_n) = <0| c (i w_n + E0 - H)^-1 c+ |0> + <0| c+ (i w_n - E0 + H)^-1 c |0>,

from continued fractions whose Lanczos runs stop once G at the lowest
frequency has converged to ``cf_tol`` (at most ``cf_steps`` steps; the T = 0 limit of
cmpy/exactdiag.py:110-134, the same quantity as
``cmpy_amd.exactdiag.greens_cf`` for one model).
"""

from __future__ import annotations

import ctypes as C
import warnings
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _lib


def matsubara(beta: float, n: int) -> np.ndarray:
    """Fermionic Matsubara frequencies w_k = (2k + 1) pi / beta, k < n."""
    return (2 * np.arange(n) + 1) * np.pi / float(beta)


@dataclass
class BatchResult:
    e0: np.ndarray            # [M]
    g: np.ndarray             # [M, nw] complex
    iterations: np.ndarray    # [M] ground-state Lanczos steps
    ritz_residual: np.ndarray  # [M]
    residual: np.ndarray      # [M] explicit |H psi - E0 psi|
    elapsed_ms: float         # device time of the launch
    cf_steps: Optional[np.ndarray] = None  # [M, 2] particle / hole continued-fraction steps
    cf_capped: Optional[np.ndarray] = None  # [M, 2] True where a fraction stopped at the step cap unconverged


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class AndersonBatch:
    """Many Anderson impurity models of uneven bath size, solved together.

    Parameters are per model: ``u``, ``eps_imp``, ``mu`` scalars (or arrays of
    length M) and ``eps_bath`` / ``v`` as sequences of 1-d arrays (lengths may
    differ between models, 1..15).  ``n_up`` / ``n_dn`` default to half filling
    (num_bath + 1) // 2.  ``sign_mode`` is "reference" (all hopping signs +1,
    as the reference's SingleImpurityAndersonModel) or "fermion".
    """

    def __init__(self, u, eps_imp, eps_bath: Sequence, v: Sequence, mu=None, n_up=None, n_dn=None,
                 sign_mode: str = "reference", max_iter: int = 300, cf_steps: int = 1024, tol: float = 1e-10,
                 check_every: int = 8, seed: int = 0x5EED, cf_tol: float = 1e-13, keep_basis: bool = False):
        _lib.require_gpu()
        if sign_mode not in ("reference", "fermion"):
            raise ValueError("sign_mode must be 'reference' or 'fermion'")
        eps_bath = [np.atleast_1d(np.asarray(e, dtype=np.float64)) for e in eps_bath]
        v = [np.atleast_1d(np.asarray(x, dtype=np.float64)) for x in v]
        M = len(eps_bath)
        if M == 0 or len(v) != M:
            raise ValueError("eps_bath and v must list the same, non-zero number of models")
        nb = np.array([len(e) for e in eps_bath], dtype=np.int32)
        if any(len(x) != n for x, n in zip(v, nb)):
            raise ValueError("each model's v must have the length of its eps_bath")
        u = np.broadcast_to(np.asarray(u, dtype=np.float64), (M,)).copy()
        eps_imp = np.broadcast_to(np.asarray(eps_imp, dtype=np.float64), (M,)).copy()
        mu = u / 2 if mu is None else np.broadcast_to(np.asarray(mu, dtype=np.float64), (M,)).copy()
        n_up = (nb + 1) // 2 if n_up is None else np.broadcast_to(np.asarray(n_up, np.int32), (M,)).copy()
        n_dn = (nb + 1) // 2 if n_dn is None else np.broadcast_to(np.asarray(n_dn, np.int32), (M,)).copy()
        mb = int(max(nb.max(), 1))
        eb = np.zeros((M, mb))
        vv = np.zeros((M, mb))
        for i in range(M):
            eb[i, :nb[i]] = eps_bath[i]
            vv[i, :nb[i]] = v[i]
        self.num_models = M
        self.num_bath = nb
        self.n_up = np.ascontiguousarray(n_up, dtype=np.int32)
        self.n_dn = np.ascontiguousarray(n_dn, dtype=np.int32)
        self.u, self.eps_imp, self.mu = u, eps_imp, mu
        self.eps_bath, self.v = eb, vv
        self.sign_mode = sign_mode
        # keep the arrays alive while the library reads them
        self._keep = [nb, self.n_up, self.n_dn, np.ascontiguousarray(u), np.ascontiguousarray(eps_imp - mu),
                      np.ascontiguousarray(eb), np.ascontiguousarray(vv)]
        # fermion signs count the particles strictly between the impurity and
        # the bath site, so one width (the largest site count) is exact for
        # every model of the batch
        width = 0 if sign_mode == "reference" else int(nb.max()) + 1
        d = _lib.BatchDesc(M, _ptr(self._keep[0], C.c_int32), _ptr(self._keep[1], C.c_int32),
                           _ptr(self._keep[2], C.c_int32), _ptr(self._keep[3], C.c_double),
                           _ptr(self._keep[4], C.c_double), _ptr(self._keep[5], C.c_double),
                           _ptr(self._keep[6], C.c_double), mb, width, int(max_iter), int(cf_steps),
                           int(check_every), int(bool(keep_basis)), float(tol), int(seed) & 0xFFFFFFFFFFFFFFFF,
                           float(cf_tol))
        self.keep_basis = bool(keep_basis)
        self.cf_steps_cap = int(cf_steps)
        h = C.c_void_p()
        _lib.call("ed_batch_create", C.byref(h), C.byref(d), _lib.stream_ptr())
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().ed_batch_destroy(h)
            except Exception:  # pragma: no cover - interpreter shutdown
                pass
            self._h = None

    @property
    def dims(self) -> np.ndarray:
        """Ground-state sector size of every model."""
        from math import comb

        return np.array([comb(int(n) + 1, int(a)) * comb(int(n) + 1, int(b))
                         for n, a, b in zip(self.num_bath, self.n_up, self.n_dn)], dtype=np.int64)

    def config(self, wave_max_dim: int, grid_blocks: int, wave_blocks: int) -> None:
        """Fix the launch shape (otherwise the first solve autotunes it)."""
        _lib.call("ed_batch_config", self._h, int(wave_max_dim), int(grid_blocks), int(wave_blocks),
                  _lib.stream_ptr())

    def info(self) -> dict:
        wm, gb, wb, tm = C.c_int64(), C.c_int32(), C.c_int32(), C.c_double()
        _lib.call("ed_batch_info", self._h, C.byref(wm), C.byref(gb), C.byref(wb), C.byref(tm))
        return {"wave_max_dim": wm.value, "grid_blocks": gb.value, "wave_blocks": wb.value,
                "tune_ms": tm.value}

    def solve(self, wn) -> BatchResult:
        """Ground state and G(i w_n) of every model (wn: Matsubara frequencies)."""
        wn = np.ascontiguousarray(np.atleast_1d(wn), dtype=np.float64)
        M, nw = self.num_models, len(wn)
        e0 = np.empty(M)
        g = np.empty((M, nw, 2))
        it = np.empty((M, 3), dtype=np.int32)
        res = np.empty((M, 2))
        ms = C.c_double()
        _lib.call("ed_batch_solve", self._h, _ptr(wn, C.c_double), nw, _ptr(e0, C.c_double),
                  _ptr(g, C.c_double), _ptr(it, C.c_int32), _ptr(res, C.c_double), C.byref(ms),
                  _lib.stream_ptr())
        steps = it[:, 1:].copy()
        capped = steps >= self.cf_steps_cap
        if capped.any():
            warnings.warn(
                f"{int(capped.any(axis=1).sum())} of {M} models have a continued fraction stopped at the "
                f"{self.cf_steps_cap}-step cap before G(i w_0) converged to cf_tol; their G and gradients "
                "carry the truncation (BatchResult.cf_capped marks them; raise cf_steps)",
                RuntimeWarning, stacklevel=2)
        return BatchResult(e0, g[..., 0] + 1j * g[..., 1], it[:, 0].copy(), res[:, 0], res[:, 1], ms.value,
                           steps, capped)

    def set_bath(self, eps_bath, v) -> None:
        """Replace the bath parameters ([M, num_bath] arrays or per-model lists)."""
        eb = np.zeros_like(self.eps_bath)
        vv = np.zeros_like(self.v)
        for i in range(self.num_models):
            nb = self.num_bath[i]
            eb[i, :nb] = np.asarray(eps_bath[i], dtype=np.float64)[:nb]
            vv[i, :nb] = np.asarray(v[i], dtype=np.float64)[:nb]
        self.eps_bath, self.v = eb, vv
        _lib.call("ed_batch_set_bath", self._h, _ptr(eb, C.c_double), _ptr(vv, C.c_double), _lib.stream_ptr())

    def backward(self, dl_dre, dl_dim):
        """Adjoint backward of the last solve: (dL/d eps_bath, dL/d v, cg_iters,
        elapsed_ms) for a real loss with dL/dRe G, dL/dIm G of shape [M, nw]."""
        if not self.keep_basis:
            raise ValueError("create the batch with keep_basis=True to differentiate it")
        gr = np.ascontiguousarray(dl_dre, dtype=np.float64)
        gi = np.ascontiguousarray(dl_dim, dtype=np.float64)
        M, mb = self.eps_bath.shape
        ge = np.zeros((M, mb))
        gv = np.zeros((M, mb))
        cg = np.empty(M, dtype=np.int32)
        ms = C.c_double()
        _lib.call("ed_batch_backward", self._h, _ptr(gr, C.c_double), _ptr(gi, C.c_double), _ptr(ge, C.c_double),
                  _ptr(gv, C.c_double), _ptr(cg, C.c_int32), C.byref(ms), _lib.stream_ptr())
        for i in range(M):  # entries past a model's bath are not written
            ge[i, self.num_bath[i]:] = 0.0
            gv[i, self.num_bath[i]:] = 0.0
        return ge, gv, cg, ms.value


def random_batch(num_models: int, seed: int, nb_range=(5, 9), u_range=(1.0, 8.0), eps_range=(-3.0, 3.0),
                 v_range=(0.1, 1.2)):
    """Synthetic batch: nb ~ U{nb_range}, U ~ U[u_range], eps_imp = -U/2,
    eps_k ~ U[eps_range], V_k ~ U[v_range], mu = 0, half filling."""
    rng = np.random.default_rng(seed)
    nb = rng.integers(nb_range[0], nb_range[1] + 1, size=num_models)
    u = rng.uniform(*u_range, size=num_models)
    eps_bath = [rng.uniform(*eps_range, size=n) for n in nb]
    v = [rng.uniform(*v_range, size=n) for n in nb]
    return dict(u=u, eps_imp=-u / 2, eps_bath=eps_bath, v=v, mu=np.zeros(num_models))


class _AndersonG(__import__("torch").autograd.Function):
    """G(i w_n) of a batch as a differentiable function of (eps_bath, v)."""

    @staticmethod
    def forward(ctx, eps_bath, v, batch, wn):
        batch.set_bath(eps_bath.detach().cpu().numpy(), v.detach().cpu().numpy())
        res = batch.solve(wn)
        ctx.batch = batch
        import torch

        return torch.from_numpy(res.g).to(eps_bath.device)

    @staticmethod
    def backward(ctx, grad_g):
        # for a real loss, torch passes dL/dRe G + i dL/dIm G
        import torch

        gr = grad_g.real.detach().cpu().numpy()
        gi = grad_g.imag.detach().cpu().numpy()
        ge, gv, _, _ = ctx.batch.backward(gr, gi)
        dev = grad_g.device
        return torch.from_numpy(ge).to(dev), torch.from_numpy(gv).to(dev), None, None


def anderson_gf(eps_bath, v, batch: AndersonBatch, wn):
    """Differentiable G(i w_n) [M, nw] (complex128) of ``batch`` at bath
    parameters ``eps_bath``, ``v`` (torch float64 [M, max_bath]).  The batch
    must be created with ``keep_basis=True``; one forward/backward pair at a
    time per batch (the backward uses the batch's last solve)."""
    return _AndersonG.apply(eps_bath, v, batch, np.ascontiguousarray(wn, dtype=np.float64))


@dataclass
class BathFit:
    eps_bath: np.ndarray  # [M, max_bath] fitted parameters (padding entries unused)
    v: np.ndarray
    loss: np.ndarray      # [M] final per-model loss sum_n |G - G_target|^2
    history: list         # (weighted) total loss at every evaluation
    evaluations: int


def fit_bath(batch: AndersonBatch, wn, g_target, eps0=None, v0=None, max_iter: int = 100,
             tolerance_change: float = 1e-30, history_size: int = 20, normalize: bool = True) -> BathFit:
    """Gradient-driven bath fit of every model of ``batch``: minimise
    L = sum_models w_m sum_n |G(i w_n; eps_bath, v) - g_target|^2 with L-BFGS
    (strong Wolfe line search), each loss and gradient from one batched
    forward solve and one adjoint backward (``anderson_gf``).  The batch must
    be created with ``keep_basis=True``.  ``eps0`` / ``v0`` ([M, max_bath])
    default to the batch's current bath.  The models' losses are independent,
    so the joint minimisation fits each model; with ``normalize`` each is
    weighted by 1 / (its loss at the start), so one badly fitting model does
    not set the step for all.  G(i w_n) constrains some parameter
    combinations only weakly (the usual ill-conditioning of Matsubara-axis
    bath fits), so a small loss does not pin every parameter.

    The reference fits the bath of its DMFT loop with derivative-free scipy
    optimisers; this is the gradient route that the adjoint backward enables.
    """
    import torch

    if not batch.keep_basis:
        raise ValueError("fit_bath needs a batch created with keep_basis=True")
    wn = np.ascontiguousarray(wn, dtype=np.float64)
    gt = torch.as_tensor(np.asarray(g_target, dtype=np.complex128))
    if gt.shape != (batch.num_models, len(wn)):
        raise ValueError(f"g_target must have shape ({batch.num_models}, {len(wn)})")
    e = torch.tensor(batch.eps_bath if eps0 is None else np.asarray(eps0, dtype=np.float64), requires_grad=True)
    v = torch.tensor(batch.v if v0 is None else np.asarray(v0, dtype=np.float64), requires_grad=True)
    weight = torch.ones(batch.num_models, dtype=torch.float64)
    if normalize:
        with torch.no_grad():
            l0 = (anderson_gf(e.detach(), v.detach(), batch, wn) - gt).abs().pow(2).sum(dim=1)
        weight = 1.0 / torch.clamp(l0, min=1e-300)
    opt = torch.optim.LBFGS([e, v], lr=1.0, max_iter=max_iter, max_eval=2 * max_iter,
                            tolerance_grad=0.0, tolerance_change=tolerance_change, history_size=history_size,
                            line_search_fn="strong_wolfe")
    hist = []

    def closure():
        opt.zero_grad()
        g = anderson_gf(e, v, batch, wn)
        loss = ((g - gt).abs().pow(2).sum(dim=1) * weight).sum()
        loss.backward()
        hist.append(float(loss.detach()))
        return loss

    opt.step(closure)
    with torch.no_grad():
        g = anderson_gf(e.detach(), v.detach(), batch, wn)
        per_model = (g - gt).abs().pow(2).sum(dim=1).numpy()
    return BathFit(e.detach().numpy().copy(), v.detach().numpy().copy(), per_model, hist, len(hist))
