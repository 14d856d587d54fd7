"""DMFT caller of the impurity solver (cmpy/dmft/utils.py, cmpy/dmft/twosite.py).

The self-consistency arithmetic here is elementwise on arrays of frequencies
(as in the reference); the impurity Green's function it consumes comes from
the device solvers (``gf_lehmann`` at T > 0, ``greens_lehmann_t0`` at T = 0)
unless the two-site closed form is asked for (``ref=True``).  Plotting helpers
of the reference are not part of this package.
"""

from __future__ import annotations

from typing import List

import numpy as np

from .models import SingleImpurityAndersonModel

__all__ = [
    "IterationStats", "mix_values", "self_energy", "bethe_gf_omega", "quasiparticle_weight",
    "impurity_params_ref", "impurity_gf_ref", "impurity_gf0", "compute_impurity_gf", "compute_self_energy",
    "twosite_dmft_half_filling", "compute_lattice_greens_function", "bethe_gf_matsubara", "hybridization",
    "BetheDMFT", "fit_hybridization", "bethe_dmft_bath_fit",
]


class IterationStats:
    """Error history and convergence status of a self-consistency loop
    (cmpy/dmft/utils.py:68-148, without the plot method)."""

    def __init__(self, *labels):
        self._errors: List[list] = []
        self.labels = labels
        self.success = False
        self.status = ""

    @property
    def iteration(self) -> int:
        return len(self._errors)

    @property
    def errors(self):
        return self._errors[-1]

    @property
    def history(self) -> np.ndarray:
        return np.asarray(self._errors).T

    @property
    def message(self) -> str:
        msg = "Self-consistency reached" if self.success else "Self-consistency not reached"
        return msg + ": " + self.status

    def append(self, *items) -> None:
        self._errors.append(list(items))

    def set_status(self, success, status="") -> None:
        self.success = success
        self.status = status

    def set_maxiter_status(self, maxiter) -> None:
        self.success = False
        self.status = f"Maximum iteration of {maxiter} reached without converging!"

    def set_parameter_converged(self, name, value, dec=4) -> None:
        self.success = True
        self.status = f"{name} converged to {float(np.squeeze(value)):.{dec}f}"

    def __bool__(self) -> bool:
        return self.success

    def __len__(self) -> int:
        return len(self._errors)

    def __iter__(self):
        yield from np.atleast_1d(self.errors)

    def __getitem__(self, i: int) -> List[float]:
        return self._errors[i]

    def errorstr(self, frmt=".2e") -> str:
        return ", ".join(f"{label}={err:{frmt}}" for label, err in zip(self.labels, self.errors))

    def __str__(self) -> str:
        return "\n".join([f"Success:    {self.success}", f"Iterations: {self.iteration}",
                          "Errors:     " + (self.errorstr(".1e") if self._errors else ""),
                          f"Status:     {self.status}"])


def mix_values(old, new, mixing=1.0):
    """new * mixing + old * (1 - mixing); ``mixing=1`` returns ``new``
    (cmpy/dmft/utils.py:163-183)."""
    if mixing == 1:
        return new
    if not 0 < mixing < 1:
        raise ValueError("mixing must lie in (0, 1]")
    return new * mixing + old * (1.0 - mixing)


def self_energy(gf_imp0, gf_imp):
    """Sigma(z) = G0(z)^-1 - G(z)^-1 (cmpy/dmft/utils.py:186-207)."""
    return 1 / gf_imp0 - 1 / gf_imp


def bethe_gf_omega(z, t):
    """Local Green's function of the Bethe lattice with infinite coordination,
    half bandwidth 2t (cmpy/dmft/utils.py:210-226)."""
    z_rel = z / (2 * t)
    return z_rel * (1 - np.sqrt(1 - 1 / (z_rel * z_rel))) / t


def quasiparticle_weight(omegas, sigma, thresh=1e-5):
    """z_qp = 1 / (1 - dRe Sigma/d omega at omega = 0), the slope from a linear
    fit over the points within one grid step of 0; 0 below ``thresh``
    (cmpy/dmft/utils.py:229-280)."""
    omegas = np.asarray(omegas, dtype=np.float64)
    dw = omegas[1] - omegas[0]
    win = (-dw <= omegas) & (omegas <= +dw)
    try:
        dsigma = np.polyfit(omegas[win], np.real(sigma)[win], 1)[0]
    except np.linalg.LinAlgError:
        ow, sw = omegas[win], np.real(sigma)[win]
        dsigma = (sw[-1] - sw[0]) / (ow[-1] - ow[0])
    z_qp = 1 / (1 - dsigma)
    if z_qp < thresh:
        z_qp = 0
    return z_qp


# -- two-site DMFT (cmpy/dmft/twosite.py) -------------------------------------------


def impurity_params_ref(u: float, v: float):
    """Residues and poles (a1, a2, e1, e2) of the two-site impurity G at half
    filling and T = 0 (E. Lange; cmpy/dmft/twosite.py:25-60)."""
    s16 = np.sqrt(u**2 + 16 * v**2)
    s64 = np.sqrt(u**2 + 64 * v**2)
    a1 = 0.25 * (1 - (u**2 - 32 * v**2) / np.sqrt((u**2 + 64 * v**2) * (u**2 + 16 * v**2)))
    return float(a1), float(0.5 - a1), float(0.25 * (s64 - s16)), float(0.25 * (s64 + s16))


def impurity_gf_ref(z, u: float, v: float):
    """Closed-form two-site impurity G (cmpy/dmft/twosite.py:63-81)."""
    a1, a2, e1, e2 = impurity_params_ref(u, v)
    return (a1 / (z - e1) + a1 / (z + e1)) + (a2 / (z - e2) + a2 / (z + e2))


def impurity_gf0(z, siam):
    """cmpy/dmft/twosite.py:87-88."""
    return siam.impurity_gf0(z)


def compute_impurity_gf(z, siam, ref=False):
    """Interacting impurity G (cmpy/dmft/twosite.py:91-94): the closed form with
    ``ref=True``; otherwise the device solvers -- the grand-canonical Lehmann
    sum at T > 0 (``siam.impurity_gf``), and at T = 0 the Lehmann sum of the
    ground-state sector (the reference's finite-T route has no T = 0 limit:
    it divides by the temperature)."""
    if ref:
        return impurity_gf_ref(z, siam.u, float(np.atleast_1d(siam.v)[0]))
    if siam.temp > 0:
        return siam.impurity_gf(z)
    from .exactdiag import compute_groundstate, greens_lehmann_t0

    gs = compute_groundstate(siam)
    return greens_lehmann_t0(siam, z, gs.n_up, gs.n_dn, pos=0, gs=gs)


def compute_self_energy(z, siam, ref=False):
    """cmpy/dmft/twosite.py:97-100."""
    return self_energy(impurity_gf0(z, siam), compute_impurity_gf(z, siam, ref=ref))


def twosite_dmft_half_filling(z, u, t=1.0, beta=np.inf, mixing=1.0, vtol=1e-6, max_iter=1000, vthresh=1e-10,
                              verbose=True, ref=True):
    """Two-site DMFT of the half-filled Hubbard model on the Bethe lattice
    (M. Potthoff, cond-mat/0107502; cmpy/dmft/twosite.py:103-172): iterate
    V <- sqrt(z_qp m2), m2 = t^2, with z_qp the quasiparticle weight of the
    impurity self-energy on the real axis ``z``.  Returns the converged
    ``SingleImpurityAndersonModel``; its ``stats`` attribute holds the
    iteration history."""
    siam = SingleImpurityAndersonModel(u, 0.0, 0.0, [t], mu=u / 2, temp=1 / beta)
    v = siam.v[0] + 0.1  # differs from the start value, so the first error is not zero
    m2 = t**2
    it = 0
    stats = IterationStats("Δv")
    while True:
        siam.update_hybridization(v)
        sigma = self_energy(impurity_gf0(z, siam), compute_impurity_gf(z, siam, ref=ref))
        qp_weight = quasiparticle_weight(z.real, sigma, thresh=vthresh)
        v_new = mix_values(v, np.sqrt(qp_weight * m2), mixing=mixing)
        delta_v = float(np.linalg.norm(v - v_new))
        stats.append(delta_v)
        v = v_new
        if v == 0:
            stats.set_parameter_converged("Hybridization", v)
            break
        if delta_v < vtol:
            stats.set_parameter_converged("Hybridization", v)
            break
        elif qp_weight == 0:
            stats.set_parameter_converged("Quasiparticle weight", qp_weight)
            break
        elif it >= max_iter:
            stats.set_maxiter_status(max_iter)
            break
        it += 1
    siam.update_hybridization(v)
    object.__setattr__(siam, "stats", stats)
    if verbose:
        print("-" * 50)
        print(f"U:          {u:.2f}")
        print(stats)
    return siam


def compute_lattice_greens_function(z, siam, t, ref=False):
    """G_lat(z) = G_Bethe(z + mu - Sigma(z)) (cmpy/dmft/twosite.py:175-178)."""
    sigma = compute_self_energy(z, siam, ref)
    return bethe_gf_omega(z + siam.mu - sigma, t)


# -- Matsubara-axis DMFT with a fitted exact-diagonalisation bath ---------------------


def bethe_gf_matsubara(zeta, t):
    """Local Green's function of the Bethe lattice (semicircular density of
    states of half bandwidth 2t) at complex argument zeta off the real axis:
    (zeta - sqrt(zeta - 2t) sqrt(zeta + 2t)) / (2 t^2).  The product of two
    principal roots keeps the branch with G ~ 1/zeta at infinity everywhere
    off the cut [-2t, 2t]; on the real axis plus i eta it equals
    ``bethe_gf_omega`` (cmpy/dmft/utils.py:210-226)."""
    zeta = np.asarray(zeta, dtype=np.complex128)
    return (zeta - np.sqrt(zeta - 2 * t) * np.sqrt(zeta + 2 * t)) / (2 * t * t)


def hybridization(z, eps_bath, v):
    """Delta(z) = sum_k |V_k|^2 / (z - eps_k) (cmpy/models/anderson.py:127-145);
    eps_bath / v of shape [..., nb] broadcast against z of shape [nz]."""
    z = np.asarray(z)[..., None]
    eps_bath = np.asarray(eps_bath, dtype=np.float64)[..., None, :]
    v = np.asarray(v, dtype=np.float64)[..., None, :]
    return np.sum(v * v / (z - eps_bath), axis=-1)


class BetheDMFT:
    """Result of ``bethe_dmft_bath_fit``: per value of U the converged bath,
    impurity and lattice Green's functions on the Matsubara axis, the
    self-energy, and the iteration history."""

    def __init__(self, u, wn, eps_bath, v, g_imp, g_loc, sigma, stats):
        self.u, self.wn = u, wn
        self.eps_bath, self.v = eps_bath, v
        self.g_imp, self.g_loc, self.sigma = g_imp, g_loc, sigma
        self.stats = stats

    @property
    def residual(self):
        """max_n |G_imp - G_loc| per model: how well the finite bath meets
        the lattice self-consistency."""
        return np.abs(self.g_imp - self.g_loc).max(axis=1)

    @property
    def quasiparticle_weight(self):
        """Z ~ 1 / (1 - Im Sigma(i w_0) / w_0), the lowest-frequency estimate."""
        return 1.0 / (1.0 - self.sigma[:, 0].imag / self.wn[0])


def _hyb_loss_and_grad(x, iw, target, shape):
    """L = sum |Delta(i w_n; eps, V) - target|^2 and dL/d(eps, V) for the
    flattened parameters x = (eps, V), each of ``shape`` [M, nb]."""
    n = int(np.prod(shape))
    e, v = x[:n].reshape(shape), x[n:].reshape(shape)
    r = 1.0 / (iw[None, :, None] - e[:, None, :])          # [M, nw, nb]
    res = (v * v)[:, None, :] * r
    diff = res.sum(axis=-1) - target                      # [M, nw]
    c = np.conj(diff)[..., None]
    ge = 2.0 * np.real(c * res * r).sum(axis=1)           # d|diff|^2 / d eps_k
    gv = 4.0 * np.real(c * v[:, None, :] * r).sum(axis=1)  # d|diff|^2 / d V_k
    return float(np.sum(np.abs(diff) ** 2)), np.concatenate([ge.ravel(), gv.ravel()])


def fit_hybridization(wn, delta_target, eps0, v0, max_iter=200):
    """Bath parameters with Delta(i w_n) = sum_k V_k^2 / (i w_n - eps_k) closest to
    ``delta_target`` ([M, nw]) in sum_n |.|^2, per model: L-BFGS-B (scipy) on the
    closed-form Delta and its analytic gradient, no impurity solve.  Returns
    (eps [M, nb], v [M, nb]) with v >= 0."""
    from scipy.optimize import minimize

    iw = 1j * np.asarray(wn, dtype=np.float64)
    target = np.asarray(delta_target, dtype=np.complex128)
    e0 = np.asarray(eps0, dtype=np.float64)
    shape = e0.shape
    x0 = np.concatenate([e0.ravel(), np.asarray(v0, dtype=np.float64).ravel()])
    sol = minimize(_hyb_loss_and_grad, x0, args=(iw, target, shape), jac=True, method="L-BFGS-B",
                   options={"maxiter": int(max_iter), "ftol": 1e-16, "gtol": 1e-14, "maxcor": 20})
    return sol.x[: e0.size].reshape(shape).copy(), np.abs(sol.x[e0.size:].reshape(shape))


def bethe_dmft_bath_fit(u, beta, num_bath=5, t=1.0, n_matsubara=128, mixing=1.0, tol=1e-4, max_iter=30,
                        fit="delta", fit_iter=40, eps0=None, v0=None, verbose=False) -> BetheDMFT:
    """Single-site DMFT of the half-filled Hubbard model on the Bethe lattice
    on the Matsubara axis, with exact diagonalisation of an Anderson model of
    ``num_bath`` bath sites as the impurity solver.

    ``u`` may be a list: every value is one impurity model of a single
    batched solve, so a whole U sweep iterates together.  Each iteration
    solves all models (``AndersonBatch``), forms Sigma = G0^-1 - G^-1 with
    G0^-1 = i w_n - (eps_imp - mu) - Delta, the lattice function
    G_loc = G_Bethe(i w_n - (eps_imp - mu) - Sigma) = G_Bethe(1/G + Delta), and
    refits the bath, mixed with the old one.  ``fit="delta"`` (the usual
    exact-diagonalisation DMFT update): on the Bethe lattice the
    self-consistent hybridisation is t^2 G_loc, and the bath's closed-form
    Delta is fitted to it (``fit_hybridization``).  ``fit="g"``: the bath is
    fitted so that the interacting impurity G itself matches G_loc
    (``fit_bath``: L-BFGS on adjoint gradients of the exact-diagonalisation G,
    ``fit_iter`` iterations).  That minimises the self-consistency residual
    directly, and is meant to refine a solution of the hybridisation fit
    (pass its bath as ``eps0`` / ``v0``): from a generic start bath the G fit
    can settle on another solution (measured: an insulating one at U = 2).  It stops when the
    impurity G of every model changes by less than ``tol`` (max over n) from
    one iteration to the next, as the reference stops on the change of its
    hybridisation (cmpy/dmft/twosite.py:144-156).  With a finite bath G cannot
    equal G_loc exactly: the remaining max_n |G - G_loc| (the fit's
    discretisation error) is reported per model in ``BetheDMFT.residual``.  The reference
    iterates the two-site version on the real axis (cmpy/dmft/twosite.py:
    103-172); this is the same self-consistency with a fitted bath of any
    size.  Half filling: eps_imp - mu = -U/2, a particle-hole symmetric
    start bath, and an odd ``num_bath``, so that the 1 + num_bath sites hold
    the half-filled spin singlet ((num_bath + 1) / 2 electrons of each spin),
    the sector the batched solver works in."""
    from .batch import AndersonBatch, fit_bath, matsubara

    us = np.atleast_1d(np.asarray(u, dtype=np.float64))
    M, nb = len(us), int(num_bath)
    if nb < 1 or nb > 15 or nb % 2 == 0:
        raise ValueError("num_bath must be odd and lie in 1..15 (half-filled singlet ground state)")
    if not 0 < mixing <= 1:
        raise ValueError("mixing must lie in (0, 1]")
    if fit not in ("delta", "g"):
        raise ValueError("fit must be 'delta' or 'g'")
    wn = matsubara(beta, n_matsubara)
    iw = 1j * wn
    if eps0 is None:
        eps0 = np.linspace(-2 * t, 2 * t, nb) * (nb - 1) / nb if nb > 1 else np.zeros(1)
    if v0 is None:
        v0 = np.full(nb, t / np.sqrt(nb))  # sum_k V_k^2 = t^2, the Bethe lattice's hybridisation weight
    eps = np.broadcast_to(np.asarray(eps0, dtype=np.float64), (M, nb)).copy()
    v = np.broadcast_to(np.asarray(v0, dtype=np.float64), (M, nb)).copy()
    level = -us / 2  # eps_imp - mu at half filling
    # physical (anticommuting) hopping signs: with the reference's sign-free
    # hopping the U = 0 impurity is not the free resolvent 1/(i w_n - Delta)
    batch = AndersonBatch(us, 0.0, list(eps), list(v), mu=us / 2, sign_mode="fermion", keep_basis=True)
    stats = IterationStats("max|dG|", "max|G-G_loc|")
    g_prev = None
    for it in range(max_iter + 1):
        batch.set_bath(eps, v)
        g = batch.solve(wn).g
        delta = hybridization(iw, eps, v)
        sigma = (iw[None, :] - level[:, None] - delta) - 1.0 / g
        g_loc = bethe_gf_matsubara(1.0 / g + delta, t)
        change = np.inf if g_prev is None else float(np.abs(g - g_prev).max())
        stats.append(change, float(np.abs(g - g_loc).max()))
        if verbose:
            print(f"DMFT iteration {it}: " + stats.errorstr())
        if change < tol:
            stats.set_status(True, f"G changed by {change:.2e} in the last of {it} bath fits")
            break
        if it == max_iter:
            stats.set_maxiter_status(max_iter)
            break
        g_prev = g
        if fit == "delta":
            e_new, v_new = fit_hybridization(wn, t * t * g_loc, eps, v)
        else:
            res = fit_bath(batch, wn, g_loc, eps0=eps, v0=v, max_iter=fit_iter)
            e_new, v_new = res.eps_bath[:, :nb], np.abs(res.v[:, :nb])
        eps = mix_values(eps, e_new, mixing)
        v = mix_values(v, v_new, mixing)
    return BetheDMFT(us, wn, eps, v, g, g_loc, sigma, stats)
