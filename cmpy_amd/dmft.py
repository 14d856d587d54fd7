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
    "twosite_dmft_half_filling", "compute_lattice_greens_function",
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
