"""Ground state and Green's functions on the GPU (cmpy/exactdiag.py surface).

* ``groundstate``            one sector, device Lanczos (FP64, seeded start,
                              convergence on the Ritz residual, two-pass vector)
* ``compute_groundstate``    minimum over all sectors (cmpy/exactdiag.py:24-43)
* ``solve_sector``           dense eigh of a sector on the device (46-57)
* ``gf_lehmann``             grand-canonical finite-T Lehmann G with occupations
                              (``GreensFunctionMeasurement``, 137-245)
* ``greens_lehmann_t0``      T = 0 fixed-sector Lehmann G (particle + hole),
                              the ``_accumulate_sum`` formula (110-129) restricted
                              to the ground state
* ``greens_cf``              T = 0 continued-fraction G from Lanczos
                              coefficients (no counterpart in cmpy)
* ``greens_lehmann`` / ``greens_cf_ij``   off-diagonal G_ij (finite-T Lehmann,
                              lehmann_full.py:19-97; T = 0 continued fractions)
* ``gf_greater`` / ``gf_lesser`` / ``gf_tevo``   real-time G (248-307) from a
                              Lanczos quadrature of c+psi0 / c psi0
* ``iter_lanczos_coeffs`` / ``lanczos_coeffs`` / ``lanczos_matrix`` /
  ``lanczos_ground_state``   the coefficient API of 324-375
"""

from __future__ import annotations

import ctypes as C
import logging
import warnings
from dataclasses import dataclass
from typing import NamedTuple, Optional, Tuple

import numpy as np
import scipy.linalg as la
import torch

from . import _lib
from .basis import UP, DN, Sector, upper_sector, lower_sector
from .operators import SectorHamiltonOperator, CreationOperator, AnnihilationOperator, SIGN_MODES

logger = logging.getLogger(__name__)

__all__ = [
    "EigenState", "LanczosStats", "groundstate", "compute_groundstate", "solve_sector",
    "GreensFunctionMeasurement", "gf_lehmann", "greens_lehmann_t0", "greens_cf", "CFResult",
    "iter_lanczos_coeffs", "lanczos_coeffs", "lanczos_matrix", "lanczos_ground_state",
    "gf_greater", "gf_lesser", "gf_tevo", "greens_lehmann", "greens_cf_ij", "LehmannIJ",
    "occupation_up", "occupation_dn", "occupation", "double_occupation", "accumulate_gf", "fourier_t2z",
]


class EigenState(NamedTuple):
    """cmpy/matrix.py:42-47."""

    energy: float = np.inf
    state: np.ndarray = None
    n_up: int = None
    n_dn: int = None


@dataclass
class LanczosStats:
    iterations: int
    converged: bool
    breakdown: bool
    ritz_residual: float
    residual: float
    matvec_ms: float
    total_ms: float


def _operator(model_or_op, n_up=None, n_dn=None) -> SectorHamiltonOperator:
    if isinstance(model_or_op, SectorHamiltonOperator):
        return model_or_op
    if n_up is None or n_dn is None:
        raise ValueError("a model needs the sector (n_up, n_dn)")
    return model_or_op.hamilton_operator(n_up=n_up, n_dn=n_dn)


def groundstate(model_or_op, n_up: int = None, n_dn: int = None, *, tol: float = 1e-10,
                max_iter: int = 1000, min_iter: int = 0, check_every: int = 10,
                reorth: Optional[bool] = None, want_vector: bool = True, fixed_iter: bool = False,
                seed: int = 12345, v0=None, device: bool = False, return_stats: bool = False,
                time_matvec: bool = False):
    """Lowest eigenpair of one sector by device Lanczos.

    Converged when the Ritz residual |H psi - E psi| <= tol (absolute, psi
    normalised).  ``reorth`` (None, bool, 'none', 'full', 'partial') defaults
    to full reorthogonalisation up to 2e5 states, partial above while the
    Krylov basis fits in 16 GiB, and the three-vector recurrence beyond.
    Returns ``EigenState(energy, state, n_up, n_dn)``; ``state`` is a numpy array,
    or a CUDA tensor when ``device=True``.  With ``return_stats`` also the
    ``LanczosStats``, whose ``matvec_ms`` (the summed H.v time, from HIP events
    that cost a few microseconds of GPU time per step) is measured only with
    ``time_matvec=True`` and is 0 otherwise.
    """
    op = _operator(model_or_op, n_up, n_dn)
    D = op.size
    if D == 0:
        raise ValueError("empty sector")
    mi = int(min(max_iter, D))
    if reorth is None:
        # full up to 200,000 states; above, partial while the basis fits
        # (one pass instead of two for the vector), else the plain recurrence
        mode = 1 if D <= 200_000 else (2 if _default_reorth(D, mi) else 0)
    else:
        mode = _reorth_mode(reorth, D, mi)
    cfg = _lib.LanczosCfg(mi, int(min_iter), int(check_every), mode, int(bool(want_vector)),
                          int(bool(fixed_iter)), float(tol), int(seed) & 0xFFFFFFFFFFFFFFFF,
                          int(bool(time_matvec)))
    nbytes = C.c_int64()
    _lib.call("ed_lanczos_workspace", op.handle, C.byref(cfg), C.byref(nbytes))
    work = torch.empty(nbytes.value, dtype=torch.uint8, device="cuda")
    psi = torch.empty(D, dtype=torch.float64, device="cuda") if want_vector else None
    v0t = None
    if v0 is not None:
        v0t = v0 if isinstance(v0, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(v0, np.float64))
        v0t = v0t.to("cuda", torch.float64).contiguous()
    res = _lib.LanczosResult()
    _lib.call("ed_lanczos_groundstate", op.handle, C.byref(cfg), _lib.dptr(v0t) if v0t is not None else None,
              _lib.dptr(psi) if psi is not None else None, _lib.dptr(work), nbytes.value, C.byref(res),
              _lib.stream_ptr())
    del work
    state = None
    if psi is not None:
        state = psi if device else psi.cpu().numpy()
    es = EigenState(float(res.e0), state, op.n_up, op.n_dn)
    if return_stats:
        st = LanczosStats(res.iterations, bool(res.converged), bool(res.breakdown), res.ritz_residual,
                          res.residual, res.matvec_ms, res.total_ms)
        return es, st
    return es


def _dense_eigh_device(op: SectorHamiltonOperator) -> Tuple[torch.Tensor, torch.Tensor]:
    """Eigenvalues (ascending) and eigenvectors stored as rows [k][D] on the device."""
    D = op.size
    h = torch.empty(D * D, dtype=torch.float64, device="cuda")
    _lib.call("ed_op_dense", op.handle, _lib.dptr(h), _lib.stream_ptr())
    w = torch.empty(D, dtype=torch.float64, device="cuda")
    _lib.call("ed_dense_eigh", _lib.dptr(h), D, _lib.dptr(w), _lib.stream_ptr())
    return w, h.view(D, D)


def _dense_eigh_sectors(model, sectors, cache: dict, streams: int = 4) -> None:
    """Dense eigenpairs of every listed sector missing from ``cache``.  Sectors
    of equal dimension are diagonalised in one batched solver call, and the
    dimension groups are spread over up to ``streams`` HIP streams, one host
    thread each (the solver's many small kernels then overlap)."""
    by_dim = {}
    for sec in dict.fromkeys(sectors):  # each sector once, in first-seen order
        if sec in cache:
            continue
        op = model.hamilton_operator(n_up=sec[0], n_dn=sec[1])
        if op.size:
            by_dim.setdefault(op.size, []).append((sec, op))
    if not by_dim:
        return
    # largest work first, dealt to the streams greedily by O(D^3) cost
    groups = sorted(by_dim.items(), key=lambda kv: -(kv[0] ** 3) * len(kv[1]))
    k = max(1, min(int(streams), len(groups)))
    lanes = [[] for _ in range(k)]
    load = [0.0] * k
    for g in groups:
        i = load.index(min(load))
        lanes[i].append(g)
        load[i] += float(g[0]) ** 3 * len(g[1])

    def run(lane):
        out = []
        for D, group in lane:
            B = len(group)
            h = torch.empty((B, D, D), dtype=torch.float64, device="cuda")
            for j, (_, op) in enumerate(group):
                _lib.call("ed_op_dense", op.handle, _lib.dptr(h[j]), _lib.stream_ptr())
            w = torch.empty((B, D), dtype=torch.float64, device="cuda")
            # returns after its stream has finished (it reads the solver status)
            _lib.call("ed_dense_eigh_batched", _lib.dptr(h), D, B, _lib.dptr(w), _lib.stream_ptr())
            out.append((group, w, h))
        return out

    if k == 1:
        results = [run(lanes[0])]
    else:
        from concurrent.futures import ThreadPoolExecutor

        main = torch.cuda.current_stream()
        side = [torch.cuda.Stream() for _ in range(k)]
        for st in side:
            st.wait_stream(main)

        def run_on(i):
            with torch.cuda.stream(side[i]):
                return run(lanes[i])

        with ThreadPoolExecutor(k) as ex:
            results = list(ex.map(run_on, range(k)))
        for st in side:
            main.wait_stream(st)
    for lane_out in results:
        for group, w, h in lane_out:
            for j, (sec, _) in enumerate(group):
                cache[sec] = (w[j], h[j])


def compute_groundstate(model, thresh: int = 50, tol: float = 1e-10) -> EigenState:
    """Ground state over every (n_up, n_dn) sector (cmpy/exactdiag.py:24-43)."""
    gs = EigenState()
    for n_up, n_dn in model.basis.iter_fillings():
        op = model.hamilton_operator(n_up=n_up, n_dn=n_dn)
        if op.size <= thresh:
            w, v = _dense_eigh_device(op)
            energy, state = float(w[0].item()), v[0].cpu().numpy()
        else:
            r = groundstate(op, tol=tol)
            energy, state = r.energy, r.state
        if energy < gs.energy:
            gs = EigenState(energy, state, n_up, n_dn)
    return gs


def solve_sector(model, sector: Sector, cache: dict = None):
    """Dense eigendecomposition of one sector (cmpy/exactdiag.py:46-57).

    Returns numpy (eigvals, eigvecs) with eigenvectors as columns, like
    ``np.linalg.eigh``.
    """
    key = (sector.n_up, sector.n_dn)
    if cache is not None and key in cache:
        w, v = cache[key]
    else:
        w, v = _dense_eigh_device(model.hamilton_operator(sector=sector))
        if cache is not None:
            cache[key] = (w, v)
    return w.cpu().numpy(), v.cpu().numpy().T


# =============================================================================
# Observables and the Lehmann accumulation (cmpy/exactdiag.py:60-134)
# =============================================================================


def _device_states(states) -> Tuple[torch.Tensor, int]:
    a = np.asarray(states.cpu() if isinstance(states, torch.Tensor) else states).astype(np.int64).ravel()
    if a.size and (a.min() < 0 or a.max() >= 1 << 32):
        raise ValueError("spin states must lie in [0, 2**32)")
    return torch.from_numpy(a.astype(np.uint32).view(np.int32)).cuda(), int(a.size)


def _rows_device(evecs, dim: int) -> torch.Tensor:
    """Eigenvectors given as columns (dim x N, numpy or torch) -> device rows [N][dim]."""
    if isinstance(evecs, torch.Tensor):
        v = evecs.to("cuda", torch.float64)
    else:
        v = torch.from_numpy(np.ascontiguousarray(np.asarray(evecs, dtype=np.float64))).cuda()
    if v.dim() == 1:
        v = v[:, None]
    if v.dim() != 2 or v.shape[0] != dim:
        raise ValueError(f"evecs must have {dim} rows (one per sector state), got shape {tuple(v.shape)}")
    return v.t().contiguous()


def _occupation_call(up_states, dn_states, evals, evecs, beta, emin, pos, kind) -> float:
    _lib.require_gpu()
    su, nu = _device_states(up_states)
    sd, nd = _device_states(dn_states)
    pos = int(pos)
    if not 0 <= pos < 32:
        raise ValueError(f"pos={pos} outside [0, 32)")
    V = _rows_device(evecs, nu * nd)
    ev = np.atleast_1d(np.asarray(evals.cpu() if isinstance(evals, torch.Tensor) else evals, np.float64))
    if ev.size != V.shape[0]:
        raise ValueError(f"{ev.size} eigenvalues for {V.shape[0]} eigenvectors")
    with np.errstate(invalid="ignore", over="ignore"):
        boltz = np.exp(-beta * (ev - emin))  # the reference's weights, NaN/inf included
    w = torch.from_numpy(np.ascontiguousarray(boltz)).cuda()
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    _lib.call("ed_occupation_states", _lib.dptr(V), V.shape[0], _lib.dptr(su), nu, _lib.dptr(sd), nd, pos, kind,
              _lib.dptr(w), _lib.dptr(out), _lib.stream_ptr())
    return float(out.item())


def occupation_up(up_states, dn_states, evals, evecs, beta, emin, pos) -> float:
    """sum_k e^{-beta(E_k - emin)} sum_{up state has pos} |evecs[i, k]|^2
    (cmpy/exactdiag.py:60-72); evecs as columns, flat index up_idx*num_dn + dn_idx."""
    return _occupation_call(up_states, dn_states, evals, evecs, beta, emin, pos, 0)


def occupation_dn(up_states, dn_states, evals, evecs, beta, emin, pos) -> float:
    """As :func:`occupation_up` for the dn spin (cmpy/exactdiag.py:75-85)."""
    return _occupation_call(up_states, dn_states, evals, evecs, beta, emin, pos, 1)


def occupation(up_states, dn_states, evals, evecs, beta, emin=0.0, pos=0, sigma=UP) -> float:
    """cmpy/exactdiag.py:88-92."""
    if sigma == UP:
        return occupation_up(up_states, dn_states, evals, evecs, beta, emin, pos)
    return occupation_dn(up_states, dn_states, evals, evecs, beta, emin, pos)


def double_occupation(up_states, dn_states, evals, evecs, beta, emin, pos) -> float:
    """sum_k e^{-beta(E_k - emin)} sum_{both spins have pos} |evecs[i, k]|^2
    (cmpy/exactdiag.py:95-107), with the flat index up_idx*num_dn + dn_idx that
    the reference's running ``idx`` stands for (it is only right serially)."""
    return _occupation_call(up_states, dn_states, evals, evecs, beta, emin, pos, 2)


def accumulate_gf(gf, z, cdag, evals, evecs, evals_p1, evecs_p1, beta, emin=0.0) -> None:
    """Add the Lehmann pair sum of one (sector, upper sector) pair into ``gf`` in
    place (cmpy/exactdiag.py:110-134):

        gf += sum_{m,n} |<m|c+|n>|^2 (e^{-beta(E_n-emin)} + e^{-beta(E_m-emin)}) / (z - E_m + E_n)

    ``cdag`` is a :class:`CreationOperator` (applied on the device) or any
    operator with ``matmat``; eigenvectors are columns.  ``beta = inf`` gives
    unit weights, as in the reference.  ``gf`` is a complex numpy array (or a
    complex torch tensor) of the shape of ``z``.
    """
    _lib.require_gpu()
    z = np.asarray(z, dtype=np.complex128)
    ev = np.ascontiguousarray(np.asarray(evals.cpu() if isinstance(evals, torch.Tensor) else evals, np.float64))
    ev1 = np.ascontiguousarray(np.asarray(evals_p1.cpu() if isinstance(evals_p1, torch.Tensor) else evals_p1,
                                          np.float64))
    n, m = ev.size, ev1.size
    d_p1 = evecs_p1.shape[0]
    V1 = _rows_device(evecs_p1, d_p1)
    if V1.shape[0] != m:
        raise ValueError(f"{m} upper-sector eigenvalues for {V1.shape[0]} eigenvectors")
    if isinstance(cdag, CreationOperator):
        V = _rows_device(evecs, cdag.shape[1])
        if V.shape[0] != n:
            raise ValueError(f"{n} eigenvalues for {V.shape[0]} eigenvectors")
        Cm = cdag.apply_device(V)
    else:
        ce = np.asarray(cdag.matmat(evecs.cpu().numpy() if isinstance(evecs, torch.Tensor) else evecs))
        if np.iscomplexobj(ce):
            if np.any(ce.imag):
                raise ValueError("accumulate_gf takes real eigenvectors")
            ce = ce.real
        Cm = _rows_device(ce, d_p1)
        if Cm.shape[0] != n:
            raise ValueError(f"{n} eigenvalues for {Cm.shape[0]} vectors c+ evecs")
    ov = torch.empty((m, n), dtype=torch.float64, device="cuda")
    _lib.call("ed_gemm_nt", _lib.dptr(V1), _lib.dptr(Cm), _lib.dptr(ov), m, n, d_p1, _lib.stream_ptr())
    zd = torch.from_numpy(np.ascontiguousarray(np.stack([z.real.ravel(), z.imag.ravel()], axis=1).ravel())).cuda()
    g = torch.zeros(2 * z.size, dtype=torch.float64, device="cuda")
    b = float(beta) if np.isfinite(beta) else float("inf")
    evd, ev1d = torch.from_numpy(ev).cuda(), torch.from_numpy(ev1).cuda()
    _lib.call("ed_lehmann_pairs", _lib.dptr(ov), _lib.dptr(evd), n, _lib.dptr(ev1d), m, b, float(emin),
              _lib.dptr(zd), z.size, _lib.dptr(g), _lib.stream_ptr())
    gh = g.view(-1, 2).cpu().numpy()
    add = (gh[:, 0] + 1j * gh[:, 1]).reshape(z.shape)
    if isinstance(gf, torch.Tensor):
        gf += torch.from_numpy(add).to(gf.device, gf.dtype)
    else:
        gf += add


# =============================================================================
# Finite-temperature Lehmann Green's function
# =============================================================================


class GreensFunctionMeasurement:
    """Accumulator of the grand-canonical Lehmann sum (cmpy/exactdiag.py:137-212).

    Partition function, G(z) and occupations are accumulated relative to a
    reference energy.  As in the reference, that is the lowest energy seen so
    far, and whenever a lower one appears every accumulated quantity is
    rescaled by exp(-beta (E_old - E_new)).  When the caller knows the lowest
    energy of every sector involved beforehand (``set_reference_energy``; the
    device path diagonalises all sectors first), that fixed value is used
    instead: the normalised results are the same, and the weights of upper
    sectors lying below the lowest energy seen so far can no longer overflow
    at low temperature.
    """

    def __init__(self, z, beta, pos=0, sigma=UP, dtype=None, measure_occ=True):
        self.z = np.asarray(z)
        self.beta = beta
        self.pos = pos
        self.sigma = sigma
        self._measure_occ = measure_occ
        # partition function and occupations stay on the device while the
        # sectors are accumulated, so accumulation never waits for the GPU
        self._part_dev = torch.zeros((), dtype=torch.float64, device="cuda")
        self._occ_dev = torch.zeros(2, dtype=torch.float64, device="cuda")
        self._gs_energy = np.inf
        self._gf = torch.zeros(2 * self.z.size, dtype=torch.float64, device="cuda")
        self._z_dev = torch.from_numpy(np.ascontiguousarray(
            np.stack([self.z.real.ravel(), self.z.imag.ravel()], axis=1).ravel())).cuda()
        self._eref = None

    def set_reference_energy(self, e: float) -> None:
        """Fix the energy the weights are measured from (at most the lowest
        eigenvalue of every sector that will be accumulated)."""
        self._eref = float(e)

    def _shift(self) -> float:
        return self._gs_energy if self._eref is None else self._eref

    @property
    def _part(self) -> float:
        return float(self._part_dev.item())

    @property
    def part(self):
        return self._part * np.exp(-self.beta * self._shift())

    @property
    def gf(self):
        g = self._gf.view(-1, 2).cpu().numpy()
        return ((g[:, 0] + 1j * g[:, 1]) / self._part).reshape(self.z.shape)

    @property
    def occ(self):
        return float(self._occ_dev[0].item()) / self._part

    @property
    def occ_double(self):
        return float(self._occ_dev[1].item()) / self._part

    @property
    def gs_energy(self):
        return self._gs_energy

    def _weights(self, w: torch.Tensor) -> torch.Tensor:
        if np.isfinite(self.beta):
            return torch.exp(-self.beta * (w - self._shift()))
        return torch.ones_like(w)

    def accumulate(self, num_sites, sector, sector_p1, w, V, w_p1, V_p1, sign_mode="reference", min_energy=None):
        """Add one (sector, upper sector) pair; w/V, w_p1/V_p1 are device eigenpairs
        (eigenvectors as rows, eigenvalues ascending).  ``min_energy`` (w[0] as a
        host float, if the caller has it) spares a device round trip."""
        if min_energy is None:
            min_energy = float(w[0].item())
        factor = 1.0
        if min_energy < self._gs_energy:
            if self._eref is None:
                factor = float(np.exp(-self.beta * (self._gs_energy - min_energy)))
            self._gs_energy = min_energy
        boltz = self._weights(w)
        if factor != 1.0:
            self._part_dev.mul_(factor)
            self._occ_dev.mul_(factor)
            self._gf.mul_(factor)
        self._part_dev.add_(boltz.sum())
        n_up, n_dn = sector
        N = V.shape[0]
        # C = c+ V (rows), overlap[m][n] = <m_p1 | c+ | n>
        Cm = torch.empty((N, V_p1.shape[1]), dtype=torch.float64, device="cuda")
        _lib.call("ed_apply_cdag", num_sites, n_up, n_dn, self.pos, self.sigma, SIGN_MODES[sign_mode],
                  _lib.dptr(V), _lib.dptr(Cm), N, _lib.stream_ptr())
        M = V_p1.shape[0]
        ov = torch.empty((M, N), dtype=torch.float64, device="cuda")
        _lib.call("ed_gemm_nt", _lib.dptr(V_p1), _lib.dptr(Cm), _lib.dptr(ov), M, N, V_p1.shape[1],
                  _lib.stream_ptr())
        beta = float(self.beta) if np.isfinite(self.beta) else float("inf")
        _lib.call("ed_lehmann_pairs", _lib.dptr(ov), _lib.dptr(w), N, _lib.dptr(w_p1), M, beta,
                  float(self._shift()), _lib.dptr(self._z_dev), self.z.size, _lib.dptr(self._gf),
                  _lib.stream_ptr())
        if self._measure_occ:
            out = torch.empty(2, dtype=torch.float64, device="cuda")
            kind = 0 if self.sigma == UP else 1
            _lib.call("ed_occupation", _lib.dptr(V), N, num_sites, n_up, n_dn, self.pos, kind,
                      _lib.dptr(boltz), _lib.dptr(out[0:1]), _lib.stream_ptr())
            _lib.call("ed_occupation", _lib.dptr(V), N, num_sites, n_up, n_dn, self.pos, 2,
                      _lib.dptr(boltz), _lib.dptr(out[1:2]), _lib.stream_ptr())
            self._occ_dev.add_(out)


def gf_lehmann(model, z, beta, pos=0, sigma=UP, eig_cache=None, occ=True) -> GreensFunctionMeasurement:
    """Grand-canonical Lehmann Green's function (cmpy/exactdiag.py:215-245).

    Visits the fillings in ``iter_fillings`` order; every sector with an upper
    sector is diagonalised on the device (cached) and accumulated.  With
    ``beta = inf`` the Boltzmann weights are 1, as in the reference.
    """
    basis = model.basis
    ns = model.num_sites
    data = GreensFunctionMeasurement(z, beta, pos, sigma, measure_occ=occ)
    cache = eig_cache if eig_cache is not None else {}
    sign_mode = "jordan_wigner" if getattr(model, "sign_mode", "reference") == "fermion" else "reference"

    pairs = [((nu, nd), upper_sector(nu, nd, sigma, ns)) for nu, nd in basis.iter_fillings()]
    pairs = [(sec, p1) for sec, p1 in pairs if p1 is not None]
    secs = [s for pr in pairs for s in pr]
    _dense_eigh_sectors(model, secs, cache)
    uniq = sorted(set(secs))
    # lowest eigenvalue of every sector, in one device-to-host copy
    emins = dict(zip(uniq, torch.stack([cache[s][0][0] for s in uniq]).cpu().tolist())) if uniq else {}
    if uniq:
        data.set_reference_energy(min(emins.values()))

    for n_up, n_dn in basis.iter_fillings():
        p1 = upper_sector(n_up, n_dn, sigma, ns)
        if p1 is None:
            continue
        w, V = cache[(n_up, n_dn)]
        w1, V1 = cache[p1]
        data.accumulate(ns, (n_up, n_dn), p1, w, V, w1, V1, sign_mode, min_energy=emins[(n_up, n_dn)])
    return data


class LehmannIJ(NamedTuple):
    gf: np.ndarray
    part: float
    gs_energy: float


def _ij_sign_mode(model, sign_mode):
    """Physical (anticommuting) ladder operators wherever the model's own
    hopping carries fermionic signs: the Hubbard model, and the impurity
    model in its fermion mode."""
    if sign_mode is not None:
        return sign_mode
    if getattr(model, "sign_mode", None) == "reference":
        return "reference"
    return "jordan_wigner"


def _cdag_rows(V: torch.Tensor, ns, sector, pos, sigma, sign_mode, d_p1) -> torch.Tensor:
    C = torch.empty((V.shape[0], d_p1), dtype=torch.float64, device="cuda")
    _lib.call("ed_apply_cdag", ns, sector[0], sector[1], pos, sigma, SIGN_MODES[sign_mode], _lib.dptr(V),
              _lib.dptr(C), V.shape[0], _lib.stream_ptr())
    return C


def greens_lehmann(model, z, beta, i, j, sigma=UP, eig_cache=None, sign_mode=None) -> LehmannIJ:
    """Finite-T off-diagonal Green's function (the sum lehmann_full.py:19-97 sets
    out to form, ``greens_lehmann(model, z, beta, i, j)``):

        G_ij(z) = (1/Z) sum_N sum_{m in N+1, n in N} <n|c_i|m><m|c+_j|n>
                  (e^{-beta E_n} + e^{-beta E_m}) / (z - E_m + E_n)

    with Z summed over every sector (the U = 0 sum is then exactly the
    single-particle resolvent).  Every sector is diagonalised on the device
    (sectors of one dimension in one batched call); the overlaps are two
    device GEMMs per sector and the pair sum runs in ``ed_lehmann_pairs_ij``.
    ``sign_mode`` defaults to Jordan-Wigner ladder operators for models with
    fermionic hopping signs (see ``_ij_sign_mode``).
    """
    if not np.isfinite(beta):
        raise ValueError("greens_lehmann needs a finite beta (greens_cf_ij gives T = 0)")
    ns = model.num_sites
    sign_mode = _ij_sign_mode(model, sign_mode)
    cache = eig_cache if eig_cache is not None else {}
    fillings = list(model.basis.iter_fillings())
    _dense_eigh_sectors(model, fillings, cache)
    # all eigenvalues in one device-to-host copy (the spectra are small)
    spectra = [cache[sec][0] for sec in fillings if cache[sec][0].numel()]
    evals = torch.cat(spectra).cpu().numpy()
    emin = float(evals.min())
    part = float(np.exp(-beta * (evals - emin)).sum())
    zt = _z_device(z)
    g = torch.zeros(zt.numel(), dtype=torch.float64, device="cuda")
    bb = float(beta)
    for sec in fillings:
        p1 = upper_sector(sec[0], sec[1], sigma, ns)
        if p1 is None:
            continue
        w, V = cache[sec]
        w1, V1 = cache[p1]
        N, M = V.shape[0], V1.shape[0]
        if N == 0 or M == 0:
            continue
        ovs = []
        for pos in ((i,) if i == j else (i, j)):
            C = _cdag_rows(V, ns, sec, pos, sigma, sign_mode, V1.shape[1])
            ov = torch.empty((M, N), dtype=torch.float64, device="cuda")
            _lib.call("ed_gemm_nt", _lib.dptr(V1), _lib.dptr(C), _lib.dptr(ov), M, N, V1.shape[1],
                      _lib.stream_ptr())
            ovs.append(ov)
        _lib.call("ed_lehmann_pairs_ij", _lib.dptr(ovs[0]), _lib.dptr(ovs[-1]), _lib.dptr(w), N, _lib.dptr(w1), M,
                  bb, emin, _lib.dptr(zt), zt.numel() // 2, _lib.dptr(g), _lib.stream_ptr())
    gh = g.view(-1, 2).cpu().numpy()
    gf = ((gh[:, 0] + 1j * gh[:, 1]) / part).reshape(np.shape(z))
    return LehmannIJ(gf, part * np.exp(-beta * emin), emin)


# =============================================================================
# T = 0 Green's functions in a fixed sector
# =============================================================================


def _z_device(z: np.ndarray) -> torch.Tensor:
    z = np.asarray(z, dtype=np.complex128).ravel()
    return torch.from_numpy(np.ascontiguousarray(np.stack([z.real, z.imag], axis=1).ravel())).cuda()


def _ladder_sector(ns, n_up, n_dn, sigma, create):
    """Sector one particle above (create) or below in spin sigma, or None."""
    d = 1 if create else -1
    nu = n_up + (d if sigma == UP else 0)
    nd = n_dn + (d if sigma == DN else 0)
    return (nu, nd) if (0 <= nu <= ns and 0 <= nd <= ns) else None


def _ladder_device(psi: torch.Tensor, ns, n_up, n_dn, pos, sigma, create, sign_mode):
    sec = _ladder_sector(ns, n_up, n_dn, sigma, create)
    if sec is None:
        return None, None
    nu, nd = sec
    lib = _lib.lib()
    size = int(lib.ed_binom(ns, nu)) * int(lib.ed_binom(ns, nd))
    y = torch.empty(size, dtype=torch.float64, device="cuda")
    _lib.call("ed_apply_cdag" if create else "ed_apply_c", ns, n_up, n_dn, pos, sigma, SIGN_MODES[sign_mode],
              _lib.dptr(psi), _lib.dptr(y), 1, _lib.stream_ptr())
    return y, (nu, nd)


def _sign_mode(model, sign_mode):
    if sign_mode is not None:
        return sign_mode
    return "jordan_wigner" if getattr(model, "sign_mode", "reference") == "fermion" else "reference"


def _gs(model, n_up, n_dn, gs, tol):
    if gs is not None:
        psi = gs.state if isinstance(gs.state, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(gs.state))
        return float(gs.energy), psi.to("cuda", torch.float64).contiguous()
    r = groundstate(model, n_up, n_dn, tol=tol, device=True)
    return r.energy, r.state


def greens_lehmann_t0(model, z, n_up, n_dn, pos=0, sigma=UP, gs: EigenState = None, tol=1e-10,
                      sign_mode=None) -> np.ndarray:
    """T = 0 Lehmann G in sector (n_up, n_dn):

    G(z) = sum_m |<m|c+|0>|^2/(z - E_m + E0) + sum_n |<n|c|0>|^2/(z + E_n - E0)

    with the N+-1 sectors diagonalised densely on the device.
    """
    ns = model.num_sites
    sign_mode = _sign_mode(model, sign_mode)
    e0, psi = _gs(model, n_up, n_dn, gs, tol)
    zt = _z_device(z)
    ops = []
    for create in (True, False):
        sec = _ladder_sector(ns, n_up, n_dn, sigma, create)
        ops.append(model.hamilton_operator(n_up=sec[0], n_dn=sec[1]) if sec is not None else None)
    hp, hm = (op.handle if op is not None else None for op in ops)
    nbytes = C.c_int64()
    _lib.call("ed_greens_lehmann_t0_workspace", hp, hm, C.byref(nbytes))
    work = torch.empty(max(nbytes.value, 1), dtype=torch.uint8, device="cuda")
    g = torch.empty(zt.numel(), dtype=torch.float64, device="cuda")
    _lib.call("ed_greens_lehmann_t0", hp, hm, ns, n_up, n_dn, pos, sigma, SIGN_MODES[sign_mode], _lib.dptr(psi),
              float(e0), _lib.dptr(zt), zt.numel() // 2, _lib.dptr(g), _lib.dptr(work), nbytes.value,
              _lib.stream_ptr())
    gh = g.view(-1, 2).cpu().numpy()
    return (gh[:, 0] + 1j * gh[:, 1]).reshape(np.shape(z))


@dataclass
class CFResult:
    gf: np.ndarray
    e0: float
    alpha_p: np.ndarray
    beta_p: np.ndarray
    norm2_p: float
    alpha_h: np.ndarray
    beta_h: np.ndarray
    norm2_h: float
    reorth_p: str = "none"  # "none", "full" or "partial": how each fraction's basis was kept orthogonal
    reorth_h: str = "none"


REORTH_MAX_BYTES = 16 << 30  # Krylov basis always allowed for reorthogonalisation
REORTH_FREE_FRACTION = 0.4  # ... or up to this share of the free device memory, if larger
_MODE_NAMES = {0: "none", 1: "full", 2: "partial"}


PARTIAL_REORTH_MIN_SIZE = 50_000  # below this, full reorthogonalisation is as cheap


def _reorth_budget() -> int:
    budget = REORTH_MAX_BYTES
    if torch.cuda.is_available():
        free, _ = torch.cuda.mem_get_info()
        budget = max(budget, int(REORTH_FREE_FRACTION * free))
    return budget


def _default_reorth(size: int, nsteps: int) -> bool:
    need = 8 * size * (nsteps + 2)
    return nsteps <= 4096 and (need <= REORTH_MAX_BYTES or need <= _reorth_budget())


def _reorth_mode(reorth, size: int, nsteps: int, warn: bool = False) -> int:
    """0 none, 1 full (two Gram-Schmidt passes every step), 2 partial (only
    where Simon's orthogonality estimate passes sqrt(eps)).  ``None``: partial
    from PARTIAL_REORTH_MIN_SIZE states, full below, none when the Krylov basis
    would exceed both REORTH_MAX_BYTES and REORTH_FREE_FRACTION of the free
    device memory; with ``warn`` that fallback raises a RuntimeWarning, since
    an unreorthogonalised fraction loses accuracy at small broadening."""
    if reorth is None:
        if not _default_reorth(size, nsteps):
            if warn:
                warnings.warn(
                    f"continued fraction of {nsteps} steps on {size} states: the Krylov basis "
                    f"({8 * size * (nsteps + 2) / 2**30:.1f} GiB) does not fit, so it runs without "
                    "reorthogonalisation (expect ~1e-3 errors in G at Im z ~ 0.05; pass reorth= to override)",
                    RuntimeWarning, stacklevel=4)
            return 0
        return 2 if size >= PARTIAL_REORTH_MIN_SIZE else 1
    if isinstance(reorth, str):
        modes = {"none": 0, "full": 1, "partial": 2}
        if reorth not in modes:
            raise ValueError("reorth must be None, a bool, or one of 'none', 'full', 'partial'")
        return modes[reorth]
    return 1 if reorth else 0


def _cf_coeffs(op: SectorHamiltonOperator, phi: torch.Tensor, nsteps: int, reorth=None, modes: list = None):
    nsteps = int(max(1, min(nsteps, op.size)))
    mode = _reorth_mode(reorth, op.size, nsteps, warn=True)
    if modes is not None:
        modes.append(_MODE_NAMES[mode])
    nbytes = C.c_int64()
    _lib.call("ed_lanczos_coeffs_workspace", op.handle, nsteps, mode, C.byref(nbytes))
    work = torch.empty(nbytes.value, dtype=torch.uint8, device="cuda")
    a = np.zeros(nsteps)
    b = np.zeros(max(nsteps, 1))
    n2 = C.c_double()
    m = C.c_int32()
    _lib.call("ed_lanczos_coeffs", op.handle, _lib.dptr(phi), nsteps, mode,
              a.ctypes.data_as(C.POINTER(C.c_double)), b.ctypes.data_as(C.POINTER(C.c_double)),
              C.byref(n2), C.byref(m), None, _lib.dptr(work), nbytes.value, _lib.stream_ptr())
    return a[: m.value].copy(), b[: max(m.value - 1, 0)].copy(), float(n2.value)


def greens_cf(model, z, n_up, n_dn, pos=0, sigma=UP, nsteps=200, gs: EigenState = None, tol=1e-10,
              sign_mode=None, reorth=None) -> CFResult:
    """T = 0 continued-fraction G from ``nsteps`` Lanczos steps started at c+psi0
    (particle, sector N+1) and c psi0 (hole, sector N-1).  ``reorth`` (None,
    bool, 'none', 'full' or 'partial') defaults to partial reorthogonalisation
    from 50,000 states and full below, whenever the Krylov basis fits in
    16 GiB or in 40% of the free device memory; beyond that the fraction runs
    without reorthogonalisation and a RuntimeWarning says so.  The mode used
    is recorded in ``CFResult.reorth_p`` / ``reorth_h``."""
    ns = model.num_sites
    sign_mode = _sign_mode(model, sign_mode)
    e0, psi = _gs(model, n_up, n_dn, gs, tol)
    zt = _z_device(z)
    nz = zt.numel() // 2
    g = torch.zeros(zt.numel(), dtype=torch.float64, device="cuda")
    coeffs, modes = {}, {}
    # one ed_greens_cf call per part (the other operator NULL), so that each
    # part gets the reorthogonalisation mode its own sector size calls for
    for k, create in enumerate((True, False)):
        sec = _ladder_sector(ns, n_up, n_dn, sigma, create)
        if sec is None:
            coeffs[create] = (np.zeros(0), np.zeros(0), 0.0)
            continue
        op = model.hamilton_operator(n_up=sec[0], n_dn=sec[1])
        steps = int(max(1, min(nsteps, op.size)))
        mode = _reorth_mode(reorth, op.size, steps, warn=True)
        modes[create] = _MODE_NAMES[mode]
        hp, hm = (op.handle, None) if create else (None, op.handle)
        nbytes = C.c_int64()
        _lib.call("ed_greens_cf_workspace", hp, hm, steps, mode, C.byref(nbytes))
        work = torch.empty(nbytes.value, dtype=torch.uint8, device="cuda")
        a = np.zeros(steps)
        b = np.zeros(steps)
        parts = _lib.CfParts()
        parts.alpha[k] = a.ctypes.data_as(C.POINTER(C.c_double))
        parts.beta[k] = b.ctypes.data_as(C.POINTER(C.c_double))
        gk = torch.empty(zt.numel(), dtype=torch.float64, device="cuda")
        _lib.call("ed_greens_cf", hp, hm, ns, n_up, n_dn, pos, sigma, SIGN_MODES[sign_mode], _lib.dptr(psi),
                  float(e0), steps, mode, _lib.dptr(zt), nz, _lib.dptr(gk), C.byref(parts), _lib.dptr(work),
                  nbytes.value, _lib.stream_ptr())
        m = int(parts.steps[k])
        coeffs[create] = (a[:m].copy(), b[: max(m - 1, 0)].copy(), float(parts.norm2[k]))
        g += gk
    gh = g.view(-1, 2).cpu().numpy()
    gf = (gh[:, 0] + 1j * gh[:, 1]).reshape(np.shape(z))
    ap, bp, np_ = coeffs[True]
    ah, bh, nh = coeffs[False]
    return CFResult(gf, e0, ap, bp, np_, ah, bh, nh, modes.get(True, "none"), modes.get(False, "none"))


def greens_cf_ij(model, z, n_up, n_dn, i, j, sigma=UP, nsteps=200, gs: EigenState = None, tol=1e-10,
                 sign_mode=None, reorth: Optional[bool] = None) -> np.ndarray:
    """T = 0 off-diagonal G_ij(z) in sector (n_up, n_dn) from continued fractions,
    by polarisation: with c_(+-) = c_i +- c_j,

        G_ij = (G[c_+] - G[c_-]) / 4,

    where G[c] is the particle + hole continued fraction of the combined
    operator (four Lanczos runs; G_ij = G_ji for a real Hamiltonian).
    ``sign_mode`` as in :func:`greens_lehmann`."""
    ns = model.num_sites
    sign_mode = _ij_sign_mode(model, sign_mode)
    e0, psi = _gs(model, n_up, n_dn, gs, tol)
    zt = _z_device(z)
    if i == j:
        r = greens_cf(model, z, n_up, n_dn, pos=i, sigma=sigma, nsteps=nsteps, gs=EigenState(e0, psi, n_up, n_dn),
                      sign_mode=sign_mode, reorth=reorth)
        return r.gf
    g = torch.zeros(zt.numel(), dtype=torch.float64, device="cuda")
    for create in (True, False):
        ai, sec = _ladder_device(psi, ns, n_up, n_dn, i, sigma, create, sign_mode)
        if ai is None:
            continue
        aj, _ = _ladder_device(psi, ns, n_up, n_dn, j, sigma, create, sign_mode)
        op = model.hamilton_operator(n_up=sec[0], n_dn=sec[1])
        for pm in (1.0, -1.0):
            phi = ai + pm * aj
            a, b, n2 = _cf_coeffs(op, phi, nsteps, reorth)
            if len(a) == 0 or n2 == 0.0:
                continue
            _lib.call("ed_cf_eval", a.ctypes.data_as(C.POINTER(C.c_double)),
                      (b if len(b) else np.zeros(1)).ctypes.data_as(C.POINTER(C.c_double)), len(a), pm * n2 / 4.0,
                      e0, 1 if create else -1, _lib.dptr(zt), zt.numel() // 2, _lib.dptr(g), 1, _lib.stream_ptr())
    gh = g.view(-1, 2).cpu().numpy()
    return (gh[:, 0] + 1j * gh[:, 1]).reshape(np.shape(z))


# =============================================================================
# Real-time Green's functions (cmpy/exactdiag.py:248-307)
# =============================================================================


def _tevo_nsteps(op: SectorHamiltonOperator, tmax: float) -> int:
    """Lanczos steps for the real-time quadrature to reach FP64 accuracy up to
    |t| = tmax.  m Gauss nodes integrate polynomials of degree 2m - 1 exactly,
    and the Chebyshev coefficients of exp(-i x t) on a spectrum of width W
    are Bessel values J_k(W t / 2), negligible for k > W t / 2 + O((W t)^(1/3)).
    W is bounded by Gershgorin: the diagonal's range plus four times the summed
    |hopping| (a state hops along a bond at most once per spin)."""
    d = op.diagonal_device()
    width = float((d.max() - d.min()).item()) + 4.0 * float(np.abs(np.asarray(op.hops, dtype=np.float64)).sum())
    del d
    x = max(width * float(tmax), 0.0)
    return int(np.ceil(x / 4.0 + 6.0 * (x / 2.0) ** (1.0 / 3.0) + 20.0))


def _tevo(model, gs, times, pos, sigma, create, nsteps, reorth, sign_mode, tol):
    ns = model.num_sites
    sign_mode = _sign_mode(model, sign_mode)
    n_up, n_dn = gs.n_up, gs.n_dn
    e0, psi = _gs(model, n_up, n_dn, gs, tol)
    phi, sec = _ladder_device(psi, ns, n_up, n_dn, pos, sigma, create, sign_mode)
    if phi is None:
        return None
    op = model.hamilton_operator(n_up=sec[0], n_dn=sec[1])
    tmax = float(np.abs(times).max()) if len(times) else 0.0
    m = int(nsteps) if nsteps is not None else _tevo_nsteps(op, tmax)
    # the plain recurrence: its quadrature matches the reorthogonalised one to
    # ~1e-14 (lost orthogonality only splits a converged node's weight over
    # copies) at a fraction of the cost
    a, b, n2 = _cf_coeffs(op, phi, m, False if reorth is None else reorth)
    tt = torch.from_numpy(np.ascontiguousarray(times, dtype=np.float64)).cuda()
    g = torch.zeros(2 * len(times), dtype=torch.float64, device="cuda")
    if len(a) and len(times):
        _lib.call("ed_tevo_eval", a.ctypes.data_as(C.POINTER(C.c_double)),
                  (b if len(b) else np.zeros(1)).ctypes.data_as(C.POINTER(C.c_double)), len(a), n2, e0,
                  1 if create else -1, _lib.dptr(tt), len(times), _lib.dptr(g), 0, _lib.stream_ptr())
    gh = g.view(-1, 2).cpu().numpy()
    return gh[:, 0] + 1j * gh[:, 1]


def gf_greater(model, gs: EigenState, start, stop, num=1000, pos=0, sigma=UP, nsteps=None,
               reorth: Optional[bool] = None, sign_mode=None, tol=1e-10):
    """Greater real-time Green's function (cmpy/exactdiag.py:248-272):

    g>(t) = -i e^{i E0 t} <psi0| c e^{-i H t} c+ |psi0>  on linspace(start, stop, num).

    The reference propagates c+ psi0 with expm_multiply and overlaps at every
    time; here one Lanczos run from c+ psi0 in the upper sector gives the Gauss
    quadrature of its spectral measure, and the time sum runs on the device
    (``ed_tevo_eval``).  ``nsteps`` defaults to a count that reaches FP64
    accuracy up to max |t| (see ``_tevo_nsteps``); ``reorth`` defaults to the
    plain three-vector recurrence.  Returns (times, g).
    """
    times = np.linspace(start, stop, num)
    g = _tevo(model, gs, times, pos, sigma, True, nsteps, reorth, sign_mode, tol)
    if g is None:
        logger.warning("Upper sector not found!")
        return times, np.zeros_like(times)
    return times, g


def gf_lesser(model, gs: EigenState, start, stop, num=1000, pos=0, sigma=UP, nsteps=None,
              reorth: Optional[bool] = None, sign_mode=None, tol=1e-10):
    """Lesser real-time Green's function (cmpy/exactdiag.py:275-299):

    g<(t) = +i e^{-i E0 t} <psi0| c+ e^{+i H t} c |psi0>  on linspace(start, stop, num).
    Same method as :func:`gf_greater`, from c psi0 in the lower sector.
    """
    times = np.linspace(start, stop, num)
    g = _tevo(model, gs, times, pos, sigma, False, nsteps, reorth, sign_mode, tol)
    if g is None:
        logger.warning("Lower sector not found!")
        return times, np.zeros_like(times)
    return times, g


def gf_tevo(model, start, stop, num=1000, pos=0, sigma=UP, nsteps=None, reorth: Optional[bool] = None,
            sign_mode=None):
    """g>(t) - g<(t) from the ground state over all sectors (cmpy/exactdiag.py:302-306)."""
    gs = compute_groundstate(model)
    times, gf_g = gf_greater(model, gs, start, stop, num, pos, sigma, nsteps, reorth, sign_mode)
    times, gf_l = gf_lesser(model, gs, start, stop, num, pos, sigma, nsteps, reorth, sign_mode)
    return times, gf_g - gf_l


def fourier_t2z(times, gf_t, omegas, delta=1e-2, eta=None):
    """Real-time G to complex frequencies (cmpy/exactdiag.py:311-316):
    z = omegas + i eta, eta = -log(delta) / times[-1] unless given, and

        G(z) = int_{t_0}^{t_end} e^{i z t} G(t) dt

    with G linear between the samples and each interval integrated exactly on
    the device (``ed_fourier_t2z``) -- the piecewise-linear rule that
    gftool.fourier.tt2z applies by default.  Returns (z, G(z))."""
    t = np.ascontiguousarray(times, dtype=np.float64).ravel()
    g = np.ascontiguousarray(gf_t, dtype=np.complex128).ravel()
    if t.size != g.size:
        raise ValueError("times and gf_t must have the same length")
    if t.size and np.any(np.diff(t) <= 0):
        raise ValueError("times must increase")
    if eta is None:
        eta = -np.log(delta) / t[-1]
    z = np.asarray(omegas, dtype=np.float64) + 1j * eta
    zt = _z_device(z)
    tt = torch.from_numpy(t).cuda()
    gt_ = torch.from_numpy(np.ascontiguousarray(np.stack([g.real, g.imag], axis=1).ravel())).cuda()
    out = torch.empty_like(zt)
    _lib.call("ed_fourier_t2z", _lib.dptr(tt), _lib.dptr(gt_), t.size, _lib.dptr(zt), zt.numel() // 2,
              _lib.dptr(out), _lib.stream_ptr())
    oh = out.view(-1, 2).cpu().numpy()
    return z, (oh[:, 0] + 1j * oh[:, 1]).reshape(np.shape(z))


# =============================================================================
# Lanczos coefficient API (cmpy/exactdiag.py:324-375)
# =============================================================================


def iter_lanczos_coeffs(ham: SectorHamiltonOperator, size: int = 10, v0=None, seed: int = 0):
    """Yields (a_n, b_n) like cmpy (first b is None), from a normalised device
    Lanczos run.  ``v0`` defaults to a seeded uniform [0, 1) vector (cmpy draws an
    unseeded one, exactdiag.py:326)."""
    a, b = lanczos_coeffs(ham, size, v0=v0, seed=seed)
    yield a[0], None
    for k in range(1, len(a)):
        yield a[k], b[k - 1]


def lanczos_coeffs(ham: SectorHamiltonOperator, size: int = 10, v0=None, seed: int = 0, reorth=False):
    """(a[size], b[size-1]) Lanczos coefficients of ``ham`` (cmpy/exactdiag.py:350-358)."""
    if not isinstance(ham, SectorHamiltonOperator):
        raise TypeError("lanczos_coeffs needs a SectorHamiltonOperator (model.hamilton_operator(...))")
    if v0 is None:
        v0 = np.random.default_rng(seed).uniform(0, 1, size=ham.size)
    phi = torch.from_numpy(np.ascontiguousarray(v0, np.float64)).cuda()
    a, b, _ = _cf_coeffs(ham, phi, size, reorth)
    return list(a), list(b)


def lanczos_matrix(a_coeffs, b_coeffs):
    """Tridiagonal matrix of the coefficients (cmpy/exactdiag.py:361-365)."""
    mat = np.diag(np.asarray(a_coeffs, dtype=np.float64))
    np.fill_diagonal(mat[1:], b_coeffs)
    np.fill_diagonal(mat[:, 1:], b_coeffs)
    return mat


def lanczos_ground_state(a_coeffs, b_coeffs, max_eig=3):
    """Lowest eigenpair of the tridiagonal matrix (cmpy/exactdiag.py:368-375)."""
    max_eig = min(max_eig, len(a_coeffs) - 1)
    xi, vi = la.eigh_tridiagonal(a_coeffs, b_coeffs, select="i", select_range=(0, max_eig))
    idx = np.argmin(xi)
    return xi[idx], vi[:, idx]
