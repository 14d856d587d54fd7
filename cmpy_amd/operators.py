"""Linear operators of the sector Hamiltonian, backed by the HIP engine.

Mirrors cmpy's ``cmpy/operators.py``:

* the projector generators ``project_up/dn``, ``project_elements_up/dn``,
  ``project_onsite_energy``, ``project_hubbard_inter``, ``project_hopping``
  yield the same (row, col, value) triples in the same order
  (cmpy/operators.py:33-527); the hopping integer work (target index by
  combinadic rank, fermionic sign by popcount) runs in ``ed_hop_terms``;
* ``LinearOperator`` (cmpy/operators.py:535-608) and ``HamiltonOperator``
  (COO triples, cmpy/operators.py:614-646);
* ``CreationOperator`` / ``AnnihilationOperator`` (cmpy/operators.py:652-791).

New here: ``SectorHamiltonOperator``, the matrix-free operator that
``model.hamilton_operator`` returns.  It never materialises the triples: the
sweep kernel applies diagonal + up hops + dn hops directly.
"""

from __future__ import annotations

import abc
import ctypes as C
from typing import Optional, Sequence

import numpy as np
import scipy.sparse.linalg as sla
import torch

from . import _lib
from .basis import UP, DN, SPIN_CHARS, sector_states

__all__ = [
    "LinearOperator", "HamiltonOperator", "SectorHamiltonOperator", "CreationOperator",
    "AnnihilationOperator", "project_up", "project_dn", "project_elements_up",
    "project_elements_dn", "project_hubbard_inter", "project_onsite_energy", "project_hopping",
    "hopping_terms", "weighted_element",
]

SIGN_MODES = {"reference": _lib.ED_SIGN_REFERENCE, "jordan_wigner": _lib.ED_SIGN_JORDAN_WIGNER}


# =============================================================================
# Projectors (reference triple order)
# =============================================================================


def project_up(up_idx, num_dn_states, dn_indices):
    """Full-sector indices of an up state (cmpy/operators.py:33-60)."""
    return np.atleast_1d(up_idx * num_dn_states + np.asarray(dn_indices))


def project_dn(dn_idx, num_dn_states, up_indices):
    """Full-sector indices of a dn state (cmpy/operators.py:63-90)."""
    return np.atleast_1d(np.asarray(up_indices) * num_dn_states + dn_idx)


def project_elements_up(up_idx, num_dn_states, dn_indices, value, target=None):
    """cmpy/operators.py:93-155."""
    if not value:
        return
    origins = project_up(up_idx, num_dn_states, dn_indices)
    if target is None:
        for o in origins:
            yield int(o), int(o), value
    else:
        targets = project_up(target, num_dn_states, dn_indices)
        for o, t in zip(origins, targets):
            yield int(o), int(t), value


def project_elements_dn(dn_idx, num_dn_states, up_indices, value, target=None):
    """cmpy/operators.py:158-220."""
    if not value:
        return
    origins = project_dn(dn_idx, num_dn_states, up_indices)
    if target is None:
        for o in origins:
            yield int(o), int(o), value
    else:
        targets = project_dn(target, num_dn_states, up_indices)
        for o, t in zip(origins, targets):
            yield int(o), int(t), value


def weighted_element(states, values) -> np.ndarray:
    """Sum of ``values[i]`` over the set bits of each state, in site order
    (cmpy/operators.py:226-250)."""
    st = np.asarray(states, dtype=np.int64)
    out = np.zeros(st.shape, dtype=np.float64)
    for i, v in enumerate(np.asarray(values, dtype=np.float64)):
        out = np.where((st >> i) & 1 == 1, out + v, out)
    return out


def project_hubbard_inter(up_states, dn_states, u):
    """Diagonal sum_i u_i n_i,up n_i,dn; nonzero entries only (operators.py:305-356)."""
    up = np.asarray(up_states, np.int64)
    dn = np.asarray(dn_states, np.int64)
    num_dn = len(dn)
    e = weighted_element((up[:, None] & dn[None, :]).ravel(), u)
    for k in np.nonzero(e)[0]:
        yield int(k), int(k), float(e[k])


def project_onsite_energy(up_states, dn_states, eps):
    """Diagonal eps terms, up part then dn part (cmpy/operators.py:359-422)."""
    up = np.asarray(up_states, np.int64)
    dn = np.asarray(dn_states, np.int64)
    num_up, num_dn = len(up), len(dn)
    e_up = weighted_element(up, eps)
    e_dn = weighted_element(dn, eps)
    for iu in range(num_up):
        if e_up[iu]:
            for k in iu * num_dn + np.arange(num_dn):
                yield int(k), int(k), float(e_up[iu])
    for idn in range(num_dn):
        if e_dn[idn]:
            for k in np.arange(num_up) * num_dn + idn:
                yield int(k), int(k), float(e_dn[idn])


def hopping_terms(num_sites: int, n: int, width: int, site1: int, site2: int, hop: float):
    """Per-spin hop list of one bond, computed on the GPU (``ed_hop_terms``).

    Returns (origin, target, value) in ascending origin order: the entries of
    ``_compute_hopping_term`` (cmpy/operators.py:436-460).  ``width`` is the
    sign width (num_sites: fermionic signs; 0: all +1).
    """
    _lib.require_gpu()
    count = int(_lib.lib().ed_binom(num_sites, n))
    tgt = torch.empty(max(count, 1), dtype=torch.int64, device="cuda")
    val = torch.empty(max(count, 1), dtype=torch.float64, device="cuda")
    _lib.call("ed_hop_terms", num_sites, n, int(width), int(site1), int(site2), float(hop),
              _lib.dptr(tgt), _lib.dptr(val), _lib.stream_ptr())
    tgt = tgt[:count].cpu().numpy()
    val = val[:count].cpu().numpy()
    origin = np.nonzero(tgt >= 0)[0]
    return origin, tgt[origin], val[origin]


def _popcount_n(states) -> int:
    st = list(states)
    if not st:
        return 0
    return bin(int(st[0])).count("1")


def project_hopping(up_states, dn_states, num_sites, site1, site2, hop):
    """Hopping between two sites, up block then dn block (cmpy/operators.py:463-527).

    ``num_sites`` is the sign width, as in the reference (the Anderson model
    passes 0, which makes every sign +1).
    """
    if not site1 < site2:
        raise ValueError(f"site1={site1} must be smaller than site2={site2}")
    up = np.asarray(up_states, np.int64)
    dn = np.asarray(dn_states, np.int64)
    ns = int(max(int(up.max(initial=0)), int(dn.max(initial=0)), (1 << site2))).bit_length()
    ns = max(ns, site2 + 1)
    num_up, num_dn = len(up), len(dn)
    for states, is_up in ((up, True), (dn, False)):
        n = _popcount_n(states)
        full = sector_states(ns, n)
        if len(full) != len(states) or not np.array_equal(full, states):
            raise ValueError("project_hopping expects complete ascending sector state lists")
        o, t, a = hopping_terms(ns, n, num_sites, site1, site2, hop)
        if is_up:
            for oi, ti, ai in zip(o, t, a):
                for d in range(num_dn):
                    yield int(oi * num_dn + d), int(ti * num_dn + d), float(ai)
        else:
            for oi, ti, ai in zip(o, t, a):
                for u in range(num_up):
                    yield int(u * num_dn + oi), int(u * num_dn + ti), float(ai)


# =============================================================================
# Operators
# =============================================================================


class LinearOperator(sla.LinearOperator, abc.ABC):
    """scipy LinearOperator with ``toarray`` and ``trace`` (cmpy/operators.py:535-608)."""

    def __init__(self, shape, dtype=None):
        sla.LinearOperator.__init__(self, shape=shape, dtype=dtype)

    def __repr__(self):
        return f"{self.__class__.__name__}(shape: {self.shape}, dtype: {self.dtype})"

    def toarray(self) -> np.ndarray:
        return self.matmat(np.eye(self.shape[1], dtype=self.dtype))

    def _trace(self) -> float:
        return float(np.trace(self.toarray()))

    def trace(self) -> float:
        return self._trace()

    def __mul__(self, x):
        scaled = super().__mul__(x)
        try:
            scaled.trace = lambda: x * self.trace()
            scaled.toarray = lambda: x * self.toarray()
        except AttributeError:
            pass
        return scaled

    def __rmul__(self, x):
        scaled = super().__rmul__(x)
        try:
            scaled.trace = lambda: x * self.trace()
            scaled.toarray = lambda: x * self.toarray()
        except AttributeError:
            pass
        return scaled


def _as_device(x: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to("cuda", non_blocking=False)


class SectorHamiltonOperator(LinearOperator):
    """Matrix-free H of one (n_up, n_dn) sector on the GPU.

    H = sum_i eps_i (n_i,up + n_i,dn) + sum_i u_i n_i,up n_i,dn
        + sum_b hop_b sum_sigma (c+_j c_i + h.c.) * sign_width(state)

    ``width`` is the hopping-sign width: ``num_sites`` gives fermionic signs
    (cmpy's Hubbard model), 0 gives all +1 (cmpy's Anderson model).
    Accepts numpy arrays (staged through HBM) or CUDA tensors (zero copy).
    """

    def __init__(self, num_sites: int, n_up: int, n_dn: int, bonds, hops, eps, u, width: int):
        _lib.require_gpu()
        bonds = np.asarray(bonds, dtype=np.int32).reshape(-1, 2)
        hops = np.broadcast_to(np.asarray(hops, dtype=np.float64), (len(bonds),)).copy()
        eps = np.broadcast_to(np.asarray(eps, dtype=np.float64), (num_sites,)).copy()
        u = np.broadcast_to(np.asarray(u, dtype=np.float64), (num_sites,)).copy()
        self.num_sites, self.n_up, self.n_dn = int(num_sites), int(n_up), int(n_dn)
        self.bonds, self.hops, self.eps, self.u, self.width = bonds, hops, eps, u, int(width)
        bb, bptr = _lib.i32p(bonds.ravel() if len(bonds) else np.zeros(2, np.int32))
        hh, hptr = _lib.f64p(hops if len(hops) else np.zeros(1))
        ee, eptr = _lib.f64p(eps)
        uu, uptr = _lib.f64p(u)
        desc = _lib.ModelDesc(self.num_sites, self.n_up, self.n_dn, len(bonds), bptr, hptr, eptr,
                              uptr, self.width)
        h = C.c_void_p()
        _lib.call("ed_op_create", C.byref(h), C.byref(desc), _lib.stream_ptr())
        self._h = h
        nu, nd = C.c_int64(), C.c_int64()
        _lib.call("ed_op_shape", h, C.byref(nu), C.byref(nd))
        self.num_up, self.num_dn = nu.value, nd.value
        size = self.num_up * self.num_dn
        super().__init__((size, size), dtype=np.float64)

    @property
    def handle(self):
        return self._h

    @property
    def size(self) -> int:
        return self.shape[0]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().ed_op_destroy(h)
            except Exception:
                pass
            self._h = None

    # -- device API -------------------------------------------------------------
    def matvec_device(self, x: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """y = H x for CUDA float64 tensors of shape (D,) or (k, D) (vectors as rows)."""
        D = self.size
        if x.dtype != torch.float64 or not x.is_cuda:
            raise TypeError("matvec_device expects a CUDA float64 tensor")
        x = x.contiguous()
        nvec = 1 if x.dim() == 1 else x.shape[0]
        if x.numel() != nvec * D:
            raise ValueError(f"expected {D} entries per vector, got shape {tuple(x.shape)}")
        if out is None:
            out = torch.empty_like(x)
        _lib.call("ed_op_matvec", self._h, _lib.dptr(x), _lib.dptr(out), nvec, _lib.stream_ptr())
        return out

    @property
    def vector_ld(self) -> int:
        """Row stride (doubles) of the padded vector layout the library's own
        Lanczos uses for this sector: num_dn rounded up to 128-byte rows where
        the two-pass sweep runs, num_dn otherwise."""
        ld = C.c_int64()
        _lib.call("ed_op_vector_ld", self._h, C.byref(ld))
        return ld.value

    def matvec_device_ld(self, x: torch.Tensor, ld: int, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """y = H x for CUDA float64 tensors holding num_up rows of ``ld`` doubles
        each ((num_up, ld) or (k, num_up, ld)); entries past num_dn in a row are
        not read in x and are written as zeros in y."""
        if x.dtype != torch.float64 or not x.is_cuda:
            raise TypeError("matvec_device_ld expects a CUDA float64 tensor")
        x = x.contiguous()
        n = self.num_up * int(ld)
        nvec = x.numel() // n if n else 0
        if n == 0 or x.numel() != nvec * n:
            raise ValueError(f"expected rows of {ld} doubles ({n} per vector), got shape {tuple(x.shape)}")
        if out is None:
            out = torch.empty_like(x)
        _lib.call("ed_op_matvec_ld", self._h, _lib.dptr(x), _lib.dptr(out), int(ld), nvec, _lib.stream_ptr())
        return out

    def diagonal_device(self) -> torch.Tensor:
        d = torch.empty(self.size, dtype=torch.float64, device="cuda")
        _lib.call("ed_op_diagonal", self._h, _lib.dptr(d), _lib.stream_ptr())
        return d

    def diagonal(self) -> np.ndarray:
        return self.diagonal_device().cpu().numpy()

    def hop_table(self, sigma: int):
        """(cnt, col, val) ELL hop table of one spin species (host copies)."""
        w = C.c_int32()
        _lib.call("ed_op_hop_table", self._h, sigma, C.byref(w), None, None, None, _lib.stream_ptr())
        n = self.num_up if sigma == UP else self.num_dn
        cnt = torch.empty(n, dtype=torch.int32, device="cuda")
        col = torch.empty(n * w.value, dtype=torch.int32, device="cuda")
        val = torch.empty(n * w.value, dtype=torch.float64, device="cuda")
        _lib.call("ed_op_hop_table", self._h, sigma, C.byref(w), _lib.dptr(cnt), _lib.dptr(col),
                  _lib.dptr(val), _lib.stream_ptr())
        return (cnt.cpu().numpy(), col.cpu().numpy().reshape(n, w.value),
                val.cpu().numpy().reshape(n, w.value))

    # -- scipy LinearOperator protocol ---------------------------------------------
    def _matmat(self, X):
        X = np.asarray(X)
        if np.iscomplexobj(X):
            return self._matmat(X.real) + 1j * self._matmat(X.imag)
        xt = _as_device(np.ascontiguousarray(X.T))
        yt = self.matvec_device(xt)
        return yt.cpu().numpy().T.astype(np.result_type(X.dtype, np.float64), copy=False)

    def _matvec(self, x):
        x = np.asarray(x)
        shp = x.shape
        if np.iscomplexobj(x):
            return (self._matvec(x.real) + 1j * self._matvec(x.imag)).reshape(shp)
        y = self.matvec_device(_as_device(x.reshape(-1))).cpu().numpy()
        return y.reshape(shp)

    def _adjoint(self):
        return self

    def toarray(self) -> np.ndarray:
        D = self.size
        if D > 16384:
            raise ValueError(f"dense H of dimension {D} is too large")
        h = torch.empty(D * D, dtype=torch.float64, device="cuda")
        _lib.call("ed_op_dense", self._h, _lib.dptr(h), _lib.stream_ptr())
        return h.cpu().numpy().reshape(D, D)

    def _trace(self) -> float:
        return float(self.diagonal_device().sum().item())

    def __repr__(self):
        return (f"{self.__class__.__name__}(sector: ({self.n_up}, {self.n_dn}), "
                f"shape: {self.shape}, dtype: {self.dtype})")


class HamiltonOperator(LinearOperator):
    """H from COO triples (cmpy/operators.py:614-646).

    The triples are sorted by target once (duplicates kept, summed in stored
    order) and the sum ``y[col] += val * x[row]`` runs on the GPU as a
    row-per-lane CSR sweep (``ed_csr_matvec``).
    """

    def __init__(self, size, data, indices, dtype=None):
        data = np.asanyarray(data)
        indices = np.asanyarray(indices)
        if dtype is None:
            dtype = data.dtype if data.size else np.float64
        super().__init__((size, size), dtype=dtype)
        self.data = data
        self.indices = indices.T
        rows = indices[0].astype(np.int64) if indices.size else np.zeros(0, np.int64)
        cols = indices[1].astype(np.int64) if indices.size else np.zeros(0, np.int64)
        order = np.argsort(cols, kind="stable")
        self._ptr = np.zeros(size + 1, dtype=np.int64)
        np.add.at(self._ptr, cols + 1, 1)
        self._ptr = np.cumsum(self._ptr)
        self._src = rows[order]
        self._val = np.asarray(data, dtype=np.float64)[order] if data.size else np.zeros(0)
        self._dev = None

    def _device_tables(self):
        if self._dev is None:
            _lib.require_gpu()
            self._dev = (
                torch.from_numpy(self._ptr).cuda(),
                torch.from_numpy(self._src).cuda() if self._src.size else torch.zeros(1, dtype=torch.int64, device="cuda"),
                torch.from_numpy(self._val).cuda() if self._val.size else torch.zeros(1, dtype=torch.float64, device="cuda"),
            )
        return self._dev

    def _matvec(self, x):
        x = np.asarray(x)
        if np.iscomplexobj(x):
            return self._matvec(x.real) + 1j * self._matvec(x.imag)
        ptr, src, val = self._device_tables()
        xt = _as_device(x.reshape(-1))
        y = torch.empty(self.shape[0], dtype=torch.float64, device="cuda")
        _lib.call("ed_csr_matvec", _lib.dptr(ptr), _lib.dptr(src), _lib.dptr(val), self.shape[0],
                  _lib.dptr(xt), _lib.dptr(y), _lib.stream_ptr())
        return y.cpu().numpy().reshape(x.shape).astype(x.dtype, copy=False)

    def toarray(self):
        import scipy.sparse as sp

        return sp.csr_matrix((self.data, self.indices.T), shape=self.shape, dtype=self.dtype).toarray()

    def _adjoint(self):
        return self

    def _trace(self) -> float:
        idx = np.where(self.indices[:, 0] == self.indices[:, 1])[0]
        return float(np.sum(self.data[idx]))


class _LadderOperator(LinearOperator):
    _create = True

    def __init__(self, sector, sector_other, pos=0, sigma=UP, sign_mode="reference"):
        if sign_mode not in SIGN_MODES:
            raise ValueError(f"sign_mode must be one of {list(SIGN_MODES)}")
        dim_origin = sector.size
        if sigma == UP:
            dim_target = sector_other.num_up * sector.num_dn
        else:
            dim_target = sector_other.num_dn * sector.num_up
        super().__init__(shape=(dim_target, dim_origin), dtype=np.complex64)
        self.pos = pos
        self.sigma = sigma
        self.sector = sector
        self.sign_mode = sign_mode
        self.num_sites = sector.num_sites

    def __repr__(self):
        name = f"{self.__class__.__name__}_{self.pos}{SPIN_CHARS[self.sigma]}"
        return f"{name}(shape: {self.shape}, dtype: {self.dtype})"

    def apply_device(self, x: torch.Tensor) -> torch.Tensor:
        """Apply to CUDA float64 tensor (D,) or (k, D) (vectors as rows)."""
        nvec = 1 if x.dim() == 1 else x.shape[0]
        x = x.contiguous()
        out_shape = (self.shape[0],) if x.dim() == 1 else (nvec, self.shape[0])
        y = torch.empty(out_shape, dtype=torch.float64, device="cuda")
        fn = "ed_apply_cdag" if self._create else "ed_apply_c"
        _lib.call(fn, self.num_sites, self.sector.n_up, self.sector.n_dn, self.pos, self.sigma,
                  SIGN_MODES[self.sign_mode], _lib.dptr(x), _lib.dptr(y), nvec, _lib.stream_ptr())
        return y

    def _matmat(self, X):
        X = np.asarray(X)
        if np.iscomplexobj(X):
            return self._matmat(X.real) + 1j * self._matmat(X.imag)
        _lib.require_gpu()
        yt = self.apply_device(_as_device(np.ascontiguousarray(X.T)))
        return yt.cpu().numpy().T.astype(X.dtype, copy=False)

    def _matvec(self, x):
        x = np.asarray(x)
        if np.iscomplexobj(x):
            return self._matvec(x.real) + 1j * self._matvec(x.imag)
        _lib.require_gpu()
        y = self.apply_device(_as_device(x.reshape(-1))).cpu().numpy()
        return y.reshape((self.shape[0],) + x.shape[1:]).astype(x.dtype, copy=False)


class CreationOperator(_LadderOperator):
    """c^dagger_{pos,sigma}: sector -> sector_p1 (cmpy/operators.py:706-747).

    ``sign_mode="reference"`` copies blocks without a Jordan-Wigner sign, as cmpy
    does; ``"jordan_wigner"`` applies (-1)^(particles of the same spin below pos).
    """

    _create = True

    def __init__(self, sector, sector_p1, pos=0, sigma=UP, sign_mode="reference"):
        super().__init__(sector, sector_p1, pos, sigma, sign_mode)
        self.sector_p1 = sector_p1

    def _adjoint(self):
        return AnnihilationOperator(self.sector_p1, self.sector, self.pos, self.sigma, self.sign_mode)


class AnnihilationOperator(_LadderOperator):
    """c_{pos,sigma}: sector -> sector_m1 (cmpy/operators.py:750-791)."""

    _create = False

    def __init__(self, sector, sector_m1, pos=0, sigma=UP, sign_mode="reference"):
        super().__init__(sector, sector_m1, pos, sigma, sign_mode)
        self.sector_m1 = sector_m1

    def _adjoint(self):
        return CreationOperator(self.sector_m1, self.sector, self.pos, self.sigma, self.sign_mode)
