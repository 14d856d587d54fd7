"""Hubbard and Anderson-impurity models (cmpy/models/{abc,hubbard,anderson}.py).

Constructors, parameters and the ``hamiltonian_data`` / ``hamilton_operator`` /
``hamiltonian`` surface match cmpy.  ``hamilton_operator`` returns the
matrix-free :class:`~cmpy_amd.operators.SectorHamiltonOperator` (a scipy
LinearOperator, so ``eigsh``/``expm_multiply`` callers keep working) instead of
cmpy's COO ``HamiltonOperator``; ``hamiltonian_data`` still yields cmpy's exact
triples for callers that build their own COO operator.
"""

from __future__ import annotations

import json
from abc import ABC, abstractmethod
from collections import OrderedDict
from collections.abc import MutableMapping
from typing import Any, Dict, Iterator, List, Optional, Sequence, Union

import numpy as np

from .basis import UP, Basis
from .operators import (
    HamiltonOperator,
    SectorHamiltonOperator,
    project_hopping,
    project_hubbard_inter,
    project_onsite_energy,
)

__all__ = [
    "ModelParameters", "AbstractModel", "AbstractManyBodyModel", "HubbardModel",
    "SingleImpurityAndersonModel", "chain_neighbors", "square_neighbors",
]


class ModelParameters(MutableMapping):
    """Parameters accessible as attributes or items (cmpy/models/abc.py:21-133)."""

    def __init__(self, **params):
        object.__setattr__(self, "__params__", OrderedDict(params))

    @property
    def params(self) -> Dict[str, Any]:
        return self.__params__

    def set_param(self, key: str, value: Any) -> None:
        self.__params__[key] = value

    def delete_param(self, key: str) -> None:
        del self.__params__[key]

    def rename_param(self, key: str, new_key: str) -> None:
        self.__params__[new_key] = self.__params__.pop(key)

    def __len__(self) -> int:
        return len(self.__params__)

    def __getitem__(self, key):
        return self.__params__[key]

    def __setitem__(self, key, value):
        self.__params__[key] = value

    def __delitem__(self, key):
        del self.__params__[key]

    def __iter__(self) -> Iterator[str]:
        return iter(self.__params__)

    def __getattr__(self, key):
        params = object.__getattribute__(self, "__params__")
        if not key.startswith("__") and key in params:
            return params[key]
        return object.__getattribute__(self, key)

    def __setattr__(self, key, value):
        params = object.__getattribute__(self, "__params__")
        if not key.startswith("__") and key in params:
            params[key] = value
        else:
            object.__setattr__(self, key, value)

    def key(self, decimals=None, delim="; "):
        out = []
        for k, v in self.__params__.items():
            if decimals is not None and isinstance(v, (int, float)):
                v = f"{v:.{decimals}f}"
            out.append(f"{k}={v}")
        return delim.join(out)

    def json(self):
        return json.dumps({k: (v.tolist() if isinstance(v, np.ndarray) else v)
                           for k, v in self.__params__.items()})

    def pformat(self):
        return ", ".join(f"{k}={v}" for k, v in self.__params__.items())

    def __repr__(self) -> str:
        return f"{self.__class__.__name__}({dict(self.__params__)})"

    def __str__(self) -> str:
        return self.pformat()


class AbstractModel(ModelParameters, ABC):
    def __str__(self) -> str:
        return f"{self.__class__.__name__}({ModelParameters.__str__(self)})"


class AbstractManyBodyModel(AbstractModel):
    """Base class of models with a Fock basis (cmpy/models/abc.py:206-260).

    Subclasses provide ``_hamiltonian_data`` (cmpy's triple generator) and
    ``operator_spec`` (bonds, hops, eps, u, sign width) for the matrix-free
    device operator.
    """

    OPERATOR_CACHE_SIZE = 8  # sector operators kept per model (device tables of a few sectors)

    def __init__(self, num_sites: Optional[int] = 0, **params):
        super().__init__(**params)
        object.__setattr__(self, "basis", Basis())
        object.__setattr__(self, "_op_cache", OrderedDict())
        self.init_basis(num_sites)

    @property
    def num_sites(self) -> int:
        return self.basis.num_sites

    @property
    def fillings(self) -> List[int]:
        return self.basis.fillings

    def init_basis(self, num_sites, init_sectors=None):
        self.basis.init(num_sites, init_sectors)

    def iter_fillings(self):
        return self.basis.iter_fillings()

    def iter_sectors(self):
        return self.basis.iter_sectors()

    def get_sector(self, n_up=None, n_dn=None):
        return self.basis.get_sector(n_up, n_dn)

    @abstractmethod
    def _hamiltonian_data(self, up_states, dn_states):
        pass

    def operator_spec(self):
        """(bonds[nb, 2], hops[nb], eps[ns], u[ns], width) of the sector Hamiltonian."""
        raise NotImplementedError

    def hamiltonian_data(self, up_states, dn_states):
        """cmpy's (data, indices[2, nnz]) COO triples (cmpy/models/abc.py:242-248)."""
        rows, cols, data = [], [], []
        for row, col, val in self._hamiltonian_data(up_states, dn_states):
            rows.append(row)
            cols.append(col)
            data.append(val)
        return data, np.array([rows, cols], dtype=np.int64)

    def hamilton_operator(self, n_up=None, n_dn=None, sector=None, dtype=None):
        """Matrix-free device operator of one sector (cmpy/models/abc.py:250-256).

        The operator (its device tables) is cached per sector and reused while
        the model's Hamiltonian parameters are unchanged: the key holds their
        values, so a parameter update (or an in-place edit of a parameter
        array) builds a new one.
        """
        if sector is None:
            if n_up is None or n_dn is None:
                raise ValueError("hamilton_operator needs a fixed sector (n_up, n_dn)")
            sector = self.basis.get_sector(n_up, n_dn)
        bonds, hops, eps, u, width = self.operator_spec()
        key = (self.num_sites, sector.n_up, sector.n_dn, int(width),
               np.asarray(bonds, dtype=np.int64).tobytes(), np.asarray(hops, dtype=np.float64).tobytes(),
               np.asarray(eps, dtype=np.float64).tobytes(), np.asarray(u, dtype=np.float64).tobytes())
        cache = self._op_cache
        op = cache.get(key)
        if op is None:
            op = SectorHamiltonOperator(self.num_sites, sector.n_up, sector.n_dn, bonds, hops, eps, u, width)
            cache[key] = op
            while len(cache) > self.OPERATOR_CACHE_SIZE:
                cache.popitem(last=False)
        else:
            cache.move_to_end(key)
        return op

    def clear_operator_cache(self) -> None:
        """Drop the cached sector operators (and their device tables)."""
        self._op_cache.clear()

    def coo_operator(self, n_up=None, n_dn=None, sector=None, dtype=None):
        """cmpy's COO ``HamiltonOperator`` built from ``hamiltonian_data``."""
        if sector is None:
            sector = self.basis.get_sector(n_up, n_dn)
        up, dn = sector.up_states, sector.dn_states
        data, indices = self.hamiltonian_data(up, dn)
        return HamiltonOperator(len(up) * len(dn), data, indices, dtype=dtype)

    def hamiltonian(self, n_up=None, n_dn=None, sector=None, dtype=None):
        return self.hamilton_operator(n_up, n_dn, sector, dtype).toarray()


def chain_neighbors(num_sites: int, periodic: bool = False):
    """Bonds (i, i+1) of a chain, plus (0, N-1) when periodic (N > 2)."""
    bonds = [(i, i + 1) for i in range(num_sites - 1)]
    if periodic and num_sites > 2:
        bonds.append((0, num_sites - 1))
    return bonds


def square_neighbors(nx: int, ny: int, periodic: bool = True):
    """Nearest-neighbour bonds (i < j) of an nx x ny square lattice, site = x + nx*y."""
    bonds = set()
    for y in range(ny):
        for x in range(nx):
            i = x + nx * y
            for dx, dy in ((1, 0), (0, 1)):
                xx, yy = x + dx, y + dy
                if periodic:
                    xx, yy = xx % nx, yy % ny
                elif xx >= nx or yy >= ny:
                    continue
                j = xx + nx * yy
                if i != j:
                    bonds.add((min(i, j), max(i, j)))
    return sorted(bonds)


class HubbardModel(AbstractManyBodyModel):
    """Hubbard model (cmpy/models/hubbard.py:37-81).

    ``HubbardModel(num_sites, neighbors, inter=..., eps=..., hop=..., mu=...)`` or
    ``HubbardModel(lattice, ...)`` with a lattice exposing ``num_sites`` and
    ``neighbor_pairs(True)``.  Only bonds with i < j are used (hubbard.py:21) and
    hopping signs are fermionic (sign width = num_sites, hubbard.py:22).
    """

    def __init__(self, *args, inter=0.0, eps=0.0, hop=1.0, mu=0.0):
        if len(args) == 1:
            latt = args[0]
            num_sites = latt.num_sites
            neighbors = latt.neighbor_pairs(True)[0]
        else:
            num_sites, neighbors = args
        super().__init__(num_sites, inter=inter, eps=eps, hop=hop, mu=mu)
        object.__setattr__(self, "neighbors", [tuple(int(v) for v in b) for b in neighbors])

    def pformat(self):
        return f"U={self.inter}, ε={self.eps}, t={self.hop}, μ={self.mu}"

    def _bonds(self):
        return [(i, j) for i, j in self.neighbors if i < j]

    def _hamiltonian_data(self, up_states, dn_states):
        """cmpy/models/hubbard.py:13-22, 75-81."""
        ns = self.num_sites
        energy = np.broadcast_to(np.asarray(self.eps, dtype=np.float64) - self.mu, (ns,))
        interaction = np.broadcast_to(np.asarray(self.inter, dtype=np.float64), (ns,))
        yield from project_onsite_energy(up_states, dn_states, energy)
        yield from project_hubbard_inter(up_states, dn_states, interaction)
        hops = np.broadcast_to(np.asarray(self.hop, dtype=np.float64), (len(self._bonds()),))
        for (i, j), t in zip(self._bonds(), hops):
            yield from project_hopping(up_states, dn_states, ns, i, j, float(t))

    def operator_spec(self):
        ns = self.num_sites
        bonds = np.array(self._bonds(), dtype=np.int32).reshape(-1, 2)
        hops = np.broadcast_to(np.asarray(self.hop, dtype=np.float64), (len(bonds),))
        eps = np.broadcast_to(np.asarray(self.eps, dtype=np.float64) - self.mu, (ns,))
        u = np.broadcast_to(np.asarray(self.inter, dtype=np.float64), (ns,))
        return bonds, hops, eps, u, ns


class SingleImpurityAndersonModel(AbstractManyBodyModel):
    """Single-impurity Anderson model (cmpy/models/anderson.py:24-170).

    ``sign_mode="reference"`` reproduces cmpy's Hamiltonian, whose hopping signs
    are all +1 (it passes a sign width of 0, anderson.py:149);
    ``sign_mode="fermion"`` uses fermionic signs (width = num_sites).
    """

    def __init__(self, u=2.0, eps_imp=0.0, eps_bath=0.0, v=1.0, mu=None, temp=0.0,
                 sign_mode: str = "reference"):
        if sign_mode not in ("reference", "fermion"):
            raise ValueError("sign_mode must be 'reference' or 'fermion'")
        mu = u / 2 if mu is None else mu
        eps_bath = np.atleast_1d(np.asarray(eps_bath, dtype=np.float64))
        v = np.atleast_1d(np.asarray(v, dtype=np.float64))
        if len(eps_bath) > 1 and len(v) == 1:
            v = np.ones(len(eps_bath)) * v[0]
        if len(eps_bath) == 1 and len(v) > 1:
            eps_bath = np.ones(len(v)) * eps_bath[0]
        if len(eps_bath) != len(v):
            raise ValueError(f"Shape of bath on-site energy {len(eps_bath)} doesn't match "
                             f"hybridization {len(v)}!")
        super().__init__(len(eps_bath) + 1, u=u, eps_imp=eps_imp, eps_bath=eps_bath, v=v, mu=mu,
                         temp=temp)
        object.__setattr__(self, "sign_mode", sign_mode)

    @property
    def num_bath(self) -> int:
        return len(self.eps_bath)

    @property
    def num_sites(self) -> int:
        return self.num_bath + 1

    @property
    def beta(self) -> float:
        return 1 / self.temp

    def update_bath_energy(self, eps_bath) -> None:
        eps_bath = np.atleast_1d(eps_bath).astype(np.float64)
        if eps_bath.shape[0] != self.num_bath:
            raise ValueError(f"Dimension of the new bath energy {eps_bath.shape} does not match "
                             f"number of baths {self.num_bath}")
        self.eps_bath = eps_bath

    def update_hybridization(self, v) -> None:
        v = np.atleast_1d(v).astype(np.float64)
        if v.shape[0] != self.num_bath:
            raise ValueError(f"Dimension of the new hybridization {v.shape} does not match "
                             f"number of baths {self.num_bath}")
        self.v = v

    def hybridization_func(self, z: np.ndarray) -> np.ndarray:
        """Delta(z) = sum_k |V_k|^2 / (z - eps_k) (cmpy/models/anderson.py:127-145)."""
        x = np.asarray(z)[..., np.newaxis]
        return np.sum(np.square(np.abs(self.v)) / (x - self.eps_bath), axis=-1)

    def _sign_width(self) -> int:
        return 0 if self.sign_mode == "reference" else self.num_sites

    def _hamiltonian_data(self, up_states, dn_states):
        """cmpy/models/anderson.py:147-158."""
        u = np.append(self.u, np.zeros(self.num_bath))
        eps = np.append(self.eps_imp - self.mu, self.eps_bath)
        yield from project_onsite_energy(up_states, dn_states, eps)
        yield from project_hubbard_inter(up_states, dn_states, u)
        for j in range(self.num_bath):
            yield from project_hopping(up_states, dn_states, self._sign_width(), 0, j + 1,
                                       float(self.v[j]))

    def operator_spec(self):
        nb = self.num_bath
        bonds = np.array([(0, j + 1) for j in range(nb)], dtype=np.int32).reshape(-1, 2)
        eps = np.append(self.eps_imp - self.mu, self.eps_bath).astype(np.float64)
        u = np.append(self.u, np.zeros(nb)).astype(np.float64)
        return bonds, np.asarray(self.v, dtype=np.float64), eps, u, self._sign_width()

    def impurity_gf0(self, z):
        """cmpy/models/anderson.py:160-161."""
        return 1 / (z + self.mu + self.eps_imp - self.hybridization_func(z))

    def impurity_gf(self, z, sigma=UP):
        """Finite-temperature Lehmann G of the impurity (cmpy/models/anderson.py:163-164)."""
        from .exactdiag import gf_lehmann

        return gf_lehmann(self, z, beta=1 / self.temp, pos=0, sigma=sigma).gf

    def pformat(self):
        return (f"U={self.u}, ε_i={self.eps_imp}, ε_b={self.eps_bath}, v={self.v}, "
                f"μ={self.mu}, T={self.temp}")
