"""Fock basis and (n_up, n_dn) sectors.

Same public surface and semantics as cmpy's ``cmpy/basis.py`` (UP/DN constants,
bit helpers, ``Sector``, ``Basis.get_sector``), so callers can switch imports.
Sector state lists come from the native library (``ed_sector_states_host``,
next-combination enumeration) instead of cmpy's permutation enumeration
(cmpy/basis.py:655-666, O(num_sites!)); the output is the same ascending list.
"""

from __future__ import annotations

import ctypes as C
from collections import defaultdict
from dataclasses import dataclass
from itertools import product
from typing import Iterable, List, Optional, Tuple, Union

import numpy as np

from . import _lib

__all__ = [
    "UP", "DN", "SPIN_CHARS", "state_label", "bit_count", "binstr", "binarr", "binidx",
    "get_ibit", "set_ibit", "overlap", "occupations", "create", "annihilate", "SpinState",
    "State", "upper_sector", "lower_sector", "Sector", "Basis", "sector_states",
]

UP, DN = 1, 2  # cmpy/basis.py:40

EMPTY_CHAR, UP_CHAR, DN_CHAR, UD_CHAR = ".", "↑", "↓", "⇅"
SPIN_CHARS = {0: EMPTY_CHAR, UP: UP_CHAR, DN: DN_CHAR, 3: UD_CHAR}


def state_label(up_num: int, dn_num: int, digits: int = None) -> str:
    """Label of a (up, dn) state, one character per site, site 0 first
    (cmpy/basis.py:49-74)."""
    n = max(len(bin(up_num)) - 2, len(bin(dn_num)) - 2, digits or 0)
    return "".join(SPIN_CHARS[((up_num >> i) & 1) | (((dn_num >> i) & 1) << 1)] for i in range(n))


def bit_count(num: int) -> int:
    return bin(int(num)).count("1")


def binstr(num: int, width: int = 0) -> str:
    return format(int(num), f"0{width or 0}b")


def binarr(num: int, width: int = None, dtype=None) -> np.ndarray:
    """Bits of ``num`` with index 0 first (cmpy/basis.py:127-166)."""
    s = binstr(num, width or 0)
    arr = np.array([int(c) for c in reversed(s)], dtype=np.int64)
    return arr.astype(dtype) if dtype is not None else arr


def binidx(num, width: int = None) -> List[int]:
    return [i for i, b in enumerate(binarr(num, width)) if b]


def get_ibit(num: int, index: int, length: int = 1) -> int:
    return (num >> index) & ((1 << length) - 1)


def set_ibit(num: int, index: int, value: int, length: int = 1) -> int:
    mask = ((1 << length) - 1) << index
    return (num & ~mask) | ((value << index) & mask)


def overlap(num1: int, num2: int, width: int = None, dtype=None) -> np.ndarray:
    return binarr(num1 & num2, width, dtype)


def occupations(num: int, width: int = None, dtype=None) -> np.ndarray:
    return binarr(num, width, dtype)


def create(num: int, pos: int) -> Optional[int]:
    """State with a particle added at ``pos``, or None (cmpy/basis.py:346-379)."""
    bit = 1 << pos
    return None if num & bit else num | bit


def annihilate(num: int, pos: int) -> Optional[int]:
    """State with the particle at ``pos`` removed, or None (cmpy/basis.py:382-416)."""
    bit = 1 << pos
    return num & ~bit if num & bit else None


class SpinState(int):
    @property
    def n(self) -> int:
        return bit_count(self)

    def binstr(self, width: int = None) -> str:
        return binstr(self, width)

    def binarr(self, width: int = None, dtype=None) -> np.ndarray:
        return binarr(self, width, dtype)

    def occ(self, pos: int) -> int:
        return self & (1 << pos)

    def occupations(self, dtype=None) -> np.ndarray:
        return binarr(self, dtype=dtype)

    def overlap(self, other, dtype=None) -> np.ndarray:
        return overlap(self, other, dtype=dtype)

    def create(self, pos: int):
        new = create(self, pos)
        return None if new is None else self.__class__(new)

    def annihilate(self, pos: int):
        new = annihilate(self, pos)
        return None if new is None else self.__class__(new)

    def __repr__(self) -> str:
        return f"{self.__class__.__name__}({self.binstr()})"

    def __str__(self) -> str:
        return self.binstr()


@dataclass
class State:
    __slots__ = ["up", "dn", "num_sites"]

    def __init__(self, up, dn, num_sites: int = None):
        self.up = SpinState(up)
        self.dn = SpinState(dn)
        self.num_sites = num_sites

    def label(self, width: int = None) -> str:
        return state_label(self.up, self.dn, width if width is not None else self.num_sites)

    def __repr__(self) -> str:
        return f"{self.__class__.__name__}({self.up}, {self.dn})"

    def __str__(self) -> str:
        return f"{self.__class__.__name__}: {self.label()}"

    def __eq__(self, other) -> bool:
        return self.up == other.up and self.dn == other.dn


def upper_sector(n_up: int, n_dn: int, sigma: int, num_sites: int) -> Optional[Tuple[int, int]]:
    """(n_up+1, n_dn) / (n_up, n_dn+1), or None at the top (cmpy/basis.py:504-526)."""
    if sigma == UP and n_up < num_sites:
        return n_up + 1, n_dn
    if sigma == DN and n_dn < num_sites:
        return n_up, n_dn + 1
    return None


def lower_sector(n_up: int, n_dn: int, sigma: int) -> Optional[Tuple[int, int]]:
    """(n_up-1, n_dn) / (n_up, n_dn-1), or None at zero (cmpy/basis.py:529-553)."""
    if sigma == UP and n_up > 0:
        return n_up - 1, n_dn
    if sigma == DN and n_dn > 0:
        return n_up, n_dn - 1
    return None


def sector_states(num_sites: int, n: int) -> np.ndarray:
    """Ascending num_sites-bit states with popcount n (native enumeration)."""
    count = int(_lib.lib().ed_binom(num_sites, n))
    out = np.zeros(max(count, 1), dtype=np.uint32)
    _lib.call("ed_sector_states_host", num_sites, n, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out[:count].astype(np.int64)


class Sector:
    """Up/dn state lists of one sector; full index k = iu * num_dn + id."""

    __slots__ = ["num_sites", "n_up", "n_dn", "up_states", "dn_states"]

    def __init__(self, up_states, dn_states, n_up=None, n_dn=None, num_sites=0):
        self.num_sites = num_sites
        self.n_up = n_up
        self.n_dn = n_dn
        self.up_states = up_states
        self.dn_states = dn_states

    @property
    def states(self):
        for up, dn in product(self.up_states, self.dn_states):
            yield State(up, dn, num_sites=self.num_sites)

    @property
    def size(self) -> int:
        return len(self.up_states) * len(self.dn_states)

    @property
    def num_up(self) -> int:
        return len(self.up_states)

    @property
    def num_dn(self) -> int:
        return len(self.dn_states)

    @property
    def filling(self):
        return self.n_up, self.n_dn

    def state_labels(self):
        return [s.label(self.num_sites) for s in self.states]

    def __repr__(self):
        return (f"{self.__class__.__name__}(size: {self.size}, num_sites: {self.num_sites}, "
                f"filling: [{self.n_up}, {self.n_dn}])")

    def __str__(self):
        return f"{self.__class__.__name__}({self.n_up}, {self.n_dn}, size: {self.size})"

    def __iter__(self):
        return iter(self.states)


class Basis:
    """All sectors of a num_sites Hubbard-type Fock space (cmpy/basis.py:606-718)."""

    __slots__ = ["size", "num_sites", "num_spinstates", "sectors", "fillings"]

    def __init__(self, num_sites: int = 0, init_sectors: bool = False):
        self.size = 0
        self.num_sites = 0
        self.num_spinstates = 0
        self.sectors = defaultdict(list)
        self.fillings = []
        self.init(num_sites, init_sectors)

    def init(self, num_sites: int, init_sectors: bool = False):
        if num_sites < 0 or num_sites > 30:
            raise ValueError(f"num_sites={num_sites} outside [0, 30]")
        self.num_sites = num_sites
        self.num_spinstates = 2**num_sites
        self.size = self.num_spinstates**2
        self.sectors = defaultdict(list)
        self.fillings = list(range(num_sites + 1))
        if init_sectors:
            for n in range(num_sites + 1):
                self.sectors[n] = [int(s) for s in sector_states(num_sites, n)]

    def generate_states(self, n: int = None) -> Union[list, np.ndarray]:
        """Same return types as the reference: list for n in {None, 0, 1}, else ndarray."""
        if n is None:
            return list(range(2**self.num_sites))
        if n == 0:
            return [0]
        if n == 1:
            return [1 << b for b in range(self.num_sites)]
        return sector_states(self.num_sites, n)

    def get_states(self, n: int = None):
        if n in self.sectors:
            return self.sectors.get(n, list(range(self.num_spinstates)))
        states = self.generate_states(n)
        self.sectors[n] = states
        return states

    def get_sector(self, n_up: int = None, n_dn: int = None) -> Sector:
        return Sector(self.get_states(n_up), self.get_states(n_dn), n_up, n_dn, self.num_sites)

    def iter_fillings(self):
        return product(self.fillings, repeat=2)

    def iter_sectors(self):
        for n_up, n_dn in product(self.fillings, repeat=2):
            yield self.get_sector(n_up, n_dn)

    def keys(self):
        return self.fillings

    def check(self, n_up, n_dn):
        return n_up in self.fillings and n_dn in self.fillings

    def upper_sector(self, n_up, n_dn, sigma):
        f = upper_sector(n_up, n_dn, sigma, self.num_sites)
        return self.get_sector(*f) if f is not None else None

    def lower_sector(self, n_up, n_dn, sigma):
        f = lower_sector(n_up, n_dn, sigma)
        return self.get_sector(*f) if f is not None else None

    def __getitem__(self, item):
        if hasattr(item, "__len__"):
            return self.get_sector(*item)
        return self.get_states(item)

    def __repr__(self):
        return (f"{self.__class__.__name__}(size: {self.size}, num_sites: {self.num_sites}, "
                f"fillings: {self.fillings})")
