"""CPU oracle for the exact-diagonalisation hot path.

TEST INFRASTRUCTURE ONLY.  This module restates, in numpy/scipy, the algorithm
that cmpy (dylanljones/cmpy) uses for the Hubbard/Anderson sector Hamiltonian,
the ground state, the Lanczos recurrence and the Lehmann Green's function.  It
is the checker: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product package
``cmpy_amd`` never imports, calls or links anything under ``oracle/``.

Every function cites the reference file:line it follows (paths relative to
the reference checkout).  How each one is pinned (checked against something
other than a retyping of its own formula) is recorded in ``PARITY`` at the
bottom of this file; the test-suite exercises every entry.

Conventions carried over from the reference:
  * sector index  k = iu * num_dn + id  (up-major, cmpy/basis.py:570-575)
  * spin states in ascending integer order (cmpy/basis.py:655-666)
  * hopping sign (-1)^popcount(state & between(site1, site2) & lowmask(width))
    with width = num_sites for the Hubbard model and width = 0 (all signs +1)
    for the Anderson impurity model (cmpy/models/anderson.py:149)
  * creation/annihilation operators copy blocks without a Jordan-Wigner sign
    (cmpy/operators.py:652-703); ``sign_mode="jordan_wigner"`` adds the
    within-spin string as an opt-in
"""

from __future__ import annotations

import itertools
import math
from typing import Iterable, List, Sequence, Tuple

import numpy as np
import scipy.linalg as la
import scipy.sparse as sp
import scipy.sparse.linalg as sla

UP, DN = 1, 2  # cmpy/basis.py:40


# =============================================================================
# Basis
# =============================================================================


def binom(n: int, k: int) -> int:
    if k < 0 or k > n or n < 0:
        return 0
    return math.comb(n, k)


def generate_states_permutations(num_sites: int, n):
    """Literal restatement of ``Basis.generate_states`` (cmpy/basis.py:655-666).

    Enumerates permutations of a bit list, O(num_sites!) -- only usable for
    num_sites <= 9.  Kept to pin :func:`sector_states` against the reference's
    own enumeration.
    """
    total = num_sites
    if n is None:
        return list(range(2**total))
    elif n == 0:
        return [0]
    elif n == 1:
        return list(2**b for b in range(total))
    bitvals = ["0" for _ in range(total - n)] + ["1" for _ in range(n)]
    states = set(int("".join(bits), 2) for bits in itertools.permutations(bitvals))
    return np.asarray(sorted(states))


def popcount(a) -> np.ndarray:
    a = np.asarray(a, dtype=np.uint64)
    c = np.zeros(a.shape, dtype=np.int64)
    x = a.copy()
    while np.any(x):
        c += (x & np.uint64(1)).astype(np.int64)
        x >>= np.uint64(1)
    return c


def sector_states(num_sites: int, n) -> np.ndarray:
    """All num_sites-bit integers with popcount n, ascending.

    Same output contract as ``Basis.get_states`` (cmpy/basis.py:655-676), computed
    by filtering ``range(2**num_sites)`` instead of enumerating permutations.
    """
    if n is None:
        return np.arange(2**num_sites, dtype=np.int64)
    if n < 0 or n > num_sites:
        return np.zeros(0, dtype=np.int64)
    allst = np.arange(2**num_sites, dtype=np.int64)
    return allst[popcount(allst) == n]


def combinadic_rank(states: Sequence[int]) -> np.ndarray:
    """r(s) = sum_k C(p_k, k) over set bits p_1 < p_2 < ... (k is 1-based).

    Equals the position of s in the ascending list of states with the same
    popcount; used as the index lookup that replaces ``bisect_left``
    (cmpy/operators.py:276-299).
    """
    out = np.zeros(len(states), dtype=np.int64)
    for idx, s in enumerate(np.asarray(states, dtype=np.int64)):
        s = int(s)
        r, k, p = 0, 0, 0
        while s:
            if s & 1:
                k += 1
                r += binom(p, k)
            s >>= 1
            p += 1
        out[idx] = r
    return out


def upper_sector(n_up, n_dn, sigma, num_sites):
    """cmpy/basis.py:504-526."""
    if sigma == UP and n_up < num_sites:
        return n_up + 1, n_dn
    elif sigma == DN and n_dn < num_sites:
        return n_up, n_dn + 1
    return None


def lower_sector(n_up, n_dn, sigma):
    """cmpy/basis.py:529-553."""
    if sigma == UP and n_up > 0:
        return n_up - 1, n_dn
    elif sigma == DN and n_dn > 0:
        return n_up, n_dn - 1
    return None


def state_label(up_num: int, dn_num: int, digits: int = None) -> str:
    """cmpy/basis.py:49-74 (one character per site, site 0 first)."""
    chars_map = {0: ".", 1: "↑", 2: "↓", 3: "⇅"}
    digits = digits if digits is not None else 0
    min_digits = max(len(bin(up_num)[2:]), len(bin(dn_num)[2:]))
    num_chars = max(min_digits, digits)
    chars = []
    for i in range(num_chars):
        u = up_num >> i & 1
        d = dn_num >> i & 1
        chars.append(chars_map[u + (d << 1)])
    return "".join(chars[::1])


# =============================================================================
# Projectors (integer parts and FP64 matrix elements)
# =============================================================================


def bit_count_width(number: int, width: int) -> int:
    """cmpy/operators.py:253-273: popcount of the low ``width`` bits."""
    count = 0
    for i in range(width):
        if number & (1 << i):
            count += 1
    return count


def hopping_sign(initial_state: int, width: int, site1: int, site2: int) -> int:
    """cmpy/operators.py:425-433."""
    mask = 0
    for i in range(site1 + 1, site2):
        mask += 1 << i
    jump_overs = bit_count_width(initial_state & mask, width)
    return (-1) ** jump_overs


def hopping_terms(states, width: int, site1: int, site2: int, hop: float):
    """Vectorised ``_compute_hopping_term`` (cmpy/operators.py:436-460).

    Returns (origin, target, value) arrays in the reference's yield order
    (ascending origin).  ``target`` is found with a left bisection of the sorted
    state list, as the reference does (operators.py:276-299, 454/459).
    """
    assert site1 < site2
    st = np.asarray(states, dtype=np.int64)
    op1 = np.int64(1 << site1)
    op2 = np.int64(1 << site2)
    occ1 = (st & op1) != 0
    occ2 = (st & op2) != 0
    new = st ^ op1 ^ op2
    ok = occ1 ^ occ2
    mask = 0
    for i in range(site1 + 1, site2):
        mask += 1 << i
    mask &= (1 << max(width, 0)) - 1 if width > 0 else 0
    jumps = popcount(st & np.int64(mask))
    sign = np.where(jumps % 2 == 0, 1, -1)
    origin = np.nonzero(ok)[0].astype(np.int64)
    target = np.searchsorted(st, new[ok], side="left").astype(np.int64)
    value = (sign[ok] * hop).astype(np.float64)
    return origin, target, value


def weighted_element(states, values) -> np.ndarray:
    """cmpy/operators.py:226-250, summed site by site in the reference's order."""
    st = np.asarray(states, dtype=np.int64)
    acc = np.zeros(st.shape, dtype=np.float64)
    for i in range(len(values)):
        bit = (st & np.int64(1 << i)) != 0
        acc = np.where(bit, acc + float(values[i]), acc)
    return acc


def onsite_energy_triples(up_states, dn_states, eps):
    """cmpy/operators.py:359-422.  Only nonzero energies are emitted (409, 419)."""
    up = np.asarray(up_states, dtype=np.int64)
    dn = np.asarray(dn_states, dtype=np.int64)
    num_up, num_dn = len(up), len(dn)
    e_up = weighted_element(up, eps)
    e_dn = weighted_element(dn, eps)
    rows = []
    vals = []
    iu = np.nonzero(e_up)[0]
    r = (iu[:, None] * num_dn + np.arange(num_dn)[None, :]).ravel()
    rows.append(r)
    vals.append(np.repeat(e_up[iu], num_dn))
    idn = np.nonzero(e_dn)[0]
    r = (np.arange(num_up)[None, :] * num_dn + idn[:, None]).ravel()
    rows.append(r)
    vals.append(np.repeat(e_dn[idn], num_up))
    rows = np.concatenate(rows).astype(np.int64)
    return rows, rows.copy(), np.concatenate(vals)


def hubbard_inter_triples(up_states, dn_states, u):
    """cmpy/operators.py:305-356.  Only nonzero energies are emitted (354)."""
    up = np.asarray(up_states, dtype=np.int64)
    dn = np.asarray(dn_states, dtype=np.int64)
    num_dn = len(dn)
    both = (up[:, None] & dn[None, :]).ravel()
    energy = weighted_element(both, u)
    idx = np.nonzero(energy)[0].astype(np.int64)
    return idx, idx.copy(), energy[idx]


def hopping_triples(up_states, dn_states, num_sites, site1, site2, hop):
    """cmpy/operators.py:463-527 (up block T_up x I, then dn block I x T_dn)."""
    up = np.asarray(up_states, dtype=np.int64)
    dn = np.asarray(dn_states, dtype=np.int64)
    num_up, num_dn = len(up), len(dn)
    all_up, all_dn = np.arange(num_up), np.arange(num_dn)
    o, t, a = hopping_terms(up, num_sites, site1, site2, hop)
    r1 = (o[:, None] * num_dn + all_dn[None, :]).ravel()
    c1 = (t[:, None] * num_dn + all_dn[None, :]).ravel()
    v1 = np.repeat(a, num_dn)
    o, t, a = hopping_terms(dn, num_sites, site1, site2, hop)
    r2 = (all_up[None, :] * num_dn + o[:, None]).ravel()
    c2 = (all_up[None, :] * num_dn + t[:, None]).ravel()
    v2 = np.repeat(a, num_up)
    return (
        np.concatenate([r1, r2]).astype(np.int64),
        np.concatenate([c1, c2]).astype(np.int64),
        np.concatenate([v1, v2]).astype(np.float64),
    )


def _concat(parts):
    rows = np.concatenate([p[0] for p in parts]) if parts else np.zeros(0, np.int64)
    cols = np.concatenate([p[1] for p in parts]) if parts else np.zeros(0, np.int64)
    vals = np.concatenate([p[2] for p in parts]) if parts else np.zeros(0)
    return rows.astype(np.int64), cols.astype(np.int64), vals.astype(np.float64)


def hubbard_triples(up_states, dn_states, num_sites, neighbors, inter, eps, hop, mu=0.0):
    """``HubbardModel._hamiltonian_data`` (cmpy/models/hubbard.py:13-22, 75-81)."""
    energy = np.full(num_sites, eps - mu, dtype=np.float64)
    interaction = np.full(num_sites, inter, dtype=np.float64)
    parts = [
        onsite_energy_triples(up_states, dn_states, energy),
        hubbard_inter_triples(up_states, dn_states, interaction),
    ]
    for i, j in neighbors:
        if i < j:
            parts.append(hopping_triples(up_states, dn_states, num_sites, i, j, hop))
    return _concat(parts)


def siam_triples(up_states, dn_states, u, eps_imp, eps_bath, v, mu):
    """``SingleImpurityAndersonModel._hamiltonian_data`` (anderson.py:147-158).

    Note ``nsites = 0`` (anderson.py:149): the hopping-sign width is zero, so
    every bath hopping carries sign +1.
    """
    eps_bath = np.atleast_1d(np.asarray(eps_bath, dtype=np.float64))
    v = np.atleast_1d(np.asarray(v, dtype=np.float64))
    num_bath = len(eps_bath)
    nsites = 0
    uu = np.append(np.float64(u), np.zeros(num_bath))
    eps = np.append(np.float64(eps_imp) - mu, eps_bath)
    parts = [
        onsite_energy_triples(up_states, dn_states, eps),
        hubbard_inter_triples(up_states, dn_states, uu),
    ]
    for j in range(num_bath):
        parts.append(hopping_triples(up_states, dn_states, nsites, 0, j + 1, v[j]))
    return _concat(parts)


# =============================================================================
# COO Hamilton operator (reference semantics)
# =============================================================================


def coo_matvec(size, rows, cols, vals, x):
    """``HamiltonOperator._matvec`` (cmpy/operators.py:626-630):
    ``matvec[col] += val * x[row]`` for every stored triple."""
    x = np.asarray(x)
    y = np.zeros_like(x)
    np.add.at(y, cols, vals * x[rows])
    return y


def coo_dense(size, rows, cols, vals):
    """``HamiltonOperator.toarray`` (cmpy/operators.py:632-635): csr sums duplicates."""
    return sp.csr_matrix((vals, (rows, cols)), shape=(size, size)).toarray()


def coo_csr(size, rows, cols, vals):
    return sp.csr_matrix((vals, (rows, cols)), shape=(size, size))


def coo_trace(rows, cols, vals):
    """cmpy/operators.py:641-646."""
    return float(np.sum(vals[rows == cols]))


# =============================================================================
# Matrix-free factorised form used for larger sectors
# =============================================================================


class SectorOperator:
    """H = diag + T_up (x) I + I (x) T_dn for one (n_up, n_dn) sector.

    Built from the same projector restatements as the COO assembly above
    (``hopping_terms``, ``weighted_element``); pinned to it at small sizes by the
    test-suite.  ``matvec`` is exact up to floating-point summation order.
    """

    def __init__(self, num_sites, n_up, n_dn, bonds, hops, eps, u, width):
        self.num_sites = num_sites
        self.n_up, self.n_dn = n_up, n_dn
        self.up = sector_states(num_sites, n_up)
        self.dn = sector_states(num_sites, n_dn)
        self.num_up, self.num_dn = len(self.up), len(self.dn)
        self.size = self.num_up * self.num_dn
        eps = np.asarray(eps, dtype=np.float64)
        u = np.asarray(u, dtype=np.float64)
        self.d_up = weighted_element(self.up, eps)
        self.d_dn = weighted_element(self.dn, eps)
        self.u = u
        self.t_up = self._spin_matrix(self.up, bonds, hops, width)
        self.t_dn = self._spin_matrix(self.dn, bonds, hops, width)

    @staticmethod
    def _spin_matrix(states, bonds, hops, width):
        n = len(states)
        rows, cols, vals = [], [], []
        for (i, j), t in zip(bonds, hops):
            o, tg, a = hopping_terms(states, width, int(i), int(j), float(t))
            rows.append(o)
            cols.append(tg)
            vals.append(a)
        if rows:
            rows = np.concatenate(rows)
            cols = np.concatenate(cols)
            vals = np.concatenate(vals)
        else:
            rows = cols = np.zeros(0, np.int64)
            vals = np.zeros(0)
        # y[target] += val * x[origin]  ->  matrix element M[target, origin]
        return sp.csr_matrix((vals, (cols, rows)), shape=(n, n))

    def diagonal(self) -> np.ndarray:
        both = (self.up[:, None] & self.dn[None, :]).ravel()
        ud = weighted_element(both, self.u)
        return (self.d_up[:, None] + self.d_dn[None, :]).ravel() + ud

    def matvec(self, x):
        x = np.asarray(x)
        X = x.reshape(self.num_up, self.num_dn)
        Y = self.t_up @ X + (self.t_dn @ X.T).T
        return self.diagonal() * x + Y.ravel()

    def csr(self):
        n_up, n_dn = self.num_up, self.num_dn
        h = sp.kron(self.t_up, sp.identity(n_dn)) + sp.kron(sp.identity(n_up), self.t_dn)
        return (h + sp.diags(self.diagonal())).tocsr()


def hubbard_operator(num_sites, n_up, n_dn, neighbors, inter, eps, hop, mu=0.0):
    bonds = [(int(i), int(j)) for i, j in neighbors if i < j]
    return SectorOperator(
        num_sites, n_up, n_dn, bonds, [hop] * len(bonds),
        np.full(num_sites, eps - mu), np.full(num_sites, inter), num_sites,
    )


def siam_operator(n_up, n_dn, u, eps_imp, eps_bath, v, mu, width=0):
    eps_bath = np.atleast_1d(eps_bath).astype(np.float64)
    v = np.atleast_1d(v).astype(np.float64)
    nb = len(eps_bath)
    bonds = [(0, j + 1) for j in range(nb)]
    eps = np.append(eps_imp - mu, eps_bath)
    uu = np.append(u, np.zeros(nb))
    return SectorOperator(nb + 1, n_up, n_dn, bonds, list(v), eps, uu, width)


# =============================================================================
# Creation / annihilation operators
# =============================================================================


def _jw_sign(states, pos):
    low = (1 << pos) - 1
    return np.where(popcount(np.asarray(states, np.int64) & low) % 2 == 0, 1.0, -1.0)


def apply_cdag(x, num_sites, n_up, n_dn, pos, sigma, sign_mode="reference"):
    """c^dagger_{pos,sigma} from sector (n_up, n_dn) to the upper sector.

    cmpy/operators.py:652-676 (block copies, no sign) and 706-747.  With
    ``sign_mode="jordan_wigner"`` the within-spin string (-1)^popcount(s & (2^pos-1))
    is applied (not part of the reference; identical at pos = 0).
    """
    x = np.asarray(x)
    tail = x.shape[1:]  # a block of vectors (columns) is mapped column by column
    up = sector_states(num_sites, n_up)
    dn = sector_states(num_sites, n_dn)
    num_up, num_dn = len(up), len(dn)
    X = x.reshape((num_up, num_dn) + tail)
    op = 1 << pos
    ones = (1,) * len(tail)
    if sigma == UP:
        up1 = sector_states(num_sites, n_up + 1)
        Y = np.zeros((len(up1), num_dn) + tail, dtype=x.dtype)
        src = np.nonzero((up & op) == 0)[0]
        tgt = np.searchsorted(up1, up[src] ^ op)
        s = _jw_sign(up[src], pos) if sign_mode == "jordan_wigner" else np.ones(len(src))
        Y[tgt] = s.reshape((-1, 1) + ones) * X[src]
    else:
        dn1 = sector_states(num_sites, n_dn + 1)
        Y = np.zeros((num_up, len(dn1)) + tail, dtype=x.dtype)
        src = np.nonzero((dn & op) == 0)[0]
        tgt = np.searchsorted(dn1, dn[src] ^ op)
        s = _jw_sign(dn[src], pos) if sign_mode == "jordan_wigner" else np.ones(len(src))
        Y[:, tgt] = X[:, src] * s.reshape((1, -1) + ones)
    return Y.reshape((-1,) + tail)


def apply_c(x, num_sites, n_up, n_dn, pos, sigma, sign_mode="reference"):
    """c_{pos,sigma} from sector (n_up, n_dn) to the lower sector.

    cmpy/operators.py:679-703 and 750-791.
    """
    x = np.asarray(x)
    tail = x.shape[1:]  # a block of vectors (columns) is mapped column by column
    up = sector_states(num_sites, n_up)
    dn = sector_states(num_sites, n_dn)
    num_up, num_dn = len(up), len(dn)
    X = x.reshape((num_up, num_dn) + tail)
    op = 1 << pos
    ones = (1,) * len(tail)
    if sigma == UP:
        up1 = sector_states(num_sites, n_up - 1)
        Y = np.zeros((len(up1), num_dn) + tail, dtype=x.dtype)
        src = np.nonzero((up & op) != 0)[0]
        tgt = np.searchsorted(up1, up[src] ^ op)
        s = _jw_sign(up[src], pos) if sign_mode == "jordan_wigner" else np.ones(len(src))
        Y[tgt] = s.reshape((-1, 1) + ones) * X[src]
    else:
        dn1 = sector_states(num_sites, n_dn - 1)
        Y = np.zeros((num_up, len(dn1)) + tail, dtype=x.dtype)
        src = np.nonzero((dn & op) != 0)[0]
        tgt = np.searchsorted(dn1, dn[src] ^ op)
        s = _jw_sign(dn[src], pos) if sign_mode == "jordan_wigner" else np.ones(len(src))
        Y[:, tgt] = X[:, src] * s.reshape((1, -1) + ones)
    return Y.reshape((-1,) + tail)


# =============================================================================
# Ground state and Lanczos
# =============================================================================


def groundstate(op, tol=0.0, v0=None):
    """One-sector ``compute_groundstate`` (cmpy/exactdiag.py:24-43).

    Dense ``eigh`` for small sectors (<= 50, exactdiag.py:31-35; extended here to
    2000 for accuracy), otherwise ARPACK ``eigsh(k=1, which="SA")`` (37).
    """
    if op.size <= 2000:
        h = op.csr().toarray() if hasattr(op, "csr") else np.asarray(op)
        w, v = np.linalg.eigh(h)
        return float(w[0]), v[:, 0]
    lin = sla.LinearOperator((op.size, op.size), matvec=op.matvec, dtype=np.float64)
    w, v = sla.eigsh(lin, k=1, which="SA", tol=tol, v0=v0)
    return float(w[0]), v[:, 0]


def lanczos_coeffs_reference(matvec, psi, size):
    """The reference's unnormalised recurrence (cmpy/exactdiag.py:324-358) with a
    caller-supplied start vector instead of the unseeded ``np.random`` draw (326).

    a_n = <p|Hp>/<p|p>,  b_n^2 = |p_n|^2/|p_{n-1}|^2,
    p_{n+1} = H p_n - a_n p_n - b_n^2 p_{n-1}.  Returns (a[size], b[size-1]).
    """
    psi = np.asarray(psi, dtype=np.float64)
    a_list, b_list = [], []
    hp = matvec(psi)
    a = np.dot(psi, hp) / np.dot(psi, psi)
    a_list.append(a)
    psi_new = hp - a * psi
    psi_prev, psi = psi, psi_new
    for _ in range(1, size):
        hp = matvec(psi)
        a = np.dot(psi, hp) / np.dot(psi, psi)
        b2 = np.dot(psi, psi) / np.dot(psi_prev, psi_prev)
        psi_new = hp - a * psi - b2 * psi_prev
        a_list.append(a)
        b_list.append(np.sqrt(b2))
        psi_prev, psi = psi, psi_new
    return np.array(a_list), np.array(b_list)


def lanczos_ground_state(a_coeffs, b_coeffs, max_eig=3):
    """cmpy/exactdiag.py:368-375 (``eigh_tridiagonal`` with ``select="i"``)."""
    max_eig = min(max_eig, len(a_coeffs) - 1)
    xi, vi = la.eigh_tridiagonal(a_coeffs, b_coeffs, select="i", select_range=(0, max_eig))
    idx = np.argmin(xi)
    return xi[idx], vi[:, idx]


def lanczos_normalized(matvec, v0, nsteps, reorth=False):
    """Normalised Lanczos, one matvec per step.  Same coefficients as
    :func:`lanczos_coeffs_reference` in exact arithmetic.

    Returns (alpha[m], beta[m-1], norm2, V) where V holds the m basis vectors
    (columns) and norm2 = |v0|^2.  Stops early on an invariant subspace.
    """
    v0 = np.asarray(v0, dtype=np.float64)
    norm2 = float(np.dot(v0, v0))
    v = v0 / np.sqrt(norm2)
    basis = np.empty((max(int(nsteps), 1), v.size))  # row j: basis vector v_j
    basis[0] = v
    alpha, beta = [], []
    v_prev = np.zeros_like(v)
    b = 0.0
    for j in range(nsteps):
        w = matvec(v) - b * v_prev
        a = float(np.dot(w, v))
        w = w - a * v
        if reorth:
            V = basis[: j + 1]
            w = w - V.T @ (V @ w)
            w = w - V.T @ (V @ w)
        alpha.append(a)
        b = float(np.linalg.norm(w))
        if j == nsteps - 1 or b < 1e-14 * max(1.0, abs(a)):
            break
        beta.append(b)
        v_prev, v = v, w / b
        basis[j + 1] = v
    return np.array(alpha), np.array(beta), norm2, basis[: len(alpha)].T


def cf_greens(alpha, beta, norm2, z, shift, sign=+1):
    """Continued fraction  s * norm2 / (s z + shift - a_0 - b_1^2/(s z + shift - a_1 - ...)).

    Particle part: sign=+1, shift=E0 (sector N+1).  Hole part: sign=-1, shift=E0
    (sector N-1), which equals  norm2/(z - E0 + a_0 - b_1^2/(z - E0 + a_1 - ...)).
    No reference function: the reference has no continued fraction (pinned by
    equality with :func:`lehmann_t0` in the test-suite).
    """
    z = np.asarray(z, dtype=np.complex128)
    w = sign * z + shift
    m = len(alpha)
    g = np.zeros_like(w)
    for k in range(m - 1, 0, -1):
        g = beta[k - 1] ** 2 / (w - alpha[k] - g)
    return sign * norm2 / (w - alpha[0] - g)


# =============================================================================
# Green's functions
# =============================================================================


def lehmann_t0(op_n, op_p1, op_m1, psi0, e0, cdag, c, z):
    """T = 0, fixed-sector Lehmann sum from the reference's own pieces.

    ``_accumulate_sum`` (cmpy/exactdiag.py:110-129) restricted to n = ground state:
    G(z) = sum_m |<m|c^+|0>|^2/(z - E_m + E0) + sum_n |<n|c|0>|^2/(z + E_n - E0).
    ``cdag``/``c`` map psi0 to the upper/lower sector (None if that sector does
    not exist); the sectors are diagonalised densely (exactdiag.py:46-57).
    """
    z = np.asarray(z, dtype=np.complex128)
    g = np.zeros_like(z)
    if op_p1 is not None:
        h = op_p1.csr().toarray()
        w, v = np.linalg.eigh(h)
        ov = np.abs(v.T @ cdag) ** 2
        g += (ov[None, :] / (z[:, None] - w[None, :] + e0)).sum(axis=1)
    if op_m1 is not None:
        h = op_m1.csr().toarray()
        w, v = np.linalg.eigh(h)
        ov = np.abs(v.T @ c) ** 2
        g += (ov[None, :] / (z[:, None] + w[None, :] - e0)).sum(axis=1)
    return g


def lehmann_t0_ij(op_p1, op_m1, e0, cdag_i, cdag_j, c_i, c_j, z):
    """T = 0 fixed-sector G_ij: lehmann_t0 with the two operators of the
    off-diagonal element,
    G_ij(z) = sum_m <0|c_i|m><m|c+_j|0>/(z - E_m + E0) + sum_n <0|c+_j|n><n|c_i|0>/(z + E_n - E0),
    real eigenvectors (so <0|c_i|m> = (V^T c+_i psi0)_m)."""
    z = np.asarray(z, dtype=np.complex128)
    g = np.zeros_like(z)
    if op_p1 is not None:
        w, v = np.linalg.eigh(op_p1.csr().toarray())
        ov = (v.T @ cdag_i) * (v.T @ cdag_j)
        g += (ov[None, :] / (z[:, None] - w[None, :] + e0)).sum(axis=1)
    if op_m1 is not None:
        w, v = np.linalg.eigh(op_m1.csr().toarray())
        ov = (v.T @ c_i) * (v.T @ c_j)
        g += (ov[None, :] / (z[:, None] + w[None, :] - e0)).sum(axis=1)
    return g


def occupation(up_states, dn_states, evals, evecs, beta, emin=0.0, pos=0, sigma=UP):
    """cmpy/exactdiag.py:60-92."""
    up = np.asarray(up_states, np.int64)
    dn = np.asarray(dn_states, np.int64)
    num_dn = len(dn)
    boltz = np.exp(-beta * (evals - emin))
    occ = 0.0
    if sigma == UP:
        for up_idx in range(len(up)):
            if up[up_idx] & (1 << pos):
                idx = up_idx * num_dn + np.arange(num_dn)
                ov = np.sum(np.abs(evecs[idx, :]) ** 2, axis=0)
                occ += np.sum(boltz * ov)
    else:
        for dn_idx in range(num_dn):
            if dn[dn_idx] & (1 << pos):
                idx = np.arange(len(up)) * num_dn + dn_idx
                ov = np.sum(np.abs(evecs[idx, :]) ** 2, axis=0)
                occ += np.sum(boltz * ov)
    return occ


def double_occupation(up_states, dn_states, evals, evecs, beta, emin=0.0, pos=0):
    """cmpy/exactdiag.py:95-107, with the flat index computed as
    up_idx * num_dn + dn_idx (the reference's running ``idx`` is only correct
    when the prange loop runs serially)."""
    up = np.asarray(up_states, np.int64)
    dn = np.asarray(dn_states, np.int64)
    num_dn = len(dn)
    boltz = np.exp(-beta * (evals - emin))
    occ = 0.0
    for up_idx in range(len(up)):
        for dn_idx in range(num_dn):
            if up[up_idx] & dn[dn_idx] & (1 << pos):
                ov = np.abs(evecs[up_idx * num_dn + dn_idx, :]) ** 2
                occ += np.sum(boltz * ov)
    return occ


def _accumulate_sum(gf, z, evals, evals_p1, evecs_p1, cdag_evec, beta, emin):
    """cmpy/exactdiag.py:110-129 (vectorised over m and n)."""
    overlap = np.abs(evecs_p1.T.conj() @ cdag_evec) ** 2
    if np.isfinite(beta):
        exp_evals = np.exp(-beta * (evals - emin))
        exp_evals_p1 = np.exp(-beta * (evals_p1 - emin))
    else:
        exp_evals = np.ones_like(evals)
        exp_evals_p1 = np.ones_like(evals_p1)
    weights = overlap * (exp_evals[None, :] + exp_evals_p1[:, None])
    poles = evals_p1[:, None] - evals[None, :]
    gf += (weights[None, :, :] / (z[:, None, None] - poles[None, :, :])).sum(axis=(1, 2))


class GreensFunctionMeasurement:
    """cmpy/exactdiag.py:137-212."""

    def __init__(self, z, beta, pos=0, sigma=UP, measure_occ=True):
        self.z = np.asarray(z)
        self.beta = beta
        self.pos = pos
        self.sigma = sigma
        self._measure_occ = measure_occ
        self._part = 0
        self._gs_energy = np.inf
        self._gf = np.zeros_like(self.z, dtype=np.complex128)
        self._occ = 0.0
        self._occ_double = 0.0

    @property
    def part(self):
        return self._part * np.exp(-self.beta * self._gs_energy)

    @property
    def gf(self):
        return self._gf / self._part

    @property
    def occ(self):
        return self._occ / self._part

    @property
    def occ_double(self):
        return self._occ_double / self._part

    @property
    def gs_energy(self):
        return self._gs_energy

    def accumulate(self, num_sites, sector, sector_p1, evals, evecs, evals_p1, evecs_p1):
        n_up, n_dn = sector
        min_energy = min(evals)
        factor = 1.0
        if min_energy < self._gs_energy:
            factor = np.exp(-self.beta * (self._gs_energy - min_energy))
            self._gs_energy = min_energy
        up = sector_states(num_sites, n_up)
        dn = sector_states(num_sites, n_dn)
        # _acc_part (171-173)
        self._part *= factor
        self._part += np.sum(np.exp(-self.beta * (evals - self._gs_energy)))
        # _acc_gf (175-183)
        if factor != 1.0:
            self._gf *= factor
        cdag_evec = apply_cdag(evecs, num_sites, n_up, n_dn, self.pos, self.sigma)
        _accumulate_sum(self._gf, self.z, evals, evals_p1, evecs_p1, cdag_evec,
                        self.beta, self._gs_energy)
        if self._measure_occ:
            self._occ *= factor
            self._occ += occupation(up, dn, evals, evecs, self.beta, self._gs_energy,
                                    self.pos, self.sigma)
            self._occ_double *= factor
            self._occ_double += double_occupation(up, dn, evals, evecs, self.beta,
                                                  self._gs_energy, self.pos)


def gf_lehmann(make_op, num_sites, z, beta, pos=0, sigma=UP, occ=True):
    """Grand-canonical finite-T Lehmann Green's function (cmpy/exactdiag.py:215-245).

    ``make_op(n_up, n_dn)`` returns a :class:`SectorOperator` of the model.
    Fillings are visited in ``product(range(N+1), repeat=2)`` order
    (cmpy/basis.py:683-684); sectors with no upper sector are skipped.
    """
    data = GreensFunctionMeasurement(z, beta, pos, sigma, measure_occ=occ)
    cache = {}

    def solve(nu, nd):
        if (nu, nd) not in cache:
            h = make_op(nu, nd).csr().toarray()
            cache[(nu, nd)] = np.linalg.eigh(h)
        return cache[(nu, nd)]

    for n_up, n_dn in itertools.product(range(num_sites + 1), repeat=2):
        p1 = upper_sector(n_up, n_dn, sigma, num_sites)
        if p1 is None:
            continue
        evals, evecs = solve(n_up, n_dn)
        evals_p1, evecs_p1 = solve(*p1)
        data.accumulate(num_sites, (n_up, n_dn), p1, evals, evecs, evals_p1, evecs_p1)
    return data


def lehmann_ij(make_op, num_sites, z, beta, i, j, sigma=UP, sign_mode="jordan_wigner"):
    """Finite-T off-diagonal Green's function

        G_ij(z) = (1/Z) sum_{N} sum_{m in N+1, n in N} <n|c_i|m><m|c+_j|n>
                  (e^{-beta E_n} + e^{-beta E_m}) / (z - E_m + E_n),

    the sum lehmann_full.py:19-97 (``greens_lehmann(model, z, beta, i, j)``)
    sets out to form (its code keeps only the c_i factor, squared, and returns
    the sum without 1/Z), normalised by the partition function of all sectors
    (gf_lehmann, cmpy/exactdiag.py:228-243, leaves sectors without an upper
    sector out of Z; with them the U = 0 sum is exactly the resolvent).
    Energies are measured from the lowest eigenvalue of all sectors at once;
    ``sign_mode="jordan_wigner"`` gives the physical (anticommuting) c, c+.
    """
    return lehmann_ij_many(make_op, num_sites, z, beta, [(i, j)], sigma, sign_mode)[0]


def lehmann_ij_many(make_op, num_sites, z, beta, pairs, sigma=UP, sign_mode="jordan_wigner"):
    """:func:`lehmann_ij` for several (i, j) at once, every sector diagonalised
    once; returns one G per pair."""
    z = np.asarray(z, dtype=np.complex128)
    fillings = list(itertools.product(range(num_sites + 1), repeat=2))
    eig = {}
    for nu, nd in fillings:
        w, v = np.linalg.eigh(make_op(nu, nd).csr().toarray())
        eig[(nu, nd)] = (w, v)
    emin = min(float(w[0]) for w, _ in eig.values())
    part = sum(float(np.sum(np.exp(-beta * (w - emin)))) for w, _ in eig.values()) if np.isfinite(beta) else None
    gs = [np.zeros_like(z) for _ in pairs]
    sites = sorted({s for p in pairs for s in p})
    for nu, nd in fillings:
        p1 = upper_sector(nu, nd, sigma, num_sites)
        if p1 is None:
            continue
        w, v = eig[(nu, nd)]
        w1, v1 = eig[p1]
        ov = {s: v1.T @ apply_cdag(v, num_sites, nu, nd, s, sigma, sign_mode) for s in sites}
        if np.isfinite(beta):
            wt = np.exp(-beta * (w - emin))[None, :] + np.exp(-beta * (w1 - emin))[:, None]
        else:
            wt = 2.0 * np.ones((len(w1), len(w)))
        poles = w1[:, None] - w[None, :]
        for g, (i, j) in zip(gs, pairs):
            weights = ov[i] * ov[j] * wt
            for k in range(len(z)):
                g[k] += np.sum(weights / (z[k] - poles))
    return [g / part if part is not None else g for g in gs]


def resolvent_ij(h, z):
    """Non-interacting G_ij(z) = [(z - h)^{-1}]_ij for the single-particle matrix h
    (on-site energies minus mu on the diagonal, hop_b on the bonds): at U = 0 the
    grand-canonical Lehmann sum reduces to it at any temperature."""
    z = np.atleast_1d(np.asarray(z, dtype=np.complex128))
    w, v = np.linalg.eigh(h)
    return np.einsum("ik,zk,jk->zij", v, 1.0 / (z[:, None] - w[None, :]), v)


def gf0_lehmann(ham, z, mu=0.0, mode="diag"):
    """cmpy/greens.py:18-64 (einsum ``"ij,...j,ji->...i"`` for mode="diag").

    As written, the diag mode sums |V[j, i]|^2/(z - lambda_j); this equals the
    diagonal Green's function whenever |V| is symmetric (tight-binding chains),
    which is where the reference's test uses it.
    """
    eigvals, eigvecs = np.linalg.eigh(ham)
    z = np.atleast_1d(z)
    eigvecs_adj = np.conj(eigvecs).T
    sub = {"full": "ik,...k,kj->...ij", "diag": "ij,...j,ji->...i", "total": "ij,...j,ji->..."}[mode]
    arg = np.subtract.outer(z + mu, eigvals)
    return np.einsum(sub, eigvecs_adj, 1 / arg, eigvecs)


def hybridization_func(z, v, eps_bath):
    """cmpy/models/anderson.py:127-145."""
    x = np.asarray(z)[..., np.newaxis]
    return np.sum(np.square(np.abs(np.atleast_1d(v))) / (x - np.atleast_1d(eps_bath)), axis=-1)


def impurity_gf0(z, u, eps_imp, v, eps_bath, mu):
    """cmpy/models/anderson.py:160-161."""
    return 1 / (z + mu + eps_imp - hybridization_func(z, v, eps_bath))


def impurity_params_ref(u, v):
    """Two-site SIAM poles/residues at half filling, T=0 (cmpy/dmft/twosite.py:25-60)."""
    sqrt16 = np.sqrt(u**2 + 16 * v**2)
    sqrt64 = np.sqrt(u**2 + 64 * v**2)
    a1 = 0.25 * (1 - (u**2 - 32 * v**2) / np.sqrt((u**2 + 64 * v**2) * (u**2 + 16 * v**2)))
    a2 = 0.5 - a1
    e1 = 0.25 * (sqrt64 - sqrt16)
    e2 = 0.25 * (sqrt64 + sqrt16)
    return float(a1), float(a2), float(e1), float(e2)


def impurity_gf_ref(z, u, v):
    """cmpy/dmft/twosite.py:63-81."""
    a1, a2, e1, e2 = impurity_params_ref(u, v)
    return (a1 / (z - e1) + a1 / (z + e1)) + (a2 / (z - e2) + a2 / (z + e2))


# =============================================================================
# Real-time Green's functions (cmpy/exactdiag.py:248-307)
# =============================================================================


def tevo_part(op, phi, e0, times, sign=+1):
    """One part of the real-time G of cmpy/exactdiag.py:248-301.

    sign=+1 (gf_greater, 248-272): phi = c+ psi0 in the upper sector,
        g(t) = -i e^{+i E0 t} <phi| e^{-i H t} |phi>;
    sign=-1 (gf_lesser, 275-299): phi = c psi0 in the lower sector,
        g(t) = +i e^{-i E0 t} <phi| e^{+i H t} |phi>.
    The reference propagates phi with expm_multiply (cmpy/linalg/expm_multiply.py
    wraps scipy's Al-Mohy-Higham interval algorithm) and applies the phase as a
    running product; here the sector is diagonalised densely, so
    <phi|e^{-i s H t}|phi> = sum_k |(V^T phi)_k|^2 e^{-i s lambda_k t}.
    """
    times = np.asarray(times, dtype=np.float64)
    w, v = np.linalg.eigh(op.csr().toarray())
    ov = (v.T @ np.asarray(phi, dtype=np.float64)) ** 2
    ph = np.exp(-1j * sign * np.multiply.outer(times, w - e0))
    return -1j * sign * (ph @ ov)


def tevo_expm_multiply(op, phi, e0, start, stop, num, sign=+1, matvec=None):
    """The reference's own procedure (cmpy/exactdiag.py:248-299) with scipy's
    public expm_multiply: propagate phi under -i s H over linspace(start, stop,
    num), overlap with phi, then multiply by the running phase factor.

    ``matvec`` (real x -> H x, e.g. the C twin's) makes H a LinearOperator
    instead of the assembled CSR matrix; the trace expm_multiply needs is then
    the sum of the operator's diagonal."""
    times, dt = np.linspace(start, stop, num, retstep=True)
    if matvec is None:
        h = (-1j * sign) * op.csr().astype(np.complex128)
        trace = None
    else:
        def apply(x, s):
            x = np.asarray(x)
            y = np.asarray(matvec(np.ascontiguousarray(x.real))) + 1j * np.asarray(
                matvec(np.ascontiguousarray(x.imag)))
            return s * y

        n = op.size
        h = sla.LinearOperator((n, n), matvec=lambda x: apply(x.ravel(), -1j * sign),
                               rmatvec=lambda x: apply(x.ravel(), +1j * sign), dtype=np.complex128)
        trace = -1j * sign * float(np.sum(op.diagonal()))
    psi_t = sla.expm_multiply(h, np.asarray(phi, dtype=np.complex128), start=start, stop=stop, num=num,
                              endpoint=True, traceA=trace)
    overlaps = psi_t @ np.asarray(phi, dtype=np.complex128).conj()
    step = np.exp(+1j * sign * e0 * dt)
    factor = -1j * sign * np.exp(+1j * sign * e0 * times[0])
    overlaps[0] *= factor
    for n in range(1, num):
        factor *= step
        overlaps[n] *= factor
    return times, overlaps


def tevo_two_site_ref(times, u, v):
    """Closed form of gf_tevo for the two-site SIAM at half filling and T = 0:
    the Lange poles (cmpy/dmft/twosite.py:25-81) G(z) = sum_k a_k [1/(z - e_k)
    + 1/(z + e_k)] give g>(t) = -i sum_k a_k e^{-i e_k t} and
    g<(t) = +i sum_k a_k e^{+i e_k t}, so g> - g< = -2i sum_k a_k cos(e_k t)."""
    a1, a2, e1, e2 = impurity_params_ref(u, v)
    t = np.asarray(times, dtype=np.float64)
    return -2j * (a1 * np.cos(e1 * t) + a2 * np.cos(e2 * t))


def fourier_t2z(times, gf_t, omegas, delta=1e-2, eta=None):
    """fourier_t2z (cmpy/exactdiag.py:311-316): z = omegas + i eta with
    eta = -log(delta) / times[-1] by default, and G(z) = int_0^T e^{izt} G(t) dt
    by gftool.fourier.tt2z (gftool >= 0.11.0, absent here; its default is the
    piecewise-linear rule tt2z_lin).  Restated from that rule's definition:
    G(t) linear between the samples, each interval integrated exactly,
      int_{t_k}^{t_k+h} e^{a t} (g_k + (g_{k+1} - g_k) s/h) dt
        = e^{a t_k} h [g_k E1(x) + (g_{k+1} - g_k) E2(x)],  a = i z, x = a h,
      E1(x) = (e^x - 1)/x,  E2(x) = (x e^x - e^x + 1)/x^2,
    with the power series of E1, E2 for |x| < 0.1 (no cancellation)."""
    t = np.asarray(times, dtype=np.float64)
    g = np.asarray(gf_t, dtype=np.complex128)
    if eta is None:
        eta = -np.log(delta) / t[-1]
    z = np.asarray(omegas, dtype=np.float64) + 1j * eta
    out = np.zeros(z.shape, dtype=np.complex128)
    for k in range(len(t) - 1):
        h = t[k + 1] - t[k]
        x = 1j * z * h
        small = np.abs(x) < 0.1
        xs = np.where(small, 1.0, x)
        e1 = np.where(small, _series(x, 1), np.expm1(xs) / xs)
        e2 = np.where(small, _series(x, 2), (xs * np.exp(xs) - np.expm1(xs)) / xs ** 2)
        out += np.exp(1j * z * t[k]) * h * (g[k] * e1 + (g[k + 1] - g[k]) * e2)
    return z, out


def _series(x, kind, terms=12):
    """E1(x) = sum_n x^n / (n+1)!,  E2(x) = sum_n (n+1) x^n / (n+2)!."""
    acc = np.zeros_like(x, dtype=np.complex128)
    for n in range(terms - 1, -1, -1):
        c = 1.0 / math.factorial(n + 1) if kind == 1 else (n + 1) / math.factorial(n + 2)
        acc = acc * x + c
    return acc


# =============================================================================
# Lattices (replacing lattpy's neighbour lists used by the reference tests)
# =============================================================================


def chain_bonds(num_sites, periodic=False):
    bonds = [(i, i + 1) for i in range(num_sites - 1)]
    if periodic and num_sites > 2:
        bonds.append((0, num_sites - 1))
    return bonds


def square_bonds(nx, ny, periodic=True):
    """Nearest-neighbour bonds (i < j) of an nx x ny square lattice, site = x + nx*y."""
    bonds = set()
    for y in range(ny):
        for x in range(nx):
            i = x + nx * y
            for dx, dy in ((1, 0), (0, 1)):
                xx, yy = x + dx, y + dy
                if periodic:
                    xx %= nx
                    yy %= ny
                elif xx >= nx or yy >= ny:
                    continue
                j = xx + nx * yy
                if i != j:
                    bonds.add((min(i, j), max(i, j)))
    return sorted(bonds)


def eigvalsh_chain(num_sites, eps, t):
    """Eigenvalues of the open tight-binding chain (cmpy/models/tightbinding.py:16-40):
    the symmetric tridiagonal matrix with eps on the diagonal and t beside it."""
    diag = np.full(num_sites, float(eps)) if np.isscalar(eps) else np.asarray(eps, dtype=np.float64)
    return la.eigvalsh_tridiagonal(diag, np.full(num_sites - 1, float(t)))


def free_fermion_energy(num_sites, bonds, hop, eps, n_up, n_dn):
    """U = 0 closed form: sum of the lowest n_up and n_dn single-particle levels."""
    h = np.diag(np.full(num_sites, float(eps)))
    for i, j in bonds:
        h[i, j] += hop
        h[j, i] += hop
    w = np.linalg.eigvalsh(h)
    return float(np.sum(w[:n_up]) + np.sum(w[:n_dn]))


# Which check pins each function (besides retyping it):
# =============================================================================
# Bath-fit gradient (no reference function: the reference fits the bath with
# scipy optimisers, cmpy/dmft/utils.py:163-280; this is the FP64 CPU reference
# gradient for the batched adjoint backward)
# =============================================================================


def siam_gf_t0(u, eps_imp, eps_bath, v, mu, n_up, n_dn, wn, width=0):
    """Impurity G(i w_n) at T = 0 by dense Lehmann sums (lehmann_t0)."""
    ns = len(eps_bath) + 1
    mode = "reference" if width == 0 else "jordan_wigner"
    op = siam_operator(n_up, n_dn, u, eps_imp, eps_bath, v, mu, width)
    w, vv = np.linalg.eigh(op.csr().toarray())
    e0, psi = w[0], vv[:, 0]
    opp = siam_operator(n_up + 1, n_dn, u, eps_imp, eps_bath, v, mu, width) if n_up + 1 <= ns else None
    opm = siam_operator(n_up - 1, n_dn, u, eps_imp, eps_bath, v, mu, width) if n_up >= 1 else None
    cd = apply_cdag(psi, ns, n_up, n_dn, 0, UP, mode) if opp is not None else None
    c = apply_c(psi, ns, n_up, n_dn, 0, UP, mode) if opm is not None else None
    return lehmann_t0(op, opp, opm, psi, e0, cd, c, 1j * np.asarray(wn))


def bathfit_grad_dense(u, eps_imp, eps_bath, v, mu, n_up, n_dn, wn, g_target, width=0):
    """L = sum_n |G(i w_n) - g_target_n|^2 and dL/d(eps_bath, v), FP64, dense.

    Adjoint formula with exact resolvents (dense solves) and exact ground
    state (dense eigh); w_n = 2 conj(G_n - T_n):
      dL/dtheta = S+ - S- + (B - A) dE0/dtheta - 2 psi^T dH zeta,
    S+ = Re sum_n w_n x_n^T dH+ x_n, A = Re sum_n w_n x_n^T x_n
    (x_n = (z_n + E0 - H+)^-1 c+ psi), S-, B likewise with
    y_n = (z_n - E0 + H-)^-1 c psi, dE0/dtheta = psi^T dH psi, and
    (H - E0) zeta = P Re(c X + c+ Y), X = sum w_n x_n, Y = sum w_n y_n.
    dH/d eps_k and dH/d v_k are the operators of the same model with unit
    eps_k (resp. v_k) and every other parameter zero (H is linear in them).
    Pinned by central finite differences of :func:`siam_gf_t0` in the tests.
    """
    eps_bath = np.asarray(eps_bath, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    nb = len(eps_bath)
    ns = nb + 1
    z = 1j * np.asarray(wn, dtype=np.float64)
    mode = "reference" if width == 0 else "jordan_wigner"

    def dense(nu, eb, vv, uu=u, ei=eps_imp, m=mu):
        return siam_operator(nu, n_dn, uu, ei, eb, vv, m, width).csr().toarray()

    H = dense(n_up, eps_bath, v)
    w, vv = np.linalg.eigh(H)
    e0, psi = w[0], vv[:, 0]
    D = len(psi)
    has_p, has_m = n_up + 1 <= ns, n_up >= 1
    eye = np.eye(D)
    G = np.zeros(len(z), dtype=np.complex128)
    xs = ys = None
    if has_p:
        Hp = dense(n_up + 1, eps_bath, v)
        phip = apply_cdag(psi, ns, n_up, n_dn, 0, UP, mode)
        xs = np.array([np.linalg.solve((zz + e0) * np.eye(len(phip)) - Hp, phip) for zz in z])
        G += xs @ phip
        Cd = np.array([apply_cdag(e, ns, n_up, n_dn, 0, UP, mode) for e in eye]).T
    if has_m:
        Hm = dense(n_up - 1, eps_bath, v)
        phim = apply_c(psi, ns, n_up, n_dn, 0, UP, mode)
        ys = np.array([np.linalg.solve((zz - e0) * np.eye(len(phim)) + Hm, phim) for zz in z])
        G += ys @ phim
        Cm = np.array([apply_c(e, ns, n_up, n_dn, 0, UP, mode) for e in eye]).T
    diff = G - np.asarray(g_target)
    loss = float(np.sum(np.abs(diff) ** 2))
    wgt = 2 * np.conj(diff)
    r = np.zeros(D)
    A = B = 0.0
    if has_p:
        r += np.real(Cd.T @ (wgt[:, None] * xs).sum(0))
        A = float(np.real(np.sum(wgt * np.einsum("na,na->n", xs, xs))))
    if has_m:
        r += np.real(Cm.T @ (wgt[:, None] * ys).sum(0))
        B = float(np.real(np.sum(wgt * np.einsum("na,na->n", ys, ys))))
    proj = eye - np.outer(psi, psi)
    zeta = proj @ np.linalg.lstsq(H - e0 * eye, proj @ r, rcond=None)[0]
    grads = {}
    for kind in ("eps", "v"):
        g = np.zeros(nb)
        for k in range(nb):
            unit = np.zeros(nb)
            unit[k] = 1.0
            eb1, v1 = (unit, np.zeros(nb)) if kind == "eps" else (np.zeros(nb), unit)
            dHn = dense(n_up, eb1, v1, 0.0, 0.0, 0.0)
            val = (B - A) * (psi @ dHn @ psi) - 2 * psi @ dHn @ zeta
            if has_p:
                dHp = dense(n_up + 1, eb1, v1, 0.0, 0.0, 0.0)
                val += np.real(np.sum(wgt * np.einsum("na,ab,nb->n", xs, dHp, xs)))
            if has_m:
                dHm = dense(n_up - 1, eb1, v1, 0.0, 0.0, 0.0)
                val -= np.real(np.sum(wgt * np.einsum("na,ab,nb->n", ys, dHm, ys)))
            g[k] = val
        grads[kind] = g
    return loss, grads["eps"], grads["v"]


PARITY = {
    "generate_states_permutations": "test_basis.py:124-141 state tables",
    "sector_states": "generate_states_permutations for ns<=8; test_basis.py:124-141",
    "combinadic_rank": "position in sector_states for every ns<=12",
    "upper_sector/lower_sector": "test_basis.py:91-111 tables",
    "state_label": "basis.py:613-631 docstring label lists",
    "hopping_sign/hopping_terms": "operators.py:503-507 matrix; U=0 free-fermion "
    "closed form on periodic chains and the 4x4 torus (wrong mask range or width "
    "fails it); hand-computed sign cases",
    "weighted_element/onsite/inter": "operators.py:337-341, 391-395 docstrings; "
    "test_models_hubbard.py:14-25",
    "coo_matvec/coo_dense": "test_models_hubbard.py:14-25; Hermiticity",
    "SectorOperator": "COO assembly at small sizes; free-fermion closed form",
    "apply_cdag/apply_c": "test_operator.py:31-42 adjointness; Lange two-site KAT; a block of "
    "vectors maps column by column",
    "lanczos_normalized": "lanczos_coeffs_reference coefficients; dense eigh",
    "cf_greens": "lehmann_t0 equality at small size; Lange KAT",
    "lehmann_t0": "Lange two-site KAT (dmft/twosite.py:25-81); sum rule",
    "gf_lehmann": "test_exactdiag.py:24-96 (U=0 vs gf0_lehmann, atomic peaks, <n>=1/2)",
    "gf0_lehmann": "chain tight-binding resolvent computed by direct inversion",
    "impurity_gf_ref": "lehmann_t0 on the two-site SIAM",
    "siam_gf_t0": "lehmann_t0 (pinned above) on the Anderson model",
    "bathfit_grad_dense": "central finite differences (h=1e-5) of the loss built on siam_gf_t0",
    "tevo_part": "tevo_expm_multiply (scipy's expm_multiply, the algorithm the reference wraps) on "
    "Hubbard and Anderson sectors; tevo_two_site_ref at half filling",
    "tevo_expm_multiply": "tevo_part (dense eigh) and the two-site closed form",
    "tevo_two_site_ref": "impurity_params_ref poles (pinned via lehmann_t0); t = 0 value -i",
    "fourier_t2z / _series": "exact on piecewise-linear G(t) (closed-form integral); closed-form transform of "
                              "sums of exponentials to O(h^2); two-site tevo_two_site_ref transform; "
                              "series/direct branches agree",
    "lehmann_ij/lehmann_ij_many": "U = 0 resolvent_ij on rings and open chains at interior sites (a wrong string sign "
    "or a dropped factor fails it); i = j equals gf_lehmann; G_ij = G_ji",
    "resolvent_ij": "direct inversion of z - h",
    "lehmann_t0_ij": "U = 0 closed-shell ground state: resolvent_ij at interior sites; i = j equals lehmann_t0",
    "eigvalsh_chain": "eps + 2 t cos(k pi / (N + 1)), k = 1..N (open-chain closed form)",
    "binom/popcount": "through combinadic_rank and sector_states (rank = position for every ns<=12)",
    "bit_count_width": "through hopping_sign: hand cases and the free-fermion closed forms",
    "onsite_energy_triples/hubbard_inter_triples/hopping_triples/hubbard_triples/_concat":
    "operators.py:337-341, 391-395, 503-507 docstring matrices; test_models_hubbard.py:14-25 "
    "(sector (1,1)) and Hermiticity over every sector (test_models_hubbard.py:28-48)",
    "siam_triples": "factorised operator equals the COO assembly; measured reference-convention energies",
    "hubbard_operator/siam_operator": "COO assembly at small sizes; free-fermion closed forms; "
    "atomic limit; SIAM reference-convention energies and the U = 0 fermion-sign closed form",
    "coo_csr/coo_trace": "dense assembly; trace equals the sum of the factorised diagonal",
    "groundstate": "dense eigh of the same matrix; free-fermion and atomic closed forms",
    "lanczos_coeffs_reference/lanczos_ground_state": "tridiagonal spectrum gives the dense-eigh E0",
    "_jw_sign": "canonical anticommutation of the Jordan-Wigner ladder operators",
    "GreensFunctionMeasurement/_accumulate_sum": "through gf_lehmann (test_exactdiag.py:24-96)",
    "occupation": "atomic limit per site (four-state atom, t = 0) and the U = 0 Fermi sum "
    "sum_k |phi_k(pos)|^2 f(e_k) with uneven levels and hoppings, both spins",
    "double_occupation": "atomic limit e^{-b(2e+U)}/Z_i; U = 0 Wick factorisation "
    "<n_up n_dn> = <n_up><n_dn>; planted index mistakes (row-reset, dn-major, "
    "selected-only increment) fail the Wick pin",
    "hybridization_func": "U = 0 impurity G (lehmann_ij, itself the resolvent) equals "
    "1/(z - (eps_imp - mu) - Delta); Delta equals the directly solved bath block",
    "impurity_gf0": "equals the U = 0 impurity G at eps_imp = 0 and, for eps_imp != 0, that "
    "of the reflected level -eps_imp (the sign anderson.py:160-161 writes)",
    "impurity_params_ref": "through impurity_gf_ref (lehmann_t0 on the two-site SIAM)",
    "chain_bonds/square_bonds/free_fermion_energy": "tight-binding levels 2t cos(2 pi k/N) "
    "and 2t(cos kx + cos ky) summed to the free-fermion energy; bond counts per site",
}
