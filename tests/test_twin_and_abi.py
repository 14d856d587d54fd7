"""CPU tests: the C twin of the oracle, and the C ABI's host side (no GPU needed)."""

import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_array_equal

from oracle import ed_oracle as O
from oracle import twin

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("case", ["hub_ring6", "hub_sq4x2", "siam_ref", "siam_fermion"])
def test_twin_matvec_matches_oracle(case):
    rng = np.random.default_rng(11)
    if case == "hub_ring6":
        ns, nu, nd, bonds = 6, 3, 2, O.chain_bonds(6, True)
        args = (ns, nu, nd, bonds, [1.0] * len(bonds), np.full(ns, -0.7), np.full(ns, 3.0), ns)
    elif case == "hub_sq4x2":
        ns, nu, nd, bonds = 8, 4, 3, O.square_bonds(4, 2, True)
        args = (ns, nu, nd, bonds, [1.0] * len(bonds), np.zeros(ns), np.full(ns, 4.0), ns)
    else:
        eps, v = rng.uniform(-2, 2, 5), rng.uniform(0.2, 1.0, 5)
        ns = 6
        args = (ns, 3, 3, [(0, k + 1) for k in range(5)], v, np.append(-2.0, eps),
                np.append(4.0, np.zeros(5)), 0 if case == "siam_ref" else ns)
    tw = twin.TwinOperator(*args)
    ref = O.SectorOperator(*args)
    x = rng.standard_normal(ref.size)
    assert_allclose(tw.matvec(x), ref.matvec(x), rtol=1e-13, atol=1e-13)


def test_twin_lanczos_energy():
    ns, bonds = 8, O.square_bonds(4, 2, True)
    args = (ns, 3, 3, bonds, [1.0] * len(bonds), np.zeros(ns), np.full(ns, 2.0), ns)
    tw = twin.TwinOperator(*args)
    a, b = tw.lanczos(200, seed=3)
    e = np.linalg.eigvalsh(np.diag(a) + np.diag(b[:-1], 1) + np.diag(b[:-1], -1))[0]
    e_ref, _ = O.groundstate(O.SectorOperator(*args))
    assert abs(e - e_ref) < 1e-9


# -- C ABI, host side ---------------------------------------------------------------


def _header_functions():
    text = (ROOT / "include" / "cmpy_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z0-9_]+\s*\*?\s*(ed_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_header_symbol():
    from cmpy_amd import _lib

    lib = _lib.lib()
    names = _header_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_binom_and_host_states():
    from cmpy_amd import _lib

    lib = _lib.lib()
    for n in range(0, 31):
        for k in range(-1, n + 2):
            assert lib.ed_binom(n, k) == (math.comb(n, k) if 0 <= k <= n else 0)
    assert lib.ed_binom(64, 32) == math.comb(64, 32)
    for ns in range(1, 17):
        for n in range(ns + 1):
            out = np.zeros(math.comb(ns, n), dtype=np.uint32)
            _lib.call("ed_sector_states_host", ns, n, out.ctypes.data_as(C.POINTER(C.c_uint32)))
            assert_array_equal(out.astype(np.int64), O.sector_states(ns, n))


def test_host_errors_are_reported():
    from cmpy_amd import _lib

    out = np.zeros(4, dtype=np.uint32)
    with pytest.raises(_lib.EDInvalidArgument, match="filling"):
        _lib.call("ed_sector_states_host", 3, 5, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    with pytest.raises(_lib.EDInvalidArgument, match="num_sites"):
        _lib.call("ed_sector_states_host", 40, 2, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    assert _lib.lib().ed_version() >= 10000


def test_basis_surface_on_host(ref_fixtures):
    from cmpy_amd import Basis, binstr, create, annihilate, UP, DN, upper_sector, lower_sector

    for case in ref_fixtures["basis_get_states"]["cases"]:
        b = Basis(case["num_sites"])
        states = b.get_states(case["n"])
        assert [binstr(s, case["num_sites"]) for s in states] == case["states"]
    for num, pos, res in ref_fixtures["create"]["cases"]:
        assert create(int(num, 2), pos) == (None if res is None else int(res, 2))
    for num, pos, res in ref_fixtures["annihilate"]["cases"]:
        assert annihilate(int(num, 2), pos) == (None if res is None else int(res, 2))
    for case in ref_fixtures["state_labels"]["cases"]:
        sec = Basis(case["num_sites"]).get_sector(case["n_up"], case["n_dn"])
        assert sec.state_labels() == case["labels"]
    assert upper_sector(2, 3, UP, 3) == (3, 3) and upper_sector(3, 3, UP, 3) is None
    assert lower_sector(0, 1, UP) is None and lower_sector(0, 1, DN) == (0, 0)
    b = Basis(12)
    assert b.get_sector(6, 6).size == 924**2


def test_product_requires_gpu_loudly():
    """Without a HIP device the engine refuses to run (no CPU fallback)."""
    from cmpy_amd import _lib, SectorHamiltonOperator

    if _lib.lib().ed_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(_lib.EDLibraryError):
        SectorHamiltonOperator(2, 1, 1, [(0, 1)], [1.0], [0.0, 0.0], [1.0, 1.0], 2)
