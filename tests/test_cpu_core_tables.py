"""Operator construction on the CPU under AddressSanitizer and UBSan
(tests/cpu_kernels/core_check.cpp): ed_op_create with every table builder of
ed_core.hip -- states, diagonal, ELL and packed hop tables, sign-split offset
tables, the half-width row tables with their host-side bank colouring
(ed_colouring.h), sorted target, pair and quad lists -- its kernels run on the
device emulation.  The sectors are chosen so that each builder runs: the 4x4
torus at half filling (12,870 states per spin: half-width and pair tables),
4x4 (1,6) (8,008 dn states), the 12-site chain, an impurity model (width 0,
non-uniform hopping: no sign-split tables) and small lattices with site-
dependent U and eps.  A run passes with no sanitizer report; the exported
hop tables and dense H then equal the oracle's."""

import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import ed_oracle as O

HERE = Path(__file__).resolve().parent
CLANG = Path("/opt/rocm/llvm/bin/clang++")

pytestmark = pytest.mark.skipif(not CLANG.exists(), reason="host clang++ with sanitizer runtimes not found")


def _models():
    rng = np.random.default_rng(7)
    sq44 = O.square_bonds(4, 4)
    ch12 = O.chain_bonds(12)
    sq32 = O.square_bonds(3, 2)
    sq42 = O.square_bonds(4, 2)
    siam = [(0, k) for k in range(1, 6)]
    return [
        # (ns, n_up, n_dn, width, bonds, hops, eps, u)
        (16, 8, 8, 16, sq44, np.ones(len(sq44)), np.zeros(16), np.full(16, 4.0)),
        (16, 1, 6, 16, sq44, np.ones(len(sq44)), np.linspace(-1, 1, 16), np.full(16, 3.0)),
        (12, 6, 6, 12, ch12, np.ones(len(ch12)), np.full(12, -2.0), np.full(12, 4.0)),
        (6, 3, 3, 0, siam, rng.uniform(0.2, 1.0, 5), np.append(-2.0, rng.uniform(-2, 2, 5)), np.append(4.0, np.zeros(5))),
        (6, 3, 2, 6, sq32, rng.uniform(0.5, 1.5, len(sq32)), rng.uniform(-1, 1, 6), rng.uniform(0, 4, 6)),
        (8, 4, 3, 8, sq42, np.ones(len(sq42)), np.zeros(8), np.full(8, 2.0)),
    ]


def _core_check():
    import importlib.util

    spec = importlib.util.spec_from_file_location("build_harness", HERE / "cpu_kernels" / "build_harness.py")
    bh = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bh)
    return bh.build(name="core_check", target="core")


def _run_models(exe, path, models, env=None):
    import os

    with open(path, "w") as f:
        f.write(f"{len(models)}\n")
        for ns, nu, nd, w, bonds, hops, eps, u in models:
            f.write(f"{ns} {nu} {nd} {w} {len(bonds)}\n")
            f.write(" ".join(f"{i} {j}" for i, j in bonds) + "\n")
            f.write(" ".join(f"{x:.17g}" for x in hops) + "\n")
            f.write(" ".join(f"{x:.17g}" for x in eps) + "\n")
            f.write(" ".join(f"{x:.17g}" for x in u) + "\n")
    r = subprocess.run([str(exe), str(path)], capture_output=True, text=True, timeout=900,
                       env={**os.environ, **(env or {})})
    assert r.returncode == 0, r.stderr[-4000:]
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    return r


def _half_lines(out):
    """{(model, sigma): words} from the harness's half-table checks."""
    got, m = {}, -1
    for line in out.splitlines():
        t = line.split()
        if t[0] == "model":
            m = int(t[1])
        elif t[0] == "half":
            got[(m, int(t[1]))] = int(t[2])
    return got


def test_half_tables_decode_to_the_hop_lists(tmp_path):
    """The half-width row tables of the 4x4 torus sectors (12,870 and 8,008
    states): every state's decoded entries equal its packed hop list as a
    multiset of (target, sign), each slot of a half-wave reads 32 distinct
    bank pairs, and the lane-contiguous copy holds the same words (checked
    inside the harness)."""
    exe = _core_check()
    got = _half_lines(_run_models(exe, tmp_path / "half.txt", _models()[:2]).stdout)
    assert set(got) == {(0, 1), (0, 2), (1, 2)}, got
    assert all(0 < w <= 24 for w in got.values()), got


def test_operator_tables_clean_and_equal_to_oracle(tmp_path):
    exe = _core_check()
    models = _models()
    r = _run_models(exe, tmp_path / "models.txt", models)
    blocks, cur = [], None
    for line in r.stdout.splitlines():
        t = line.split()
        if t[0] == "model":
            cur = {"D": int(t[2]), "hop": {1: {}, 2: {}}}
            blocks.append(cur)
        elif t[0] == "dense":
            cur["dense"] = np.array(t[1:], dtype=np.float64)
        elif t[0] == "hop":
            cnt = int(t[3])
            cur["hop"][int(t[1])][int(t[2])] = [(int(t[4 + 2 * e]), float(t[5 + 2 * e])) for e in range(cnt)]
    assert len(blocks) == len(models)
    for blk, (ns, nu, nd, w, bonds, hops, eps, u) in zip(blocks, models):
        assert blk["D"] == O.binom(ns, nu) * O.binom(ns, nd)
        for sigma, n in ((1, nu), (2, nd)):
            got = blk["hop"][sigma]
            if not got:
                continue
            st = O.sector_states(ns, n)
            for k in range(0, len(st), max(1, len(st) // 400)):
                exp = []
                for b, (i, j) in enumerate(bonds):
                    o, _, a = O.hopping_terms(st[k:k + 1], w, i, j, hops[b])
                    if len(o):
                        exp.append((int(np.searchsorted(st, int(st[k]) ^ (1 << i) ^ (1 << j))), float(a[0])))
                assert got[k] == exp, (ns, nu, nd, sigma, k)
        if "dense" in blk:
            ref = O.SectorOperator(ns, nu, nd, bonds, hops, eps, u, w).csr().toarray()
            D = blk["D"]
            assert np.array_equal(blk["dense"].reshape(D, D), ref), (ns, nu, nd)
