"""Bank-conflict-free slot order of the half-width row-pass tables (host code).

cmpy_amd/csrc/ed_colouring.h orders each 32-state half-wave's hop sources so
that every gather slot reads 32 distinct LDS bank pairs (index mod 32).  A
small g++ driver runs the header on the CPU, under AddressSanitizer and UBSan;
the checks are the colouring's
defining properties: every slot is a permutation of the 32 bank pairs, every
lane keeps exactly its own sources, pads are zero-pad indices D..D+31, and
the slot count equals the largest lane or bank-pair degree (the Konig bound
for bipartite edge colouring).
"""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import ed_oracle as O

ROOT = Path(__file__).resolve().parents[1]
L = 32  # lanes served together by one 64-bit LDS read = bank pairs

DRIVER = r"""
#include <cstdio>
#include "ed_colouring.h"
int main() {
  long long D; int nchunk;
  if (scanf("%lld %d", &D, &nchunk) != 2) return 1;
  for (int c = 0; c < nchunk; ++c) {
    std::vector<std::vector<int32_t>> lanes(32);
    for (int l = 0; l < 32; ++l) {
      int n; if (scanf("%d", &n) != 1) return 1;
      lanes[l].resize(n);
      for (int k = 0; k < n; ++k) if (scanf("%d", &lanes[l][k]) != 1) return 1;
    }
    ed::ChunkColouring cc = ed::colour_chunk(lanes, D);
    printf("%d\n", cc.colours);
    for (size_t i = 0; i < cc.target.size(); ++i) printf("%d ", cc.target[i]);
    printf("\n");
  }
  return 0;
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("colour")
    src = d / "drv.cpp"
    src.write_text(DRIVER)
    exe = d / "drv"
    # AddressSanitizer and UBSan on the host table builder: every test below
    # runs the colouring under both and fails on any report
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                    "-fno-omit-frame-pointer", f"-I{ROOT / 'cmpy_amd' / 'csrc'}", str(src), "-o", str(exe)],
                   check=True)
    return exe


def run_colouring(exe, D, chunks):
    lines = [f"{D} {len(chunks)}"]
    for lanes in chunks:
        for l in range(L):
            v = lanes[l] if l < len(lanes) else []
            lines.append(" ".join([str(len(v))] + [str(t) for t in v]))
    out = subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    tok = out.stdout.split("\n")
    res = []
    for c in range(len(chunks)):
        K = int(tok[2 * c])
        if K < 0:
            res.append((K, None))
            continue
        tg = np.array(tok[2 * c + 1].split(), dtype=np.int64).reshape(K, L) if K else np.zeros((0, L), np.int64)
        res.append((K, tg))
    return res


def check_chunk(D, lanes, K, tg):
    lanes = [list(v) for v in lanes] + [[] for _ in range(L - len(lanes))]
    deg_lane = max(len(v) for v in lanes)
    bank_deg = np.bincount([t % L for v in lanes for t in v], minlength=L)
    assert K == max(deg_lane, int(bank_deg.max()) if bank_deg.size else 0)
    for k in range(K):
        banks = tg[k] % L
        assert sorted(banks.tolist()) == list(range(L)), f"slot {k} repeats a bank pair"
    for l in range(L):
        col = tg[:, l]
        real = sorted(int(t) for t in col if t < D)
        assert real == sorted(lanes[l]), f"lane {l} sources changed"
        pads = [int(t) for t in col if t >= D]
        assert all(D <= t < D + L for t in pads)


def hop_lists(ns, n, bonds, width):
    """Per state: (targets with sign +1, targets with sign -1), from the oracle."""
    states = O.sector_states(ns, n)
    pos = [[] for _ in states]
    neg = [[] for _ in states]
    for (i, j) in bonds:
        a, b = min(i, j), max(i, j)
        for src, tgt, val in zip(*O.hopping_terms(states, width, a, b, 1.0)):
            (neg if val < 0 else pos)[src].append(int(tgt))
    return len(states), pos, neg


def test_colouring_4x4_dn_basis(driver):
    """Every half-wave of the 4x4 periodic (8 of 16) basis, both sign lists."""
    D, pos, neg = hop_lists(16, 8, O.square_bonds(4, 4), 16)
    for lists in (pos, neg):
        chunks = [lists[c:c + L] for c in range(0, D, L)]
        for lanes, (K, tg) in zip(chunks, run_colouring(driver, D, chunks)):
            check_chunk(D, lanes, K, tg)


def test_colouring_skewed_random(driver):
    """Random multigraphs with heavy bank skew and repeated targets."""
    rng = np.random.default_rng(7)
    D = 5000
    chunks = []
    for _ in range(40):
        hot = rng.integers(0, D, 4)
        lanes = []
        for _ in range(L):
            n = int(rng.integers(0, 24))
            t = np.where(rng.random(n) < 0.4, rng.choice(hot, n), rng.integers(0, D, n))
            lanes.append(t.tolist())
        chunks.append(lanes)
    chunks.append([[] for _ in range(L)])           # no entries at all
    chunks.append([[3] * 5] + [[] for _ in range(10)])  # a short last chunk
    for lanes, (K, tg) in zip(chunks, run_colouring(driver, D, chunks)):
        check_chunk(D, lanes, K, tg)


def test_colouring_rejects_out_of_range(driver):
    """A target outside [0, D) is reported (colours = -1), not written past the table."""
    D = 100
    res = run_colouring(driver, D, [[[5, 100]], [[-1]], [[99]]])
    assert res[0][0] == -1 and res[1][0] == -1
    check_chunk(D, [[99]], *res[2])
