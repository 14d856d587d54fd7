"""The batched impurity solver's kernels run on the CPU under AddressSanitizer
and UBSan (tests/cpu_kernels): the library's own k_batch_solve and
k_batch_backward, every GPU thread a fiber, every LDS and private array at
its exact size, every group's workspace slot followed by poisoned guard
bytes.  A run passes only with no sanitizer report and no divergence at a
collective point; its results are then checked against the oracle, so the
emulation itself is pinned (ground-state energies by the oracle's eigensolver,
G(i w_n) by its exact Lehmann sums, the adjoint gradient by central
differences of those sums).  A planted slot-sizing mistake must be reported.
"""

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import ed_oracle as O

HERE = Path(__file__).resolve().parent
CLANG = Path("/opt/rocm/llvm/bin/clang++")

pytestmark = pytest.mark.skipif(not CLANG.exists(), reason="host clang++ with sanitizer runtimes not found")


def _build_harness():
    import importlib.util

    spec = importlib.util.spec_from_file_location("build_harness", HERE / "cpu_kernels" / "build_harness.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="module")
def harness():
    return _build_harness().build()


def _models(sel_nb, per=1, seed=2024):
    """Models of random_batch(512, seed) (the benchmark batch): the first
    ``per`` of each bath size in sel_nb."""
    from cmpy_amd.batch import random_batch

    p = random_batch(512, seed)
    nb = np.array([len(e) for e in p["eps_bath"]])
    idx = []
    for n in sel_nb:
        idx += list(np.nonzero(nb == n)[0][:per])
    out = []
    for i in idx:
        n = (nb[i] + 1) // 2
        out.append(dict(nb=int(nb[i]), n=int(n), u=float(p["u"][i]), eimp=float(-p["u"][i] / 2),
                        eb=np.asarray(p["eps_bath"][i]), v=np.asarray(p["v"][i])))
    return out


def _write(path, models, wn, wave_max_dim, grid, wave_blocks, keep=0, max_iter=300, cf_steps=1024):
    with open(path, "w") as f:
        f.write(f"{len(models)} 9 {wave_max_dim} {grid} {wave_blocks} {max_iter} {cf_steps} {keep} {len(wn)}\n")
        f.write(" ".join(f"{w:.17g}" for w in wn) + "\n")
        for m in models:
            f.write(f"{m['nb']} {m['n']} {m['n']} {m['u']:.17g} {m['eimp']:.17g} "
                    + " ".join(f"{x:.17g}" for x in m["eb"]) + " " + " ".join(f"{x:.17g}" for x in m["v"]) + "\n")


def _run(exe, path, timeout=600, env=None):
    import os

    r = subprocess.run([str(exe), str(path)], capture_output=True, text=True, timeout=timeout,
                       env={**os.environ, **(env or {})})
    return r


def _parse(out, M, nw):
    e0 = np.zeros(M)
    res = np.zeros((M, 2))
    g = np.zeros((M, nw), dtype=np.complex128)
    grads = {}
    for line in out.splitlines():
        t = line.split()
        if t[0] == "e0":
            e0[int(t[1])] = float(t[2])
        elif t[0] == "res":
            res[int(t[1])] = [float(t[2]), float(t[3])]
        elif t[0] == "g":
            g[int(t[1]), int(t[2])] = float(t[3]) + 1j * float(t[4])
        elif t[0] == "grad":
            grads[(int(t[1]), int(t[2]))] = (float(t[3]), float(t[4]))
    return e0, res, g, grads


def _clean(r):
    assert r.returncode == 0, r.stderr[-4000:]
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]


def test_wave_mode_benchmark_models(harness, tmp_path):
    """Two models of every bath size of the benchmark batch (5..9, the
    largest D = 63,504), all in wave mode (the autotuner's largest split):
    clean, and E0 equal to the oracle's eigensolver."""
    models = _models((9, 8, 7, 6, 5), per=2)
    wn = (2 * np.arange(16) + 1) * np.pi / 50.0
    path = tmp_path / "wave.txt"
    _write(path, models, wn, 63504, 1, 1)
    r = _run(harness, path)
    _clean(r)
    e0, res, g, _ = _parse(r.stdout, len(models), len(wn))
    assert np.all(res[:, 1] <= 1e-9), res
    for i, m in enumerate(models):
        op = O.siam_operator(m["n"], m["n"], m["u"], m["eimp"], m["eb"], m["v"], 0.0)
        ref = O.groundstate(op, tol=1e-14)[0]
        assert abs(e0[i] - ref) <= 1e-10 * abs(ref), (i, e0[i], ref)
    assert np.all(np.isfinite(g))


def test_block_and_wave_split_against_lehmann(harness, tmp_path):
    """The default kind of split: the 5- and 6-bath models one per wave, the
    7-bath ones one per workgroup (staged rows, LDS hop caches).  Clean, and
    G(i w_n) equal to the oracle's exact Lehmann sums."""
    models = _models((7, 6, 5), per=2)
    wn = (2 * np.arange(16) + 1) * np.pi / 50.0
    path = tmp_path / "split.txt"
    _write(path, models, wn, 1225, 3, 1)
    r = _run(harness, path)
    _clean(r)
    e0, res, g, _ = _parse(r.stdout, len(models), len(wn))
    for i, m in enumerate(models):
        ref = O.siam_gf_t0(m["u"], m["eimp"], m["eb"], m["v"], 0.0, m["n"], m["n"], wn)
        assert np.abs(g[i] - ref).max() <= 1e-9 * np.abs(ref).max(), i


def test_nine_bath_block_rows_against_wave_mode(harness, tmp_path):
    """A 9-bath model (252 dn states: four columns per lane in the
    wave-private row sweep, 210 and 252 up states) solved by one workgroup
    with the LDS paths and by one wave without them: clean, and E0 and
    G(i w_n) the same."""
    models = _models((9,), per=1)
    wn = (2 * np.arange(16) + 1) * np.pi / 50.0
    blk, wav = tmp_path / "block.txt", tmp_path / "wave.txt"
    _write(blk, models, wn, -1, 1, 0)
    _write(wav, models, wn, 63504, 1, 1)
    rb, rw = _run(harness, blk), _run(harness, wav)
    _clean(rb)
    _clean(rw)
    e0b, resb, gb, _ = _parse(rb.stdout, 1, len(wn))
    e0w, resw, gw, _ = _parse(rw.stdout, 1, len(wn))
    assert resb[0, 1] <= 1e-9 and resw[0, 1] <= 1e-9, (resb, resw)
    assert abs(e0b[0] - e0w[0]) <= 1e-12 * abs(e0w[0]), (e0b, e0w)
    assert np.abs(gb - gw).max() <= 1e-9 * np.abs(gw).max()


@pytest.mark.parametrize("split_m", ["96", "0"])
def test_backward_against_finite_differences(harness, tmp_path, split_m):
    """keep_basis and the adjoint backward (FP64 matrix-core products
    emulated with the instruction's lane layout): clean, and dL/d(eps_k, V_k)
    for L = sum_n Re G + Im G equal to central differences of the oracle's
    Lehmann G.  split_m "0" sends every part's K products through the
    device-wide kernels (grouped GEMM, row-block sums, per-model combine),
    the default only parts of at least 96 steps."""
    models = _models((5, 6), per=1)
    wn = (2 * np.arange(8) + 1) * np.pi / 50.0
    path = tmp_path / "keep.txt"
    _write(path, models, wn, 1225, 2, 1, keep=1)
    r = _run(harness, path, env={"CMPY_AMD_BWD_SPLIT_M": split_m})
    _clean(r)
    _, _, _, grads = _parse(r.stdout, len(models), len(wn))
    h = 1e-5
    for i, m in enumerate(models):
        def loss(eb, v):
            gg = O.siam_gf_t0(m["u"], m["eimp"], eb, v, 0.0, m["n"], m["n"], wn)
            return float(np.sum(gg.real + gg.imag))

        for k in range(m["nb"]):
            de = np.zeros(m["nb"])
            de[k] = h
            fd_e = (loss(m["eb"] + de, m["v"]) - loss(m["eb"] - de, m["v"])) / (2 * h)
            fd_v = (loss(m["eb"], m["v"] + de) - loss(m["eb"], m["v"] - de)) / (2 * h)
            ge, gv = grads[(i, k)]
            assert abs(ge - fd_e) <= 1e-6 * max(1.0, abs(fd_e)), (i, k, ge, fd_e)
            assert abs(gv - fd_v) <= 1e-6 * max(1.0, abs(fd_v)), (i, k, gv, fd_v)


def test_planted_slot_overrun_is_reported(tmp_path):
    """A slot that is 256 bytes too short (its last piece, the scalar
    scratch, then runs into the next slot's guard) must be reported: the
    harness finds what a wrong workspace size would do on the GPU."""
    exe = _build_harness().build(
        mutate=("4 * coef_bytes(miter) + 4 * coef_bytes(csteps) + 256;", "4 * coef_bytes(miter) + 4 * coef_bytes(csteps);"),
        name="batch_check_planted")
    models = _models((5,), per=1)
    wn = (2 * np.arange(4) + 1) * np.pi / 50.0
    path = tmp_path / "planted.txt"
    _write(path, models, wn, 1225, 1, 1)
    r = _run(exe, path)
    assert r.returncode != 0
    assert "AddressSanitizer" in r.stderr and "use-after-poison" in r.stderr, r.stderr[-3000:]


def test_chunked_store_matches_single_launch(harness, tmp_path):
    """A ground-state vector store larger than the device offers (here 30 MB
    of "free" memory: a budget of 18 MB against 31 MB for the six models)
    splits the batch into chunks of models, one launch each: clean, and every
    result identical to the single launch."""
    models = _models((7, 6, 5), per=2)
    wn = (2 * np.arange(8) + 1) * np.pi / 50.0
    path = tmp_path / "chunks.txt"
    _write(path, models, wn, 1225, 2, 1)
    one = _run(harness, path)
    _clean(one)
    many = _run(harness, path, env={"EMU_FREE_BYTES": str(30_000_000)})
    _clean(many)
    r1 = _parse(one.stdout, len(models), len(wn))
    r2 = _parse(many.stdout, len(models), len(wn))
    assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[2], r2[2])
    assert "chunks 1" in one.stdout and "chunks 1" not in many.stdout
