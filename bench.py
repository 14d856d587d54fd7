#!/usr/bin/env python3
"""Benchmark of the exact-diagonalisation hot path on one MI355X.

Headline workload: 2D Hubbard model on the 4x4 periodic square lattice, t=1,
U=4, sector n_up = n_dn = 8 (D = 12870^2 = 165,636,900), FP64.  A step is one
sector H.v on a Lanczos-sized vector already resident in HBM; the metric is
H.v GB/s with the algorithmic traffic 16 D bytes per H.v (x read once, y
written once).  The same run also
  * solves the 4x4 ground state with the three-vector Lanczos recurrence
    (up to 300 steps, Ritz residual 1e-10, explicit residual reported),
  * times the CPU restatement (oracle C twin) of the same H.v on the host
    cores as the CPU baseline, and checks the GPU result against it,
  * runs the secondary workloads (1+5-site impurity T = 0 Lehmann G, 12-site
    chain GS + CF G, 12-site chain real-time G, finite-T Lehmann G of a
    1+6-site impurity model, off-diagonal G_ij of a 7-site ring, batched
    Anderson solves, bath-fit gradients, Bethe-lattice DMFT sweep), each with
    its CPU baseline, unless --no-secondary is given.

CPU baselines run on every core this process may use (capped by
OMP_NUM_THREADS): the oracle's C twin with OpenMP, or the oracle's numpy/scipy
port in a pool of single-threaded processes over independent work (models,
frequency chunks); each is one warm-up and the median of 5 timed runs.

Prints ONE JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent


# secondary workloads, in the order they run
SECONDARY = {"siam6_lehmann": "bench_cfg1", "chain12_solves": "bench_chain12", "chain12_realtime_g": "bench_tevo12",
             "finite_t_lehmann": "bench_finite_t", "offdiagonal_gij": "bench_gij",
             "batched_anderson": "bench_batched", "bathfit_gradients": "bench_bathfit",
             "dmft_bethe_sweep": "bench_dmft"}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--lanczos-iters", type=int, default=300)
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-lanczos", action="store_true")
    p.add_argument("--no-secondary", action="store_true")
    p.add_argument("--size", choices=["4x4", "4x3"], default="4x4")
    p.add_argument("--only-secondary", choices=sorted(SECONDARY), default=None,
                   help="run one secondary workload alone (no headline), e.g. for a profiler")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run device copies of one sector vector (the reference kernel for "
                        "scaling the FETCH_SIZE / WRITE_SIZE counters to bytes)")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def model_spec(size):
    from cmpy_amd.models import square_neighbors

    nx, ny = (4, 4) if size == "4x4" else (4, 3)
    ns = nx * ny
    bonds = square_neighbors(nx, ny, periodic=True)
    n = ns // 2
    return ns, n, bonds


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    import numpy as np
    import torch

    import cmpy_amd
    from cmpy_amd import SectorHamiltonOperator, groundstate

    if args.only_secondary:
        name = args.only_secondary
        print(json.dumps({"secondary": {name: globals()[SECONDARY[name]](args, log)}}), flush=True)
        return
    ns, n, bonds = model_spec(args.size)
    if args.copy_calibration:
        from math import comb

        D = comb(ns, n) ** 2
        x = torch.rand(D, dtype=torch.float64, device="cuda") - 0.5
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"dim": D, "bytes_read": 8 * D, "bytes_written": 8 * D,
                                               "copies": 3}}), flush=True)
        return
    U, t = 4.0, 1.0
    log(f"building the {args.size} operator (sector ({n},{n}))")
    t0 = time.perf_counter()
    op = SectorHamiltonOperator(ns, n, n, bonds, [t] * len(bonds), np.zeros(ns), np.full(ns, U), ns)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    D = op.size
    log(f"D = {D}, tables built in {build_s:.3f} s")

    g = torch.Generator(device="cuda")
    g.manual_seed(1234)
    x = torch.rand(D, dtype=torch.float64, device="cuda", generator=g) - 0.5
    x /= torch.linalg.vector_norm(x)
    # the vector layout the library's Lanczos uses: rows of num_dn states
    # padded to whole 128-byte lines (zeros), ld doubles per row
    ld = op.vector_ld
    xp = torch.zeros((op.num_up, ld), dtype=torch.float64, device="cuda")
    xp[:, :op.num_dn] = x.view(op.num_up, op.num_dn)
    yp = torch.empty_like(xp)

    for _ in range(args.warmup):
        op.matvec_device_ld(xp, ld, out=yp)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record()
        op.matvec_device_ld(xp, ld, out=yp)
        ends[k].record()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kern_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    avg_kern_ms = float(np.mean(kern_ms))
    alg_bytes = 16.0 * D
    value = alg_bytes * args.steps / elapsed / 1e9
    achieved = alg_bytes / (avg_kern_ms * 1e-3) / 1e9
    peak = 8000.0
    log(f"H.v: {elapsed / args.steps * 1e3:.3f} ms/step wall, {avg_kern_ms:.3f} ms kernel, {achieved:.0f} GB/s")

    traffic = None
    pmc = ROOT / "profiles" / f"hv_{args.size}_pmc.json"
    if pmc.exists():
        try:
            traffic = json.loads(pmc.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    result = {
        "metric": "H·v HBM GB/s (4x4 periodic Hubbard, sector (8,8), FP64)" if args.size == "4x4"
        else "H·v HBM GB/s (4x3 periodic Hubbard, sector (6,6), FP64)",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (random unit vector, seed 1234)",
        "config": {"workload": f"hubbard-{args.size}-periodic-sector-hv", "num_sites": ns,
                   "n_up": n, "n_dn": n, "dim": D, "U": U, "t": t, "bonds": len(bonds),
                   "row_stride": ld},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": peak, "unit": "GB/s",
                     "frac": round(achieved / peak, 4), "traffic": traffic,
                     "alg_bytes_per_launch": alg_bytes, "kernel_ms": round(avg_kern_ms, 4)},
    }

    # -- ground state: three-vector Lanczos, up to 300 steps -------------------------
    if not args.no_lanczos:
        log("Lanczos ground state")
        gs, st = groundstate(op, tol=1e-10, max_iter=args.lanczos_iters, reorth=False, check_every=10,
                             want_vector=True, device=True, return_stats=True, seed=99, time_matvec=True)
        result["lanczos"] = {
            "e0": gs.energy, "iterations": st.iterations, "converged": st.converged,
            "ritz_residual": st.ritz_residual, "residual": st.residual,
            "ms_per_iteration": round(st.total_ms / max(1, 2 * st.iterations), 4),
            "total_s": round(st.total_ms / 1e3, 3),
            "hv_gbs_in_solver": round(alg_bytes * st.iterations / (st.matvec_ms * 1e-3) / 1e9, 2)
            if st.matvec_ms > 0 else None,
        }
        log(f"E0 = {gs.energy:.12f}, {st.iterations} steps, residual {st.residual:.2e}")
        del gs

    # -- CPU baseline: the oracle's C twin on the host cores -------------------------
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT))
        from oracle import twin

        cores = _cpu_threads()
        _cpu_all_cores()  # OMP_NUM_THREADS is read when the twin is first loaded
        log(f"CPU baseline on {cores} cores")
        tw = twin.TwinOperator(ns, n, n, bonds, [t] * len(bonds), np.zeros(ns), np.full(ns, U), ns)
        xh = x.cpu().numpy()
        yh = np.empty_like(xh)
        tw.matvec(xh, out=yh)  # warm-up (page faults)
        ts, t_cpu = [], 0.0
        while t_cpu < args.cpu_seconds or len(ts) < 5:
            t0 = time.perf_counter()
            tw.matvec(xh, out=yh)
            ts.append(time.perf_counter() - t0)
            t_cpu += ts[-1]
            if len(ts) >= 25:
                break
        t_med = float(np.median(ts))
        op.matvec_device_ld(xp, ld, out=yp)
        yg = yp[:, :op.num_dn].cpu().numpy().ravel()
        rel = float(np.max(np.abs(yg - yh)) / np.max(np.abs(yh)))
        result["cpu_baseline"] = {
            "value": round(alg_bytes / t_med / 1e9, 3), "unit": "GB/s", "cores": twin.threads(),
            "kind": "port",
            "sample": f"median of {len(ts)} full H.v of the same sector (D={D}) by the oracle's C twin "
                      f"(OpenMP, one thread per core), after one warm-up",
            "ms_per_hv": round(t_med * 1e3, 2),
        }
        result["parity"] = {"hv_max_rel_diff_vs_cpu_twin": rel}
        log(f"CPU {t_med * 1e3:.1f} ms/H.v (median of {len(ts)}); max rel diff GPU vs CPU {rel:.2e}")
        del tw, xh, yh

    if not args.no_cpu:
        result["host"] = _host_info()

    if not args.no_secondary:
        result["secondary"] = {}
        for name, fn_name in SECONDARY.items():
            try:
                result["secondary"][name] = globals()[fn_name](args, log)
            except Exception as exc:  # report, never hide the headline
                result["secondary"][name] = {"error": repr(exc)}

    print(json.dumps(result), flush=True)
    if _POOL is not None:
        _POOL.shutdown()


def _host_info():
    """CPU model, cores available to this process, and the BLAS numpy uses
    (the context of every cpu_baseline figure)."""
    info = {"cores_available": len(os.sched_getaffinity(0)), "cpu_model": None, "blas": None}
    try:
        with open("/proc/cpuinfo", encoding="utf-8") as f:
            for line in f:
                if line.startswith("model name"):
                    info["cpu_model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        from threadpoolctl import threadpool_info

        info["blas"] = sorted({f"{p.get('internal_api')} {p.get('version')}" for p in threadpool_info()
                               if p.get("user_api") == "blas"})
    except ImportError:  # pragma: no cover
        pass
    return info


def _cpu_threads() -> int:
    """Host threads the CPU baselines use: the cores this process may run on,
    capped by OMP_NUM_THREADS when that is set (the GPU host's CPU share)."""
    n = len(os.sched_getaffinity(0))
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, n)


def _cpu_all_cores():
    """OpenMP and BLAS threads pinned to _cpu_threads() for a CPU baseline."""
    n = _cpu_threads()
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ[k] = str(n)
    try:
        from threadpoolctl import threadpool_limits

        return threadpool_limits(limits=n)
    except ImportError:  # pragma: no cover
        import contextlib

        return contextlib.nullcontext()


def _median_time(fn, runs=5, warm=None):
    """One warm-up call (``warm`` or ``fn``), then the median wall time of
    ``runs`` calls; returns (median seconds, all run times, last result)."""
    import numpy as np

    (warm or fn)()
    ts, out = [], None
    for _ in range(runs):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), [round(t, 4) for t in ts], out


_POOL = None


def _worker_init():
    # one thread per worker process: the pool itself spreads the work over the cores
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ[k] = "1"
    if str(ROOT) not in sys.path:
        sys.path.insert(0, str(ROOT))


def _pool():
    """Process pool of _cpu_threads() single-threaded workers (spawned, so
    they import only numpy/scipy and the oracle, never the GPU runtime)."""
    global _POOL
    if _POOL is None:
        import multiprocessing as mp
        from concurrent.futures import ProcessPoolExecutor

        _POOL = ProcessPoolExecutor(max_workers=_cpu_threads(), mp_context=mp.get_context("spawn"),
                                    initializer=_worker_init)
    return _POOL


def _pool_map(fn, items):
    return list(_pool().map(fn, items))


def _chunks(a, n):
    import numpy as np

    return [c for c in np.array_split(np.asarray(a), n) if len(c)]


# -- CPU-baseline work units (run in the pool's worker processes) -----------------

def _w_cfg1(seed):
    """The oracle's port of one 1+5-site impurity solve: dense ground state of
    (3,3), then the T = 0 Lehmann G over (4,3) and (2,3) on 1000 frequencies."""
    import numpy as np

    from oracle import ed_oracle as O

    eps, v = _cfg1_bath(seed)
    mk = lambda nu, nd: O.siam_operator(nu, nd, 4.0, -2.0, eps, v, 0.0)
    op = mk(3, 3)
    e0, psi = O.groundstate(op)
    z = np.linspace(-8, 8, 1000) + 0.05j
    cd = O.apply_cdag(psi, 6, 3, 3, 0, O.UP)
    cc = O.apply_c(psi, 6, 3, 3, 0, O.UP)
    return O.lehmann_t0(op, mk(4, 3), mk(2, 3), psi, e0, cd, cc, z)


def _w_impurity(args):
    """One impurity solve of the oracle's port (see _cpu_impurity_solve)."""
    from oracle import ed_oracle as O

    u, eimp, eb, v, nu, nd, wn, steps = args
    return _cpu_impurity_solve(O, u, eimp, eb, v, nu, nd, wn, steps)


def _w_bathfit_fd(args):
    """Central differences (h = 1e-5) of one model's bath-fit loss in its 14
    parameters with the oracle's port of the forward (28 solves)."""
    import numpy as np

    from oracle import ed_oracle as O

    eb, v, target, wn, steps = args
    nb, h = len(eb), 1e-5

    def loss(e, vv):
        _, g = _cpu_impurity_solve(O, 3.0, -1.5, e, vv, 4, 4, wn, steps)
        return float(np.sum(np.abs(g - target) ** 2))

    out = np.zeros(2 * nb)
    for k in range(2 * nb):
        d = np.zeros(2 * nb)
        d[k] = h
        out[k] = (loss(eb + d[:nb], v + d[nb:]) - loss(eb - d[:nb], v - d[nb:])) / (2 * h)
    return out


def _w_finite_t(zs):
    """The oracle's gf_lehmann (every sector diagonalised) on a chunk of the
    frequencies; returns (G, occ, occ_double)."""
    import numpy as np

    from oracle import ed_oracle as O

    eps, v = _finite_t_bath()
    mk = lambda nu, nd: O.siam_operator(nu, nd, 4.0, -2.0, eps, v, 0.0)
    m = O.gf_lehmann(mk, 7, np.asarray(zs), 20.0)
    return m.gf, m.occ, m.occ_double


def _w_gij(zs):
    """The oracle's off-diagonal G_0j, j = 0..3, of the 7-site ring on a chunk
    of the frequencies, sector eigenpairs shared between the four elements."""
    import numpy as np

    from oracle import ed_oracle as O

    bonds = O.chain_bonds(7, periodic=True)
    mk = lambda nu, nd: O.hubbard_operator(7, nu, nd, bonds, 4.0, 0.0, 1.0, mu=2.0)
    return O.lehmann_ij_many(mk, 7, np.asarray(zs), 10.0, [(0, j) for j in range(4)])


def _w_gij_timed(zs):
    """_w_gij and its own wall time."""
    t0 = time.perf_counter()
    g = _w_gij(zs)
    return g, time.perf_counter() - t0


def _w_tevo_part(args):
    """One part (greater / lesser) of the 12-site chain's real-time G by the
    oracle's port of the reference procedure (scipy expm_multiply)."""
    from oracle import ed_oracle as O

    u, create, e0, psi, stop, num = args
    ns, n = 12, 6
    bonds = O.chain_bonds(ns)
    phi = (O.apply_cdag if create else O.apply_c)(psi, ns, n, n, 0, O.UP)
    opx = O.hubbard_operator(ns, n + (1 if create else -1), n, bonds, u, 0.0, 1.0)
    return O.tevo_expm_multiply(opx, phi, e0, 0.0, stop, num, +1 if create else -1)


def _w_dmft_model(args):
    """One impurity solve of the DMFT sweep by the oracle (fermionic signs,
    full-Krylov reorthogonalised continued fractions)."""
    import numpy as np

    from oracle import ed_oracle as O

    u, eb, v, nb, wn = args
    n = (nb + 1) // 2
    mk = lambda nu, nd: O.siam_operator(nu, nd, u, 0.0, eb, v, u / 2, width=nb + 1)
    e0, psi = O.groundstate(mk(n, n), tol=1e-12)
    g = np.zeros(len(wn), dtype=np.complex128)
    for create, sign in ((True, 1), (False, -1)):
        phi = (O.apply_cdag if create else O.apply_c)(psi, nb + 1, n, n, 0, O.UP, "jordan_wigner")
        opx = mk(n + 1 if create else n - 1, n)
        a, b, n2, _ = O.lanczos_normalized(opx.matvec, phi, opx.size, reorth=True)
        g += O.cf_greens(a, b, n2, 1j * wn, e0, sign)
    return g


def _cfg1_bath(seed):
    import numpy as np

    rng = np.random.default_rng(seed)
    eps = rng.uniform(-2, 2, 5)
    v = rng.uniform(0.2, 1.0, 5)
    return eps, v


def _finite_t_bath():
    import numpy as np

    rng = np.random.default_rng(0)
    return rng.uniform(-2, 2, 6), rng.uniform(0.2, 1.0, 6)


def _cpu_impurity_solve(O, u, eimp, eb, v, nu, nd, wn, cf_steps=(200, 200)):
    """The oracle's port of one impurity solve: Lanczos/dense ground state,
    then particle and hole continued fractions at the Matsubara points, with
    cf_steps (particle, hole) Lanczos steps (the device's adaptive counts when
    it is the baseline of a device solve, so both do the same work)."""
    import numpy as np

    ns = len(eb) + 1
    op = O.siam_operator(nu, nd, u, eimp, eb, v, 0.0)
    e0, psi = O.groundstate(op, tol=1e-10)
    g = np.zeros(len(wn), dtype=np.complex128)
    for k, (create, sign) in enumerate(((True, 1), (False, -1))):
        nn = nu + 1 if create else nu - 1
        if nn < 0 or nn > ns or cf_steps[k] <= 0:
            continue
        phi = (O.apply_cdag if create else O.apply_c)(psi, ns, nu, nd, 0, O.UP)
        opx = O.siam_operator(nn, nd, u, eimp, eb, v, 0.0)
        a, b, n2, _ = O.lanczos_normalized(opx.matvec, phi, int(min(cf_steps[k], opx.size)))
        g += O.cf_greens(a, b, n2, 1j * wn, e0, sign)
    return e0, g


def bench_cfg1(args, log):
    """Single-orbital Anderson impurity, 1 + 5 sites, sector (3,3) (D = 400),
    U = 4, eps_imp = -2, mu = 0, bath eps_k ~ U[-2, 2], V_k ~ U[0.2, 1] from
    seeds 0..9: Lanczos ground state, then the T = 0 Lehmann G of site 0 (up)
    over the (4,3) and (2,3) sectors on 1000 real frequencies in [-8, 8] + 0.05i
    (cmpy/exactdiag.py:24-57, 110-134 restricted to the ground state).  A solve
    is one model's ground state + G."""
    import numpy as np
    import torch

    from cmpy_amd import map_streams
    from cmpy_amd.exactdiag import greens_lehmann_t0, groundstate
    from cmpy_amd.models import SingleImpurityAndersonModel

    seeds = list(range(10))
    z = np.linspace(-8, 8, 1000) + 0.05j
    models = [SingleImpurityAndersonModel(4.0, -2.0, *_cfg1_bath(sd), mu=0.0) for sd in seeds]

    def solve(m):
        gs = groundstate(m, 3, 3, tol=1e-10, device=True)
        return greens_lehmann_t0(m, z, 3, 3, gs=gs)

    def solve_all():
        gl = [solve(m) for m in models]
        torch.cuda.synchronize()
        return gl

    def solve_concurrent():
        gl = map_streams(solve, models, num_streams=5)
        torch.cuda.synchronize()
        return gl

    t, ts, g_gpu = _median_time(solve_all)
    tcc, tscc, g_cc = _median_time(solve_concurrent)
    cc_diff = float(max(np.abs(a - b).max() for a, b in zip(g_gpu, g_cc)))
    out = {"metric": "solves/s", "value": round(len(seeds) / tcc, 2), "unit": "solves/s", "models": len(seeds),
           "dim": 400, "frequencies": len(z), "seconds_per_solve": round(tcc / len(seeds), 5), "runs_s": tscc,
           "concurrent": {"streams": 5, "max_abs_diff_vs_sequential": cc_diff},
           "sequential": {"value": round(len(seeds) / t, 2), "runs_s": ts},
           "data": "synthetic (bath from numpy default_rng(seed), seeds 0..9)"}
    log(f"1+5-site impurity, T = 0 Lehmann: {out['value']} solves/s on 5 streams, "
        f"{out['sequential']['value']} one at a time")
    if not args.no_cpu:
        cores = _cpu_threads()
        with _cpu_all_cores():
            tc, tsc, g_cpu = _median_time(lambda: _pool_map(_w_cfg1, seeds))
        out["parity"] = {"g_max_abs_diff": float(max(np.abs(a - b).max() for a, b in zip(g_gpu, g_cpu))),
                         "g_max_abs": float(max(np.abs(b).max() for b in g_cpu))}
        out["cpu_baseline"] = {"value": round(len(seeds) / tc, 3), "unit": "solves/s",
                               "cores": min(cores, len(seeds)), "kind": "port", "runs_s": tsc,
                               "sample": "the same 10 solves per run, one per single-threaded process: the "
                                         "oracle's port (dense eigh of (3,3), (4,3), (2,3), ladder operators, "
                                         "T = 0 Lehmann sum); median of 5 after a warm-up"}
    return out


def bench_chain12(args, log):
    """12-site open Hubbard chain, t = 1, U in {2, 4, 8}, sector (6,6)
    (D = 853,776): ground state to residual 1e-10, then the 200-step particle
    and hole continued fractions of G_00 on 2048 real frequencies in [-10, 10]
    with eta = 0.05.  The three solves are timed one after another (per_u,
    sequential) and together, each on a HIP stream of its own (map_streams):
    the leg's value is the concurrent rate."""
    import time

    import numpy as np
    import torch

    from cmpy_amd import map_streams
    from cmpy_amd.exactdiag import greens_cf, groundstate
    from cmpy_amd.models import HubbardModel, chain_neighbors

    ns, n = 12, 6
    z = np.linspace(-10, 10, 2048) + 0.05j
    per_u, models, g_seq = {}, {}, {}
    for U in (2.0, 4.0, 8.0):
        model = HubbardModel(ns, chain_neighbors(ns, periodic=False), inter=U, eps=0.0, hop=1.0, mu=0.0)
        models[U] = model
        greens_cf(model, z, n, n, nsteps=200, gs=groundstate(model, n, n, tol=1e-10, device=True))  # warm
        reps = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            gs, st = groundstate(model, n, n, tol=1e-10, return_stats=True, device=True)
            cf = greens_cf(model, z, n, n, nsteps=200, gs=gs)
            torch.cuda.synchronize()
            reps.append(time.perf_counter() - t0)
        dt = float(np.median(reps))
        g_seq[U] = cf.gf
        # H.v bytes actually streamed at 16 D each (read x, write y): the
        # ground-state steps plus its residual check in (6,6), D = 853,776,
        # and the particle and hole fractions in (7,6) / (5,6), D = 731,808
        n_cf = len(cf.alpha_p) + len(cf.alpha_h)
        hv_bytes = 16 * (853776 * (st.iterations + 1) + 731808 * n_cf)
        per_u[U] = {"seconds": round(dt, 4), "repeats": [round(r, 4) for r in reps], "e0": gs.energy,
                    "gs_steps": st.iterations, "cf_steps": n_cf, "residual": st.residual,
                    "hv_bytes": hv_bytes, "hbm_fraction": round(hv_bytes / dt / 8.0e12, 5)}
        if U == 4.0:
            # parity inputs (untimed): coefficients, and G on a broadened axis
            # (eta = 1), where 200 steps have converged whether or not the
            # recurrence is reorthogonalised
            cf_u4 = cf
            zb = np.linspace(-10, 10, 512) + 1.0j
            g_u4_b = greens_cf(model, zb, n, n, nsteps=200, gs=gs).gf
    total = sum(v["seconds"] for v in per_u.values())

    def solve(U):
        gs = groundstate(models[U], n, n, tol=1e-10, device=True)
        return greens_cf(models[U], z, n, n, nsteps=200, gs=gs).gf

    us = list(models)
    map_streams(solve, us)  # warm (per-stream workspaces)
    reps_c = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g_conc = map_streams(solve, us)
        torch.cuda.synchronize()
        reps_c.append(time.perf_counter() - t0)
    t_conc = float(np.median(reps_c))
    # the same kernels in the same order on each stream: identical results
    conc_diff = max(float(np.abs(g - g_seq[U]).max()) for U, g in zip(us, g_conc))
    out = {"metric": "solves/s", "value": round(len(us) / t_conc, 3), "unit": "solves/s", "dim": 853776,
           "concurrent": {"streams": len(us), "seconds": round(t_conc, 4), "repeats": [round(r, 4) for r in reps_c],
                          "max_abs_diff_vs_sequential": conc_diff},
           "sequential": {"value": round(len(per_u) / total, 3), "seconds": round(total, 4)},
           "hbm_fraction": round(sum(v["hv_bytes"] for v in per_u.values()) / t_conc / 8.0e12, 5),
           "hbm_fraction_note": "H.v bytes (16 D per product) per second over 8 TB/s; the 6.8 MB vectors "
                                "are served from L2/MALL, so this workload is latency-bound",
           "per_u": {str(k): v for k, v in per_u.items()}, "cf_steps": 200, "frequencies": 2048,
           "data": "deterministic model (no random input beyond the seeded Lanczos start)"}
    log(f"12-site chain: {out['value']} solves/s on {len(us)} streams, {out['sequential']['value']} one at a time "
        f"(max |G_concurrent - G_sequential| = {conc_diff:.1e})")
    if not args.no_cpu:
        import sys

        sys.path.insert(0, str(ROOT))
        from oracle import ed_oracle as O

        import scipy.sparse.linalg as sla

        from oracle import twin

        bonds = O.chain_bonds(ns)
        with _cpu_all_cores():
            # the oracle's procedure with the C twin's OpenMP H.v as the matvec
            tw = {d: twin.TwinOperator(ns, n + d, n, bonds, [1.0] * len(bonds), np.zeros(ns), np.full(ns, 4.0), ns)
                  for d in (0, 1, -1)}

            def cpu_solve():
                lin = sla.LinearOperator((tw[0].size,) * 2, matvec=tw[0].matvec, dtype=np.float64)
                w, vec = sla.eigsh(lin, k=1, which="SA", tol=1e-10)
                e0, psi = float(w[0]), vec[:, 0]
                coef = {}
                for create, sign in ((True, 1), (False, -1)):
                    # pos 0, spin up: the reference's c+/c and the Jordan-Wigner
                    # form coincide
                    phi = (O.apply_cdag if create else O.apply_c)(psi, ns, n, n, 0, O.UP)
                    a, b, n2, _ = O.lanczos_normalized(tw[1 if create else -1].matvec, phi, 200)
                    coef[create] = (a, b, n2)
                    O.cf_greens(a, b, n2, z, e0, sign)
                return e0, coef

            t, ts, (e0, coef_cpu) = _median_time(cpu_solve)
            threads = twin.threads()
        # the leading coefficients agree between the device's reorthogonalised
        # and the port's plain recurrence; later ones amplify the last digits of
        # the two ground-state vectors (the moment problem is ill-conditioned)
        k = 4
        coef_diff = max(
            float(np.max(np.abs(cf_u4.alpha_p[:k] - coef_cpu[True][0][:k]))),
            float(np.max(np.abs(cf_u4.beta_p[:k] - coef_cpu[True][1][:k]))),
            float(np.max(np.abs(cf_u4.alpha_h[:k] - coef_cpu[False][0][:k]))),
            float(np.max(np.abs(cf_u4.beta_h[:k] - coef_cpu[False][1][:k]))))
        g_cpu_b = sum(O.cf_greens(*coef_cpu[c], zb, e0, sg) for c, sg in ((True, 1), (False, -1)))
        out["parity_u4"] = {
            "e0_rel_diff": abs(per_u[4.0]["e0"] - e0) / abs(e0),
            "cf_coeff_max_abs_diff_first4": coef_diff,
            "g_eta1_max_abs_diff": float(np.abs(g_u4_b - g_cpu_b).max()),
            "g_eta1_max_abs": float(np.abs(g_cpu_b).max())}
        out["cpu_baseline"] = {"value": round(1.0 / t, 4), "unit": "solves/s", "cores": threads, "kind": "port",
                               "runs_s": ts,
                               "sample": "median of 5 solves (U = 4) after a warm-up, by the oracle's port with "
                                         "the C twin's OpenMP H.v: eigsh ground state + 2 x 200 Lanczos steps + "
                                         "continued fractions"}
    return out


def bench_tevo12(args, log):
    """Real-time G of the 12-site open chain at U = 4, sector (6,6): ground
    state to 1e-10, then g>(t) and g<(t) of site 0 (spin up) on 1000 times in
    [0, 20], as gf_tevo does (cmpy/exactdiag.py:248-307)."""
    import time

    import numpy as np
    import torch

    from cmpy_amd.exactdiag import gf_greater, gf_lesser, groundstate
    from cmpy_amd.models import HubbardModel, chain_neighbors

    ns, n, U = 12, 6, 4.0
    start, stop, num = 0.0, 20.0, 1000
    model = HubbardModel(ns, chain_neighbors(ns, periodic=False), inter=U, eps=0.0, hop=1.0, mu=0.0)

    def solve():
        gs = groundstate(model, n, n, tol=1e-10)
        times, g_gt = gf_greater(model, gs, start, stop, num)
        _, g_lt = gf_lesser(model, gs, start, stop, num)
        return gs, times, g_gt, g_lt

    solve()  # warm (kernels, solver handles)
    reps = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gs, times, g_gt, g_lt = solve()
        torch.cuda.synchronize()
        reps.append(time.perf_counter() - t0)
    dt = float(np.median(reps))
    out = {"metric": "solves/s", "value": round(1.0 / dt, 3), "unit": "solves/s", "dim": 853776,
           "seconds": round(dt, 4), "repeats": [round(r, 4) for r in reps], "times": num, "t_range": [start, stop],
           "g_at_0": [float(g_gt[0].imag), float(g_lt[0].imag)],
           "data": "deterministic model (no random input beyond the seeded Lanczos start)"}
    log(f"12-site chain real-time G: {dt:.3f} s per solve")
    if not args.no_cpu:
        import sys

        sys.path.insert(0, str(ROOT))
        from oracle import ed_oracle as O

        import scipy.sparse.linalg as sla

        from oracle import twin

        # bounded sample of the reference's procedure: the ground state (eigsh
        # on the C twin's OpenMP H.v), and the greater and lesser parts
        # (scipy expm_multiply + overlap; its cost is its own vector work, one
        # core each, run side by side) over the first s_stop of the time
        # range, scaled linearly to [start, stop] (expm_multiply's fixed
        # norm-estimation cost makes that scaling an overestimate)
        s_stop, s_num = 0.25, 13
        bonds = O.chain_bonds(ns)
        with _cpu_all_cores():
            tw0 = twin.TwinOperator(ns, n, n, bonds, [1.0] * len(bonds), np.zeros(ns), np.full(ns, U), ns)

            def cpu_gs():
                lin = sla.LinearOperator((tw0.size,) * 2, matvec=tw0.matvec, dtype=np.float64)
                w, vec = sla.eigsh(lin, k=1, which="SA", tol=1e-10)
                return float(w[0]), vec[:, 0]

            t_gs, ts_gs, (e0, psi) = _median_time(cpu_gs)
            parts = [(U, True, e0, psi, s_stop, s_num), (U, False, e0, psi, s_stop, s_num)]
            t_part, ts_part, res = _median_time(lambda: _pool_map(_w_tevo_part, parts),
                                                warm=lambda: _pool_map(_w_tevo_part, [p[:4] + (0.01, 2) for p in parts]))
        ov = res[0][1]
        est = t_gs + t_part * (stop - start) / (s_stop - start)
        _, g_s = gf_greater(model, gs, start, s_stop, s_num)
        out["parity_sample"] = {"max_abs_diff_vs_expm_multiply": float(np.abs(g_s - ov).max()),
                                "max_abs": float(np.abs(ov).max())}
        out["cpu_baseline"] = {
            "value": round(1.0 / est, 5), "unit": "solves/s", "cores": _cpu_threads(), "kind": "port",
            "runs_s": {"ground_state": ts_gs, "parts": ts_part},
            "sample": f"the oracle's port of the reference procedure, medians of 5 runs after a warm-up: ground "
                      f"state by eigsh on the C twin's OpenMP H.v ({t_gs:.2f} s), greater and lesser parts "
                      f"(scipy expm_multiply + overlap, one process each, side by side) over t in [0, {s_stop}] "
                      f"({s_num} points, {t_part:.2f} s), scaled linearly to [0, {stop}]"}
    return out


def bench_finite_t(args, log):
    """Grand-canonical finite-T Lehmann G of a 1+6-site Anderson impurity model
    (U = 4, eps_imp = -2, mu = 0, bath from seed 0), beta = 20, G_00 up on
    1000 frequencies in [-8, 8] + 0.05i, every sector diagonalised densely
    (cmpy/exactdiag.py:137-245; largest sector D = 1225)."""
    import time

    import numpy as np
    import torch

    from cmpy_amd.exactdiag import gf_lehmann
    from cmpy_amd.models import SingleImpurityAndersonModel

    rng = np.random.default_rng(0)
    nb = 6
    eps, v = rng.uniform(-2, 2, nb), rng.uniform(0.2, 1.0, nb)
    U, eimp, beta = 4.0, -2.0, 20.0
    z = np.linspace(-8, 8, 1000) + 0.05j
    model = SingleImpurityAndersonModel(U, eimp, eps, v, mu=0.0)
    gf_lehmann(model, z[:8], beta)  # warm (kernels, solver handle)
    reps = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        data = gf_lehmann(model, z, beta)
        g = data.gf
        torch.cuda.synchronize()
        reps.append(time.perf_counter() - t0)
    dt = float(np.median(reps))
    out = {"metric": "solves/s", "value": round(1.0 / dt, 4), "unit": "solves/s", "seconds": round(dt, 4),
           "repeats": [round(r, 4) for r in reps],
           "sites": nb + 1, "sectors": (nb + 2) ** 2, "max_dim": 1225, "beta": beta, "frequencies": len(z),
           "occ": data.occ, "occ_double": data.occ_double,
           "data": "synthetic (bath from numpy default_rng(0))"}
    log(f"finite-T Lehmann (1+{nb} sites): {dt:.3f} s per solve")
    if not args.no_cpu:
        import sys

        sys.path.insert(0, str(ROOT))
        from oracle import ed_oracle as O

        cores = _cpu_threads()
        with _cpu_all_cores():
            chunks = _chunks(z, cores)
            t, ts, res = _median_time(lambda: _pool_map(_w_finite_t, chunks),
                                      warm=lambda: _pool_map(_w_finite_t, _chunks(z[:cores], cores)))
        g_cpu = np.concatenate([r[0] for r in res])
        occ_cpu, dbl_cpu = res[0][1], res[0][2]
        out["parity_sample"] = {"max_abs_diff": float(np.abs(g - g_cpu).max()),
                                "max_abs": float(np.abs(g_cpu).max()),
                                "occ_diff": abs(data.occ - occ_cpu), "occ_double_diff": abs(data.occ_double - dbl_cpu)}
        out["cpu_baseline"] = {
            "value": round(1.0 / t, 5), "unit": "solves/s", "cores": cores, "kind": "port", "runs_s": ts,
            "sample": f"the whole solve, median of 5 after a warm-up: the oracle's gf_lehmann (numpy eigh of every "
                      f"sector + vectorised pair sum) in {cores} single-threaded processes, each on 1/{cores} of "
                      f"the {len(z)} frequencies (each process diagonalises every sector itself)"}
    return out


def bench_gij(args, log):
    """Off-diagonal finite-T Green's functions of a 7-site Hubbard ring (U = 4,
    mu = U/2, beta = 10): G_0j for j = 0..3 on 1000 frequencies in [-8, 8] +
    0.05i, every sector diagonalised once and shared (lehmann_full.py:19-97)."""
    import time

    import numpy as np
    import torch

    from cmpy_amd.exactdiag import greens_lehmann
    from cmpy_amd.models import HubbardModel, chain_neighbors

    ns, U, beta = 7, 4.0, 10.0
    z = np.linspace(-8, 8, 1000) + 0.05j
    model = HubbardModel(ns, chain_neighbors(ns, periodic=True), inter=U, eps=0.0, hop=1.0, mu=U / 2)
    greens_lehmann(model, z[:8], beta, 0, 1)  # warm
    reps = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cache = {}
        gs = [greens_lehmann(model, z, beta, 0, j, eig_cache=cache).gf for j in range(4)]
        torch.cuda.synchronize()
        reps.append(time.perf_counter() - t0)
    dt = float(np.median(reps))
    out = {"metric": "solves/s", "value": round(1.0 / dt, 4), "unit": "solves/s", "seconds": round(dt, 4),
           "repeats": [round(r, 4) for r in reps],
           "sites": ns, "elements": 4, "beta": beta, "frequencies": len(z), "max_dim": 1225,
           "data": "deterministic model"}
    log(f"G_ij (7-site ring, 4 elements): {dt:.3f} s per solve")
    if not args.no_cpu:
        import sys

        sys.path.insert(0, str(ROOT))
        from oracle import ed_oracle as O

        cores = _cpu_threads()
        # sample: every 4th frequency (250), all four elements, split over the
        # processes.  The warm-up (one frequency per process) times the shared
        # part (diagonalising every sector, overlaps); the estimate scales the
        # rest of the slowest process's time to the full 1000 frequencies
        zs = z[::4]
        with _cpu_all_cores():
            chunks = _chunks(zs, cores)
            warm = []
            t, ts, res = _median_time(lambda: _pool_map(_w_gij_timed, chunks),
                                      warm=lambda: warm.extend(_pool_map(_w_gij_timed, _chunks(zs[:cores], cores))))
        g_cpu = [np.concatenate([r[0][j] for r in res]) for j in range(4)]
        shared = float(np.median([w[1] for w in warm]))
        pair_part = max(0.0, max(r[1] for r in res) - shared)
        est = t + (len(z) / len(zs) - 1.0) * pair_part
        out["parity_sample"] = {"max_abs_diff": float(max(np.abs(gs[j][::4] - g_cpu[j]).max() for j in range(4))),
                                "max_abs": float(max(np.abs(g_cpu[j]).max() for j in range(4)))}
        out["cpu_baseline"] = {
            "value": round(1.0 / est, 5), "unit": "solves/s", "cores": cores, "kind": "port", "runs_s": ts,
            "sample": f"the oracle's lehmann_ij (numpy eigh of every sector, overlaps, pair sums; four elements "
                      f"sharing the eigenpairs) in {cores} single-threaded processes on 250 of the {len(z)} "
                      f"frequencies, median of 5 after a warm-up ({t:.2f} s), the slowest process's pair-sum "
                      f"time ({pair_part:.2f} s) scaled to all {len(z)} frequencies"}
    return out


def bench_batched(args, log):
    """512 Anderson impurity models, bath sizes 5..9, ground state to 1e-10 and
    G(i w_n) at 256 Matsubara frequencies (beta = 50), one persistent launch."""
    import time

    import numpy as np

    from cmpy_amd.batch import AndersonBatch, matsubara, random_batch
    from cmpy_amd.exactdiag import greens_cf, groundstate
    from cmpy_amd.models import SingleImpurityAndersonModel

    M, seed = 512, 2024
    p = random_batch(M, seed)
    wn = matsubara(50.0, 256)
    b = AndersonBatch(**p)
    t0 = time.perf_counter()
    r = b.solve(wn)  # first solve autotunes the launch shape
    tune_s = time.perf_counter() - t0
    ms = []
    for _ in range(5):
        r = b.solve(wn)
        ms.append(r.elapsed_ms)
    ms_med = float(np.median(ms))
    out = {"metric": "solves/s", "value": round(M / (ms_med * 1e-3), 1), "unit": "solves/s", "models": M,
           "matsubara": 256, "beta": 50.0, "ms_per_batch": round(ms_med, 3), "tune_s": round(tune_s, 2),
           "launch": b.info(), "max_residual": float(r.residual.max()),
           "ground_state_steps_mean": float(r.iterations.mean()),
           "cf_steps_mean": [float(x) for x in r.cf_steps.mean(axis=0)],
           "data": f"synthetic (random_batch seed {seed})"}
    # algorithmic bytes: per model (GS steps * D + CF steps * D_cf) * 40 bytes
    # (16 D for the H.v, 24 D for the vector update), particle sector
    # (n_up + 1, n_dn) and hole sector (n_up - 1, n_dn)
    from math import comb

    ns = np.asarray(b.num_bath) + 1
    d_p = np.array([comb(int(n), int(a) + 1) * comb(int(n), int(c)) for n, a, c in zip(ns, b.n_up, b.n_dn)])
    d_h = np.array([comb(int(n), int(a) - 1) * comb(int(n), int(c)) if a > 0 else 0
                    for n, a, c in zip(ns, b.n_up, b.n_dn)])
    alg = 40.0 * float(np.sum(b.dims * r.iterations + d_p * r.cf_steps[:, 0] + d_h * r.cf_steps[:, 1]))
    out["algorithmic_bytes"] = alg
    out["hbm_fraction"] = round(alg / (ms_med * 1e-3) / 8.0e12, 5)
    log(f"batched Anderson: {out['value']} solves/s ({ms_med:.1f} ms per 512)")
    # plain per-model GPU path on a sample: one model, one call chain at a
    # time (models built and warmed outside the timer; their sector
    # operators stay cached on the model)
    sample = list(range(0, M, M // 16))
    models = []
    for i in sample:
        m = SingleImpurityAndersonModel(u=p["u"][i], eps_imp=p["eps_imp"][i], eps_bath=p["eps_bath"][i],
                                        v=p["v"][i], mu=0.0)
        nu = (len(p["eps_bath"][i]) + 1) // 2
        models.append((m, nu))

    def plain():
        for m, nu in models:
            gs = groundstate(m, nu, nu, tol=1e-10)
            greens_cf(m, 1j * wn, nu, nu, gs=gs)

    t_plain, ts_plain, _ = _median_time(plain, runs=3)
    out["plain_gpu"] = {"value": round(len(sample) / t_plain, 2), "unit": "solves/s", "sample": len(sample),
                        "runs_s": ts_plain,
                        "note": "groundstate + greens_cf per model on prebuilt, warmed models (median of 3)"}
    out["speedup_vs_plain"] = round(out["value"] / out["plain_gpu"]["value"], 1)
    if not args.no_cpu:
        cores = _cpu_threads()
        csample = list(range(0, M, max(1, M // (2 * cores))))
        work = [(p["u"][i], p["eps_imp"][i], p["eps_bath"][i], p["v"][i], (len(p["eps_bath"][i]) + 1) // 2,
                 (len(p["eps_bath"][i]) + 1) // 2, wn, tuple(int(x) for x in r.cf_steps[i])) for i in csample]
        with _cpu_all_cores():
            t, ts, _ = _median_time(lambda: _pool_map(_w_impurity, work),
                                    warm=lambda: _pool_map(_w_impurity, work[:cores]))
        out["cpu_baseline"] = {"value": round(len(work) / t, 3), "unit": "solves/s", "cores": cores, "kind": "port",
                               "runs_s": ts,
                               "sample": f"{len(work)} models of the same batch (every {M // len(csample)}th) per run, "
                                         f"median of 5 after a warm-up, by the oracle's numpy/scipy port (ground "
                                         f"state + continued fractions with the device's step counts) in {cores} "
                                         f"single-threaded processes"}
    return out


def bench_dmft(args, log):
    """Matsubara-axis DMFT of the half-filled Bethe-lattice Hubbard model
    (half bandwidth 2): a sweep of 8 U values in [0.5, 7.5] iterated together,
    5 bath sites (D = 400), beta = 50, 128 Matsubara frequencies, bath update
    by hybridisation fit, to max |dG| < 1e-5 (cmpy/dmft/twosite.py:103-172 is
    the reference's two-site loop).  Unit: DMFT iterations per second over a
    fixed 30 iterations, one iteration = 8 impurity solves + the
    self-consistency and bath fits."""
    import time

    import numpy as np

    from cmpy_amd.dmft import bethe_dmft_bath_fit

    us = np.linspace(0.5, 7.5, 8)
    beta, nb, nw = 50.0, 5, 128
    bethe_dmft_bath_fit(us, beta, num_bath=nb, n_matsubara=nw, max_iter=1)  # warm
    t0 = time.perf_counter()
    r = bethe_dmft_bath_fit(us, beta, num_bath=nb, n_matsubara=nw, tol=0.0, max_iter=29)
    dt = time.perf_counter() - t0
    its = r.stats.iteration
    out = {"metric": "DMFT iterations/s", "value": round(its / dt, 3), "unit": "iterations/s", "models": len(us),
           "bath_sites": nb, "beta": beta, "matsubara": nw, "iterations": its, "seconds": round(dt, 3),
           "last_max_dG": r.stats.errors[0], "residual": [round(float(x), 5) for x in r.residual],
           "quasiparticle_weight": [round(float(z), 5) for z in r.quasiparticle_weight],
           "data": "deterministic (particle-hole symmetric start bath)"}
    log(f"DMFT sweep: {its} iterations in {dt:.2f} s")
    if not args.no_cpu:
        import sys

        sys.path.insert(0, str(ROOT))
        from oracle import ed_oracle as O

        # the oracle's impurity solves for one iteration (8 models, fermionic
        # signs as the device loop uses), timed; parity on the converged bath
        wn = r.wn
        n = (nb + 1) // 2
        work = [(u, r.eps_bath[i], r.v[i], nb, wn) for i, u in enumerate(us)]
        with _cpu_all_cores():
            t, ts, g_cpu = _median_time(lambda: _pool_map(_w_dmft_model, work))
        # a model whose sector ground state is (nearly) degenerate -- the Mott
        # side, where a bath level decouples -- has no unique T = 0 G, so the
        # parity figure covers the models with a gap above 1e-6
        gaps = []
        for i, u in enumerate(us):
            ev = np.linalg.eigvalsh(O.siam_operator(n, n, u, 0.0, r.eps_bath[i], r.v[i], u / 2,
                                                    width=nb + 1).csr().toarray())
            gaps.append(float(ev[1] - ev[0]))
        diff = np.abs(np.array(g_cpu) - r.g_imp).max(axis=1)
        ok = np.array(gaps) > 1e-6
        out["parity_sample"] = {"max_abs_diff_g_imp": float(diff[ok].max()) if ok.any() else None,
                                "models_compared": int(ok.sum()),
                                "sector_gaps": [round(g, 8) for g in gaps]}
        out["cpu_baseline"] = {"value": round(1.0 / t, 4), "unit": "iterations/s", "cores": min(len(us), _cpu_threads()),
                               "kind": "port", "runs_s": ts,
                               "sample": f"the oracle's impurity solves of one iteration (8 models, D = 400, "
                                         f"full-Krylov continued fractions, one model per single-threaded "
                                         f"process; median of 5 after a warm-up, {t:.2f} s), bath fits not "
                                         f"included"}
    return out


def bench_bathfit(args, log):
    """Bath-fit gradients: 256 Anderson models, 7 bath sites, sector (4,4)
    (D = 4,900), U = 3, 512 Matsubara frequencies at beta = 100; loss
    sum |G - G_target|^2 with G_target from a second seed's parameters."""
    import time

    import numpy as np

    from cmpy_amd.batch import AndersonBatch, matsubara

    M, nb = 256, 7
    ra, rb = np.random.default_rng(101), np.random.default_rng(202)
    u = np.full(M, 3.0)
    ea, va = ra.uniform(-3, 3, (M, nb)), ra.uniform(0.1, 1.2, (M, nb))
    eb_t, vb_t = rb.uniform(-3, 3, (M, nb)), rb.uniform(0.1, 1.2, (M, nb))
    wn = matsubara(100.0, 512)
    tb = AndersonBatch(u=u, eps_imp=-u / 2, eps_bath=list(eb_t), v=list(vb_t), mu=0.0)
    target = tb.solve(wn).g
    del tb
    b = AndersonBatch(u=u, eps_imp=-u / 2, eps_bath=list(ea), v=list(va), mu=0.0, keep_basis=True)

    def step(batch):
        r = batch.solve(wn)
        d = r.g - target
        ge, gv, cg, ms_b = batch.backward(2 * d.real, 2 * d.imag)
        return r, ms_b, ge, gv

    step(b)  # tune + warm
    fw, bw = [], []
    for _ in range(5):
        r, ms_b, ge, gv = step(b)
        fw.append(r.elapsed_ms)
        bw.append(ms_b)
    ms = float(np.median(np.array(fw) + np.array(bw)))
    out = {"metric": "grads/s", "value": round(M / (ms * 1e-3), 1), "unit": "grads/s", "models": M,
           "bath_sites": nb, "dim": 4900, "matsubara": 512, "beta": 100.0,
           "ms_forward": round(float(np.median(fw)), 3), "ms_backward": round(float(np.median(bw)), 3),
           "cf_steps_max": [int(x) for x in r.cf_steps.max(axis=0)], "cf_capped": int(r.cf_capped.any(axis=1).sum()),
           "data": "synthetic (theta from seed 101, target from seed 202)"}
    log(f"bath-fit gradients: {out['value']} grads/s")
    # gradient check (untimed): central differences (h = 1e-5) of every model's
    # loss, 28 batched solves with a tighter ground state, against the timed
    # run's adjoint gradients
    fb = AndersonBatch(u=u, eps_imp=-u / 2, eps_bath=list(ea), v=list(va), mu=0.0, tol=1e-12)
    h = 1e-5
    fd = np.zeros((M, 2 * nb))
    for kind in range(2):
        for k in range(nb):
            lp = []
            for sgn in (1, -1):
                e2, v2 = ea.copy(), va.copy()
                (e2 if kind == 0 else v2)[:, k] += sgn * h
                fb.set_bath(e2, v2)
                lp.append(np.sum(np.abs(fb.solve(wn).g - target) ** 2, axis=1))
            fd[:, kind * nb + k] = (lp[0] - lp[1]) / (2 * h)
    del fb
    grad = np.concatenate([ge, gv], axis=1)
    rel = np.abs(grad - fd).max(axis=1) / np.abs(fd).max(axis=1)
    out["fd_check"] = {"max_rel_err": float(rel.max()), "models_above_1e-6": int((rel > 1e-6).sum()),
                       "h": h, "note": "per model: max |adjoint - central difference| / max |central difference|"}
    # plain per-model GPU path: each model alone (one model per launch),
    # batches built and warmed (launch shape tuned) outside the timer
    ns = 8
    singles = []
    for i in range(ns):
        bi = AndersonBatch(u=u[i:i + 1], eps_imp=-u[i:i + 1] / 2, eps_bath=[ea[i]], v=[va[i]], mu=0.0,
                           keep_basis=True)
        singles.append((bi, target[i:i + 1]))

    def plain():
        for bi, ti in singles:
            rr = bi.solve(wn)
            d = rr.g - ti
            bi.backward(2 * d.real, 2 * d.imag)

    t_plain, ts_plain, _ = _median_time(plain)
    out["plain_gpu"] = {"value": round(ns / t_plain, 2), "unit": "grads/s", "sample": ns, "runs_s": ts_plain,
                        "note": "one model per batch (solve + backward), prebuilt and warmed, median of 5"}
    out["speedup_vs_plain"] = round(out["value"] / out["plain_gpu"]["value"], 1)
    if not args.no_cpu:
        # CPU: central differences of the oracle's forward (the derivative a
        # CPU fit of the reference would use), one model per process
        cores = _cpu_threads()
        work = [(ea[i], va[i], target[i], wn, tuple(int(x) for x in r.cf_steps[i])) for i in range(cores)]
        with _cpu_all_cores():
            t, ts, _ = _median_time(lambda: _pool_map(_w_bathfit_fd, work))
        out["cpu_baseline"] = {"value": round(len(work) / t, 4), "unit": "grads/s", "cores": cores, "kind": "port",
                               "runs_s": ts,
                               "sample": f"{len(work)} models per run, one per single-threaded process: central "
                                         "differences (28 solves each) with the oracle's numpy/scipy port of the "
                                         "forward (the device's continued-fraction step counts); median of 5 "
                                         "after a warm-up"}
    return out


if __name__ == "__main__":
    main()
