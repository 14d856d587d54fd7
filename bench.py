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
  * runs the secondary workloads (12-site chain GS + CF G, 12-site chain
    real-time G, finite-T Lehmann G of a 1+6-site impurity model, off-diagonal
    G_ij of a 7-site ring, batched Anderson solves, bath-fit gradients), each
    with its CPU baseline, unless --no-secondary is given.

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
SECONDARY = {"chain12_solves": "bench_chain12", "chain12_realtime_g": "bench_tevo12",
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
    p.add_argument("--secondary-seconds", type=float, default=10.0,
                   help="CPU seconds per secondary workload's baseline sample")
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
    y = torch.empty_like(x)

    for _ in range(args.warmup):
        op.matvec_device(x, out=y)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record()
        op.matvec_device(x, out=y)
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
                   "n_up": n, "n_dn": n, "dim": D, "U": U, "t": t, "bonds": len(bonds)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": peak, "unit": "GB/s",
                     "frac": round(achieved / peak, 4), "traffic": traffic,
                     "alg_bytes_per_launch": alg_bytes, "kernel_ms": round(avg_kern_ms, 4)},
    }

    # -- ground state: three-vector Lanczos, up to 300 steps -------------------------
    if not args.no_lanczos:
        log("Lanczos ground state")
        gs, st = groundstate(op, tol=1e-10, max_iter=args.lanczos_iters, reorth=False, check_every=10,
                             want_vector=True, device=True, return_stats=True, seed=99)
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

        cores = len(os.sched_getaffinity(0))
        os.environ.setdefault("OMP_NUM_THREADS", str(cores))
        log(f"CPU baseline on {cores} cores")
        tw = twin.TwinOperator(ns, n, n, bonds, [t] * len(bonds), np.zeros(ns), np.full(ns, U), ns)
        xh = x.cpu().numpy()
        yh = np.empty_like(xh)
        tw.matvec(xh, out=yh)  # warm-up (page faults)
        runs, t_cpu = 0, 0.0
        while t_cpu < args.cpu_seconds or runs < 2:
            t0 = time.perf_counter()
            tw.matvec(xh, out=yh)
            t_cpu += time.perf_counter() - t0
            runs += 1
            if runs >= 20:
                break
        op.matvec_device(x, out=y)
        yg = y.cpu().numpy()
        rel = float(np.max(np.abs(yg - yh)) / np.max(np.abs(yh)))
        result["cpu_baseline"] = {
            "value": round(alg_bytes * runs / t_cpu / 1e9, 3), "unit": "GB/s", "cores": twin.threads(),
            "kind": "port",
            "sample": f"{runs} full H.v of the same sector (D={D}) by the oracle's C twin (OpenMP)",
            "ms_per_hv": round(t_cpu / runs * 1e3, 2),
        }
        result["parity"] = {"hv_max_rel_diff_vs_cpu_twin": rel}
        log(f"CPU {t_cpu / runs * 1e3:.1f} ms/H.v; max rel diff GPU vs CPU {rel:.2e}")
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


def _cpu_single_thread():
    """BLAS limited to one thread for the CPU baselines (cores = 1)."""
    try:
        from threadpoolctl import threadpool_limits

        return threadpool_limits(limits=1)
    except ImportError:  # pragma: no cover
        import contextlib

        return contextlib.nullcontext()


def _cpu_impurity_solve(O, u, eimp, eb, v, nu, nd, wn, cf_steps=200):
    """The oracle's port of one impurity solve: Lanczos/dense ground state,
    then particle and hole continued fractions at the Matsubara points."""
    import numpy as np

    ns = len(eb) + 1
    op = O.siam_operator(nu, nd, u, eimp, eb, v, 0.0)
    e0, psi = O.groundstate(op, tol=1e-10)
    g = np.zeros(len(wn), dtype=np.complex128)
    for create, sign in ((True, 1), (False, -1)):
        nn = nu + 1 if create else nu - 1
        if nn < 0 or nn > ns:
            continue
        phi = (O.apply_cdag if create else O.apply_c)(psi, ns, nu, nd, 0, O.UP)
        opx = O.siam_operator(nn, nd, u, eimp, eb, v, 0.0)
        a, b, n2, _ = O.lanczos_normalized(opx.matvec, phi, min(cf_steps, opx.size))
        g += O.cf_greens(a, b, n2, 1j * wn, e0, sign)
    return e0, g


def bench_chain12(args, log):
    """12-site open Hubbard chain, t = 1, U in {2, 4, 8}, sector (6,6)
    (D = 853,776): ground state to residual 1e-10, then the 200-step particle
    and hole continued fractions of G_00 on 2048 real frequencies in [-10, 10]
    with eta = 0.05."""
    import time

    import numpy as np
    import torch

    from cmpy_amd.exactdiag import greens_cf, groundstate
    from cmpy_amd.models import HubbardModel, chain_neighbors

    ns, n = 12, 6
    z = np.linspace(-10, 10, 2048) + 0.05j
    per_u = {}
    for U in (2.0, 4.0, 8.0):
        model = HubbardModel(ns, chain_neighbors(ns, periodic=False), inter=U, eps=0.0, hop=1.0, mu=0.0)
        greens_cf(model, z, n, n, nsteps=200, gs=groundstate(model, n, n, tol=1e-10))  # warm (tables, kernels)
        reps = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            gs, st = groundstate(model, n, n, tol=1e-10, return_stats=True)
            cf = greens_cf(model, z, n, n, nsteps=200, gs=gs)
            torch.cuda.synchronize()
            reps.append(time.perf_counter() - t0)
        dt = float(np.median(reps))
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
    out = {"metric": "solves/s", "value": round(len(per_u) / total, 3), "unit": "solves/s", "dim": 853776,
           "hbm_fraction": round(sum(v["hv_bytes"] for v in per_u.values()) / total / 8.0e12, 5),
           "hbm_fraction_note": "H.v bytes (16 D per product) per second over 8 TB/s; the 6.8 MB vectors "
                                "are served from L2/MALL, so this workload is latency-bound",
           "per_u": {str(k): v for k, v in per_u.items()}, "cf_steps": 200, "frequencies": 2048,
           "data": "deterministic model (no random input beyond the seeded Lanczos start)"}
    log(f"12-site chain: {out['value']} solves/s")
    if not args.no_cpu:
        import sys

        sys.path.insert(0, str(ROOT))
        from oracle import ed_oracle as O

        with _cpu_single_thread():
            t0 = time.perf_counter()
            bonds = O.chain_bonds(ns)
            op = O.hubbard_operator(ns, n, n, bonds, 4.0, 0.0, 1.0)
            e0, psi = O.groundstate(op, tol=1e-10)
            coef_cpu = {}
            for create, sign in ((True, 1), (False, -1)):
                # pos 0, spin up: the reference's c+/c and the Jordan-Wigner
                # form coincide
                phi = (O.apply_cdag if create else O.apply_c)(psi, ns, n, n, 0, O.UP)
                opx = O.hubbard_operator(ns, n + (1 if create else -1), n, bonds, 4.0, 0.0, 1.0)
                a, b, n2, _ = O.lanczos_normalized(opx.matvec, phi, 200)
                coef_cpu[create] = (a, b, n2)
                O.cf_greens(a, b, n2, z, e0, sign)
            t = time.perf_counter() - t0
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
        out["cpu_baseline"] = {"value": round(1.0 / t, 4), "unit": "solves/s", "cores": 1, "kind": "port",
                               "sample": "1 solve (U = 4) by the oracle's numpy/scipy port: eigsh ground state "
                                         "+ 2 x 200 Lanczos steps + continued fractions"}
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

        # bounded sample of the reference's procedure: the greater part over
        # the first 5% of the time range (expm_multiply cost grows linearly
        # with the range), scaled by 20 x 2 parts
        s_stop, s_num = 1.0, 50
        with _cpu_single_thread():
            bonds = O.chain_bonds(ns)
            op = O.hubbard_operator(ns, n, n, bonds, U, 0.0, 1.0)
            t0 = time.perf_counter()
            e0, psi = O.groundstate(op, tol=1e-10)
            t_gs = time.perf_counter() - t0
            phi = O.apply_cdag(psi, ns, n, n, 0, O.UP)
            opx = O.hubbard_operator(ns, n + 1, n, bonds, U, 0.0, 1.0)
            opx.csr()
            t0 = time.perf_counter()
            ts, ov = O.tevo_expm_multiply(opx, phi, e0, start, s_stop, s_num, +1)
            t_part = time.perf_counter() - t0
        est = t_gs + 2 * t_part * (stop - start) / (s_stop - start)
        _, g_s = gf_greater(model, gs, start, s_stop, s_num)
        out["parity_sample"] = {"max_abs_diff_vs_expm_multiply": float(np.abs(g_s - ov).max()),
                                "max_abs": float(np.abs(ov).max())}
        out["cpu_baseline"] = {
            "value": round(1.0 / est, 5), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"the oracle's port of the reference procedure (scipy expm_multiply + overlap, 1 core): "
                      f"ground state ({t_gs:.1f} s) and the greater part over t in [0, {s_stop}] "
                      f"({s_num} points, {t_part:.1f} s), scaled to both parts over [0, {stop}]"}
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

        mk = lambda nu, nd: O.siam_operator(nu, nd, U, eimp, eps, v, 0.0)
        # bounded samples: every sector, on 10 and on 20 of the frequencies;
        # the cost is fixed (dense eigh, overlaps) plus linear in the count
        ts = {}
        with _cpu_single_thread():
            for nz in (10, 20):
                zs = z[:: len(z) // nz][:nz]
                t0 = time.perf_counter()
                ref = O.gf_lehmann(mk, nb + 1, zs, beta)
                ts[nz] = time.perf_counter() - t0
        per_z = max(ts[20] - ts[10], 0.0) / 10
        fixed = max(ts[10] - 10 * per_z, 0.0)
        est = fixed + per_z * len(z)
        zs = z[:: len(z) // 20][:20]
        out["parity_sample"] = {"max_abs_diff": float(np.abs(g[:: len(z) // 20][:20] - ref.gf).max()),
                                "max_abs": float(np.abs(ref.gf).max()),
                                "occ_diff": abs(data.occ - ref.occ), "occ_double_diff": abs(data.occ_double - ref.occ_double)}
        out["cpu_baseline"] = {
            "value": round(1.0 / est, 5), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"the oracle's gf_lehmann (numpy eigh + vectorised pair sum, 1 core) on every sector at 10 "
                      f"and 20 of the frequencies ({ts[10]:.1f} s, {ts[20]:.1f} s), extrapolated linearly to "
                      f"{len(z)}"}
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

        bonds = O.chain_bonds(ns, periodic=True)
        mk = lambda nu, nd: O.hubbard_operator(ns, nu, nd, bonds, U, 0.0, 1.0, mu=U / 2)
        ts = {}
        with _cpu_single_thread():
            for nz in (10, 20):
                zs = z[:: len(z) // nz][:nz]
                t0 = time.perf_counter()
                ref = O.lehmann_ij(mk, ns, zs, beta, 0, 1)
                ts[nz] = time.perf_counter() - t0
        per_z = max(ts[20] - ts[10], 0.0) / 10
        fixed = max(ts[10] - 10 * per_z, 0.0)
        # one diagonalisation of every sector, four elements' sums
        est = fixed + 4 * per_z * len(z)
        out["parity_sample"] = {"max_abs_diff": float(np.abs(gs[1][:: len(z) // 20][:20] - ref).max()),
                                "max_abs": float(np.abs(ref).max())}
        out["cpu_baseline"] = {
            "value": round(1.0 / est, 5), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"the oracle's lehmann_ij (numpy eigh + pair sums, 1 core) for G_01 on 10 and 20 of the "
                      f"frequencies ({ts[10]:.1f} s, {ts[20]:.1f} s); estimate = fixed part (eigh, overlaps) once "
                      f"+ 4 elements x {len(z)} frequencies"}
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
    # plain per-model GPU path on a sample: one model, one call chain at a time
    sample = list(range(0, M, M // 16))
    t0 = time.perf_counter()
    for i in sample:
        m = SingleImpurityAndersonModel(u=p["u"][i], eps_imp=p["eps_imp"][i], eps_bath=p["eps_bath"][i],
                                        v=p["v"][i], mu=0.0)
        nu = (len(p["eps_bath"][i]) + 1) // 2
        gs = groundstate(m, nu, nu, tol=1e-10)
        greens_cf(m, 1j * wn, nu, nu, gs=gs)
    t_plain = time.perf_counter() - t0
    out["plain_gpu"] = {"value": round(len(sample) / t_plain, 2), "unit": "solves/s", "sample": len(sample)}
    out["speedup_vs_plain"] = round(out["value"] / out["plain_gpu"]["value"], 1)
    if not args.no_cpu:
        import sys

        sys.path.insert(0, str(ROOT))
        from oracle import ed_oracle as O

        done, t = 0, 0.0
        with _cpu_single_thread():
            for i in range(0, M, 7):
                nu = (len(p["eps_bath"][i]) + 1) // 2
                t1 = time.perf_counter()
                _cpu_impurity_solve(O, p["u"][i], p["eps_imp"][i], p["eps_bath"][i], p["v"][i], nu, nu, wn)
                t += time.perf_counter() - t1
                done += 1
                if t > args.secondary_seconds:
                    break
        out["cpu_baseline"] = {"value": round(done / t, 3), "unit": "solves/s", "cores": 1, "kind": "port",
                               "sample": f"{done} models of the same batch (every 7th) by the oracle's numpy/scipy "
                                         "port: ground state + continued fractions"}
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
        with _cpu_single_thread():
            t0 = time.perf_counter()
            g_cpu = []
            for i, u in enumerate(us):
                mk = lambda nu, nd: O.siam_operator(nu, nd, u, 0.0, r.eps_bath[i], r.v[i], u / 2, width=nb + 1)
                e0, psi = O.groundstate(mk(n, n), tol=1e-12)
                g = np.zeros(len(wn), dtype=np.complex128)
                for create, sign in ((True, 1), (False, -1)):
                    phi = (O.apply_cdag if create else O.apply_c)(psi, nb + 1, n, n, 0, O.UP, "jordan_wigner")
                    opx = mk(n + 1 if create else n - 1, n)
                    a, b, n2, _ = O.lanczos_normalized(opx.matvec, phi, opx.size, reorth=True)
                    g += O.cf_greens(a, b, n2, 1j * wn, e0, sign)
                g_cpu.append(g)
            t = time.perf_counter() - t0
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
        out["cpu_baseline"] = {"value": round(1.0 / t, 4), "unit": "iterations/s", "cores": 1, "kind": "port",
                               "sample": f"the oracle's impurity solves of one iteration (8 models, D = 400, "
                                         f"full-Krylov continued fractions; {t:.2f} s), bath fits not included"}
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
    for _ in range(3):
        t0 = time.perf_counter()
        r, ms_b, ge, gv = step(b)
        fw.append(r.elapsed_ms)
        bw.append(ms_b)
    ms = float(np.median(np.array(fw) + np.array(bw)))
    out = {"metric": "grads/s", "value": round(M / (ms * 1e-3), 1), "unit": "grads/s", "models": M,
           "bath_sites": nb, "dim": 4900, "matsubara": 512, "beta": 100.0,
           "ms_forward": round(float(np.median(fw)), 3), "ms_backward": round(float(np.median(bw)), 3),
           "data": "synthetic (theta from seed 101, target from seed 202)"}
    log(f"bath-fit gradients: {out['value']} grads/s")
    # plain per-model GPU path: each model alone (its own launches)
    t0 = time.perf_counter()
    ns = 8
    for i in range(ns):
        bi = AndersonBatch(u=u[i:i + 1], eps_imp=-u[i:i + 1] / 2, eps_bath=[ea[i]], v=[va[i]], mu=0.0,
                           keep_basis=True)
        bi.config(-1, 1, 0)
        r = bi.solve(wn)
        d = r.g - target[i:i + 1]
        bi.backward(2 * d.real, 2 * d.imag)
    t_plain = time.perf_counter() - t0
    out["plain_gpu"] = {"value": round(ns / t_plain, 2), "unit": "grads/s", "sample": ns}
    out["speedup_vs_plain"] = round(out["value"] / out["plain_gpu"]["value"], 1)
    if not args.no_cpu:
        import sys

        sys.path.insert(0, str(ROOT))
        from oracle import ed_oracle as O

        # CPU: central differences of the oracle's forward (the derivative a
        # CPU fit of the reference would use), one model
        h = 1e-5
        t0 = time.perf_counter()
        with _cpu_single_thread():
            def loss(e, v):
                _, g = _cpu_impurity_solve(O, 3.0, -1.5, e, v, 4, 4, wn)
                return float(np.sum(np.abs(g - target[0]) ** 2))

            for k in range(2 * nb):
                d = np.zeros(2 * nb)
                d[k] = h
                loss(ea[0] + d[:nb], va[0] + d[nb:])
                loss(ea[0] - d[:nb], va[0] - d[nb:])
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(1.0 / t, 4), "unit": "grads/s", "cores": 1, "kind": "port",
                               "sample": "1 model: central differences (28 solves) with the oracle's numpy/scipy "
                                         "port of the forward"}
    return out


if __name__ == "__main__":
    main()
