"""Benchmark: DFN solves per second on one MI355X.

A step is one complete solve of the seeded 5,000-fracture network (500 m cube, radii on
[10, 150] m, 20 coefficients / 40 collocation points per element): discharge-matrix assembly,
LU factorisation, the fixed-point iterations with coefficient growth, and the final head
residual, with the network already resident in device memory when the timed region starts.
The solve stops as soon as the head residual (max |A y - b(c)| over the head equations) is at
most 1e-6 m.  ``value`` is the FP64 form (``--precision``; the batched cfg5 workload defaults
to its specified mixed form); ``forms`` adds the other form (the
mixed-precision form: FP32 factors refined in FP64, FP32 far field and fits) with its error
against the FP64 result; ``adjoint`` times the implicit adjoint (gradient of the x = 0 boundary
discharge plus 100 observation heads with respect to every log10 transmissivity and both
boundary heads) against one forward solve.

Prints one JSON line.  ``roofline`` describes the dominant kernel chain, the blocked triangular
solves (HBM bound; one per iteration); ``roofline_sweep`` the collocation sweep kernel (FP64
vector ALU bound); ``cpu_baseline`` times the CPU FP64 oracle (a restatement of AnDFN's Numba
path with SciPy's SuperLU) on a bounded sample of the same solve on this host.
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP64 = 78.6e12  # MI355X dense FP64 (vector and matrix), flop/s
PEAK_HBM = 8.0e12  # MI355X HBM3E, bytes/s
HEAD_TOL = 1e-6
ADJOINT_TOL = 1e-10


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cfg2")
    ap.add_argument("--precision", default=None, choices=["fp64", "mixed"],
                    help="form timed for value (default: fp64; mixed for cfg5, which is specified mixed)")
    ap.add_argument("--lu-mode", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="time only the --precision form")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--realisations", type=int, default=64, help="cfg5: number of realisations")
    ap.add_argument("--max-iterations", type=int, default=0, help="override MAX_ITERATIONS (short profiling runs)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--cpu-iterations", type=int, default=2, help="oracle iterations in the CPU sample")
    args = ap.parse_args()
    if args.precision is None:
        args.precision = "mixed" if args.workload == "cfg5" else "fp64"
    return args


def traffic_from_profiles(name, workload):
    """HBM bytes per launch (``sweep``) or per solve (``solve``) measured with rocprofv3 PMC
    counters (profiles/<name>_<workload>_traffic.json), if present."""
    p = ROOT / "profiles" / f"{name}_{workload}_traffic.json"
    if p.exists():
        try:
            d = json.loads(p.read_text())
            return d.get("bytes_per_launch", d.get("bytes_per_solve"))
        except Exception:
            return None
    return None


def cpu_baseline(fa, ea, K, consts, nit_gpu, n_iter):
    """Oracle on this host: SuperLU factorisation + n_iter iterations, scaled to nit_gpu iterations."""
    from oracle import oracle

    fo, eo = fa.copy(), ea.copy()
    O = oracle.OracleHandle(fo, eo, K)
    lib = O.lib
    t0 = time.perf_counter()
    A = O.discharge_matrix()
    import scipy.sparse.linalg as spla

    t1 = time.perf_counter()
    lu = spla.splu(A)
    t2 = time.perf_counter()
    lib.oracle_init(O.h)
    y = np.zeros(O.n)
    lib.oracle_old_discharges(O.h, oracle._ptr(y))
    t3 = time.perf_counter()
    for nit in range(1, n_iter + 1):
        b = O.head_rhs()
        y = lu.solve(b)
        O.set_discharges(y)
        lib.oracle_element_solver2(O.h, float(consts["MAX_ERROR"]), nit, float(consts["COEF_RATIO"]),
                                   int(consts["MAX_NCOEF"]), int(consts["COEF_INCREASE"]), float(consts["DAMPING"]),
                                   None)
    t4 = time.perf_counter()
    O.close()
    t_iter = (t4 - t3) / n_iter
    t_solve = (t1 - t0) + (t2 - t1) + nit_gpu * t_iter
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": 1.0 / t_solve,
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"same network: assembly {t1 - t0:.2f}s + SuperLU splu {t2 - t1:.2f}s + {n_iter} oracle "
                   f"iterations ({t_iter:.2f}s each, ncoef growing from the initial counts) scaled to the GPU "
                   f"run's {nit_gpu} iterations; SuperLU single-threaded, element kernels on {threads} OpenMP threads"),
    }


def _obs_points(fa, n, seed=3):
    rng = np.random.default_rng(seed)
    of = rng.integers(0, len(fa), n)
    oz = fa["radius"][of] * rng.uniform(0.1, 0.7, n) * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    return of, oz


def _time_form(net, precision, lu_mode, warmup, steps, label):
    for _ in range(warmup):
        net.run(precision, lu_mode, HEAD_TOL)
    infos, walls = [], []
    for k in range(steps):
        t0 = time.perf_counter()
        info = net.run(precision, lu_mode, HEAD_TOL)
        walls.append(time.perf_counter() - t0)
        infos.append(info)
        print(f"# {label} step {k}: {walls[-1]:.3f}s nit={info['nit']} head_residual={info['head_residual']:.2e} "
              f"factor={info['t_factor']:.3f}s iterate={info['t_iterate']:.3f}s", file=sys.stderr, flush=True)
    return infos, walls


def bench_batch(args):
    """Batched realisations: 64 seeded DFNs (fracture counts uniform on [500, 4000], the 200-disk law
    at constant density, 10 to 20 coefficients) as one network; the mixed form to head residual
    1e-6 and the adjoint of the summed x = 0 boundary discharge w.r.t. every log10 T."""
    from andfn_amd.autograd import boundary_groups
    from andfn_amd.batch import build_batch, realisation_specs
    from andfn_amd.hpc_solve import DeviceNetwork

    count = args.realisations
    t0 = time.perf_counter()
    specs = realisation_specs(count)
    dfns, fa, ea, off, K, consts = build_batch(specs, dict(MAX_ERROR=1e-9, MAX_NCOEF=10, MAX_ITERATIONS=500))
    print(f"# batch: {count} realisations, {len(fa)} fractures, {len(ea)} elements, built in "
          f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    net = DeviceNetwork(fa, ea, K, consts)
    dq = np.zeros(len(ea))
    for r, d in enumerate(dfns):
        dq[off[r][1] + boundary_groups(d)["x0"]] = 1.0
    res = {}
    for prec in (args.precision, "fp64" if args.precision == "mixed" else "mixed"):
        if prec != args.precision and args.no_compare:
            continue
        infos, walls = _time_form(net, prec, 0, args.warmup, args.steps, prec)
        aw = []
        for k in range(1 + args.steps):
            t1 = time.perf_counter()
            gt, gh, ai = net.adjoint(dq=dq, tol=ADJOINT_TOL)
            aw.append(time.perf_counter() - t1)
            print(f"# {prec} adjoint {k}: {aw[-1]:.3f}s iterations={ai['nit']}", file=sys.stderr, flush=True)
        res[prec] = {"ms_per_batch": 1e3 * sum(walls) / args.steps, "nit": int(infos[-1]["nit"]),
                     "head_residual": infos[-1]["head_residual"], "solves_fp32": int(infos[-1]["solves_fp32"]),
                     "nit_fp32": int(infos[-1]["nit_fp32"]), "t_factor_s": infos[-1]["t_factor"],
                     "adjoint_ms": 1e3 * sum(aw[1:]) / args.steps, "adjoint_iterations": int(ai["nit"])}
        res[prec]["adjoint_forward_ratio"] = res[prec]["adjoint_ms"] / res[prec]["ms_per_batch"]
    net.close()
    main_r = res[args.precision]
    out = {
        "metric": "DFN realisation solves/sec at max head residual 1e-6 (batched, one network)",
        "value": count / (main_r["ms_per_batch"] / 1e3),
        "unit": "realisations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_r["ms_per_batch"],
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else "f64",
        "data": "synthetic (seeded counter-based generator)",
        "config": {"workload": f"cfg5: {count} realisations, fracture counts uniform on [500, 4000], 200-disk law "
                               f"at constant density, 10-20 coefficients",
                   "fractures": int(len(fa)), "elements": int(len(ea))},
        "forms": res,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def main():
    args = parse()
    if args.workload == "cfg5":
        if args.gpus != 1:
            print("bench.py runs on one GPU only", file=sys.stderr)
            sys.exit(2)
        bench_batch(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); multi-GPU runs are not part of this build",
              file=sys.stderr)
        sys.exit(2)
    from andfn_amd import _lib
    from andfn_amd.autograd import boundary_groups
    from andfn_amd.generator import build_network, workload
    from andfn_amd.hpc_solve import DeviceNetwork

    if _lib.load().andfn_mi_device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.calibrate_copy_mib > 0:
        import torch

        nb = args.calibrate_copy_mib << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"# calibration copy: {nb} bytes read, {nb} bytes written", file=sys.stderr, flush=True)
        del src, dst
    spec = workload(args.workload, constants={"MAX_ITERATIONS": args.max_iterations}) if args.max_iterations \
        else workload(args.workload)
    dfn = build_network(spec)
    dfn.consolidate_dfn()
    fa, ea = dfn.fractures_struc_array_hpc, dfn.elements_struc_array_hpc
    fa0, ea0 = fa.copy(), ea.copy()
    net = DeviceNetwork(fa, ea, dfn.discharge_int, dfn.constants)
    print(f"# network: {len(fa)} fractures, {len(ea)} elements; warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)
    other = "mixed" if args.precision == "fp64" else "fp64"
    forms = {}
    states = {}
    obs_frac, obs_z = _obs_points(fa, 100)
    for prec in ([args.precision] + ([] if args.no_compare else [other])):
        infos, walls = _time_form(net, prec, args.lu_mode, args.warmup, args.steps, prec)
        q, c = net.state_discharges()
        net.download()
        states[prec] = (ea["coef"].copy(), q, c, net.state_heads(obs_frac, obs_z))
        last = infos[-1]
        forms[prec] = {"solves_per_s": args.steps / sum(walls), "ms_per_solve": 1e3 * sum(walls) / args.steps,
                       "nit": int(last["nit"]), "nit_fp32": int(last["nit_fp32"]),
                       "head_residual": last["head_residual"], "t_factor_s": last["t_factor"],
                       "t_iterate_s": last["t_iterate"], "info": last, "walls": walls}
    main_form = forms[args.precision]
    last = main_form.pop("info")
    walls = main_form.pop("walls")
    if other in forms:
        forms[other].pop("info")
        forms[other].pop("walls")
        a, b = states["fp64"], states["mixed"]

        def rel(x, y):
            return float(np.max(np.abs(x - y)) / max(np.max(np.abs(y)), 1e-300))

        forms["mixed"]["error_vs_fp64"] = {"coef_rel": rel(b[0], a[0]), "q_rel": rel(b[1], a[1]),
                                           "constant_rel": rel(b[2], a[2]),
                                           "heads_abs_m": float(np.max(np.abs(b[3] - a[3]))),
                                           "note": "both forms stopped at head residual <= 1e-6 m; free-running "
                                                   "coefficient growth; heads at 100 observation points"}
    adjoint = None
    if not args.no_adjoint:
        # gradient of Q(x=0) + sum of 100 observation heads w.r.t. every log10 T and the two boundary heads
        net.run("fp64", args.lu_mode, HEAD_TOL)
        t_fwd = forms["fp64"]["ms_per_solve"] / 1e3 if "fp64" in forms else None
        groups = boundary_groups(dfn)
        dq = np.zeros(len(ea))
        dq[groups["x0"]] = 1.0
        aw = []
        for k in range(1 + args.steps):
            t0 = time.perf_counter()
            gt, gh, ai = net.adjoint(dq=dq, obs_frac=obs_frac, obs_z=obs_z, dhead=np.ones(len(obs_frac)),
                                     tol=ADJOINT_TOL)
            aw.append(time.perf_counter() - t0)
            print(f"# adjoint {k}: {aw[-1]:.3f}s iterations={ai['nit']} residual={ai['delta']:.2e}",
                  file=sys.stderr, flush=True)
        t_adj = sum(aw[1:]) / args.steps
        adjoint = {"objective": "Q through x=0 + sum of heads at 100 observation points",
                   "gradient": f"d/d log10 T for all {len(fa)} fractures and d/d head of the 2 boundary planes",
                   "method": "implicit fixed-point adjoint, GMRES on the transposed iteration (FP64)",
                   "tol": ADJOINT_TOL, "iterations": int(ai["nit"]), "residual": ai["delta"],
                   "ms": 1e3 * t_adj, "first_call_ms": 1e3 * aw[0],
                   "adjoint_forward_ratio": (t_adj / t_fwd) if t_fwd else None}
    net.close()
    total = sum(walls)
    sweep_ms = last["k_sweep_ms"] / max(last["n_sweep"], 1)
    flops_launch = last["flops_sweep"] / max(last["n_sweep"], 1)
    achieved = flops_launch / (sweep_ms * 1e-3) if sweep_ms > 0 else 0.0
    # the dominant kernel chain: the blocked triangular solves (one HIP graph per solve), timed
    # by HIP events around every graph launch.  One connected part, so one dense block of n
    # unknowns: the chain reads the strictly triangular panels of L and U and the triangles of
    # the diagonal-block inverses, n^2 + n values.
    n_solve = max(int(last["n_rhs"]), 1)
    solve_ms = last["k_lu_ms"] / n_solve
    n_unk = int(last["n_unknowns"])
    solve_bytes = (8.0 if args.precision == "fp64" else 4.0) * (n_unk * n_unk + n_unk)
    solve_bw = solve_bytes / (solve_ms * 1e-3) if solve_ms > 0 else 0.0
    out = {
        "metric": "DFN solves/sec at max head residual 1e-6 (FP64 vs mixed GPU); adjoint/forward ratio",
        "value": args.steps / total,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded counter-based generator)",
        "config": {
            "workload": f"{args.workload}: {spec.n_fractures} disks, {spec.cube:g} m cube, radii [{spec.rmin:g}, "
                        f"{spec.rmax:g}] m, pdf r^-{spec.alpha:g}, ncoef {spec.ncoef}, nint {spec.nint}, "
                        f"K {spec.discharge_int}",
            "fractures_kept": int(len(fa)),
            "elements": int(len(ea)),
            "unknowns": int(last["n_unknowns"]),
            "precision": args.precision,
            "stop": f"head residual <= {HEAD_TOL:g} m (max |A y - b(c)| over the head equations), "
                    f"MAX_ITERATIONS {int(dfn.constants['MAX_ITERATIONS'])}",
            "filters": f"rim clearance {spec.rim_clearance:g} R, near-coincident pairs < {spec.min_separation:g} L "
                       f"at < {spec.min_angle_deg:g} deg",
        },
        "nit": int(last["nit"]),
        "head_residual": last["head_residual"],
        "bc_residual": last["bc_residual"],
        "converged": bool(last["converged"]),
        "t_factor_s": last["t_factor"],
        "t_iterate_s": last["t_iterate"],
        "forms": forms,
        "adjoint": adjoint,
        "roofline": {
            "bound": "hbm",
            "kernel": "blocked triangular solves (gemv_cm_kernel + gemv_diag_kernel chain, one HIP graph per "
                      "solve)",
            "share_of_step": (last["k_lu_ms"] / (1e3 * total / args.steps)) if total > 0 else None,
            "achieved": solve_bw / 1e9,
            "peak": PEAK_HBM / 1e9,
            "unit": "GB/s",
            "frac": solve_bw / PEAK_HBM,
            "traffic": traffic_from_profiles("solve", args.workload) if args.precision == "fp64" else None,
            "avg_launch_ms": solve_ms,
            "bytes_per_launch": solve_bytes,
            "launches": n_solve,
            "unit_of_launch": "one triangular-solve chain (graph launch)",
        },
        "roofline_sweep": {
            "bound": "mfma",
            "kernel": "sweep_kernel<MODE_COLLOC> (FP64 vector ALU; MI355X FP64 vector peak = FP64 matrix peak)",
            "achieved": achieved / 1e12,
            "peak": PEAK_FP64 / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64,
            "traffic": traffic_from_profiles("sweep", args.workload),
            "avg_launch_ms": sweep_ms,
            "flops_per_launch": flops_launch,
            "launches": int(last["n_sweep"]),
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(fa0, ea0, dfn.discharge_int, dfn.constants, int(last["nit"]),
                                           args.cpu_iterations)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
