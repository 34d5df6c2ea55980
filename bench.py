"""Benchmark: DFN solves per second on one MI355X.

A step is one complete solve of the seeded 5,000-fracture network (500 m cube, radii on
[10, 150] m, 20 coefficients / 40 collocation points per element): discharge-matrix assembly,
LU factorisation, the fixed-point iterations with coefficient growth, and the final head
residual, with the network already resident in device memory when the timed region starts.
The solve stops on AnDFN's rule (coefficient and discharge change below MAX_ERROR twice) or as
soon as the head residual (max |A y - b(c)| over the head equations) is at most 1e-6 m.

Prints one JSON line.  ``roofline`` describes the collocation sweep kernel (FP64 vector ALU
bound); ``cpu_baseline`` times the CPU FP64 oracle (a restatement of AnDFN's Numba path with
SciPy's SuperLU) on a bounded sample of the same solve on this host.
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
HEAD_TOL = 1e-6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cfg2")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--lu-mode", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-iterations", type=int, default=2, help="oracle iterations in the CPU sample")
    return ap.parse_args()


def traffic_from_profiles(workload):
    """HBM bytes per sweep launch measured with rocprofv3 PMC counters (profiles/*.json), if present."""
    p = ROOT / "profiles" / f"sweep_{workload}_traffic.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("bytes_per_launch")
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


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); multi-GPU runs are not part of this build",
              file=sys.stderr)
        sys.exit(2)
    from andfn_amd import _lib
    from andfn_amd.generator import build_network, workload
    from andfn_amd.hpc_solve import DeviceNetwork

    if _lib.load().andfn_mi_device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    spec = workload(args.workload)
    dfn = build_network(spec)
    dfn.consolidate_dfn()
    fa, ea = dfn.fractures_struc_array_hpc, dfn.elements_struc_array_hpc
    fa0, ea0 = fa.copy(), ea.copy()
    net = DeviceNetwork(fa, ea, dfn.discharge_int, dfn.constants)
    print(f"# network: {len(fa)} fractures, {len(ea)} elements; warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        net.run(args.precision, args.lu_mode, HEAD_TOL)
    infos = []
    walls = []
    for k in range(args.steps):
        t0 = time.perf_counter()
        info = net.run(args.precision, args.lu_mode, HEAD_TOL)
        walls.append(time.perf_counter() - t0)
        infos.append(info)
        print(f"# step {k}: {walls[-1]:.3f}s nit={info['nit']} head_residual={info['head_residual']:.2e} "
              f"factor={info['t_factor']:.3f}s iterate={info['t_iterate']:.3f}s", file=sys.stderr, flush=True)
    net.close()
    total = sum(walls)
    last = infos[-1]
    sweep_ms = last["k_sweep_ms"] / max(last["n_sweep"], 1)
    flops_launch = last["flops_sweep"] / max(last["n_sweep"], 1)
    achieved = flops_launch / (sweep_ms * 1e-3) if sweep_ms > 0 else 0.0
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
            "stop": f"AnDFN rule (MAX_ERROR {float(dfn.constants['MAX_ERROR']):g}) or head residual <= "
                    f"{HEAD_TOL:g} m, MAX_ITERATIONS {int(dfn.constants['MAX_ITERATIONS'])}",
            "filters": f"rim clearance {spec.rim_clearance:g} R, near-coincident pairs < {spec.min_separation:g} L "
                       f"at < {spec.min_angle_deg:g} deg",
        },
        "nit": int(last["nit"]),
        "head_residual": last["head_residual"],
        "bc_residual": last["bc_residual"],
        "converged": bool(last["converged"]),
        "t_factor_s": last["t_factor"],
        "t_iterate_s": last["t_iterate"],
        "roofline": {
            "bound": "mfma",
            "kernel": "sweep_kernel<MODE_COLLOC> (FP64 vector ALU; MI355X FP64 vector peak = FP64 matrix peak)",
            "achieved": achieved / 1e12,
            "peak": PEAK_FP64 / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64,
            "traffic": traffic_from_profiles(args.workload),
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
