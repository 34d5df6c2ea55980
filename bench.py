"""Benchmark: DFN solves per second on one MI355X.

The default workload is the 20,000-fracture network (1,000 m cube, radii on [10, 300] m, 20
coefficients / 40 collocation points per element, 50 wells).  A step is one complete solve:
discharge-matrix assembly, sparse LU factorisation, the fixed-point iterations with coefficient
growth, stopping as soon as the head residual (max |A y - b(c)| over the head equations) is at
most 1e-6 m.  ``value`` is the specified form of the workload (mixed precision for the
20,000-fracture network and the batch: FP32 far field, fit and factors, FP64 residual, near
field and correction; FP64 otherwise) on a handle whose network is resident in device memory
(the numeric factorisation is redone every step; the symbolic analysis of the unchanged pattern
is kept, as SuperLU's ``SamePattern`` reuse keeps it).  ``forms`` adds the other form with its
error against the FP64 result; ``drop_in`` times the drop-in ``hpc_solve.solve`` call, which
pays the record parse, upload, symbolic analysis and every pivot-delay round, and downloads the
records; ``adjoint`` times the implicit adjoint against one forward solve; ``flow`` times the
post-processing of the solved network on the device (streamline tracking of 4,096 lines and the
complex potential on every fracture's flow-net pattern) beside the oracle on bounded samples.

Prints one JSON line.  ``roofline`` is the dominant kernel, the collocation sweep (FP64 vector
ALU bound); ``roofline_solve`` the sparse triangular solves (HBM bound, bytes by the
12 nnz(L+U) + 16 n convention per pass through the factors); both from the FP64 form's run,
whose kernels are the same throughout.  ``cpu_baseline`` times the CPU
FP64 oracle (a restatement of AnDFN's Numba path with SciPy's SuperLU) on this host in a child
process started before any GPU call and run after the GPU timing, on a bounded sample (element
iterations, assembly, the SuperLU factorisation up to a wall-clock cap); a capped factorisation
makes the CPU time a lower bound and the speed-up a lower bound.
"""

import argparse
import json
import os
import subprocess
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
MIXED_WORKLOADS = ("cfg3", "cfg5")  # workloads specified in the mixed-precision form
BATCH_CONSTANTS = dict(MAX_ERROR=1e-9, MAX_NCOEF=10, MAX_ITERATIONS=500)  # the batched realisations


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cfg3")
    ap.add_argument("--precision", default=None, choices=["fp64", "mixed"],
                    help="form timed for value (default: mixed for cfg3 and cfg5, which are specified mixed)")
    ap.add_argument("--lu-mode", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="time only the --precision form")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--dropin-steps", type=int, default=3, help="drop-in hpc_solve.solve calls timed")
    ap.add_argument("--realisations", type=int, default=64, help="cfg5: number of realisations")
    ap.add_argument("--max-iterations", type=int, default=0, help="override MAX_ITERATIONS (short profiling runs)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--cpu-iterations", type=int, default=2, help="oracle element iterations in the CPU sample")
    ap.add_argument("--cpu-cap", type=float, default=90.0, help="wall-clock cap (s) of the CPU sample")
    ap.add_argument("--flow-lines", type=int, default=4096,
                    help="streamlines traced on the solved network in the flow block (0: no flow block)")
    ap.add_argument("--cpu-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.precision is None:
        args.precision = "mixed" if args.workload in MIXED_WORKLOADS else "fp64"
    return args


def traffic_from_profiles(name, workload):
    """HBM bytes per launch (a sweep launch; a pass through the factors) measured with rocprofv3
    PMC counters (profiles/<name>_<workload>_traffic.json), if present."""
    p = ROOT / "profiles" / f"{name}_{workload}_traffic.json"
    if p.exists():
        try:
            d = json.loads(p.read_text())
            return d.get("bytes_per_launch", d.get("bytes_per_solve"))
        except Exception:
            return None
    return None


# ---- CPU baseline: a child process that never touches the GPU -----------------------------------

def cpu_child(args):
    """Build the same network, report ready, wait for ``go``, then time the oracle's pieces one
    by one (each printed as soon as it is known, so a capped run still reports the finished ones):
    element iterations (the Numba path's per-iteration work without the triangular solve),
    assembly of the discharge matrix, SuperLU factorisation, one SuperLU solve."""
    from andfn_amd.generator import build_network, workload
    from oracle import oracle

    if args.workload == "cfg5":
        from andfn_amd.batch import build_batch, realisation_specs

        _, fa, ea, _, K, consts = build_batch(realisation_specs(args.realisations), BATCH_CONSTANTS)
    else:
        d = build_network(workload(args.workload))
        d.consolidate_dfn()
        fa, ea, K, consts = d.fractures_struc_array_hpc, d.elements_struc_array_hpc, d.discharge_int, d.constants
    fa, ea = fa.copy(), ea.copy()
    O = oracle.OracleHandle(fa, ea, K)
    print("ready", flush=True)
    line = sys.stdin.readline()
    if not line.startswith("go"):
        return
    lib = O.lib
    lib.oracle_init(O.h)
    y = np.zeros(O.n)
    lib.oracle_old_discharges(O.h, oracle._ptr(y))
    O.set_discharges(y)
    c = {k: consts[k].item() for k in consts.dtype.names}
    for it in range(1, args.cpu_iterations + 1):
        t0 = time.perf_counter()
        b = O.head_rhs()
        lib.oracle_element_solver2(O.h, float(c["MAX_ERROR"]), it, float(c["COEF_RATIO"]), int(c["MAX_NCOEF"]),
                                   int(c["COEF_INCREASE"]), float(c["DAMPING"]), None)
        print(f"iteration {time.perf_counter() - t0:.6f}", flush=True)
    t0 = time.perf_counter()
    A = O.discharge_matrix()
    print(f"assembly {time.perf_counter() - t0:.6f}", flush=True)
    import scipy.sparse.linalg as spla

    print("splu_start", flush=True)
    t0 = time.perf_counter()
    lu = spla.splu(A)
    print(f"splu {time.perf_counter() - t0:.6f}", flush=True)
    t0 = time.perf_counter()
    lu.solve(b)
    print(f"lusolve {time.perf_counter() - t0:.6f}", flush=True)
    O.close()


def start_cpu_child(args):
    cmd = [sys.executable, "-u", str(ROOT / "bench.py"), "--cpu-child", "--workload", args.workload,
           "--cpu-iterations", str(args.cpu_iterations), "--realisations", str(args.realisations)]
    return subprocess.Popen(cmd, stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True, bufsize=1, cwd=str(ROOT))


def finish_cpu_child(proc, args, nit_gpu):
    """Send go, collect the timings within the cap, stop the child, and assemble cpu_baseline."""
    import selectors

    t_start = time.perf_counter()
    fd = proc.stdout.fileno()
    sel = selectors.DefaultSelector()
    sel.register(fd, selectors.EVENT_READ)
    buf = b""

    def lines_until(deadline):
        """complete lines from the child's unbuffered pipe, until EOF or the deadline"""
        nonlocal buf
        while True:
            while b"\n" in buf:
                line, buf = buf.split(b"\n", 1)
                yield line.decode()
            left = deadline - time.perf_counter()
            if left <= 0 or not sel.select(timeout=min(left, 5.0)):
                if left <= 0:
                    return
                continue
            chunk = os.read(fd, 4096)
            if not chunk:
                return
            buf += chunk

    ready = next(lines_until(time.perf_counter() + 600.0), "")
    if ready.strip() != "ready":
        proc.kill()
        proc.wait()
        return None
    proc.stdin.write("go\n")
    proc.stdin.flush()
    t_go = time.perf_counter()
    its, got, splu_started = [], {}, None
    for line in lines_until(t_go + args.cpu_cap):
        key, *val = line.split()
        if key == "iteration":
            its.append(float(val[0]))
        elif key == "splu_start":
            splu_started = time.perf_counter()
        elif val:
            got[key] = float(val[0])
        if key == "lusolve":
            break
    capped = "lusolve" not in got
    if proc.poll() is None:
        proc.kill()  # the child started above (its own PID), capped or done
    proc.wait()
    if not its or "assembly" not in got:
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t_iter = float(np.mean(its))
    if "splu" in got:
        t_splu, splu_note = got["splu"], f"SuperLU splu {got['splu']:.1f} s"
    else:
        t_splu = time.perf_counter() - splu_started if splu_started else 0.0
        splu_note = f"SuperLU splu stopped at the {args.cpu_cap:g} s cap after {t_splu:.1f} s (unfinished)"
    t_lus = got.get("lusolve", 0.0)
    t_solve = got["assembly"] + t_splu + nit_gpu * (t_iter + t_lus)
    lower = capped or "splu" not in got
    return {
        "value": 1.0 / t_solve,
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "bound": "upper (CPU time is a lower bound)" if lower else "measured pieces",
        "seconds_per_solve": t_solve,
        "sample": (f"same network, FP64 oracle on the host: {len(its)} element iterations ({t_iter:.2f} s each on "
                   f"{threads} OpenMP threads: head RHS, element sweep and fit), discharge-matrix assembly "
                   f"{got['assembly']:.2f} s, {splu_note}"
                   + (f", one SuperLU solve {t_lus:.3f} s" if t_lus else ", SuperLU solves not reached (omitted)")
                   + f"; the solve time counts assembly + factorisation + {nit_gpu} iterations (the GPU run's count)"),
        "wall_s": time.perf_counter() - t_start,
    }


# ---- GPU forms -----------------------------------------------------------------------------------

def _obs_points(fa, n, seed=3):
    rng = np.random.default_rng(seed)
    of = rng.integers(0, len(fa), n)
    oz = fa["radius"][of] * rng.uniform(0.1, 0.7, n) * np.exp(1j * rng.uniform(0, 2 * np.pi, n))
    return of, oz


def _launch_info(info):
    """How the iterations ran: the sweep's tile shape (autotuned) and the run-ahead graph launches."""
    return {"sweep_points_per_lane": int(info["sweep_points_per_lane"]),
            "sweep_persistent": int(info["sweep_persistent"]),
            "runahead_iterations": int(info["runahead_iterations"]),
            "graph_iterations": int(info["graph_iterations"])}


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


def _time_dropin(fa0, ea0, dfn, precision, steps):
    """The drop-in entry point on the unsolved records, records in and out, a new handle per call."""
    from andfn_amd.hpc_solve import last_solve_info, solve

    walls, infos = [], []
    for k in range(steps):
        fa, ea = fa0.copy(), ea0.copy()
        t0 = time.perf_counter()
        solve(fa, ea, dfn.discharge_int, dfn.constants, precision=precision, head_tol=HEAD_TOL)
        walls.append(time.perf_counter() - t0)
        infos.append(last_solve_info())
        print(f"# drop-in {precision} call {k}: {walls[-1]:.3f}s nit={infos[-1]['nit']} "
              f"factor={infos[-1]['t_factor']:.3f}s rounds={infos[-1]['factor_rounds']}", file=sys.stderr, flush=True)
    return walls, infos


def _roofline_sweep(last, fp64=True, workload=None):
    n = max(int(last["n_sweep"]), 1)
    ms = last["k_sweep_ms"] / n
    fl = last["flops_sweep"] / n
    achieved = fl / (ms * 1e-3) if ms > 0 else 0.0
    return {
        "bound": "mfma",
        "kernel": ("sweep_kernel<MODE_COLLOC, FP64> collocation sweep of the FP64 form" if fp64 else
                   "sweep_kernel<MODE_COLLOC> collocation sweep of the mixed form (FP32 far field in its FP32 "
                   "phase)") + " (FP64 vector ALU; MI355X FP64 vector peak = FP64 matrix peak)",
        "form": "fp64" if fp64 else "mixed",
        "achieved": achieved / 1e12,
        "peak": PEAK_FP64 / 1e12,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP64,
        "traffic": traffic_from_profiles("sweep", workload) if (fp64 and workload) else None,
        "avg_launch_ms": ms,
        "flops_per_launch": fl,
        "launches": n,
    }


def _roofline_solve(last, f32, workload=None):
    passes = max(int(last["solve_passes"]), 1)
    ms = last["k_lu_ms"] / passes  # per pass through the factors (correction solves: one each)
    nnz, n = float(last["nnz_factors"]), float(last["n_unknowns"])
    conv = 12.0 * nnz + 16.0 * n  # bytes per pass by the 12 nnz(L+U) + 16 n convention
    streamed = (4.0 if f32 else 8.0) * float(last["solve_values"]) + 16.0 * n
    bw = conv / (ms * 1e-3) if ms > 0 else 0.0
    return {
        "bound": "hbm",
        "kernel": "multifrontal triangular solves (mfs_panel_kernel, four launches per level of the elimination tree, "
                  "one HIP graph per solve)",
        "form": "mixed (FP32 tiles)" if f32 else "fp64",
        "achieved": bw / 1e9,
        "peak": PEAK_HBM / 1e9,
        "unit": "GB/s",
        "frac": bw / PEAK_HBM,
        "traffic": traffic_from_profiles("solve", workload) if (not f32 and workload) else None,
        "avg_pass_ms": ms,
        "bytes_per_pass": conv,
        "bytes_streamed_per_pass": streamed,
        "streamed_GBps": streamed / (ms * 1e-3) / 1e9 if ms > 0 else 0.0,
        "passes": passes,
        "note": ("bytes_per_pass follows the 12 nnz(L+U) + 16 n convention (values plus 32-bit indices); the "
                 "panel tiles hold dense blocks without indices, so a pass streams bytes_streamed_per_pass"
                 + (" (tiles in FP32)" if f32 else "")),
    }


FLOW_DS, FLOW_MAX_LENGTH, FLOW_APERTURE, FLOW_NET_POINTS = 0.01, 500, 1e-4, (900, 100)


def _flow_starts(fa, n, seed=5):
    rng = np.random.default_rng(seed)
    fr = rng.choice(len(fa), n)
    z0 = 0.5 * fa["radius"][fr] * np.sqrt(rng.uniform(size=n)) * np.exp(2j * np.pi * rng.uniform(size=n))
    return fr, z0


def _flow_bench(net, fa, n_lines):
    """Post-processing of the solved network: streamlines (DFN.streamline_tracking for n_lines start
    points) and the complex potential on every fracture's flow-net pattern (get_flow_nets)."""
    from andfn_amd.hpc_fracture import sunflower_spiral

    fr, z0 = _flow_starts(fa, n_lines)
    ap = np.full(len(fa), FLOW_APERTURE)
    el = np.full(n_lines, 0.5)
    net.streamlines(fr[:64], z0[:64], el[:64], ap, FLOW_DS, FLOW_MAX_LENGTH)  # warm-up
    t0 = time.perf_counter()
    pts, vel, npts, sf, sp, nseg, ex = net.streamlines(fr, z0, el, ap, FLOW_DS, FLOW_MAX_LENGTH)
    t_sl = time.perf_counter() - t0
    zu = sunflower_spiral(*FLOW_NET_POINTS)
    ffr = np.repeat(np.arange(len(fa)), len(zu))
    zz = (zu[None, :] * fa["radius"][:, None]).ravel()
    net.state_omega(ffr[:4096], zz[:4096])  # warm-up
    t0 = time.perf_counter()
    om = net.state_omega(ffr, zz)
    t_om = time.perf_counter() - t0
    return {
        "streamlines": {"lines": int(n_lines), "points": int(npts.sum()), "segments": int(nseg.sum()),
                        "s": t_sl, "points_per_s": float(npts.sum() / t_sl),
                        "setup": f"start at 0.5 r in random fractures, elevation 0.5, ds {FLOW_DS} r, max_length "
                                 f"{FLOW_MAX_LENGTH}, aperture {FLOW_APERTURE}; andfn_mi_state_streamlines + fetch"},
        "flow_nets": {"points": int(len(zz)), "s": t_om, "points_per_s": float(len(zz) / t_om),
                      "nan_points": int(np.isnan(om).sum()),
                      "setup": f"Omega at sunflower_spiral{FLOW_NET_POINTS} x radius on every fracture "
                               "(andfn_mi_state_omega, points in and values out of host memory)"},
        "cpu_baseline": None,
    }


def _flow_cpu(fa, ea, K, n_lines, cap_s=10.0):
    """The oracle on the same work, bounded samples: lines traced one after another for cap_s
    seconds (its C++ w behind a Python restatement of the tracking), Omega at the first 200,000
    flow-net points (C++, OpenMP)."""
    from andfn_amd.hpc_fracture import sunflower_spiral
    from oracle import oracle as orc
    from oracle import streamlines as osl

    O = orc.OracleHandle(fa, ea, K)
    try:
        fr, z0 = _flow_starts(fa, n_lines)
        T = osl.Tracker(O, fa, ea, np.full(len(fa), FLOW_APERTURE))
        t0 = time.perf_counter()
        pts = lines = 0
        while lines < n_lines and time.perf_counter() - t0 < cap_s:
            s, _, _, _ = T.track(complex(z0[lines]), int(fr[lines]), 0.5, FLOW_DS, FLOW_MAX_LENGTH, False)
            pts += sum(len(x) for x in s)
            lines += 1
        t_sl = time.perf_counter() - t0
        zu = sunflower_spiral(*FLOW_NET_POINTS)
        ffr = np.repeat(np.arange(len(fa)), len(zu))[:200000]
        zz = (zu[None, :] * fa["radius"][:, None]).ravel()[:200000]
        t0 = time.perf_counter()
        O.omega_at(ffr, zz)
        t_om = time.perf_counter() - t0
    finally:
        O.close()
    return {"streamlines_points_per_s": pts / t_sl, "streamline_sample": f"{lines} lines, {pts} points",
            "flow_net_points_per_s": len(zz) / t_om, "flow_net_sample": f"{len(zz)} points",
            "cores_streamlines": 1, "cores_flow_nets": os.cpu_count(), "kind": "port"}


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
    dfns, fa, ea, off, K, consts = build_batch(specs, BATCH_CONSTANTS)
    print(f"# batch: {count} realisations, {len(fa)} fractures, {len(ea)} elements, built in "
          f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    net = DeviceNetwork(fa, ea, K, consts)
    dq = np.zeros(len(ea))
    for r, d in enumerate(dfns):
        dq[off[r][1] + boundary_groups(d)["x0"]] = 1.0
    res = {}
    lasts = {}
    for prec in (args.precision, "fp64" if args.precision == "mixed" else "mixed"):
        if prec != args.precision and args.no_compare:
            continue
        infos, walls = _time_form(net, prec, 0, args.warmup, args.steps, prec)
        lasts[prec] = infos[-1]
        aw = []
        for k in range(1 + min(args.steps, 3)):
            t1 = time.perf_counter()
            gt, gh, ai = net.adjoint(dq=dq, tol=ADJOINT_TOL)
            aw.append(time.perf_counter() - t1)
            print(f"# {prec} adjoint {k}: {aw[-1]:.3f}s iterations={ai['nit']}", file=sys.stderr, flush=True)
        res[prec] = {"ms_per_batch": 1e3 * sum(walls) / args.steps, "nit": int(infos[-1]["nit"]),
                     "head_residual": infos[-1]["head_residual"], "solves_fp32": int(infos[-1]["solves_fp32"]),
                     "nit_fp32": int(infos[-1]["nit_fp32"]), "t_factor_s": infos[-1]["t_factor"],
                     "adjoint_ms": 1e3 * sum(aw[1:]) / (len(aw) - 1), "adjoint_iterations": int(ai["nit"]),
                     "k_sweep_ms": infos[-1]["k_sweep_ms"], "k_rhs_ms": infos[-1]["k_rhs_ms"],
                     "k_fit_ms": infos[-1]["k_fit_ms"], "k_lu_ms": infos[-1]["k_lu_ms"], **_launch_info(infos[-1])}
        res[prec]["adjoint_forward_ratio"] = res[prec]["adjoint_ms"] / res[prec]["ms_per_batch"]
    net.close()
    main_r = res[args.precision]
    last = lasts[args.precision]
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
        "nit": int(last["nit"]),
        "roofline": _roofline_sweep(lasts["fp64"] if "fp64" in lasts else last, "fp64" in lasts, "cfg5"),
        "roofline_solve": _roofline_solve(lasts["fp64"] if "fp64" in lasts else last,
                                          False if "fp64" in lasts else (args.precision == "mixed"
                                                                         and last["solves_fp32"] > 0), "cfg5"),
        "cpu_baseline": None,
    }
    return out


def main():
    args = parse()
    if args.cpu_child:
        cpu_child(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); multi-GPU runs are not part of this build",
              file=sys.stderr)
        sys.exit(2)
    # the CPU baseline's child starts before anything touches the GPU and waits for the GPU timing
    child = None if args.no_cpu_baseline else start_cpu_child(args)
    try:
        out = bench_batch(args) if args.workload == "cfg5" else gpu_part(args)
        if child is not None:
            out["cpu_baseline"] = finish_cpu_child(child, args, int(out["nit"]))
            child = None
            cb = out["cpu_baseline"]
            if cb and args.workload == "cfg5":  # one solve of the batch network = all realisations
                cb["value"] *= args.realisations
                cb["unit"] = "realisations/s"
            recs = out.pop("_flow_records", None)
            if out.get("flow") is not None and recs is not None:
                fl = out["flow"]
                fl["cpu_baseline"] = _flow_cpu(*recs, fl["streamlines"]["lines"])
                fl["speedup_vs_cpu"] = {
                    "streamlines": fl["streamlines"]["points_per_s"] / fl["cpu_baseline"]["streamlines_points_per_s"],
                    "flow_nets": fl["flow_nets"]["points_per_s"] / fl["cpu_baseline"]["flow_net_points_per_s"]}
            if cb:
                out["speedup_vs_cpu"] = {"value": out["value"] / cb["value"],
                                         "bound": "lower" if cb["bound"].startswith("upper") else "measured",
                                         "drop_in": out["drop_in"]["solves_per_s"] / cb["value"]
                                         if out.get("drop_in") else None}
        out.pop("_flow_records", None)
        print(json.dumps(out))
    finally:
        if child is not None and child.poll() is None:
            child.kill()


def gpu_part(args):
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
    # the network built twice: pair search, chords and connectivity on the GPU / in the library
    # (the network solved below), and by the host builder (numpy, SciPy) for the time beside it
    t0 = time.perf_counter()
    dfn = build_network(spec, device=True)
    dfn.consolidate_dfn()
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    dh = build_network(spec, device=False)
    dh.consolidate_dfn()
    t_build_host = time.perf_counter() - t0
    same_structure = bool(np.array_equal(dh.elements_struc_array_hpc["_type"], dfn.elements_struc_array_hpc["_type"])
                          and np.array_equal(dh.fractures_struc_array_hpc["elements"],
                                             dfn.fractures_struc_array_hpc["elements"]))
    endpoint_diff = float(np.max(np.abs(dh.elements_struc_array_hpc["endpoints0"]
                                        - dfn.elements_struc_array_hpc["endpoints0"]))) if same_structure else None
    del dh
    build = {"device_s": t_build, "host_s": t_build_host, "same_records_structure": same_structure,
             "max_endpoint_diff_m": endpoint_diff,
             "note": "build_network(device=True): candidate pairs and chords on the GPU, connectivity in the "
                     "library's C++; host: numpy KD-tree search and chords, SciPy components; both include the "
                     "Python objects and consolidation"}
    fa, ea = dfn.fractures_struc_array_hpc, dfn.elements_struc_array_hpc
    fa0, ea0 = fa.copy(), ea.copy()  # the unsolved records (the handle's downloads overwrite fa, ea)
    net = DeviceNetwork(fa, ea, dfn.discharge_int, dfn.constants)
    print(f"# network: {len(fa)} fractures, {len(ea)} elements (built in {t_build:.1f}s); warmup {args.warmup}, "
          f"steps {args.steps}", file=sys.stderr, flush=True)
    other = "mixed" if args.precision == "fp64" else "fp64"
    forms, states, lasts = {}, {}, {}
    obs_frac, obs_z = _obs_points(fa, 100)
    for prec in ([args.precision] + ([] if args.no_compare else [other])):
        infos, walls = _time_form(net, prec, args.lu_mode, args.warmup, args.steps, prec)
        q, c = net.state_discharges()
        net.download()
        states[prec] = (ea["coef"].copy(), q, c, net.state_heads(obs_frac, obs_z))
        last = infos[-1]
        lasts[prec] = last
        forms[prec] = {"solves_per_s": args.steps / sum(walls), "ms_per_solve": 1e3 * sum(walls) / args.steps,
                       "nit": int(last["nit"]), "nit_fp32": int(last["nit_fp32"]),
                       "solves_fp32": int(last["solves_fp32"]), "refine_rate": last["refine_rate"],
                       "head_residual": last["head_residual"], "t_factor_s": last["t_factor"],
                       "t_iterate_s": last["t_iterate"], "k_sweep_ms": last["k_sweep_ms"],
                       "k_rhs_ms": last["k_rhs_ms"], "k_fit_ms": last["k_fit_ms"], "k_lu_ms": last["k_lu_ms"],
                       **_launch_info(last), "walls": walls}
    main_form = forms[args.precision]
    last = lasts[args.precision]
    walls = main_form.pop("walls")
    if other in forms:
        forms[other].pop("walls")
        a, b = states["fp64"], states["mixed"]

        def rel(x, y):
            return float(np.max(np.abs(x - y)) / max(np.max(np.abs(y)), 1e-300))

        forms["mixed"]["error_vs_fp64"] = {"coef_rel": rel(b[0], a[0]), "q_rel": rel(b[1], a[1]),
                                           "constant_rel": rel(b[2], a[2]),
                                           "heads_abs_m": float(np.max(np.abs(b[3] - a[3]))),
                                           "note": "both forms stopped at head residual <= 1e-6 m; heads at 100 "
                                                   "observation points"}
        forms["mixed_over_fp64"] = forms["mixed"]["solves_per_s"] / forms["fp64"]["solves_per_s"]
    fixed_point = None
    if not args.no_compare and not args.max_iterations:
        # how far the states at head residual 1e-6 are from the fixed point: an FP64 solve of the
        # same network to head residual 1e-9 (MAX_ITERATIONS raised for it)
        cref = dfn.constants.copy()
        cref["MAX_ITERATIONS"] = 5000
        fr_, er_ = fa0.copy(), ea0.copy()
        nref = DeviceNetwork(fr_, er_, dfn.discharge_int, cref)
        iref = nref.run("fp64", args.lu_mode, 1e-9)
        qr, cr = nref.state_discharges()
        nref.download()
        ref = (er_["coef"].copy(), qr, cr, nref.state_heads(obs_frac, obs_z))
        nref.close()

        def rel_(x, y):
            return float(np.max(np.abs(x - y)) / max(np.max(np.abs(y)), 1e-300))

        for prec, st in states.items():
            forms[prec]["error_vs_fixed_point"] = {
                "coef_rel": rel_(st[0], ref[0]), "q_rel": rel_(st[1], ref[1]), "constant_rel": rel_(st[2], ref[2]),
                "heads_abs_m": float(np.max(np.abs(st[3] - ref[3])))}
        fixed_point = {"head_residual": iref["head_residual"], "nit": int(iref["nit"]),
                       "note": "FP64 solve to head residual 1e-9 m; forms.*.error_vs_fixed_point compare the timed "
                               "forms' states (head residual <= 1e-6 m) with it"}
    adjoint = None
    if not args.no_adjoint:
        # gradient of Q(x=0) + sum of 100 observation heads w.r.t. every log10 T and the two boundary heads
        net.run("fp64", args.lu_mode, HEAD_TOL)
        t_fwd = forms["fp64"]["ms_per_solve"] / 1e3 if "fp64" in forms else None
        groups = boundary_groups(dfn)
        dq = np.zeros(len(ea))
        dq[groups["x0"]] = 1.0
        aw = []
        for k in range(1 + min(args.steps, 3)):
            t0 = time.perf_counter()
            gt, gh, ai = net.adjoint(dq=dq, obs_frac=obs_frac, obs_z=obs_z, dhead=np.ones(len(obs_frac)),
                                     tol=ADJOINT_TOL)
            aw.append(time.perf_counter() - t0)
            print(f"# adjoint {k}: {aw[-1]:.3f}s iterations={ai['nit']} residual={ai['delta']:.2e}",
                  file=sys.stderr, flush=True)
        t_adj = sum(aw[1:]) / (len(aw) - 1)
        adjoint = {"objective": "Q through x=0 + sum of heads at 100 observation points",
                   "gradient": f"d/d log10 T for all {len(fa)} fractures and d/d head of the 2 boundary planes",
                   "method": "implicit fixed-point adjoint, GMRES on the transposed iteration (FP64)",
                   "tol": ADJOINT_TOL, "iterations": int(ai["nit"]), "residual": ai["delta"],
                   "ms": 1e3 * t_adj, "first_call_ms": 1e3 * aw[0],
                   "adjoint_forward_ratio": (t_adj / t_fwd) if t_fwd else None}
    flow = _flow_bench(net, fa, args.flow_lines) if args.flow_lines > 0 else None
    net.close()
    drop_in = None
    if args.dropin_steps > 0:
        dw, dinfo = _time_dropin(fa0, ea0, dfn, args.precision, args.dropin_steps)
        drop_in = {"solves_per_s": len(dw) / sum(dw), "ms_per_solve": 1e3 * sum(dw) / len(dw), "calls": len(dw),
                   "t_setup_s": dinfo[-1]["t_setup"], "t_factor_s": dinfo[-1]["t_factor"],
                   "t_iterate_s": dinfo[-1]["t_iterate"], "factor_rounds": int(dinfo[-1]["factor_rounds"]),
                   "note": "andfn_amd.hpc_solve.solve per call: records in, new handle (upload, symbolic analysis, "
                           "every pivot-delay round, factorisation), iterations, records out"}
    total = sum(walls)
    f32 = args.precision == "mixed" and last["solves_fp32"] > 0
    return {
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
                        f"K {spec.discharge_int}" + (f", {spec.n_wells} wells" if spec.n_wells else ""),
            "fractures_kept": int(len(fa)),
            "elements": int(len(ea)),
            "unknowns": int(last["n_unknowns"]),
            "precision": args.precision,
            "stop": f"head residual <= {HEAD_TOL:g} m (max |A y - b(c)| over the head equations), "
                    f"MAX_ITERATIONS {int(dfn.constants['MAX_ITERATIONS'])}",
            "timed": "DeviceNetwork.run on a resident handle: assembly, numeric factorisation and the iterations "
                     "every step (drop_in times the full drop-in call)",
        },
        "nit": int(last["nit"]),
        "head_residual": last["head_residual"],
        "bc_residual": last["bc_residual"],
        "converged": bool(last["converged"]),
        "t_factor_s": last["t_factor"],
        "t_iterate_s": last["t_iterate"],
        "nnz_factors": int(last["nnz_factors"]),
        "forms": forms,
        "fixed_point": fixed_point,
        "build": build,
        "drop_in": drop_in,
        "adjoint": adjoint,
        "flow": flow,
        "_flow_records": (fa, ea, dfn.discharge_int) if flow is not None else None,  # for the CPU leg
        # the FP64 form's kernels (one sweep kernel throughout; the mixed form switches between
        # its FP32-far-field and FP64 sweeps and between FP32 and FP64 tiles)
        "roofline": _roofline_sweep(lasts["fp64"] if "fp64" in lasts else last, "fp64" in lasts, args.workload),
        "roofline_solve": _roofline_solve(lasts["fp64"] if "fp64" in lasts else last,
                                          False if "fp64" in lasts else f32, args.workload),
        "cpu_baseline": None,
    }


if __name__ == "__main__":
    main()
