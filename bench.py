"""Benchmark: three-phonon scattering rates on one MI355X.

Workload (default, --workload cell10): the 10-atom cell (30 branches), 20x20x20 Gamma-centred
mesh with time reversal (4004 irreducible q0, 8000 q1 each, 30^3 branch triplets per pair),
Gaussian sigma = 0.1 rad/ps, T = 300 K.  A step is one slice of the irreducible q0; the `steps`
slices cover the wedge once, so the timed region yields every linewidth and the RTA kappa.
FP64 is the headline; FP32 and split-BF16 matrix-core runs report their kappa and linewidth
errors against it.  --workload rocksalt16 is the 2-atom 16^3 line (one step = every q0).

Prints one JSON line:
  metric/value   |V3|^2 triplets per second (sum over irreducible q0 of Nq * Ndof^3 per
                 sweep, divided by the sweep time), FP64 path
  roofline       the dominant kernel's algorithmic FLOP rate (HIP events on the library's
                 stream) against MI355X's peak for its precision
  cpu_baseline   the FP64 CPU oracle timed on a bounded sample of the workload on the host cores
  fp32, bf16     the reduced-precision paths: rate, speedup, error vs FP64
  kappa_tts_s    wall time from FC + structure in memory to kappa (phonons included)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 (vector and matrix are the same rate)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--workload", default="cell10",
                   choices=["rocksalt16", "cell10", "learned10", "fourph10", "fourph_tetra", "bte16",
                            "bte10", "phonons10", "tetra16", "tetra10"],
                   help="cell10 (default, the headline line): 10-atom cell, 20^3 mesh with time "
                        "reversal, FP64 vs FP32 vs BF16; rocksalt16: 2-atom cell, 16^3 mesh; learned10: the "
                        "learned importance-sampling proposal on the cell10 workload; fourph10: "
                        "four-phonon linewidths, 2-atom cell with FC4, 10^3 mesh; fourph_tetra: the "
                        "same with the tetrahedron method on a --mesh4 mesh; tetra16: three-phonon "
                        "linewidths with the tetrahedron method (the reference's default delta) "
                        "on the 2-atom --mesh mesh; tetra10: the same on the cell10 workload (10-atom "
                        "cell, --mesh10 mesh with time reversal); bte16: iterative "
                        "BTE of the 2-atom cell on a 16^3 mesh (--mesh); bte10: the same for the 10-atom "
                        "cell on a --mesh-bte10 mesh (W on the matrix cores, timed against the "
                        "generic-kernel store); phonons10: harmonic phonons "
                        "(D(q), eigensolve, group velocities) of the 10-atom cell on its 20^3 mesh")
    p.add_argument("--samples", type=int, default=2_000_000, help="learned10: training samples")
    p.add_argument("--budget", type=float, default=0.05, help="learned10: sampled fraction of q1")
    p.add_argument("--pair-prec", default="fp32", choices=["fp64", "fp32", "bf16"],
                   help="learned10: precision of the sampled pair terms")
    p.add_argument("--head", type=float, default=0.5,
                   help="learned10: fraction of the budget evaluated exactly on the highest-scored q1 "
                        "(the '+head' estimators)")
    p.add_argument("--mesh10", type=int, default=20)
    p.add_argument("--mesh4", type=int, default=8, help="fourph_tetra: mesh size")
    p.add_argument("--mesh-bte10", type=int, default=8, help="bte10: mesh size")
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--mesh", type=int, default=16)
    p.add_argument("--sigma", type=float, default=0.1)
    p.add_argument("--T", type=float, default=300.0)
    p.add_argument("--cpu-q0", type=int, default=145, help="irreducible q0 in one CPU pass")
    p.add_argument("--cpu-seconds", type=float, default=15.0,
                   help="repeat CPU passes until this much wall time is spent")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--precisions", default="fp64,fp32,bf16",
                   help="cell10: the precisions to sweep (comma list; fp64 is the reference of the errors "
                        "and always runs)")
    p.add_argument("--no-tts", action="store_true",
                   help="cell10: skip the end-to-end kappa time-to-solution runs (profiling)")
    p.add_argument("--phonons", default="device", choices=["device", "host"],
                   help="cell10: phonons of the mesh from spr_get_phonon (device) or the host LAPACK "
                        "path; both feed the FP64 reference, the reduced-precision runs and the CPU leg")
    return p.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "cell10":
        return cell10(args)
    if args.workload == "learned10":
        return learned10(args)
    if args.workload == "fourph10":
        return fourph10(args)
    if args.workload == "fourph_tetra":
        return fourph_tetra(args)
    if args.workload == "tetra16":
        return tetra16(args)
    if args.workload == "tetra10":
        return tetra16(args, cell10=True)
    if args.workload == "bte16":
        return bte16(args)
    if args.workload == "bte10":
        return bte16(args, cell10=True)
    if args.workload == "phonons10":
        return phonons10(args)
    from spring_amd import model, mesh, phonons, lib

    t_in0 = time.perf_counter()
    m = model.rocksalt2()
    t_model = time.perf_counter() - t_in0
    n = args.mesh
    t0 = time.perf_counter()
    qm = mesh.make_mesh(m.crystal, (n, n, n))
    ph = phonons.solve(m.crystal, m.fc2, qm.q_cart, qm.q_int, args.T)
    ctx = lib.GpuContext(m, qm, ph)
    lw, kap, _ = ctx.lw3_kappa_all(args.sigma, "fp64", args.T)     # also builds the pair lists
    kappa_tts = time.perf_counter() - t0

    nd = m.crystal.ndof
    triplets = qm.nq0 * qm.nq * nd ** 3

    def run(prec, steps, warmup):
        # one step: linewidths of every irreducible q0 and kappa, LW_Allq0 kept on the device
        for _ in range(warmup):
            ctx.lw3_kappa_all(args.sigma, prec, args.T)
        ctx.kernel_time_reset()
        t = time.perf_counter()
        for _ in range(steps):
            l, k, _ = ctx.lw3_kappa_all(args.sigma, prec, args.T)
        dt = time.perf_counter() - t
        nl, kms = ctx.kernel_time()
        return dt / steps, kms / max(nl, 1), l, k

    step64, kern64, lw64, k64 = run("fp64", args.steps, args.warmup)
    step32, kern32, lw32, k32 = run("fp32", args.steps, args.warmup)

    # algorithmic FLOP per executed pair: phase sum over the nonzero FC3 blocks (4 flop per
    # real x complex MAC, 27 per block) + three complex mode products (24 Ndof^4) + ~30 per
    # branch triplet for |V|^2, 1/(w w w) and the three Gaussians
    nblocks = int(np.count_nonzero(np.abs(m.fc3.fc).sum(axis=(3, 4, 5))))
    flop_pair = 4 * 27 * nblocks + 24 * nd ** 4 + 30 * nd ** 3
    npairs = sum(len(mesh.pair_list(qm, int(i))[0]) for i in qm.irr_idx)
    achieved = flop_pair * npairs / (kern64 * 1e-3) / 1e12

    err_k = float(np.max(np.abs(k32 - k64)) / np.max(np.abs(k64)))
    sig = np.abs(lw64) > 1e-6 * np.abs(lw64).max()
    err_lw = float(np.max(np.abs(lw32 - lw64)[sig] / np.abs(lw64)[sig]))

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        inp = O.Inputs(m, qm, ph)
        threads, _, _, _ = _host_cpu()
        rng = np.random.default_rng(0)
        pos = rng.choice(qm.nq0, size=min(args.cpu_q0, qm.nq0), replace=False)
        sample = qm.irr_idx[pos]
        reps, dt = 0, 0.0
        t = time.perf_counter()
        while dt < args.cpu_seconds and reps < 64:
            lw_cpu = _cpu_lw3_pairs(O, inp, qm, sample, args.sigma, threads)
            reps += 1
            dt = time.perf_counter() - t
        gs = np.abs(lw_cpu) > 1e-6 * np.abs(lw_cpu).max()
        cpu_err = float(np.max(np.abs(lw64[pos] - lw_cpu)[gs] / np.abs(lw_cpu)[gs]))
        cpu = {"value": reps * len(sample) * qm.nq * nd ** 3 / dt, "unit": "triplets/s",
               "cores": threads, "kind": "port",
               "sample": f"{reps} passes over {len(sample)} seeded irreducible q0 of the same "
                         f"workload, every q1 of each through its pair-permutation reduced list (each "
                         f"pair evaluated once for both orderings, as on the GPU), FP64 C "
                         f"restatement, {threads} threads, {dt:.1f} s wall; GPU FP64 linewidths "
                         f"agree to max rel err {cpu_err:.1e}"}

    traffic = None
    # HBM bytes per launch of the FP64 kernel, from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE summary of this workload (counters scaled by their copy-kernel calibration)
    tf = os.path.join(ROOT, "profiles", "k_lw3_small_rocksalt16.md")
    if os.path.exists(tf):
        for line in open(tf):
            if line.startswith("traffic_bytes_per_launch_fp64:"):
                traffic = int(line.split(":")[1])

    out = {
        "metric": "|V3|^2 triplets/s and kappa time-to-solution at <=1e-3 rel err vs FP64 full sum",
        "value": triplets / step64,
        "unit": "triplets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step64 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded rocksalt spring model + random permutation-symmetric ASR FC3",
        "config": {"workload": f"3-phonon Gaussian linewidths + RTA kappa, 2-atom rocksalt, {n}^3 mesh, "
                               f"{qm.nq0} irreducible q0 x {qm.nq} q1 x {nd}^3 branches",
                   "mesh": n, "nq0": qm.nq0, "ndof": nd, "sigma_rad_per_ps": args.sigma, "T_K": args.T,
                   "pairs_executed_per_step": npairs, "triplets_per_step": triplets},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "compute_unit": "vector ALU (FP64 FMA; MI355X's FP64 vector and matrix peaks are equal)",
                     "kernel": "k_lw3_small<double,2>", "kernel_ms": kern64,
                     "flop_per_pair": flop_pair, "pairs_per_launch": npairs},
        "cpu_baseline": cpu,
        "fp32": {"value": triplets / step32, "ms_per_step": step32 * 1e3, "kernel_ms": kern32,
                 "speedup_vs_fp64": step64 / step32, "kappa_rel_err_vs_fp64": err_k,
                 "linewidth_max_rel_err_vs_fp64": err_lw},
        "kappa_tts_s": kappa_tts,
        "kappa_W_mK": [float(k64[0, 0]), float(k64[1, 1]), float(k64[2, 2])],
        "host_model_s": t_model,
    }
    print(json.dumps(out), flush=True)


FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0


def _host_cpu():
    """Host threads to use (all cores this process may run on, within OMP_NUM_THREADS when the
    host sets it), the CPU model and the oracle's compiler flags."""
    cores = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    threads = min(cores, int(env)) if env and env.isdigit() and int(env) > 0 else cores
    model_name = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model_name = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    flags = "unknown"
    try:
        for line in open(os.path.join(ROOT, "oracle", "Makefile")):
            if line.startswith("CFLAGS"):
                flags = "gcc " + line.split("=", 1)[1].strip()
    except OSError:
        pass
    return threads, cores, model_name, flags


def _cpu_lw3_pairs(O, inp, qm, grids, sigma, threads):
    """FP64 CPU linewidths of the given q0 (grid indices) over their pair-permutation reduced
    lists (each pair evaluated once for both orderings, as the device does), the pairs of all q0
    cut into one chunk per thread and the chunks summed per q0 in order."""
    from concurrent.futures import ThreadPoolExecutor
    from spring_amd import mesh
    work = []
    for j, g in enumerate(grids):
        q1s, q2s, _ = mesh.pair_list(qm, int(g))
        pr = np.stack([q1s, q2s], axis=1)
        for c in np.array_split(pr, max(1, threads // len(grids))):
            if len(c):
                work.append((j, int(g), c))
    with ThreadPoolExecutor(threads) as ex:
        parts = list(ex.map(lambda w: O.lw3_gauss_pairs(inp, w[1], sigma, w[2]), work))
    lw = np.zeros((len(grids), inp.nd))
    for (j, _, _), p in zip(work, parts):
        lw[j] += p
    return lw


def _profile_value(name, key):
    """An integer recorded as `key: value` in a committed profile summary, or None."""
    tf = os.path.join(ROOT, "profiles", name)
    if os.path.exists(tf):
        for line in open(tf):
            if line.startswith(key + ":"):
                return int(line.split(":")[1])
    return None


def cell10(args):
    """The 10-atom workload (30 branches) on a mesh10^3 mesh with time reversal: all irreducible
    q0, every q1, Gaussian sigma, 300 K, FP64 against the FP32 and split-BF16 matrix-core paths.

    A step is one slice of the irreducible q0 (spr_lw3_slice, DistributeQpoints' (offset, count)
    split): the wedge is cut into `steps` slices, so the timed steps cover every irreducible q0
    exactly once and their linewidths give kappa.  `warmup` slices run untimed before each
    precision (tuning, graph capture).  The kappa time to solution of each precision is measured
    end to end (mesh, device phonons, upload, pair lists, linewidths, kappa) in a fresh context."""
    from spring_amd import model, mesh, phonons, lib
    t0 = time.perf_counter()
    m = model.rocksalt10()
    t_model = time.perf_counter() - t0
    n = args.mesh10
    log = lambda *a: print("#", *a, file=sys.stderr, flush=True)

    def setup():
        t = time.perf_counter()
        qm = mesh.make_mesh(m.crystal, (n, n, n), time_rev=1)
        t_mesh = time.perf_counter() - t
        if args.phonons == "device":
            ph = lib.get_phonon(m.crystal, m.fc2, qm.q_cart, qm.q_int, args.T)
        else:
            ph = phonons.solve(m.crystal, m.fc2, qm.q_cart, qm.q_int, args.T)
        t_ph = time.perf_counter() - t - t_mesh
        return qm, ph, t_mesh, t_ph

    qm, ph, t_mesh, t_ph = setup()
    ctx = lib.GpuContext(m, qm, ph)
    log(f"cell10: model {t_model:.1f} s, mesh {t_mesh:.2f} s, phonons ({args.phonons}) {t_ph:.2f} s, nq0 {qm.nq0}")
    nd = m.crystal.ndof
    steps = max(1, min(args.steps, qm.nq0))
    bounds = np.linspace(0, qm.nq0, steps + 1).round().astype(int)
    triplets = qm.nq0 * qm.nq * nd ** 3
    nblocks = int(np.count_nonzero(np.abs(m.fc3.fc).sum(axis=(3, 4, 5))))
    flop_pair = 4 * 27 * nblocks + 24 * nd ** 4 + 30 * nd ** 3
    npairs = sum(len(mesh.pair_list(qm, int(i))[0]) for i in qm.irr_idx)

    def sweep(prec):
        for k in range(args.warmup):
            a, b = bounds[k % steps], bounds[k % steps + 1]
            ctx.lw3_slice(int(a), int(b - a), args.sigma, prec)
        ctx.kernel_time_reset()
        lw = np.zeros((qm.nq0, nd))
        t = time.perf_counter()
        for k in range(steps):
            a, b = bounds[k], bounds[k + 1]
            lw[a:b] = ctx.lw3_slice(int(a), int(b - a), args.sigma, prec)
        dt = time.perf_counter() - t
        nl, kms = ctx.kernel_time()
        k3, _ = ctx.kappa(lw, args.T, symm=True)
        log(f"cell10 {prec}: {steps} steps {dt:.2f} s, kernel {kms:.0f} ms over {nl} launches")
        return dict(step=dt / steps, total=dt, kern=kms / max(nl, 1), launches=nl, lw=lw, k=k3)

    precs = ["fp64"] + [x for x in ("fp32", "bf16") if x in args.precisions.split(",")]
    res = {p_: sweep(p_) for p_ in precs}
    r64 = res["fp64"]
    sig = np.abs(r64["lw"]) > 1e-6 * np.abs(r64["lw"]).max()
    peaks = {"fp64": FP64_PEAK_TFLOPS, "fp32": FP32_PEAK_TFLOPS, "bf16": BF16_PEAK_TFLOPS}
    pairs_per_launch = npairs / steps

    def kappa_err(prec):
        r = res[prec]
        return float(np.max(np.abs(r["k"] - r64["k"])) / np.max(np.abs(r64["k"])))

    # kappa time to solution, end to end in a fresh context (the compiled kernels, the tuning
    # table and the graphs of the sweeps above are not reused: a new context captures anew)
    tts = {}
    for prec in ([] if args.no_tts else [x for x in ("bf16", "fp32") if x in res]):
        t = time.perf_counter()
        qm2, ph2, _, _ = setup()
        c2 = lib.GpuContext(m, qm2, ph2)
        _, k2, _ = c2.lw3_kappa_all(args.sigma, prec, args.T)
        tts[prec] = time.perf_counter() - t
        c2.close()
        log(f"cell10 kappa time to solution {prec}: {tts[prec]:.2f} s")

    def entry(prec):
        r = res[prec]
        achieved = flop_pair * pairs_per_launch / (r["kern"] * 1e-3) / 1e12
        e = {"value": triplets / r["total"], "ms_per_step": r["step"] * 1e3, "sweep_s": r["total"],
             "kernel_ms_per_launch": r["kern"], "speedup_vs_fp64": r64["total"] / r["total"],
             "roofline": {"bound": "mfma", "achieved": achieved, "peak": peaks[prec], "unit": "TFLOP/s",
                          "frac": achieved / peaks[prec]}}
        if prec != "fp64":
            e["kappa_rel_err_vs_fp64"] = kappa_err(prec)
            e["linewidth_max_rel_err_vs_fp64"] = float(
                np.max(np.abs(r["lw"] - r64["lw"])[sig] / np.abs(r64["lw"])[sig]))
            e["linewidth_mean_rel_err_vs_fp64"] = float(
                np.mean(np.abs(r["lw"] - r64["lw"])[sig] / np.abs(r64["lw"])[sig]))
            e["kappa_tts_s"] = tts.get(prec)
        else:
            # FP64 is the reference of the error: its time to solution is composed of the measured
            # parts (mesh, phonons, the timed sweep); the reduced precisions are measured whole
            e["kappa_tts_s_composed"] = t_mesh + t_ph + r["total"]
        return e

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        inp = O.Inputs(m, qm, ph)
        threads, cores, cpu_model, cflags = _host_cpu()
        rng = np.random.default_rng(0)
        qpos = rng.choice(qm.nq0, size=qm.nq0, replace=False)
        # calibrate one thread on a few pairs, then size the sample (whole q0, each through its
        # reduced pair list of about Nq/2 pairs) to ~args.cpu_seconds
        g0 = int(qm.irr_idx[qpos[0]])
        q1s, q2s, _ = mesh.pair_list(qm, g0)
        t = time.perf_counter()
        O.lw3_gauss_pairs(inp, g0, args.sigma, np.stack([q1s[:8], q2s[:8]], axis=1))
        per_pair = (time.perf_counter() - t) / 8
        pairs = max(threads, int(args.cpu_seconds * threads / per_pair))
        nq0_s = max(1, min(len(qpos), pairs // max(1, len(q1s))))
        t = time.perf_counter()
        lw_cpu = _cpu_lw3_pairs(O, inp, qm, qm.irr_idx[qpos[:nq0_s]], args.sigma, threads)
        dt = time.perf_counter() - t
        gs = np.abs(lw_cpu) > 1e-6 * np.abs(lw_cpu).max()
        cpu_err = float(np.max(np.abs(r64["lw"][qpos[:nq0_s]] - lw_cpu)[gs] / np.abs(lw_cpu)[gs]))
        cpu = {"value": nq0_s * qm.nq * nd ** 3 / dt, "unit": "triplets/s", "cores": threads,
               "kind": "port", "host_cores_visible": cores, "cpu_model": cpu_model, "compiler": cflags,
               "sample": f"{nq0_s} seeded irreducible q0 of the same workload, every q1 of each "
                         f"through its pair-permutation reduced list (each pair evaluated once for both "
                         f"orderings, as on the GPU), pairs split over {threads} threads, FP64 C "
                         f"restatement, {dt:.1f} s wall; GPU FP64 linewidths of those q0 agree to max rel "
                         f"err {cpu_err:.1e}"}
    traffic = _profile_value("k_lw3_mm_cell10.md", f"traffic_bytes_per_launch_fp64_{n}_{steps}")
    e64 = entry("fp64")
    out = {
        "metric": "|V3|^2 triplets/s and kappa time-to-solution at <=1e-3 rel err vs FP64 full sum",
        "value": triplets / r64["total"], "unit": "triplets/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": r64["step"] * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded jittered 10-atom rocksalt supercell, spring FC2, random "
                "permutation-symmetric ASR FC3 within the 3rd shell",
        "config": {"workload": f"3-phonon Gaussian linewidths + RTA kappa, 10-atom cell, {n}^3 mesh "
                               f"with time reversal, {qm.nq0} irreducible q0 x {qm.nq} q1 x {nd}^3 branches; "
                               f"a step is one of {steps} slices of the irreducible q0",
                   "mesh": n, "nq0": qm.nq0, "ndof": nd, "sigma_rad_per_ps": args.sigma, "T_K": args.T,
                   "pairs_executed_per_sweep": npairs, "triplets_per_sweep": triplets,
                   "flop_per_pair": flop_pair},
        "roofline": {"bound": "mfma", "achieved": e64["roofline"]["achieved"], "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": e64["roofline"]["frac"], "traffic": traffic,
                     "kernel": "k_lw3_mm<0,10> (FP64; tile shape picked by the launcher tuner)", "kernel_ms": r64["kern"], "flop_per_pair": flop_pair,
                     "pairs_per_launch": pairs_per_launch},
        "cpu_baseline": cpu,
        **{p_: (e64 if p_ == "fp64" else entry(p_)) for p_ in precs},
        "kappa_tts_s": min(tts.values()) if tts else None,
        "kappa_tts_precision": min(tts, key=tts.get) if tts else None,
        "kappa_tts_rel_err_vs_fp64": kappa_err(min(tts, key=tts.get)) if tts else None,
        "phonons_s": t_ph, "mesh_s": t_mesh, "phonons": args.phonons, "host_model_s": t_model,
        "kappa_W_mK": [float(r64["k"][i, i]) for i in range(3)],
    }
    print(json.dumps(out), flush=True)


def learned10(args):
    """Learned importance-sampling proposal on the 10-atom cell.  The pair-level network
    (features of (q0, q1) -> log of the pair's share of the linewidths) is trained on the pair
    terms of args.samples / Nq random training q0 (every q1 of each, computed by the method's
    FP32 pair path), scores every (q0, q1) on the GPU, and K = budget * Nq sampled q1 per q0 (all
    branches of each sampled pair, pair terms in args.pair_prec) estimate every linewidth; kappa
    and the per-mode linewidths are judged against the FP64 full sum.  The G-only proposal (the
    same pair weights without the network) and uniform sampling are reported beside it, and each
    proposal's kappa error is swept over budgets to find the smallest that reaches 1e-3."""
    from spring_amd import model, mesh, phonons, lib
    from spring_amd import learned as Lrn
    n = args.mesh10
    m = model.rocksalt10()
    t0 = time.perf_counter()
    qm = mesh.make_mesh(m.crystal, (n, n, n), time_rev=1)
    ph = phonons.solve(m.crystal, m.fc2, qm.q_cart, qm.q_int, args.T)
    t_ph = time.perf_counter() - t0
    ctx = lib.GpuContext(m, qm, ph)
    nd = m.crystal.ndof
    log = lambda *a: print("#", *a, file=sys.stderr, flush=True)
    t = time.perf_counter()
    lw_full = ctx.lw3_all(args.sigma, "fp64")
    k_full, _ = ctx.kappa(lw_full, args.T, symm=True)
    t_full = time.perf_counter() - t
    log(f"learned10: FP64 full sum {t_full:.2f} s, nq0 {qm.nq0}")
    wmax = float(np.abs(ph.omega).max())
    n_train = max(1, min(qm.nq0 // 2, int(round(args.samples / qm.nq))))
    q0_train = np.random.default_rng(0).permutation(qm.nq0)[:n_train]
    t = time.perf_counter()
    X, y = Lrn.pair_training_set(ctx, qm, ph, q0_train, args.sigma, prec="fp32")
    t_set = time.perf_counter() - t
    t = time.perf_counter()
    prop = Lrn.PairProposal.train(X, y, wmax, epochs=10)
    t_train = time.perf_counter() - t
    log(f"learned10: {y.size} pair samples from {n_train} q0 in {t_set:.1f} s, trained in {t_train:.1f} s, "
        f"loss {prop.losses[0]:.3f} -> {prop.losses[-1]:.3f}")
    qis = np.arange(qm.nq0)
    sig = np.abs(lw_full) > 1e-6 * np.abs(lw_full).max()
    scores = {}
    for name in ("learned", "g_only", "uniform"):
        t = time.perf_counter()
        if name == "uniform":
            S = np.ones((qm.nq0, qm.nq))
        else:
            S = Lrn.pair_proposal_scores(prop if name == "learned" else None, ctx, qis, args.sigma, wmax)
        scores[name] = (S, time.perf_counter() - t)

    def estimate(name, budget, seed):
        # "+head": the highest-scored half of the budget is evaluated exactly, the rest sampled
        # from the remaining q1 (learned.head_tail_draw)
        S, t_score = scores[name.split("+")[0]]
        K = max(1, int(round(budget * qm.nq)))
        rng = np.random.default_rng(seed)
        t = time.perf_counter()
        q0s, q1s, ws = [], [], []
        for qi in range(qm.nq0):
            if name.endswith("+head"):
                pick, w = Lrn.head_tail_draw(S[qi], K, args.head, rng)
            else:
                pr = Lrn.mixture(S[qi])
                pick = rng.choice(qm.nq, size=K, replace=True, p=pr)
                w = 1.0 / (K * pr[pick])
            q0s.append(np.full(pick.size, qi))
            q1s.append(pick)
            ws.append(w)
        q0s, q1s, ws = np.concatenate(q0s), np.concatenate(q1s), np.concatenate(ws)
        t_draw = time.perf_counter() - t
        t = time.perf_counter()
        est = ctx.lw3_sampled(q0s, q1s, ws, args.sigma, args.pair_prec)
        k_est, _ = ctx.kappa(est, args.T, symm=True)
        t_pairs = time.perf_counter() - t
        err = np.abs(est - lw_full)[sig] / np.abs(lw_full)[sig]
        return {"kappa_rel_err_vs_fp64": float(np.max(np.abs(k_est - k_full)) / np.max(np.abs(k_full))),
                "linewidth_rel_err_median": float(np.median(err)), "linewidth_rel_err_mean": float(err.mean()),
                "linewidth_rel_err_p90": float(np.quantile(err, 0.9)), "score_s": t_score, "draw_s": t_draw,
                "sampled_pairs_s": t_pairs, "total_s": t_score + t_draw + t_pairs,
                "speedup_vs_fp64_full": t_full / (t_score + t_draw + t_pairs)}

    res, sweep = {}, {}
    budgets = [0.01, 0.02, 0.05, 0.1, 0.2, 0.4]
    for name in ("learned", "learned+head", "g_only", "g_only+head", "uniform"):
        res[name] = estimate(name, args.budget, 1)
        log(f"learned10 {name} at {args.budget:.0%}: {res[name]}")
        sweep[name] = {}
        reach = None
        for b in budgets:
            e = estimate(name, b, 2)
            sweep[name][f"{b:g}"] = {"kappa_rel_err": e["kappa_rel_err_vs_fp64"], "total_s": e["total_s"],
                                     "linewidth_rel_err_median": e["linewidth_rel_err_median"]}
            if reach is None and e["kappa_rel_err_vs_fp64"] <= 1e-3:
                reach = b
        res[name]["budget_reaching_kappa_1e-3"] = reach
        log(f"learned10 {name} sweep: {sweep[name]}")
    out = {
        "metric": "learned importance sampling of the three-phonon linewidth sum: kappa and linewidth "
                  "error vs the FP64 full sum, time vs the FP64 GPU path",
        "config": {"workload": f"10-atom cell, {n}^3 mesh with time reversal, {qm.nq0} irreducible q0, "
                               f"{int(round(args.budget * qm.nq))} sampled q1 per q0 ({args.budget:.0%} of Nq), "
                               f"all branches per pair, {args.pair_prec} pair terms",
                   "sigma_rad_per_ps": args.sigma, "T_K": args.T, "training_samples": int(y.size),
                   "training_q0": int(n_train),
                   "network": "12 pair features -> 128 -> 128 -> 1 tanh, FP32 Adam training, BF16 MFMA scoring"},
        "fp64_full_s": t_full, "phonons_s": t_ph, "training_set_s": t_set, "training_s": t_train,
        "training_loss": [float(prop.losses[0]), float(prop.losses[-1])],
        "proposals": res, "budget_sweep": sweep,
    }
    print(json.dumps(out), flush=True)


def fourph10(args):
    """Four-phonon Gaussian linewidths of the 2-atom model (FC4 within the 1st shell) on a
    10^3 mesh: FP64 against FP32, and the FP64 CPU oracle on a sample of q0."""
    from spring_amd import model, mesh, phonons, lib
    m = model.rocksalt2(with_fc4=True)
    n = 10
    qm = mesh.make_mesh(m.crystal, (n, n, n))
    ph = phonons.solve(m.crystal, m.fc2, qm.q_cart, qm.q_int, args.T)
    ctx = lib.GpuContext(m, qm, ph)
    nd = m.crystal.ndof
    quartets = qm.nq0 * qm.nq * qm.nq * nd ** 4
    res = {}
    for prec in ("fp64", "fp32"):
        ctx.lw4_all(args.sigma, prec)
        ctx.kernel_time_reset()
        t = time.perf_counter()
        lw = ctx.lw4_all(args.sigma, prec)
        dt = time.perf_counter() - t
        nl, kms = ctx.kernel_time()
        res[prec] = dict(step=dt, kern=kms / max(nl, 1), lw=lw)
    l64, l32 = res["fp64"]["lw"], res["fp32"]["lw"]
    sig = np.abs(l64) > 1e-6 * np.abs(l64).max()
    # algorithmic FLOP per (q0, q1, q2) triple: phase sums 4*81*sum Natm4^3 over the nonzero
    # clusters, four complex mode products 32*Ndof^5, ~30 per branch quartet
    nclus = int(np.count_nonzero(np.abs(m.fc4.fc).sum(axis=(4, 5, 6, 7))))
    flop = 4 * 81 * nclus + 32 * nd ** 5 + 30 * nd ** 4
    triples = qm.nq0 * qm.nq * qm.nq
    # the kernels evaluate one ordered triple per permutation class (q1 <= q2 <= q3)
    q12 = np.arange(qm.nq)
    classes = 0
    for g in qm.irr_idx:
        k0 = qm.q_int[int(g)].astype(np.int64)
        q3 = qm.index(-k0[None, None, :] - qm.q_int[q12][:, None, :].astype(np.int64)
                      - qm.q_int[q12][None, :, :].astype(np.int64))
        classes += int(np.count_nonzero((q12[:, None] <= q12[None, :]) & (q12[None, :] <= q3)))
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        inp = O.Inputs(m, qm, ph, which=4)   # the FC4 arrays (orc_lw4_gauss reads fC(3,3,3,3,...))
        threads, _, _, _ = _host_cpu()
        from concurrent.futures import ThreadPoolExecutor
        i0 = int(qm.irr_idx[min(1, qm.nq0 - 1)])
        per = 8                                        # q1 per thread, every q2
        t = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda k: O.lw4_gauss(inp, i0, args.sigma, k * per, (k + 1) * per), range(threads)))
        dt = time.perf_counter() - t
        cpu = {"value": threads * per * qm.nq * nd ** 4 / dt, "unit": "quartets/s", "cores": threads,
               "kind": "port", "sample": f"one irreducible q0, {threads * per} q1 x all {qm.nq} q2, FP64 C "
                                         f"restatement, {threads} threads, {dt:.1f} s"}
    out = {
        "metric": "four-phonon |V4|^2 quartets/s (Gaussian delta)", "unit": "quartets/s",
        "value": quartets / res["fp64"]["step"], "dtype": "f64",
        "config": {"workload": f"4-phonon Gaussian linewidths, 2-atom rocksalt with FC4 (1st shell), {n}^3 mesh, "
                               f"{qm.nq0} irreducible q0 x {qm.nq}^2 (q1, q2) x {nd}^4 branches",
                   "sigma_rad_per_ps": args.sigma, "T_K": args.T, "flop_per_triple": flop,
                   "fc4_clusters": nclus},
        "classes_executed": classes,
        "fp64": {"s": res["fp64"]["step"], "kernel_ms": res["fp64"]["kern"],
                 "effective_TFLOPs": flop * triples / (res["fp64"]["kern"] * 1e-3) / 1e12,
                 "executed_TFLOPs": flop * classes / (res["fp64"]["kern"] * 1e-3) / 1e12},
        "fp32": {"s": res["fp32"]["step"], "kernel_ms": res["fp32"]["kern"],
                 "speedup_vs_fp64": res["fp64"]["step"] / res["fp32"]["step"],
                 "linewidth_max_rel_err_vs_fp64": float(np.max(np.abs(l32 - l64)[sig] / np.abs(l64)[sig]))},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def fourph_tetra(args):
    """Four-phonon tetrahedron-method linewidths (TetrahedronIntegration4th) of the 2-atom model
    with FC4 on a --mesh4 mesh, FP64: every irreducible q0, W4 of all ordered (q1, q2) then the
    tetrahedron sums; the CPU leg runs the oracle on a sample of q1 for one q0."""
    from spring_amd import model, mesh, phonons, lib
    m = model.rocksalt2(with_fc4=True)
    n = args.mesh4
    qm = mesh.make_mesh(m.crystal, (n, n, n))
    ph = phonons.solve(m.crystal, m.fc2, qm.q_cart, qm.q_int, args.T)
    ctx = lib.GpuContext(m, qm, ph)
    nd = m.crystal.ndof
    for _ in range(max(0, args.warmup)):
        ctx.lw4_tetra(min(1, qm.nq0 - 1))
    steps = max(1, args.steps)
    ctx.kernel_time_reset()
    t = time.perf_counter()
    for _ in range(steps):
        lw = ctx.lw4_tetra_all()
    dt = (time.perf_counter() - t) / steps
    # one evaluation unit: (q0, q1, tetrahedron, s0 s1 s2 s3), three processes at four vertices
    units = qm.nq0 * qm.nq * 6 * qm.nq * nd ** 4
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        from concurrent.futures import ThreadPoolExecutor
        inp = O.Inputs(m, qm, ph, which=4)
        threads = max(1, min(16, os.cpu_count() or 1))
        i0 = int(qm.irr_idx[min(1, qm.nq0 - 1)])

        per = max(1, min(32, qm.nq // threads))       # q1 rows per thread (about 6 s at 8^3)

        def part(k):
            parts = []
            for q1 in range(k * per, (k + 1) * per):
                W = O.w4_rows(inp, i0, q1, q1 + 1)
                parts.append(O.lw4_tetra_from_w(inp.mesh, inp.q_int, qm.tetra, i0, q1, W, inp.omega, inp.nbe))
            return parts

        t = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(part, range(threads)))
        dtc = time.perf_counter() - t
        cpu = {"value": threads * per * 6 * qm.nq * nd ** 4 / dtc, "unit": "units/s", "cores": threads,
               "kind": "port", "sample": f"one irreducible q0, {threads * per} q1 ({per} per thread) x all "
                                         f"{6 * qm.nq} tetrahedra, W4 rows and tetrahedron sums, FP64 C "
                                         f"restatement, {dtc:.1f} s"}
    out = {
        "metric": "four-phonon tetrahedron units/s (q0, q1, tetrahedron, branch quartet)", "unit": "units/s",
        "value": units / dt, "higher_is_better": True, "steps": steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"4-phonon tetrahedron linewidths, 2-atom rocksalt with FC4 (1st shell), "
                               f"{n}^3 mesh, {qm.nq0} irreducible q0 x {qm.nq} q1 x {6 * qm.nq} "
                               f"tetrahedra x {nd}^4 branches", "T_K": args.T},
        "lw_sum": float(lw.sum()),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def tetra16(args, cell10=False):
    """Three-phonon tetrahedron-method linewidths (CalculateLwTetrahedron, the reference's default
    delta) of every irreducible q0, FP64: W of every q1 for a batch of q0 (the small-cell kernel's
    W mode for the 2-atom model, the matrix-core kernel's for the 10-atom cell) then the
    tetrahedron sums; the CPU leg runs the oracle's tetrahedron sum on the same W for a few q0.
    cell10: the 10-atom cell on its --mesh10 mesh with time reversal, else the 2-atom model on
    --mesh."""
    from spring_amd import model, mesh, phonons, lib
    m = model.rocksalt10() if cell10 else model.rocksalt2()
    n = args.mesh10 if cell10 else args.mesh
    qm = mesh.make_mesh(m.crystal, (n, n, n), time_rev=1 if cell10 else 0)
    ph = phonons.solve(m.crystal, m.fc2, qm.q_cart, qm.q_int, args.T)
    ctx = lib.GpuContext(m, qm, ph)
    nd = m.crystal.ndof
    for _ in range(max(0, args.warmup)):
        ctx.lw3_tetra(min(1, qm.nq0 - 1))
    steps = max(1, args.steps)
    ctx.kernel_time_reset()
    t = time.perf_counter()
    for _ in range(steps):
        lw, p3 = ctx.lw3_tetra_all()
    dt = (time.perf_counter() - t) / steps
    nl, kms = ctx.kernel_time()
    # one unit: (q0, tetrahedron, s0 s1 s2), three processes at four vertices
    units = qm.nq0 * 6 * qm.nq * nd ** 3
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        inp = O.Inputs(m, qm, ph)
        k = min(1 if cell10 else 4, qm.nq0)
        # W of the sampled q0: the oracle's own (2-atom), or the device's (10-atom: 8000 q1 x 30^3)
        Ws = [(int(qm.irr_idx[i]), ctx.v3_w(i) if cell10 else O.lw3_gauss(inp, int(qm.irr_idx[i]), 1.0, want_w=True)[1])
              for i in range(k)]
        t = time.perf_counter()
        for g, W in Ws:
            O.lw3_tetra(qm, ph, g, W)
        dtc = time.perf_counter() - t
        cpu = {"value": k * 6 * qm.nq * nd ** 3 / dtc, "unit": "units/s", "cores": 1, "kind": "port",
               "sample": f"tetrahedron sums of {k} irreducible q0 from precomputed W, FP64 C restatement, "
                         f"1 thread, {dtc:.1f} s"}
    out = {
        "metric": "three-phonon tetrahedron units/s (q0, tetrahedron, branch triplet)", "unit": "units/s",
        "value": units / dt, "higher_is_better": True, "steps": steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3-phonon tetrahedron linewidths, {'10-atom cell' if cell10 else '2-atom rocksalt'}, "
                               f"{n}^3 mesh, {qm.nq0} irreducible q0 x {6 * qm.nq} tetrahedra x {nd}^3 branches",
                   "T_K": args.T},
        "tetra_kernel_ms_per_q0": kms / max(nl, 1),
        "lw_sum": float(lw.sum()), "p3_sum": float(p3.sum()),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bte16(args, cell10=False):
    """Iterative BTE (kappa_iter.x) of the 2-atom model (or, with cell10, the 10-atom cell): W of
    every time-reversal representative stored once on the device, then 1/tau and the
    self-consistent F updates; the CPU leg times the oracle's F update on a sample of
    representatives.  For the 10-atom cell the W store runs on the matrix cores; the line also
    times the generic-kernel store of the same W (SPR_KERNEL=generic) in a second context."""
    from spring_amd import model, mesh, phonons, lib
    m = model.rocksalt10() if cell10 else model.rocksalt2()
    n = args.mesh_bte10 if cell10 else args.mesh
    qm = mesh.make_mesh(m.crystal, (n, n, n))
    ph = phonons.solve(m.crystal, m.fc2, qm.q_cart, qm.q_int, args.T)
    nd = m.crystal.ndof
    t_generic = None
    if cell10:
        os.environ["SPR_KERNEL"] = "generic"
        try:
            cg = lib.GpuContext(m, qm, ph)
            t = time.perf_counter()
            cg.bte_irr()
            t_generic = time.perf_counter() - t
            cg.close()
        finally:
            del os.environ["SPR_KERNEL"]
    ctx = lib.GpuContext(m, qm, ph)
    t = time.perf_counter()
    irr, deg = ctx.bte_irr()                       # first call: W store
    t_setup = time.perf_counter() - t
    ctx.kernel_time_reset()
    t = time.perf_counter()
    itau = ctx.bte_tau_inv()
    t_tau = time.perf_counter() - t
    ctx.kernel_time_reset()
    t = time.perf_counter()
    out = ctx.bte_solve(args.T, eps_k=1e-5, max_iter=200)
    t_solve = time.perf_counter() - t
    nl, kms = ctx.kernel_time()
    # one evaluation = (representative q, tetrahedron, s, s1, s2): both processes, four vertices
    items = len(irr) * 6 * qm.nq * nd ** 3
    per_it_ms = kms / max(nl, 1)
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        from concurrent.futures import ThreadPoolExecutor
        inp = O.Inputs(m, qm, ph)
        oirr, map_irr, map_sign, odeg = O.tr_irreducible(qm)
        threads = max(1, min(16, os.cpu_count() or 1))
        per = 1 if cell10 else 8
        rng = np.random.default_rng(1)
        sample = rng.choice(len(oirr), size=min(len(oirr), threads * per), replace=False)
        with ThreadPoolExecutor(threads) as ex:      # W of the sample (not timed)
            W = np.array(list(ex.map(lambda k: O.lw3_gauss(inp, int(oirr[k]), 1.0, want_w=True)[1], sample)))
        Wp, Wm = O.bte_w_pm(qm, ph, oirr[sample], W)
        F = out["F"]

        def one(j):
            sl = slice(j * per, (j + 1) * per)
            return O.bte_sums(qm, ph, oirr[sample[sl]], map_irr, map_sign, Wp[sl], Wm[sl], F)

        nth = (len(sample) + per - 1) // per
        t = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            res = np.concatenate(list(ex.map(one, range(nth))))
        dt = time.perf_counter() - t
        gdF = ctx.bte_delta_f(F)[sample]
        agree = float(np.max(np.abs(res - gdF)) / np.max(np.abs(gdF)))
        cpu = {"value": len(sample) * 6 * qm.nq * nd ** 3 / dt, "unit": "evaluations/s", "cores": threads,
               "kind": "port", "sample": f"F update of {len(sample)} seeded time-reversal representatives (all "
                                         f"tetrahedra, all branch triplets), FP64 C restatement, {threads} "
                                         f"threads, {dt:.1f} s; GPU dF agrees to {agree:.1e} normwise"}
    R = len(irr) * nd
    a_bytes = R * R * 8
    if cpu is not None:
        # the reference recomputes the tetrahedron sums every iteration: its time to solution
        # from stored W is (1 + iterations) full passes over the evaluations
        cpu["time_to_solution_s_extrapolated"] = (1 + out["iterations"]) * items / cpu["value"]
    res = {
        "metric": "iterative BTE kappa time-to-solution from stored W (collision-matrix build + iterations)",
        "value": t_tau + t_solve, "unit": "s", "higher_is_better": False, "dtype": "f64",
        "config": {"workload": f"iterative BTE, {'10-atom cell' if cell10 else '2-atom rocksalt'}, {n}^3 "
                               f"mesh, {len(irr)} time-reversal representatives x {6 * qm.nq} tetrahedra x "
                               f"{nd}^3 branches",
                   "T_K": args.T, "eps_k_W_mK": 1e-5},
        "W_store_s": t_setup, "W_store_GB": len(irr) * qm.nq * nd ** 3 * 8 / 1e9,
        "W_store_generic_s": t_generic,
        "build_s": t_tau, "evaluations_in_build": items, "solve_s": t_solve, "iterations": out["iterations"],
        "F_update_ms": per_it_ms,
        "roofline_F_update": {"bound": "hbm", "achieved": a_bytes / (per_it_ms * 1e-3) / 1e9, "peak": 8000.0,
                              "unit": "GB/s", "frac": a_bytes / (per_it_ms * 1e-3) / 8e12,
                              "traffic": None, "bytes": a_bytes},
        "kappa_rta_W_mK": np.diag(out["kappa0"]).tolist(), "kappa_iter_W_mK": np.diag(out["kappa"]).tolist(),
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


def phonons10(args):
    """Harmonic phonons of every q of the 10-atom cell's 20^3 mesh (set_Phon_mp / get_phonon):
    spr_get_phonon on the device against the host LAPACK path (one zheevd-type eigh per q,
    batched over 16 threads) and the oracle's per-q restatement on a sample."""
    from spring_amd import model, mesh, phonons, lib
    m = model.rocksalt10()
    n = args.mesh10
    qm = mesh.make_mesh(m.crystal, (n, n, n), time_rev=1)
    nd = m.crystal.ndof
    for _ in range(args.warmup):
        lib.get_phonon(m.crystal, m.fc2, qm.q_cart, qm.q_int, args.T)
    t = time.perf_counter()
    for _ in range(args.steps):
        g = lib.get_phonon(m.crystal, m.fc2, qm.q_cart, qm.q_int, args.T)
    dt = (time.perf_counter() - t) / args.steps
    threads = max(1, min(16, os.cpu_count() or 1))
    t = time.perf_counter()
    h = phonons.solve(m.crystal, m.fc2, qm.q_cart, qm.q_int, args.T, chunk=256, threads=threads)
    dt_host = time.perf_counter() - t
    lam_g = np.sign(g.omega) * g.omega ** 2
    lam_h = np.sign(h.omega) * h.omega ** 2
    agree = float(np.abs(lam_g - lam_h).max() / np.abs(lam_h).max())
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        k = min(qm.nq, 2000)
        t = time.perf_counter()
        O.phonons(m.crystal, m.fc2, qm.q_cart[:k], qm.q_int[:k], args.T)
        dtc = time.perf_counter() - t
        cpu = {"value": k / dtc, "unit": "q-points/s", "cores": 1, "kind": "port",
               "sample": f"first {k} q of the mesh through the oracle's per-q restatement (numpy LAPACK "
                         f"eigh per q), 1 thread, {dtc:.1f} s"}
    res = {
        "metric": "harmonic phonons (D(q), eigenvectors, omega, n, group velocities) per second",
        "value": qm.nq / dt, "unit": "q-points/s", "higher_is_better": True, "dtype": "f64",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
        "config": {"workload": f"10-atom cell, {n}^3 mesh ({qm.nq} q, Ndof {nd}), host buffers in and out "
                               f"(the call includes the upload of FC2 and q and the download of "
                               f"{qm.nq * nd * nd * 16 / 1e6:.0f} MB of eigenvectors)", "T_K": args.T},
        "host_lapack_16_threads": {"value": qm.nq / dt_host, "unit": "q-points/s", "cores": threads,
                                   "eigenvalues_agree": agree},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
