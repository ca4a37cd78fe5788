"""Benchmark: one warm KIFMM3d evaluate on one MI355X.

Headline workload: 8,388,608 nonuniform points (90% from 16 Gaussians, 10%
uniform), Laplace single layer, np 6, ptsMax 128, maxLevel 14; error against
the FP64 direct sum at 4,096 sampled targets.  Points, tree, lists, operators
and densities are resident in HBM before the timed region; one step is one
evaluate (device pointers in, device pointers out).  The mixed-precision and
the plain FP64 paths are both timed, and the line carries a roofline fraction
for every phase (reference PETSc event names) plus the whole evaluation.

Beside the headline (extra keys of the same JSON line): the same workload at
np 7, the smallest np whose discretisation reaches 1e-6; the 32,768-point
uniform case (FP64); the Stokes double layer on 32 spheres at 1,048,576 points
(FP64 and mixed); the 1,048,576-point gradient workload (forward, transposed
and field passes).

Prints one JSON line.  --gpus N with N > 1 exits: this build runs on one GPU.
"""
import argparse
import gc
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fmm-g_amd"))

METRIC = "Mpoints/s per FMM eval at ≤1e-6 rel err, mixed vs FP64 GPU; % roofline"
HBM_GBS = 8000.0
PEAK_TF = {"fp64": 78.6, "fp32": 157.3}

# name: kernel, coefs, np, ptsMax, maxLevel, generator, n, config index (seed stream)
CONFIGS = {
    "c1": (111, [1.0, 0.25], 4, 64, 30, "uniform", 32768, 0),
    "c2": (321, [1.0, 0.25], 6, 64, 30, "spheres", 1048576, 1),
    "c3": (111, [1.0, 0.25], 6, 128, 14, "gaussians", 8388608, 2),
    "c4": (111, [1.0, 0.25], 6, 64, 30, "uniform", 1048576, 3),
}
WORKLOAD_NAME = {
    "c1": "Laplace-S uniform N=32768 np=4 ptsMax=64",
    "c2": "Stokes-D (321) on 32 spheres N=1048576 np=6 ptsMax=64",
    "c3": "Laplace-S nonuniform (16 Gaussians + 10% uniform) N=8388608 np=6 ptsMax=128 maxLevel=14",
    "c4": "Laplace-S uniform N=1048576 np=6 ptsMax=64 (forward, transposed and field passes)",
}
# flops per (target, source) pair of the user kernel: Laplace single layer 11
# (3 sub, 5 for r^2, 1 rsqrt, 2 FMA); Stokes double layer 28 (r^2, rsqrt, r^-5,
# r.n, then 9 products and FMAs into 3 outputs per 3-vector density)
PAIR_FLOPS = {111: 11.0, 321: 28.0, 511: 27.0}
MM = {111: (111, 1, 1), 321: (301, 4, 3), 511: (511, 3, 3)}


def make_points(gen, n, cfg_index):
    from kifmm import workloads as W
    if gen == "uniform":
        return W.uniform(n, W.seed(cfg_index, 0)), None
    if gen == "gaussians":
        return W.gaussians(n, W.seed(cfg_index, 0)), None
    if gen == "spheres":
        return W.spheres(n, W.seed(cfg_index, 0), nsph=32)
    raise ValueError(gen)


def traffic_bytes(kernel):
    """HBM bytes per launch of `kernel` on the 8M-point workload from the
    committed rocprofv3 counter summary (profiles/hadamard_c3_traffic.json:
    FETCH_SIZE and WRITE_SIZE in KiB per dispatch, scaled by the same counters
    read on a 512 MiB device copy); None when no summary covers the kernel."""
    path = os.path.join(ROOT, "profiles", "hadamard_c3_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    k = d.get("kernels", {}).get(kernel)
    if not k:
        return None
    return int((k["fetch_kib"] * d["fetch_scale"] + k["write_kib"] * d["write_scale"]) * 1024)


def host_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "nproc_available": avail}


def cpu_baseline(cfg, pinv, n_sample):
    """The CPU oracle (reference algorithm restated in C, 1 thread) on a bounded
    sample of the same distribution: the tree and operators are built first,
    one cold evaluate, then the timed warm second evaluate (tt0.cpp:501-524)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from kifmm import workloads as W
    kt, coefs, np_, pm, ml, gen, _, ci = cfg
    pos, nor = make_points(gen, n_sample, ci)
    nor = np.zeros_like(pos) if nor is None else nor
    den = W.densities(n_sample, 1, W.seed(ci, 1))
    pu, pd = pinv
    ops = O.Ops(O.mm_kernel(kt), coefs, np_, pu, pd)
    tree = O.Tree(pos, pts_max=pm, max_level=ml)
    t0 = time.perf_counter()
    O.evaluate(tree, ops, kt, coefs, pos, nor, pos, den)
    cold = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.evaluate(tree, ops, kt, coefs, pos, nor, pos, den)
    warm = time.perf_counter() - t0
    return n_sample / warm / 1e6, warm, cold


def spectrum_len(np_, mixed):
    """Frequencies per V-list spectrum: the r2c half spectrum (np+1)(2np)^2 in
    FP64; in the mixed path the packed form, which drops the conjugate mirrors
    in the planes kc = 0 and kc = np, 4 np^3 + 4."""
    return 4 * np_ ** 3 + 4 if mixed else (np_ + 1) * (2 * np_) ** 2


def phase_roofline(kt, np_, w, phases, mixed, n_src, n_trg):
    """Per-phase roofline: T_roof = max(F / peak, B / HBM) from the work counts of
    the built tree, fraction = T_roof / T_measured, and sum(T_roof) / T_eval.
    F and B are algorithmic: pair flops at c_K per pair (S2T, S2M, S2L, M2T,
    L2T); 2 M K per GEMM column; 2.5 G log2 G per FFT line set of a box and
    DOF; 8 flops per complex multiply-add of the Hadamard.  B counts every
    operand once.  Peaks: FP32 157.3 TFLOP/s (mixed near field, surfaces,
    FFTs, Hadamard), FP64 78.6 TFLOP/s (GEMMs; everything in the FP64 path),
    HBM 8 TB/s."""
    km, sd, td = MM[kt]
    SD = {111: 1, 321: 3, 511: 3}[kt]
    TD = {111: 1, 321: 3, 511: 3}[kt]
    cK = PAIR_FLOPS[kt]
    nue = np_ ** 3 - (np_ - 2) ** 3
    nuc = (np_ + 2) ** 3 - np_ ** 3
    ue, uc, de, dc = nue * sd, nuc * td, nue * sd, nue * td
    G = (2 * np_) ** 3
    H = spectrum_len(np_, mixed)
    vec = PEAK_TF["fp32" if mixed else "fp64"] * 1e12
    f64 = PEAK_TF["fp64"] * 1e12
    bw = HBM_GBS * 1e9
    cw = 8 if mixed else 16  # bytes per complex spectrum entry

    def troof(parts):
        return sum(max(F / P, B / bw) for F, P, B in parts)

    parents, contrib = w["m2m_cols"], w["up_cols"]
    leaves = contrib - parents
    if mixed:
        up_gemm = [(2.0 * ue * 8 * ue * parents, f64, 8.0 * (ue * 8 * ue + parents * (8 * ue + ue))),
                   (2.0 * ue * uc * leaves, f64, 8.0 * (ue * uc + leaves * (uc + ue)))]
    else:
        up_gemm = [(2.0 * uc * 8 * ue * parents, f64, 8.0 * (uc * 8 * ue + parents * (8 * ue + uc))),
                   (2.0 * ue * uc * contrib, f64, 8.0 * (ue * uc + contrib * (uc + ue)))]
    s2m = [(cK * w["s2m_pairs"], vec, 8.0 * (n_src * (3 + SD)) + 8.0 * leaves * uc)]
    fft = [(2.5 * G * math.log2(G) * (sd * w["n_vsrc"] + td * w["n_vtrg"]), vec,
            8.0 * ue * w["n_vsrc"] + cw * H * (sd * w["n_vsrc"] + td * w["n_vtrg"]) + 16.0 * dc * w["n_vtrg"])]
    had = [(8.0 * H * sd * td * w["vlist_pairs"], vec,
            cw * H * (sd * w["n_vsrc"] + td * w["n_vtrg"]) + cw * 316 * sd * td * H)]
    s2l = [(cK * w["s2l_pairs"], vec, 8.0 * n_src * (3 + SD))]
    down = [(2.0 * dc * de * w["l2l_cols"], f64, 8.0 * (8 * dc * de + w["l2l_cols"] * (de + dc))),
            (2.0 * de * dc * w["dn_cols"], f64, 8.0 * (de * dc + w["dn_cols"] * (dc + de)))]
    near = [(cK * w["near_pairs"], vec, 8.0 * (n_src * (3 + SD) + n_trg * (3 + TD)))]
    # density load into leaf order and the write-back to the caller's order:
    # each value read once and written once
    ini = [(0.0, vec, 16.0 * n_src * SD)]
    # (the mixed Laplace near kernel writes the caller's order itself)
    fin = [] if (mixed and kt == 111) else [(0.0, vec, 16.0 * n_trg * TD)]
    out = {}
    rows = {"EvlIni": ini, "EvlUpwComp": s2m + up_gemm, "EvlVList": fft + had, "EvlXList": s2l,
            "EvlCombine": down, "EvlUList": near, "EvlFinalize": fin}
    total_roof = 0.0
    for name, parts in rows.items():
        t = troof(parts)
        total_roof += t
        meas = phases.get(name, 0.0) * 1e-3
        out[name] = {"t_roof_ms": round(t * 1e3, 4), "t_ms": round(meas * 1e3, 4),
                     "frac": round(t / meas, 4) if meas > 0 else None}
    out["VListHadamard"] = {"t_roof_ms": round(troof(had) * 1e3, 4),
                            "t_ms": round(phases.get("VListHadamard", 0.0), 4),
                            "frac": round(troof(had) / (phases["VListHadamard"] * 1e-3), 4)
                            if phases.get("VListHadamard") else None}
    return out, total_roof


class Runner:
    def __init__(self, dev, steps, warmup, numchk, autotune=True):
        self.dev, self.steps, self.warmup, self.numchk = dev, steps, warmup, numchk
        self.autotune = autotune

    def inputs(self, cfg, n):
        from kifmm import workloads as W
        import torch
        kt, coefs, np_, pm, ml, gen, _, ci = cfg
        pos, nor = make_points(gen, n, ci)
        sd = {111: 1, 321: 3, 511: 3}[kt]
        den = W.densities(n, sd, W.seed(ci, 1), kind="uniform" if sd == 1 else "normal")
        rng = np.random.default_rng(W.seed(ci, 3))
        idx = rng.choice(n, size=min(self.numchk, n), replace=False)
        return pos, nor, den, torch.from_numpy(den).to(self.dev), idx

    def run(self, cfg, n, precision, pos, nor, d_den, idx, steps=None, warmup=None, np_override=None, tag=""):
        """Setup, warm-up, K timed evaluates (plain stream, then a graph replay
        on a stream of its own), a few evaluates with phase timing, error vs
        the FP64 direct sum at the sampled targets."""
        import torch
        import kifmm
        kt, coefs, np_, pm, ml, _, _, _ = cfg
        if np_override:
            np_ = np_override
        steps = steps or self.steps
        warmup = self.warmup if warmup is None else warmup
        t0 = time.perf_counter()
        f = kifmm.FMM3d(kt, coefs, np_, pm, ml, precision=precision)
        f.set_points(pos, nor)
        f.setup()
        setup_s = time.perf_counter() - t0
        setup_parts = f.setup_times()
        d_val = torch.empty(n * f.trg_dof, dtype=torch.float64, device=self.dev)
        tuning = None
        if precision == "mixed" and self.autotune:
            # launch shapes chosen on this tree (untimed, like the setup)
            t0 = time.perf_counter()
            tuning = f.autotune(d_den, d_val)
            tuning["seconds"] = round(time.perf_counter() - t0, 3)
        torch.cuda.synchronize()
        for _ in range(warmup):
            f.evaluate_device(d_den, d_val)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            f.evaluate_device(d_den, d_val)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        # the same pass replayed from a captured hipGraph on a stream of its own;
        # the first call captures and launches, then the same number of untimed
        # replays as the direct loop's warmup (the kernels run a few percent
        # slower in the first replays after the capture pause)
        st = torch.cuda.Stream(device=self.dev)
        f.set_graph(True)
        with torch.cuda.stream(st):
            for _ in range(1 + warmup):
                f.evaluate_device(d_den, d_val, stream=st.cuda_stream)
            st.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                f.evaluate_device(d_den, d_val, stream=st.cuda_stream)
            st.synchronize()
            dt_graph = (time.perf_counter() - t0) / steps
        f.set_graph(False)
        torch.cuda.synchronize()
        # phase breakdown: HIP events on the evaluate stream (timing mode
        # synchronises after each evaluate)
        f.set_timing(True)
        nph = max(1, min(steps, 5))
        phase_sum = {}
        for _ in range(nph):
            f.evaluate_device(d_den, d_val)
            for k, v in f.phase_times().items():
                phase_sum[k] = phase_sum.get(k, 0.0) + v
        f.set_timing(False)
        phases = {k: v / nph for k, v in phase_sum.items()}
        torch.cuda.synchronize()
        d_chk = torch.from_numpy(np.ascontiguousarray(pos[idx])).to(self.dev)
        d_ref = torch.empty(len(idx) * f.trg_dof, dtype=torch.float64, device=self.dev)
        t0 = time.perf_counter()
        f.direct_device(d_den, d_chk, d_ref)
        torch.cuda.synchronize()
        direct_ms = (time.perf_counter() - t0) * 1e3
        ref = d_ref.cpu().numpy().reshape(len(idx), -1)
        val = d_val.cpu().numpy()
        td = f.trg_dof
        got = val.reshape(-1, td)[idx]
        rel_err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        print("[%s %s np=%d] setup %.2fs  ms/step %.3f (graph %.3f)  rel_err %.3e  direct %.1f ms  phases %s"
              % (tag, precision, np_, setup_s, dt * 1e3, dt_graph * 1e3, rel_err, direct_ms,
                 json.dumps({k: round(v, 3) for k, v in phases.items()})), file=sys.stderr, flush=True)
        return {"f": f, "dt": dt, "dt_graph": dt_graph, "phases": phases, "rel_err": rel_err, "val": val,
                "setup_s": setup_s, "setup_parts": setup_parts, "direct_ms": direct_ms, "tuning": tuning}


def pair_summary(mx, f64, n):
    d = float(np.linalg.norm(mx["val"] - f64["val"]) / np.linalg.norm(f64["val"]))
    return {"mixed_ms": round(mx["dt"] * 1e3, 4), "fp64_ms": round(f64["dt"] * 1e3, 4),
            "mixed_graph_ms": round(mx["dt_graph"] * 1e3, 4),
            "mixed_mpoints_s": round(n / mx["dt"] / 1e6, 3), "fp64_mpoints_s": round(n / f64["dt"] / 1e6, 3),
            "speedup_vs_fp64": round(f64["dt"] / mx["dt"], 3),
            "rel_err_mixed": mx["rel_err"], "rel_err_fp64": f64["rel_err"], "rel_diff_vs_fp64": d}


def release(*rs):
    for r in rs:
        if r and "f" in r:
            r["f"].close()
            r.pop("f", None)
    gc.collect()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-extra", action="store_true", help="headline workload only")
    ap.add_argument("--n", type=int, default=0, help="override the point count (diagnostics only)")
    ap.add_argument("--cpu-sample", type=int, default=262144)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--numchk", type=int, default=4096)
    ap.add_argument("--only", choices=["c2", "c4", "c5"], default=None,
                    help="run one secondary workload alone and print its JSON (profiling runs)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time the mixed path with its default launch shapes (profiling runs)")
    ap.add_argument("--c5-problems", type=int, default=512,
                    help="problems of the batched Navier workload (the first k of the 512)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="copy 512 MiB device-to-device first (calibrates FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU (multi-GPU sharding is a design, "
              "not a build); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    dev = torch.device("cuda:0")
    if args.calib_copy:
        src = torch.ones(64 * 2 ** 20, dtype=torch.float64, device=dev)  # 512 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    R = Runner(dev, args.steps, args.warmup, args.numchk, autotune=not args.no_autotune)
    if args.only:
        # one secondary workload alone (profiling runs); prints its JSON object
        if args.only == "c2":
            res = bench_spheres(R, min(args.steps, 5))
        elif args.only == "c4":
            res = bench_gradient(R, min(args.steps, 5))
        else:
            res = bench_batched(R, args.c5_problems)
        print(json.dumps({args.only: res}), flush=True)
        return

    # ---- headline: the 8M-point clustered Laplace workload, np 6 ----
    cfg = CONFIGS["c3"]
    kt, coefs, np_, pm, ml, gen, n, ci = cfg
    if args.n:
        n = args.n
    pos, nor, den, d_den, idx = R.inputs(cfg, n)
    r64 = None
    if args.precision == "mixed" and not args.no_fp64:
        r64 = R.run(cfg, n, "fp64", pos, nor, d_den, idx, tag="c3")
        r64["f"].close()
    r = R.run(cfg, n, args.precision, pos, nor, d_den, idx, tag="c3")
    f = r["f"]
    mixed = args.precision == "mixed"
    w = f.work_counts()
    phases = r["phases"]
    per_phase, total_roof = phase_roofline(kt, np_, w, phases, mixed, n, n)

    # roofline of the dominant kernel, the V-list Hadamard (k_m2l_tile in the
    # mixed path, k_hadamard in FP64).  Algorithmic work: one complex
    # multiply-add (8 flops) per V-list (target, source) pair, frequency and
    # sd*td component.  Compulsory bytes: every source and target spectrum once
    # and the 316 translation spectra, (n_vsrc*sd + n_vtrg*td)*H*2w +
    # 316*sd*td*H*2w with w = 4 (FP32) or 8 (FP64).  The intensity (about 76
    # flop/byte in FP32) is far above the machine balance (157.3 TFLOP/s over
    # 8 TB/s, about 20): the kernel is bound by its vector (VALU) issue, and
    # the chip's FP32 peak is the roof.
    had_ms = phases["VListHadamard"]
    wbytes = 4 if mixed else 8
    H = spectrum_len(np_, mixed)
    alg_bytes = (w["n_vsrc"] + w["n_vtrg"]) * H * 2 * wbytes + 316 * H * 2 * wbytes
    alg_flops = 8.0 * w["vlist_pairs"] * H
    had_tf = alg_flops / (had_ms * 1e-3) / 1e12
    peak = PEAK_TF["fp32" if mixed else "fp64"]
    kname = "k_m2l_tile" if mixed else "k_hadamard"
    roofline = {"bound": "valu", "kernel": kname, "achieved": round(had_tf, 3), "peak": peak,
                "unit": "TFLOP/s", "frac": round(had_tf / peak, 4), "traffic": traffic_bytes(kname),
                "kernel_ms": round(had_ms, 4), "algorithmic_flops": alg_flops, "algorithmic_bytes": alg_bytes,
                "intensity_flop_per_byte": round(alg_flops / alg_bytes, 2),
                "hbm_gbs_at_algorithmic_bytes": round(alg_bytes / (had_ms * 1e-3) / 1e9, 1)}
    near_ms = phases["EvlUList"]
    near_tf = w["near_pairs"] * 11.0 / (near_ms * 1e-3) / 1e12
    near = {"bound": "valu", "kernel": "k_near_lap_mx" if mixed else "k_near", "achieved": round(near_tf, 3),
            "peak": peak, "unit": "TFLOP/s", "frac": round(near_tf / peak, 4), "kernel_ms": round(near_ms, 4)}

    cpu = None
    if not args.no_cpu:
        nue = np_ ** 3 - (np_ - 2) ** 3
        nuc = (np_ + 2) ** 3 - np_ ** 3
        pinv = (f.operator(0).reshape(nuc, nue).T, f.operator(1).reshape(nue, nue).T)
        sample = min(args.cpu_sample, n)
        mps, warm, cold = cpu_baseline(cfg, pinv, sample)
        hi = host_info()
        cpu = {"value": round(mps, 5), "unit": "Mpoints/s", "cores": 1, "kind": "port",
               "sample": "%d points of the same distribution (np 6, ptsMax 128), warm second evaluate %.1f s "
                         "(cold first %.1f s), 1 thread" % (sample, warm, cold),
               "cpu_model": hi["cpu_model"], "nproc": hi["nproc"], "nproc_available": hi["nproc_available"]}
    val = r["val"]
    diff_fp64 = float(np.linalg.norm(val - r64["val"]) / np.linalg.norm(r64["val"])) if r64 else None
    dt, dt64 = r["dt"], (r64["dt"] if r64 else None)
    release(r)

    out = {
        "metric": METRIC,
        "value": round(n / dt / 1e6, 3),
        "unit": "Mpoints/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": WORKLOAD_NAME["c3"] if not args.n else "c3 (n=%d)" % n,
                   "kernel": kt, "np": np_, "ptsMax": pm, "maxLevel": ml, "precision": args.precision,
                   "n_points": n},
        "rel_err": r["rel_err"],
        "graph_ms_per_step": round(r["dt_graph"] * 1e3, 4),
        "fp64_ms_per_step": round(dt64 * 1e3, 4) if dt64 else None,
        "speedup_vs_fp64": round(dt64 / dt, 3) if dt64 else None,
        "rel_diff_vs_fp64": diff_fp64,
        "setup_s": round(r["setup_s"], 3),
        "setup_parts_ms": {k: round(v, 2) for k, v in r["setup_parts"].items()},
        "tuning": r["tuning"],
        "direct_sum_ms": round(r["direct_ms"], 2),
        "phases_ms": {k: round(v, 4) for k, v in phases.items()},
        "roofline": roofline,
        "phase_roofline": per_phase,
        "eval_roofline_frac": round(total_roof / (sum(phases[k] for k in phases if k != "VListHadamard") * 1e-3), 4),
        "near_field": near,
        "cpu_baseline": cpu,
    }

    if not args.no_extra and not args.n:
        xs = min(args.steps, 5)
        extra = {}
        # the same workload at np 7: the smallest np that reaches 1e-6
        r7_64 = R.run(cfg, n, "fp64", pos, nor, d_den, idx, steps=xs, warmup=1, np_override=7, tag="c3")
        r7_64["f"].close()
        r7 = R.run(cfg, n, "mixed", pos, nor, d_den, idx, steps=xs, warmup=1, np_override=7, tag="c3")
        s = pair_summary(r7, r7_64, n)
        s["meets_1e-6"] = bool(r7["rel_err"] <= 1e-6)
        extra["c3_np7"] = s
        release(r7, r7_64)
        del pos, nor, den, d_den
        gc.collect()
        torch.cuda.empty_cache()
        # the 32,768-point uniform case, FP64 (the CPU-runnable case)
        c = CONFIGS["c1"]
        p1, n1_, dn1, dd1, i1 = R.inputs(c, c[6])
        a = R.run(c, c[6], "fp64", p1, n1_, dd1, i1, steps=xs, warmup=1, tag="c1")
        extra["c1"] = {"workload": WORKLOAD_NAME["c1"], "fp64_ms": round(a["dt"] * 1e3, 4),
                       "fp64_graph_ms": round(a["dt_graph"] * 1e3, 4),
                       "mpoints_s": round(c[6] / a["dt"] / 1e6, 3), "rel_err": a["rel_err"]}
        release(a)
        # Stokes double layer on 32 spheres, FP64 and mixed
        extra["c2"] = bench_spheres(R, xs)
        # the gradient workload: forward, transposed and field passes
        extra["c4"] = bench_gradient(R, xs)
        # the batched Navier workload: 512 problems x 8 right-hand sides
        extra["c5"] = bench_batched(R, args.c5_problems)
        out.update(extra)
    print(json.dumps(out), flush=True)


def bench_spheres(R, xs):
    """Stokes double layer on 32 spheres, 1,048,576 points, np 6: FP64 and mixed."""
    c = CONFIGS["c2"]
    p2, n2_, dn2, dd2, i2 = R.inputs(c, c[6])
    b64 = R.run(c, c[6], "fp64", p2, n2_, dd2, i2, steps=xs, warmup=1, tag="c2")
    b64["f"].close()
    b = R.run(c, c[6], "mixed", p2, n2_, dd2, i2, steps=xs, warmup=1, tag="c2")
    s = pair_summary(b, b64, c[6])
    s["workload"] = WORKLOAD_NAME["c2"]
    s["phases_ms_mixed"] = {k: round(v, 4) for k, v in b["phases"].items()}
    s["phases_ms_fp64"] = {k: round(v, 4) for k, v in b64["phases"].items()}
    wb = b["f"].work_counts()
    pr, tr = phase_roofline(321, 6, wb, b["phases"], True, c[6], c[6])
    s["phase_roofline_mixed"] = pr
    s["tuning"] = b["tuning"]
    s["eval_roofline_frac_mixed"] = round(
        tr / (sum(b["phases"][k] for k in b["phases"] if k != "VListHadamard") * 1e-3), 4)
    release(b, b64)
    return s


def bench_gradient(R, steps):
    import torch
    import kifmm
    from kifmm import workloads as W
    c = CONFIGS["c4"]
    kt, coefs, np_, pm, ml, gen, n, ci = c
    pos, _ = make_points(gen, n, ci)
    res = {"workload": WORKLOAD_NAME["c4"]}
    for prec in ("fp64", "mixed"):
        f = kifmm.FMM3d(kt, coefs, np_, pm, ml, precision=prec)
        f.set_points(pos)
        f.setup()
        sig = torch.from_numpy(W.densities(n, 1, W.seed(ci, 1))).to(R.dev)
        out = torch.empty(n, dtype=torch.float64, device=R.dev)
        adj = torch.empty(n, dtype=torch.float64, device=R.dev)
        grad = torch.empty(3 * n, dtype=torch.float64, device=R.dev)
        times = {}
        for name, fn in (("forward", lambda: f.evaluate_device(sig, out)),
                         ("adjoint", lambda: f.evaluate_adjoint_device(out, adj)),
                         ("field", lambda: f.field_device(sig, grad))):
            fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                fn()
            torch.cuda.synchronize()
            times[name] = round((time.perf_counter() - t0) / steps * 1e3, 4)
        res[prec + "_ms"] = times
        print("[c4 %s] %s" % (prec, json.dumps(times)), file=sys.stderr, flush=True)
        f.close()
    return res


C5_POINTS_PER_FOREST = 3_500_000


def c5_problems(nprob=512):
    """Point counts and points of the batched workload: N_p = round(exp(U[ln 2048,
    ln 262144])), uniform in [0,1)^3, one seeded stream per problem."""
    from kifmm import workloads as W
    rng = np.random.default_rng(W.seed(4, 0))
    sizes = np.rint(np.exp(rng.uniform(np.log(2048), np.log(262144), 512))).astype(np.int64)[:nprob]
    return sizes, [W.uniform(int(m), W.seed(4, 0) * 1000 + p) for p, m in enumerate(sizes)]


def bench_batched(R, nprob):
    """Navier single layer (mu 1, nu 0.3), np 8, ptsMax 120, maxLevel 10, 8
    right-hand sides of 3-vector N(0,1) densities per problem.  The problems go
    into forests of at most 3.5M points (the V-list spectra of 8 right-hand
    sides of all 512 problems, about 0.7 TB, do not fit in HBM at once); each
    forest is one handle whose every pass is one launch over all its problems
    and right-hand sides.  The time of one evaluation of the whole workload is
    the sum over the forests of a warm 8-RHS evaluate (mixed path, after the
    handle's autotune); the FP64 path is timed on the first forest.  Error:
    FP64 direct sum at 256 targets of the first problem, first right-hand
    side."""
    import torch
    import kifmm
    from kifmm import workloads as W
    sizes, pts = c5_problems(nprob)
    forests, cur, acc = [], [], 0
    for p, m in enumerate(sizes):
        if cur and acc + m > C5_POINTS_PER_FOREST:
            forests.append(cur)
            cur, acc = [], 0
        cur.append(p)
        acc += int(m)
    if cur:
        forests.append(cur)
    NRHS = 8
    res = {"workload": "Navier-S (511, nu 0.3) np=8 ptsMax=120 maxLevel=10, %d problems (N_p log-uniform in "
                       "[2048, 262144], total %d points) x %d RHS" % (len(sizes), int(sizes.sum()), NRHS),
           "n_points": int(sizes.sum()), "n_forests": len(forests), "nrhs": NRHS}
    t_mixed, t_setup, tune = 0.0, 0.0, None
    for fi, probs in enumerate(forests):
        for prec in (("mixed", "fp64") if fi == 0 else ("mixed",)):
            f = kifmm.FMM3d(511, [1.0, 0.3], 8, 120, 10, precision=prec)
            f.set_points_batch([pts[p] for p in probs])
            t0 = time.perf_counter()
            f.setup()
            ts = time.perf_counter() - t0
            n = f.n_src
            den = np.concatenate([W.densities(n, 3, W.seed(4, 1) * 100 + 10 * fi + r, kind="normal")
                                  for r in range(NRHS)])
            d_den = torch.from_numpy(den).to(R.dev)
            d_val = torch.empty(NRHS * n * 3, dtype=torch.float64, device=R.dev)
            if prec == "mixed" and R.autotune:
                tune = f.autotune(d_den, d_val, nrhs=NRHS)["tuning"] if fi == 0 else tune
                if fi > 0:
                    f.set_tuning(**tune)
            f.evaluate_multi_device(d_den, d_val, NRHS)  # warm
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            f.evaluate_multi_device(d_den, d_val, NRHS)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            if fi == 0:
                m0 = int(sizes[probs[0]])
                idx = np.random.default_rng(W.seed(4, 3)).choice(m0, 256, replace=False)
                g = kifmm.FMM3d(511, [1.0, 0.3], 8, 120, 10)
                g.set_points(pts[probs[0]])
                g.setup()
                d_chk = torch.from_numpy(np.ascontiguousarray(pts[probs[0]][idx])).to(R.dev)
                d_ref = torch.empty(256 * 3, dtype=torch.float64, device=R.dev)
                g.direct_device(d_den[:m0 * 3].contiguous(), d_chk, d_ref)
                ref = d_ref.cpu().numpy().reshape(256, 3)
                got = d_val[:m0 * 3].cpu().numpy().reshape(m0, 3)[idx]
                res["rel_err_%s" % prec] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
                g.close()
            if prec == "mixed":
                t_mixed += dt
                t_setup += ts
            else:
                res["fp64_first_forest_ms"] = round(dt * 1e3, 3)
                res["fp64_first_forest_setup_s"] = round(ts, 3)
            if fi == 0:
                res["%s_first_forest_ms" % prec] = round(dt * 1e3, 3)
                res["first_forest"] = {"problems": len(probs), "points": int(n)}
                f.set_timing(True)
                f.evaluate_multi_device(d_den, d_val, NRHS)
                ph = f.phase_times()
                res["%s_first_forest_phases_ms" % prec] = {k: round(v, 3) for k, v in ph.items()}
                f.set_timing(False)
                # Hadamard roofline: 8 flops per complex multiply-add, 9 components
                # (sd = td = 3), per V pair, frequency and RHS
                w = f.work_counts()
                flops = 8.0 * spectrum_len(8, prec == "mixed") * 9 * w["vlist_pairs"] * NRHS
                peak = PEAK_TF["fp32" if prec == "mixed" else "fp64"]
                tf = flops / (ph["VListHadamard"] * 1e-3) / 1e12
                res["%s_first_forest_hadamard" % prec] = {
                    "vlist_pairs": w["vlist_pairs"], "algorithmic_flops": flops, "achieved_tflops": round(tf, 3),
                    "peak_tflops": peak, "frac": round(tf / peak, 4)}
            print("[c5 forest %d/%d %s] %d problems %d points setup %.2fs evaluate(8 RHS) %.1f ms"
                  % (fi + 1, len(forests), prec, len(probs), n, ts, dt * 1e3), file=sys.stderr, flush=True)
            f.close()
            del d_den, d_val
            gc.collect()
            torch.cuda.empty_cache()
    res["mixed_ms"] = round(t_mixed * 1e3, 3)
    res["mixed_mpoints_s"] = round(res["n_points"] * NRHS / t_mixed / 1e6, 3)  # point-RHS per second
    res["setup_s"] = round(t_setup, 3)
    res["tuning"] = tune
    return res


if __name__ == "__main__":
    main()
