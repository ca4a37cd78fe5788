"""Benchmark: one warm KIFMM3d evaluate on one MI355X.

Workload (default): 8,388,608 nonuniform points (90% from 16 Gaussians, 10%
uniform), Laplace single layer, np 6, ptsMax 128, maxLevel 14; error against
the FP64 direct sum at 4,096 sampled targets.  Points, tree, lists, operators
and densities are resident in HBM before the timed region; one step is one
evaluate (device pointers in, device pointers out).

Prints one JSON line.  --gpus N with N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fmm-g_amd"))

METRIC = "Mpoints/s per FMM eval at ≤1e-6 rel err, mixed vs FP64 GPU; % roofline"
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}

CONFIGS = {
    # name: (kernel, coefs, np, ptsMax, maxLevel, generator, n)
    "c1": (111, [1.0, 0.25], 4, 64, 30, "uniform", 32768),
    "c3": (111, [1.0, 0.25], 6, 128, 14, "gaussians", 8388608),
}
WORKLOAD_NAME = {
    "c1": "Laplace-S uniform N=32768 np=4 ptsMax=64",
    "c3": "Laplace-S nonuniform (16 Gaussians + 10% uniform) N=8388608 np=6 ptsMax=128 maxLevel=14",
}


def make_points(gen, n, cfg_index):
    from kifmm import workloads as W
    if gen == "uniform":
        return W.uniform(n, W.seed(cfg_index, 0)), None
    if gen == "gaussians":
        return W.gaussians(n, W.seed(cfg_index, 0)), None
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


def cpu_baseline(cfg, precision_ops, n_sample, cfg_index):
    """The CPU oracle (reference algorithm restated in C, 1 thread) on a bounded
    sample of the same distribution; returns Mpoints/s and the sample size."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from kifmm import workloads as W
    kt, coefs, np_, pm, ml, gen, _ = cfg
    pos, _ = make_points(gen, n_sample, cfg_index)
    den = W.densities(n_sample, 1, W.seed(cfg_index, 1))
    pu, pd = precision_ops
    ops = O.Ops(O.mm_kernel(kt), coefs, np_, pu, pd)
    tree = O.Tree(pos, pts_max=pm, max_level=ml)
    t0 = time.perf_counter()
    O.evaluate(tree, ops, kt, coefs, pos, np.zeros_like(pos), pos, den)
    dt = time.perf_counter() - t0
    return n_sample / dt / 1e6, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--n", type=int, default=0, help="override the point count (diagnostics only)")
    ap.add_argument("--cpu-sample", type=int, default=131072)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--numchk", type=int, default=4096)
    ap.add_argument("--calib-copy", action="store_true",
                    help="copy 512 MiB device-to-device first (calibrates FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU (multi-GPU sharding is a design, "
              "not a build); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import kifmm
    from kifmm import workloads as W

    cfg = CONFIGS[args.config]
    cfg_index = {"c1": 0, "c3": 2}[args.config]
    kt, coefs, np_, pm, ml, gen, n = cfg
    if args.n:
        n = args.n
    dev = torch.device("cuda:0")
    pos, nor = make_points(gen, n, cfg_index)
    den = W.densities(n, 1, W.seed(cfg_index, 1))
    d_den = torch.from_numpy(den).to(dev)
    if args.calib_copy:
        src = torch.ones(64 * 2 ** 20, dtype=torch.float64, device=dev)  # 512 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    rng = np.random.default_rng(W.seed(cfg_index, 3))
    idx = rng.choice(n, size=min(args.numchk, n), replace=False)

    def run(precision):
        """Setup, warm-up, K timed evaluates; returns timings, phases and error."""
        t0 = time.perf_counter()
        f = kifmm.FMM3d(kt, coefs, np_, pm, ml, precision=precision)
        f.set_points(pos, nor)
        f.setup()
        print("[%s] setup %.2fs n=%d" % (precision, time.perf_counter() - t0, n), file=sys.stderr, flush=True)
        d_val = torch.empty(n * f.trg_dof, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        for _ in range(args.warmup):
            f.evaluate_device(d_den, d_val)
        torch.cuda.synchronize()
        f.set_timing(True)
        phase_sum = {}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            f.evaluate_device(d_den, d_val)
            for k, v in f.phase_times().items():
                phase_sum[k] = phase_sum.get(k, 0.0) + v
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        f.set_timing(False)
        phases = {k: v / args.steps for k, v in phase_sum.items()}
        # accuracy vs the FP64 direct sum at sampled targets (fmm3d_check_mpi.cpp:159-165)
        d_chk = torch.from_numpy(np.ascontiguousarray(pos[idx])).to(dev)
        d_ref = torch.empty(len(idx) * f.trg_dof, dtype=torch.float64, device=dev)
        f.direct_device(d_den, d_chk, d_ref)
        torch.cuda.synchronize()
        ref = d_ref.cpu().numpy()
        val = d_val.cpu().numpy()
        rel_err = float(np.linalg.norm(val[idx] - ref) / np.linalg.norm(ref))
        print("[%s] ms/step %.3f rel_err %.3e phases %s" % (precision, dt * 1e3, rel_err,
              json.dumps({k: round(v, 3) for k, v in phases.items()})), file=sys.stderr, flush=True)
        return f, dt, phases, rel_err, val

    f64 = dt64 = None
    if args.precision == "mixed" and not args.no_fp64:
        f64, dt64, ph64, err64, val64 = run("fp64")
        del f64
    f, dt, phases, rel_err, val = run(args.precision)
    diff_fp64 = None
    if dt64 is not None:
        diff_fp64 = float(np.linalg.norm(val - val64) / np.linalg.norm(val64))

    # roofline of the dominant kernel: the V-list Hadamard (k_m2l_tile in the
    # mixed path, k_hadamard in FP64).  Algorithmic work: one complex
    # multiply-add (8 flops) per V-list (target, source) pair, frequency and
    # sd*td component.  Compulsory bytes: every source and target spectrum once
    # and the 316 translation spectra, (n_vsrc*sd + n_vtrg*td)*H*2w +
    # 316*sd*td*H*2w with w = 4 (FP32) or 8 (FP64).  The intensity (about 78
    # flop/byte in FP32) is far above the machine balance (157.3 TFLOP/s over
    # 8 TB/s, about 20), so the compute peak is the bound.
    w = f.work_counts()
    mixed = args.precision == "mixed"
    had_ms = phases["VListHadamard"]
    wbytes = 4 if mixed else 8
    H = (np_ + 1) * (2 * np_) ** 2
    sd = td = 1  # equivalent-density kernel 111
    alg_bytes = (w["n_vsrc"] * sd + w["n_vtrg"] * td) * H * 2 * wbytes + 316 * sd * td * H * 2 * wbytes
    alg_flops = 8.0 * w["vlist_pairs"] * H * sd * td
    had_tf = alg_flops / (had_ms * 1e-3) / 1e12
    had_peak = PEAK_TFLOPS["fp32" if mixed else "fp64"]
    kname = "k_m2l_tile" if mixed else "k_hadamard"
    roofline = {"bound": "mfma", "kernel": kname, "achieved": round(had_tf, 3), "peak": had_peak,
                "unit": "TFLOP/s", "frac": round(had_tf / had_peak, 4), "traffic": traffic_bytes(kname),
                "kernel_ms": round(had_ms, 4), "algorithmic_flops": alg_flops, "algorithmic_bytes": alg_bytes,
                "intensity_flop_per_byte": round(alg_flops / alg_bytes, 2),
                "hbm_gbs_at_algorithmic_bytes": round(alg_bytes / (had_ms * 1e-3) / 1e9, 1)}
    # the near field (U/W/X S2T, M2T, L2T), the second-largest kernel: FP32/FP64
    # vector bound, 11 flops per Laplace pair
    near_ms = phases["EvlUList"]
    near_tf = w["near_pairs"] * 11.0 / (near_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS["fp32" if mixed else "fp64"]
    near = {"bound": "valu", "kernel": "k_near_lap_mx" if mixed else "k_near", "achieved": round(near_tf, 3),
            "peak": peak, "unit": "TFLOP/s", "frac": round(near_tf / peak, 4), "kernel_ms": round(near_ms, 4)}

    cpu = None
    if not args.no_cpu:
        nue = np_ ** 3 - (np_ - 2) ** 3
        nuc = (np_ + 2) ** 3 - np_ ** 3
        pu = f.operator(0).reshape(nuc, nue).T
        pd = f.operator(1).reshape(nue, nue).T
        sample = min(args.cpu_sample, n)
        mps, secs = cpu_baseline(cfg, (pu, pd), sample, cfg_index)
        cpu = {"value": round(mps, 5), "unit": "Mpoints/s", "cores": 1, "kind": "port",
               "sample": "%d points of the same distribution, one evaluate, %.1f s" % (sample, secs)}

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
        "config": {"workload": WORKLOAD_NAME[args.config] if not args.n else "%s (n=%d)" % (args.config, n),
                   "kernel": kt, "np": np_, "ptsMax": pm, "maxLevel": ml, "precision": args.precision,
                   "n_points": n},
        "rel_err": rel_err,
        "fp64_ms_per_step": round(dt64 * 1e3, 4) if dt64 else None,
        "speedup_vs_fp64": round(dt64 / dt, 3) if dt64 else None,
        "rel_diff_vs_fp64": diff_fp64,
        "phases_ms": {k: round(v, 4) for k, v in phases.items()},
        "roofline": roofline,
        "near_field": near,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
