"""Benchmark: pair Coulomb potentials/s on one MI355X at 81^3.

Workload: 81^3 sinc-DVR grid, spacing 0.25 bohr, 16 hermnorm orbitals (each a
seeded sum of 6 Gaussians, Gram-Schmidt orthonormalised) -> 136 unique / 256
ordered orbital pairs, circulant edge M = 162.  A step delivers the full
twoereduced(81^3, 16, 16) array (every pair potential, mirrors filled) from
orbitals already resident in HBM.  `value` is the plain FP64 form; the mixed
form (FP32 butterflies, FP64 spectrum construction and reductions) is reported
beside it with its measured error against FP64.  Two-electron matrix
elements and the kinetic apply are timed separately.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "v1-1_amd"))

METRIC = "pair Coulomb potentials/s at 81^3 (mixed vs FP64), % HBM peak, p99 latency"
HBM_PEAK_GBS = 8000.0


def alg_bytes_per_pair(stage, n, m, s_elem, mirror_frac):
    """Minimum HBM bytes one pair moves through each stage of the pipeline
    (complex elements of s_elem bytes for the intermediate planes, 16 bytes
    for FP64 orbitals and outputs)."""
    plane_set = 2 * m * n * n * s_elem          # X[pair][kx][z][y], 2m planes of n^2
    if stage == "xfwd":
        return 2 * n ** 3 * 16 + plane_set       # read two orbitals, write planes
    if stage == "yz":
        return 2 * plane_set                     # read + write planes
    if stage == "xinv":
        return plane_set + (1 + mirror_frac) * n ** 3 * 16
    raise ValueError(stage)


def time_steps(fn, steps, warmup, torch):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def cpu_baseline(n, dx, phi, cores):
    """The oracle (a C port of the reference's CPU path) on one spf2b batch
    of call_twoe_matel00 -- 16 densities convolved by 3 x (1-D FFT +
    transpose) at M = 2n = 162 with the kernel FFT recomputed, as the
    reference does -- scaled to the full twoereduced (16 batches)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle
    oracle.set_threads(cores)
    kern = oracle.coulomb_kernel(n, dx)
    N = phi.shape[1]
    den = np.asfortranarray(phi * np.conj(phi[:, [0]]))   # densities (:, a) * conj(phi(:, 1))
    t0 = time.perf_counter()
    oracle.op_tinv(n, kern, den)
    t = time.perf_counter() - t0
    full = t * N                                          # N batches deliver all N^2 ordered pairs
    return {"value": (N * (N + 1) // 2) / full, "unit": "pair Coulomb potentials/s", "cores": cores,
            "kind": "port",
            "sample": "one spf2b batch of call_twoe_matel00 at 81^3 (16 densities + kernel FFT, M=162, "
                      "%.1f s), scaled x16 to the full twoereduced" % t}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding is described in DESIGN.md",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from sincdvr import SincDVR, inputs

    cfg = inputs.CONFIGS["c3_81"]
    n, dx, N = cfg["n"], cfg["spacing"], cfg["norb"]
    npairs = N * (N + 1) // 2
    print(f"[bench] generating {N} orbitals on {n}^3", file=sys.stderr, flush=True)
    phi_h = inputs.gaussian_orbitals(n, dx, N, cfg["ngauss"], cfg["seed"])
    phi = torch.from_numpy(np.ascontiguousarray(phi_h.T)).cuda()
    print("[bench] building kernel spectrum", file=sys.stderr, flush=True)
    s = SincDVR(n, dx, N, precision=0)
    m = s.circulant_edge // 2
    tr = torch.zeros((N, N, n ** 3), dtype=torch.complex128, device="cuda")
    tm = torch.zeros((N, N, N, N), dtype=torch.complex128, device="cuda")
    mirror_frac = (npairs - N) / npairs

    def run_form(prec):
        s.set_precision(prec)
        step = lambda: s.twoereduced(1, N, phi, out=tr)        # noqa: E731
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        s.profile(True)
        t = time_steps(step, args.steps, 0, torch)
        prof = s.profile_read()
        s.profile(False)
        return t, prof

    print("[bench] FP64 form", file=sys.stderr, flush=True)
    t64, prof64 = run_form(0)
    tr64 = tr.clone()
    print("[bench] matrix elements (FP64)", file=sys.stderr, flush=True)
    s.set_precision(0)
    tmat = time_steps(lambda: s.call_twoe_matel00(1, N, phi, out=(tm, tr)), max(2, args.steps // 2), 1, torch)
    tm64 = tm.clone()
    print("[bench] mixed form", file=sys.stderr, flush=True)
    t32, prof32 = run_form(1)
    s.call_twoe_matel00(1, N, phi, out=(tm, tr))
    torch.cuda.synchronize()
    tm32 = tm.clone()
    s.set_precision(0)
    d = (tm32 - tm64).abs()
    rel_elem = d / tm64.abs().clamp_min(1e-300)
    big = tm64.abs() > 1e-6 * tm64.abs().max()
    err = {"max_abs_over_max": float(d.max() / tm64.abs().max()),
           "max_rel_elem": float(rel_elem[big].max()),
           "median_rel_elem": float(rel_elem[big].median()),
           "twoereduced_max_rel": float((tr - tr64).abs().max() / tr64.abs().max())}
    print("[bench] kinetic apply", file=sys.stderr, flush=True)
    ke_out = torch.empty_like(phi)
    s.profile(True)
    tke = time_steps(lambda: s.mult_ke(phi), args.steps, 1, torch)
    kprof = s.profile_read()
    s.profile(False)
    del ke_out

    # roofline of the dominant kernel of the FP64 form
    stages = ["xfwd", "yz", "xinv"]
    dom = max(stages, key=lambda k: prof64[k][0])
    ms_tot, launches = prof64[dom]
    avg_ms = ms_tot / launches
    pairs_per_launch = npairs * args.steps / launches
    alg = alg_bytes_per_pair(dom, n, m, 16, mirror_frac) * pairs_per_launch
    achieved = alg / (avg_ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_k_yz_81.json")
    if dom == "yz" and os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")
    path_bytes = sum(alg_bytes_per_pair(k, n, m, 16, mirror_frac) for k in stages) * npairs
    out = {
        "metric": METRIC,
        "value": npairs / t64,
        "unit": "pair Coulomb potentials/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t64 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "twoereduced for all orbital pairs: 81^3 grid, spacing 0.25, 16 hermnorm orbitals "
                               "(136 unique / 256 ordered pairs), circulant edge %d" % s.circulant_edge,
                   "grid": n, "orbitals": N, "unique_pairs": npairs, "circulant_edge": s.circulant_edge,
                   "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "kernel": "k_" + dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "alg_bytes_per_launch": alg, "avg_launch_ms": avg_ms},
        "path_hbm_frac": path_bytes / t64 / 1e9 / HBM_PEAK_GBS,
        "stage_ms_per_step_fp64": {k: prof64[k][0] / args.steps for k in stages},
        "mixed": {"value": npairs / t32, "ms_per_step": t32 * 1e3, "speedup_vs_fp64": t64 / t32,
                  "dtype": "f32 butterflies, f64 spectrum construction and reductions",
                  "twoematel_error_vs_fp64": err,
                  "stage_ms_per_step": {k: prof32[k][0] / args.steps for k in stages}},
        "twoematel_ms_fp64": tmat * 1e3,
        "mult_ke": {"orbitals_per_s": N / tke, "ms_per_call_16_orbitals": tke * 1e3},
        "serving": None,
    }
    if not args.no_cpu:
        cores = min(os.cpu_count() or 1, 16)
        print(f"[bench] CPU baseline on {cores} threads", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(n, dx, phi_h, cores)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
