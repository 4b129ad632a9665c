"""Benchmark: pair Coulomb potentials/s on one MI355X at 81^3.

Workload: 81^3 sinc-DVR grid, spacing 0.25 bohr, 16 hermnorm orbitals (each a
seeded sum of 6 Gaussians, Gram-Schmidt orthonormalised) -> 136 unique / 256
ordered orbital pairs, circulant edge M = 162.  A step delivers the full
twoereduced(81^3, 16, 16) array (every pair potential, mirrors filled) from
orbitals already resident in HBM.  `value` is the plain FP64 form; beside it:
  * mixed: FP32 butterflies, FP64 spectrum construction and reductions, with
    its two-electron matrix-element error against FP64 per element;
  * batched_mixed_81: the same pairs through the batch API (one hipGraph per
    step), mixed precision;
  * batch: 64 orbital sets of uneven grids (edges 25..61, 4..14 orbitals),
    FP32 and FP64, grouped schedule, one hipGraph per step;
  * serving: Poisson request streams from 256 simulated propagations (edges
    31..81, 6..16 orbitals) against the request engine, mixed and FP64, with
    p50/p99 latency at 70% of each form's measured capacity;
  * two-electron matrix elements and the kinetic apply, timed separately.

Roofline of the dominant kernel, the fused y/z kernel k_yz: its algorithmic
HBM bytes are what the fused kernel must move, read the M n^2 x-forward
output once and write the M n^2 input of the x-inverse once, s (2 M n^2) per
convolution (s = 16 FP64, 8 FP32; 34.0 MB FP64 at 81^3), plus the spectrum
slices once per launch; `traffic` is what it does move (calibrated rocprofv3
counters, profiles/), and `bound` is what limits it by the SQ counters
(VALU issue and LDS exchanges, not HBM).  The path's effective bandwidth
against a streaming pencil implementation, B = s (2 n^3 + 4 M n^2 + 4 M^2 n)
per convolution, is reported beside it as path_effective_pencil_frac.  The
136 pair potentials take 128 convolutions: the 16 diagonal pair densities
|phi_a|^2 are real and run two per complex convolution.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--quick]
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
MFMA_PEAK_F64 = 78.6    # TFLOP/s dense, MI355X
MFMA_PEAK_F32 = 157.3


def fused_bytes_per_conv(stage, n, M, s_elem):
    """Algorithmic HBM bytes of one convolution in each fused kernel: the
    inputs it must read once and the outputs it must write once."""
    if stage == "xfwd":        # two orbital columns (n^3 each, FP64) -> M n^2 x-frequency lines
        return 2 * 16 * n ** 3 + s_elem * M * n * n
    if stage == "yz":          # M n^2 lines in, M n^2 lines out (the spectrum once per launch: separate)
        return 2 * s_elem * M * n * n
    if stage == "xinv":        # M n^2 lines -> one or two output columns (n^3, FP64)
        return s_elem * M * n * n + 2 * 16 * n ** 3
    raise ValueError(stage)


def alg_bytes_per_pair(stage, n, M, s_elem):
    """Pencil-model HBM bytes of one pair potential in each kernel's passes."""
    if stage == "xfwd":
        return s_elem * (n ** 3 + M * n * n)
    if stage == "yz":
        return s_elem * (2 * M * n * n + 4 * M * M * n)
    if stage == "xinv":
        return s_elem * (M * n * n + n ** 3)
    if stage == "path":
        return s_elem * (2 * n ** 3 + 4 * M * n * n + 4 * M * M * n)
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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, dx, phi, cores, nbatch_ref=4, nbatch_port=4):
    """The reference's own CPU convolution -- circ3d_sub, myzfft3d and the
    DFFTPACK backend of myzfft1d compiled from the reference sources into
    oracle/_ref/libref_conv.so, framed as op_tinv_notscaled -- on `nbatch_ref`
    spf2b batches of call_twoe_matel00 (each N densities phi(:,a) conj(phi(:,b))
    at M = 2n = 162 with the kernel FFT recomputed, as the reference does),
    scaled to the full twoereduced (N batches); beside it the oracle's C
    restatement of the same algorithm.  Falls back to the restatement alone
    (kind "port") when oracle/_ref is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import oracle
    oracle.set_threads(cores)
    kern = oracle.coulomb_kernel(n, dx)
    N = phi.shape[1]
    pairs = N * (N + 1) // 2

    def timed(fn, nb):
        t = 0.0
        for b in range(nb):
            den = np.asfortranarray(phi * np.conj(phi[:, [b]]))
            t0 = time.perf_counter()
            fn(den)
            t += time.perf_counter() - t0
        return t

    t_port = timed(lambda d: oracle.op_tinv(n, kern, d), nbatch_port)
    port = {"value": pairs / (t_port * N / nbatch_port), "unit": "pair Coulomb potentials/s", "cores": cores,
            "sample": "%d of the %d spf2b batches, %.1f s, scaled to the full twoereduced" % (nbatch_port, N, t_port)}
    host = {"nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "threads_note": "threads = the GPU host's CPU share for one GPU (OMP_NUM_THREADS there)"}
    if oracle.conv_lib() is None:
        return dict(port, kind="port", **host)
    t_ref = timed(lambda d: oracle.ref_op_tinv(n, kern, d, threads=cores), nbatch_ref)
    return dict({"value": pairs / (t_ref * N / nbatch_ref), "unit": "pair Coulomb potentials/s", "cores": cores,
                 "kind": "reference",
                 "sample": "%d of the %d spf2b batches of call_twoe_matel00 at 81^3 (each %d densities + the "
                           "kernel FFT, M=162, through the reference's circ3d_sub / myzfft3d / DFFTPACK built from "
                           "its sources), %.1f s, scaled to the full twoereduced" % (nbatch_ref, N, N, t_ref),
                 "port": port}, **host)


def bench_small_configs(torch, steps):
    """Pair potentials/s of the single-set path on the 21^3 / 4-orbital and
    41^3 / 10-orbital workloads (FP64 and mixed), beside the 81^3 headline."""
    from sincdvr import SincDVR, inputs
    res = {}
    for key in ("grid21_4orb", "grid41_10orb"):
        cfg = inputs.CONFIGS[key]
        n, dx, N = cfg["n"], cfg["spacing"], cfg["norb"]
        phi = inputs.gaussian_orbitals_device(n, dx, N, cfg["ngauss"], cfg["seed"])
        s = SincDVR(n, dx, N)
        tr = torch.zeros((N, N, n ** 3), dtype=torch.complex128, device="cuda")
        ent = {"grid": n, "orbitals": N, "unique_pairs": N * (N + 1) // 2, "circulant_edge": s.circulant_edge}
        for prec, name in ((0, "fp64"), (1, "mixed")):
            s.set_precision(prec)
            t = time_steps(lambda: s.twoereduced(1, N, phi, out=tr), max(steps, 20), 3, torch)
            ent[name] = {"pairs_per_s": N * (N + 1) // 2 / t, "ms_per_step": t * 1e3}
        res[key] = ent
    return res


def bench_batched_81(torch, s, phi, npairs, steps):
    """The 81^3 / 16-orbital pairs through the batch API, mixed precision,
    replayed as one hipGraph per step."""
    from sincdvr import Batch
    b = Batch([(s, phi)], precision=1)
    side = torch.cuda.Stream()
    b.capture(side)
    for _ in range(2):
        b.run(side)
    side.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run(side)
    side.synchronize()
    t = (time.perf_counter() - t0) / steps
    b.close()
    return {"value": npairs / t, "ms_per_step": t * 1e3, "precision": "f32 butterflies", "graph": True}


def bench_batch(torch, quick):
    """64 orbital sets of uneven grids, grouped schedule, one hipGraph per step."""
    from sincdvr import Batch, SincDVR, inputs
    descs = inputs.batch_sets(nsets=16 if quick else 64)
    sets = []
    for d in descs:
        phi = inputs.gaussian_orbitals_device(d["n"], d["spacing"], d["norb"], d["ngauss"], d["seed"])
        sets.append((SincDVR(d["n"], d["spacing"], d["norb"]), phi))
    npairs = sum(d["norb"] * (d["norb"] + 1) // 2 for d in descs)
    out = {"sets": len(descs), "unique_pairs": npairs, "grid_edges": sorted({d["n"] for d in descs}),
           "schedule": "grouped (one launch per stage per circulant size), hipGraph per step"}
    for prec, name in ((1, "mixed"), (0, "fp64")):
        b = Batch(sets, precision=prec)
        side = torch.cuda.Stream()
        b.capture(side)
        b.run(side)
        side.synchronize()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            b.run(side)
        side.synchronize()
        t = (time.perf_counter() - t0) / reps
        out[name] = {"pairs_per_s": npairs / t, "ms_per_step": t * 1e3,
                     "work_buffer_gb": b.stats()["work_bytes"] / 1e9}
        b.close()
    del sets
    torch.cuda.empty_cache()
    return out


def bench_serving(torch, quick):
    """Request engine under Poisson load from simulated propagations."""
    from sincdvr import serving
    nprop = 32 if quick else 256
    descs = serving.propagations(nprop=nprop)
    ctxs, props = serving.build_propagations(descs, steps=2)
    # 1 GiB ticks: the same capacity as 4 GiB ticks for these requests, with a
    # shorter wait for a request that arrives while a tick runs (Coulomb p99
    # at 70 % load over repeated runs: mixed 8.9-13.2 ms with 4 GiB ticks,
    # 7.2-9.7 ms with 1 GiB; FP64 14.0-14.6 ms -> 11.5-11.8 ms)
    budget, tick = 48 << 30, 1 << 30
    res = {"propagations": nprop, "grid_edges": list(serving.SERVING_EDGES), "orbitals": "6..16",
           "hbm_budget_gb": budget / 2 ** 30, "tick_work_gb": tick / 2 ** 30,
           "requests": "alternating Coulomb (all pair potentials) and kinetic apply per propagation",
           "latency": "submit to device result ready, observed by the host, ms"}
    for prec, name in ((1, "mixed"), (0, "fp64")):
        srv = serving.Server(budget, tick, precision=prec)
        # capacity: every request submitted at once
        burst = serving.request_stream(props, nreq=nprop, rate=1e9, seed=11)
        serving.run_stream(srv, props, burst[: min(16, len(burst))])          # warm-up
        lat, q, wall = serving.run_stream(srv, props, burst)
        cap = len(burst) / wall
        pairs = serving.pair_count(props, burst)
        rate = 0.7 * cap
        nreq = 2 * nprop
        stream = serving.request_stream(props, nreq=nreq, rate=rate, seed=12)
        lat, q, wall = serving.run_stream(srv, props, stream)
        ent = {"capacity_requests_per_s": cap, "capacity_pair_potentials_per_s": pairs / (len(burst) / cap),
               "offered_requests_per_s": rate, "requests": nreq, "achieved_requests_per_s": nreq / wall,
               "pair_potentials_per_s": serving.pair_count(props, stream) / wall,
               "coulomb_latency_ms": serving.percentiles(lat[serving.COULOMB]),
               "kinetic_latency_ms": serving.percentiles(lat[serving.KE]),
               "queue_ms": serving.percentiles(q), "arena_peak_gb": srv.stats()["arena_peak"] / 2 ** 30,
               "ticks": srv.stats()["ticks"]}
        res[name] = ent
        srv.close()
    del props, ctxs
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # steady state: the GPU's clock settles over the first steps (81^3 FP64:
    # 3.86 ms per step after 3 warmup steps, 3.77 after 20)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="smaller batch and serving sections")
    ap.add_argument("--no-extra", action="store_true", help="headline forms only")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this library runs the pair potentials on one GPU; --gpus must be 1 "
              "(pairs shard across GPUs by unique-pair ranges, which this benchmark does not run)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from sincdvr import SincDVR, inputs

    cfg = inputs.CONFIGS["grid81_16orb"]
    n, dx, N = cfg["n"], cfg["spacing"], cfg["norb"]
    npairs = N * (N + 1) // 2
    print(f"[bench] generating {N} orbitals on {n}^3", file=sys.stderr, flush=True)
    phi_h = inputs.gaussian_orbitals(n, dx, N, cfg["ngauss"], cfg["seed"])
    phi = torch.from_numpy(np.ascontiguousarray(phi_h.T)).cuda()
    print("[bench] building kernel spectrum", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s = SincDVR(n, dx, N, precision=0)           # GETINVERSE + kernel placement + spectra, on the device
    t_setup = time.perf_counter() - t0
    M = s.circulant_edge
    tr = torch.zeros((N, N, n ** 3), dtype=torch.complex128, device="cuda")
    tm = torch.zeros((N, N, N, N), dtype=torch.complex128, device="cuda")

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
    bm = None
    if not args.no_extra:  # right after the single-set mixed form, under the same conditions
        print("[bench] batched mixed form at 81^3", file=sys.stderr, flush=True)
        bm = bench_batched_81(torch, s, phi, npairs, args.steps)
        bm["speedup_vs_fp64"] = bm["value"] * t64 / npairs
        s.set_precision(1)
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
           "p99_rel_elem": float(torch.quantile(rel_elem[big].flatten()[:2 ** 24], 0.99)),
           "elements_counted": int(big.sum()), "elements": int(tm64.numel()),
           "twoereduced_max_rel": float((tr - tr64).abs().max() / tr64.abs().max())}
    del tr64
    print("[bench] kinetic apply", file=sys.stderr, flush=True)
    ke_res = {}
    ke_out = {}
    for prec, name, peak in ((0, "fp64", MFMA_PEAK_F64), (1, "mixed", MFMA_PEAK_F32)):
        s.set_precision(prec)
        tk = time_steps(lambda: s.mult_ke(phi), args.steps, 1, torch)
        ke_out[name] = s.mult_ke(phi)
        useful = 12.0 * n ** 4 * N                                    # 3 axes x 2 (re, im) x 2 n^4
        # matrix-core work as issued: the x/y pass runs nt^2 16 x 16 tiles per
        # plane (2 terms, re and im, np/4 steps of 16x16x4), the z pass nq
        # 16-row tiles per 16 columns; with n = 16k + 1 the last row/column is
        # on the VALU and not counted
        npad = (n + 3) // 4 * 4
        nt = n // 16 if n % 16 == 1 else (n + 15) // 16
        issued_xy = nt * nt * 2 * 2 * (npad // 4) * 2048.0 * n * N
        issued_z = ((n * n + 15) // 16) * nt * 2 * (npad // 4) * 2048.0 * N
        issued = issued_xy + issued_z
        hbm = 5 * 16.0 * n ** 3 * N   # z pass: in read, out written; x/y pass: in read, out read and written
        ke_res[name] = {"ms_per_call_16_orbitals": tk * 1e3, "orbitals_per_s": N / tk,
                        "mfma_tflops_useful": useful / tk / 1e12, "mfma_tflops_issued": issued / tk / 1e12,
                        "mfma_peak_tflops": peak, "mfma_utilisation": issued / tk / 1e12 / peak,
                        "hbm_bytes_per_call": hbm, "hbm_frac": hbm / tk / 1e9 / HBM_PEAK_GBS}
    s.set_precision(0)
    ke_res["mixed"]["error_vs_fp64_max_abs_over_max"] = float(
        (ke_out["mixed"] - ke_out["fp64"]).abs().max() / ke_out["fp64"].abs().max())
    del ke_out
    tke = ke_res["fp64"]["ms_per_call_16_orbitals"] / 1e3
    print("[bench] mean field", file=sys.stderr, flush=True)
    # reducedpot = twoereduced . twoden (S/mean.f90:556-558), twoden a seeded N^2 x N^2 complex matrix
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    twoden = torch.complex(torch.randn((N * N, N * N), dtype=torch.float64, device="cuda", generator=gen),
                           torch.randn((N * N, N * N), dtype=torch.float64, device="cuda", generator=gen))
    tmf = time_steps(lambda: s.get_reducedpot(tr, twoden), args.steps, 1, torch)
    mf_flops = 8.0 * n ** 3 * N ** 4
    mf_issued = 6.0 * n ** 3 * N ** 4           # three real products per complex one on the matrix cores
    mean_field = {"ms_per_call": tmf * 1e3, "tflops": mf_flops / tmf / 1e12, "peak_tflops_f64": MFMA_PEAK_F64,
                  "frac": mf_flops / tmf / 1e12 / MFMA_PEAK_F64,
                  "mfma_tflops_issued": mf_issued / tmf / 1e12,
                  "mfma_utilisation": mf_issued / tmf / 1e12 / MFMA_PEAK_F64,
                  "shape": "(n^3 x N^2) . (N^2 x N^2) complex, 8 n^3 N^4 real flops (tflops, frac); issued as "
                           "three real products per complex one (Gauss), 6 n^3 N^4 on the matrix cores"}
    del twoden

    # roofline of the dominant kernel of the FP64 form
    stages = ["xfwd", "yz", "xinv"]
    dom = max(stages, key=lambda k: prof64[k][0])
    ms_tot, launches = prof64[dom]
    avg_ms = ms_tot / launches
    nconv = s.convolutions(1, N)                 # 16 diagonal pairs run two per convolution
    pairs_per_launch = npairs * args.steps / launches
    conv_per_launch = nconv * args.steps / launches
    m = M // 2
    alg = fused_bytes_per_conv(dom, n, M, 16) * conv_per_launch
    if dom == "yz":
        alg += 8.0 * (m + 1) ** 3                 # the FP64 spectrum octant, read once per launch
    achieved = alg / (avg_ms * 1e-3) / 1e9
    traffic = None
    counters = {}
    tfile = os.path.join(ROOT, "profiles", "traffic_k_yz_81.json")
    if dom == "yz" and os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        traffic = tj.get("bytes_per_launch")
        if traffic is not None and tj.get("convolutions_per_launch"):
            traffic = traffic * conv_per_launch / tj["convolutions_per_launch"]
        else:
            traffic = None                            # counters of another work split
        counters = {k: tj[k] for k in ("valu_busy", "valu_busy_source", "lds_bank_conflict_frac", "wait_any_frac",
                                       "waves_per_cu", "source") if k in tj}
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
                               "(136 unique / 256 ordered pairs), circulant edge %d" % M,
                   "grid": n, "orbitals": N, "unique_pairs": npairs, "circulant_edge": M,
                   "parallelism": "1 GPU"},
        "roofline": {"bound": "valu" if dom == "yz" else "hbm", "kernel": "k_" + dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": min(1.0, achieved / HBM_PEAK_GBS),
                     "traffic": traffic,
                     "traffic_frac": (traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None,
                     "alg_bytes_per_launch": alg, "alg_bytes": "fused kernel's floor: read the x-forward "
                     "output once, write the x-inverse input once (+ spectrum octant)",
                     "limiter": "FP64 VALU issue, with the LDS exchanges' writes overlapping it only in part "
                     "(rocprofv3 VALUBusy and SQ counters, profiles/)" if dom == "yz" else "HBM",
                     "sq_counters": counters or None,
                     "pairs_per_launch": pairs_per_launch,
                     "convolutions_per_launch": conv_per_launch, "avg_launch_ms": avg_ms},
        "path_effective_pencil_frac": alg_bytes_per_pair("path", n, M, 16) * nconv / t64 / 1e9 / HBM_PEAK_GBS,
        "path_effective_note": "pencil-model bytes B = s(2n^3 + 4Mn^2 + 4M^2n) per convolution / step time / "
                               "8 TB/s: the path's speed against a streaming 3-pass pencil FFT at HBM peak; the "
                               "fused kernels move less than B",
        "stage_ms_per_step_fp64": {k: prof64[k][0] / args.steps for k in stages},
        "mixed": {"value": npairs / t32, "ms_per_step": t32 * 1e3, "speedup_vs_fp64": t64 / t32,
                  "path_effective_pencil_frac": alg_bytes_per_pair("path", n, M, 8) * nconv / t32 / 1e9
                  / HBM_PEAK_GBS,
                  "dtype": "f32 butterflies, f64 spectrum construction and reductions",
                  "twoematel_error_vs_fp64": err,
                  "stage_ms_per_step": {k: prof32[k][0] / args.steps for k in stages}},
        "twoematel_ms_fp64": tmat * 1e3,
        "mean_field_fp64": mean_field,
        "spectrum_setup_ms": t_setup * 1e3,
        "mult_ke": dict(ke_res, note="two passes on the matrix cores (FP64 v_mfma_f64_16x16x4f64, mixed "
                                     "v_mfma_f32_16x16x4f32): z contraction from registers, then the x and y "
                                     "contractions of each z plane from one staged copy; utilisation = issued "
                                     "matrix flops / peak"),
    }
    del tm64, tm32, tm, tr
    torch.cuda.empty_cache()
    if not args.no_extra:
        out["batched_mixed_81"] = bm
        print("[bench] 21^3 and 41^3 workloads", file=sys.stderr, flush=True)
        out["smaller_grids"] = bench_small_configs(torch, args.steps)
        print("[bench] 64-set batch", file=sys.stderr, flush=True)
        out["batch"] = bench_batch(torch, args.quick)
        print("[bench] serving", file=sys.stderr, flush=True)
        out["serving"] = bench_serving(torch, args.quick)
    if not args.no_cpu:
        cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
        print(f"[bench] CPU baseline on {cores} threads", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(n, dx, phi_h, cores)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
