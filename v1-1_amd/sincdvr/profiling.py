"""Profiling workload and rocprofv3 summariser.

run: the 81^3 / 16-orbital pair-potential step (the benchmark workload) a few
     times, plus a plain streaming kernel of a known size (torch.neg from one
     COPY_BYTES buffer into another: reads and writes COPY_BYTES once), so
     that rocprofv3 counters can be calibrated: FETCH_SIZE and WRITE_SIZE of
     that kernel against its known bytes give the factors that turn the
     counters of the convolution kernels into HBM bytes.

summarize: reads rocprofv3 output (kernel-trace stats CSV, counter CSVs of a
     FETCH_SIZE pass and a WRITE_SIZE pass) and writes a short markdown summary
     and the traffic figure the benchmark reports.

    rocprofv3 --kernel-trace --stats --output-format csv -d OUT/trace -o run -- python3 bench.py --no-cpu --no-extra --steps 20 --warmup 20
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT/fetch -o run -- python3 v1-1_amd/sincdvr/profiling.py run
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT/write -o run -- python3 v1-1_amd/sincdvr/profiling.py run
    python3 v1-1_amd/sincdvr/profiling.py summarize OUT profiles

sq: the SQ counter table (VALU and LDS instructions, waits, LDS bank
     conflicts, VALU busy) of the convolution kernels from one counter pass
     per precision (summarize_sq gives the commands).
"""
import argparse
import csv
import glob
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

COPY_BYTES = 1 << 30          # read and written once by the calibration copy
PAIRS_PER_LAUNCH = 136        # every pair of the 16-orbital step in one launch (the default)
CONV_PER_LAUNCH = 128         # its convolutions: 120 off-diagonal pairs + 16 diagonal ones two per convolution
KERNELS = ("k_xfwd", "k_yz", "k_xinv", "k_ke_xy", "k_ke_z")


def run_ke(precision, steps):
    """The kinetic apply (mult_ke) of the 81^3 / 16-orbital orbitals, plain and
    with ECS (the 41^3 workload's ramp parameters)."""
    import numpy as np
    import torch

    from sincdvr import SincDVR, inputs
    cfg = inputs.CONFIGS["grid81_16orb"]
    n, dx, N = cfg["n"], cfg["spacing"], cfg["norb"]
    phi = torch.from_numpy(np.ascontiguousarray(
        inputs.gaussian_orbitals(n, dx, N, cfg["ngauss"], cfg["seed"]).T)).cuda()
    for sc in (None, inputs.ECS_41):
        s = SincDVR(n, dx, N, scaling=sc, precision=precision)
        for i in range(steps):
            s.mult_ke(phi)
        torch.cuda.synchronize()
        print(f"[profiling] mult_ke ecs={sc is not None} done", flush=True)


def run(precision, steps):
    import numpy as np
    import torch

    from sincdvr import SincDVR, inputs
    cfg = inputs.CONFIGS["grid81_16orb"]
    n, dx, N = cfg["n"], cfg["spacing"], cfg["norb"]
    phi = torch.from_numpy(np.ascontiguousarray(
        inputs.gaussian_orbitals(n, dx, N, cfg["ngauss"], cfg["seed"]).T)).cuda()
    s = SincDVR(n, dx, N, precision=precision)
    s.set_pair_batch(PAIRS_PER_LAUNCH)
    assert s.convolutions(1, N) == CONV_PER_LAUNCH, s.convolutions(1, N)
    tr = torch.zeros((N, N, n ** 3), dtype=torch.complex128, device="cuda")
    for i in range(steps):
        s.twoereduced(1, N, phi, out=tr)
        torch.cuda.synchronize()
        print(f"[profiling] step {i} done", flush=True)
    src = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.neg(src, out=dst)   # an elementwise kernel (a same-type copy_ may go to a DMA engine)
    torch.cuda.synchronize()
    print("[profiling] copy done", flush=True)


def _rows(path_glob):
    out = []
    for f in sorted(glob.glob(path_glob, recursive=True)):
        with open(f) as fh:
            out.extend(csv.DictReader(fh))
    return out


def _kernel_of(name):
    for k in KERNELS:
        if f"sdv::{k}<" in name:
            return name.split("(")[0].replace("void ", "")
    if "elementwise" in name and "neg" in name.lower():
        return "copy"
    return None


def counter_per_dispatch(rows, counter):
    """{kernel: value per dispatch} over the dispatches of the largest grid of
    each kernel (the full pair batches)."""
    per = {}
    for r in rows:
        if r.get("Counter_Name") != counter:
            continue
        k = _kernel_of(r.get("Kernel_Name", ""))
        if k is None:
            continue
        d = int(r.get("Dispatch_Id", 0))
        grid = int(r.get("Grid_Size", 0) or 0)
        per.setdefault(k, {}).setdefault(d, [grid, 0.0])
        per[k][d][1] += float(r["Counter_Value"])
    res = {}
    for k, ds in per.items():
        gmax = max(g for g, _ in ds.values())
        vals = [v for g, v in ds.values() if g == gmax]
        res[k] = sum(vals) / len(vals)
    return res


def summarize(out_dir, prof_dir, tag="pairpot_81_n16"):
    stats = _rows(os.path.join(out_dir, "trace", "**", "*kernel_stats.csv"))
    fetch = counter_per_dispatch(_rows(os.path.join(out_dir, "fetch", "**", "*counter_collection.csv")),
                                 "FETCH_SIZE")
    write = counter_per_dispatch(_rows(os.path.join(out_dir, "write", "**", "*counter_collection.csv")),
                                 "WRITE_SIZE")
    lines = []
    if stats:
        lines.append("| kernel | calls | avg us | total ms | % |")
        lines.append("|---|---:|---:|---:|---:|")
        for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:12]:
            lines.append("| %s | %s | %.1f | %.3f | %.1f |" % (
                r["Name"].split("(")[0].replace("void ", "")[:70], r["Calls"], float(r["AverageNs"]) / 1e3,
                float(r["TotalDurationNs"]) / 1e6, float(r["Percentage"])))
    traffic = {}
    if "copy" in fetch and "copy" in write:
        fk = COPY_BYTES / 1024.0 / fetch["copy"]
        wk = COPY_BYTES / 1024.0 / write["copy"]
        traffic["fetch_factor"] = fk
        traffic["write_factor"] = wk
        lines.append("")
        lines.append("Counter calibration: a %d-byte device copy reads FETCH_SIZE %.0f KiB and WRITE_SIZE %.0f KiB "
                     "per dispatch; factors (bytes per counted KiB / 1024) fetch %.3f, write %.3f." % (
                         COPY_BYTES, fetch["copy"], write["copy"], fk, wk))
        lines.append("")
        lines.append("| kernel (full dispatch: %d pair potentials, %d convolutions) | FETCH KiB | WRITE KiB | "
                     "HBM bytes (calibrated) |" % (PAIRS_PER_LAUNCH, CONV_PER_LAUNCH))
        lines.append("|---|---:|---:|---:|")
        for k in sorted(set(fetch) & set(write)):
            if k == "copy":
                continue
            b = (fetch[k] * fk + write[k] * wk) * 1024.0
            traffic[k] = {"fetch_kib": fetch[k], "write_kib": write[k], "bytes_per_launch": b,
                          "pairs_per_launch": PAIRS_PER_LAUNCH, "convolutions_per_launch": CONV_PER_LAUNCH}
            lines.append("| %s | %.0f | %.0f | %.3e |" % (k, fetch[k], write[k], b))
    os.makedirs(prof_dir, exist_ok=True)
    with open(os.path.join(prof_dir, f"{tag}_profile_summary.md"), "w") as f:
        f.write(f"# rocprofv3 summary: {tag}\n\n")
        f.write("Commands (repository root, on the GPU):\n\n")
        f.write("    rocprofv3 --kernel-trace --stats --output-format csv -d OUT/trace -o run -- python3 bench.py --no-cpu --no-extra --steps 20 --warmup 20\n")
        f.write("    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT/fetch -o run -- python3 v1-1_amd/sincdvr/profiling.py run\n")
        f.write("    rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT/write -o run -- python3 v1-1_amd/sincdvr/profiling.py run\n")
        f.write("    python3 v1-1_amd/sincdvr/profiling.py summarize OUT profiles\n\n")
        f.write("\n".join(lines) + "\n")
    for k, v in traffic.items():
        if isinstance(v, dict) and k.startswith("sdv::k_yz<double"):
            with open(os.path.join(prof_dir, "traffic_k_yz_81.json"), "w") as f:
                json.dump(dict(v, kernel=k, fetch_factor=traffic["fetch_factor"],
                               write_factor=traffic["write_factor"]), f, indent=1)
    return lines


SQ_COUNTERS = ("SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT",
               "SQ_LDS_IDX_ACTIVE", "SQ_WAIT_ANY", "SQ_WAIT_INST_ANY", "SQ_WAVE_CYCLES")
SIMDS = 1024                  # 256 CUs x 4 SIMDs
# resident waves per SIMD of the 9x9 y/z kernel (768-thread workgroups): FP64 one
# workgroup per CU (LDS), FP32 two (compiled for 6 waves per SIMD)
YZ_WAVES_PER_SIMD = {0: 3, 1: 6}


def summarize_sq(out_dir, prof_dir):
    """Table of SQ counters per full dispatch from the passes OUT/sq0 (FP64)
    and OUT/sq1 (mixed).  SQ_WAVE_CYCLES counts quad cycles; one VALU
    instruction of a wavefront takes one quad cycle on its SIMD (FP64 and
    unpacked FP32 alike), so VALU busy = VALU instructions per SIMD over the
    SIMD quad cycles of the dispatch (its wavefront cycles over SIMDS x
    resident waves per SIMD).  Each pass lists the kernels of its own element
    type (the mixed pass also runs FP64 convolutions to build its carrier)."""
    lines = [
        "# rocprofv3 SQ counters: k_yz / k_xfwd / k_xinv at 81^3, 16 orbitals (%d convolutions per dispatch)"
        % CONV_PER_LAUNCH, "",
        "Commands (repository root, on the GPU; one pass per precision P = 0 FP64, 1 mixed):", "",
        "    rocprofv3 --pmc %s --output-format csv -d OUT/sqP -o run -- python3 v1-1_amd/sincdvr/profiling.py "
        "run --precision P --steps 1" % " ".join(SQ_COUNTERS),
        "    python3 v1-1_amd/sincdvr/profiling.py sq OUT profiles", "",
        "Per dispatch of the full %d-convolution grid (the largest grid of each kernel), summed over the" % CONV_PER_LAUNCH,
        "device.  Waits are fractions of wavefront cycles; bank conflicts are per LDS-active cycle.",
        "VALU rate = VALU instructions per SIMD quad cycle, with the",
        "dispatch's wavefront cycles spread over %d SIMDs at the resident waves per SIMD (for k_yz %d in"
        % (SIMDS, YZ_WAVES_PER_SIMD[0]), "FP64, %d in FP32).  An FP64 instruction takes a quad cycle, so in FP64 the rate is the" % YZ_WAVES_PER_SIMD[1],
        "VALU busy fraction; unpacked FP32 issues at up to twice that rate on gfx950, so the FP32",
        "rate is no busy fraction (rocprofv3's VALUBusy, profiles/utilisation_81_n16.md, is).", "",
        "| kernel | VALU insts | LDS insts | wait any / wave cycles | wait issue / wave cycles | "
        "LDS bank conflict / LDS active |",
        "|---|---:|---:|---:|---:|---:|"]
    yz64 = None
    for p, label in ((0, "FP64"), (1, "FP32")):
        rows = _rows(os.path.join(out_dir, f"sq{p}", "**", "*counter_collection.csv"))
        c = {name: counter_per_dispatch(rows, name) for name in SQ_COUNTERS}
        busy = None
        elem = "<double," if p == 0 else "<float,"
        for k in sorted(c["SQ_INSTS_VALU"]):
            if elem not in k:
                continue
            wc = c["SQ_WAVE_CYCLES"].get(k, 0.0)
            lds_act = c["SQ_LDS_IDX_ACTIVE"].get(k, 0.0)
            lines.append("| %s | %.3e | %.3e | %.2f | %.2f | %.2f |" % (
                k, c["SQ_INSTS_VALU"][k], c["SQ_INSTS_LDS"].get(k, 0.0),
                c["SQ_WAIT_ANY"].get(k, 0.0) / wc if wc else 0.0,
                c["SQ_WAIT_INST_ANY"].get(k, 0.0) / wc if wc else 0.0,
                c["SQ_LDS_BANK_CONFLICT"].get(k, 0.0) / lds_act if lds_act else 0.0))
            if k.startswith("sdv::k_yz<") and wc:
                busy = (c["SQ_INSTS_VALU"][k] / SIMDS) / (wc / (SIMDS * YZ_WAVES_PER_SIMD[p]))
                if p == 0:
                    lds_act = c["SQ_LDS_IDX_ACTIVE"].get(k, 0.0)
                    yz64 = {"valu_busy": busy, "waves_per_cu": 4 * YZ_WAVES_PER_SIMD[0],
                            "lds_bank_conflict_frac": (c["SQ_LDS_BANK_CONFLICT"].get(k, 0.0) / lds_act
                                                       if lds_act else None),
                            "wait_any_frac": c["SQ_WAIT_ANY"].get(k, 0.0) / wc,
                            "source": "rocprofv3 SQ counters (profiles/k_yz_81_n16_sq_counters.md)"}
        if busy is not None:
            lines.append("|  (k_yz VALU instructions per SIMD quad cycle, %s) | %.2f | | | | |" % (label, busy))
    os.makedirs(prof_dir, exist_ok=True)
    with open(os.path.join(prof_dir, "k_yz_81_n16_sq_counters.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    tj = os.path.join(prof_dir, "traffic_k_yz_81.json")
    if yz64 and os.path.exists(tj):  # the limiter figures the benchmark reports beside the traffic
        with open(tj) as f:
            t = json.load(f)
        t.update(yz64)
        with open(tj, "w") as f:
            json.dump(t, f, indent=1)
    return lines


def summarize_util(out_dir, prof_dir):
    """VALUBusy and MfmaUtil (rocprofv3's derived counters: VALU active
    quad-cycles, and matrix-core busy cycles, over the dispatch's GPU-active
    cycles) of the convolution kernels (OUT/util0, OUT/util1: one pass per
    precision over the pair-potential workload) and of the kinetic kernels
    (OUT/ke0, OUT/ke1: the kinetic workload); writes
    profiles/utilisation_81_n16.md and adds the FP64 y/z kernel's VALUBusy to
    profiles/traffic_k_yz_81.json."""
    lines = ["# VALUBusy / MfmaUtil at 81^3, 16 orbitals", "",
             "Commands (repository root, on the GPU; P = 0 FP64, 1 mixed):", "",
             "    rocprofv3 --pmc VALUBusy MfmaUtil --output-format csv -d OUT/utilP -o run -- python3 "
             "v1-1_amd/sincdvr/profiling.py run --precision P --steps 1",
             "    rocprofv3 --pmc VALUBusy MfmaUtil --output-format csv -d OUT/keP -o run -- python3 "
             "v1-1_amd/sincdvr/profiling.py run --what ke --precision P --steps 3",
             "    python3 v1-1_amd/sincdvr/profiling.py util OUT profiles", "",
             "Per dispatch (the largest grid of each kernel; kinetic kernels: the mean over the plain and",
             "ECS dispatches of the workload).", "",
             "| kernel | VALUBusy % | MfmaUtil % |", "|---|---:|---:|"]
    yz64 = None
    # each pass lists the kernels of its own element type and family (the
    # mixed passes also run FP64 convolutions, for the charge carrier)
    for sub, elem, fam in (("util0", "<double,", "k_x"), ("util0", "<double,", "k_yz"),
                           ("util1", "<float,", "k_x"), ("util1", "<float,", "k_yz"),
                           ("ke0", "<double,", "k_ke"), ("ke1", "<float,", "k_ke")):
        rows = _rows(os.path.join(out_dir, sub, "**", "*counter_collection.csv"))
        vb = counter_per_dispatch(rows, "VALUBusy")
        mu = counter_per_dispatch(rows, "MfmaUtil")
        for k in sorted(vb):
            if k == "copy" or elem not in k or ("sdv::" + fam) not in k:
                continue
            lines.append("| %s | %.1f | %.1f |" % (k, vb[k], mu.get(k, 0.0)))
            if sub == "util0" and k.startswith("sdv::k_yz<double"):
                yz64 = vb[k] / 100.0
    os.makedirs(prof_dir, exist_ok=True)
    with open(os.path.join(prof_dir, "utilisation_81_n16.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    tj = os.path.join(prof_dir, "traffic_k_yz_81.json")
    if yz64 is not None and os.path.exists(tj):
        with open(tj) as f:
            t = json.load(f)
        t["valu_busy"] = yz64
        t["valu_busy_source"] = "rocprofv3 VALUBusy (profiles/utilisation_81_n16.md)"
        with open(tj, "w") as f:
            json.dump(t, f, indent=1)
    return lines


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--precision", type=int, default=0)
    r.add_argument("--steps", type=int, default=3)
    r.add_argument("--what", choices=("pairpot", "ke"), default="pairpot")
    sm = sub.add_parser("summarize")
    sm.add_argument("out_dir")
    sm.add_argument("prof_dir")
    sm.add_argument("--tag", default="pairpot_81_n16")
    sq = sub.add_parser("sq")
    sq.add_argument("out_dir")
    sq.add_argument("prof_dir")
    ut = sub.add_parser("util")
    ut.add_argument("out_dir")
    ut.add_argument("prof_dir")
    a = ap.parse_args()
    if a.cmd == "run":
        (run_ke if a.what == "ke" else run)(a.precision, a.steps)
    elif a.cmd == "sq":
        print("\n".join(summarize_sq(a.out_dir, a.prof_dir)))
    elif a.cmd == "util":
        print("\n".join(summarize_util(a.out_dir, a.prof_dir)))
    else:
        print("\n".join(summarize(a.out_dir, a.prof_dir, a.tag)))


if __name__ == "__main__":
    main()
