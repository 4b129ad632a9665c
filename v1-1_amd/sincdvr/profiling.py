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

    rocprofv3 --kernel-trace --stats --output-format csv -d OUT/trace -o run -- python3 bench.py --no-cpu --no-extra
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT/fetch -o run -- python3 v1-1_amd/sincdvr/profiling.py run
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT/write -o run -- python3 v1-1_amd/sincdvr/profiling.py run
    python3 v1-1_amd/sincdvr/profiling.py summarize OUT profiles
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
KERNELS = ("k_xfwd", "k_yz", "k_xinv")


def run(precision, steps):
    import numpy as np
    import torch

    from sincdvr import SincDVR, inputs
    cfg = inputs.CONFIGS["c3_81"]
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
        f.write("    rocprofv3 --kernel-trace --stats --output-format csv -d OUT/trace -o run -- python3 bench.py --no-cpu --no-extra\n")
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


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--precision", type=int, default=0)
    r.add_argument("--steps", type=int, default=3)
    sm = sub.add_parser("summarize")
    sm.add_argument("out_dir")
    sm.add_argument("prof_dir")
    sm.add_argument("--tag", default="pairpot_81_n16")
    a = ap.parse_args()
    if a.cmd == "run":
        run(a.precision, a.steps)
    else:
        print("\n".join(summarize(a.out_dir, a.prof_dir, a.tag)))


if __name__ == "__main__":
    main()
