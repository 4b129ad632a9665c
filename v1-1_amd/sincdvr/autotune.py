"""Autotuner for the convolution schedules.

What it picks, by timing on the device (HIP events on the current stream):
  * the circulant half-edge m >= n of each grid, which fixes the radix pair
    P x Q of the two-step line FFTs (e.g. n = 61: m = 63 = 9x7, 64 = 8x8 or
    72 = 9x8) -- any M = 2m >= 2n - 1 gives the same linear convolution, so
    only speed differs;
  * the pair batch (pair potentials per launch of the single-set pipeline);
  * the batch schedule (grouped per-size launches and the number of forked
    streams, or the fused persistent launch and its grid).

Results are cached in a JSON file keyed by device, grid, orbital count and
precision (SDV_TUNE_CACHE, default ~/.cache/sincdvr/tune.json), so a later
process applies them without timing again.
"""
import json
import os

import torch

from .ops import Batch, SincDVR, circulant_sizes

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "sincdvr", "tune.json")


def _time_ms(fn, reps=3, warmup=1):
    for _ in range(warmup):
        fn()
    st = torch.cuda.current_stream()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(st)
    for _ in range(reps):
        fn()
    b.record(st)
    b.synchronize()
    return a.elapsed_time(b) / reps


class TuneCache:
    def __init__(self, path=None):
        self.path = path or os.environ.get("SDV_TUNE_CACHE", DEFAULT_CACHE)
        self.table = {}
        if os.path.exists(self.path):
            try:
                with open(self.path) as f:
                    self.table = json.load(f)
            except (OSError, ValueError):
                self.table = {}

    @staticmethod
    def key(n, norb, precision):
        dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "none"
        return f"{dev}|n={n}|N={norb}|p={precision}"

    def get(self, n, norb, precision):
        return self.table.get(self.key(n, norb, precision))

    def put(self, n, norb, precision, entry):
        self.table[self.key(n, norb, precision)] = entry

    def save(self):
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.table, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def candidate_half_edges(n, extra=2):
    """The smallest supported half-edge >= n and the next `extra` ones."""
    ms = [m for m in circulant_sizes() if m >= n]
    return ms[: 1 + extra]


def tune_context(s: SincDVR, phi, precision=None, half_edges=None, pair_batches=(16, 32, 64, 136),
                 reps=3, cache=None):
    """Time s.twoereduced(1, N, phi) over half-edges x pair batches, leave the
    context on the fastest and return {"m", "pair_batch", "ms", "table"}.
    phi: device tensor (numspf, totpoints)."""
    if precision is not None:
        s.set_precision(precision)
    prec = s.precision
    if cache is not None:
        hit = cache.get(s.n, s.numspf, prec)
        if hit is not None:
            s.set_circulant(hit["m"])
            s.set_pair_batch(hit["pair_batch"])
            return hit
    N = s.numspf
    npairs = N * (N + 1) // 2
    out = torch.empty((N, N, s.totpoints), dtype=torch.complex128, device=phi.device)
    table = []
    best = None
    for m in (half_edges or candidate_half_edges(s.n)):
        s.set_circulant(m)
        seen = set()
        for pb in pair_batches:
            pb = min(pb, npairs)
            if pb in seen:
                continue
            seen.add(pb)
            s.set_pair_batch(pb)
            ms = _time_ms(lambda: s.twoereduced(1, N, phi, out=out), reps)
            table.append({"m": m, "pair_batch": pb, "ms": ms})
            if best is None or ms < best["ms"]:
                best = {"m": m, "pair_batch": pb, "ms": ms}
    s.set_circulant(best["m"])
    s.set_pair_batch(best["pair_batch"])
    best = dict(best, table=table)
    if cache is not None:
        cache.put(s.n, s.numspf, prec, {k: best[k] for k in ("m", "pair_batch", "ms")})
    return best


BATCH_SCHEDULES = (("grouped", 1, 0), ("grouped", 1, 2), ("grouped", 1, 4), ("fused", 1, 0), ("fused", 2, 0))


def tune_batch(sets, precision=1, tune_sizes=True, schedules=BATCH_SCHEDULES, reps=3, cache=None):
    """Tune a batch: for every distinct grid edge pick the circulant half-edge
    on one representative set (applied to every context of that edge), then
    build the Batch and pick its schedule.  sets: [(SincDVR, phi)] as for
    Batch.  Returns (batch, report)."""
    report = {"sizes": {}, "schedules": []}
    if tune_sizes:
        by_n = {}
        for s, p in sets:
            by_n.setdefault(s.n, []).append((s, p))
        for n, group in sorted(by_n.items()):
            rep_s, rep_p = max(group, key=lambda sp: sp[0].numspf)
            r = tune_context(rep_s, rep_p, precision, pair_batches=(10 ** 9,), reps=reps, cache=cache)
            report["sizes"][n] = r["m"]
            for s, _ in group:
                s.set_precision(precision)
                s.set_circulant(r["m"])
    b = Batch(sets, precision=precision)
    best = None
    for sch in schedules:
        b.set_schedule(*sch)
        ms = _time_ms(lambda: b.run(), reps)
        report["schedules"].append({"schedule": list(sch), "ms": ms})
        if best is None or ms < best[1]:
            best = (sch, ms)
    b.set_schedule(*best[0])
    report["best"] = {"schedule": list(best[0]), "ms": best[1]}
    return b, report
