"""Independent problems of uneven size analysed side by side on one GPU.

The reference analyses one structure at a time, point by point (the sweep of
src/PSSFSS.jl:230-305).  Here each problem's points go through the batched sheet
path (fill, LU and solves of a chunk of points per launch sequence), problems are
dealt largest first (unknowns^3 x points) to the first free worker, and each worker
drives a library context of its own (its own stream, workspaces and look-ahead LU
streams) from a host thread.  One worker is the default: with several, every
release of a per-call device temporary (hipFree) waits for the whole device, so
side-by-side contexts stall each other and four workers run at about half the rate
of one.

ChunkTuner picks the points per chunk for each class of unknown count (rounded up
to a multiple of 512) and precision by timing candidate chunk sizes on the first
points of the first problem of that class, and records the choice (optionally in
a JSON file that later runs read back).

merge=True instead puts problems of similar size into shared batches: the sheet
Gblocks of all problems are sorted by unknown count and cut into size classes
(largest / smallest <= class_ratio, equal end-region mode counts), and each class
goes through pssfss_sheets_gsm_batch, whose batches are filled sheet by sheet and
then factored and solved together, every system padded to the batch's largest
unknown count (padding work <= class_ratio^3 - 1 of the smallest member's).
"""
import json
import os
import queue
import threading
import time
from typing import List, Optional

from . import hip
from .analyze import prepare


def _cost(setup) -> float:
    n = max((r.nbf for r in setup.rwgdat if r is not None), default=1)
    return float(n) ** 3 * len(list(setup.points()))


class ChunkTuner:
    """Points per chunk by (precision, unknown-count class), measured once."""

    TIE = 0.95      # a larger chunk within 5 % of the best rate is taken

    def __init__(self, candidates=(16, 32, 64), path: Optional[str] = None):
        self.candidates = tuple(int(c) for c in candidates)
        self.path = path
        self.table = {}
        self.lock = threading.Lock()
        if path and os.path.exists(path):
            with open(path) as f:
                self.table = json.load(f)

    @staticmethod
    def key(nbf: int, precision: str) -> str:
        return f"{precision}:{512 * ((int(nbf) + 511) // 512)}"

    def chunk(self, ctx, setup, precision: str, mem_fraction: float) -> int:
        nbf = max(r.nbf for r in setup.rwgdat if r is not None)
        k = self.key(nbf, precision)
        with self.lock:
            if k in self.table:
                return int(self.table[k]["chunk"])
        # the same first points of the sweep under every candidate chunking, each on
        # freshly uploaded sheets (every spatial sum computed, none from a cache); the
        # warm-up runs the largest candidate, so that no timed run pays for growing the
        # context's workspaces (tens of GB of Z at the largest chunk)
        limit = max(self.candidates)
        rates = {}
        for k_run, c in enumerate((max(self.candidates),) + self.candidates):
            ctx.set_chunk(c, mem_fraction)
            with hip.SheetSolver(ctx, setup, precision) as s:
                s.upload()
                t0 = time.perf_counter()
                res, _ = s.run(limit=limit)
                if k_run > 0:                       # the first run is the warm-up
                    rates[c] = len(res) / (time.perf_counter() - t0)
        # near-ties go to the larger chunk (fewer factorisation chains, whose panel
        # latency a larger batch amortises); single timings differ by a few per cent
        top = max(rates.values())
        best = max(c for c, r in rates.items() if r >= self.TIE * top)
        with self.lock:
            self.table[k] = {"chunk": best, "points_per_s": {str(c): round(r, 3)
                                                             for c, r in rates.items()}}
            if self.path:
                with open(self.path, "w") as f:
                    json.dump(self.table, f, indent=1, sort_keys=True)
        return best


def size_classes(entries, ratio: float):
    """Cut (key, unknowns, mode count) entries into classes of equal mode count whose
    largest / smallest unknown count stays within ratio, each class in order of
    increasing size; returns lists of keys."""
    out = []
    by_modes = {}
    for key, n, nc in entries:
        by_modes.setdefault(nc, []).append((n, key))
    for nc in sorted(by_modes):
        cur, lo = [], None
        for n, key in sorted(by_modes[nc], key=lambda t: t[0]):
            if cur and n > ratio * lo:
                out.append(cur)
                cur, lo = [], None
            if lo is None:
                lo = n
            cur.append(key)
        if cur:
            out.append(cur)
    return out


def _analyze_merged(ctx, setups, precision: str, ratio: float):
    solvers = [hip.SheetSolver(ctx, su, precision) for su in setups]
    try:
        t0 = time.perf_counter()
        reqs = [s.requests() for s in solvers]
        entries = []
        for i, (_, _, rq) in enumerate(reqs):
            for ig, lst in rq.items():
                if not lst:
                    continue
                pt = lst[0][1].point
                entries.append(((i, ig), int(solvers[i].sheet_of(ig).rwgdat.nbf),
                                int(pt.nmodes[0] + pt.nmodes[1])))
        solved = [dict() for _ in setups]
        infos = [[] for _ in setups]
        classes = size_classes(entries, ratio)
        for cls in classes:
            groups = [(solvers[i].sheet_of(ig), [b for _, b in reqs[i][2][ig]]) for i, ig in cls]
            out = hip.sheets_gsm_batch(ctx, groups, precision)
            for (i, ig), (gsms, inf) in zip(cls, out):
                for (ip, _), g in zip(reqs[i][2][ig], gsms):
                    solved[i][(ip, ig)] = g
                infos[i] += inf
        results = []
        for i, s in enumerate(solvers):
            points, live, _ = reqs[i]
            results.append(s.finish(points, live, solved[i], infos[i])[0])
        wall = time.perf_counter() - t0
    finally:
        for s in solvers:
            s.close()
    npts = sum(len(r) for r in results)
    sizes = {k: n for k, n, _ in entries}
    return results, {"seconds": wall, "points": npts, "points_per_s": npts / wall,
                     "classes": [[sizes[k] for k in c] for c in classes]}


def analyze_batch(problems: List[tuple], precision: str = "mixed", workers: int = 1,
                  tuner: Optional[ChunkTuner] = None, chunk: int = 0, setups=None,
                  device: int = 0, merge: bool = False, class_ratio: float = 1.1):
    """Analyze independent problems, `workers` at a time on one GPU.

    problems: (strata, freqs, steering) tuples as analyze takes them; setups: their
    prepare() results if already made.  chunk: points per chunk for every problem
    (0: the library default, unless a tuner chooses).  One worker runs in the calling
    thread on the process context (device 0), whose workspaces, grown once, serve
    every problem; several workers each open a context on `device`.  Returns
    (results per problem in input order, stats: wall seconds, points, per-problem
    seconds and chunks).  merge: problems of similar size share batches (one
    worker, the process context; chunk sizes from the library's memory rule unless
    chunk is given); stats then lists the size classes."""
    ctx0 = hip.get_context()
    if setups is None:
        setups = [prepare(st, fr, sr, classify=lambda sh: hip.face_pair_classes(ctx0, sh))
                  for st, fr, sr in problems]
    if merge:
        ctx0.set_chunk(chunk, 0.45)
        try:
            return _analyze_merged(ctx0, setups, precision, class_ratio)
        finally:
            ctx0.set_chunk(0, 0.0)
    order = sorted(range(len(setups)), key=lambda i: -_cost(setups[i]))
    share = min(0.45, 0.8 / max(1, workers))
    chunks = [chunk] * len(setups)
    if tuner is not None:       # measured alone, before the workers share the GPU
        for i in order:
            chunks[i] = tuner.chunk(ctx0, setups[i], precision, share)
        ctx0.set_chunk(0, 0.0)
    work = queue.Queue()
    for i in order:
        work.put(i)
    results = [None] * len(setups)
    per = [None] * len(setups)
    errors = []

    def run_one(ctx, i):
        c = chunks[i]
        ctx.set_chunk(c, share)
        t1 = time.perf_counter()
        with hip.SheetSolver(ctx, setups[i], precision) as s:
            res, _ = s.run()
        results[i] = res
        per[i] = {"seconds": time.perf_counter() - t1, "chunk": c, "points": len(res)}

    def worker():
        ctx = hip.Context(device)
        try:
            while True:
                try:
                    i = work.get_nowait()
                except queue.Empty:
                    return
                run_one(ctx, i)
        except Exception as e:          # reported after the join
            errors.append(e)
        finally:
            ctx.close()

    t0 = time.perf_counter()
    if workers <= 1:
        try:
            for i in order:
                run_one(ctx0, i)
        finally:
            ctx0.set_chunk(0, 0.0)
    else:
        threads = [threading.Thread(target=worker) for _ in range(workers)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    wall = time.perf_counter() - t0
    if errors:
        raise errors[0]
    npts = sum(len(r) for r in results)
    return results, {"seconds": wall, "points": npts, "points_per_s": npts / wall,
                     "problems": per}
