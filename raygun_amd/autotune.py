"""Launch-shape autotuner: picks rgrt_tuning (wavefront passes before the
drain, persistent grid sizes per CU, host lag) for one scene by timing its
frames with hipEvents. Coordinate descent: each parameter in turn takes the
value from its candidate list that gives the fastest frame with the others
fixed, for `rounds` rounds. Every candidate is a documented launch shape, so
a sweep only changes timing, never results (the tests render the same frame
under different tunings and compare).

No counterpart in the reference (its pipeline's launch shape is the
driver's).
"""
from .hiptime import HipEvents, time_ms

CANDIDATES = {
    "wavefrontPasses": [2, 4, 6, 8, 12],
    "traceBlocksPerCU": [4, 6, 8],
    "shadeBlocksPerCU": [4, 8, 12],
    "drainBlocksPerCU": [2, 4, 8],
    "hostLag": [1, 2, 4],
}


def tune(ctx, frame, candidates=None, rounds=1, reps=2, log=None, contexts=None, timer=None, confirm=0, confirm_reps=3):
    """frame(): enqueues one frame on ctx. Returns (best tuning, best ms,
    table of (tuning, ms)). Leaves ctx set to the best tuning.

    For frames in flight on several contexts sharing a scene: `contexts`
    lists every context the tuning is applied to, and timer() returns the
    ms per frame of the whole pipeline under the current tuning (frame is
    then unused).

    confirm > 1: the descent keeps whichever shape had the luckiest single
    timing, so afterwards the `confirm` fastest distinct shapes of the table
    are timed again `confirm_reps` times each, in turn, and the lowest mean
    wins (its mean is the returned ms; the table keeps the descent's
    timings)."""
    cand = candidates or CANDIDATES
    ev = HipEvents()
    ctxs = contexts or [ctx]
    measure = timer or (lambda: time_ms(ctx, frame, reps=reps, warmup=1, events=ev))

    def apply(t):
        for c in ctxs:
            c.set_tuning(**t)

    best = dict(ctx.tuning())
    apply(best)
    best_ms = measure()
    table = [(dict(best), best_ms)]
    for _ in range(rounds):
        for name, values in cand.items():
            current = best[name]  # already timed with the other parameters as they are
            for v in values:
                if v == current:
                    continue
                t = dict(best)
                t[name] = v
                apply(t)
                ms = measure()
                table.append((t, ms))
                if log:
                    log(f"[autotune] {t} {ms:.3f} ms")
                if ms < best_ms:
                    best, best_ms = t, ms
            apply(best)
    if confirm > 1:
        finalists, seen = [], set()
        for t, _ in sorted(table, key=lambda e: e[1]):
            key = tuple(sorted(t.items()))
            if key not in seen:
                seen.add(key)
                finalists.append(t)
            if len(finalists) == confirm:
                break
        sums = [0.0] * len(finalists)
        for _ in range(confirm_reps):
            for i, t in enumerate(finalists):
                apply(t)
                sums[i] += measure()
        i = min(range(len(finalists)), key=lambda k: sums[k])
        best, best_ms = finalists[i], sums[i] / confirm_reps
        if log:
            log(f"[autotune] confirmed {best} {best_ms:.3f} ms over {confirm_reps} timings of {len(finalists)} shapes")
        apply(best)
    return best, best_ms, table
