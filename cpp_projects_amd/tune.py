"""Autotuner for the merge schedule (how many FDMT iterations each launch
fuses).  Candidates are timed with the stage events of the search handle on
the real input; the fastest is kept on the handle."""


def candidate_schedules(n, max_depth=3):
    c = set()
    for d in range(1, max_depth + 1):
        s = [d] * (n // d)
        r = n - sum(s)
        if r:
            c.add(tuple(s + [r]))
            c.add(tuple([r] + s))
        else:
            c.add(tuple(s))
    for k3 in range(0, n // 3 + 1):
        rem = n - 3 * k3
        if rem % 2 == 0:
            c.add(tuple([3] * k3 + [2] * (rem // 2)))
            c.add(tuple([2] * (rem // 2) + [3] * k3))
        c.add(tuple([3] * k3 + [1] * rem))
        c.add(tuple([1] * rem + [3] * k3))
    return sorted(s for s in c if sum(s) == n and all(1 <= d <= max_depth for d in s))


def tune_search(search, run_once, reps=2, log=None):
    """Times every candidate schedule on `search` (a SingleBeamSearch) with
    run_once() issuing one search step; keeps and returns the fastest."""
    n = sum(search.get_schedule())
    results = []
    search.enable_timing(True)
    try:
        for s in candidate_schedules(n):
            search.set_schedule(list(s))
            run_once()  # builds the pass tables
            tot = 0.0
            for _ in range(reps):
                run_once()
                st = search.stage_times_ms()
                tot += sum(st)
            results.append((tot / reps, s))
            if log:
                log(f"[tune] {list(s)}: {tot / reps:.3f} ms")
    finally:
        search.enable_timing(False)
    results.sort()
    best = list(results[0][1])
    search.set_schedule(best)
    return best, results
