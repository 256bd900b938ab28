"""Autotuner for the merge passes: which kernel family runs them (direct tree
passes, or LDS-staged group passes) and how many FDMT iterations each pass
fuses.  Candidates are timed with the stage events of the search handle on
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
    for big in range(2, max_depth + 1):
        for kb in range(0, n // big + 1):
            rem = n - big * kb
            for small in range(1, big):
                if rem % small == 0:
                    c.add(tuple([big] * kb + [small] * (rem // small)))
                    c.add(tuple([small] * (rem // small) + [big] * kb))
    # passes of depth 1 only pay off at the ends of a schedule: keep at most two
    return sorted(s for s in c if sum(s) == n and all(1 <= d <= max_depth for d in s) and s.count(1) <= 2)


def tune_search(search, run_once, reps=2, log=None, modes=(False, True)):
    """Times every (staged, schedule) candidate on `search` (a SingleBeamSearch)
    with run_once() issuing one search step; keeps the fastest on the handle
    and returns ({"staged": bool, "schedule": list}, results)."""
    n = sum(search.get_schedule())
    results = []
    search.enable_timing(True)
    try:
        for staged in modes:
            search.set_staged(staged)
            for s in candidate_schedules(n, 4 if staged else 3):
                search.set_schedule(list(s))
                run_once()  # builds the pass tables
                tot = 0.0
                for _ in range(reps):
                    run_once()
                    tot += sum(search.stage_times_ms())
                results.append((tot / reps, staged, s))
                if log:
                    log(f"[tune] {'staged' if staged else 'tree'} {list(s)}: {tot / reps:.3f} ms")
    finally:
        search.enable_timing(False)
    results.sort()
    _, staged, sched = results[0]
    search.set_staged(staged)
    search.set_schedule(list(sched))
    return {"staged": staged, "schedule": list(sched)}, results
