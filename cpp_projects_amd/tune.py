"""Autotuner for the merge passes: which kernel family runs them (direct tree
passes, or LDS-staged group passes with or without a staged final pass), the
staged group shape (LDS window, rows per group, tiles per block), and how many
FDMT iterations each pass fuses.  Candidates are timed with the stage events
of the search handle on the real input; the fastest is kept on the handle."""

DEFAULT_SHAPES = ((6144, 16, 16), (5120, 16, 16), (7168, 16, 16), (6144, 16, 32), (6144, 16, 64), (7168, 16, 32),
                  (5120, 16, 32), (6144, 32, 32), (8192, 32, 16), (9216, 32, 16), (8192, 64, 16))


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


def tune_search(search, run_once, reps=2, log=None, modes=(0, 1, 3, 4), shapes=DEFAULT_SHAPES):
    """Times every (staging mode, group shape, schedule) candidate on
    `search` (a SingleBeamSearch) with run_once() issuing one search step;
    keeps the fastest on the handle and returns
    ({"staged": mode, "shape": shape or None, "schedule": list}, results).
    Modes: 0 tree passes, 1 staged passes with a tree final pass, 3 all
    passes staged; 2 and 4 are 1 and 3 with the mean init fused into the
    first pass."""
    n = sum(search.get_schedule())
    results = []
    search.enable_timing(True)
    try:
        for mode in modes:
            for shape in ((None,) if mode == 0 else shapes):
                search.set_staged(mode)
                if shape is not None:
                    search.set_group_shape(*shape)
                for s in candidate_schedules(n, 4 if mode in (3, 4) else 3):
                    search.set_schedule(list(s))
                    run_once()  # builds the pass tables
                    tot = 0.0
                    for _ in range(reps):
                        run_once()
                        tot += sum(search.stage_times_ms())
                    results.append((tot / reps, mode, shape, s))
                    if log:
                        log(f"[tune] mode {mode} shape {shape} {list(s)}: {tot / reps:.3f} ms")
    finally:
        search.enable_timing(False)
    results.sort(key=lambda r: r[0])
    _, mode, shape, sched = results[0]
    search.set_staged(mode)
    if shape is not None:
        search.set_group_shape(*shape)
    search.set_schedule(list(sched))
    return {"staged": mode, "shape": shape, "schedule": list(sched)}, results
