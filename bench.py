"""Single-beam search benchmark on one MI355X.

Workload (one step = one beam through the whole hot path, inputs resident in
HBM): a 4096-channel x 262144-sample uint8 filterbank over 1200-1600 MHz
(N(128,16) clipped to 0..255, 20 injected dispersed pulses, sweeps up to 4096
samples, widths 1-256), normalised with the reference's global mean/std and
outlier-row cleaning, FDMT with dt_max 4095 (4096 DM trials), S/N
normalisation by the transform of the channel mask, and boxcar S/N over
widths 1,2,..,256 with a strict 6-sigma threshold and candidate compaction.

Prints ONE JSON line: DM-trials x samples / s, the merge kernel's achieved
HBM bandwidth against 8 TB/s (HIP events on the launch stream, algorithmic
bytes 4*nsamps*(rows_{i-1}+rows_i) per iteration), and the CPU oracle timed
on a bounded time slice of the same beam.

The "batched" object measures a ragged batch of 64 float32 beams (channel
counts 1024/2048/4096, 2^14..2^18 samples) replayed as one captured HIP graph
per step, next to the same beams run one at a time through the single-beam
search with a count readback after each.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DM-trials×samples/s (FDMT+boxcar) and % of HBM peak at 1 GPU, batched vs plain"
HBM_PEAK_GBS = 8000.0

CFG = dict(f_min=1200.0, f_max=1600.0, nchans=4096, nsamps=1 << 18, dt_max=4095, box_q=8, threshold=6.0,
           npulses=20, seed=2)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(x_u8, ndt, sample_cols, threads):
    """Times the CPU oracle chain (normalise -> FDMT -> mask FDMT -> S/N ->
    boxcar) on the first sample_cols samples of every channel."""
    import numpy as np

    from oracle import orc

    orc.lib().orc_set_num_threads(threads)
    xs = np.ascontiguousarray(x_u8[:, :sample_cols])
    c = CFG
    t0 = time.perf_counter()
    xn, mask, _ = orc.normalise(xs, 1)
    p = orc.Plan(c["f_min"], c["f_max"], c["nchans"], sample_cols, 1.0, c["dt_max"])
    img = p.execute(xn)
    norm = p.execute(mask)
    snr = orc.fdmt_normalization(img, norm)
    cands = orc.boxcar_pow2(snr, c["threshold"], c["box_q"])
    dt = time.perf_counter() - t0
    return ndt * sample_cols / dt, dt, len(cands)


BATCH = dict(nbeams=64, nch_choices=(1024, 2048, 4096), n_lo=1 << 14, n_hi=1 << 18, seed=3, box_q=8, threshold=6.0)


def batch_specs(nbeams):
    """Ragged batch: channel count from {1024, 2048, 4096}, length uniform in
    2^14..2^18 samples, 0-5 pulses, beam b seeded 3 + b; dt_max = nch - 1."""
    import numpy as np

    rng = np.random.default_rng(BATCH["seed"])
    specs = []
    for b in range(nbeams):
        nch = int(rng.choice(BATCH["nch_choices"]))
        ns = int(rng.integers(BATCH["n_lo"], BATCH["n_hi"] + 1))
        npulses = int(rng.integers(0, 6))
        specs.append(dict(nch=nch, ns=ns, npulses=npulses, seed=BATCH["seed"] + b))
    return specs


def live_rows(plan, out_rows):
    """Rows of each FDMT state that some delivered output row depends on (the
    library computes only these)."""
    n = plan.get_niters()
    live = [set() for _ in range(n + 1)]
    live[n] = set(range(out_rows))
    for i in range(n, 0, -1):
        for o, t, h, _ in plan.rows(i):
            if int(o) in live[i]:
                live[i - 1].add(int(t))
                if h >= 0:
                    live[i - 1].add(int(h))
    return [len(v) for v in live]


def plan_rows(nch):
    return plan_rows_for(nch, BATCH["n_lo"], nch - 1)


def alg_bytes(nch, ns, rows):
    """Unfused streaming bytes of one beam (fp32): stats read, init read +
    state-0 write, every merge's input read + output write, boxcar read."""
    merges = sum(rows[i - 1] + rows[i] for i in range(1, len(rows)))
    return 4.0 * ns * (nch + (nch + rows[0]) + merges + rows[-1])


def bench_batch(args):
    """Batched multi-beam search (one captured graph per step) against the
    same beams run one at a time through the single-beam search, each
    followed by its candidate-count readback."""
    import torch

    from cpp_projects_amd import synth
    from cpp_projects_amd.batch import BatchSearch, Beam
    from cpp_projects_amd.search import SingleBeamSearch
    from cpp_projects_amd.tune import tune_search

    specs = batch_specs(args.batch_beams)
    rows = {nch: plan_rows(nch) for nch in {s["nch"] for s in specs}}
    cap = 1 << 16
    log(f"[batch] generating {len(specs)} beams")
    ins, cands, counts = [], [], []
    for s in specs:
        x, _ = synth.f32_filterbank(s["nch"], s["ns"], CFG["f_min"], CFG["f_max"], s["seed"], npulses=s["npulses"],
                                    max_sweep=s["nch"] - 1, width_range=(1.0, 16.0))
        ins.append(x)
        cands.append(torch.empty((cap, 4), dtype=torch.int32, device="cuda"))
        counts.append(torch.zeros(1, dtype=torch.int64, device="cuda"))
    torch.cuda.synchronize()
    mk = dict(in_dtype="f32", box_q=BATCH["box_q"], threshold=BATCH["threshold"], cand_capacity=cap)
    # schedule per channel count, tuned on the longest beam of that count
    sched = {}
    for nch in sorted(rows):
        i = max((k for k, s in enumerate(specs) if s["nch"] == nch), key=lambda k: specs[k]["ns"])
        s = SingleBeamSearch(CFG["f_min"], CFG["f_max"], nch, specs[i]["ns"], nch - 1, **mk)
        s.run(ins[i], None, cands[i], counts[i])
        best, _ = tune_search(s, lambda: s.run(ins[i], None, cands[i], counts[i]), reps=1)
        sched[nch] = best
        log(f"[batch] nch {nch}: schedule {best}")
        del s
    work = sum((s["nch"]) * s["ns"] for s in specs)  # DM trials x samples (ndt = nch)
    tot_bytes = sum(alg_bytes(s["nch"], s["ns"], rows[s["nch"]]) for s in specs)

    # plain: one beam per launch sequence, synchronised per beam
    plain_s, plain_counts = 0.0, []
    for i, s in enumerate(specs):
        sb = SingleBeamSearch(CFG["f_min"], CFG["f_max"], s["nch"], s["ns"], s["nch"] - 1, **mk)
        sb.set_staged(sched[s["nch"]]["staged"])
        if sched[s["nch"]]["shape"]:
            sb.set_group_shape(*sched[s["nch"]]["shape"])
        sb.set_schedule(sched[s["nch"]]["schedule"])
        sb.run(ins[i], None, cands[i], counts[i])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sb.run(ins[i], None, cands[i], counts[i])
        plain_counts.append(int(counts[i].item()))
        plain_s += time.perf_counter() - t0
        del sb
    log(f"[batch] plain: {plain_s * 1e3:.1f} ms")

    beams = [Beam(CFG["f_min"], CFG["f_max"], s["nch"], s["ns"], s["nch"] - 1, ins[i], cands[i], counts[i])
             for i, s in enumerate(specs)]
    b = BatchSearch(beams, in_dtype="f32", box_q=BATCH["box_q"], threshold=BATCH["threshold"],
                    nworkers=args.batch_workers)
    for i, s in enumerate(specs):
        b.set_staged(i, sched[s["nch"]]["staged"])
        if sched[s["nch"]]["shape"]:
            b.set_group_shape(i, *sched[s["nch"]]["shape"])
        b.set_schedule(i, sched[s["nch"]]["schedule"])
    b.launch()  # capture + first replay
    torch.cuda.synchronize()
    batch_counts = [int(c.item()) for c in counts]
    t0 = time.perf_counter()
    for _ in range(args.batch_steps):
        b.launch()
    torch.cuda.synchronize()
    batch_s = (time.perf_counter() - t0) / args.batch_steps
    log(f"[batch] graph: {batch_s * 1e3:.1f} ms per batch")
    t0 = time.perf_counter()
    b.run()
    torch.cuda.synchronize()
    eager_s = time.perf_counter() - t0
    same = batch_counts == plain_counts
    del b
    return {
        "workload": f"{len(specs)} beams, nch in {{1024,2048,4096}}, nsamps U[2^14,2^18], f32 N(0,1) + 0-5 pulses, "
                    "dt_max nch-1, global norm + S/N norm, boxcar widths 1..256 pow2, thr 6",
        "value": work / batch_s,
        "unit": "DM-trials*samples/s",
        "ms_per_batch": batch_s * 1e3,
        "steps": args.batch_steps,
        "plain_value": work / plain_s,
        "plain_ms_per_batch": plain_s * 1e3,
        "eager_ms_per_batch": eager_s * 1e3,
        "speedup_vs_plain": plain_s / batch_s,
        "alg_bytes_per_batch": tot_bytes,
        "effective_GBs": tot_bytes / batch_s / 1e9,
        "frac_of_hbm_peak": tot_bytes / batch_s / 1e9 / HBM_PEAK_GBS,
        "workers": args.batch_workers,
        "candidates": sum(batch_counts),
        "candidate_counts_equal_plain": same,
        "schedules": {str(k): v for k, v in sched.items()},
    }


LARGE = dict(nchans=8192, nsamps=1 << 20, dt_max=8191, box_q=9, threshold=6.0, npulses=50, seed=5, rfi_frac=0.1,
             rfi_sigma=5.0)


def bench_large(args):
    """The large single-GPU share: 8192 ch x 2^20 float32 samples (32 GiB),
    10% of the channels replaced by N(0, 25) interference, 50 pulses (sweep up
    to 8192 samples, widths 1-512), dt_max 8191, widths 1..512."""
    import torch

    from cpp_projects_amd import synth
    from cpp_projects_amd.search import SingleBeamSearch
    from cpp_projects_amd.tune import tune_search

    L = LARGE
    R, N = L["nchans"], L["nsamps"]
    log(f"[large] generating {R}x{N} f32 filterbank")
    x, _ = synth.f32_filterbank(R, N, CFG["f_min"], CFG["f_max"], L["seed"], npulses=L["npulses"],
                                max_sweep=L["dt_max"], width_range=(1.0, 512.0), rfi_frac=L["rfi_frac"],
                                rfi_sigma=L["rfi_sigma"])
    cap = 1 << 24
    s = SingleBeamSearch(CFG["f_min"], CFG["f_max"], R, N, L["dt_max"], in_dtype="f32", box_q=L["box_q"],
                         threshold=L["threshold"], cand_capacity=cap)
    cands = torch.empty((cap, 4), dtype=torch.int32, device="cuda")
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    s.run(x, None, cands, count)
    torch.cuda.synchronize()
    best, _ = tune_search(s, lambda: s.run(x, None, cands, count), reps=1)
    log(f"[large] schedule {best}")
    s.run(x, None, cands, count)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.large_steps):
        s.run(x, None, cands, count)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.large_steps
    ncand = int(count.item())
    ndt = s.ndt
    rows = plan_rows_for(R, N, L["dt_max"])
    b = alg_bytes(R, N, rows)
    del s, cands, count, x
    return {
        "workload": "8192ch x 2^20 f32 (1200-1600 MHz), 10% channels N(0,25), 50 pulses, dt_max 8191, "
                    "global norm + S/N norm, boxcar widths 1..512 pow2, thr 6",
        "value": ndt * N / dt,
        "unit": "DM-trials*samples/s",
        "ms_per_step": dt * 1e3,
        "steps": args.large_steps,
        "alg_bytes_per_step": b,
        "effective_GBs": b / dt / 1e9,
        "frac_of_hbm_peak": b / dt / 1e9 / HBM_PEAK_GBS,
        "candidates": ncand,
        "candidates_per_s": ncand / dt,
        "merge_schedule": best,
    }


def plan_rows_for(nch, ns, dt_max):
    from cpp_projects_amd.fdmt import FDMTGPU

    f = FDMTGPU(CFG["f_min"], CFG["f_max"], nch, ns, 1.0, dt_max)
    return [int(f.state_shape(i)[3]) for i in range(f.get_niters() + 1)]


def load_traffic(kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 counters
    (profiles/merge_traffic.json), or None when the profile covers another
    kernel."""
    path = os.path.join(ROOT, "profiles", "merge_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("traffic_bytes_per_launch") if d.get("kernel") == kernel else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-cols", type=int, default=1 << 17)
    # the defaults are the autotuner's choice for this workload on MI355X
    ap.add_argument("--schedule", type=str, default="3,3,3,3", help="comma-separated merge fusion depths")
    ap.add_argument("--staged", type=int, default=3,
                    help="staging mode 0..4 (0 tree passes, 1 staged passes with a tree final pass, 3 all "
                         "staged; 2, 4 add the fused init)")
    ap.add_argument("--group-shape", type=str, default="6144,16,32",
                    help="staged passes: LDS floats per tile buffer, rows per group, tiles per block")
    ap.add_argument("--tune", action="store_true", help="autotune mode, shape and schedule during warmup instead")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-beam measurement")
    ap.add_argument("--batch-beams", type=int, default=BATCH["nbeams"])
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--batch-workers", type=int, default=4)
    ap.add_argument("--no-large", action="store_true", help="skip the 8192 x 2^20 large-share measurement")
    ap.add_argument("--large-steps", type=int, default=3)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one plain device copy of this size first (counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import torch

    from cpp_projects_amd import synth
    from cpp_projects_amd.search import SingleBeamSearch

    c = CFG
    N, R = c["nsamps"], c["nchans"]
    log(f"[bench] generating {R}x{N} u8 filterbank")
    x, pulses = synth.u8_filterbank(R, N, c["f_min"], c["f_max"], c["seed"], npulses=c["npulses"])
    cap = 1 << 22
    s = SingleBeamSearch(c["f_min"], c["f_max"], R, N, c["dt_max"], in_dtype="u8", box_q=c["box_q"],
                         threshold=c["threshold"], cand_capacity=cap)
    ndt = s.ndt
    dmt = torch.empty((ndt, N), dtype=torch.float32, device="cuda")
    cands = torch.empty((cap, 4), dtype=torch.int32, device="cuda")
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    if args.calib_copy_mib > 0:
        nb = args.calib_copy_mib << 20
        a = torch.empty(nb // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {args.calib_copy_mib} MiB done")
        del a, b
    if not args.tune:  # the measured configuration is also the warmed-up one
        s.set_staged(int(args.staged))
        if int(args.staged):
            s.set_group_shape(*[int(v) for v in args.group_shape.split(",")])
        s.set_schedule([int(v) for v in args.schedule.split(",")])
    log(f"[bench] ndt={ndt}; warmup {args.warmup}")
    for _ in range(args.warmup):
        s.run(x, dmt, cands, count)
    torch.cuda.synchronize()
    shape = None
    if args.tune:
        from cpp_projects_amd.tune import tune_search

        best, _ = tune_search(s, lambda: s.run(x, dmt, cands, count), reps=2, log=log)
        shape = best["shape"]
        staged = int(best["staged"])
        log(f"[bench] tuned schedule {best}")
    else:
        staged = int(args.staged)
        s.set_staged(staged)
        if staged:
            shape = [int(v) for v in args.group_shape.split(",")]
            s.set_group_shape(*shape)
        s.set_schedule([int(v) for v in args.schedule.split(",")])
    s.run(x, dmt, cands, count)
    torch.cuda.synchronize()
    schedule = s.get_schedule()
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.run(x, dmt, cands, count)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = ndt * N * args.steps / (t1 - t0)
    ncand = int(count.item())

    # stage timing with HIP events on the same stream (separate, untimed steps)
    s.enable_timing(True)
    stage_sum = None
    reps = max(3, min(args.steps, 5))
    for _ in range(reps):
        s.run(x, dmt, cands, count)
        st = s.stage_times_ms()
        stage_sum = st if stage_sum is None else [a + b for a, b in zip(stage_sum, st)]
    s.enable_timing(False)
    stage_ms = [v / reps for v in stage_sum]
    sbytes = s.stage_bytes()  # per pass: compulsory I/O, one read of its input state + one write
    npass = len(schedule)
    # algorithmic bytes of the unfused model: 4*N*(R_{i-1} + R_i) per iteration,
    # credited to the pass that performs iteration i
    from cpp_projects_amd.fdmt import FDMTGPU

    plan = FDMTGPU(c["f_min"], c["f_max"], R, N, 1.0, c["dt_max"])
    rows = [int(plan.state_shape(i)[3]) for i in range(sum(schedule) + 1)]
    live = live_rows(plan, ndt)
    alg, io, i0 = [], [], 1
    for L in schedule:
        alg.append(sum(4.0 * N * (rows[i - 1] + rows[i]) for i in range(i0, i0 + L)))
        # compulsory traffic: the live rows of the pass's input and output states
        io.append(4.0 * N * (live[i0 - 1] + live[i0 + L - 1]))
        i0 += L
    pass_ms = stage_ms[3:3 + npass]
    fused_init = staged in (2, 4) and npass >= 2 and stage_ms[2] < 0.05
    if fused_init:
        # the first pass also performs the init: u8 input read, no state 0
        alg[0] += 4.0 * N * (R + rows[0])
        io[0] = 1.0 * N * R + 4.0 * N * live[schedule[0]]

    def kname(k):
        L, last = schedule[k], k == npass - 1
        if k == 0 and fused_init:
            return f"k_merge_group_init<{L}, 1, true>"
        if staged and not (last and staged in (1, 2) and L <= 3):
            # a staged pass is two launches: <L, SNR, true> over the whole
            # tiles past every row's shifts, <L, SNR, false> over the rest
            return f"k_merge_group<{L}, {'true' if last else 'false'}, *>"
        return f"k_merge_fused<{L}, true, {'true' if last else 'false'}>"

    # dominant kernel: the non-final passes of the kernel instantiation with
    # the most time (the final pass carries the S/N division)
    by_kernel = {}
    for k in range(npass - 1):
        by_kernel.setdefault(kname(k), []).append(k)
    if not by_kernel:
        by_kernel = {kname(npass - 1): [npass - 1]}
    dom = max(by_kernel, key=lambda d: sum(pass_ms[k] for k in by_kernel[d]))
    ks = by_kernel[dom]
    avg_ms = sum(pass_ms[k] for k in ks) / len(ks)
    avg_bytes = sum(alg[k] for k in ks) / len(ks)
    avg_io = sum(io[k] for k in ks) / len(ks)
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    traffic = load_traffic(dom)
    names = (["stats", "mask_fdmt", "init"] + [f"pass{k}(L={schedule[k]})" for k in range(npass - 1)] +
             [f"pass{npass - 1}(L={schedule[-1]})+snr"] + ["snr_norm_separate", "boxcar"])
    stages = {n: round(v, 4) for n, v in zip(names, stage_ms)}
    total_alg = sum(sbytes[:2 if fused_init else 3]) + sum(alg) + sum(sbytes[3 + npass:])

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DM-trials*samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "single-beam search 4096ch x 262144 u8 (1200-1600 MHz), dt_max 4095, "
                        "global norm + S/N norm, boxcar widths 1..256 pow2, thr 6",
            "ndt": ndt, "nsamps": N, "nchans": R, "candidates": ncand, "path": "plain (one beam per launch sequence)",
            "merge_schedule": schedule,
            "staged_passes": staged,
            "group_shape": shape if staged else None,
            "live_rows": live,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_GBs": traffic / (avg_ms * 1e-3) / 1e9 if traffic else None,
            "traffic_frac": traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None,
            "alg_bytes_per_launch": avg_bytes,
            "avg_launch_ms": avg_ms,
            "launches_per_step": len(ks),
            "compulsory_io_bytes_per_launch": avg_io,
            "compulsory_io_GBs": avg_io / (avg_ms * 1e-3) / 1e9,
            "note": "achieved credits the unfused byte model (every state row read and written once per "
                    "iteration); traffic is the counter-measured HBM bytes of one launch (rocprofv3, "
                    "profiles/merge_traffic.json); for a staged pass a 'launch' is its two kernel launches "
                    "(fast-path tiles and the first tiles), timed together by the stage events",
        },
        "stages_ms": stages,
        "effective_GBs_whole_step": total_alg / (sum(stage_ms) * 1e-3) / 1e9,
    }
    xs = x[:, : args.cpu_sample_cols].cpu().numpy() if not args.no_cpu_baseline else None
    if not args.no_batch:
        del s, dmt, cands, count, x
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        out["batched"] = bench_batch(args)
        torch.cuda.empty_cache()
    if not args.no_large:
        if args.no_batch:
            del s, dmt, cands, count, x
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        out["large"] = bench_large(args)
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"[bench] cpu baseline on {R}x{args.cpu_sample_cols} with {threads} threads")
        v, dt, nc = cpu_baseline(xs, ndt, args.cpu_sample_cols, threads)
        out["cpu_baseline"] = {"value": v, "unit": "DM-trials*samples/s", "cores": threads, "kind": "port",
                               "sample": f"{R}ch x {args.cpu_sample_cols} samples (first slice of the beam), "
                                         f"same chain, {dt:.2f} s"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
