"""Single-beam search benchmark on one MI355X.

Workload (one step = one beam through the whole hot path, inputs resident in
HBM): a 4096-channel x 262144-sample uint8 filterbank over 1200-1600 MHz
(N(128,16) clipped to 0..255, 20 injected dispersed pulses, sweeps up to 4096
samples, widths 1-256), normalised with the reference's global mean/std and
outlier-row cleaning, FDMT with dt_max 4095 (4096 DM trials), S/N
normalisation by the transform of the channel mask, and boxcar S/N over
widths 1,2,..,256 with a strict 6-sigma threshold and candidate compaction.

Prints ONE JSON line: DM-trials x samples / s; the dominant merge pass's
achieved HBM bandwidth against 8 TB/s (HIP events on the launch stream over
its compulsory bytes: one read of the live rows of the pass's input state and
one write of those of its output state, 4*nsamps bytes per row); a per-stage
table of the same figures; and the CPU oracle chain timed on a bounded time
slice of the same beam.  Every candidate buffer is sized from a sizing run of
its input, and the line reports found and stored counts.

The "batched" object measures a ragged batch of 64 float32 beams (channel
counts 1024/2048/4096, 2^14..2^18 samples) replayed as one captured HIP graph
per step, next to the same beams run one at a time through the single-beam
search with a count readback after each, overall and per channel count.
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

from cpp_projects_amd import workloads  # noqa: E402

CFG = dict(workloads.SINGLE_BEAM, f_min=workloads.F_MIN, f_max=workloads.F_MAX)
BATCH = workloads.BATCH
LARGE = workloads.LARGE


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Marker:
    """A named roctx range around a region (rocprofv3 --marker-trace shows it
    beside the kernel trace); does nothing unless enabled."""

    lib = None

    @classmethod
    def enable(cls):
        import ctypes

        cls.lib = ctypes.CDLL("librocprofiler-sdk-roctx.so")
        cls.lib.roctxRangePushA.argtypes = [ctypes.c_char_p]

    def __init__(self, name):
        self.name = name

    def __enter__(self):
        if Marker.lib is not None:
            Marker.lib.roctxRangePushA(self.name.encode())

    def __exit__(self, *exc):
        if Marker.lib is not None:
            Marker.lib.roctxRangePop()


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


def cpu_reference_fdmt(x_u8, ndt, sample_cols, threads):
    """Times the reference's own FDMTCPU (built from its sources with its
    OpenMP switch, oracle/_ref) on the same slice: one fresh instance, one
    execute of the normalised waterfall.  None when that build is absent."""
    import numpy as np

    from oracle import orc

    if not os.path.exists(os.path.join(orc.REF_DIR, "libref_pravir_omp.so")):
        return None
    c = CFG
    xn, _, _ = orc.normalise(np.ascontiguousarray(x_u8[:, :sample_cols]), 1)
    orc.ref_lib("pravir_omp").ref_fdmt_set_num_threads(threads)
    r = orc.RefFDMT("pravir_omp", c["f_min"], c["f_max"], c["nchans"], sample_cols, 1.0, c["dt_max"])
    t0 = time.perf_counter()
    r.execute(xn)
    dt = time.perf_counter() - t0
    return {"value": ndt * sample_cols / dt, "unit": "DM-trials*samples/s", "cores": threads, "kind": "reference",
            "seconds": dt, "what": "reference FDMTCPU (FDMT_PRAVIR, USE_OPENMP) alone, one execute, same slice"}


def batch_specs(nbeams):
    return workloads.batch_specs(nbeams)


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


def sized_capacity(n):
    """Candidate capacity for a count found by a sizing run (counts are
    deterministic for a given input, so the exact count suffices; a little
    slack keeps the buffer reusable)."""
    return max(1024, int(n) + (int(n) >> 6) + 1024)


def bench_batch(args):
    """Batched multi-beam search (one captured graph per step) against the
    same beams run one at a time through the single-beam search, each
    followed by its candidate-count readback; plus the same comparison for
    the beams of each channel count alone."""
    import torch

    from cpp_projects_amd.batch import BatchSearch, Beam
    from cpp_projects_amd.search import SingleBeamSearch
    from cpp_projects_amd.tune import tune_search

    specs = batch_specs(args.batch_beams)
    log(f"[batch] generating {len(specs)} beams")
    ins = []
    for s in specs:
        ins.append(workloads.batch_beam_input(s))
    torch.cuda.synchronize()
    mk = dict(in_dtype="f32", box_q=BATCH["box_q"], threshold=BATCH["threshold"])
    nchs = sorted({s["nch"] for s in specs})
    # schedule per channel count, tuned on the longest beam of that count
    sched = {}
    tmp_c = torch.empty((1 << 16, 4), dtype=torch.int32, device="cuda")
    tmp_n = torch.zeros(1, dtype=torch.int64, device="cuda")
    for nch in nchs:
        if args.batch_no_tune:  # the autotuner's earlier choice (profiling runs)
            sched[nch] = dict(workloads.BATCH_TUNED[nch])
            continue
        i = max((k for k, s in enumerate(specs) if s["nch"] == nch), key=lambda k: specs[k]["ns"])
        s = SingleBeamSearch(CFG["f_min"], CFG["f_max"], nch, specs[i]["ns"], nch - 1, cand_capacity=1 << 16, **mk)
        s.run(ins[i], None, tmp_c, tmp_n)
        best, _ = tune_search(s, lambda: s.run(ins[i], None, tmp_c, tmp_n), reps=1)
        sched[nch] = best
        log(f"[batch] nch {nch}: schedule {best}")
        del s
    del tmp_c, tmp_n

    # plain: one beam per launch sequence, synchronised per beam; the first
    # (untimed) run of each beam sizes its candidate buffer
    plain_ms, counts_found, cands, counts, comp_bytes = [], [], [], [], []
    for i, s in enumerate(specs):
        sc = sched[s["nch"]]
        cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
        probe = torch.empty((1 << 16, 4), dtype=torch.int32, device="cuda")
        sb = SingleBeamSearch(CFG["f_min"], CFG["f_max"], s["nch"], s["ns"], s["nch"] - 1, cand_capacity=1 << 16, **mk)
        sb.run(ins[i], None, probe, cnt)
        cap = sized_capacity(int(cnt.item()))
        del sb, probe
        sb = SingleBeamSearch(CFG["f_min"], CFG["f_max"], s["nch"], s["ns"], s["nch"] - 1, cand_capacity=cap, **mk)
        sb.set_staged(sc["staged"])
        if sc["shape"]:
            sb.set_group_shape(*sc["shape"])
        sb.set_schedule(sc["schedule"])
        cb = torch.empty((cap, 4), dtype=torch.int32, device="cuda")
        sb.run(ins[i], None, cb, cnt)
        torch.cuda.synchronize()
        with Marker(f"plain_{i}_{s['nch']}x{s['ns']}"):
            t0 = time.perf_counter()
            sb.run(ins[i], None, cb, cnt)
            counts_found.append(int(cnt.item()))
            plain_ms.append((time.perf_counter() - t0) * 1e3)
        comp_bytes.append(sum(sb.stage_bytes()))
        cands.append(cb)
        counts.append(cnt)
        del sb
    plain_s = sum(plain_ms) / 1e3
    caps = [int(c.shape[0]) for c in cands]
    log(f"[batch] plain: {plain_s * 1e3:.1f} ms")

    def make_batch(idx):
        beams = [Beam(CFG["f_min"], CFG["f_max"], specs[i]["nch"], specs[i]["ns"], specs[i]["nch"] - 1, ins[i],
                      cands[i], counts[i]) for i in idx]
        b = BatchSearch(beams, in_dtype="f32", box_q=BATCH["box_q"], threshold=BATCH["threshold"],
                        nworkers=args.batch_workers)
        for k, i in enumerate(idx):
            sc = sched[specs[i]["nch"]]
            b.set_staged(k, sc["staged"])
            if sc["shape"]:
                b.set_group_shape(k, *sc["shape"])
            b.set_schedule(k, sc["schedule"])
        return b

    def time_batch(b, steps, mark="batch"):
        b.launch()  # capture + first replay
        torch.cuda.synchronize()
        with Marker(mark):
            t0 = time.perf_counter()
            for _ in range(steps):
                b.launch()
            torch.cuda.synchronize()
            t = (time.perf_counter() - t0) / steps
        return t

    all_idx = list(range(len(specs)))
    b = make_batch(all_idx)
    batch_s = time_batch(b, args.batch_steps, "batch_all_beams")
    batch_counts = [int(c.item()) for c in counts]
    sorted_ok = True
    for k in all_idx:
        try:
            b.sort_candidates(k)
        except Exception:
            sorted_ok = False
    log(f"[batch] graph: {batch_s * 1e3:.1f} ms per batch")
    t0 = time.perf_counter()
    b.run()
    torch.cuda.synchronize()
    eager_s = time.perf_counter() - t0
    del b
    # per channel count: the beams of one count alone, batched vs plain
    by_nch = {}
    for nch in nchs:
        idx = [i for i in all_idx if specs[i]["nch"] == nch]
        bb = make_batch(idx)
        t = time_batch(bb, args.batch_steps, f"batch_nch_{nch}")
        del bb
        p = sum(plain_ms[i] for i in idx) / 1e3
        by_nch[str(nch)] = {"beams": len(idx), "samples": sum(specs[i]["ns"] for i in idx),
                            "batched_ms": t * 1e3, "plain_ms": p * 1e3, "speedup": p / t,
                            "compulsory_frac": sum(comp_bytes[i] for i in idx) / t / 1e9 / HBM_PEAK_GBS}
    # per length: plain time per beam against its compulsory bytes, and the
    # short and the long beams batched alone
    short = [i for i in all_idx if specs[i]["ns"] < (1 << 16)]
    long_ = [i for i in all_idx if specs[i]["ns"] >= (1 << 16)]
    by_len = {}
    for name, idx in (("short_lt_2e16", short), ("long_ge_2e16", long_)):
        if not idx:
            continue
        bb = make_batch(idx)
        t = time_batch(bb, args.batch_steps, f"batch_{name}")
        del bb
        p = sum(plain_ms[i] for i in idx) / 1e3
        by_len[name] = {"beams": len(idx), "samples": sum(specs[i]["ns"] for i in idx), "batched_ms": t * 1e3,
                        "plain_ms": p * 1e3, "speedup": p / t,
                        "compulsory_frac": sum(comp_bytes[i] for i in idx) / t / 1e9 / HBM_PEAK_GBS,
                        "plain_compulsory_frac": sum(comp_bytes[i] for i in idx) / p / 1e9 / HBM_PEAK_GBS}

    def plain_rate(idx):
        if not idx:
            return None
        return sum(comp_bytes[i] for i in idx) / (sum(plain_ms[i] for i in idx) * 1e-3) / 1e9

    work = sum(s["nch"] * s["ns"] for s in specs)  # DM trials x samples (ndt = nch)
    tot_comp = sum(comp_bytes)
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
        "compulsory_bytes_per_batch": tot_comp,
        "compulsory_GBs": tot_comp / batch_s / 1e9,
        "compulsory_frac": tot_comp / batch_s / 1e9 / HBM_PEAK_GBS,
        "plain_compulsory_frac": tot_comp / plain_s / 1e9 / HBM_PEAK_GBS,
        "by_nchans": by_nch,
        "by_length": by_len,
        "plain_GBs_short_beams_lt_2e16": plain_rate(short),
        "plain_GBs_long_beams": plain_rate(long_),
        "workers": args.batch_workers,
        "candidates_found": sum(batch_counts),
        "cand_capacity_ok": all(n <= c for n, c in zip(batch_counts, caps)) and sorted_ok,
        "candidate_counts_equal_plain": batch_counts == counts_found,
        "schedules": {str(k): v for k, v in sched.items()},
    }


def bench_coherent(args):
    """The coherent front-end on one GUPPI-like chunk (cpp_projects_amd/
    workloads.py COHERENT): 8-bit unpack, forward FFT, then per coherent DM
    the chirp, the sub-band inverse FFTs and the intensity waterfall that the
    search takes.  Compulsory bytes: unpack reads the raw chunk and writes the
    blocks; the forward FFT reads and writes the chunk once; per DM the fused
    pass reads the spectra once and writes the waterfall.  Where the chunk
    does not take the fused pass, per
    DM the chirp reads and writes the chunk, the inverse FFT reads and writes
    it, the intensity pass reads the kept bins and writes the waterfall."""
    import torch

    from cpp_projects_amd import unpack
    from cpp_projects_amd.coherent import CoherentChunk

    C = workloads.COHERENT
    npol, nchan, nbin, nfft, nov = C["npol"], C["nchan"], C["nbin"], C["nfft"], C["noverlap"]
    nsamp = (nbin - 2 * nov) * nfft
    g = torch.Generator(device="cuda")
    g.manual_seed(C["seed"])
    raw = torch.randint(-30, 31, (nchan * nsamp * npol,), dtype=torch.int8, device="cuda", generator=g)
    cc = CoherentChunk(C["f_min"], C["f_max"], npol, nchan, C["len_sft"], nbin, nfft, nov, C["tsamp"])
    blocks = torch.empty(unpack.out_size(nbin, nfft, npol, nchan), dtype=torch.complex64, device="cuda")
    wf = torch.empty((cc.rows, cc.cols), dtype=torch.float32, device="cuda")
    dms = C["dms"]

    def run():
        unpack.unpack_chunk_guppi(raw, nsamp, nov, nbin, nfft, npol, nchan, blocks)
        cc.forward(blocks)
        for dm in dms:
            cc.waterfall(blocks, dm, wf)

    run()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    steps = 3
    t_unpack = t_fwd = t_dm = 0.0
    for _ in range(steps):
        ev[0].record()
        unpack.unpack_chunk_guppi(raw, nsamp, nov, nbin, nfft, npol, nchan, blocks)
        ev[1].record()
        cc.forward(blocks)
        ev[2].record()
        for dm in dms:
            cc.waterfall(blocks, dm, wf)
        ev[3].record()
        torch.cuda.synchronize()
        t_unpack += ev[0].elapsed_time(ev[1]) / steps
        t_fwd += ev[1].elapsed_time(ev[2]) / steps
        t_dm += ev[2].elapsed_time(ev[3]) / steps / len(dms)
    ok = bool(torch.isfinite(wf).all().item()) and float(wf.max().item()) > 0.0
    # the whole of CChunk_gpu::process from the raw bytes: unpack, then the
    # forward FFT and per DM the waterfall, statistics, two CFdmt transforms
    # and the S/N normalisation, then the clustering (which synchronises)
    img = torch.empty((len(dms) * C["len_sft"], cc.cols), dtype=torch.float32, device="cuda")

    def chunk_process():
        unpack.unpack_chunk_guppi(raw, nsamp, nov, nbin, nfft, npol, nchan, blocks)
        return cc.process(blocks, dms, C["sigma_bound"], d_image=img, bwrite=False)

    chunk_process()
    torch.cuda.synchronize()
    t_proc = 0.0
    for _ in range(steps):
        ev[0].record()
        _, _, dele, _ = chunk_process()
        ev[1].record()
        torch.cuda.synchronize()
        t_proc += ev[0].elapsed_time(ev[1]) / steps
    chunk = 8.0 * cc.spectra
    kept = chunk * (nbin // C["len_sft"] - 2 * (nov // C["len_sft"])) / (nbin // C["len_sft"])
    b_unpack = raw.numel() + chunk
    b_fwd = 2 * chunk
    if cc.fused:
        b_dm = chunk + 4.0 * cc.rows * cc.cols
    else:
        b_dm = 2 * chunk + 2 * chunk + kept + 4.0 * cc.rows * cc.cols
    gbs = lambda b, ms: b / (ms * 1e-3) / 1e9
    return {
        "workload": f"{nchan} coarse ch x {C['len_sft']} sub-ch, {npol // 2} pol, {nfft} blocks of {nbin} bins "
                    f"({nov} overlap), 8-bit GUPPI, {len(dms)} coherent DMs -> {cc.rows} x {cc.cols} waterfalls",
        "fused": cc.fused, "ms_per_chunk": t_unpack + t_fwd + t_dm * len(dms), "steps": steps,
        "unpack_ms": t_unpack, "forward_fft_ms": t_fwd, "ms_per_dm": t_dm,
        "waterfall_samples_per_s": cc.rows * cc.cols * len(dms) / ((t_unpack + t_fwd + t_dm * len(dms)) * 1e-3),
        "process_ms_per_chunk": t_proc, "process_delegates": int(dele.numel()),
        "compulsory_GBs": {"unpack": gbs(b_unpack, t_unpack), "forward_fft": gbs(b_fwd, t_fwd),
                           "per_dm": gbs(b_dm, t_dm)},
        "frac_of_peak_per_dm": gbs(b_dm, t_dm) / 8000.0,
        "output_ok": ok,
    }


def bench_large(args):
    """The large single-GPU share: 8192 ch x 2^20 float32 samples (32 GiB),
    10% of the channels replaced by N(0, 25) interference, 50 pulses (sweep up
    to 8192 samples, widths 1-512), dt_max 8191, widths 1..512."""
    import torch

    from cpp_projects_amd.search import SingleBeamSearch
    from cpp_projects_amd.tune import tune_search

    L = LARGE
    R, N = L["nchans"], L["nsamps"]
    log(f"[large] generating {R}x{N} f32 filterbank")
    x, _ = workloads.large_input()
    mk = dict(in_dtype="f32", box_q=L["box_q"], threshold=L["threshold"])
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    probe = torch.empty((1 << 16, 4), dtype=torch.int32, device="cuda")
    s = SingleBeamSearch(CFG["f_min"], CFG["f_max"], R, N, L["dt_max"], cand_capacity=1 << 16, **mk)
    s.run(x, None, probe, count)
    torch.cuda.synchronize()
    if args.large_no_tune:  # the autotuner's earlier choice (profiling runs)
        best = dict(workloads.LARGE_TUNED)
    else:
        best, _ = tune_search(s, lambda: s.run(x, None, probe, count), reps=1)
    log(f"[large] schedule {best}")
    cap = sized_capacity(int(count.item()))
    del s, probe
    s = SingleBeamSearch(CFG["f_min"], CFG["f_max"], R, N, L["dt_max"], cand_capacity=cap, **mk)
    s.set_staged(best["staged"])
    if best["shape"]:
        s.set_group_shape(*best["shape"])
    s.set_schedule(best["schedule"])
    cands = torch.empty((cap, 4), dtype=torch.int32, device="cuda")
    s.run(x, None, cands, count)
    torch.cuda.synchronize()
    with Marker("large_steps"):
        t0 = time.perf_counter()
        for _ in range(args.large_steps):
            s.run(x, None, cands, count)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.large_steps
    ncand = int(count.item())
    n_sorted = s.sort_candidates(cands, count)  # raises on overflow
    ndt = s.ndt
    rows = plan_rows_for(R, N, L["dt_max"])
    b = alg_bytes(R, N, rows)
    comp = sum(s.stage_bytes())
    del s, cands, count, x
    return {
        "workload": "8192ch x 2^20 f32 (1200-1600 MHz), 10% channels N(0,25), 50 pulses, dt_max 8191, "
                    "global norm + S/N norm, boxcar widths 1..512 pow2, thr 6",
        "value": ndt * N / dt,
        "unit": "DM-trials*samples/s",
        "ms_per_step": dt * 1e3,
        "steps": args.large_steps,
        "compulsory_bytes_per_step": comp,
        "compulsory_GBs": comp / dt / 1e9,
        "compulsory_frac": comp / dt / 1e9 / HBM_PEAK_GBS,
        "unfused_equiv_GBs": b / dt / 1e9,
        "candidates_found": ncand,
        "candidates_stored": n_sorted,
        "cand_capacity": cap,
        "cand_capacity_ok": ncand <= cap,
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
    ap.add_argument("--schedule", type=str, default=",".join(map(str, workloads.SINGLE_BEAM_TUNED["schedule"])),
                    help="comma-separated merge fusion depths")
    ap.add_argument("--staged", type=int, default=workloads.SINGLE_BEAM_TUNED["staged"],
                    help="staging mode 0..4 (0 tree passes, 1 staged passes with a tree final pass, 3 all "
                         "staged; 2, 4 add the fused init)")
    ap.add_argument("--group-shape", type=str, default=",".join(map(str, workloads.SINGLE_BEAM_TUNED["shape"])),
                    help="staged passes: LDS floats per tile buffer, rows per group, tiles per block")
    ap.add_argument("--tune", action="store_true", help="autotune mode, shape and schedule during warmup instead")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-beam measurement")
    ap.add_argument("--batch-beams", type=int, default=BATCH["nbeams"])
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--batch-workers", type=int, default=4)
    ap.add_argument("--batch-only", action="store_true", help="run the batched multi-beam measurement alone")
    ap.add_argument("--large-only", action="store_true", help="run the large-share measurement alone")
    ap.add_argument("--large-no-tune", action="store_true",
                    help="large share: take the stored schedule instead of tuning")
    ap.add_argument("--roctx", action="store_true", help="mark the timed batched replays as roctx ranges")
    ap.add_argument("--batch-no-tune", action="store_true",
                    help="batched measurement: take the stored per-channel-count schedules instead of tuning")
    ap.add_argument("--no-large", action="store_true", help="skip the 8192 x 2^20 large-share measurement")
    ap.add_argument("--no-coherent", action="store_true", help="skip the coherent front-end measurement")
    ap.add_argument("--coherent-only", action="store_true", help="run the coherent front-end measurement alone")
    ap.add_argument("--large-steps", type=int, default=3)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one plain device copy of this size first (counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import torch

    from cpp_projects_amd.search import SingleBeamSearch

    if args.roctx:
        Marker.enable()
    if args.batch_only:
        print(json.dumps({"metric": METRIC, "n_gpus": 1, "batched": bench_batch(args)}))
        return
    if args.large_only:
        print(json.dumps({"metric": METRIC, "n_gpus": 1, "large": bench_large(args)}))
        return
    if args.coherent_only:
        print(json.dumps({"metric": METRIC, "n_gpus": 1, "coherent": bench_coherent(args)}))
        return
    c = CFG
    N, R = c["nsamps"], c["nchans"]
    log(f"[bench] generating {R}x{N} u8 filterbank")
    x, pulses = workloads.single_beam_input()
    mk = dict(in_dtype="u8", box_q=c["box_q"], threshold=c["threshold"])
    if args.calib_copy_mib > 0:
        nb = args.calib_copy_mib << 20
        a = torch.empty(nb // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {args.calib_copy_mib} MiB done")
        del a, b
    # sizing run: the candidate count of this input fixes the buffer
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    probe = torch.empty((1 << 16, 4), dtype=torch.int32, device="cuda")
    s = SingleBeamSearch(c["f_min"], c["f_max"], R, N, c["dt_max"], cand_capacity=1 << 16, **mk)
    s.run(x, None, probe, count)
    cap = sized_capacity(int(count.item()))
    del s, probe
    s = SingleBeamSearch(c["f_min"], c["f_max"], R, N, c["dt_max"], cand_capacity=cap, **mk)
    ndt = s.ndt
    dmt = torch.empty((ndt, N), dtype=torch.float32, device="cuda")
    cands = torch.empty((cap, 4), dtype=torch.int32, device="cuda")

    def configure(staged, shape, sched):
        s.set_staged(int(staged))
        if int(staged) and shape:
            s.set_group_shape(*shape)
        s.set_schedule(sched)

    shape = [int(v) for v in args.group_shape.split(",")]
    staged = int(args.staged)
    if not args.tune:  # the measured configuration is also the warmed-up one
        configure(staged, shape, [int(v) for v in args.schedule.split(",")])
    log(f"[bench] ndt={ndt}; candidate capacity {cap}; warmup {args.warmup}")
    for _ in range(args.warmup):
        s.run(x, dmt, cands, count)
    torch.cuda.synchronize()
    if args.tune:
        from cpp_projects_amd.tune import tune_search

        best, _ = tune_search(s, lambda: s.run(x, dmt, cands, count), reps=2, log=log)
        shape, staged = best["shape"], int(best["staged"])
        configure(staged, shape, best["schedule"])
        log(f"[bench] tuned schedule {best}")
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
    n_sorted = s.sort_candidates(cands, count)  # raises CapacityError if the buffer overflowed

    # stage timing with HIP events on the launch stream (separate, untimed steps)
    s.enable_timing(True)
    stage_sum = None
    reps = max(3, min(args.steps, 5))
    for _ in range(reps):
        s.run(x, dmt, cands, count)
        st = s.stage_times_ms()
        stage_sum = st if stage_sum is None else [a + b for a, b in zip(stage_sum, st)]
    s.enable_timing(False)
    stage_ms = [v / reps for v in stage_sum]
    sbytes = s.stage_bytes()  # compulsory bytes per stage (live rows, actual input element size)
    npass = len(schedule)
    # unfused byte model credited to each pass (informational only: fusion and
    # row pruning make it exceed the traffic)
    from cpp_projects_amd.fdmt import FDMTGPU

    plan = FDMTGPU(c["f_min"], c["f_max"], R, N, 1.0, c["dt_max"])
    rows = [int(plan.state_shape(i)[3]) for i in range(sum(schedule) + 1)]
    live = live_rows(plan, ndt)
    unfused, i0 = [], 1
    for L in schedule:
        unfused.append(sum(4.0 * N * (rows[i - 1] + rows[i]) for i in range(i0, i0 + L)))
        i0 += L

    fused0 = staged in (2, 4) and sbytes[2] == 0.0  # the first pass built state 0 itself

    def kname(k):
        L, last = schedule[k], k == npass - 1
        if k == 0 and fused0:
            return f"k_merge_group_init<{L}, 1, true, *>"  # u8 input, normalised
        if staged and not (last and staged in (1, 2) and L <= 3):
            # a staged pass is two launches: the fast kernel over the whole
            # tiles past every row's shifts (the producer/consumer form for
            # windows above 6144 floats), <L, SNR, false> over the rest
            snr = 'true' if last else 'false'
            if shape[0] > 6144:
                return f"k_merge_group_pc<{L}, {snr}> + k_merge_group<{L}, {snr}, false, *>"
            return f"k_merge_group<{L}, {snr}, *>"
        return f"k_merge_fused<{L}, true, {'true' if last else 'false'}>"

    stage_names = (["stats", "mask_transform", "init"] +
                   [f"pass{k}(L={schedule[k]})" + ("+init" if k == 0 and fused0 else "") for k in range(npass - 1)] +
                   [f"pass{npass - 1}(L={schedule[-1]})+snr"] + ["snr_norm_separate", "boxcar"])
    stage_kernels = (["k_row_stats<true> + k_stats_finalize", "k_mask_counts (counted mask transform)",
                      "k_init_consec<1, true>"] + [kname(k) for k in range(npass)] +
                     ["k_snr_normalise", "k_boxcar_pow2_pf<%d>" % min(max(c["box_q"], 8), 9)])
    kernels = {}
    for name, kn, ms, b in zip(stage_names, stage_kernels, stage_ms, sbytes):
        if ms <= 0.0:
            continue
        kernels[name] = {"kernel": kn, "ms": round(ms, 4), "compulsory_bytes": b,
                         "GBs": round(b / (ms * 1e-3) / 1e9, 1),
                         "frac_of_peak": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    # dominant kernel: the non-final passes of the kernel instantiation with
    # the most time (the final pass carries the S/N division)
    pass_ms = stage_ms[3:3 + npass]
    pass_io = sbytes[3:3 + npass]
    by_kernel = {}
    for k in range(npass - 1):
        by_kernel.setdefault(kname(k), []).append(k)
    if not by_kernel:
        by_kernel = {kname(npass - 1): [npass - 1]}
    dom = max(by_kernel, key=lambda d: sum(pass_ms[k] for k in by_kernel[d]))
    ks = by_kernel[dom]
    avg_ms = sum(pass_ms[k] for k in ks) / len(ks)
    avg_io = sum(pass_io[k] for k in ks) / len(ks)
    avg_unfused = sum(unfused[k] for k in ks) / len(ks)
    achieved = avg_io / (avg_ms * 1e-3) / 1e9
    traffic = load_traffic(dom)
    step_comp = sum(sbytes)
    step_ms = sum(stage_ms)

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
            "ndt": ndt, "nsamps": N, "nchans": R, "path": "plain (one beam per launch sequence)",
            "candidates_found": ncand, "candidates_stored": n_sorted, "cand_capacity": cap,
            "cand_capacity_ok": ncand <= cap,
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
            "compulsory_bytes_per_launch": avg_io,
            "avg_launch_ms": avg_ms,
            "launches_per_step": len(ks),
            "unfused_equiv_GBs": avg_unfused / (avg_ms * 1e-3) / 1e9,
            "note": "achieved = compulsory bytes of one pass (one read of the live rows of its input state, "
                    "one write of the live rows of its output state) / the pass's event-timed duration; "
                    "traffic = counter-measured HBM bytes of one pass (rocprofv3, profiles/merge_traffic.json); "
                    "unfused_equiv_GBs credits the unfused byte model (every row of every intermediate state "
                    "read and written once) and is not a fraction of peak; a staged pass is two kernel "
                    "launches (whole tiles past every row's shifts, and the first tiles), timed together",
        },
        "kernels": kernels,
        "step_compulsory_bytes": step_comp,
        "step_compulsory_frac": step_comp / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }
    xs = x[:, : args.cpu_sample_cols].cpu().numpy() if not args.no_cpu_baseline else None
    del s, dmt, cands, count, x
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    if not args.no_batch:
        out["batched"] = bench_batch(args)
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    if not args.no_large:
        out["large"] = bench_large(args)
        torch.cuda.empty_cache()
    if not args.no_coherent:
        out["coherent"] = bench_coherent(args)
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"[bench] cpu baseline on {R}x{args.cpu_sample_cols} with {threads} threads")
        v, dt, nc = cpu_baseline(xs, ndt, args.cpu_sample_cols, threads)
        out["cpu_baseline"] = {"value": v, "unit": "DM-trials*samples/s", "cores": threads, "kind": "port",
                               "cpu_model": cpu_model(), "nproc": os.cpu_count(),
                               "sample": f"{R}ch x {args.cpu_sample_cols} samples (first slice of the beam), "
                                         f"same chain, {dt:.2f} s",
                               "reference_fdmt": cpu_reference_fdmt(xs, ndt, args.cpu_sample_cols, threads)}
    print(json.dumps(out))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
