"""Writes the golden fixtures under tests/golden/ from the reference build.

Run in a checkout that has the reference sources: it builds oracle/_ref (the
reference's own CPU FDMT classes compiled from source by oracle/Makefile),
runs them on deterministic inputs and stores plan tables, outputs (full, or a
SHA-256 of the fp32 bytes for the larger cases) and the reference's own
clustering fixture files.  Usage:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import orc  # noqa: E402
from fixture_inputs import noise_image, u8_image  # noqa: E402

REF = os.environ.get("REFROOT", "/root/reference")

# name, fmin, fmax, nchans, nsamps, dt_max, dt_step, dt_min, input, seed, full
FDMT_CASES = [
    ("c64", 1200.0, 1600.0, 64, 1024, 63, 1, 0, "u8", 11, True),
    ("c100f", 1200.0, 1600.0, 100, 512, 99, 1, 0, "noise", 12, True),
    ("c37odd", 400.0, 800.0, 37, 300, 36, 1, 0, "noise", 13, True),
    ("c512file", 1200.0, 1600.0, 512, 512, 255, 1, 0, "file512", 0, False),
    ("c1000odd", 1200.0, 1600.0, 1000, 2048, 999, 1, 0, "u8", 14, False),
    ("c1024", 1200.0, 1600.0, 1024, 4096, 1023, 1, 0, "u8", 15, False),
    ("c256step2", 1200.0, 1600.0, 256, 2048, 255, 2, 10, "noise", 16, False),
    ("c3000", 1100.0, 1700.0, 3000, 4096, 2047, 1, 0, "u8", 17, False),
]

# name, fmin, fmax, nchan, cols, imaxdt, input, seed, full
CFDMT_CASES = [
    ("lc64", 1200.0, 1600.0, 64, 1024, 64, "u8", 11, True),
    ("lc33", 400.0, 800.0, 33, 300, 50, "noise", 18, True),
    ("lc512file", 1200.0, 1600.0, 512, 512, 256, "file512", 0, False),
    ("lc1000", 1200.0, 1600.0, 1000, 2048, 1000, "u8", 19, False),
]


def file512():
    p = os.path.join(REF, "fdmt_pr_restr", "file512.bin")
    return np.fromfile(p, dtype=np.float32).reshape(512, 512)


def make_input(kind, rows, cols, seed):
    if kind == "u8":
        return u8_image(rows, cols, seed).astype(np.float32)
    if kind == "noise":
        return noise_image(rows, cols, seed)
    if kind == "file512":
        return file512()
    raise ValueError(kind)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def plan_tables(r):
    d = {}
    for i in range(r.niters + 1):
        off, g = r.grid(i)
        m, c = r.maps(i)
        d[f"shape{i}"] = r.state_shape(i)
        d[f"gridoff{i}"] = off
        d[f"grid{i}"] = g
        d[f"map{i}"] = m
        d[f"copy{i}"] = c
        d[f"df{i}"] = np.array(r.df(i), np.float32)
    return d


def tables_sha(d):
    h = hashlib.sha256()
    for k in sorted(d):
        h.update(k.encode())
        h.update(np.ascontiguousarray(d[k]).tobytes())
    return h.hexdigest()


def main():
    orc.build(ref=True)
    manifest = {"fdmt": {}, "cfdmt": {}}
    arrays = {}
    for (name, fmin, fmax, nch, ns, dtm, step, dmin, kind, seed, full) in FDMT_CASES:
        x = make_input(kind, nch, ns, seed)
        r = orc.RefFDMT("pravir", fmin, fmax, nch, ns, 1.0, dtm, step, dmin)
        out = r.execute(x)
        tabs = plan_tables(r)
        ent = {"fmin": fmin, "fmax": fmax, "nchans": nch, "nsamps": ns, "dt_max": dtm, "dt_step": step,
               "dt_min": dmin, "input": kind, "seed": seed, "niters": int(r.niters),
               "out_shape": list(out.shape), "out_sha256": sha(out), "plan_sha256": tables_sha(tabs),
               "dm_grid_sha256": sha(r.dm_grid())}
        if step == 1:
            rr = orc.RefFDMT("restrict", fmin, fmax, nch, ns, 1.0, dtm, step, dmin)
            ent["restrict_out_sha256"] = sha(rr.execute(x))
        if full:
            arrays[f"{name}__out"] = out
            for k, v in tabs.items():
                arrays[f"{name}__{k}"] = v
        manifest["fdmt"][name] = ent
        print(name, out.shape, ent["out_sha256"][:12], flush=True)
    for (name, fmin, fmax, nch, cols, imax, kind, seed, full) in CFDMT_CASES:
        x = make_input(kind, nch, cols, seed)
        r = orc.RefCFdmt(fmin, fmax, nch, cols, imax)
        out = r.process(x)
        ones = r.process(None, True)
        ent = {"fmin": fmin, "fmax": fmax, "nchan": nch, "cols": cols, "imaxdt": imax, "input": kind,
               "seed": seed, "niter": int(r.niter), "out_sha256": sha(out), "ones_sha256": sha(ones)}
        cfg = {}
        for i in range(r.niter + 1):
            cs, fr = r.config(i)
            cfg[f"cumsum{i}"] = cs
            cfg[f"freq{i}"] = fr
        ent["config_sha256"] = tables_sha(cfg)
        if full:
            arrays[f"{name}__out"] = out
            arrays[f"{name}__ones"] = ones
            for k, v in cfg.items():
                arrays[f"{name}__{k}"] = v
        manifest["cfdmt"][name] = ent
        print(name, out.shape, ent["out_sha256"][:12], flush=True)
    # the reference's data file, stored losslessly as int8 (values -14..24)
    f5 = file512()
    assert np.array_equal(f5, np.round(f5)) and f5.min() >= -128 and f5.max() <= 127
    arrays["file512_int8"] = f5.astype(np.int8)
    np.savez_compressed(os.path.join(HERE, "fdmt_golden.npz"), **arrays)
    with open(os.path.join(HERE, "fdmt_golden.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    for fn in ("Candidates1.csv", "Candidates2.csv"):
        shutil.copyfile(os.path.join(REF, "clusterization_forpravir", fn), os.path.join(HERE, fn))
    print("wrote", os.path.join(HERE, "fdmt_golden.npz"))


if __name__ == "__main__":
    main()
