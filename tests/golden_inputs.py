"""Rebuilds the inputs behind each golden fixture (see tests/golden/make_golden.py)."""
import hashlib

import numpy as np

from fixture_inputs import noise_image, u8_image


def make_input(kind, rows, cols, seed, arr):
    if kind == "u8":
        return u8_image(rows, cols, seed).astype(np.float32)
    if kind == "noise":
        return noise_image(rows, cols, seed)
    if kind == "file512":
        return arr["file512_int8"].astype(np.float32)
    raise ValueError(kind)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def tables_sha(d):
    h = hashlib.sha256()
    for k in sorted(d):
        h.update(k.encode())
        h.update(np.ascontiguousarray(d[k]).tobytes())
    return h.hexdigest()


def plan_tables(p):
    """Same canonical table dict as make_golden.plan_tables, from any plan
    object exposing grid/maps/state_shape/df."""
    d = {}
    for i in range(p.niters + 1):
        off, g = p.grid(i)
        m, c = p.maps(i)
        d[f"shape{i}"] = np.asarray(p.state_shape(i), np.int64)
        d[f"gridoff{i}"] = np.asarray(off, np.int64)
        d[f"grid{i}"] = np.asarray(g, np.int64)
        d[f"map{i}"] = np.asarray(m, np.int64).reshape(-1, 7)
        d[f"copy{i}"] = np.asarray(c, np.int64).reshape(-1, 4)
        d[f"df{i}"] = np.array(p.df(i), np.float32)
    return d
