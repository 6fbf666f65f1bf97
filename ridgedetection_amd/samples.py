"""Text sample files in the reference's format (ridge_detection/rd/utils/samples_set.hpp).

Writing (samples_set.hpp:342-372): an optional first line "%" + comment, then one line per
sample with every coordinate printed fixed-point with 4 decimals, right-aligned in 10
characters and followed by two spaces -- so values are rounded to 1e-4 on the way out.
Reading (samples_set.hpp:254-291, 444-462): the dimension is the number of values on the
first sample line; empty lines are skipped.  Lines starting with "%" (the comment written
above, or the "-dim=... -np=..." descriptor of generated-sample files,
samples_set.hpp:147-241) are not samples; `read_descriptor` parses the latter.
"""

from __future__ import annotations

import numpy as np


def save_samples(path: str, P, comment: str = "") -> None:
    P = np.asarray(P)
    if P.ndim != 2:
        raise ValueError("samples must be an (n, dim) array")
    with open(path, "w") as f:
        if comment:
            f.write("%" + comment + "\n")
        for row in P:
            f.write("".join(f"{float(v):10.4f}  " for v in row) + "\n")


def load_samples(path: str, dtype=np.float32) -> np.ndarray:
    rows = []
    dim = None
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("%"):
                continue
            vals = [float(t) for t in s.split()]
            if dim is None:
                dim = len(vals)
            elif len(vals) < dim:
                raise ValueError(f"{path}: a sample line has {len(vals)} values, expected {dim}")
            rows.append(vals[:dim])
    if dim is None:
        return np.zeros((0, 0), dtype=dtype)
    return np.ascontiguousarray(np.array(rows, dtype=dtype))


def read_descriptor(path: str) -> dict:
    """Key/value pairs of a "%-dim=3 -np=1000 -a=..." descriptor line (numbers as float)."""
    with open(path) as f:
        first = f.readline().strip()
    out = {}
    if not first.startswith("%"):
        return out
    for tok in first[1:].split():
        tok = tok.lstrip("-")
        if "=" in tok:
            k, v = tok.split("=", 1)
            try:
                out[k] = float(v)
            except ValueError:
                out[k] = v
    return out
