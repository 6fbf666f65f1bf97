"""Text sample files in the reference's format (ridge_detection/rd/utils/samples_set.hpp).

Writing, rd::Samples<T>::saveToFile (samples_set.hpp:342-372; double specialisation
:474-503): an optional first line "%" + comment, then one line per sample with every
coordinate printed fixed-point, right-aligned and followed by two spaces -- 4 decimals in a
field of 10 for float, 12 decimals in a field of 16 for double.  So float files round values
to 1e-4 on the way out, double files to 1e-12.

Reading: loadFromFile (samples_set.hpp:254-291) takes a file of samples only, the dimension
being the number of values on its first line; loadGeneratedSamples (:147-241) takes a file
whose first line is a "% --dim=.. --np=.. --a=.. --b=.. --s=.." descriptor, requires those keys
(b only for dim 2 and 3), optionally keeps only the first `downgradeToDim` coordinates, and
needs exactly np samples.  Both skip empty lines and parse every value as the value type
(decimal to float rounding for float files, not decimal to double to float).

`load_samples` accepts either kind of file (it skips every "%" line; the reference's
loadFromFile cannot read a file whose first line is a comment) and raises ValueError where the
reference's readers throw.
"""

from __future__ import annotations

from fractions import Fraction

import numpy as np


def _fmt(dtype) -> str:
    return "{:16.12f}  " if np.dtype(dtype) == np.float64 else "{:10.4f}  "


def save_samples(path: str, P, comment: str = "") -> None:
    """Write the (n, dim) samples P as Samples<T>::saveToFile does for T = P's dtype (float64
    gives the double format, anything else the float format)."""
    P = np.asarray(P)
    if P.ndim != 2:
        raise ValueError("samples must be an (n, dim) array")
    fmt = _fmt(P.dtype)
    with open(path, "w") as f:
        if comment:
            f.write("%" + comment + "\n")
        for row in P:
            f.write("".join(fmt.format(float(v)) for v in row) + "\n")


def _parse(tokens, dtype) -> np.ndarray:
    """Decimal strings to `dtype`, each rounded once to the nearest value (ties to even)."""
    x = np.array(tokens, dtype=np.float64)
    if np.dtype(dtype) != np.float32:
        return x.astype(dtype)
    f = x.astype(np.float32)
    # decimal -> double -> float rounds twice; the result can differ from a single rounding
    # only where the double lands exactly on the midpoint of two adjacent floats
    lo = f.astype(np.float64)
    nb = np.nextafter(f, np.where(x > lo, np.float32(np.inf), np.float32(-np.inf))).astype(np.float64)
    mid = np.isfinite(lo) & np.isfinite(nb) & (x != lo) & (x == (lo + nb) / 2)
    for i in np.flatnonzero(mid):
        exact = Fraction(tokens[i])
        a, b = Fraction(float(lo[i])), Fraction(float(nb[i]))
        da, db = abs(exact - a), abs(exact - b)
        if da != db:
            f[i] = np.float32(lo[i] if da < db else nb[i])
        # an exact tie keeps the double-rounded value: ties-to-even either way
    return f


def _sample_lines(path: str):
    with open(path) as fh:
        for line in fh:
            s = line.strip()
            if s and not s.startswith("%"):
                yield s.split()


def load_samples(path: str, dtype=np.float32, dim: int | None = None) -> np.ndarray:
    """Samples of a text file as an (n, dim) array.  dim defaults to the number of values on
    the first sample line; longer lines keep their first `dim` values."""
    toks = []
    n = 0
    for vals in _sample_lines(path):
        if dim is None:
            dim = len(vals)
        if len(vals) < dim:
            raise ValueError(f"{path}: a sample line has {len(vals)} values, expected {dim}")
        toks.extend(vals[:dim])
        n += 1
    if dim is None:
        return np.zeros((0, 0), dtype=dtype)
    return np.ascontiguousarray(_parse(toks, dtype).reshape(n, dim))


def read_descriptor(path: str) -> dict:
    """Key/value pairs of a "% --dim=3 --np=1000 --a=.." descriptor line (numbers as float);
    keys with one or two leading dashes are accepted."""
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


def load_generated_samples(path: str, dtype=np.float32, downgrade_to_dim: int = 0):
    """A descriptor file as rd::Samples<T>::loadGeneratedSamples reads it: returns (P, descriptor).
    P has descriptor np samples of descriptor dim (or downgrade_to_dim, when smaller)."""
    d = read_descriptor(path)
    for key in ("dim", "np", "a"):
        if key not in d:
            raise ValueError(f'Missing "{key}" parameter in descriptor!')
    dim = int(d["dim"])
    if dim in (2, 3) and "b" not in d:
        raise ValueError('Missing "b" parameter in descriptor!')
    if "s" not in d:
        raise ValueError('Missing "s" parameter in descriptor!')
    if downgrade_to_dim and downgrade_to_dim < dim:
        dim = int(downgrade_to_dim)
    P = load_samples(path, dtype=dtype, dim=dim)
    if len(P) != int(d["np"]):
        raise ValueError("Incorrect points count read!")
    return P, d
