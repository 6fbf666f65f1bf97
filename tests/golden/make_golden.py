"""Write the golden fixtures under tests/golden/ from the compiled reference CPU path.

The outputs come from oracle/_ref/libref.so: the reference's own header-only CPU
implementation (serial build, no RD_USE_OPENMP) behind oracle/ref_driver.cpp, built by
`make -C oracle` where the reference sources are present.  Inputs are generated from
seeded numpy generators and stored in the fixtures as binary, so checking a fixture does
not depend on libm.  Run: python tests/golden/make_golden.py
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import Oracle  # noqa: E402
from ridgedetection_amd import synth  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def brute_case(ref, name, P, r1, r2):
    d = {"P": P, "r1": np.array(r1, P.dtype), "r2": np.array(r2, P.dtype)}
    S0, idx = ref.choose(P, r1)
    d["choose_S"], d["choose_idx"] = S0, idx.astype(np.int32)
    S1, flag, counts, sums, _, _ = ref.evolve_iter(P, S0, r1)
    d["iter_S"], d["iter_flag"] = S1, np.array(flag, np.int32)
    d["iter_counts"], d["iter_sums"] = counts, sums
    SE, _ = ref.evolve(P, S0, r1)
    d["evolve_S"] = SE
    SD, kept = ref.decimate(SE, r2)
    d["decimate_S"], d["decimate_kept"] = SD, kept.astype(np.int32)
    SR, _ = ref.ridge_detection(P, r1, r2)
    d["rd_S"] = SR
    # chain ordering of the ridge (r2 condition) and of the raw chosen set (r1)
    d["order_idx"], d["order_off"] = ref.order(SR, r2)
    d["order0_idx"], d["order0_off"] = ref.order(S0, r1)
    # quality of the ridge against a pattern (every 7th sample stands in for the ideal curve)
    ideal = np.ascontiguousarray(P[::7])
    d["quality_hausdorff"] = np.array(ref.hausdorff(SR, ideal), P.dtype) if len(SR) else np.array(-1, P.dtype)
    if len(SR):
        d["quality_stats"] = ref.distance_stats(SR, ideal)
    np.savez_compressed(os.path.join(OUT, f"brute_{name}.npz"), **d)
    print(f"brute_{name}: n={len(P)} K0={len(S0)} K={len(SR)}")


def nocopy_case(ref, name, S, r2, sizes):
    k, alive = ref.decimate_nocopy(S, r2, sizes)
    np.savez_compressed(os.path.join(OUT, f"nocopy_{name}.npz"), S=S, r2=np.array(r2, S.dtype),
                        sizes=np.asarray(sizes, np.int64), alive=alive, count=np.array(k))
    print(f"nocopy_{name}: K={len(S)} -> {k}")


def tiled_case(ref, name, P, r1, r2, tpd, cap):
    d = {"P": P, "r1": np.array(r1, P.dtype), "r2": np.array(r2, P.dtype),
         "tpd": np.asarray(tpd, np.int64), "cap": np.array(cap, np.int64)}
    for kind in (0, 1, 2):
        leaves = ref.tree(kind, P, r1, cap, tpd)
        d[f"tree{kind}_own_counts"] = np.array([len(l["own"]) for l in leaves], np.int32)
        d[f"tree{kind}_own"] = np.concatenate([l["own"] for l in leaves]).astype(np.int32)
        d[f"tree{kind}_nb_counts"] = np.array([len(l["nb"]) for l in leaves], np.int32)
        d[f"tree{kind}_nb"] = (np.concatenate([l["nb"] for l in leaves]) if leaves else np.zeros(0)).astype(np.int32)
        d[f"tree{kind}_cap"] = np.array([l["cap"] for l in leaves], np.int64)
        d[f"tree{kind}_bounds"] = np.stack([l["bounds"] for l in leaves])
    for alg in (0, 1, 2, 3, 4, 5):
        for refine in (0, 1):
            d[f"alg{alg}_ref{refine}_S"] = ref.tiled(alg, P, r1, r2, cap, tpd, refinement=bool(refine))
    np.savez_compressed(os.path.join(OUT, f"tiled_{name}.npz"), **d)
    print(f"tiled_{name}: n={len(P)} " + " ".join(
        f"{a}:{len(d[f'alg{a}_ref0_S'])}" for a in (0, 1, 2, 4, 5)))


SAMPLES_IO = os.path.join(ROOT, "oracle", "_ref", "ref_samples_io")
REF_DIR = os.path.dirname(SAMPLES_IO)


def _samples_io(*args):
    r = subprocess.run([SAMPLES_IO, *[str(a) for a in args]], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"ref_samples_io {args}: {r.returncode} {r.stderr}")


def _read_loaded(name):
    with open(os.path.join(REF_DIR, name + ".bin"), "rb") as f:
        n, dim = np.frombuffer(f.read(16), dtype=np.int64)
        return np.frombuffer(f.read(), dtype=np.float64).reshape(int(n), int(dim))


def samples_case():
    """Text sample files written by the reference's Samples<T>::saveToFile (float: fixed, 4
    decimals, width 10; double: 12 decimals, width 16) and the values its readers return."""
    g = np.random.Generator(np.random.PCG64(2024))
    base = synth.helix3d(37, seed=21, dtype=np.float64) * 7.0
    special = np.array([[0.0, -0.0, 2.5e-5], [-2.5e-5, 1.00005, -1.00005], [12345678.9, -1e-9, 1e10],
                        [0.123456789012345, -987.654321098765, 3.0]], dtype=np.float64)
    P64 = np.concatenate([base, special, g.normal(0.0, 100.0, (9, 3))])
    d = {}
    # the descriptor as the reference's generator programs write it
    # (tests/test_gpu_samples_set/samples_generator.cu:56): leading space, "--" keys, a
    # trailing newline (so a blank line follows it in the file)
    desc = f" --dim=3 --np={len(P64)} --a=31.4159 --b=0.15  --s=0.05\n"
    for dt, isd, sfx in ((np.float32, 0, "f32"), (np.float64, 1, "f64")):
        P = P64.astype(dt)
        blob = os.path.join(REF_DIR, f"in_{sfx}.bin")
        with open(blob, "wb") as f:
            f.write(np.array([len(P), 3, isd], dtype=np.int64).tobytes() + P.tobytes())
        for name, comment in ((f"samples_{sfx}", desc), (f"samples_plain_{sfx}", "")):
            _samples_io("save", blob, name, *([comment] if comment else []))
            shutil.copy(os.path.join(REF_DIR, name + ".txt"), os.path.join(OUT, name + ".txt"))
        _samples_io("loadgen", f"samples_{sfx}.txt", isd)
        d[f"loadgen_{sfx}"] = _read_loaded(f"samples_{sfx}.txt").astype(dt)
        _samples_io("loadgen", f"samples_{sfx}.txt", isd, 2)
        d[f"loadgen_dim2_{sfx}"] = _read_loaded(f"samples_{sfx}.txt").astype(dt)
        _samples_io("load", f"samples_plain_{sfx}.txt", isd)
        d[f"load_plain_{sfx}"] = _read_loaded(f"samples_plain_{sfx}.txt").astype(dt)
        d[f"P_{sfx}"] = P
    np.savez_compressed(os.path.join(OUT, "samples_io.npz"), **d)
    print(f"samples: {len(P64)} points, text files for f32 and f64")


def main():
    samples_case()
    ref = Oracle("ref")
    for dt, sfx in ((np.float32, "f32"), (np.float64, "f64")):
        brute_case(ref, f"spiral20k_{sfx}", synth.spiral2d(20000, seed=1, dtype=dt), 0.45, 0.9)
        brute_case(ref, f"helix5k_{sfx}", synth.helix3d(5000, seed=5, dtype=dt), 0.2, 0.4)
        brute_case(ref, f"bezier6d_4k_{sfx}", synth.bezier_union(4000, seed=7, dtype=dt), 0.25, 0.5)
        tiled_case(ref, f"helix10k_{sfx}", synth.helix3d(10000, seed=2, dtype=dt), 0.2, 0.4, [3, 3, 3], 800)
        tiled_case(ref, f"spiral10k_{sfx}", synth.spiral2d(10000, seed=3, dtype=dt), 0.45, 0.9, [3, 3], 1500)
        # decimateNoCopy over four uneven lists of evolved chosen points
        P = synth.spiral2d(20000, seed=11, dtype=dt)
        S0, _ = ref.choose(P, 0.3)
        SE, _ = ref.evolve(P, S0, 0.3)
        q = len(SE) // 4
        nocopy_case(ref, f"spiral_{sfx}", SE, 0.6, [q, q + 3, q - 3, len(SE) - 3 * q])
    # tiny edge cases (f32): 1..6 points, collinear / clustered, r = 0
    g = np.random.Generator(np.random.PCG64(123))
    for n in (1, 2, 3, 4, 5, 6):
        P = g.uniform(-1, 1, (n, 2)).astype(np.float32)
        brute_case(ref, f"tiny{n}_f32", P, 0.5, 1.0)
    P = g.uniform(-1, 1, (64, 3)).astype(np.float32)
    brute_case(ref, "r0_f32", P, 0.0, 0.0)


if __name__ == "__main__":
    main()
