"""MI355X-native (gfx950) ridge detection: choose / evolve / decimate in brute-force and
tiled-tree forms behind a C ABI (include/rdhip.h, librdhip.so).

The C ABI is the product; `ridgedetection_amd._lib.Context` is a thin ctypes binding used by
the tests and the benchmark.  `ridgedetection_amd.synth` holds seeded cloud generators.
"""

__all__ = ["synth"]
