"""Writes tests/golden/downsample_numpy.npz: outputs of the reference's own
NumPy model of Downsample (test/downsample_reference.py in the reference
tree, numpy only) on two small float 4:4:4 frames, for every siting, kernel
and quality at strength 1 and a few at strength 0.5.  The guided internals
of Downsample (direction classes, gates, candidate scoring) have no closed
form or invariant that fixes them, so these vectors pin the oracle's
restatement.  Run in a container that holds the reference tree:

    python3 tests/golden/make_downsample_vectors.py [/root/reference]
"""
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
sys.path.insert(0, os.path.join(REF, "test"))
from downsample_reference import downsample_reference  # noqa: E402

LOCS = ("left", "center", "topleft", "top", "bottomleft", "bottom")
KERNELS = ("spline36", "lanczos3", "binomial")


def frames():
    # a: one diagonal edge shared by luma and chroma, weak luma texture
    h, w = 16, 20
    yy, xx = np.mgrid[:h, :w]
    side = xx > yy + 2
    a = ((0.2 + 0.5 * side + 0.01 * np.sin(xx)).astype(np.float32),
         (-0.2 + 0.4 * side).astype(np.float32),
         (0.2 - 0.35 * side).astype(np.float32))
    # b: a vertical and a diagonal edge, chroma texture, one isoluminant blob
    h, w = 22, 28
    yy, xx = np.mgrid[:h, :w]
    side = (xx > w // 2) | ((yy < h // 2) & (xx > yy + 5))
    tex = 0.02 * np.sin(xx * 0.9) * np.cos(yy * 0.6)
    blob = ((xx - 6) ** 2 + (yy - 16) ** 2 < 10).astype(np.float64)
    b = ((0.25 + 0.45 * side + tex).astype(np.float32),
         (-0.3 + 0.5 * side + 0.5 * tex + 0.1 * blob).astype(np.float32),
         (0.3 - 0.55 * side - 0.3 * tex).astype(np.float32))
    return {"a": a, "b": b}


def main():
    out = {}
    for name, (y, u, v) in frames().items():
        out[f"{name}_y"], out[f"{name}_u"], out[f"{name}_v"] = y, u, v
        for loc in LOCS:
            for kernel in KERNELS:
                for quality in (0, 1, 2):
                    strengths = (1.0, 0.5) if (loc, kernel) == ("left", "spline36") else (1.0,)
                    for strength in strengths:
                        ru, rv = downsample_reference(y, u, v, quality=quality, kernel=kernel,
                                                      strength=strength, loc=loc)
                        key = f"{name}_{loc}_{kernel}_q{quality}_s{strength:g}"
                        out[key + "_u"] = np.asarray(ru, np.float32)
                        out[key + "_v"] = np.asarray(rv, np.float32)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "downsample_numpy.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays")


if __name__ == "__main__":
    main()
