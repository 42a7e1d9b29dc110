"""Writes tests/golden/eval_gates.npz from the reference's gate study,
test/eval_gates.py (in the reference tree).

That script evaluates the plugin on 156 synthetic 128x128 scenes, each under
three 4:2:0 degradations (box, triangle, bicubic), and records the results in
test/results/eval_gates_latest.md.  It also holds NumPy models of two
heuristic maps of the Recon path: ms_map (the mutual-structure gate,
maps.cpp:831-948) and q_maps (the algo-6 affine credibility, algos.cpp
1382-1526).  Neither map has a closed form or invariant that pins it, so this
script writes, with the reference's own functions:

* the SHA-256 of every scene (Y, U, V as float32, band mask) and of every
  degraded chroma pair (U, V at 4:2:0 as the plugin receives them, float32),
  which pin the tests' own scene generator (tests/eval_gates_scene.py);
* for a subset of observations, ms_map, the trace of its structure tensor
  and q_maps' q_prod at the chroma samples of the evaluation band.

Only the module's NumPy function definitions are loaded (parsed with ast and
executed on their own); the module itself, which imports VapourSynth and
loads the plugin, is never imported.  Run in a container that holds the
reference tree:

    python3 tests/golden/make_gate_fixtures.py [/root/reference]
"""
import ast
import hashlib
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
FUNCS = ("edge_sample", "gen_samples", "sited_box", "_axis_weights", "degrade_kernel", "sobel", "box3", "bil",
         "normals_from", "band_chroma", "ms_map", "local_mean", "q_maps")
KINDS = ("box", "tri", "bic")


def load_reference_functions():
    path = os.path.join(REF, "test", "eval_gates.py")
    tree = ast.parse(open(path).read(), path)
    defs = [n for n in tree.body if isinstance(n, ast.FunctionDef) and n.name in FUNCS]
    assert sorted(n.name for n in defs) == sorted(FUNCS), [n.name for n in defs]
    W, H = 128, 128
    YY, XX = np.mgrid[0:H, 0:W]
    ns = {"np": np, "W": W, "H": H, "YY": YY, "XX": XX}
    exec(compile(ast.Module(body=defs, type_ignores=[]), path, "exec"), ns)
    return ns


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def vector_subset(names):
    """Every non-edge scene and every fifth edge scene."""
    edge = [i for i, n in enumerate(names) if n.startswith("edge_")]
    other = [i for i, n in enumerate(names) if not n.startswith("edge_")]
    return sorted(set(edge[::5]) | set(other))


def main():
    ref = load_reference_functions()
    samples = ref["gen_samples"]()
    names = [s[0] for s in samples]
    scene_sha, chroma_sha = [], []
    subset = set(vector_subset(names))
    out = {"names": np.array(names), "kinds": np.array(KINDS)}
    vec_keys = []
    for si, (name, Y, U, V, band) in enumerate(samples):
        scene_sha.append(sha(Y.astype(np.float32), U.astype(np.float32), V.astype(np.float32),
                             band.astype(np.uint8)))
        y64, u64, v64 = Y.astype(np.float64), U.astype(np.float64), V.astype(np.float64)
        yc_box = ref["sited_box"](y64)
        bcm = ref["band_chroma"](band)
        row = []
        for kind in KINDS:
            u420 = ref["degrade_kernel"](u64, kind)
            v420 = ref["degrade_kernel"](v64, kind)
            row.append(sha(u420.astype(np.float32), v420.astype(np.float32)))
            if si in subset:
                ms = ref["ms_map"](yc_box, u420, v420, bcm)
                _, _, trace = ref["normals_from"](yc_box)
                q = ref["q_maps"](y64, u420, v420)["q_prod"]
                key = f"s{si}_{kind}"
                out[key + "_bcm"] = bcm
                out[key + "_ms"] = ms[bcm].astype(np.float64)
                out[key + "_trace"] = trace[bcm].astype(np.float64)
                out[key + "_qprod"] = q[bcm].astype(np.float64)
                vec_keys.append(key)
        chroma_sha.append(row)
    out["scene_sha256"] = np.array(scene_sha)
    out["chroma_sha256"] = np.array(chroma_sha)
    out["vector_keys"] = np.array(vec_keys)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "eval_gates.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(names)} scenes, {len(vec_keys)} observations with map vectors")


if __name__ == "__main__":
    main()
