"""Regenerate the committed fixtures of tests/golden/.

reference_cases.json  literal inputs and expected outputs of the reference's own tests
                      (test/lbvh_test.cu, test/approx_math_test.cu), transcribed as data
validate_points.npy   the lbvh_validate ideal gas (gen_validate_points.cpp, libstdc++ mt19937)
oracle_n1000.npz      oracle build/compress/traverse vectors for 1000 points with duplicate
                      Morton codes (a drift check of the oracle itself)

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

S = 0x7FFFFFFF

REFERENCE_CASES = {
    "source": "mphowardlab/neighbor test/lbvh_test.cu and test/approx_math_test.cu",
    "basic": {
        "cite": "test/lbvh_test.cu:17-315",
        "points": [[2.5, 0.0, 0.0], [1.5, 0.0, 0.0], [0.5, 0.0, 0.0]],
        "lo": [0.0, 0.0, 0.0],
        "hi": [1024.0, 1024.0, 1024.0],
        "spheres": [[0.5, 0, 0, 0.5], [1.5, 0, 0, 0.5], [2.5, 0, 0, 0.5], [1.0, 0, 0, 1.0],
                    [2.0, 0, 0, 1.0], [1.5, 0, 0, 1.5], [-0.5, 0, 0, 0.5]],
        "root": 0,
        "parent": [S, 0, 1, 1, 0],
        "left": [1, 2],
        "right": [4, 3],
        "primitive": [2, 1, 0],
        "leaf_x": {"2": [0.5, 0.5], "3": [1.5, 1.5], "4": [2.5, 2.5]},
        "internal_x": {"1": [0.5, 1.5], "0": [0.5, 2.5]},
        "records_x_bins_hip_path": [[0, 0], [0, 511], [0, 1022], [511, 511], [1022, 0]],
        "records_z": [1, 2, -3, -2, -1],
        "records_w": [S, 4, 3, 4, S],
        "hits": [1, 1, 1, 2, 2, 3, 0],
        "max_neigh": 2,
        "list_rows": [[2], [1], [0], [2, 1], [1, 0], [2, 1], []],
        "map": [2, 1, 0],
        "list_rows_mapped": [[0], [1], [2], [0, 1], [1, 2], [0, 1], []],
    },
    "periodic": {
        "cite": "test/lbvh_test.cu:318-388",
        "points": [[1.9, 1.9, 1.9], [0.0, 0.0, 0.0], [-1.9, -1.9, -1.9]],
        "lo": [-2.0, -2.0, -2.0],
        "hi": [2.0, 2.0, 2.0],
        "spheres": [[-1.9, 1.9, 1.9, 0.5], [1.9, -1.9, -1.9, 0.5]],
        "box": [4.0, 4.0, 4.0],
        "hits_self": [0, 0],
        "hits_27": [2, 2],
    },
    "validate": {
        "cite": "test/lbvh_test.cu:391-503",
        "box": [20.0, 15.0, 25.0],
        "N": 7500,
        "rcut": 1.0,
        "check": "hits[i] >= brute-force minimum-image float count",
    },
    "single": {
        "cite": "test/lbvh_test.cu:506-582",
        "points": [[2.5, 0.0, 0.0]],
        "lo": [0.0, 0.0, 0.0],
        "hi": [1024.0, 1024.0, 1024.0],
        "spheres": [[2.5, 0, 0, 0.5], [-0.5, 0, 0, 0.5]],
        "parent": [S],
        "primitive": [0],
        "leaf_x": [2.5, 2.5],
        "records_z": [-1],
        "records_w": [S],
        "hits": [1, 0],
    },
    "approx_math": {
        "cite": "test/approx_math_test.cu:24-131",
        "cases": [
            ["double2float_rd", [1.000000000001], "==", 1.0],
            ["double2float_ru", [1.000000000001], ">", 1.0],
            ["double2float_rd", [-1.000000000001], "<", -1.0],
            ["double2float_ru", [-1.000000000001], "==", -1.0],
            ["fadd_rd", [1.0, 1e-12], "<=", 1.0],
            ["fadd_ru", [1.0, 1e-12], ">", 1.0],
            ["fadd_rd", [-1.0, -1e-12], "<", -1.0],
            ["fadd_ru", [-1.0, -1e-12], ">=", -1.0],
            ["fsub_rd", [1.0, 1e-12], "<", 1.0],
            ["fsub_ru", [1.0, 1e-12], ">=", 1.0],
            ["fsub_rd", [-1.0, -1e-12], "<=", -1.0],
            ["fsub_ru", [-1.0, -1e-12], ">", -1.0],
            ["fmul_rd", [0.1, 0.2], "<=", 0.02],
            ["fmul_ru", [0.1, 0.2], ">", 0.02],
            ["fmul_rd", [-0.1, 0.2], "<", -0.02],
            ["fmul_ru", [-0.1, 0.2], ">=", -0.02],
            ["fdiv_rd", [1.0, 3.0], "<", 1.0 / 3.0],
            ["fdiv_ru", [1.0, 3.0], ">", 1.0 / 3.0],
            ["fdiv_rd", [-1.0, 3.0], "<", -1.0 / 3.0],
            ["fdiv_ru", [-1.0, 3.0], ">", -1.0 / 3.0],
            ["frcp_rd", [3.0], "<", 1.0 / 3.0],
            ["frcp_ru", [3.0], ">", 1.0 / 3.0],
            ["frcp_rd", [-3.0], "<", -1.0 / 3.0],
            ["frcp_ru", [-3.0], ">", -1.0 / 3.0],
            ["fmaf_rd", [2.0, 1.0, 1e-12], "<=", 2.0],
            ["fmaf_ru", [2.0, 1.0, 1e-12], ">", 2.0],
            ["fmaf_rd", [-2.0, 1.0, -1e-12], "<", -2.0],
            ["fmaf_ru", [-2.0, 1.0, -1e-12], ">=", -2.0],
        ],
        "note": "float comparisons against the float32 literal (0.02f, 1.0f, 2.0f); "
                "division/reciprocal cases compare against the double 1.0/3.0",
    },
}


def write_reference_cases():
    with open(os.path.join(HERE, "reference_cases.json"), "w") as f:
        json.dump(REFERENCE_CASES, f, indent=1)


def write_validate_points():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "gen")
        out = os.path.join(d, "pts.f32")
        subprocess.run(["g++", "-O2", "-std=c++14", os.path.join(HERE, "gen_validate_points.cpp"), "-o", exe], check=True)
        subprocess.run([exe, out], check=True)
        pts = np.fromfile(out, dtype=np.float32).reshape(-1, 3)
    np.save(os.path.join(HERE, "validate_points.npy"), pts)


def write_oracle_vectors():
    import oracle
    rng = np.random.default_rng(1234)
    # points on a coarse lattice so that many Morton codes repeat
    pts = (rng.integers(0, 40, size=(1000, 3)).astype(np.float32) * np.float32(0.25))
    lo, hi = [0.0, 0.0, 0.0], [10.0, 10.0, 10.0]
    t = oracle.build(pts, lo, hi)
    rec, rlo, rhi, bins = oracle.compress(t)
    sph = np.concatenate([pts, np.full((1000, 1), 0.6, np.float32)], axis=1)
    images = oracle.images27([10.0, 10.0, 10.0])
    tr = oracle.traverse(1000, (rec, rlo, rhi, bins), sph, images, max_neigh=64)
    np.savez_compressed(os.path.join(HERE, "oracle_n1000.npz"), points=pts, lo=np.float32(lo), hi=np.float32(hi),
                        codes=t["codes"], primitive=t["primitive"], parent=t["parent"], left=t["left"],
                        right=t["right"], node_lo=t["lo"], node_hi=t["hi"], records=rec, root_lo=rlo, root_hi=rhi,
                        bins=bins, spheres=sph, images=images, counts=tr["counts"], nlist=tr["list"],
                        visits=np.uint64(tr["visits"]))


if __name__ == "__main__":
    write_reference_cases()
    write_validate_points()
    write_oracle_vectors()
    print("fixtures written to", HERE)
