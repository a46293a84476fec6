"""Device-side pins of the directed rounding (ApproximateMath.h), the bounding volumes
(BoundingVolumes.h) and the compression on degenerate axes, against the oracle.

- tests/device/device_math.bin runs approx::*, BoundingSphere (float and double ctors) and its
  overlap test inside a gfx950 kernel on the 28 probes of the reference's
  test/approx_math_test.cu:24-131, on IEEE special values (+-0, +-denorm_min, +-FLT_MAX, +-inf,
  NaN) crossed with each other, and on random values over 60 decades; every result is compared
  bit for bit with the oracle (NaN compared as NaN: the payload of a NaN that an operation makes
  is not specified).
- Full int4 records (all four words), root box and bin sizes on the reference's basic and single
  cases and on collinear (y = z = 0), coplanar (z = 0) and coincident point sets, where
  fsub_ru(0, 0) is a denormal, bins are 2.9e-39 and floorf gives -1 that saturates to 0
  (kernels/LBVHTraverser.cuh:64-110); counts of the three reference-result traversals as well.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "device", "device_math.bin")
F32 = np.float32

CASE = np.dtype([("a", "<f4"), ("b", "<f4"), ("c", "<f4"), ("pad", "<f4"), ("o", "<f8", 3), ("r", "<f8"),
                 ("lo", "<f4", 3), ("hi", "<f4", 3)])
assert CASE.itemsize == 72
OPS = [("next_down", "a"), ("next_up", "a"), ("fadd_rd", "ab"), ("fadd_ru", "ab"), ("fsub_rd", "ab"), ("fsub_ru", "ab"),
       ("fmul_rd", "ab"), ("fmul_ru", "ab"), ("fdiv_rd", "ab"), ("fdiv_ru", "ab"), ("frcp_rd", "b"), ("frcp_ru", "b"),
       ("fmaf_rd", "abc"), ("fmaf_ru", "abc"), ("double2float_rd", "o"), ("double2float_ru", "o")]


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def nb():
    import torch
    assert torch.cuda.is_available(), "gpu tests need a GPU"
    import neighbor_amd
    neighbor_amd.lib()
    return neighbor_amd


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _cases(golden_dir):
    with open(os.path.join(golden_dir, "reference_cases.json")) as f:
        probes = json.load(f)["approx_math"]["cases"]
    rows = []
    # the reference's 28 probes: their arguments as (a, b, c) and, for the double conversions, o.x
    for name, args, _, _ in probes:
        a = list(args) + [0.0] * (3 - len(args))
        if name.startswith("double2float"):
            rows.append((1.0, 1.0, 0.0, (args[0], 0.5, -0.5), 1.0))
        elif name.startswith("frcp"):
            rows.append((1.0, a[0], 0.0, (0.1, 0.2, 0.3), 2.8))
        else:
            rows.append((a[0], a[1], a[2], (0.1, 0.2, 0.3), 2.8))
    fmax = float(np.finfo(np.float32).max)
    dmin = float(F32(1.4e-45))
    specials = [0.0, -0.0, dmin, -dmin, fmax, -fmax, np.inf, -np.inf, np.nan, 1.0, -1.0, 2.9e-39, -1e-30, 3.0, 0.1]
    for i, a in enumerate(specials):
        for j, b in enumerate(specials):
            c = specials[(i * 7 + j * 3) % len(specials)]
            rows.append((a, b, c, (a if np.isfinite(a) else 0.0, b if np.isfinite(b) else 1.0, 1e-320), 1.0))
    rng = np.random.default_rng(31)
    n = 6000
    mag = 10.0 ** rng.integers(-30, 30, size=(n, 3))
    abc = (rng.standard_normal((n, 3)) * mag).astype(np.float32)
    o = rng.uniform(-1, 1, size=(n, 3)) * 10.0 ** rng.integers(-3, 6, size=(n, 1))
    r = rng.uniform(0, 3, size=n) * 10.0 ** rng.integers(-2, 3, size=n)
    for k in range(n):
        rows.append((abc[k, 0], abc[k, 1], abc[k, 2], tuple(o[k]), r[k]))
    rec = np.zeros(len(rows), CASE)
    for k, (a, b, c, oo, rr) in enumerate(rows):
        rec[k]["a"], rec[k]["b"], rec[k]["c"] = a, b, c
        rec[k]["o"] = oo
        rec[k]["r"] = rr
    # boxes around the double sphere's center: overlapping, touching and apart
    cen = np.nan_to_num(rec["o"], nan=0.0, posinf=0.0, neginf=0.0).astype(np.float32)
    off = rng.standard_normal((len(rows), 3)).astype(np.float32) * np.float32(1.5)
    half = np.abs(rng.standard_normal((len(rows), 3))).astype(np.float32) * np.float32(0.5)
    rec["lo"] = cen + off - half
    rec["hi"] = cen + off + half
    return rec


def _same(exp, got):
    e = np.asarray(exp, np.float32)
    g = np.asarray(got, np.float32)
    return bool((e.view(np.uint32) == g.view(np.uint32)) | (np.isnan(e) & np.isnan(g)))


def test_device_rounding_and_spheres_match_oracle(nb, orc, golden_dir, tmp_path):
    assert os.path.exists(EXE), "tests/device/device_math.bin is built by __graft_entry__.build()"
    rec = _cases(golden_dir)
    rec.tofile(tmp_path / "in.bin")
    log(f"device math probe: {len(rec)} cases")
    p = subprocess.run([EXE, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr
    out = np.fromfile(tmp_path / "out.bin", np.uint32).reshape(-1, 22)
    assert out.shape[0] == len(rec)
    bad = []
    with np.errstate(all="ignore"):
        for k, c in enumerate(rec):
            a, b, cc, ox = float(c["a"]), float(c["b"]), float(c["c"]), float(c["o"][0])
            args = {"a": (a,), "b": (b,), "ab": (a, b), "abc": (a, b, cc), "o": (ox,)}
            for w, (name, which) in enumerate(OPS):
                exp = orc.approx(name, *args[which])
                if not _same(exp, out[k, w:w + 1].view(np.float32)[0]):
                    bad.append((k, name, args[which], exp, out[k, w:w + 1].view(np.float32)[0]))
            if orc.sat_f2u(a) != int(out[k, 16]):
                bad.append((k, "sat_f2u", a, orc.sat_f2u(a), int(out[k, 16])))
            origin, rsq = orc.sphere_double(c["o"], float(c["r"]))
            got = out[k, 17:21].view(np.float32)
            if not (_same(rsq, got[0]) and all(_same(origin[i], got[1 + i]) for i in range(3))):
                bad.append((k, "BoundingSphere(double3, double)", tuple(c["o"]), float(c["r"]), (origin, rsq), got))
            ov = orc.sphere_overlap(origin, rsq, c["lo"], c["hi"])
            f_origin, f_rsq = orc.sphere_float([a, b, cc], float(F32(c["r"])))
            fv = orc.sphere_overlap(f_origin, f_rsq, c["lo"], c["hi"])
            if int(out[k, 21]) != (int(ov) | (int(fv) << 1)):
                bad.append((k, "overlap", int(out[k, 21]), ov, fv))
            if len(bad) > 20:
                break
    assert not bad, bad[:20]


def _full_record_parity(nb, orc, pts, lo, hi, spheres, images=None):
    lb = nb.LBVH()
    lb.build(dev(pts), lo, hi)
    assert lb.error_word() == 0
    t = orc.build(pts, lo, hi)
    arr = {k: host(v) for k, v in lb.arrays().items()}
    assert np.array_equal(arr["primitive"].view(np.uint32), t["primitive"])
    assert np.array_equal(arr["lo"], t["lo"]) and np.array_equal(arr["hi"], t["hi"])
    rec, rlo, rhi, bins = orc.compress(t)
    tr = nb.LBVHTraverser()
    tr.setup(lb)
    d = tr.data(lb)
    assert np.array_equal(host(d["records"]), rec)                     # all four words of every record
    assert host(d["root_lo"]).tobytes() == rlo.tobytes() and host(d["root_hi"]).tobytes() == rhi.tobytes()
    assert host(d["bins"]).tobytes() == bins.tobytes()
    ref = orc.traverse(len(pts), (rec, rlo, rhi, bins), spheres, images)["counts"]
    for variant in (0, 1, 4):
        trv = nb.LBVHTraverser(variant=variant)
        got = host(trv.count(lb, dev(spheres), images=None if images is None else dev(images)))
        assert np.array_equal(got.view(np.uint32), ref), variant
    return bins


def test_reference_cases_full_records(nb, orc, golden_dir):
    """The basic and single cases of test/lbvh_test.cu (collinear points on the x axis): every
    record word, the root box and the bins, bit for bit.  The y and z axes are degenerate, so
    their bins are fsub_ru(0,0) = 1.4e-45 turned into FLT_MAX and back: 2.9e-39."""
    with open(os.path.join(golden_dir, "reference_cases.json")) as f:
        cases = json.load(f)
    for name in ("basic", "single"):
        c = cases[name]
        log(f"full records: {name}")
        pts = np.array(c["points"], np.float32)
        bins = _full_record_parity(nb, orc, pts, c["lo"], c["hi"], np.array(c["spheres"], np.float32))
        if name == "basic":
            assert bins[1] == bins[2] and 2.8e-39 < bins[1] < 3.0e-39


@pytest.mark.parametrize("shape", ["collinear", "coplanar", "coincident", "collinear_negative"])
def test_degenerate_axes_records(nb, orc, shape):
    log(f"degenerate axes: {shape}")
    rng = np.random.default_rng({"collinear": 1, "coplanar": 2, "coincident": 3, "collinear_negative": 4}[shape])
    N = 3000
    pts = rng.uniform(0.0, 50.0, size=(N, 3)).astype(np.float32)
    if shape == "collinear":
        pts[:, 1:] = 0.0
    elif shape == "coplanar":
        pts[:, 2] = 0.0
    elif shape == "coincident":
        pts[:] = np.float32([12.5, 3.25, 40.0])
    else:
        pts[:, 0] = np.float32(-4.0)
        pts[:, 2] = np.float32(-7.5)
        pts[:, 1] = -pts[:, 1]
    lo, hi = [-60.0, -60.0, -60.0], [60.0, 60.0, 60.0]
    sph = np.concatenate([pts[:400], np.full((400, 1), 2.0, np.float32)], axis=1)
    sph[::7, :3] += np.float32(0.75)        # off-line queries
    _full_record_parity(nb, orc, pts, lo, hi, sph)
    # and over 27 periodic images of a 120-wide box
    from neighbor_amd.workloads import images27
    _full_record_parity(nb, orc, pts, lo, hi, sph, images27(120.0))
