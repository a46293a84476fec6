"""CPU check of the product's device math headers: ApproximateMath.h and
BoundingVolumes.h are __host__ __device__, so a host build of them is compared
here with the oracle on the same inputs (the GPU build is compared in the gpu tests)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdio>
#include <cstdint>
#include <cstring>
#include "neighbor/BoundingVolumes.h"
using namespace neighbor;
int main(int argc, char** argv) {
    FILE* fi = std::fopen(argv[1], "rb"); FILE* fo = std::fopen(argv[2], "wb");
    float v[3];
    while (std::fread(v, sizeof(float), 3, fi) == 3) {
        float r[12] = {approx::next_down(v[0]), approx::next_up(v[0]), approx::fadd_rd(v[0], v[1]), approx::fsub_ru(v[0], v[1]),
                       approx::fmul_rd(v[0], v[1]), approx::fdiv_rd(v[0], v[1]), approx::frcp_rd(v[1]),
                       approx::fmaf_rd(v[0], v[1], v[2]), approx::fmaf_ru(v[0], v[1], v[2]),
                       approx::double2float_rd((double)v[0] * 1.0000001), approx::double2float_ru((double)v[0] * 1.0000001),
                       (float)approx::sat_f2u(v[0])};
        std::fwrite(r, sizeof(float), 12, fo);
    }
    std::fclose(fi); std::fclose(fo);
    return 0;
}
"""


@pytest.fixture(scope="module")
def host_exe():
    d = tempfile.mkdtemp()
    src = os.path.join(d, "m.hip")
    exe = os.path.join(d, "m")
    open(src, "w").write(PROGRAM)
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950",
           "-I", os.path.join(ROOT, "neighbor_amd", "include"), src, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr[-2000:])
    return exe, d


def test_host_build_of_device_math_matches_oracle(orc, host_exe):
    exe, d = host_exe
    rng = np.random.default_rng(11)
    vals = np.concatenate([
        np.array([[0, 1, 2], [-0.0, 3, 1], [np.inf, 2, 0], [-np.inf, 2, 0], [3.4028235e38, 2, 1],
                  [-3.4028235e38, 2, 1], [1.4e-45, 3, 0], [-1.4e-45, 3, 0], [-1.0, 7, 0.5], [5e9, 1, 1]], np.float32),
        (rng.standard_normal((3000, 3)) * 10.0 ** rng.integers(-30, 30, (3000, 1))).astype(np.float32),
    ])
    vals.tofile(os.path.join(d, "in.f32"))
    subprocess.run([exe, os.path.join(d, "in.f32"), os.path.join(d, "out.f32")], check=True)
    out = np.fromfile(os.path.join(d, "out.f32"), np.float32).reshape(-1, 12)
    with np.errstate(all="ignore"):
        for (x, y, z), r in zip(vals, out):
            exp = [orc.next_down(x), orc.next_up(x), orc.approx("fadd_rd", x, y), orc.approx("fsub_ru", x, y),
                   orc.approx("fmul_rd", x, y), orc.approx("fdiv_rd", x, y), orc.approx("frcp_rd", y),
                   orc.approx("fmaf_rd", x, y, z), orc.approx("fmaf_ru", x, y, z),
                   orc.approx("double2float_rd", float(x) * 1.0000001), orc.approx("double2float_ru", float(x) * 1.0000001),
                   float(orc.sat_f2u(x))]
            e = np.array(exp, np.float32)
            same = (e.view(np.uint32) == r.view(np.uint32)) | (np.isnan(e) & np.isnan(r))
            assert same.all(), (x, y, z, e, r)


PREFILTER_PROGRAM = r"""
#include <cstdio>
#include "neighbor/QueryOps.h"
using namespace neighbor;
int main(int argc, char** argv) {
    FILE* fi = std::fopen(argv[1], "rb"); FILE* fo = std::fopen(argv[2], "wb");
    float q[4]; double t[3]; float p[3];
    long bad = 0;
    while (std::fread(q, 4, 4, fi) == 4 && std::fread(t, 8, 3, fi) == 3 && std::fread(p, 4, 3, fi) == 3) {
        const ExactSphere v = ExactSphereQueryOp::make(make_float4(q[0], q[1], q[2], q[3]), make_double3(t[0], t[1], t[2]));
        const float4 pp = make_float4(p[0], p[1], p[2], 0.f);
        unsigned char r[2] = {(unsigned char)ExactSphereQueryOp::refine_point(v, pp), (unsigned char)ExactSphereQueryOp::refine_fp64(v, pp)};
        std::fwrite(r, 1, 2, fo);
    }
    std::fclose(fi); std::fclose(fo);
    return 0;
}
"""


def test_float_pretest_agrees_with_fp64_near_the_sphere(tmp_path):
    """Adversarial points within a few float ulps of |q + t - p| = R: the banded float
    pre-test of ExactSphereQueryOp::refine_point must give the FP64 answer every time."""
    src = tmp_path / "p.hip"
    exe = tmp_path / "p"
    src.write_text(PREFILTER_PROGRAM)
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950",
           "-I", os.path.join(ROOT, "neighbor_amd", "include"), str(src), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    rng = np.random.default_rng(5)
    L = 105.56672
    n = 200000
    q = rng.uniform(0, L, size=(n, 3)).astype(np.float32)
    R = rng.choice(np.array([2.8, 1.0, 5.9, 0.3], np.float32), size=n)
    img = rng.integers(-1, 2, size=(n, 3)).astype(np.float64) * L
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    scale = 1.0 + rng.choice([0.0, 1e-7, -1e-7, 3e-7, -3e-7, 1e-6, -1e-6, 1e-5, -1e-5], size=n)
    c = q.astype(np.float64) + img
    p = (c - d * (R.astype(np.float64) * scale)[:, None]).astype(np.float32)
    rec = np.zeros(n, dtype=[("q", np.float32, 4), ("t", np.float64, 3), ("p", np.float32, 3)])
    rec["q"][:, :3] = q
    rec["q"][:, 3] = R
    rec["t"] = img
    rec["p"] = p
    rec.tofile(tmp_path / "in.bin")
    subprocess.run([str(exe), str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], check=True)
    out = np.fromfile(tmp_path / "out.bin", np.uint8).reshape(-1, 2)
    assert out.shape[0] == n
    assert np.array_equal(out[:, 0], out[:, 1])
    # both outcomes occur, so the test exercises the boundary
    assert 0.2 < out[:, 1].mean() < 0.8
