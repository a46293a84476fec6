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
