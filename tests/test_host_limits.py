"""CPU checks of the template API's argument limits: they throw before any device work, so a
host build of a caller runs them here without a GPU."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include "neighbor/neighbor.h"
#include <cstdio>
#include <stdexcept>
int main() {
    neighbor::LBVH lbvh;
    const neighbor::PointInsertOp insert(nullptr, 1u << 30);
    try {
        lbvh.build(insert, make_float3(0.f, 0.f, 0.f), make_float3(1.f, 1.f, 1.f));
    } catch (const std::invalid_argument& e) {
        std::printf("invalid_argument: %s\n", e.what());
        return lbvh.getN() == 0 ? 0 : 3;
    } catch (...) {
        return 2;
    }
    return 1;
}
"""


def test_lbvh_rejects_two_to_the_thirty_primitives(tmp_path):
    """LBVH::build with N >= 2^30 throws std::invalid_argument (the sort's 30-bit tile counts and
    the int node ids of 2N - 1 nodes) before allocating or launching anything."""
    src = tmp_path / "limits.hip"
    exe = tmp_path / "limits"
    src.write_text(PROGRAM)
    cmd = ["/opt/rocm/bin/hipcc", "-O1", "-std=c++17", "--offload-arch=gfx950", "-ffp-contract=off",
           "-I", os.path.join(ROOT, "neighbor_amd", "include"), str(src), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, (p.returncode, p.stdout, p.stderr)
    assert "2^30" in p.stdout
