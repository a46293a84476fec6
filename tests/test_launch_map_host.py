"""CPU check of the launch map of the compacted exact walks (wave::tail_block in
kernels/LBVHTraverse.h, __host__ __device__): a host build enumerates every block and
lane of the grid the traverser launches and the test checks that the live lanes cover
each query position exactly once, that a block's lead is the number of positions of its
group of 256 that precede it, that every XCD's blocks (b % 8) serve one contiguous
stretch of groups in dispatch order, and that the split blocks come last."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include "neighbor/LBVHTraverser.h"
using neighbor::gpu::kernel::wave::tail_block;
using neighbor::gpu::kernel::wave::BlockSpan;
int main(int argc, char** argv) {
    const unsigned int n = std::atoi(argv[1]), T = std::atoi(argv[2]), S = std::atoi(argv[3]);
    const unsigned int G = (n + 255u) / 256u;
    const unsigned int x0 = G / 8u + (G % 8u ? 1u : 0u);
    // the grid of LBVHTraverser::traverse_wave
    const unsigned int grid = T ? 8u * (x0 + (S - 1u) * std::min(x0, T)) : G;
    std::printf("%u\n", grid);
    for (unsigned int b = 0; b < grid; ++b) {
        const BlockSpan s = tail_block(b, G, T, S);
        std::printf("%u %u %u %u %d\n", b, s.first, s.lead, s.per_wave, s.live ? 1 : 0);
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def map_exe():
    d = tempfile.mkdtemp()
    src = os.path.join(d, "m.hip")
    exe = os.path.join(d, "m")
    open(src, "w").write(PROGRAM)
    cmd = ["/opt/rocm/bin/hipcc", "-O1", "-std=c++17", "--offload-arch=gfx950",
           "-I", os.path.join(ROOT, "neighbor_amd", "include"), src, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr[-2000:])
    return exe


def run_map(exe, n, T, S):
    out = subprocess.run([exe, str(n), str(T), str(S)], check=True, capture_output=True, text=True).stdout.split("\n")
    grid = int(out[0])
    rows = np.array([[int(v) for v in line.split()] for line in out[1:] if line.strip()], dtype=np.int64)
    assert rows.shape == (grid, 5)
    return rows


@pytest.mark.parametrize("n", [1, 63, 256, 257, 2000, 7 * 256, 9 * 256 + 5, 100_000, 1_000_000])
@pytest.mark.parametrize("T,S", [(0, 2), (1, 2), (3, 4), (32, 2), (64, 4), (10 ** 6, 2), (5, 1)])
def test_tail_map_covers_each_position_once(map_exe, n, T, S):
    rows = run_map(map_exe, n, T, S)
    G = (n + 255) // 256
    seen = np.zeros(256 * G, np.int64)
    order = {x: [] for x in range(8)}        # groups of each XCD in dispatch order
    for b, first, lead, per_wave, live in rows:
        if not live:
            continue
        assert per_wave in (64, 32, 16)
        size = 4 * per_wave                   # positions of the block
        assert first % 256 == lead and lead % size == 0
        pos = first + np.arange(4)[:, None] * per_wave + np.arange(per_wave)[None, :]
        seen[pos.ravel()] += 1
        order[b % 8].append((first, per_wave))
    assert (seen == 1).all(), "positions not covered exactly once"
    # each XCD: one contiguous, ascending stretch of positions; full blocks first, split blocks last
    for x, blocks in order.items():
        firsts = [f for f, _ in blocks]
        assert firsts == sorted(firsts)
        sizes = [p for _, p in blocks]
        assert sizes == sorted(sizes, reverse=True)
        for (f0, p0), (f1, _) in zip(blocks, blocks[1:]):
            assert f1 == f0 + 4 * p0
    # T = 0 is the plain map: no block is split
    if T == 0:
        assert (rows[:, 3] == 64).all() and rows.shape[0] == G
