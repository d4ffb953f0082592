"""Brick shapes of the sparse pair kernel (CPU only).

psim_brick_shape.hpp is compiled with g++ into a small program that prints
the shape chosen for a grid of (cells per dimension, occupancy, target); the
test checks what the kernel relies on: bricks tile the S^3 grid with no empty
brick, every halo and home-row table fits its LDS capacity, and the home
population lands near the target where the caps allow it."""
import os
import shutil
import subprocess

import pytest

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "proteinsim2015_amd", "csrc")

PROGRAM = r"""
#include <cstdio>
#include "psim_brick_shape.hpp"
int main() {
    const double occs[] = {0.02, 0.1, 0.3, 0.6, 1.0, 1.9, 4.0};
    const double targets[] = {51.2, 96, 192, 204.8, 384};
    for (int S = 3; S <= 200; ++S)
        for (double o : occs)
            for (double t : targets) {
                psim::BrickGeo b = psim::brick_shape(S, o, t);
                std::printf("%d %g %g %d %d %d %d %d %d %d\n", S, o, t, b.bx, b.by, b.bz,
                            b.nbx, b.nby, b.nbz, b.nbricks);
            }
    std::printf("caps %d %d %d %d\n", psim::kBrickHaloCap, psim::kBrickRowCap,
                psim::kBrickMaxEdge, psim::kBrickStageCap);
}
"""


@pytest.fixture(scope="module")
def shapes(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("brick")
    src, exe = d / "shape.cpp", d / "shape"
    src.write_text(PROGRAM)
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-I", CSRC, str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)], text=True).split("\n")
    caps = [int(v) for v in out[-2].split()[1:]]
    rows = [line.split() for line in out[:-2]]
    return caps, [(int(r[0]), float(r[1]), float(r[2]), *map(int, r[3:])) for r in rows]


def test_bricks_tile_the_grid(shapes):
    caps, rows = shapes
    for S, occ, t, bx, by, bz, nbx, nby, nbz, nb in rows:
        for b, n in ((bx, nbx), (by, nby), (bz, nbz)):
            assert 1 <= b <= S
            assert n * b >= S and (n - 1) * b < S, (S, b, n)   # no empty brick
        assert nb == nbx * nby * nbz


def test_tables_fit_their_caps(shapes):
    (halo_cap, row_cap, max_edge, _), rows = shapes
    for S, occ, t, bx, by, bz, *_ in rows:
        assert (bx + 2) * (by + 2) * (bz + 2) <= halo_cap
        assert by * bz <= row_cap
        assert max(bx, by, bz) <= max_edge


def test_home_population_near_target(shapes):
    (halo_cap, _, _, stage_cap), rows = shapes
    for S, occ, t, bx, by, bz, *_ in rows:
        home = bx * by * bz * occ
        halo_pop = (bx + 2) * (by + 2) * (bz + 2) * occ
        # reachable: the halo cap bounds a brick to ~360 home cells
        if S >= 16 and 0.1 <= occ <= 1.0 and t / occ <= 300 and \
                t * (3.4 if t < 100 else 2.6) <= 0.5 * stage_cap:
            assert 0.4 * t <= home <= 2.0 * t, (S, occ, t, bx, by, bz)
        if S >= 16 and 0.1 <= occ <= 1.0:
            assert halo_pop <= 0.5 * stage_cap or bx * by * bz == 1, (S, occ, t, bx, by, bz)
