"""The oracle's index arithmetic against the reference's own unit tests.

Each test restates one reference test: its closed-form expectation (written
independently of the implementation by the reference authors) or its
hard-coded golden table (tests/golden/reference_vectors.json, extracted by
tests/golden/extract_reference_vectors.py).  Sweeps are numbered 1..8 (3D)
and 1..4 (2D) as in SweepNumber3D/SweepNumber2D.
"""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden",
                      "reference_vectors.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def _neg(sweep, axis):
    """Does sweep (1-based) run towards -axis (axis 0: x, 1: y, 2: z)?"""
    return ((sweep - 1) >> axis) & 1


# ---------------------------------------------------------------- 3D -------
def test_sweep_signs3d(oracle):
    """testing/solver3d.cpp:64-115."""
    expected = {1: (-1, -1, -1, 1, 1, 1), 2: (1, -1, -1, -1, 1, 1),
                3: (-1, 1, -1, 1, -1, 1), 4: (1, 1, -1, -1, -1, 1),
                5: (-1, -1, 1, 1, 1, -1), 6: (1, -1, 1, -1, 1, -1),
                7: (-1, 1, 1, 1, -1, -1), 8: (1, 1, 1, -1, -1, -1)}
    for sweep, ref in expected.items():
        assert oracle.sweep_signs3d(sweep) == ref


def test_slowness_indices3d(oracle):
    """testing/solver3d.cpp:117-306: the 7 cells around every node of a
    25 x 26 x 27 grid, for all 8 sweeps, against the test's clamped
    closed form."""
    nx, ny, nz = 25, 26, 27
    ncx, ncy, ncz = nx - 1, ny - 1, nz - 1

    def g(x, y, z):
        return (z * ncy + y) * ncx + x

    def clamp(v, n):
        return max(0, min(n - 1, v))

    for sweep in range(1, 9):
        ixs, iys, izs, _, _, _ = oracle.sweep_signs3d(sweep)
        sgx = 0 if _neg(sweep, 0) else 1
        sgy = 0 if _neg(sweep, 1) else 1
        sgz = 0 if _neg(sweep, 2) else 1
        for iz in range(nz):
            for iy in range(ny):
                for ix in range(nx):
                    cx = clamp(ix - sgx, ncx)
                    cy = clamp(iy - sgy, ncy)
                    cz = clamp(iz - sgz, ncz)
                    cxs, cys = clamp(cx - ixs, ncx), clamp(cy - iys, ncy)
                    czs = clamp(cz - izs, ncz)
                    ref = (g(cx, cy, cz), g(cxs, cy, cz), g(cxs, cys, cz),
                           g(cx, cys, cz), g(cx, cy, czs), g(cxs, cy, czs),
                           g(cx, cys, czs))
                    got = oracle.slowness_indices3d(sweep, ix, iy, iz, ncx,
                                                    ncy, ncz)
                    assert tuple(got) == ref, (sweep, ix, iy, iz)


def test_travel_time_indices3d(oracle):
    """testing/solver3d.cpp:308-484 on a 32 x 43 x 56 grid (every node, all
    8 sweeps)."""
    nx, ny, nz = 32, 43, 56

    def g(x, y, z):
        return (z * ny + y) * nx + x

    def c(v, n):
        return max(0, min(n - 1, v))

    for sweep in range(1, 9):
        ixs, iys, izs, _, _, _ = oracle.sweep_signs3d(sweep)
        for iz in range(nz):
            for iy in range(ny):
                for ix in range(nx):
                    x1, y1, z1 = c(ix + ixs, nx), c(iy + iys, ny), c(iz + izs,
                                                                      nz)
                    ref = (g(ix, iy, iz), g(x1, iy, iz), g(x1, y1, iz),
                           g(ix, y1, iz), g(ix, iy, z1), g(x1, iy, z1),
                           g(x1, y1, z1), g(ix, y1, z1))
                    got = oracle.travel_time_indices3d(sweep, ix, iy, iz, nx,
                                                       ny, nz)
                    assert tuple(got) == ref, (sweep, ix, iy, iz)


def test_loop_limits3d(oracle):
    """testing/solver3d.cpp:486-555 (10 x 11 x 9)."""
    n = (10, 11, 9)
    for sweep in range(1, 9):
        got = oracle.loop_limits3d(sweep, *n)
        for axis in range(3):
            start, stop, step = got[axis], got[3 + axis], got[6 + axis]
            if _neg(sweep, axis):
                assert (start, stop, step) == (n[axis] - 2, -1, -1)
            else:
                assert (start, stop, step) == (1, n[axis], 1)


def test_loop_limits_init3d_cover_octant(oracle):
    """The initialisation ranges (solverUtilities3d.hpp:1075-1184) start at
    the source cell on the sweep's downwind side: every node a sweep may
    update lies in its octant, and the 8 octants together cover every node
    off the source cell's own upwind boundary."""
    n = (9, 12, 10)
    for src in ((0, 0, 0), (3, 7, 4), (7, 10, 8), (4, 0, 8)):
        covered = np.zeros(n[::-1], dtype=np.int32)
        for sweep in range(1, 9):
            lim = oracle.loop_limits3d(sweep, *n, init_source=src)
            rng = [range(lim[a], lim[3 + a], lim[6 + a]) for a in range(3)]
            for a in range(3):
                if _neg(sweep, a):
                    assert all(i <= src[a] + 1 for i in rng[a])
                else:
                    assert all(i >= src[a] for i in rng[a])
            for iz in rng[2]:
                for iy in rng[1]:
                    for ix in rng[0]:
                        covered[iz, iy, ix] += 1
        assert covered[1:-1, 1:-1, 1:-1].min() >= 1


def test_permute3d(oracle):
    """testing/solver3d.cpp:557-613 (4 x 5 x 6)."""
    n = (4, 5, 6)
    for sweep in range(1, 9):
        for iz in range(n[2]):
            for iy in range(n[1]):
                for ix in range(n[0]):
                    got = oracle.permute3d(sweep, ix, iy, iz, *n)
                    ref = tuple(n[a] - 1 - v if _neg(sweep, a) else v
                                for a, v in enumerate((ix, iy, iz)))
                    assert got == ref
                    # an involution
                    assert oracle.permute3d(sweep, *got, *n) == (ix, iy, iz)


def test_is_sweep_boundary3d(oracle):
    """testing/solver3d.cpp:615-695 (5 x 6 x 7)."""
    n = (5, 6, 7)
    for sweep in range(1, 9):
        for iz in range(n[2]):
            for iy in range(n[1]):
                for ix in range(n[0]):
                    ref = any(
                        v == (n[a] - 1 if _neg(sweep, a) else 0)
                        for a, v in enumerate((ix, iy, iz)))
                    assert oracle.is_sweep_boundary3d(sweep, ix, iy, iz,
                                                      *n) == ref


def test_graph3d_golden(oracle, golden):
    """testing/graph.cpp:40-265: the level lists of all 8 sweeps on a
    4 x 3 x 5 grid, the level pointer, 10 levels, largest level 11."""
    nx, ny, nz = 4, 3, 5
    assert oracle.graph3d_number_of_levels(nx, ny, nz) == 10
    for sweep in range(1, 9):
        start, nodes = oracle.graph3d(sweep, nx, ny, nz)
        assert start.tolist() == golden["levelStartPtrRef"]
        assert np.diff(start).max() == 11
        assert nodes.tolist() == golden["nodeInLevelToGridPointRef%d" % sweep]


def test_graph3d_levels_are_hyperplanes(oracle):
    """Every node of level L has jx + jy + jz = L in the sweep's frame, and
    every node appears once: the closed form the GPU kernels use instead of
    stored lists."""
    nx, ny, nz = 7, 5, 6
    assert oracle.graph3d_number_of_levels(5, 3, 7) == 13   # graph.cpp:31-38
    for sweep in range(1, 9):
        start, nodes = oracle.graph3d(sweep, nx, ny, nz)
        assert sorted(nodes.tolist()) == list(range(nx * ny * nz))
        for level in range(len(start) - 1):
            for idx in nodes[start[level]:start[level + 1]]:
                iz, rem = divmod(int(idx), nx * ny)
                iy, ix = divmod(rem, nx)
                j = oracle.permute3d(sweep, ix, iy, iz, nx, ny, nz)
                assert sum(j) == level


# ---------------------------------------------------------------- 2D -------
def test_sweep_signs2d_and_levels(oracle):
    """testing/solver2d.cpp:19-59."""
    assert oracle.number_of_levels2d(5, 7) == 11
    expected = {1: (-1, -1, 1, 1), 2: (1, -1, -1, 1), 3: (-1, 1, 1, -1),
                4: (1, 1, -1, -1)}
    for sweep, ref in expected.items():
        assert oracle.sweep_signs2d(sweep) == ref


@pytest.mark.parametrize("nx,nz,ref", [
    (3, 5, [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (2, 1), (1, 2),
            (0, 3), (2, 2), (1, 3), (0, 4), (2, 3), (1, 4), (2, 4)]),
    (5, 3, [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1),
            (1, 2), (4, 0), (3, 1), (2, 2), (4, 1), (3, 2), (4, 2)]),
])
def test_level_set_indices2d(oracle, nx, nz, ref):
    """testing/solver2d.cpp:245-440: the nodes of each level, per sweep."""
    for sweep in range(1, 5):
        j = 0
        for level in range(oracle.number_of_levels2d(nx, nz)):
            i0, i1 = oracle.level_start_stop2d(nx, nz, level)
            for indx in range(i0, i1):
                ix, iz = oracle.sweep_level_index_to_grid2d(sweep, level,
                                                            indx, nx, nz)
                rx, rz = ref[j]
                if _neg(sweep, 0):
                    rx = nx - 1 - rx
                if _neg(sweep, 1):
                    rz = nz - 1 - rz
                assert (ix, iz) == (rx, rz)
                j += 1
        assert j == len(ref)


@pytest.mark.parametrize("nx,nz", [(12, 12), (52, 71)])
def test_grid_sweep_to_level_index2d(oracle, nx, nz):
    """testing/solver2d.cpp:442-562: level/index <-> grid round trip."""
    for sweep in range(1, 5):
        for level in range(oracle.number_of_levels2d(nx, nz)):
            i0, i1 = oracle.level_start_stop2d(nx, nz, level)
            for indx in range(i0, i1):
                ix, iz = oracle.sweep_level_index_to_grid2d(sweep, level,
                                                            indx, nx, nz)
                assert oracle.grid_to_level_index2d(sweep, ix, iz, nx,
                                                    nz) == (level, indx)


def test_level_travel_time_indices2d_golden(oracle, golden):
    """testing/solver2d.cpp:564-710: the four travel times each level node
    reads on a 5 x 3 grid holding 1..15, per sweep (tRef0..tRef3)."""
    nx, nz = 5, 3
    t = golden["tTimes"]
    for sweep in range(1, 5):
        ref = golden["tRef%d" % (sweep - 1)]
        j = 0
        for level in range(oracle.number_of_levels2d(nx, nz)):
            i0, i1 = oracle.level_start_stop2d(nx, nz, level)
            for indx in range(i0, i1):
                it = oracle.level_travel_time_indices2d(sweep, level, indx,
                                                        nx, nz)
                assert [t[k] for k in it] == ref[j:j + 4]
                j += 4
        assert j == len(ref)


def test_sweep_slowness2d_golden(oracle, golden):
    """testing/solver2d.cpp:712-863: the cells s0, s1, s3 each level node
    reads on a 5 x 4 grid with slowness 1..12, per sweep
    (sweepSlowRef1..4; s2 is unused by the reference)."""
    nx, nz = 5, 4
    slow = golden["slow"]
    for sweep in range(1, 5):
        ref = golden["sweepSlowRef%d" % sweep]
        j = 0
        for level in range(oracle.number_of_levels2d(nx, nz)):
            i0, i1 = oracle.level_start_stop2d(nx, nz, level)
            for indx in range(i0, i1):
                ix, iz = oracle.sweep_level_index_to_grid2d(sweep, level,
                                                            indx, nx, nz)
                c0, c1, c3 = oracle.slowness_indices2d(sweep, ix, iz, nx - 1,
                                                       nz - 1)
                assert (slow[c0], slow[c1], slow[c3]) == (
                    ref[j], ref[j + 1], ref[j + 3])
                j += 4
        assert j == len(ref)


def test_sweep_slowness2d_large(oracle):
    """testing/solver2d.cpp:864-1027: the same gather on 242 x 323 against
    the test's explicit case analysis (sampled every 7th node)."""
    nx, nz = 242, 323
    ncx, ncz = nx - 1, nz - 1
    for sweep in range(1, 5):
        negx, negz = _neg(sweep, 0), _neg(sweep, 1)
        for k, (ix, iz) in enumerate(
                (ix, iz) for iz in range(nz) for ix in range(nx)):
            if k % 7:
                continue
            if not negx:
                c0x = max(0, ix - 1)
                c1x = 0 if ix == 0 else min(ncx - 1, c0x + 1)
            else:
                c0x, c1x = ix, max(0, ix - 1)
                if ix == ncx:
                    c0x = c1x = ix - 1
            if not negz:
                c0z = max(0, iz - 1)
                c2z = 0 if iz == 0 else min(ncz - 1, c0z + 1)
            else:
                c0z, c2z = iz, max(0, iz - 1)
                if iz == ncz:
                    c0z = c2z = iz - 1
            ref = (c0z * ncx + c0x, c0z * ncx + c1x, c2z * ncx + c0x)
            got = oracle.slowness_indices2d(sweep, ix, iz, ncx, ncz)
            assert tuple(got) == ref, (sweep, ix, iz)


def test_loop_limits2d(oracle):
    """testing/solver2d.cpp:1029-1118: the refinement loops visit every
    node off the sweep's upwind boundary exactly once."""
    nx, nz = 10, 11
    for sweep in range(1, 5):
        ix0, iz0, ix1, iz1, dx, dz = oracle.loop_limits2d(sweep, nx, nz)
        seen = np.zeros((nz, nx), dtype=int)
        for iz in range(iz0, iz1, dz):
            for ix in range(ix0, ix1, dx):
                seen[iz, ix] += 1
        ref = np.ones((nz, nx), dtype=int)
        ref[:, nx - 1 if _neg(sweep, 0) else 0] = 0
        ref[nz - 1 if _neg(sweep, 1) else 0, :] = 0
        assert np.array_equal(seen, ref)
