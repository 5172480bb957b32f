"""The dependency rule of the sweep overlap (option sweep_overlap), checked
by brute force on small grids: a column chunk C of sweep k may start once
every chunk of sweep k - 1 that writes a node C reads or writes, or reads a
node C writes, has finished.  The library waits for sweep k - 1's diagonal
(max column x) + (max column y) + (max chunk z) over C's node box grown by
one node; this test restates that closed form and checks that it is never
earlier than the brute-force requirement over nodes."""
import itertools

import pytest

TILE = 8


def perm(i, n, neg):
    return n - 1 - i if neg else i


def tiles(n, size):
    return (n + size - 1) // size


def closed_form(nx, ny, nz, chunk, k):
    """req[d] for sweep k (previous sweep k - 1 mod 8), as the library
    computes it."""
    kp = (k + 7) % 8
    neg = [(k >> a) & 1 for a in range(3)]
    negp = [(kp >> a) & 1 for a in range(3)]
    ns = (nx, ny, nz)
    sizes = (TILE, TILE, chunk)
    nb = [tiles(n, s) for n, s in zip(ns, sizes)]
    ndiag = sum(nb) - 2
    req = [-1] * ndiag

    def grown(j0, j1, n, ng):
        j0, j1 = max(j0, 0), min(j1, n - 1)
        lo, hi = (n - 1 - j1, n - 1 - j0) if ng else (j0, j1)
        return max(lo - 1, 0), min(hi + 1, n - 1)

    def last_tile(lo, hi, n, ng, size):
        return (n - 1 - lo if ng else hi) // size

    for b in itertools.product(*(range(m) for m in nb)):
        t = 0
        for a in range(3):
            if a < 2:
                j0, j1 = b[a] * TILE, b[a] * TILE + 7
            else:
                j0, j1 = b[a] * chunk - 1, (b[a] + 1) * chunk - 1
            lo, hi = grown(j0, j1, ns[a], neg[a])
            t += last_tile(lo, hi, ns[a], negp[a], sizes[a])
        d = sum(b)
        req[d] = max(req[d], min(t, ndiag - 1))
    return req


def brute_force(nx, ny, nz, chunk, k):
    """The latest sweep-(k-1) diagonal that writes a node an item of sweep k
    reads or writes (its nodes, the read-only plane before its chunk and
    its upwind halo), or reads a node it writes (downwind neighbours in
    sweep k - 1), per diagonal of sweep k."""
    kp = (k + 7) % 8
    neg = [(k >> a) & 1 for a in range(3)]
    negp = [(kp >> a) & 1 for a in range(3)]
    ns = (nx, ny, nz)
    sizes = (TILE, TILE, chunk)
    nb = [tiles(n, s) for n, s in zip(ns, sizes)]
    ndiag = sum(nb) - 2

    def diag_prev(node):
        return sum(perm(node[a], ns[a], negp[a]) // sizes[a]
                   for a in range(3))

    def inside(node):
        return all(0 <= node[a] < ns[a] for a in range(3))

    req = [-1] * ndiag
    for b in itertools.product(*(range(m) for m in nb)):
        need = -1
        jr = []
        for a in range(3):
            lo = b[a] * sizes[a]
            hi = min(lo + sizes[a], ns[a])
            if a == 2:
                lo = max(lo - 1, 0)        # the read-only first plane
            jr.append(range(lo, hi))
        for j in itertools.product(*jr):
            node = tuple(perm(j[a], ns[a], neg[a]) for a in range(3))
            # written or read by this item: itself and its upwind
            # neighbours in sweep k
            for off in itertools.product((0, 1), repeat=3):
                m = tuple(node[a] + (off[a] if neg[a] else -off[a])
                          for a in range(3))
                if inside(m):
                    need = max(need, diag_prev(m))
            # readers in sweep k - 1: its downwind neighbours there
            for off in itertools.product((0, 1), repeat=3):
                m = tuple(node[a] + (-off[a] if negp[a] else off[a])
                          for a in range(3))
                if inside(m):
                    need = max(need, diag_prev(m))
        d = sum(b)
        req[d] = max(req[d], need)
    return req


@pytest.mark.parametrize("dims,chunk", [((17, 11, 5), 5), ((9, 20, 6), 6),
                                        ((13, 13, 12), 4), ((8, 9, 7), 3)])
def test_overlap_rule_is_never_early(dims, chunk):
    nx, ny, nz = dims
    for k in range(8):
        cf = closed_form(nx, ny, nz, chunk, k)
        bf = brute_force(nx, ny, nz, chunk, k)
        assert len(cf) == len(bf)
        for d, (a, b) in enumerate(zip(cf, bf)):
            assert a >= b, (k, d, a, b)


def test_overlap_rule_overlaps_on_a_wide_grid():
    """On a grid of many columns a sweep that reverses x starts while the
    previous sweep still has about half of its diagonals to go."""
    cf = closed_form(161, 161, 9, 9, 1)
    assert cf[0] < 0.6 * len(cf)
