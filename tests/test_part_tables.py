"""Domain-part tables (gbx_part_tables, the host side of gbx_geom_create_part),
checked on the CPU against the periodic lattice itself: for every part of a
process grid, every slot and every leg, the site a leg reads (a local slot,
or a ghost slot filled from the neighbouring part's face) must be the
global neighbour x +/- mu.

The reference runs on one rank (TableGenerate.cc:84 asserts that every
stencil neighbour is local), so there is no reference decomposition to
compare with: the periodic neighbour relation is the ground truth.
"""
import numpy as np
import pytest

CASES = [
    ((4, 4, 4, 8), (1, 1, 1, 2)),
    ((4, 4, 4, 8), (1, 1, 1, 4)),
    ((4, 4, 4, 8), (1, 1, 1, 8)),   # one layer per part: both faces are the same sites
    ((8, 4, 4, 4), (2, 1, 1, 1)),
    ((4, 6, 4, 4), (1, 3, 1, 1)),
    ((4, 4, 4, 4), (1, 1, 2, 1)),   # two layers per part: no interior
    ((6, 4, 2, 10), (1, 1, 1, 5)),
    ((4, 4, 6, 6), (1, 1, 2, 2)),   # two cut dimensions
    ((4, 6, 6, 6), (2, 1, 3, 2)),   # three
    ((4, 4, 4, 4), (2, 2, 2, 2)),   # every dimension cut, two layers each
    ((6, 3, 3, 4), (3, 3, 1, 2)),   # one layer along y
]


def coords(site, dims):
    c = []
    for d in dims:
        c.append(site % d)
        site //= d
    return c


def lex(c, dims):
    s, stride = 0, 1
    for x, d in zip(c, dims):
        s += x * stride
        stride *= d
    return s


def global_nbr(site, mu, dims):
    c = coords(site, dims)
    d = mu & 3
    c[d] = (c[d] + (1 if mu < 4 else -1)) % dims[d]
    return lex(c, dims)


def step(c, d, k, grid):
    c = list(c)
    c[d] = (c[d] + k) % grid[d]
    return tuple(c)


@pytest.mark.parametrize("dims,grid", CASES)
def test_legs_read_the_global_neighbour(gbx, dims, grid):
    parts = {c: gbx.part_tables(dims, grid, c) for c in gbx.grid_coords(grid)}
    ld = [d // n for d, n in zip(dims, grid)]
    t0 = next(iter(parts.values()))
    vol = t0["vol"]
    assert vol * len(parts) == int(np.prod(dims))
    seen = np.zeros(int(np.prod(dims)), np.int64)
    for c, t in parts.items():
        sos, gs, nbr = t["site_of_slot"], t["gsite"], t["nbr"]
        assert sorted(sos.tolist()) == list(range(vol))
        for q in range(vol):
            x = coords(int(sos[q]), ld)
            assert gs[q] == lex([x[d] + c[d] * ld[d] for d in range(4)], dims)
        seen[gs] += 1
        # which ghost range (dimension, side) a slot index falls in
        ranges = []
        for d, base in t["ghost_base"].items():
            f = t["face"][d]
            ranges.append((base, base + f, d, 0))
            ranges.append((base + f, base + 2 * f, d, 1))
        nslots = vol + sum(2 * f for f in t["face"].values())
        for q in range(vol):
            for mu in range(8):
                want = global_nbr(int(gs[q]), mu, dims)
                n = int(nbr[q, mu])
                assert n < nslots
                if n < vol:
                    got = gs[n]
                else:
                    lo, _, d, side = next(r for r in ranges if r[0] <= n < r[1])
                    assert d == (mu & 3) and side == (0 if mu >= 4 else 1), (c, q, mu)
                    # minus ghosts hold the previous part's top face, plus ghosts the next part's bottom face
                    src = parts[step(c, d, -1 if side == 0 else 1, grid)]
                    got = src["gsite"][src["face_slots"][d][1 - side][n - lo]]
                assert got == want, (c, q, mu, n)
    assert (seen == 1).all()


@pytest.mark.parametrize("dims,grid", CASES)
def test_interior_and_faces(gbx, dims, grid):
    ld = [d // n for d, n in zip(dims, grid)]
    cut = [d for d in range(4) if grid[d] > 1]
    for c in gbx.grid_coords(grid):
        t = gbx.part_tables(dims, grid, c)
        vol, nint = t["vol"], t["ninterior"]
        interior = int(np.prod([ld[d] - 2 if d in cut else ld[d] for d in range(4)])) if all(
            ld[d] > 2 for d in cut) else 0
        assert nint == interior
        ghost_legs = (t["nbr"] >= vol).sum(axis=1)
        assert (ghost_legs[:nint] == 0).all() and (ghost_legs[nint:] >= 1).all()
        for q in range(vol):
            x = coords(int(t["site_of_slot"][q]), ld)
            want = sum((x[d] == 0) + (x[d] == ld[d] - 1) for d in cut)
            assert ghost_legs[q] == want
        for d in cut:
            for side, layer in ((0, 0), (1, ld[d] - 1)):
                for f, q in enumerate(t["face_slots"][d][side]):
                    x = coords(int(t["site_of_slot"][q]), ld)
                    assert x[d] == layer
                    assert lex([x[e] for e in range(4) if e != d], [ld[e] for e in range(4) if e != d]) == f


@pytest.mark.parametrize("args", [((4, 4, 4, 8), (1, 1, 1, 3), (0, 0, 0, 0)),   # 3 does not divide 8
                                  ((4, 4, 4, 8), (1, 1, 1, 1), (0, 0, 0, 0)),   # nothing cut
                                  ((4, 4, 4, 8), (1, 1, 1, 2), (0, 0, 0, 2)),   # coord out of range
                                  ((4, 4, 4, 8), (1, 1, 1, 2), (0, 0, 0, -1)),
                                  ((4, 4, 4, 8), (0, 1, 1, 2), (0, 0, 0, 0))])
def test_rejects_bad_decompositions(gbx, args):
    with pytest.raises(gbx.GbxError) as e:
        gbx.part_tables(*args)
    assert e.value.status == 1  # GBX_EINVAL
