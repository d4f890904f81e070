"""Red-black (even-odd) host helpers: the checkerboarded lexicographic order
((x/2) fastest) used by half-lattice fields, checked on the CPU."""
import numpy as np
import pytest


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (8, 2, 6, 4), (2, 2, 2, 2)])
def test_rb_split_merge_roundtrip_and_parity(gbx, dims):
    V = int(np.prod(dims))
    rng = np.random.default_rng(1)
    lex = rng.standard_normal((2, V, 3, 24))
    e, o = gbx.rb_split(dims, lex)
    assert e.shape == (2, V // 2, 3, 24) and o.shape == e.shape
    assert np.array_equal(gbx.rb_merge(dims, e, o), lex)
    # element h of a half is the site with x/2 = h % (Lx/2) and (y, z, t) from h // (Lx/2),
    # x chosen so that the site has that half's parity
    L0 = dims[0]
    for p, half in ((0, e), (1, o)):
        for h in (0, 1, V // 2 - 1, V // 4):
            rest = h // (L0 // 2)
            y, z, t = rest % dims[1], (rest // dims[1]) % dims[2], rest // (dims[1] * dims[2])
            x = 2 * (h % (L0 // 2))
            if (x + y + z + t) % 2 != p:
                x += 1
            site = x + L0 * (y + dims[1] * (z + dims[2] * t))
            assert np.array_equal(half[:, h], lex[:, site])
