import numpy as np
import pytest

from oracle import dga_oracle as O


@pytest.mark.parametrize("n,count", [(4, 6), (16, 45), (24, 91), (32, 153), (48, 325), (64, 561)])
def test_irreducible_counts(n, count):
    m = O.kgrid_irr_maps((n, n, 1))
    assert len(m["irrk_ind"]) == count
    assert m["irrk_count"].sum() == n * n
    assert np.array_equal(m["irrk_inv"][m["irrk_ind"]], np.arange(count))


@pytest.mark.parametrize("n", [4, 6, 16])
def test_orbits_are_the_square_group(n):
    """Brute force: every k maps to the representative of its D4 orbit."""
    m = O.kgrid_irr_maps((n, n, 1))
    rep = m["irrq_list"][m["irrk_inv"]]
    for k in range(n * n):
        kx, ky = divmod(k, n)
        orbit = set()
        for (a, b) in [(kx, ky), (ky, kx)]:
            for sx in (1, -1):
                for sy in (1, -1):
                    orbit.add(((sx * a) % n, (sy * b) % n))
        assert (rep[k][0], rep[k][1]) in orbit


@pytest.mark.parametrize("nk,sym", [((4, 4, 1), "two_dimensional_square"), ((24, 24, 1), "two_dimensional_square"),
                                    ((7, 7, 1), "two_dimensional_square"), ((6, 4, 1), "two_dimensional_nematic"),
                                    ((4, 4, 4), "three_dimensional_cubic"), ((5, 5, 3), "quasi_two_dimensional_square")])
def test_product_kgrid_matches_oracle(nk, sym):
    from dgamore_amd.brillouin_zone import KGrid

    g = KGrid(nk, sym)
    m = O.kgrid_irr_maps(nk, sym)
    assert np.array_equal(g.irrk_ind, m["irrk_ind"])
    assert np.array_equal(g.irrk_inv, m["irrk_inv"])
    assert np.array_equal(g.irrk_count, m["irrk_count"])
    assert np.array_equal(g.get_irrq_list(), m["irrq_list"])
