"""The uneven-size batch workload (16 models, seeded) and the orbital representations that the
irreducible-zone reduction of its multi-orbital models uses (CPU only)."""
import numpy as np

from dgamore_amd import synthetic as S
from dgamore_amd.brillouin_zone import KGrid, group_index_maps, orbital_transforms


def test_batch_draw_spans_the_size_range():
    cfgs = S.batch_configs(seed=180)
    assert len(cfgs) == 16
    sizes = sorted(c["nb"] ** 2 * 2 * c["niv"] for c in cfgs)
    assert sizes[0] == 16 and sizes[-1] == 1080
    for c in cfgs:
        assert c["nb"] in (1, 2, 3) and c["beta"] in (5.0, 10.0, 20.0, 40.0)
        assert c["niv"] == c["niw"] == round(1.5 * c["beta"])
        assert c["nk"][0] in (16, 24, 32) and 2.0 <= c["U"] <= 8.0 and 0.8 <= c["n"] <= 1.0
    assert S.batch_configs(seed=180) == cfgs  # seeded


def test_square_lattice_group_and_orbital_representations():
    nk = (6, 6, 1)
    maps = group_index_maps(nk)
    assert len(maps) == 8 and np.array_equal(maps[0], np.arange(36))
    # every element is a permutation of the grid and the set is closed under composition
    keys = {m.tobytes() for m in maps}
    for a in maps:
        assert sorted(a) == list(range(36))
        for b in maps:
            assert b[a].tobytes() in keys
    assert orbital_transforms(S.dispersion_1band(nk, 1.0, -0.2, 0.1), KGrid(nk)) is None
    for ek in (S.dispersion_eg(nk), S.dispersion_t2g(nk)):
        grid = KGrid(nk)
        us, conjs = orbital_transforms(ek, grid)
        e = ek.reshape(36, ek.shape[-1], -1)
        rep = grid.irrk_ind[grid.irrk_inv]
        assert np.abs(e - np.einsum("kab,kbc,kdc->kad", us, e[rep], us.conj())).max() < 1e-12
        assert not conjs.any()
        assert np.allclose(us[grid.irrk_ind], np.eye(ek.shape[-1]))
