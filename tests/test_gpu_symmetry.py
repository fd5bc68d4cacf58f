"""Irreducible-to-full-zone map with per-k orbital transforms (the reference's auto symmetry mode)
on the device: the expansion kernel against the oracle for general unitaries with conjugation,
and the whole SDE step of an orbital-symmetric multi-orbital model, reduced to the irreducible
zone with the transforms, against the same step computed on every q of the full zone."""
import sys

import numpy as np
import pytest

from oracle import dga_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
    b = b.detach().cpu().numpy() if hasattr(b, "detach") else np.asarray(b)
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture(scope="module")
def K():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dgamore_amd import kernels

    return kernels


@pytest.mark.parametrize("nb,nk", [(2, (6, 6, 1)), (3, (4, 4, 1))])
def test_kernel_orbital_expand_general_unitaries(K, nb, nk):
    say(f"kernel_orbital_expand nb={nb} nk={nk}: random unitaries, random conjugation")
    rng = np.random.default_rng(nb)
    maps = O.kgrid_irr_maps(nk)
    nq, N, nw, nv = len(maps["irrk_ind"]), int(np.prod(nk)), 3, 4
    k_irr = rng.standard_normal((nq,) + (nb,) * 4 + (nw, nv)) + 1j * rng.standard_normal((nq,) + (nb,) * 4 + (nw, nv))
    us = np.linalg.qr(rng.standard_normal((N, nb, nb)) + 1j * rng.standard_normal((N, nb, nb)))[0]
    us[maps["irrk_ind"]] = np.eye(nb)
    conjs = rng.random(N) < 0.3
    conjs[maps["irrk_ind"]] = False
    ref = O.orbital_expand_fbz(k_irr, maps["irrk_inv"], us, conjs)
    src, u, cj, inv_ext = K.orbital_expand_tables(maps["irrk_inv"], us, conjs)
    ext = K.kernel_orbital_expand(K.ref4_to_qf(K.to_dev(k_irr)), K.to_dev(src, torch.int32), K.to_dev(u),
                                  K.to_dev(cj, torch.int32))
    torch.cuda.synchronize()
    got = K.qf_to_ref4(ext, nb).cpu().numpy()[inv_ext]
    assert rel(got, ref) < 1e-13


@pytest.mark.parametrize("cfg", [4, 2])
def test_reduced_step_with_transforms_equals_full_zone_step(K, cfg):
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import SDEPipeline

    say(f"SDE step, orbital-symmetric config-{cfg} model 6x6: irreducible zone + transforms vs every q")
    m = S.make_model(cfg, nk=(6, 6, 1), niw=2, niv=3, symmetric=True)
    assert m.orbital_transforms is not None
    red = SDEPipeline(m)
    full = SDEPipeline(m, symmetries="none", orbital_transforms=None)
    plain = SDEPipeline(m, orbital_transforms=None)
    assert red.nq < full.nq
    for p in (red, full, plain):
        p.step(keep=True)
    torch.cuda.synchronize()
    assert rel(red.last["sigma_nl"], full.last["sigma_nl"]) < 1e-10
    assert rel(red.last["sigma"], full.last["sigma"]) < 1e-10
    # the plain gather (the reference's named-symmetry mode) misses the orbital permutation
    assert rel(plain.last["sigma_nl"], full.last["sigma_nl"]) > 1e-4


def test_pairing_setup_with_transforms_equals_expanded_vertex(K):
    """The Eliashberg operator built from the irreducible pp vertex with orbital transforms equals
    the one built from the vertex expanded on the host by the oracle (every k its own entry)."""
    say("pairing_setup with orbital transforms, nb=2 nk=6x6 np2=6")
    from dgamore_amd import synthetic as S
    from dgamore_amd.brillouin_zone import KGrid, orbital_transforms

    nk, nb, np2 = (6, 6, 1), 2, 6
    grid = KGrid(nk)
    us, conjs = orbital_transforms(S.dispersion_eg(nk), grid)
    rng = np.random.default_rng(8)
    shape = (grid.nk_irr,) + (nb,) * 4 + (np2, np2)
    gam = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    full = O.orbital_expand_fbz(gam, grid.irrk_inv, us, conjs)
    ident = K.to_dev(np.arange(grid.nk_tot), torch.int32)
    for ch in ("sing", "trip"):
        a = K.pairing_setup(K.to_dev(gam), K.to_dev(grid.irrk_inv, torch.int32), nk, ch, orbital_transforms=(us, conjs))
        b = K.pairing_setup(K.to_dev(full), ident, nk, ch)
        torch.cuda.synchronize()
        assert rel(a[0], b[0]) < 1e-13 and rel(a[1], b[1]) < 1e-13
