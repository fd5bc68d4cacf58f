"""Self-energy mixing, restart files and the reference-shaped containers of the mirror (CPU only).

The Pulay / Anderson cases follow the behaviours the reference pins in tests/test_mixing.py:170-405
(fixed point, no mutation of sigma_old, tails from sigma_new, the returned object, history order),
with the history read through the module-level read_last_n_sigmas_from_files as the reference
reads it. On top of those, the least-squares steps are checked against what the mathematics fixes:
on a linear self-consistency map the accelerated iterations must reach the fixed point much faster
than linear mixing, and Pulay with a full history reaches it exactly.
"""
from copy import deepcopy
from unittest.mock import patch

import numpy as np
import pytest

from dgamore_amd import config, nonlocal_sde
from dgamore_amd.four_point import FourPoint, LocalFourPoint, SpinChannel
from dgamore_amd.self_energy import SelfEnergy
from oracle import dga_oracle as O
from tests.conftest import golden

NIV, NIV_CORE = 8, 4


@pytest.fixture(autouse=True)
def mixing_config():
    saved = dict(vars(config.self_consistency)), dict(vars(config.box)), dict(vars(config.output))
    config.box.niv_core = NIV_CORE
    yield
    for ns, d in zip((config.self_consistency, config.box, config.output), saved):
        ns.__dict__.clear()
        ns.__dict__.update(d)


def sigma(value, nk=1, nb=1, niv=NIV):
    return SelfEnergy(np.full((nk, nb, nb, 2 * niv), value, dtype=complex), (nk, 1, 1), True, True)


def mat(value, nk=1, nb=1, niv=NIV):
    return np.full((nk, nb, nb, 2 * niv), value, dtype=complex)


def run(strategy, new, old, dmft, files, mixing=0.5, n_hist=3, current_iter=10):
    config.self_consistency.mixing_strategy = strategy
    config.self_consistency.mixing = mixing
    config.self_consistency.mixing_history_length = n_hist
    with patch("dgamore_amd.nonlocal_sde.read_last_n_sigmas_from_files", return_value=files):
        return nonlocal_sde.apply_mixing_strategy(new, old, dmft, current_iter)


CORE = slice(NIV - NIV_CORE, NIV + NIV_CORE)


@pytest.mark.parametrize("strategy", ["pulay", "anderson"])
def test_returns_sigma_new_itself(strategy):
    new = sigma(2.0)
    out = run(strategy, new, sigma(1.0), sigma(0.0), [mat(v) for v in (0.5, 0.8, 1.0)])
    assert out is new and isinstance(out, SelfEnergy) and out.mat.shape == (1, 1, 1, 2 * NIV)


@pytest.mark.parametrize("strategy", ["pulay", "anderson"])
def test_fixed_point(strategy):
    """The last proposal reproduced by the new result (zero residual) is returned unchanged. (The
    reference's all-equal history has no usable direction and takes the linear branch on the device;
    that case runs in the GPU tests.)"""
    v = 3.0 + 1.0j
    out = run(strategy, sigma(v), sigma(v), sigma(0.0), [mat(v + d) for d in (0.5, 0.2j, 0.0)])
    np.testing.assert_allclose(out.mat[..., CORE], v, atol=1e-12)


@pytest.mark.parametrize("strategy", ["pulay", "anderson"])
def test_sigma_old_untouched_and_tails_from_sigma_new(strategy):
    old = sigma(99.0)
    before = old.mat.copy()
    out = run(strategy, sigma(2.0), old, sigma(0.0), [mat(v) for v in (1.0, 1.5, 1.8)])
    np.testing.assert_array_equal(old.mat, before)
    np.testing.assert_allclose(out.mat[..., :NIV - NIV_CORE], 2.0)
    np.testing.assert_allclose(out.mat[..., NIV + NIV_CORE:], 2.0)
    assert np.all(np.isfinite(out.mat))


def test_anderson_differs_from_linear_and_depends_on_history_order():
    new, old = sigma(2.0 + 0.5j), sigma(1.0 + 0.2j)
    files = [mat(complex(v + 0.3j * v)) for v in (0.5, 1.0, 1.5)]
    a = run("anderson", deepcopy(new), old, sigma(0.0), files)
    lin = 0.5 * new.mat + 0.5 * old.mat
    assert not np.allclose(a.mat[..., CORE], lin[..., CORE], atol=1e-6)
    files = [mat(float(v)) for v in (0.5, 1.2, 1.5)]
    fwd = run("anderson", sigma(2.0), sigma(1.0), sigma(0.0), files)
    rev = run("anderson", sigma(2.0), sigma(1.0), sigma(0.0), files[::-1])
    assert not np.allclose(fwd.mat[..., CORE], rev.mat[..., CORE], atol=1e-6)


def _linear_history(seed, n_hist, scale):
    """A linear self-consistency map F(x) = A x + b on the core window (two complex unknowns, one
    momentum, one band: four real dimensions) and the history the reference's scheme pairs up:
    proposals [dmft, F(dmft), ..., F^n(dmft)] with results [F(dmft), ..., F^(n+1)(dmft)], i.e. the
    saved self-energies of n unmixed iterations started from the DMFT self-energy. Returns
    (dmft, files, sigma_new, fixed point of F on the core window)."""
    rng = np.random.default_rng(seed)
    nc, niv = 1, 3
    config.box.niv_core = nc
    core = slice(niv - nc, niv + nc)
    a = scale * np.linalg.qr(rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2)))[0]
    b = rng.standard_normal(2) + 1j * rng.standard_normal(2)
    tail = rng.standard_normal(2 * niv) + 1j * rng.standard_normal(2 * niv)

    def full(vals):
        x = tail.copy()
        x[core] = vals
        return x.reshape(1, 1, 1, 2 * niv)

    xs = [full(rng.standard_normal(2) + 1j * rng.standard_normal(2))]
    for _ in range(n_hist + 1):
        xs.append(full(a @ xs[-1][0, 0, 0, core] + b))
    fixed = np.linalg.solve(np.eye(2) - a, b)
    return xs[0], xs[1:n_hist + 1], xs[n_hist + 1], fixed, core


@pytest.mark.parametrize("strategy,alpha,scale", [("pulay", 0.3, 0.9), ("pulay", 1.0, 0.5), ("anderson", 1.0, 0.3)])
def test_exact_on_a_linear_map_with_a_spanning_history(strategy, alpha, scale):
    """With as many history differences as real unknowns, the least squares of both schemes sees the
    whole linear map: Pulay lands on the fixed point for any mixing; undamped Anderson (alpha = 1)
    too, up to its Tikhonov term (cutoff 1e-5 of the largest singular value)."""
    dmft, files, new, fixed, core = _linear_history(seed=3, n_hist=4, scale=scale)
    config.self_consistency.mixing_strategy = strategy
    config.self_consistency.mixing = alpha
    config.self_consistency.mixing_history_length = 4
    out = nonlocal_sde.apply_mixing_strategy(SelfEnergy(new.copy(), (1, 1, 1), True, True),
                                             SelfEnergy(files[-1], (1, 1, 1), True, True),
                                             SelfEnergy(dmft, (1, 1, 1), True, True), 5, history=files)
    tol = 1e-10 if strategy == "pulay" else 1e-6
    assert np.abs(out.mat[0, 0, 0, core] - fixed).max() < tol * np.abs(fixed).max()
    np.testing.assert_array_equal(np.delete(out.mat[0, 0, 0], core), np.delete(new[0, 0, 0], core))


def test_pulay_short_history_is_not_exact():
    """A planted mistake in the scheme's bookkeeping would not be caught if every history were
    exact: with fewer differences than unknowns the step is a genuine least-squares compromise."""
    dmft, files, new, fixed, core = _linear_history(seed=3, n_hist=2, scale=0.9)
    config.self_consistency.mixing_strategy = "pulay"
    config.self_consistency.mixing = 1.0
    config.self_consistency.mixing_history_length = 2
    out = nonlocal_sde.apply_mixing_strategy(SelfEnergy(new.copy(), (1, 1, 1), True, True),
                                             SelfEnergy(files[-1], (1, 1, 1), True, True),
                                             SelfEnergy(dmft, (1, 1, 1), True, True), 3, history=files)
    err = np.abs(out.mat[0, 0, 0, core] - fixed).max() / np.abs(fixed).max()
    assert 1e-6 < err < np.abs(new[0, 0, 0, core] - fixed).max() / np.abs(fixed).max()


def test_singular_history_falls_back_to_linear_core():
    """An all-equal history has no usable direction: the core window must not be touched by the
    least squares (the device linear mix runs instead, tested on the GPU)."""
    for strategy in ("pulay", "anderson"):
        assert nonlocal_sde._history_mixing(strategy, mat(2.0), mat(2.0), [mat(2.0)] * 3, 3, 0.5) is None


# -- restart files (nonlocal_sde.py:677-716) ----------------------------------------------------
def test_read_last_sigmas_orders_by_iteration_across_directories(tmp_path):
    out, prev = tmp_path / "out", tmp_path / "prev"
    out.mkdir()
    prev.mkdir()
    for it in (1, 2, 9, 10, 11):
        np.save(out / f"sigma_dga_iteration_{it}.npy", mat(float(it)))
    for it in (3, 12):
        np.save(prev / f"sigma_dga_iteration_{it}.npy", mat(float(it)))
    np.save(out / "sigma_dga_iteration_x.npy", mat(-1.0))  # not an iteration file
    got = nonlocal_sde.read_last_n_sigmas_from_files(4, str(out), str(prev))
    assert [g[0, 0, 0, 0].real for g in got] == [9.0, 10.0, 11.0, 12.0]
    got = nonlocal_sde.read_last_n_sigmas_from_files(3, str(out), "")
    assert [g[0, 0, 0, 0].real for g in got] == [9.0, 10.0, 11.0]


def test_get_starting_sigma_resumes_from_the_last_iteration(tmp_path):
    default = sigma(0.0)
    assert nonlocal_sde.get_starting_sigma(str(tmp_path / "missing"), default) == (default, 0)
    assert nonlocal_sde.get_starting_sigma(str(tmp_path), default) == (default, 0)
    for it in (2, 7, 10):
        SelfEnergy(mat(float(it), nk=4), (2, 2, 1), True, True).save(str(tmp_path), f"sigma_dga_iteration_{it}")
    config.lattice.nk = (2, 2, 1)
    s, it = nonlocal_sde.get_starting_sigma(str(tmp_path), default)
    assert it == 10 and isinstance(s, SelfEnergy) and s.mat.shape == (4, 1, 1, 2 * NIV)
    assert np.all(s.mat == 10.0) and s.has_compressed_q_dimension and s.full_niv_range


# -- lambda correction from the reference's files (nonlocal_sde.py:339-401) ----------------------
@pytest.mark.parametrize("kind", ["spch", "sp"])
def test_lambda_correction_reads_the_reference_files(tmp_path, kind):
    from dgamore_amd import lambda_correction as lc
    from dgamore_amd.brillouin_zone import KGrid

    names = {"chi_dens_loc": f"lc_{kind}_chi_dens_loc.npy", "chi_magn_loc": f"lc_{kind}_chi_magn_loc.npy",
             "chi_phys_q_dens": f"lc_{kind}_chi_phys_q_dens_before.npy"}
    for dst, src in names.items():
        np.save(tmp_path / f"{dst}.npy", golden(src))
    config.lattice.q_grid = KGrid((4, 4, 1))
    config.sys.beta = 12.5
    config.lambda_correction.type = kind
    config.output.output_path = str(tmp_path)

    def load(name, ch):
        return FourPoint.load(str(tmp_path / name) if name.endswith(".npy") else name, ch, nq=(4, 4, 1),
                              num_vn_dimensions=0)

    magn = FourPoint(golden(f"lc_{kind}_chi_phys_q_magn_before.npy"), SpinChannel.MAGN, (4, 4, 1), 1, 0, False,
                     has_compressed_q_dimension=True)
    dens = FourPoint(golden(f"lc_{kind}_chi_phys_q_dens_before.npy"), SpinChannel.DENS, (4, 4, 1), 1, 0, False,
                     has_compressed_q_dimension=True)
    dens_before = dens.mat.copy()
    magn_c = lc.perform_lambda_correction(magn)
    dens_c = lc.perform_lambda_correction(dens)
    ref = golden(f"lc_{kind}_chi_phys_q_magn.npy")
    assert np.abs(magn_c.mat - ref).max() / np.abs(ref).max() < 1e-5  # test_lambda_correction.py:124-125
    lines = open(tmp_path / f"lambda_{kind}.txt").read().split()
    if kind == "sp":
        assert np.array_equal(dens_c.mat, dens_before)
        assert lines == ["lambda_magn:", lines[1]] and np.isclose(float(lines[1]), 4.281153)
    else:
        lam = dict(zip(lines[::2], map(float, lines[1::2])))
        assert np.isclose(lam["lambda_magn:"], 4.328781)  # test_lambda_correction.py:71
        # The reference's density value (-37.450340) is where its complex64 iteration is left after
        # maxiter steps without meeting eps: the starting pole -min(1/Re chi(w=0)), at which the sum
        # rule is off by ~1e3. In FP64 the same iteration converges to a root of the sum rule.
        before = O.to_full_niw(dens_before, 0).squeeze()
        assert np.isclose(lc.get_lambda_start(before), -37.450340)
        counts = config.lattice.q_grid.irrk_count[:, None]
        loc_sum = O.to_full_niw(golden("lc_spch_chi_dens_loc.npy"), 0).sum() / 12.5

        def rule(lam_):
            return (counts * lc.apply_lambda(before, lam_)).sum() / 12.5 / 16 - loc_sum

        assert abs(rule(lam["lambda_dens:"]).real) < 1e-10 < 1.0 < abs(rule(-37.450340))


# -- containers (n_point_base.py:107-187, four_point.py:49-71, 303-327, self_energy.py:109-168) ---
def test_fourpoint_algebra_and_files(tmp_path):
    rng = np.random.default_rng(5)
    shape = (3, 2, 2, 2, 2, 5, 4)  # q, o1..o4, w (full range), v
    a = FourPoint(rng.standard_normal(shape) + 1j * rng.standard_normal(shape), SpinChannel.NONE, (3, 1, 1), 1, 1,
                  True, True, True)
    b = FourPoint(rng.standard_normal(shape) + 1j * rng.standard_normal(shape), SpinChannel.NONE, (3, 1, 1), 1, 1,
                  True, True, True)
    k = -a
    k += b
    k += 3 * b
    np.testing.assert_allclose(k.mat, -a.mat + 4 * b.mat)
    np.testing.assert_allclose((a - b).mat, a.mat - b.mat)
    half = deepcopy(b).to_half_niw_range()
    np.testing.assert_allclose((a + half).mat, a.mat + deepcopy(half).to_full_niw_range().mat)  # a's full range
    with pytest.raises(ValueError):
        a + FourPoint(np.ones((3, 2, 2, 2, 2, 5)), SpinChannel.NONE, (3, 1, 1), 1, 0, True, True, True)
    p = a.permute_orbitals("abcd->cbad")
    np.testing.assert_allclose(p.mat, np.einsum("qabcd...->qcbad...", a.mat))
    a.save(str(tmp_path), "obj")
    assert a.full_niw_range  # saving in the half range leaves the object as it was
    saved = np.load(tmp_path / "obj.npy")
    np.testing.assert_allclose(saved, a.mat[..., 2:, :])
    back = FourPoint.load(str(tmp_path / "obj.npy"), num_vn_dimensions=1, nq=(3, 1, 1)).to_full_niw_range()
    np.testing.assert_allclose(back.mat[..., 2:, :], a.mat[..., 2:, :])
    a.free()
    assert a.mat is None


def test_local_fourpoint_load_and_permute(tmp_path):
    g = golden("srvo3_gamma_dens_loc.npy")[..., :3, :4, :4]
    np.save(tmp_path / "gamma_dens_loc.npy", g)
    loc = LocalFourPoint.load(str(tmp_path / "gamma_dens_loc.npy"), SpinChannel.DENS)
    assert loc.n_bands == 3 and not loc.full_niw_range and loc.channel == SpinChannel.DENS
    f = 2 * loc.permute_orbitals("abcd->cbad")
    np.testing.assert_allclose(f.mat, 2 * np.einsum("abcd...->cbad...", g))


def test_self_energy_ranges_and_sums():
    rng = np.random.default_rng(1)
    pos = rng.standard_normal((4, 2, 2, 3)) + 1j * rng.standard_normal((4, 2, 2, 3))
    s = SelfEnergy(pos.copy(), (2, 2, 1), full_niv_range=False, has_compressed_q_dimension=True)
    s.to_full_niv_range()
    np.testing.assert_allclose(s.mat[..., 3:], pos)
    np.testing.assert_allclose(s.mat[..., :3], np.conj(pos[..., ::-1]))
    hart = rng.standard_normal((1, 2, 2, 1))
    fock = rng.standard_normal((4, 2, 2, 1))
    t = s + hart + fock
    np.testing.assert_allclose(t.mat, s.mat + hart + fock)
    dmft = SelfEnergy(rng.standard_normal((1, 2, 2, 10)) + 0j, (1, 1, 1), True, True)
    c = t.concatenate_self_energies(dmft)
    assert c.mat.shape == (4, 2, 2, 10)
    np.testing.assert_allclose(c.mat[..., 2:8], t.mat)
    np.testing.assert_allclose(c.mat[..., :2], np.broadcast_to(dmft.mat[..., :2], (4, 2, 2, 2)))
    np.testing.assert_allclose(c.cut_niv(3).mat, t.mat)
    np.testing.assert_allclose((-(0.5 * c) / 2).mat, -c.mat / 4)
