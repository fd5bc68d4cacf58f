"""Uneven-size batches: every kernel the autotuner may pick agrees with the oracle, the tuner and
its cache, and a batch of independent models advanced in one step and in one HIP graph gives
each model's own single-model result."""
import sys

import numpy as np
import pytest

from oracle import dga_oracle as O
from tests.conftest import ladder_case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

RTOL = 1e-10


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture(scope="module")
def K():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dgamore_amd import kernels

    return kernels


@pytest.fixture()
def A(K):
    from dgamore_amd import autotune

    autotune.clear()
    yield autotune
    autotune.clear()


def _solve(K, case, precision):
    chi0_inv, gamma, u, v, beta = case
    out, info = K.ladder_solve_sum(K.ref4_to_qf(K.to_dev(chi0_inv)), K.vertex_to_cmp(K.to_dev(gamma))[None].contiguous(),
                                   K.orb_to_m(K.to_dev(u))[None].contiguous(),
                                   K.orb_to_m(K.to_dev(v))[None].contiguous(), beta, precision)
    torch.cuda.synchronize()
    assert int(info.abs().sum()) == 0
    return K.qf_to_ref4(out[0], chi0_inv.shape[1])


@pytest.mark.parametrize("nb,niv,niw,nq", [
    (1, 8, 2, 3),    # n = 16
    (2, 5, 2, 3),    # n = 40, 4 right-hand sides
    (3, 4, 1, 2),    # n = 72, 9 right-hand sides
    (1, 60, 1, 2),   # n = 120
    (3, 12, 1, 1),   # n = 216
    (3, 40, 1, 1),   # n = 720
    (3, 60, 0, 1),   # n = 1080
])
def test_every_tuner_candidate_matches_oracle(K, A, nb, niv, niw, nq):
    from dgamore_amd import _lib

    n = nb * nb * 2 * niv
    case = ladder_case(nb, niv, niw, nq, seed=900 + n)
    ref = O.ladder_sum_solve(*case)
    for kid in A.candidates(n, nb * nb, _lib.PLAIN_FP64):
        say(f"tuner candidate {A.kernel_name(kid)} plain n={n} nb={nb} niw={niw} nq={nq}")
        A.set_choice(n, nb * nb, _lib.PLAIN_FP64, kid)
        assert rel(_solve(K, case, _lib.PLAIN_FP64), ref) < RTOL
    if n >= 216:
        for kid in A.candidates(n, nb * nb, _lib.MIXED_BF16X3_IR):
            say(f"tuner candidate {A.kernel_name(kid)} bf16x3 n={n}")
            A.set_choice(n, nb * nb, _lib.MIXED_BF16X3_IR, kid)
            assert rel(_solve(K, case, _lib.MIXED_BF16X3_IR), ref) < 1e-8


def test_tune_ladder_times_and_caches(K, A, tmp_path, monkeypatch):
    monkeypatch.setenv(A.CACHE_ENV, str(tmp_path / "tune.json"))
    say("tune_ladder n=40 (10 candidates) and n=216 (2), then from the cache")
    first = [A.tune_ladder(2, 5, 3, 3), A.tune_ladder(3, 12, 2, 2)]
    for r in first:
        assert r["source"] == "timed" and r["choice"] in A.candidates(r["n"], r["nrhs"])
        assert A.get_choice(r["n"], r["nrhs"]) == r["choice"] and len(r["times_ms"]) >= 2
    assert (tmp_path / "tune.json").exists()
    A.clear()
    again = [A.tune_ladder(2, 5, 3, 3), A.tune_ladder(3, 12, 2, 2)]
    assert [r["source"] for r in again] == ["cache", "cache"]
    assert [r["choice"] for r in again] == [r["choice"] for r in first]
    assert A.get_choice(216, 9) == first[1]["choice"]


def _small_batch():
    from dgamore_amd import synthetic as S

    cfgs = S.batch_configs(count=4, seed=11, betas=(5.0, 8.0), grids=(6, 8))
    return [S.make_model(c, niw=min(c["niw"], 5), niv=min(c["niv"], 6), device=torch.device("cuda")) for c in cfgs]


def test_model_batch_step_matches_single_models(K, A):
    from dgamore_amd.pipeline import ModelBatch, SDEPipeline

    models = _small_batch()
    say("model batch step vs single-model steps: " + ", ".join(f"{m.name} n={m.nb**2*2*m.niv}" for m in models))
    mb = ModelBatch(models, mixing=0.3)
    outs = mb.step()
    torch.cuda.synchronize()
    for m, out in zip(models, outs):
        ref = SDEPipeline(m, mixing=0.3).step()
        torch.cuda.synchronize()
        assert rel(out, ref.cpu().numpy()) < RTOL, m.name


def test_model_batch_graph_chain_matches_eager_iterations(K, A):
    from dgamore_amd.pipeline import ModelBatch

    models = _small_batch()
    say("model batch HIP graph, chained replays vs eager iterations (mixing 0.3, 3 iterations)")
    mb = ModelBatch(models, mixing=0.3)
    mb.tune()
    mb.capture(chain=True)
    start = [s.clone() for s in mb.graph_sigmas]
    sig = [s.clone() for s in start]
    for it in range(3):
        mb.replay()
        torch.cuda.synchronize()
        sig = mb.step(sig)  # eager iteration from the previous eager Sigma
        torch.cuda.synchronize()
        for g, e, m in zip(mb.graph_sigmas, sig, models):
            assert rel(g, e.cpu().numpy()) < 1e-12, (it, m.name)
    # the iteration moved Sigma
    assert max(float((g - s).abs().max()) for g, s in zip(mb.graph_sigmas, start)) > 1e-8


def test_config3_batch_at_full_size(K):
    """The 16-model batch of the benchmark (seed 180: 1-3 orbitals, beta 5-40, niv = niw =
    round(1.5 beta), k-grids 16^2-32^2, n = 16..1080), orbital-symmetric inputs as benchmarked:
    10 chained replays of the captured step, each equal to the eager step from the same Sigma,
    and every model's ladder sums on one irreducible q-point (both channels, every w) equal to
    the oracle's LU solves, the n = 1080 model included."""
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import ModelBatch, u_as_channel

    cfgs = S.batch_configs(seed=180)
    sizes = [c["nb"] ** 2 * 2 * c["niv"] for c in cfgs]
    say(f"config-3 batch seed 180: 16 models, n = {sorted(sizes)}")
    models = [S.make_model(c, device=torch.device("cuda"), symmetric=True) for c in cfgs]
    mb = ModelBatch(models, mixing=0.3)
    mb.capture(chain=True)
    sig = [s.clone() for s in mb.graph_sigmas]
    for it in range(10):
        mb.replay()
        torch.cuda.synchronize()
        sig = mb.step(sig, keep=(it == 0))
        torch.cuda.synchronize()
        for g, e, m in zip(mb.graph_sigmas, sig, models):
            err = float((g - e).abs().max() / e.abs().max())
            assert err < 1e-12, (it, m.name, err)
        if it == 0:
            for p in mb.pipes:
                m, last = p.m, p.last
                q_sel = [p.nq - 1]
                inv = K.qf_to_ref4(last["chi0_inv"], m.nb).cpu().numpy()[q_sel]
                zero_v = np.zeros((1,) + (m.nb,) * 4)
                for c, (ch, _) in enumerate(p.CHANNELS):
                    ref = O.ladder_sum_solve(inv, m.gamma[ch], u_as_channel(m.u, ch), zero_v, m.beta)
                    got = K.qf_to_ref4(last["aux"][c], m.nb).cpu().numpy()[q_sel]
                    assert rel(got, ref) < RTOL, (m.name, ch)
            say("ladder sums of every model match the oracle on one q-point; replays continue")
    assert all(bool(torch.isfinite(s).all()) for s in sig)


def test_model_batch_scf_graph_matches_eager_iterations(K, A):
    """The captured step of every model as a whole self-consistency iteration (mu secant and
    occupations on the device): chained replays equal eager scf steps, Sigma and every model's mu."""
    from dgamore_amd.pipeline import ModelBatch

    models = _small_batch()
    say("model batch HIP graph with the device mu secant, chained replays vs eager iterations (3)")
    mb = ModelBatch(models, mixing=0.3)
    mb.capture(chain=True, scf=True)
    eager = ModelBatch(models, mixing=0.3)
    sig = [s.clone() for s in mb.graph_sigmas]
    eager.start_scf(sig)
    mu0 = [float(p.mu_dev.item()) for p in mb.pipes]
    for it in range(3):
        mb.replay()
        torch.cuda.synchronize()
        sig = eager.step(sig, scf=True)
        torch.cuda.synchronize()
        for g, e, pg, pe, m in zip(mb.graph_sigmas, sig, mb.pipes, eager.pipes, models):
            assert rel(g, e.cpu().numpy()) < 1e-12, (it, m.name)
            assert abs(float(pg.mu_dev.item()) - float(pe.mu_dev.item())) < 1e-12, (it, m.name)
    assert max(abs(float(p.mu_dev.item()) - m) for p, m in zip(mb.pipes, mu0)) > 1e-8


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_mixed_model_batch_matches_plain_and_replays(K, A, precision):
    """A batch whose blocked-LU models (n > 64) take the complex64 LU with FP64 refinement: each
    model's new Sigma within 1e-8 of the complex128 batch, and a captured whole self-consistency
    iteration (mu secant on the device) replays equal to the eager mixed iteration."""
    from dgamore_amd import _lib
    from dgamore_amd.pipeline import ModelBatch

    prec = {"fp32": _lib.MIXED_FP32_IR, "bf16x3": _lib.MIXED_BF16X3_IR}[precision]
    models = _small_batch()
    sizes = [m.nb ** 2 * 2 * m.niv for m in models]
    assert max(sizes) > 64
    say(f"mixed ({precision}) model batch vs plain batch, n = {sizes}")
    plain = ModelBatch(models, mixing=0.3).step()
    mb = ModelBatch(models, mixing=0.3, precision=prec)
    mixed = mb.step()
    torch.cuda.synchronize()
    for m, a, b in zip(models, mixed, plain):
        err = float((a - b).abs().max() / b.abs().max())
        assert err < 1e-8, (m.name, err)
    mb.capture(chain=True, scf=True)
    eager = ModelBatch(models, mixing=0.3, precision=prec)
    sig = [s.clone() for s in mb.graph_sigmas]
    eager.start_scf(sig)
    for it in range(2):
        mb.replay()
        torch.cuda.synchronize()
        sig = eager.step(sig, scf=True)
        torch.cuda.synchronize()
        for g, e, pg, pe, m in zip(mb.graph_sigmas, sig, mb.pipes, eager.pipes, models):
            assert rel(g, e.cpu().numpy()) < 1e-12, (it, m.name)
            assert abs(float(pg.mu_dev.item()) - float(pe.mu_dev.item())) < 1e-12, (it, m.name)
        assert all(int(torch.count_nonzero(p.last_info)) == 0 for p in mb.pipes)


def test_config3_batch_mixed_matches_plain(K):
    """The 16-model batch of the benchmark (seed 180) with its blocked-LU models (n = 72..1080) in
    split-BF16 + FP64 refinement: one eager self-consistency iteration gives every model's Sigma
    within 1e-8 of the complex128 iteration, and every model's mu within 1e-10."""
    from dgamore_amd import _lib
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import ModelBatch

    cfgs = S.batch_configs(seed=180)
    say("config-3 batch seed 180: split-BF16 iteration vs complex128 iteration")
    models = [S.make_model(c, device=torch.device("cuda"), symmetric=True) for c in cfgs]
    mb = ModelBatch(models, mixing=0.3, precision=_lib.MIXED_BF16X3_IR)
    mb.start_scf()
    state0 = [(p.mu_dev.clone(), p.occ.clone(), p.occ_k.clone()) for p in mb.pipes]
    mixed = mb.step(scf=True)
    mu_mixed = [float(p.mu_dev.item()) for p in mb.pipes]
    for p, (mu, occ, occ_k) in zip(mb.pipes, state0):
        p.mu_dev.copy_(mu)
        p.occ.copy_(occ)
        p.occ_k.copy_(occ_k)
    mb.precision = _lib.PLAIN_FP64
    plain = mb.step(scf=True)
    torch.cuda.synchronize()
    errs = [float((a - b).abs().max() / b.abs().max()) for a, b in zip(mixed, plain)]
    dmu = [abs(a - float(p.mu_dev.item())) for a, p in zip(mu_mixed, mb.pipes)]
    say(f"  max Sigma rel. error {max(errs):.2e}, max mu difference {max(dmu):.2e}")
    assert max(errs) < 1e-8
    assert max(dmu) < 1e-10
