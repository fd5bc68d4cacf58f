"""HIP-graph capture of one plain-FP64 SDE step: every replay equals the eager step on the same
input, with eager work and synchronisations between the replays."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def test_graph_replays_match_eager_steps():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import SDEPipeline

    say("graph capture of the SDE step, config-0 model 8x8 niw=4 niv=6, mixing 0.3, 4 replays")
    m = S.make_model(0, nk=(8, 8, 1), niw=4, niv=6)
    pipe = SDEPipeline(m, mixing=0.3)
    pipe.capture()
    for it in range(4):
        inp = pipe.graph_sigma.clone()
        out = pipe.replay().clone()
        torch.cuda.synchronize()
        ref = pipe.step(inp, m.mu)  # eager, same input
        torch.cuda.synchronize()
        err = float((out - ref).abs().max() / ref.abs().max())
        assert err < 1e-13, f"replay {it}: {err}"
        # linear-mixing iteration: the new Sigma is the next input
        pipe.graph_sigma.copy_(out)
        torch.cuda.synchronize()
    assert np.isfinite(out.abs().max().item())


def test_replay_after_workspace_growth():
    """A captured step keeps its workspace buffers when a later eager call needs larger ones:
    capture a small model, step a larger model eagerly (every workspace the bubble, the ladder
    and the self-energy use grows), then replay and compare with the small model's eager step."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import SDEPipeline

    say("workspace growth under a live graph: capture 4x4 niw=2 niv=72 (blocked LU), eager 8x8 niw=3 niv=80")
    small = S.make_model(0, nk=(4, 4, 1), niw=2, niv=72)
    pipe = SDEPipeline(small, mixing=0.3)
    pipe.capture()
    inp = pipe.graph_sigma.clone()
    big = S.make_model(0, nk=(8, 8, 1), niw=3, niv=80)
    SDEPipeline(big, mixing=0.3).step()
    torch.cuda.synchronize()
    out = pipe.replay().clone()
    torch.cuda.synchronize()
    ref = pipe.step(inp, small.mu)
    torch.cuda.synchronize()
    err = float((out - ref).abs().max() / ref.abs().max())
    assert err < 1e-13, err
