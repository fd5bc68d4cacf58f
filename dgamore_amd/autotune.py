"""Autotuner of the batched ladder solves.

The library can run one ladder shape -- compound size n = nb^2 * 2 niv, nrhs = nb^2 right-hand
sides, a precision mode -- with several kernels: register-tile variants and column-cyclic
variants for small n (plain FP64 only), and blocked-LU panel shapes (panel width 32 or 16,
rows per panel thread) for any n up to 1536. tune_ladder() times every candidate on a
synthetic, well-conditioned batch of the caller's shape, drops any whose result differs from
the library's default kernel, stores the fastest in the library's process-wide table (which
every later ladder call of that shape reads) and records it in a JSON cache keyed by the GPU
name, so that a later process loads the choice without timing again.

The reference has no counterpart: NumPy's inverse and solve pick their LAPACK kernels
themselves (four_point.py:620-684).
"""

import ctypes
import json
import os
import sys
import time

import numpy as np
import torch

from dgamore_amd import _lib
from dgamore_amd import kernels as K

FAMILIES = {1: "reg", 2: "col", 3: "blocked"}
PRECISION_NAMES = {_lib.PLAIN_FP64: "plain", _lib.MIXED_FP32_IR: "mixed", _lib.MIXED_BF16X3_IR: "bf16x3"}
CACHE_ENV = "DGAMI_TUNE_CACHE"

# tuning batches are cut to about this many flops per timed call (about 0.1-0.3 s of FP64 work)
TUNE_FLOPS = 2.0e13


def kernel_name(kid):
    """'reg1', 'col0', 'blocked1', ... for a kernel id (family * 256 + variant)."""
    return f"{FAMILIES.get(kid // 256, '?')}{kid % 256}"


def _status(rc, what):
    if rc < 0:
        raise ValueError(f"{what}: {_lib.last_error()}")
    return rc


def candidates(n, nrhs, precision=_lib.PLAIN_FP64):
    """Kernel ids that can run (n, nrhs, precision)."""
    lib = _lib.load()
    cnt = _status(lib.dgami_ladder_tune_candidates(n, nrhs, precision, None, 0), "ladder_tune_candidates")
    ids = (ctypes.c_int * max(cnt, 1))()
    _status(lib.dgami_ladder_tune_candidates(n, nrhs, precision, ctypes.cast(ids, ctypes.c_void_p), cnt),
            "ladder_tune_candidates")
    return list(ids[:cnt])


def set_choice(n, nrhs, precision, kid):
    """Store (or with kid=None remove) the kernel of a shape in the library's table."""
    _status(_lib.load().dgami_ladder_tune_set(n, nrhs, precision, -1 if kid is None else int(kid)),
            "ladder_tune_set")


def get_choice(n, nrhs, precision=_lib.PLAIN_FP64):
    """Stored kernel id of a shape, or None."""
    kid = _lib.load().dgami_ladder_tune_get(n, nrhs, precision)
    return None if kid < 0 else kid


def clear():
    _lib.load().dgami_ladder_tune_clear()


# ---------------------------------------------------------------------------
# JSON cache
# ---------------------------------------------------------------------------
def cache_path():
    return os.environ.get(CACHE_ENV) or os.path.join(os.path.expanduser("~"), ".cache", "dgamore_amd",
                                                     "ladder_tune.json")


def _device_key():
    return torch.cuda.get_device_name(torch.cuda.current_device())


def _key(n, nrhs, precision, nitems):
    # problems per call enter as a power-of-two bucket: a full GPU and a nearly empty one may prefer
    # different kernels
    return f"{n}:{nrhs}:{PRECISION_NAMES[precision]}:{1 << max(0, int(nitems) - 1).bit_length()}"


def _read_cache(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _write_cache(path, data):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home leaves the choice in this process only


# ---------------------------------------------------------------------------
# timing
# ---------------------------------------------------------------------------
def _problem(nb, niv, nq, nw, nchan, seed=0):
    """A well-conditioned ladder batch on the device: chi0^-1 = 2 (orbital identity) + noise,
    Gamma of spectral norm ~0.5, small U (beta = 1, so A = chi0^-1 delta + Gamma - U)."""
    g = torch.Generator(device=K.device()).manual_seed(seed)
    dev = K.device()
    n = nb * nb * 2 * niv

    def crandn(*shape):
        return torch.complex(torch.randn(shape, generator=g, device=dev, dtype=torch.float64),
                             torch.randn(shape, generator=g, device=dev, dtype=torch.float64))

    inv = 0.05 * crandn(nw, nb**4, 2 * niv, nq)
    for o1 in range(nb):
        for o2 in range(nb):
            inv[:, ((o1 * nb + o2) * nb + o1) * nb + o2] += 2.0
    gamma = (0.25 / np.sqrt(n)) * crandn(nchan, nw, n, n)
    u_r = (0.1 / n) * crandn(nchan, nb**4)
    return inv.contiguous(), gamma.contiguous(), u_r.contiguous()


def _time_ms(fn, reps):
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        fn()
    end.record()
    end.synchronize()
    return start.elapsed_time(end) / reps


def tune_ladder(nb, niv, nq, nw, nchan=2, precision=_lib.PLAIN_FP64, reps=3, use_cache=True, verbose=False):
    """Pick the kernel of the ladder shape (nb, niv) at nchan * nw * nq problems per call.

    Returns {"n", "nrhs", "choice", "name", "times_ms", "source"}; the choice is also stored in
    the library's table. With use_cache, a choice cached for this GPU is applied without timing.
    """
    n, nrhs = nb * nb * 2 * niv, nb * nb
    items = nchan * nw * nq
    key = _key(n, nrhs, precision, items)
    path = cache_path()
    cands = candidates(n, nrhs, precision)
    if not cands:
        raise ValueError(f"no ladder kernel can run n={n} with {nrhs} right-hand sides")
    cached = _read_cache(path).get(_device_key(), {}).get(key) if use_cache else None
    if cached is not None and cached["choice"] in cands:
        set_choice(n, nrhs, precision, cached["choice"])
        return dict(cached, n=n, nrhs=nrhs, source="cache")
    if len(cands) == 1:
        set_choice(n, nrhs, precision, cands[0])
        return dict(n=n, nrhs=nrhs, choice=cands[0], name=kernel_name(cands[0]), times_ms={}, source="only")

    # cut the batch to a bounded amount of work: fewer q-points, the same frequencies and channels
    flops = 8.0 / 3.0 * n**3 + 8.0 * nrhs * n**2
    nq_t = int(max(1, min(nq, TUNE_FLOPS // max(flops * nchan * nw, 1.0))))
    inv, gamma, u_r = _problem(nb, niv, nq_t, nw, nchan)
    set_choice(n, nrhs, precision, None)
    ref, _ = K.ladder_solve_sum(inv, gamma, u_r, None, 1.0, precision, check_info=False)
    tol = 1e-9 if precision == _lib.PLAIN_FP64 else 1e-7
    scale = float(ref.abs().max())
    times = {}
    for kid in cands:
        set_choice(n, nrhs, precision, kid)
        out, info = K.ladder_solve_sum(inv, gamma, u_r, None, 1.0, precision, check_info=False)
        torch.cuda.synchronize()
        err = float((out - ref).abs().max()) / max(scale, 1e-300)
        if int(info.abs().sum()) != 0 or not err < tol:
            if verbose:
                print(f"[tune] n={n} {kernel_name(kid)} rejected (err {err:.2e})", file=sys.stderr, flush=True)
            continue
        times[kid] = _time_ms(lambda: K.ladder_solve_sum(inv, gamma, u_r, None, 1.0, precision, check_info=False),
                              reps)
        if verbose:
            print(f"[tune] n={n} nrhs={nrhs} {PRECISION_NAMES[precision]} {kernel_name(kid)}: "
                  f"{times[kid]:.3f} ms ({nchan * nw * nq_t} problems)", file=sys.stderr, flush=True)
    if not times:
        raise RuntimeError(f"every ladder kernel candidate failed for n={n}")
    best = min(times, key=times.get)
    set_choice(n, nrhs, precision, best)
    entry = dict(choice=best, name=kernel_name(best), times_ms={kernel_name(k): round(v, 4) for k, v in times.items()},
                 problems_timed=nchan * nw * nq_t, tuned_at=time.strftime("%Y-%m-%d %H:%M:%S"))
    if use_cache:
        data = _read_cache(path)
        data.setdefault(_device_key(), {})[key] = entry
        _write_cache(path, data)
    return dict(entry, n=n, nrhs=nrhs, source="timed")


def tune_models(models, nchan=2, precision=_lib.PLAIN_FP64, **kw):
    """Tune every distinct ladder shape of a list of synthetic models (objects with nb, niv, niw and
    an nq attribute or k-grid); returns the list of tune_ladder results."""
    seen, out = set(), []
    for m in models:
        nq = m.nq if hasattr(m, "nq") else m.grid.nk_irr
        key = (m.nb, m.niv, nq, m.niw + 1)
        if key in seen:
            continue
        seen.add(key)
        out.append(tune_ladder(m.nb, m.niv, nq, m.niw + 1, nchan, precision, **kw))
    return out
