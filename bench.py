"""Benchmark of the ladder-DGA hot path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--config 1] [--precision plain|fp32|bf16x3]

A step is one full non-local SDE iteration (G(k) -> bubble -> double-counting
kernel -> ladder solves of both channels -> three-leg vertices and
susceptibilities -> kernel -> Sigma(k)) on synthetic inputs already resident
in HBM. The JSON line reports (q, w) ladder solves per second of the whole
step, the step wall time, the roofline of the dominant kernel (the batched
ladder solve) measured with HIP events on the launch stream, and a CPU
baseline: the NumPy restatement of the reference's per-q solve path timed on
the host on a bounded sample of the same workload.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--precision", default="plain", choices=["plain", "fp32", "bf16x3"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline sample time")
    return ap.parse_args()


def cpu_baseline(model, target_s):
    """Time the reference's per-q solve path (NumPy/LAPACK, complex128) on a bounded sample."""
    import numpy as np

    from dgamore_amd.brillouin_zone import KGrid
    from oracle import dga_oracle as O

    nb, niv, niw, beta = model.nb, model.niv, model.niw, model.beta
    g = model.giwk()
    grid = KGrid(model.nk)
    # bubble for a few irreducible q-points only (direct k-sum form), then the ladder solves
    nq_s = 2
    q_list = grid.get_irrq_list()[:nq_s]
    chi0 = O.bubble_direct(g, niw, niv, q_list, beta)
    inv = O.block_invert_vn1(chi0)
    zero_v = np.zeros((nq_s,) + (nb,) * 4)
    solves = 0
    t0 = time.perf_counter()
    while True:
        for ch in ("dens", "magn"):
            ur = O.u_as_channel(model.u, ch)
            O.ladder_sum_solve(inv[:1], model.gamma[ch], ur, zero_v[:1], beta)
            solves += niw + 1
        el = time.perf_counter() - t0
        if el > target_s:
            break
    threads = len(os.sched_getaffinity(0))
    return {"value": solves / el, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{solves} (q,w,channel) ladder solves of n={nb * nb * 2 * niv} via numpy.linalg.solve "
                      f"(complex128, reference v3 path) in {el:.1f} s, BLAS threads <= {threads}"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}); the multi-GPU sharding of the q-work "
              "list is described in DESIGN.md but not run.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from dgamore_amd import _lib
    from dgamore_amd import kernels as K
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import SDEPipeline

    prec = {"plain": _lib.PLAIN_FP64, "fp32": _lib.MIXED_FP32_IR, "bf16x3": _lib.MIXED_BF16X3_IR}[args.precision]
    t_setup = time.perf_counter()
    model = S.make_model(args.config)
    pipe = SDEPipeline(model, prec)
    print(f"[bench] config {args.config} ({model.name}) nq_irr={pipe.nq} niw={model.niw} niv={model.niv} "
          f"n={model.nb**2 * 2 * model.niv} set-up {time.perf_counter() - t_setup:.1f}s", file=sys.stderr, flush=True)

    stream = torch.cuda.current_stream()
    for i in range(args.warmup):
        pipe.step(keep=(i == args.warmup - 1))
    torch.cuda.synchronize()
    print("[bench] warm-up done", file=sys.stderr, flush=True)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        pipe.step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = ev0.elapsed_time(ev1) / args.steps
    print(f"[bench] {args.steps} steps: {ms_step:.3f} ms/step (host wall {wall / args.steps * 1e3:.3f} ms)",
          file=sys.stderr, flush=True)

    # dominant kernel: the batched ladder solve, timed alone on the same stream
    last = pipe.last
    nrep = max(3, args.steps)
    la, lb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    la.record(stream)
    for _ in range(nrep):
        K.ladder_solve_sum(last["chi0_inv"], pipe.gamma, pipe.u_r, None, model.beta, prec, check_info=False)
    lb.record(stream)
    torch.cuda.synchronize()
    ms_ladder = la.elapsed_time(lb) / nrep
    solves = pipe.solves_per_step()
    flops = solves * pipe.ladder_flops_per_solve()
    achieved = flops / (ms_ladder * 1e-3) / 1e12
    peak = PEAK_FP64_TFLOPS if prec == _lib.PLAIN_FP64 else PEAK_FP32_TFLOPS
    info_bad = int(torch.count_nonzero(last["info"]).item())

    cpu = None if args.no_cpu_baseline else cpu_baseline(model, args.cpu_seconds)
    line = {
        "metric": "(q,w) ladder solves/s + Sigma wall time (full non-local SDE iteration)",
        "value": solves / (ms_step * 1e-3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128" if prec == _lib.PLAIN_FP64 else ("c64+c128-IR" if prec == _lib.MIXED_FP32_IR else "bf16x3+c128-IR"),
        "data": "synthetic (seeded DMFT-like model, random-perturbation local vertex)",
        "config": {"workload": f"config {args.config}: {model.name}", "nk": list(model.nk), "nq_irr": pipe.nq,
                   "niw": model.niw, "niv": model.niv, "n_compound": model.nb**2 * 2 * model.niv,
                   "solves_per_step": solves, "precision": args.precision},
        "sigma_wall_ms": ms_step,
        "ladder_ms": ms_ladder,
        "ladder_solves_per_s": solves / (ms_ladder * 1e-3),
        "singular_solves": info_bad,
        "roofline": {"bound": "mfma", "kernel": "ladder_solve_sum", "achieved": achieved,
                     "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                     "flops_per_solve": pipe.ladder_flops_per_solve()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
