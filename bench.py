"""Benchmark of the ladder-DGA hot path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--config 2] [--precision bf16x3|mixed|plain] [--no-compare]

A step is one full non-local SDE iteration (G(k) -> bubble -> double-counting
kernel -> ladder solves of both channels -> three-leg vertices and
susceptibilities -> kernel -> Sigma(k)) on synthetic inputs already resident
in HBM. The default workload is the 3-orbital t2g model on a 24x24 k-grid
(91 irreducible q, niw=30, niv=40: 5642 ladder solves of n=720 per step), its
local inputs averaged over the lattice's orbital symmetry and its kernel mapped
to the full zone with the per-k orbital transforms.
The JSON line reports (q, w) ladder solves per second of the whole step in
the headline form (bf16x3: complex64 LU whose trailing updates run as three
split-BF16 matrix-core products, refined in FP64; mixed: the same with FP32
matrix-core updates; plain: complex128), the step wall time, the same for the
plain complex128 form with the relative errors of chi*, chi and Sigma between
the two, the roofline of the dominant kernel (the trailing-matrix update of
the blocked LU) timed with HIP events on its launch stream, and a CPU
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
PEAK_FP32_TFLOPS = 157.3  # dense FP32 matrix cores
PEAK_BF16_TFLOPS = 2500.0  # dense BF16 matrix cores; the split-BF16 form needs three products
PEAK_HBM_GBS = 8000.0
# HBM bytes per trailing-update launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# (scaled by the copy-kernel factors), one "precision workload GB_per_launch" line each
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "trailing_update_traffic.txt")
# the headline metric, as the benchmark definition names it
METRIC = "(q,\u03c9) ladder solves/s + \u03a3 wall time: mixed vs FP64 on GPU at \u22641e-8 rel err"


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--precision", default="bf16x3", choices=["plain", "mixed", "bf16x3"],
                    help="form whose throughput is the headline value (mixed: complex64 LU on the FP32 matrix "
                         "cores + FP64 refinement; bf16x3: the same with split-BF16 MFMA trailing updates)")
    ap.add_argument("--no-compare", action="store_true", help="skip the plain FP64 comparison run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU-baseline sample time")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (3 times), the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counts to bytes")
    ap.add_argument("--ladder-only", action="store_true",
                    help="after set-up and one warm-up step, run only the batched ladder solve (for counter runs)")
    ap.add_argument("--batch", action="store_true",
                    help="instead of one model: 16 independent models of uneven sizes (1-3 orbitals, beta 5-40, "
                         "niv = niw = round(1.5 beta), k-grids 16^2-32^2, n = 16..1080) advanced together, the "
                         "ladder kernels autotuned, the whole step captured as one HIP graph and replayed "
                         "--steps times (default 10) as chained iterations with linear mixing 0.3")
    ap.add_argument("--batch-seed", type=int, default=180)
    ap.add_argument("--no-symmetric", action="store_true",
                    help="keep the local inputs of multi-orbital models as drawn (not averaged over the lattice's "
                         "orbital symmetry) and map the kernel to the full zone by the plain gather")
    ap.add_argument("--eliashberg", action="store_true",
                    help="instead of the SDE step: the linearised-gap matvec at the 2-orbital 48x48, 2np=128 "
                         "workload with the vertex stored in complex128 and in complex64")
    ap.add_argument("--eliashberg-solve", action="store_true",
                    help="instead of the SDE step: the whole Eliashberg solve of the 2-orbital 48x48, 2np=128 "
                         "workload from the ladder (bubble and ladder of both channels, three-leg vertices, "
                         "pairing vertices from the full vertices with the complex128 and the split-BF16 "
                         "factorisation, the four leading eigenpairs by Arnoldi), each stage timed")
    return ap.parse_args()


def eliashberg_matvec(steps, warmup):
    """Pairing matvec of the 2-orbital e_g model on a 48x48 k-grid (325 irreducible q), 2np = 128
    (n = nb^2 2np = 512, 2304 lattice vectors): set-up once, then `steps` applications of one and
    of four gap vectors, the vertex stored in complex128 and in complex64. HBM roofline: every
    application reads both transformed matrices once (2 N n^2 element sizes)."""
    import numpy as np
    import torch

    from dgamore_amd import kernels as K
    from dgamore_amd import synthetic as S
    from dgamore_amd.brillouin_zone import KGrid

    np2 = 128
    m = S.make_model(4, with_vertices=False)
    grid = KGrid(m.nk)
    N, nb, nq = int(np.prod(m.nk)), m.nb, grid.nk_irr
    n = nb * nb * np2
    dev = torch.device("cuda")
    gen = torch.Generator(device=dev).manual_seed(4004)
    # synthetic pairing vertex on the irreducible q, decaying in both fermionic frequencies
    v = torch.arange(np2, device=dev, dtype=torch.float64) - np2 / 2 + 0.5
    decay = 1.0 / (1.0 + (v[:, None] ** 2 + v[None, :] ** 2) / 400.0)
    gam = torch.randn((nq,) + (nb,) * 4 + (np2, np2), dtype=torch.complex128, device=dev, generator=gen) * decay
    g = K.g_full(K.to_dev(m.ek.reshape(N, nb, nb)), K.to_dev(m.sigma_loc[None]), m.mu, m.beta, np2 // 2 + 8)
    chi0pp = K.chi0_pp(g, np2 // 2)
    irrk_inv = K.to_dev(grid.irrk_inv, torch.int32)
    x = torch.randn((4, N, nb, nb, np2), dtype=torch.complex128, device=dev, generator=gen)
    out = {}
    results = {}
    for name, storage in (("c128", torch.complex128), ("c64", torch.complex64)):
        g_hat, g_flip = K.pairing_setup(gam, irrk_inv, m.nk, "sing", storage)
        es = torch.finfo(torch.float64 if storage == torch.complex128 else torch.float32).bits // 4
        for nvec in (1, 4):
            xv = x[:nvec].reshape((nvec,) + tuple(m.nk) + (nb, nb, np2)).contiguous()
            for _ in range(warmup):
                y = K.pairing_matvec(g_hat, g_flip, chi0pp, xv, m.nk, "sing", m.beta)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(steps):
                y = K.pairing_matvec(g_hat, g_flip, chi0pp, xv, m.nk, "sing", m.beta)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / steps
            mat_bytes = 2.0 * N * n * n * es
            results[f"{name}_nvec{nvec}"] = {"ms": ms, "matrix_GB": mat_bytes / 1e9,
                                             "achieved_GBps": mat_bytes / (ms * 1e-3) / 1e9,
                                             "frac_hbm_peak": mat_bytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS}
            out[(name, nvec)] = y
        del g_hat, g_flip
        torch.cuda.empty_cache()
    a, b = out[("c64", 4)], out[("c128", 4)]
    err = float((a - b).abs().max() / b.abs().max())
    print(json.dumps({"workload": f"eliashberg matvec: 2 orbitals, k-grid 48x48x1 ({nq} irreducible q), 2np={np2}, "
                                  f"n={n}, {N} lattice vectors", "results": results,
                      "rel_err_c64_vs_c128": err}), flush=True)


def eliashberg_solve():
    """The Eliashberg equation of the 2-orbital e_g model (U = 5, J = 0.8, beta = 20, 48x48 k-grid,
    325 irreducible q, niw = 127, niv = np = 64: 2np = 128, operator dimension 1,179,648) from the
    ladder, stage by stage: the SDE front (G, bubble, chi0^-1), the ladder sums of both channels,
    the three-leg vertices and physical susceptibilities, the singlet pairing vertex from the full
    vertices (complex128 LU, then split-BF16 LU), the R-space operator, and the four leading
    eigenpairs by Arnoldi with their residuals, for the pairing vertex of each factorisation."""
    import numpy as np
    import torch

    from dgamore_amd import eliashberg_solver as es
    from dgamore_amd import kernels as K
    from dgamore_amd import synthetic as S
    from dgamore_amd.brillouin_zone import KGrid
    from dgamore_amd.pipeline import SDEPipeline

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t

    nk, niw, niv = (48, 48, 1), 127, 64
    (m, pipe), t_model = timed(lambda: (lambda mm: (mm, SDEPipeline(mm)))(
        S.make_model(4, nk=nk, niw=niw, niv=niv, device=torch.device("cuda"))))
    front, t_front = timed(lambda: pipe.front(pipe.sigma_dmft))
    (aux, _), t_ladder = timed(lambda: K.ladder_solve_sum(front["chi0_inv"], pipe.gamma, pipe.u_r, pipe.v_r, m.beta))
    from dgamore_amd import _lib

    (aux_m, _), t_ladder_m = timed(lambda: K.ladder_solve_sum(front["chi0_inv"], pipe.gamma, pipe.u_r, pipe.v_r, m.beta,
                                                             _lib.MIXED_BF16X3_IR))
    ladder_err = float((aux_m - aux).abs().max() / aux.abs().max())
    del aux_m

    def three_leg():
        out = [K.vrg_chi_phys(front["chi0_inv"], aux[c], front["full"], front["core"], pipe.u_r[c], None, m.nb, m.beta)
               for c in range(2)]
        return [o[0] for o in out], [o[1] for o in out]

    (vrgs, chis), t_vrg = timed(three_leg)
    chi0pp = K.chi0_pp(front["g"], niv)
    solves = 2 * pipe.nq * (niw + 1)
    stages = {"model_setup_s": t_model, "front_s": t_front, "ladder_s": t_ladder, "ladder_solves": solves,
              "ladder_bf16x3_s": t_ladder_m, "ladder_bf16x3_rel_err": ladder_err, "three_leg_s": t_vrg}
    results = {}
    v0 = np.random.default_rng(7).standard_normal(48 * 48 * 4 * 2 * niv) + 0j
    for name, prec in (("c128", 0), ("bf16x3", 2)):
        (gs, _), t_pp = timed(lambda: es.pairing_vertices_from_ladder(front["chi0_inv"], pipe.gamma, pipe.u_r,
                                                                       pipe.v_r, vrgs, chis, m.beta, m.nb, niv,
                                                                       precision=prec))
        op, t_op = timed(lambda: es.PairingOperator(gs, chi0pp, KGrid(nk), m.beta, "sing", "fp64"))
        del gs
        t0 = time.perf_counter()
        lams, _, res = es.solve_eliashberg_arnoldi(op, 4, tol=1e-12, v0=v0)
        t_arn = time.perf_counter() - t0
        t0 = time.perf_counter()
        lks, _, rks = es.solve_eliashberg_krylov_schur(op, 4, tol=1e-12, v0=v0)
        t_ks = time.perf_counter() - t0
        results[name] = {"pairing_vertex_s": t_pp, "operator_setup_s": t_op, "arnoldi_s": t_arn,
                         "lambda": [[float(x.real), float(x.imag)] for x in lams], "residuals": res.tolist(),
                         "krylov_schur_s": t_ks, "krylov_schur_residuals": rks.tolist(),
                         "krylov_schur_vs_arnoldi": float(np.abs(lks - lams).max() / np.abs(lams).max())}
        del op
        torch.cuda.empty_cache()
        print(f"[bench] eliashberg {name}: pairing vertex {t_pp:.2f} s, operator {t_op:.2f} s, Arnoldi {t_arn:.2f} s, "
              f"device Krylov-Schur {t_ks:.2f} s, lambda {np.round(lams, 6).tolist()}", file=sys.stderr, flush=True)
    la = np.array([complex(*x) for x in results["c128"]["lambda"]])
    lb = np.array([complex(*x) for x in results["bf16x3"]["lambda"]])
    print(json.dumps({"workload": "eliashberg solve: 2 orbitals (e_g), U=5, J=0.8, beta=20, k-grid 48x48 "
                                  f"({pipe.nq} irreducible q), niw={niw}, 2np={2 * niv}, singlet, operator "
                                  f"dimension {48 * 48 * 4 * 2 * niv}",
                      "stages": stages, "results": results,
                      "lambda_rel_err_bf16x3_vs_c128": float(np.abs(la - lb).max() / np.abs(la).max())}), flush=True)


def batch_graph(steps, warmup, seed, symmetric=True, precision="plain"):
    """16 seeded models of uneven sizes in one ModelBatch: autotune every ladder shape, time one eager
    step, capture the step of all models as one HIP graph, check one replay against the eager step,
    then time `steps` chained replays (whole self-consistency iterations with mixing 0.3: every
    model's mu secant and occupations run inside the graph)."""
    import torch

    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import ModelBatch

    t0 = time.perf_counter()
    cfgs = S.batch_configs(seed=seed)
    models = []
    for c in cfgs:
        models.append(S.make_model(c, device=torch.device("cuda"), symmetric=symmetric))
        print(f"[bench] built {c['name']} (n={c['nb'] ** 2 * 2 * c['niv']}) {time.perf_counter() - t0:.1f}s",
              file=sys.stderr, flush=True)
    from dgamore_amd import _lib

    prec = {"plain": _lib.PLAIN_FP64, "mixed": _lib.MIXED_FP32_IR, "bf16x3": _lib.MIXED_BF16X3_IR}[precision]
    mb = ModelBatch(models, mixing=0.3, precision=prec)
    t1 = time.perf_counter()
    tuned = mb.tune(verbose=True)
    t_tune = time.perf_counter() - t1
    print(f"[bench] autotuned {len(tuned)} ladder shapes in {t_tune:.1f}s", file=sys.stderr, flush=True)
    stream = torch.cuda.current_stream()
    # every step is a whole self-consistency iteration: each model's mu secant and occupations run on
    # the device after its new Sigma
    mb.start_scf()
    state0 = [(p.mu_dev.clone(), p.occ.clone(), p.occ_k.clone()) for p in mb.pipes]

    def restore():
        for p, (mu, occ, occ_k) in zip(mb.pipes, state0):
            p.mu_dev.copy_(mu)
            p.occ.copy_(occ)
            p.occ_k.copy_(occ_k)

    for _ in range(max(1, warmup)):
        restore()
        eager = mb.step(scf=True)
    restore()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    eager = mb.step(scf=True)
    e1.record(stream)
    torch.cuda.synchronize()
    ms_eager = e0.elapsed_time(e1)
    eager_mu = [float(p.mu_dev.item()) for p in mb.pipes]
    print(f"[bench] eager step of all models {ms_eager:.2f} ms", file=sys.stderr, flush=True)
    compare = None
    if prec != _lib.PLAIN_FP64:
        # the same iteration with every ladder solve in complex128
        mb.precision = _lib.PLAIN_FP64
        restore()
        torch.cuda.synchronize()
        e0.record(stream)
        plain = mb.step(scf=True)
        e1.record(stream)
        torch.cuda.synchronize()
        ms_plain = e0.elapsed_time(e1)
        plain_mu = [float(p.mu_dev.item()) for p in mb.pipes]
        mb.precision = prec
        restore()
        torch.cuda.synchronize()
        compare = {"eager_ms_per_step_plain": ms_plain, "eager_speedup_vs_plain": ms_plain / ms_eager,
                   "sigma_rel_err_vs_plain": max(float((e - q).abs().max() / q.abs().max())
                                                 for e, q in zip(eager, plain)),
                   "mu_err_vs_plain": max(abs(a - b) for a, b in zip(eager_mu, plain_mu))}
        print(f"[bench] plain eager step {ms_plain:.2f} ms; {compare}", file=sys.stderr, flush=True)
    mb.capture(chain=True, scf=True)
    start = [s.clone() for s in mb.graph_sigmas]
    state0 = [(p.mu_dev.clone(), p.occ.clone(), p.occ_k.clone()) for p in mb.pipes]
    outs = mb.replay()
    torch.cuda.synchronize()
    err = max(float(torch.linalg.vector_norm(o - e) / torch.linalg.vector_norm(e)) for o, e in zip(outs, eager))
    mu_err = max(abs(float(p.mu_dev.item()) - m) for p, m in zip(mb.pipes, eager_mu))
    for s, s0 in zip(mb.graph_sigmas, start):
        s.copy_(s0)
    restore()
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(steps):
        mb.replay()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    finite = all(bool(torch.isfinite(o).all()) for o in mb.graph_outs)
    delta = max(float(torch.linalg.vector_norm(o - s) / torch.linalg.vector_norm(o))
                for o, s in zip(mb.graph_outs, start))
    solves = mb.solves_per_step()
    sizes = [p.m.nb ** 2 * 2 * p.m.niv for p in mb.pipes]
    print(json.dumps({
        "metric": METRIC,
        "value": solves / (ms * 1e-3), "unit": "solves/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128" if prec == _lib.PLAIN_FP64 else f"{precision} LU + c128 refinement (n > 64), c128 (n <= 64)",
        "data": "synthetic (16 seeded DMFT-like models, random-perturbation local vertices)",
        "config": {"workload": f"batch of {len(models)} models (seed {seed}): 1-3 orbitals, beta 5-40, niv=niw="
                               f"round(1.5 beta), k-grids 16^2-32^2, compound sizes {min(sizes)}..{max(sizes)}; "
                               "one HIP graph per SCF iteration of all models (mu secant and occupations "
                               "on the device), linear mixing 0.3",
                   "models": [p.m.name for p in mb.pipes], "n_compound": sizes, "solves_per_step": solves},
        "eager_ms_per_step": ms_eager, "graph_speedup": ms_eager / ms,
        "ladder_tflops_per_s": mb.ladder_flops_per_step() / (ms * 1e-3) / 1e12,
        "replay_vs_eager_rel_err": err, "replay_vs_eager_mu_err": mu_err, "finite": finite,
        "mu_after_replays": [float(p.mu_dev.item()) for p in mb.pipes],
        "sigma_change_over_replays": delta,
        "autotune": {f"n={t['n']},nrhs={t['nrhs']}": {"choice": t["name"], "source": t["source"],
                                                      "times_ms": t.get("times_ms", {})} for t in tuned},
        "autotune_seconds": t_tune,
        "precision": precision, "compare_plain": compare,
    }), flush=True)


def copy_calibration():
    import torch

    n = (1 << 30) // 16
    a = torch.ones(n, dtype=torch.complex128, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(3):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 3
    print(json.dumps({"copy_bytes_read": n * 16, "copy_bytes_written": n * 16, "copies": 4,
                      "copy_ms": ms, "copy_GBps": 2 * n * 16 / (ms * 1e-3) / 1e9}), flush=True)


def cpu_baseline(model, target_s):
    """Time the reference's per-q path restated in NumPy/LAPACK (the oracle) on bounded samples of
    the workload, on this host's cores:
    * ladder solves (numpy.linalg.solve with the nb^2 indicator right-hand sides, the reference's
      v2 path): the headline value runs one single-threaded solver per core in parallel, as the
      reference runs one MPI rank per core with OMP_NUM_THREADS=1 (README.md:125), in complex128 and
      in complex64 (the precision the reference stores); beside it one solver with BLAS threads =
      the cores, and one single-threaded solver;
    * the whole per-q chain of one SDE iteration for one irreducible q-point (direct bubble,
      double-counting kernel, both channels' ladders, three-leg vertex, susceptibility, kernel)
      and the Sigma sum over the full zone, extrapolated to the step (nq times the per-q chain
      plus one Sigma sum)."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from dgamore_amd.brillouin_zone import KGrid
    from oracle import dga_oracle as O

    nb, niv, niw, beta = model.nb, model.niv, model.niw, model.beta
    g = model.giwk()
    grid = KGrid(model.nk)
    q_list = grid.get_irrq_list()[1:2]
    chi0 = O.bubble_direct(g, niw, niv, q_list, beta)
    inv = O.block_invert_vn1(chi0)
    zero_v = np.zeros((1,) + (nb,) * 4)

    rhs = O.ladder_rhs(nb, 2 * niv)

    def timed(seconds, dtype):
        """(q, w, channel) solves of one q-point: the ladder matrices of both channels assembled and
        solved with the indicator right-hand sides in `dtype`, repeated for `seconds`."""
        solves = 0
        b = np.broadcast_to(rhs.astype(dtype), (niw + 1,) + rhs.shape)
        t0 = time.perf_counter()
        while True:
            for ch in ("dens", "magn"):
                a = O.ladder_matrix(inv[0], model.gamma[ch], O.u_as_channel(model.u, ch), zero_v[0], beta)
                np.linalg.solve(a.astype(dtype), b)
                solves += niw + 1
            el = time.perf_counter() - t0
            if el > seconds:
                return solves, el

    threads = min(len(os.sched_getaffinity(0)), 16)

    def q_parallel(seconds, dtype):
        """The reference's q-parallel layout on this host: `threads` independent single-threaded
        solvers (its MPI ranks with OMP_NUM_THREADS=1), here as host threads, each assembling and
        solving its own ladder matrices (LAPACK releases the interpreter lock)."""
        from concurrent.futures import ThreadPoolExecutor

        with threadpool_limits(limits=1), ThreadPoolExecutor(threads) as ex:
            res = list(ex.map(lambda _: timed(seconds, dtype), range(threads)))
        return sum(r[0] for r in res), max(r[1] for r in res)

    solves, el = q_parallel(target_s, np.complex128)
    solves32, el32 = q_parallel(target_s / 3, np.complex64)
    with threadpool_limits(limits=threads):
        solves_blas, el_blas = timed(target_s / 3, np.complex128)
    with threadpool_limits(limits=1):
        solves1, el1 = timed(target_s / 3, np.complex128)
    # one q-point's chain of the SDE iteration and the Sigma sum (all threads)
    with threadpool_limits(limits=threads):
        t0 = time.perf_counter()
        c0 = O.bubble_direct(g, niw, niv, q_list, beta)
        iv = O.block_invert_vn1(c0)
        kern = -O.dc_kernel(model.f_magn, c0, model.u, beta, niv)
        for ch, wgt in (("dens", 1.0), ("magn", 3.0)):
            ur = O.u_as_channel(model.u, ch)
            aux = O.ladder_sum_solve(iv, model.gamma[ch], ur, zero_v, beta)
            vrg = O.vrg(iv, aux, beta)
            chi = O.chi_phys_shell(O.sum_over_vn(aux, beta), c0.sum(-1) / beta**2, c0.sum(-1) / beta**2, ur, zero_v)
            kern = kern + wgt * O.kernel_r(vrg, chi, ur, zero_v, ch)
        t_q = time.perf_counter() - t0
        kern_all = np.zeros((grid.nk_irr,) + kern.shape[1:], dtype=complex)
        t0 = time.perf_counter()
        O.sigma_fft(kern_all, g, grid.irrk_inv, beta, niv, niw)
        t_sigma = time.perf_counter() - t0
    n = nb * nb * 2 * niv
    return {"value": solves / el, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{solves} (q,w,channel) ladder solves of n={n} via numpy.linalg.solve (complex128, the "
                      f"reference's solve-based path) in {el:.1f} s by {threads} single-threaded solvers in parallel "
                      f"(the reference's q-parallel ranks); complex64 the same way: {solves32} solves in {el32:.1f} s; "
                      f"one solver with BLAS threads = {threads}: {solves_blas} solves in {el_blas:.1f} s; one "
                      f"thread: {solves1} solves in {el1:.1f} s; the per-q chain of one SDE iteration (1 of "
                      f"{grid.nk_irr} irreducible q) in {t_q:.1f} s and the Sigma sum in {t_sigma:.1f} s",
            "value_1_thread": solves1 / el1, "value_1_thread_x_cores": solves1 / el1 * threads,
            "value_blas_threads": solves_blas / el_blas,
            "value_c64": solves32 / el32,
            "sde_step_s_extrapolated": t_q * grid.nk_irr + t_sigma,
            "per_q_chain_s": t_q, "sigma_sum_s": t_sigma}


def lambda_error(prec, name):
    """The lambda correction (spch) of the single-band workload (64x64, niw = niv = 60, n = 120),
    where it is defined, in the plain form and in the mixed form `prec`: both channels' lambda and
    the corrected Sigma of one step, and their relative differences. The local susceptibility
    sums are set below the uncorrected lattice sums (0.9 and 0.8 of them), so that both sum rules
    have a root."""
    import numpy as np
    import torch

    from dgamore_amd import _lib
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import SDEPipeline

    m = S.make_model(1)
    p0 = SDEPipeline(m)
    p0.step(keep=True)
    torch.cuda.synchronize()
    counts, N = p0.grid.irrk_count, p0.N

    def lattice_sum(x):
        half = x[:, 0, :].T.cpu().numpy()
        full = np.concatenate([np.conj(half[:, :0:-1]), half], axis=1)
        return (counts[:, None] * full).sum() / N

    sums = {"dens": 0.9 * lattice_sum(p0.last["chi_phys"][0]), "magn": 0.8 * lattice_sum(p0.last["chi_phys"][1])}
    out = {}
    for tag, pr in (("plain", _lib.PLAIN_FP64), (name, prec)):
        p = SDEPipeline(m, pr, lambda_correction="spch", chi_loc_sums=sums)
        p.step(keep=True)
        torch.cuda.synchronize()
        out[tag] = (dict(p.lambdas), p.last["sigma_nl"].clone())
    (lp, sp), (lm, sm) = out["plain"], out[name]
    return {"workload": f"{m.name}: 1 orbital, k-grid {m.nk[0]}x{m.nk[1]}, niw={m.niw}, niv={m.niv}, n={2 * m.niv}",
            "lambda_plain": {k: float(np.real(v)) for k, v in lp.items()},
            f"lambda_{name}": {k: float(np.real(v)) for k, v in lm.items()},
            "lambda_rel_err": max(abs(lm[k] - lp[k]) / abs(lp[k]) for k in lp),
            "sigma_rel_err": rel_err(sm, sp)}


def time_steps(pipe, steps, warmup, stream):
    import torch

    for i in range(warmup):
        pipe.step(keep=(i == warmup - 1))
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        pipe.step()
    ev1.record(stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps


def time_ladder(pipe, nrep, stream):
    """The batched ladder solve of both channels alone, on the launch stream."""
    import torch

    from dgamore_amd import kernels as K

    last = pipe.last
    la, lb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    la.record(stream)
    for _ in range(nrep):
        K.ladder_solve_sum(last["chi0_inv"], pipe.gamma, pipe.u_r, None, pipe.m.beta, pipe.precision,
                           check_info=False)
    lb.record(stream)
    torch.cuda.synchronize()
    return la.elapsed_time(lb) / nrep


def time_sigma(pipe, nrep, stream):
    """The Sigma sum (irr -> FBZ kernel, k-convolution with G, nu sum) alone."""
    import torch

    from dgamore_amd import kernels as K

    last, m = pipe.last, pipe.m
    la, lb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    la.record(stream)
    for _ in range(nrep):
        K.sigma_from_kernel(last["kernel"], pipe.irrk_inv, last["g"], m.nk, m.niw, m.beta)
    lb.record(stream)
    torch.cuda.synchronize()
    return la.elapsed_time(lb) / nrep


def dominant_kernel(pipe, stream, ms_ladder, precision_name):
    """Roofline of the dominant kernel of the ladder solve, timed with HIP events on its stream.

    Compound sizes above 128 run the blocked LU, whose trailing-matrix update dominates: the
    library brackets each of its launches with events (dgami_profile_*) and returns the
    algorithmic flops (8 m n k per problem) and the bytes of the C tile read and written plus the
    A and B panels read once. The ladder is matrix-core work (8/3 n^3 + 8 nb^2 n^2 flops per
    solve against kilobytes of operands), so the bound is the matrix-core peak of the precision
    (split-BF16: the BF16 peak / 3 products). `traffic` is the HBM bytes per launch measured with
    rocprofv3 FETCH_SIZE / WRITE_SIZE (profiles/trailing_update_traffic.txt).
    Smaller sizes run one register-resident kernel per call, timed as the whole ladder call.
    """
    import torch

    from dgamore_amd import _lib
    from dgamore_amd import kernels as K

    m = pipe.m
    peaks = {_lib.PLAIN_FP64: PEAK_FP64_TFLOPS, _lib.MIXED_FP32_IR: PEAK_FP32_TFLOPS,
             _lib.MIXED_BF16X3_IR: PEAK_BF16_TFLOPS / 3}
    peak_tf = peaks[pipe.precision]

    def profiled(lanes):
        # lanes: the ladder's problems on `lanes` streams (1: every launch alone on the GPU)
        old = os.environ.get("DGAMI_LADDER_STREAMS")
        os.environ["DGAMI_LADDER_STREAMS"] = str(lanes)
        try:
            _lib.profile_enable(True)
            K.ladder_solve_sum(pipe.last["chi0_inv"], pipe.gamma, pipe.u_r, None, m.beta, pipe.precision,
                               check_info=False)
            torch.cuda.synchronize()
            return _lib.profile_read()
        finally:
            _lib.profile_enable(False)
            if old is None:
                os.environ.pop("DGAMI_LADDER_STREAMS", None)
            else:
                os.environ["DGAMI_LADDER_STREAMS"] = old

    # the kernel's own rate is measured with the ladder on one stream: the default two lanes run
    # update launches beside the other lane's kernels, which stretches each launch's duration
    prof = profiled(1)
    # and, unless the whole run is on one stream, the launches as the default two lanes run them
    prof2 = None if os.environ.get("DGAMI_LADDER_STREAMS") == "1" else profiled(2)
    if prof["launches"] == 0:
        flops = pipe.solves_per_step() * pipe.ladder_flops_per_solve()
        tf = flops / (ms_ladder * 1e-3) / 1e12
        return {"bound": "mfma", "kernel": "register-resident ladder kernel (one launch per call)",
                "achieved": tf, "peak": peak_tf, "unit": "TFLOP/s", "frac": tf / peak_tf, "traffic": None,
                "launches": 1, "avg_launch_ms": ms_ladder}
    avg_ms = prof["ms"] / prof["launches"]
    tf = prof["flops"] / (prof["ms"] * 1e-3) / 1e12
    gbs = prof["bytes"] / (prof["ms"] * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        for line in open(TRAFFIC_FILE):
            f = line.split("#", 1)[0].split()
            if len(f) == 3 and f[0] == precision_name and f[1] == m.name:
                traffic = float(f[2]) * 1e9
    return {"kernel": "trailing-matrix update of the blocked LU (blk_gemm_*)", "launches": prof["launches"],
            "avg_launch_ms": avg_ms, "bound": "mfma", "achieved": tf, "peak": peak_tf, "unit": "TFLOP/s",
            "frac": tf / peak_tf, "traffic": traffic,
            "algorithmic_flops_per_launch": prof["flops"] / prof["launches"],
            "c_tile_and_panel_bytes_per_launch": prof["bytes"] / prof["launches"],
            "hbm_GBps_of_measured_traffic": None if traffic is None else traffic / (avg_ms * 1e-3) / 1e9,
            "hbm_frac_of_measured_traffic": None if traffic is None else traffic / (avg_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "measured_with": "the ladder on one stream (DGAMI_LADDER_STREAMS=1); the timed steps use two",
            "two_lanes": None if prof2 is None else {"launches": prof2["launches"],
                                                     "avg_launch_ms": prof2["ms"] / max(1, prof2["launches"])}}


def ladder_algorithmic_bytes(pipe):
    """HBM bytes one ladder call must move at least (complex128): Gamma_r(w) of both channels read,
    chi0^-1 read, chi*_sum of both channels written."""
    m = pipe.m
    n = m.nb**2 * 2 * m.niv
    nw = m.niw + 1
    nch = len(pipe.CHANNELS)
    return 16.0 * (nch * nw * n * n + nw * m.nb**4 * 2 * m.niv * pipe.nq * (1 + nch))


def rel_err(a, b):
    import torch

    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}); the multi-GPU sharding of the q-work "
              "list is described in the design notes but not run.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.eliashberg:
        eliashberg_matvec(args.steps, args.warmup)
        return
    if args.eliashberg_solve:
        eliashberg_solve()
        return
    if args.batch:
        batch_graph(args.steps if "--steps" in sys.argv else 10, args.warmup, args.batch_seed, not args.no_symmetric,
                    args.precision)
        return
    import torch

    from dgamore_amd import _lib
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import SDEPipeline

    prec = {"plain": _lib.PLAIN_FP64, "mixed": _lib.MIXED_FP32_IR, "bf16x3": _lib.MIXED_BF16X3_IR}[args.precision]
    t_setup = time.perf_counter()
    model = S.make_model(args.config, symmetric=not args.no_symmetric)
    pipe = SDEPipeline(model, prec)
    n = model.nb**2 * 2 * model.niv
    print(f"[bench] config {args.config} ({model.name}) nq_irr={pipe.nq} niw={model.niw} niv={model.niv} "
          f"n={n} set-up {time.perf_counter() - t_setup:.1f}s", file=sys.stderr, flush=True)
    stream = torch.cuda.current_stream()
    if args.ladder_only:
        pipe.step(keep=True)
        torch.cuda.synchronize()
        ms = time_ladder(pipe, 1, stream)
        print(json.dumps({"ladder_ms": ms, "solves": pipe.solves_per_step(), "precision": args.precision}), flush=True)
        return

    ms_step = time_steps(pipe, args.steps, args.warmup, stream)
    print(f"[bench] {args.precision}: {args.steps} steps, {ms_step:.3f} ms/step", file=sys.stderr, flush=True)
    ms_ladder = time_ladder(pipe, max(2, args.steps), stream)
    print(f"[bench] {args.precision}: ladder {ms_ladder:.3f} ms", file=sys.stderr, flush=True)
    solves = pipe.solves_per_step()
    flops = solves * pipe.ladder_flops_per_solve()
    achieved = flops / (ms_ladder * 1e-3) / 1e12
    peak = {_lib.PLAIN_FP64: PEAK_FP64_TFLOPS, _lib.MIXED_FP32_IR: PEAK_FP32_TFLOPS,
            _lib.MIXED_BF16X3_IR: PEAK_BF16_TFLOPS / 3}[prec]
    info_bad = int(torch.count_nonzero(pipe.last["info"]).item())
    ms_sigma = time_sigma(pipe, max(2, args.steps), stream)
    roof = dominant_kernel(pipe, stream, ms_ladder, args.precision)
    print(f"[bench] {args.precision}: Sigma sum {ms_sigma:.3f} ms; dominant kernel {roof['kernel']}: "
          f"{roof['launches']} launches, {roof['avg_launch_ms']:.3f} ms avg, {roof['frac']:.3f} of {roof['bound']} peak",
          file=sys.stderr, flush=True)

    compare = None
    if not args.no_compare and prec != _lib.PLAIN_FP64:
        ref = SDEPipeline(model, _lib.PLAIN_FP64)
        ms_ref = time_steps(ref, args.steps, args.warmup, stream)
        ms_ref_ladder = time_ladder(ref, max(2, args.steps), stream)
        print(f"[bench] plain: {ms_ref:.3f} ms/step, ladder {ms_ref_ladder:.3f} ms", file=sys.stderr, flush=True)
        a, b = pipe.last, ref.last
        compare = {
            "fp64_ms_per_step": ms_ref,
            "fp64_solves_per_s": solves / (ms_ref * 1e-3),
            "fp64_ladder_ms": ms_ref_ladder,
            "speedup_step": ms_ref / ms_step,
            "speedup_ladder": ms_ref_ladder / ms_ladder,
            "rel_err_vs_fp64": {
                "chi_aux_sum": rel_err(a["aux"], b["aux"]),
                "chi_phys": max(rel_err(x, y) for x, y in zip(a["chi_phys"], b["chi_phys"])),
                "sigma": rel_err(a["sigma_nl"], b["sigma_nl"]),
                "sigma_with_hf_and_tails": rel_err(a["sigma"], b["sigma"]),
            },
        }

    lam_check = None
    if not args.no_compare and prec != _lib.PLAIN_FP64:
        lam_check = lambda_error(prec, args.precision)
        print(f"[bench] lambda check: {lam_check}", file=sys.stderr, flush=True)

    cpu = None if args.no_cpu_baseline else cpu_baseline(model, args.cpu_seconds)
    line = {
        "metric": METRIC,
        "value": solves / (ms_step * 1e-3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {0: "c128", 1: "c64-LU+c128-refinement", 2: "bf16x3-LU+c128-refinement"}[prec],
        "data": "synthetic (seeded DMFT-like model, random-perturbation local vertex"
                + (")" if args.no_symmetric or model.orbital_transforms is None
                   else ", orbital-symmetric, full-zone map with per-k orbital transforms)"),
        "config": {"workload": f"{model.name}: {model.nb} orbitals, k-grid {'x'.join(map(str, model.nk))}, "
                               f"{pipe.nq} irreducible q, niw={model.niw}, niv={model.niv}, n={n}",
                   "nq_irr": pipe.nq, "niw": model.niw, "niv": model.niv, "n_compound": n,
                   "solves_per_step": solves, "precision": args.precision,
                   "ladder_streams": int(os.environ.get("DGAMI_LADDER_STREAMS") or 2)},
        "sigma_wall_ms": ms_step,
        "sigma_sum_ms": ms_sigma,
        "ladder_ms": ms_ladder,
        "ladder_solves_per_s": solves / (ms_ladder * 1e-3),
        "ladder_roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                            "frac": achieved / peak, "flops_per_solve": pipe.ladder_flops_per_solve()},
        "ladder_traffic": None if roof.get("traffic") is None else {
            "trailing_update_GB_per_call": roof["traffic"] * roof["launches"] / 1e9,
            "algorithmic_GB_per_call": ladder_algorithmic_bytes(pipe) / 1e9,
            "ratio": roof["traffic"] * roof["launches"] / ladder_algorithmic_bytes(pipe)},
        "singular_solves": info_bad,
        "compare_fp64": compare,
        "lambda_check": lam_check,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
