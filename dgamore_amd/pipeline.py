"""Device-resident non-local SDE iteration on one GPU.

One call of SDEPipeline.step runs the hot loop of calculate_self_energy_q
(nonlocal_sde.py:777-890) from the self-energy to the new self-energy, with
every intermediate kept in HBM in the library's q-fastest layouts:

    G(k)            dgami_g_full            (greens_function.py:125-142)
    chi0(q_irr)     dgami_bubble_irrq       (bubble_gen.py:33-69)
    chi0^-1, sums   dgami_chi0_prepare      (nonlocal_sde.py:824-835)
    kernel = -K_dc  dgami_dc_kernel         (nonlocal_sde.py:185-198, 817-820)
    chi*_sum        dgami_ladder_solve_sum  (both channels in one launch; nonlocal_sde.py:85-156)
    vrg, chi_r      dgami_vrg_chi_phys      (nonlocal_sde.py:159-182)
    kernel += K_d + 3 K_m   dgami_kernel_r  (nonlocal_sde.py:201-215, 838-857)
    Sigma(k)        dgami_sigma_from_kernel (nonlocal_sde.py:544-594, 874-879)
    Hartree, Fock   dgami_hartree_fock      (nonlocal_sde.py:31-69)
    new Sigma       dgami_sigma_assemble    (+ H + F + delta_sigma, DMFT tails, linear mixing;
                                             nonlocal_sde.py:884-893, 1145-1151)

iterate() adds the chemical-potential update and the new occupations of the
self-consistency loop (nonlocal_sde.py:895-914): the frequency sums run in
dgami_fill_sums, the secant iteration on mu on the host. iterate_device() and
capture_iteration() run the same loop with mu, its secant (dgami_update_mu), the
local moments (dgami_fit_smom) and the occupations (dgami_occupations) on the
device, so that one graph replay is one whole iteration.

The MPI distribution of the reference (irreducible q over ranks) becomes the
(channel, w, q) work list of a single ladder launch.
"""

import numpy as np
import torch

from dgamore_amd import _lib
from dgamore_amd import kernels as K
from dgamore_amd.brillouin_zone import KGrid


def u_as_channel(u, channel):
    """Local interaction in a spin channel (interaction.py:43-65)."""
    ut = np.einsum("abcd->adcb", u)
    return {"dens": 2 * u - ut, "magn": -ut, "sing": u + ut, "trip": u - ut}[channel]


class SDEPipeline:
    """Inputs of one model resident on the device; step() runs one SDE iteration."""

    CHANNELS = (("dens", 1.0), ("magn", 3.0))

    def __init__(self, model, precision=_lib.PLAIN_FP64, symmetries="two_dimensional_square", mixing=1.0,
                 orbital_transforms="model", lambda_correction=None, chi_loc_sums=None):
        """orbital_transforms: (us [N,nb,nb], conjs [N]) of the irreducible-to-full-zone map (the
        reference's auto symmetry mode), None for the plain gather, or "model" to take
        model.orbital_transforms when the model carries them.
        lambda_correction: None, "spch" (both channels) or "sp" (magnetic channel) correction of the
        physical susceptibility before the kernel (single band, nonlocal_sde.py:319-320, 339-401);
        chi_loc_sums: {"dens": ..., "magn": ...} sums of the local physical susceptibilities over the
        full bosonic range."""
        if lambda_correction is not None:
            if lambda_correction not in ("spch", "sp"):
                raise ValueError("Lambda correction type must be either 'spch' or 'sp'.")
            if model.nb != 1:
                raise ValueError("the lambda correction is defined for single-band models")
            if chi_loc_sums is None:
                raise ValueError("the lambda correction needs the local susceptibility sums")
        self.lambda_correction = lambda_correction
        self.chi_loc_sums = chi_loc_sums
        self.lambdas = {}
        self.m = model
        self.mixing = mixing
        self.precision = precision
        self.grid = KGrid(model.nk, symmetries)
        self.nq = self.grid.nk_irr
        self.nk = model.nk
        self.N = int(np.prod(model.nk))
        nb = model.nb
        self.irrk_ind = K.to_dev(self.grid.irrk_ind, torch.int32)
        self.irrk_inv = K.to_dev(self.grid.irrk_inv, torch.int32)
        if isinstance(orbital_transforms, str) and orbital_transforms == "model":
            orbital_transforms = getattr(model, "orbital_transforms", None)
        self.expand = None
        if orbital_transforms is not None:
            us, conjs = orbital_transforms
            src, u, cj, inv_ext = K.orbital_expand_tables(self.grid.irrk_inv, us, conjs)
            self.expand = (K.to_dev(src, torch.int32), K.to_dev(u), K.to_dev(cj, torch.int32))
            self.irrk_inv_ext = K.to_dev(inv_ext, torch.int32)
        self.ek = K.to_dev(model.ek.reshape(self.N, nb, nb))
        self.gamma = torch.stack([K.vertex_to_cmp(K.to_dev(model.gamma[ch])) for ch, _ in self.CHANNELS]).contiguous()
        self.u_r = torch.stack([K.orb_to_m(K.to_dev(u_as_channel(model.u, ch))) for ch, _ in self.CHANNELS])
        # non-local V(q) [nx,ny,nz,nb,nb,nb,nb] (reference order) when the model has one: V_dens = 2V,
        # V_magn = 0 on the irreducible q (interaction.py:160-182), the full grid for Fock
        v = getattr(model, "v", None)
        self.v_full = self.v_r = None
        if v is not None:
            v = np.asarray(v).reshape((self.N,) + (nb,) * 4)
            self.v_full = K.to_dev(v)
            v_irr = K.orb_to_m(K.to_dev(v[self.grid.irrk_ind]))
            self.v_r = torch.stack([2.0 * v_irr, torch.zeros_like(v_irr)]).contiguous()
        self.u_ref = K.to_dev(model.u)
        self.f_magn = K.to_dev(model.f_magn)
        self.sigma_dmft = K.to_dev(model.sigma_loc[None])  # local Sigma as [1][nb][nb][2 niv_s]
        self.niv_s = model.sigma_loc.shape[-1] // 2
        self.delta_sigma = None  # sigma_dmft - sigma_local on the core window (zero for the synthetic models)
        self.occ = self.occ_k = None
        self.last = {}

    # -- pieces -------------------------------------------------------------------------------
    def g_full(self, sigma, mu, niv):
        """sigma: [1 or N][nb][nb][2 niv_sigma] device tensor."""
        return K.g_full(self.ek, sigma, mu, self.m.beta, niv)

    def update_occupation(self, sigma, mu):
        """n, occ and occ_k of G(sigma, mu) with the k-dependent model (greens_function.py:212-240)."""
        from dgamore_amd import greens_function as gf
        from dgamore_amd.self_energy import fit_smom

        smom0, _ = fit_smom(sigma, self.m.beta)
        n, occ, occ_k = gf.get_fill_nonlocal(sigma, self.m.ek, mu, self.m.beta, smom0)
        self.occ = K.to_dev(occ)
        self.occ_k = K.to_dev(occ_k.reshape(self.N, self.m.nb, self.m.nb))
        return n

    def step(self, sigma=None, mu=None, keep=False, sigma_old=None):
        """One non-local SDE iteration from Sigma to the new, mixed Sigma [N,nb,nb,2 niv_s] on the device
        (the DMFT self-energy outside the core window). sigma, sigma_old: [1 or N][nb][nb][2 niv_s];
        sigma_old defaults to sigma. Eager steps check the ladder's info (mixed-precision items that
        did not converge are solved again in FP64; a singular matrix raises LinAlgError); a captured
        step leaves info in self.last_info for the caller to read after a replay."""
        sigma = self.sigma_dmft if sigma is None else sigma
        front = self.front(sigma, mu)
        aux, info = K.ladder_solve_sum(front["chi0_inv"], self.gamma, self.u_r, self.v_r, self.m.beta, self.precision,
                                       check_info=not K.capturing())
        self.last_info = info
        return self.back(front, aux, info, sigma, sigma_old, keep)

    def front(self, sigma, mu=None):
        """G(k), the bubble, chi0^-1 with its sums and the double-counting kernel: everything the
        ladder solves need (a dict of device tensors)."""
        m = self.m
        nb, beta = m.nb, m.beta
        mu = m.mu if mu is None else mu
        if self.occ is None:
            self.update_occupation(sigma, mu)
        g = self.g_full(sigma, mu, m.niv_g).reshape(m.nk + (nb, nb, 2 * m.niv_g))
        chi0 = K.bubble_irrq(g, m.nk, m.niw, m.niv_full, beta, self.irrk_ind)
        inv, full, core = K.chi0_prepare(chi0, nb, m.niv, beta)
        kern = K.dc_kernel(chi0, self.u_ref, self.f_magn, nb, m.niv, beta)
        return dict(g=g, chi0=chi0, chi0_inv=inv, full=full, core=core, kernel=kern)

    def ladder_model(self, front):
        """The ladder arguments of this model for kernels.ladder_solve_sum_multi."""
        return (front["chi0_inv"], self.gamma, self.u_r, self.v_r, self.m.beta)

    def back(self, front, aux, info, sigma, sigma_old=None, keep=False):
        """From the ladder sums to the new, mixed Sigma: vrg and chi, the kernel, the Sigma
        convolution, Hartree-Fock and the assembly with mixing."""
        m = self.m
        nb, beta = m.nb, m.beta
        g, inv, full, core, kern = (front[k] for k in ("g", "chi0_inv", "full", "core", "kernel"))
        chi0 = front["chi0"]
        vrgs, chis = [], []
        for c, (ch, weight) in enumerate(self.CHANNELS):
            v_c = None if self.v_r is None else self.v_r[c]
            vrg, chi = K.vrg_chi_phys(inv, aux[c], full, core, self.u_r[c], v_c, nb, beta)
            if self.lambda_correction == "spch" or (self.lambda_correction == "sp" and ch == "magn"):
                self.lambdas[ch] = self._lambda_correct(ch, chi, chis)
            K.kernel_r(vrg, chi, self.u_r[c], v_c, nb, ch, weight, kernel=kern, accumulate=True)
            vrgs.append(vrg)
            chis.append(chi)
        if self.expand is None:
            sig = K.sigma_from_kernel(kern, self.irrk_inv, g, m.nk, m.niw, beta)
        else:
            sig = K.sigma_from_kernel(K.kernel_orbital_expand(kern, *self.expand), self.irrk_inv_ext, g, m.nk, m.niw,
                                      beta)
        hart, fock = K.hartree_fock(self.u_ref, self.v_full, self.occ, self.occ_k, m.nk)
        old = sigma if sigma_old is None else sigma_old
        if self.mixing != 1.0 and old.shape[0] != self.N:
            old = old.expand(self.N, -1, -1, -1).contiguous()
        new = K.sigma_assemble(sig.reshape(self.N, nb, nb, -1), hart, fock, self.delta_sigma, self.sigma_dmft[0],
                               self.niv_s, old if self.mixing != 1.0 else None, self.mixing)
        if keep:
            self.last = dict(g=g, chi0=chi0, chi0_inv=inv, full=full, core=core, aux=aux, info=info, vrg=vrgs,
                             chi_phys=chis, kernel=kern, sigma_nl=sig, hartree=hart, fock=fock, sigma=new)
        return new

    def _lambda_correct(self, ch, chi, chis):
        """Replace chi (QW [nw][1][nq], in place) by 1/(1/chi + lambda); lambda from the FP64 root
        find on the host (nonlocal_sde.py:339-401, lambda_correction.py:28-73)."""
        from dgamore_amd.lambda_correction import apply_lambda, find_lambda_weighted

        def full_range(x):  # [nq, 2niw+1], chi(-w) = conj chi(w)
            half = x[:, 0, :].T.cpu().numpy()
            return np.concatenate([np.conj(half[:, :0:-1]), half], axis=1)

        counts, beta = self.grid.irrk_count, self.m.beta
        if self.lambda_correction == "spch":
            target = self.chi_loc_sums[ch] / beta
        else:
            dens = full_range(chis[0])
            target = (self.chi_loc_sums["dens"] + self.chi_loc_sums["magn"]
                      - (counts[:, None] * dens).sum() / self.N) / beta
        lam = find_lambda_weighted(full_range(chi), target, counts, beta, self.N)
        chi.copy_(K.to_dev(apply_lambda(chi.cpu().numpy(), lam)))
        return lam

    # -- HIP graph of one step ------------------------------------------------------------------
    def capture(self, mu=None, chain=False):
        """Capture one plain-FP64 SDE step as a HIP graph (torch.cuda.CUDAGraph). The captured step
        reads Sigma from the fixed buffer `graph_sigma` [N,nb,nb,2 niv_s] and writes the new, mixed
        Sigma to `graph_out`; replay() runs it again without host work. With chain, the graph ends
        by copying graph_out into graph_sigma, so that consecutive replays are consecutive
        iterations. mu, the occupations of the Hartree-Fock terms and the model are fixed at
        capture time. Every precision can be captured: the mixed forms keep their refinement
        status on the device, and after a replay `last_info` holds the ladder's info (nonzero
        entries: singular matrices or refinements that did not converge, which a captured step
        cannot re-solve)."""
        if self.lambda_correction is not None:
            raise ValueError("the lambda correction finds lambda on the host and cannot be captured as a graph")
        mu = self.m.mu if mu is None else mu
        nb = self.m.nb
        self.graph_sigma = self.sigma_dmft.expand(self.N, nb, nb, 2 * self.niv_s).contiguous().clone()
        # warm-up outside the capture: library workspaces and occupations are allocated here
        self.step(self.graph_sigma, mu)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        _lib.hold_workspaces(self.graph)
        with torch.cuda.graph(self.graph):
            self.graph_out = self.step(self.graph_sigma, mu)
            if chain:
                self.graph_sigma.copy_(self.graph_out)
        torch.cuda.synchronize()
        return self.graph

    def replay(self):
        """Run the captured step on the current contents of graph_sigma; returns graph_out."""
        self.graph.replay()
        return self.graph_out

    def iterate(self, n_iter, sigma=None, mu=None, n_target=None):
        """Self-consistency loop: SDE step, mu update to the target filling, new occupations
        (nonlocal_sde.py:777-914). Returns (sigma [N,nb,nb,2 niv_s], mu history)."""
        from dgamore_amd import greens_function as gf
        from dgamore_amd.self_energy import fit_smom

        sigma = self.sigma_dmft if sigma is None else sigma
        mu = self.m.mu if mu is None else mu
        if n_target is None:
            n_target = self.update_occupation(sigma, mu)
        history = [mu]
        for _ in range(n_iter):
            sigma = self.step(sigma, mu)
            smom0, _ = fit_smom(sigma, self.m.beta)
            new_mu = gf.update_mu(mu, n_target, self.m.ek, sigma, self.m.beta, smom0)
            if not np.isnan(new_mu):
                mu = self.mixing * new_mu + (1 - self.mixing) * mu
            history.append(mu)
            self.update_occupation(sigma, mu)
        return sigma, history

    # -- the self-consistency loop on the device ---------------------------------------------------
    def device_loop_state(self, sigma=None, mu=None, n_target=None):
        """Start the device-resident loop: mu as a float64 device scalar, the occupations at
        (sigma, mu) and the target filling (that of the start, as iterate() takes it)."""
        sigma = self.sigma_dmft if sigma is None else sigma
        mu = self.m.mu if mu is None else mu
        n = self.update_occupation(sigma, mu)
        self.n_target = n if n_target is None else n_target
        self.mu_dev = torch.tensor([float(mu)], dtype=torch.float64, device=self.ek.device)
        self.mu_info = torch.zeros(3, dtype=torch.float64, device=self.ek.device)
        self.hloc = K.to_dev(np.asarray(self.m.ek).reshape(self.N, self.m.nb, self.m.nb).mean(axis=0))
        return self.mu_dev

    def device_iteration(self, sigma):
        """One iteration of the self-consistency loop with no host work (nonlocal_sde.py:777-914):
        the SDE step at the device mu, the local moments of the new Sigma (dgami_fit_smom), the secant
        update of mu to the target filling with mixing (dgami_update_mu) and the occupations of the
        next Hartree-Fock terms (dgami_occupations). Returns the new Sigma; mu_dev, occ and occ_k are
        updated in place, mu_info holds the secant's result."""
        return self.scf_update(self.step(sigma, self.mu_dev))

    def scf_update(self, new):
        """The part of the self-consistency loop after the SDE step, on the device: local moments of
        the new Sigma, the secant update of mu_dev with mixing, and the occupations at the new mu."""
        beta = self.m.beta
        smom0, _ = K.fit_smom(new, beta)
        K.update_mu(self.mu_dev, self.n_target, self.ek, new, self.hloc, smom0, beta, self.mixing, info=self.mu_info)
        K.occupations(self.ek, new, smom0, self.mu_dev, beta, self.occ_k, self.occ)
        return new

    def iterate_device(self, n_iter, sigma=None, mu=None, n_target=None):
        """iterate() with the chemical potential and the occupations kept on the device. Returns
        (sigma, mu history)."""
        sigma = self.sigma_dmft if sigma is None else sigma
        self.device_loop_state(sigma, mu, n_target)
        history = [float(self.mu_dev.item())]
        for _ in range(n_iter):
            sigma = self.device_iteration(sigma)
            history.append(float(self.mu_dev.item()))
        return sigma, history

    def capture_iteration(self, mu=None, n_target=None, chain=True):
        """Capture one whole self-consistency iteration (device_iteration) as a HIP graph: Sigma from
        graph_sigma to graph_out, mu in mu_dev and the occupations all updated by the graph, so that
        chained replays are consecutive iterations of the reference's loop without host work."""
        if self.lambda_correction is not None:
            raise ValueError("the lambda correction finds lambda on the host and cannot be captured as a graph")
        nb = self.m.nb
        self.graph_sigma = self.sigma_dmft.expand(self.N, nb, nb, 2 * self.niv_s).contiguous().clone()
        self.device_loop_state(self.graph_sigma, mu, n_target)
        mu0 = self.mu_dev.clone()
        occ0, occ_k0 = self.occ.clone(), self.occ_k.clone()
        self.device_iteration(self.graph_sigma)  # warm-up: workspaces are allocated outside the capture
        torch.cuda.synchronize()
        self.mu_dev.copy_(mu0)
        self.occ.copy_(occ0)
        self.occ_k.copy_(occ_k0)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        _lib.hold_workspaces(self.graph)
        with torch.cuda.graph(self.graph):
            self.graph_out = self.device_iteration(self.graph_sigma)
            if chain:
                self.graph_sigma.copy_(self.graph_out)
        torch.cuda.synchronize()
        return self.graph

    # -- accounting ---------------------------------------------------------------------------
    def solves_per_step(self):
        """(q, w, channel) ladder solves in one iteration: 2 channels x nq_irr x (niw+1)."""
        return len(self.CHANNELS) * self.nq * (self.m.niw + 1)

    def ladder_flops_per_solve(self):
        """Algorithmic flops of one solve: complex LU (8/3 n^3) + nb^2 triangular RHS pairs (8 nb^2 n^2)."""
        n = self.m.nb**2 * 2 * self.m.niv
        return 8.0 / 3.0 * n**3 + 8.0 * self.m.nb**2 * n**2


class ModelBatch:
    """Independent models advanced together. One step runs, for every model, the front part of the
    SDE iteration (G, bubble, chi0^-1, double counting), then the ladder solves of all models in one
    dgami_ladder_solve_sum_multi call (the small compound sizes share register-kernel launches, the
    large ones take the blocked LU, every shape in the kernel the autotuner picked), then each
    model's back part to its new, mixed Sigma. capture() records the whole step of all models as
    one HIP graph; with chain, consecutive replays are consecutive self-consistency iterations, at
    fixed mu and occupations, or (scf) with every model's mu secant and occupations on the device
    in the graph. The reference runs one model per invocation
    (nonlocal_sde.py:777-914); this replaces a loop of such runs."""

    def __init__(self, models, mixing=0.3, symmetries="two_dimensional_square", precision=_lib.PLAIN_FP64):
        """precision: the form of the blocked-LU models' ladder solves (the register-resident ones,
        compound size <= 64 by default, stay FP64)."""
        self.precision = precision
        self.pipes = [SDEPipeline(m, precision, symmetries, mixing) for m in models]

    def tune(self, **kw):
        """Autotune the ladder kernel of every distinct shape of the batch: the FP64 choice (which
        also decides the register-resident models), and with a mixed precision the blocked-LU
        variant of the shapes the FP64 choice does not put in the register kernel."""
        from dgamore_amd import autotune

        out, seen = [], set()
        for p in self.pipes:
            key = (p.m.nb, p.m.niv, p.nq, p.m.niw + 1)
            if key not in seen:
                seen.add(key)
                r = autotune.tune_ladder(*key, nchan=len(p.CHANNELS), **kw)
                out.append(r)
                if self.precision != _lib.PLAIN_FP64 and autotune.FAMILIES.get(r["choice"] // 256) != "reg":
                    out.append(autotune.tune_ladder(*key, nchan=len(p.CHANNELS), precision=self.precision, **kw))
        return out

    def step(self, sigmas=None, mus=None, keep=False, scf=False):
        """One step of every model. scf: each model's mu is its device mu (start_scf) and the step
        ends with its secant update of mu and its new occupations, a whole self-consistency
        iteration of every model."""
        sigmas = [p.sigma_dmft for p in self.pipes] if sigmas is None else sigmas
        if scf:
            mus = [p.mu_dev for p in self.pipes]
        mus = [None] * len(self.pipes) if mus is None else mus
        fronts = [p.front(s, mu) for p, s, mu in zip(self.pipes, sigmas, mus)]
        outs = K.ladder_solve_sum_multi([p.ladder_model(f) for p, f in zip(self.pipes, fronts)],
                                        check_info=not K.capturing(), precision=self.precision)
        for p, (_, info) in zip(self.pipes, outs):
            p.last_info = info
        news = [p.back(f, aux, info, s, None, keep) for p, f, (aux, info), s in zip(self.pipes, fronts, outs, sigmas)]
        if scf:
            news = [p.scf_update(n) for p, n in zip(self.pipes, news)]
        return news

    def start_scf(self, sigmas=None):
        """Device mu, occupations and target filling of every model at its start (SDEPipeline.device_loop_state)."""
        sigmas = [None] * len(self.pipes) if sigmas is None else sigmas
        for p, s in zip(self.pipes, sigmas):
            p.device_loop_state(s)

    def capture(self, chain=True, scf=False):
        """Record one step of every model as one HIP graph reading graph_sigmas and writing
        graph_outs (and, with chain, copying graph_outs back into graph_sigmas). scf: the step is a
        whole self-consistency iteration with each model's mu and occupations on the device."""
        self.graph_sigmas = [p.sigma_dmft.expand(p.N, p.m.nb, p.m.nb, 2 * p.niv_s).contiguous().clone()
                             for p in self.pipes]
        if scf:
            self.start_scf(self.graph_sigmas)
            saved = [(p.mu_dev.clone(), p.occ.clone(), p.occ_k.clone()) for p in self.pipes]
        self.step(self.graph_sigmas, scf=scf)  # warm-up: workspaces and occupations
        torch.cuda.synchronize()
        if scf:  # the warm-up moved mu and the occupations; the graph starts from the initial state
            for p, (mu, occ, occ_k) in zip(self.pipes, saved):
                p.mu_dev.copy_(mu)
                p.occ.copy_(occ)
                p.occ_k.copy_(occ_k)
            torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        _lib.hold_workspaces(self.graph)
        with torch.cuda.graph(self.graph):
            self.graph_outs = self.step(self.graph_sigmas, scf=scf)
            if chain:
                for s, o in zip(self.graph_sigmas, self.graph_outs):
                    s.copy_(o)
        torch.cuda.synchronize()
        return self.graph

    def replay(self):
        self.graph.replay()
        return self.graph_outs

    def solves_per_step(self):
        return sum(p.solves_per_step() for p in self.pipes)

    def ladder_flops_per_step(self):
        return sum(p.solves_per_step() * p.ladder_flops_per_solve() for p in self.pipes)

