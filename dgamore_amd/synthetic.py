"""Synthetic DMFT-like inputs for the ladder-DGA hot path (host side, NumPy).

The reference reads its inputs (G, Sigma, the local vertices Gamma_r and
F_magn, U, V) from w2dynamics files (dmft_interface.py). Those files are not
available, so the workloads are built from a seed:

* lattice dispersion e_k on an nx x ny x 1 grid (hamiltonian.py:188-209 for
  one band: -2t(cos kx + cos ky) - 4t' cos kx cos ky - 2t''(cos 2kx + cos 2ky));
* local self-energy Sigma(v) = U^2 n(1-n)/(i v) (1 + 0.05 smooth noise);
* chemical potential mu from the filling of the lattice Green's function;
* local bubble chi0_loc(w,v) = -beta G_ad(v) G_cb(v-w) (bubble_gen.py:21-30),
  generalised susceptibility chi_r = chi0_loc + a Hermitian random
  perturbation of relative norm 0.1 (inter-orbital couplings 0.05), and from
  it Gamma_r = beta^2 (chi_r^-1 - chi0_loc^-1) (local_sde.py:32-37) and
  F_r = beta^2 (chi0^-1 - chi0^-1 chi_r chi0^-1) (local_sde.py:85-90);
* Kanamori U (hamiltonian.py:128-186), V^q = 0.

Everything is complex128. Arrays use the reference's layouts.
"""

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Model:
    name: str
    nb: int
    nk: tuple
    beta: float
    niw: int
    niv: int
    niv_shell: int
    mu: float
    ek: np.ndarray  # [nx,ny,nz,nb,nb]
    sigma_loc: np.ndarray  # [nb,nb,2*niv_sigma]
    u: np.ndarray  # [nb]*4 (reference order)
    gamma: dict = field(default_factory=dict)  # channel -> [nb]*4+[niw+1, 2niv, 2niv]
    f_magn: np.ndarray = None  # [nb]*4+[niw+1, 2niv_full, 2niv_full]
    seed: int = 0
    orbital_transforms: tuple = None  # (us [N,nb,nb], conjs [N]) of the irreducible-zone map, or None
    v: np.ndarray = None  # non-local V(q) [nx,ny,nz,nb,nb,nb,nb] (reference order), or None

    @property
    def niv_full(self):
        return self.niv + self.niv_shell

    @property
    def niv_g(self):
        """Fermionic box of G(k) that the bubble and Sigma need (reference: niv_cut, nonlocal_sde.py:754)."""
        return self.niw + self.niv_full + 10

    def giwk(self):
        """Lattice G(k,v) = [(iv+mu) - e_k - Sigma(v)]^-1 on the niv_g box (greens_function.py:125-142)."""
        return lattice_g(self.ek, self.sigma_loc, self.mu, self.beta, self.niv_g)


def vn(niv, beta):
    return np.pi / beta * (2 * np.arange(-niv, niv) + 1)


def kanamori(nb, U, J, Up=None):
    """U_abcd of the Kanamori model: U_llll=U, U_lmml=U_llmm=J, U_lmlm=U' (hamiltonian.py:148-184)."""
    Up = U - 2 * J if Up is None else Up
    u = np.zeros((nb,) * 4)
    for a in range(nb):
        for b in range(nb):
            for c in range(nb):
                for d in range(nb):
                    if a == b == c == d:
                        u[a, b, c, d] = U
                    elif (a == d and b == c) or (a == b and c == d):
                        u[a, b, c, d] = J
                    elif a == c and b == d:
                        u[a, b, c, d] = Up
    return u


def dispersion_1band(nk, t, tp, tpp):
    kx = 2 * np.pi * np.arange(nk[0]) / nk[0]
    ky = 2 * np.pi * np.arange(nk[1]) / nk[1]
    cx, cy = np.cos(kx)[:, None], np.cos(ky)[None, :]
    e = -2 * t * (cx + cy) - 4 * tp * cx * cy - 2 * tpp * (np.cos(2 * kx)[:, None] + np.cos(2 * ky)[None, :])
    return e.reshape(nk[0], nk[1], 1, 1, 1).astype(np.complex128)


def dispersion_t2g(nk, txy=1.0, txz=0.8, tyz=0.8, hyb=0.1):
    """3-orbital t2g-like nearest-neighbour model with a uniform inter-orbital hybridisation."""
    kx = 2 * np.pi * np.arange(nk[0]) / nk[0]
    ky = 2 * np.pi * np.arange(nk[1]) / nk[1]
    cx, cy = np.cos(kx)[:, None], np.cos(ky)[None, :]
    e = np.zeros((nk[0], nk[1], 1, 3, 3), dtype=np.complex128)
    e[..., 0, 0] = (-2 * txy * (cx + cy))[..., None]
    e[..., 1, 1] = (-2 * txz * cx - 2 * 0.1 * cy)[..., None]
    e[..., 2, 2] = (-2 * tyz * cy - 2 * 0.1 * cx)[..., None]
    off = (-2 * hyb * (cx + cy))[..., None]
    for a in range(3):
        for b in range(3):
            if a != b:
                e[..., a, b] = off
    return e


def dispersion_eg(nk, t=1.0, hyb=0.2):
    """2-orbital e_g-like model: two NN bands with a (cos kx - cos ky) inter-orbital term."""
    kx = 2 * np.pi * np.arange(nk[0]) / nk[0]
    ky = 2 * np.pi * np.arange(nk[1]) / nk[1]
    cx, cy = np.cos(kx)[:, None], np.cos(ky)[None, :]
    e = np.zeros((nk[0], nk[1], 1, 2, 2), dtype=np.complex128)
    e[..., 0, 0] = (-1.5 * t * (cx + cy))[..., None]
    e[..., 1, 1] = (-0.5 * t * (cx + cy))[..., None]
    e[..., 0, 1] = e[..., 1, 0] = (hyb * (cx - cy))[..., None]
    return e


def smooth_noise(rng, n, width=3.0):
    """N(0,1) noise convolved with a Gaussian of `width` points, renormalised to unit variance."""
    x = rng.standard_normal(n + 8 * int(width))
    k = np.exp(-0.5 * (np.arange(-4 * int(width), 4 * int(width) + 1) / width) ** 2)
    y = np.convolve(x, k / np.linalg.norm(k), mode="valid")[:n]
    return y / max(y.std(), 1e-12)


def local_sigma(nb, U, n, beta, niv, rng):
    """Sigma_ll(v) = U^2 n(1-n)/(iv) (1 + 0.05 noise), Sigma(-v) = conj Sigma(v)."""
    v = vn(niv, beta)[niv:]
    s = np.zeros((nb, nb, 2 * niv), dtype=np.complex128)
    for l in range(nb):
        pos = U * U * n * (1 - n) / (1j * v) * (1 + 0.05 * smooth_noise(rng, niv))
        s[l, l, niv:] = pos
        s[l, l, :niv] = np.conj(pos[::-1])
    return s


def lattice_g(ek, sigma_loc, mu, beta, niv):
    nb = ek.shape[-1]
    ns = sigma_loc.shape[-1] // 2
    sig = sigma_loc[..., ns - niv: ns + niv]
    iv = 1j * vn(niv, beta)
    m = (iv[None, None, :] + mu) * np.eye(nb)[..., None]
    mat = m[None, None, None] - ek[..., None] - sig[None, None, None]
    return np.linalg.inv(mat.transpose(0, 1, 2, 5, 3, 4)).transpose(0, 1, 2, 4, 5, 3)


def filling(ek, sigma_loc, mu, beta, niv=2048):
    """Spin-summed filling per orbital from the lattice G with the 1/(iv) tail removed."""
    nb = ek.shape[-1]
    ns = sigma_loc.shape[-1] // 2
    v = vn(ns, beta)[ns:]
    pos = sigma_loc[..., ns:]
    # extend Sigma beyond its box with its own 1/(iv) tail
    nv_ext = max(niv, ns)
    vv = np.pi / beta * (2 * np.arange(nv_ext) + 1)
    sig = np.zeros((nb, nb, nv_ext), dtype=np.complex128)
    sig[..., :ns] = pos
    tail = pos[..., -1] * v[-1]
    sig[..., ns:] = tail[..., None] / vv[ns:]
    iv = 1j * vv
    m = (iv[None, None, :] + mu) * np.eye(nb)[..., None]
    ekf = ek.reshape(-1, nb, nb)
    g = np.linalg.inv((m[None] - ekf[..., None] - sig[None]).transpose(0, 3, 1, 2))  # [k, v, a, b]
    gloc = g.mean(axis=0)
    occ = 0.5 + (2.0 / beta) * np.sum(np.real(np.einsum("vaa->va", gloc) - 1.0 / iv[:, None]), axis=0)
    return 2.0 * occ.mean()


def find_mu(ek, sigma_loc, beta, n_per_orbital):
    """mu such that the spin-summed filling per orbital equals n_per_orbital (Brent's method)."""
    from scipy.optimize import brentq

    return brentq(lambda mu: filling(ek, sigma_loc, mu, beta, niv=1024) - n_per_orbital, -30.0, 30.0, xtol=1e-12)


def local_bubble(g_loc, niw, niv, beta):
    """chi0_loc[a,b,c,d,w,v] = -beta G_ad(v) G_cb(v-w), w in [0, niw] (bubble_gen.py:21-30)."""
    ng = g_loc.shape[-1] // 2
    vi = np.arange(-niv, niv)
    w = np.arange(niw + 1)
    gl = g_loc[:, None, None, :, ng + vi]  # [a,1,1,d,v]
    gr = g_loc.swapaxes(0, 1)[None, :, :, None][..., ng + vi[None, :] - w[:, None]]  # [1,b,c,1,w,v]
    return -beta * gl[..., None, :] * gr


def _compound(x):
    nb, nw, nv = x.shape[0], x.shape[4], x.shape[5]
    return x.transpose(4, 0, 1, 5, 3, 2, 6).reshape(nw, nb * nb * nv, nb * nb * nv)


def _uncompound(m, nb, nv):
    nw = m.shape[0]
    return m.reshape(nw, nb, nb, nv, nb, nb, nv).transpose(1, 2, 5, 4, 0, 3, 6)


def _inverse_products(c0, chi, beta, want_f, device):
    """beta^2 (chi^-1 - c0^-1) or beta^2 (c0^-1 - c0^-1 chi c0^-1); with a device, the per-w inverses
    run in the library's blocked LU (those of large boxes take minutes in NumPy)."""
    if device is None:
        c0inv = np.linalg.inv(c0)
        return beta**2 * (c0inv - c0inv @ chi @ c0inv) if want_f else beta**2 * (np.linalg.inv(chi) - c0inv)
    import torch

    from dgamore_amd import kernels as K

    c0d, chid = torch.as_tensor(c0, device=device), torch.as_tensor(chi, device=device)
    c0inv = K.invert_dense(c0d)  # the library's blocked LU
    out = c0inv - c0inv @ chid @ c0inv if want_f else K.invert_dense(chid) - c0inv
    return beta**2 * out.cpu().numpy()


def local_vertices(g_loc, niw, niv, niv_full, beta, rng, rel=0.1, inter=0.05, device=None):
    """(Gamma_dens, Gamma_magn, F_magn) from chi_r = chi0_loc + Hermitian perturbation."""
    nb = g_loc.shape[0]
    out = {}
    for ch in ("dens", "magn", "f_magn"):
        nvb = niv if ch != "f_magn" else niv_full
        chi0 = local_bubble(g_loc, niw, nvb, beta)  # [a,b,c,d,w,v]
        diag = np.einsum("abcdwv,vp->abcdwvp", chi0, np.eye(2 * nvb))
        c0 = _compound(diag)
        n = c0.shape[-1]
        if ch == "f_magn":
            rng_ch = np.random.default_rng(rng.integers(1 << 31))
            key = "magn"
        else:
            rng_ch = np.random.default_rng(rng.integers(1 << 31))
            key = ch
        chi = np.empty_like(c0)
        orb = np.repeat(np.arange(nb * nb), 2 * nvb)
        same = (orb[:, None] == orb[None, :])
        for w in range(niw + 1):
            p = rng_ch.standard_normal((n, n)) + 1j * rng_ch.standard_normal((n, n))
            p = 0.5 * (p + p.conj().T)
            p = np.where(same, p, inter / rel * p)
            scale = rel * np.linalg.norm(np.diag(c0[w])) / max(np.linalg.norm(p), 1e-300)
            chi[w] = c0[w] + scale * p
        res = _inverse_products(c0, chi, beta, ch == "f_magn", device)
        out["f_magn" if ch == "f_magn" else key] = _uncompound(res, nb, 2 * nvb)
    return out


CONFIGS = {
    # 0: 1 band, 16x16, beta=10, niv=20, niw=20 (small enough for the CPU restatement)
    0: dict(name="1band_16x16_b10_niv20", nb=1, nk=(16, 16, 1), beta=10.0, niw=20, niv=20, U=8.0, n=0.85,
            disp=("1band", 1.0, -0.2, 0.1)),
    # 1: 1 band, 64x64, beta=25, niv=60, niw=60
    1: dict(name="1band_64x64_b25_niv60", nb=1, nk=(64, 64, 1), beta=25.0, niw=60, niv=60, U=8.0, n=0.85,
            disp=("1band", 1.0, -0.2, 0.1)),
    # 2: 3 orbitals (t2g-like), 24x24, beta=15, niv=40, niw=30, Kanamori U=4, J=0.6 (the benchmark workload)
    2: dict(name="3orb_24x24_b15_niv40", nb=3, nk=(24, 24, 1), beta=15.0, niw=30, niv=40, U=4.0, J=0.6, n=0.85,
            disp=("t2g",)),
    # 4: 2 orbitals (e_g-like), 48x48, beta=20, U=5, J=0.8 (the Eliashberg workload)
    4: dict(name="2orb_48x48_b20", nb=2, nk=(48, 48, 1), beta=20.0, niw=64, niv=64, U=5.0, J=0.8, n=0.85,
            disp=("eg",)),
}


def batch_configs(count=16, seed=3, nb_choices=(1, 2, 3), betas=(5.0, 10.0, 20.0, 40.0), grids=(16, 24, 32)):
    """Parameters of a batch of independent models with uneven sizes: orbital count, beta
    (niv = niw = round(1.5 beta)), k-grid, U in [2, 8] (J = 0.15 U) and filling in [0.8, 1.0],
    drawn from a seed. Compound sizes n = nb^2 * 2 niv range over 16..1080 for the defaults."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nb = int(rng.choice(nb_choices))
        beta = float(rng.choice(betas))
        nx = int(rng.choice(grids))
        U = float(rng.uniform(2.0, 8.0))
        fill = float(rng.uniform(0.8, 1.0))
        niv = int(round(1.5 * beta))
        disp = {1: ("1band", 1.0, -0.2, 0.1), 2: ("eg",), 3: ("t2g",)}[nb]
        out.append(dict(name=f"batch{i}_{nb}orb_{nx}x{nx}_b{beta:g}_niv{niv}", nb=nb, nk=(nx, nx, 1), beta=beta,
                        niw=niv, niv=niv, U=U, J=0.15 * U, n=fill, disp=disp, seed=int(rng.integers(1 << 30))))
    return out


def orbital_average(x, us, nidx):
    """Average of a local object over orbital transformations: nidx = 2 (X -> U X U^dagger) or
    4 (the vertex rule U, U*, U, U* on its first four axes)."""
    out = np.zeros_like(x)
    for u in us:
        if nidx == 2:
            out += np.einsum("ap,bq,pq...->ab...", u, u.conj(), x)
        else:
            out += np.einsum("ap,bq,cr,ds,pqrs...->abcd...", u, u.conj(), u, u.conj(), x, optimize=True)
    return out / len(us)


def make_model(cfg, seed=None, nk=None, niw=None, niv=None, niv_shell=0, with_vertices=True, device=None,
               symmetric=False):
    """Build the synthetic model of one of the workloads in CONFIGS, or of a parameter dict from
    batch_configs (optionally with smaller boxes/grids for tests).
    device: torch device for the per-frequency inverses of the local vertices (None: NumPy).
    symmetric: average the local self-energy and vertices over the orbital representation of the
    lattice's point group (the signed orbital permutations under which e_k is invariant), so
    that the whole model has the symmetry the irreducible-zone reduction assumes; the per-k
    orbital transforms of the reduction are then stored in model.orbital_transforms."""
    c = dict(CONFIGS[cfg]) if isinstance(cfg, int) else dict(cfg)
    seed = (1000 + cfg if isinstance(cfg, int) else c.get("seed", 1000)) if seed is None else seed
    rng = np.random.default_rng(seed)
    nk = tuple(nk or c["nk"])
    niw = c["niw"] if niw is None else niw
    niv = c["niv"] if niv is None else niv
    nb, beta = c["nb"], c["beta"]
    d = c["disp"]
    if d[0] == "1band":
        ek = dispersion_1band(nk, *d[1:])
    elif d[0] == "t2g":
        ek = dispersion_t2g(nk)
    else:
        ek = dispersion_eg(nk)
    u = kanamori(nb, c["U"], c.get("J", 0.0)) if nb > 1 else np.full((1, 1, 1, 1), c["U"])
    nsig = niw + niv + niv_shell + 10 + 64
    sig = local_sigma(nb, c["U"], c["n"], beta, nsig, rng)
    group_us, transforms = None, None
    if symmetric:
        from dgamore_amd.brillouin_zone import KGrid, orbital_representations, orbital_transforms

        _, reps = orbital_representations(ek, nk)
        group_us = list({u.tobytes(): u for u in reps}.values())
        sig = orbital_average(sig, group_us, 2)
        transforms = orbital_transforms(ek, KGrid(nk))
    mu = find_mu(ek, sig, beta, c["n"])
    m = Model(name=c["name"], nb=nb, nk=nk, beta=beta, niw=niw, niv=niv, niv_shell=niv_shell, mu=mu, ek=ek,
              sigma_loc=sig, u=u.astype(np.complex128), seed=seed)
    if with_vertices:
        g = lattice_g(ek, sig, mu, beta, niw + niv + niv_shell + 10)
        g_loc = g.mean(axis=(0, 1, 2))
        v = local_vertices(g_loc, niw, niv, niv + niv_shell, beta, rng, device=device)
        if group_us is not None:
            v = {key: orbital_average(val, group_us, 4) for key, val in v.items()}
        m.gamma = {"dens": v["dens"], "magn": v["magn"]}
        m.f_magn = v["f_magn"]
    m.orbital_transforms = transforms
    return m
