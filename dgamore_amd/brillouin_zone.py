"""Momentum grid and its irreducible wedge (host side, index arithmetic only).

Mirrors KGrid of the reference (brillouin_zone.py:308-581): the full-zone
index grid is symmetrised by the lattice operations in order, and
numpy.unique of the result gives the irreducible representatives
(irrk_ind), the map back to the full zone (irrk_inv) and the multiplicities
(irrk_count). The maps are uploaded once and drive the device gathers.
"""

import numpy as np

SYMMETRY_SETS = {
    "two_dimensional_square": ("x_inv", "y_inv", "x_y_sym"),
    "three_dimensional_cubic": ("x_inv", "y_inv", "z_inv", "x_y_sym", "x_z_sym", "y_z_sym"),
    "two_dimensional_nematic": ("x_inv", "y_inv"),
    "quasi_two_dimensional_square": ("x_inv", "y_inv", "z_inv", "x_y_sym"),
    "quasi_one_dimensional_square": ("x_inv", "y_inv"),
    "none": (),
}


def _apply(grid, op):
    if op.endswith("_inv") and len(op) == 5:  # x_inv / y_inv / z_inv: k -> -k along one axis
        ax = "xyz".index(op[0])
        n = grid.shape[ax]
        # representative of k and n-k is the one stored first (k <= n/2)
        idx = np.arange(n)
        mirror = (n - idx) % n
        take = np.where(idx <= n // 2, idx, mirror)
        grid[...] = np.take(grid, take, axis=ax)
    elif op.endswith("_sym"):  # exchange of two equal axes
        a, b = "xyz".index(op[0]), "xyz".index(op[2])
        if grid.shape[a] == grid.shape[b]:
            grid[...] = np.minimum(grid, grid.swapaxes(a, b))
    else:
        raise ValueError(f"unknown symmetry {op}")


class KGrid:
    """nk = (nx, ny, nz); symmetries: name of a symmetry set or a sequence of operation names."""

    def __init__(self, nk, symmetries="two_dimensional_square"):
        self.nk = tuple(int(n) for n in nk)
        ops = SYMMETRY_SETS[symmetries] if isinstance(symmetries, str) else tuple(symmetries)
        grid = np.arange(int(np.prod(self.nk))).reshape(self.nk)
        for op in ops:
            _apply(grid, op)
        self.fbz2irrk = grid
        _, self.irrk_ind, inv, self.irrk_count = np.unique(grid, return_index=True, return_inverse=True,
                                                           return_counts=True)
        self.irrk_inv = inv.ravel()
        self.kmesh_ind = np.array(np.meshgrid(*[np.arange(n) for n in self.nk], indexing="ij"))

    @property
    def nk_tot(self):
        return int(np.prod(self.nk))

    @property
    def nk_irr(self):
        return len(self.irrk_ind)

    def get_q_list(self):
        return np.array([self.kmesh_ind[i].flatten() for i in range(3)]).T

    def get_irrq_list(self):
        return np.array([self.kmesh_ind[i].flatten()[self.irrk_ind] for i in range(3)]).T
