"""DMFT input from w2dynamics output files (dmft_interface.py:125-330).

W2dynInterface keeps the reference's method names, arguments and return shapes; it reads the
files with the package's own HDF5 parser (hdf5_read), so that h5py is not a dependency.
The reference wraps G and Sigma in GreensFunction / SelfEnergy objects; here they are the
arrays those objects hold: get_giw -> [nb, nb, 2 niv_dmft] and get_siw -> [1, 1, 1, nb, nb,
2 niv_dmft] (double counting added), both diagonal in the orbitals, spin-averaged.
"""
import os

import numpy as np

from dgamore_amd import config
from dgamore_amd import hdf5_read


class DMFTInterface:
    """Abstract interface for DMFT output (dmft_interface.py:20-122)."""

    def get_beta(self):
        raise NotImplementedError()

    def get_mu(self):
        raise NotImplementedError()

    def get_nd(self, ineq=1):
        raise NotImplementedError()

    def get_totdens(self):
        raise NotImplementedError()

    def get_occ(self, ineq=1):
        raise NotImplementedError()

    def get_udd(self, ineq=1):
        raise NotImplementedError()

    def get_jdd(self, ineq=1):
        raise NotImplementedError()

    def get_vdd(self, ineq=1):
        raise NotImplementedError()

    def get_dc(self, ineq=1):
        raise NotImplementedError()

    def get_giw(self, ineq=1):
        raise NotImplementedError()

    def get_siw(self, ineq=1):
        raise NotImplementedError()


class W2dynInterface(DMFTInterface):
    """w2dynamics output (dmft_interface.py:125-315). The file names default to
    config.dmft.input_path / fname_1p (and fname_2p, opened only when present: the two-particle
    Green's function is read by get_g2iw only)."""

    def __init__(self, file_1p=None, file_2p=None):
        self.path_1p = file_1p or os.path.join(config.dmft.input_path, config.dmft.fname_1p)
        self.path_2p = file_2p or os.path.join(config.dmft.input_path, config.dmft.fname_2p)
        self.file_1p = None
        self.file_2p = None
        self._open()

    def get_beta(self):
        """dmft_interface.py:137-141."""
        return float(self.file_1p[".config"].attrs["general.beta"])

    def get_mu(self, dmft_iter="dmft-last"):
        """dmft_interface.py:143-148."""
        return float(self.file_1p[dmft_iter + "/mu/value"][()])

    def get_nd(self, ineq=1):
        """dmft_interface.py:150-155."""
        return int(self._from_ineq_config("nd", ineq=ineq))

    def get_totdens(self, dmft_iter="dmft-last"):
        """dmft_interface.py:157-162."""
        return float(self.file_1p[".config"].attrs["general.totdens"])

    def get_occ(self, ineq=1, dmft_iter="dmft-last"):
        """Orbital-resolved occupation 2 mean_spin rho1 (dmft_interface.py:164-171)."""
        rho1 = self.file_1p[self._ineq_group(ineq, dmft_iter) + "/rho1/value"][()]
        return 2 * np.mean(rho1, axis=(1, 3))

    def get_udd(self, ineq=1):
        """dmft_interface.py:173-179."""
        return float(self._from_ineq_config("udd", ineq=ineq))

    def get_jdd(self, ineq=1):
        """dmft_interface.py:181-187."""
        return float(self._from_ineq_config("jdd", ineq=ineq))

    def get_vdd(self, ineq=1):
        """dmft_interface.py:189-196."""
        return float(self._from_ineq_config("vdd", ineq=ineq))

    def get_dc(self, ineq=1, dmft_iter="dmft-last"):
        """dmft_interface.py:198-204."""
        return self.file_1p[self._ineq_group(ineq, dmft_iter) + "/dc/value"][()]

    def get_giw(self, ineq=1, dmft_iter="dmft-last"):
        """G(iv) [nb, nb, 2 niv_dmft], spin-averaged, on the orbital diagonal (dmft_interface.py:206-216)."""
        giw = self.file_1p[self._ineq_group(ineq, dmft_iter) + "/giw/value"][()]  # [band, spin, niv]
        return self._extend_orbital(np.mean(giw, axis=1))

    def get_siw(self, ineq=1, dmft_iter="dmft-last"):
        """Sigma(iv) + double counting [1, 1, 1, nb, nb, 2 niv_dmft] (dmft_interface.py:218-232)."""
        siw = self.file_1p[self._ineq_group(ineq, dmft_iter) + "/siw/value"][()]  # [band, spin, niv]
        siw = self._extend_orbital(np.mean(siw, axis=1))[None, None, None, ...]
        siw_dc = np.mean(self.get_dc(ineq, dmft_iter), axis=-1)  # [band, spin] -> spin mean
        return siw + self._extend_orbital(siw_dc)[None, None, None, ..., None]

    def _extend_orbital(self, obj):
        """The first axis onto the orbital diagonal (dmft_interface.py:263-268)."""
        return np.einsum("i...,ij->ij...", obj, np.eye(obj.shape[0]))

    def _ineq_group(self, ineq=1, dmft_iter="dmft-last"):
        return dmft_iter + f"/ineq-{ineq:03}"

    def _from_ineq_config(self, key, ineq=1):
        return self.file_1p[".config"].attrs[f"atoms.{ineq:1}.{key}"]

    def _open(self):
        self.file_1p = hdf5_read.File(self.path_1p)
        self.file_2p = hdf5_read.File(self.path_2p) if os.path.exists(self.path_2p) else None

    def _close(self):
        self.file_1p = None
        self.file_2p = None

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc_val, exc_tb):
        self._close()
