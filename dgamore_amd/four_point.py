"""FourPoint: the reference's non-local four-point container, with its hot methods on the GPU.

Same constructor arguments, attribute names and layouts as the reference
(four_point.py:18-46, local_n_point.py:13-108): ``mat`` is a C-order numpy
array [q, o1, o2, o3, o4, (w), (v), (v')] (compressed q) or with [qx, qy, qz]
leading axes. Storage is complex128 (the reference casts to complex64 at
n_point_base.py:64-74).

GPU-backed methods (one code path, errors raise):
* ``invert`` for num_vn = 1 (per-(q,w,v) blocks) and num_vn = 0 (per-(q,w) blocks)
  (four_point.py:590-618);
* ``invert_and_sum_over_last_vn`` / ``_v2`` (four_point.py:620-684).
Bookkeeping methods (frequency cuts, half/full bosonic range, sums) are plain
array operations with the reference's semantics.
"""

from enum import Enum

import numpy as np

from dgamore_amd import _lib


class SpinChannel(Enum):
    DENS = "dens"
    MAGN = "magn"
    SING = "sing"
    TRIP = "trip"
    UU = "uu"
    UD = "ud"
    UD_BAR = "ud_bar"
    NONE = "none"


class FrequencyNotation(Enum):
    PH = "ph"
    PH_BAR = "ph_bar"
    PP = "pp"


def _torch():
    import torch

    return torch


class FourPoint:
    def __init__(self, mat, channel=SpinChannel.NONE, nq=(1, 1, 1), num_wn_dimensions=1, num_vn_dimensions=2,
                 full_niw_range=True, full_niv_range=True, has_compressed_q_dimension=False,
                 frequency_notation=FrequencyNotation.PH):
        if num_vn_dimensions not in (0, 1, 2):
            raise ValueError("0 - 2 fermionic frequency dimensions are supported.")
        if num_wn_dimensions not in (0, 1):
            raise ValueError("0 or 1 bosonic frequency dimensions are supported.")
        self.mat = np.asarray(mat, dtype=np.complex128)
        self.channel = channel
        self._nq = tuple(nq)
        self._num_wn_dimensions = num_wn_dimensions
        self._num_vn_dimensions = num_vn_dimensions
        self._full_niw_range = full_niw_range
        self._full_niv_range = full_niv_range
        self._has_compressed_q_dimension = has_compressed_q_dimension
        self.frequency_notation = frequency_notation

    # -- properties (local_n_point.py:46-108, n_point_base.py:306-352) ---------------------------
    @property
    def nq(self):
        return self._nq

    @property
    def has_compressed_q_dimension(self):
        return self._has_compressed_q_dimension

    @property
    def num_wn_dimensions(self):
        return self._num_wn_dimensions

    @property
    def num_vn_dimensions(self):
        return self._num_vn_dimensions

    @property
    def full_niw_range(self):
        return self._full_niw_range

    @property
    def full_niv_range(self):
        return self._full_niv_range

    @property
    def n_bands(self):
        return self.mat.shape[1] if self._has_compressed_q_dimension else self.mat.shape[3]

    @property
    def nq_tot(self):
        return self.mat.shape[0] if self._has_compressed_q_dimension else int(np.prod(self._nq))

    @property
    def niw(self):
        if self._num_wn_dimensions == 0:
            return 0
        return self.mat.shape[-(1 + self._num_vn_dimensions)] // 2

    @property
    def niv(self):
        return 0 if self._num_vn_dimensions == 0 else self.mat.shape[-1] // 2

    @property
    def current_shape(self):
        return self.mat.shape

    # -- bookkeeping --------------------------------------------------------------------------
    def compress_q_dimension(self):
        if not self._has_compressed_q_dimension:
            self.mat = self.mat.reshape((int(np.prod(self._nq)),) + self.mat.shape[3:])
            self._has_compressed_q_dimension = True
        return self

    def decompress_q_dimension(self):
        if self._has_compressed_q_dimension:
            self.mat = self.mat.reshape(tuple(self._nq) + self.mat.shape[1:])
            self._has_compressed_q_dimension = False
        return self

    def to_half_niw_range(self):
        """w >= 0 part (local_n_point.py:227-240)."""
        if self._num_wn_dimensions == 0 or not self._full_niw_range:
            return self
        ax = self.mat.ndim - 1 - self._num_vn_dimensions
        n = self.mat.shape[ax]
        self.mat = np.take(self.mat, np.arange(n // 2, n), axis=ax)
        self._full_niw_range = False
        return self

    def to_full_niw_range(self):
        """F(-w) = conj F(w) with all frequency axes flipped (local_n_point.py:191-225)."""
        if self._num_wn_dimensions == 0 or self._full_niw_range:
            return self
        ax = self.mat.ndim - 1 - self._num_vn_dimensions
        axes = tuple(range(ax, self.mat.ndim))
        neg = np.conj(np.flip(np.take(self.mat, np.arange(1, self.mat.shape[ax]), axis=ax), axis=axes))
        self.mat = np.concatenate([neg, self.mat], axis=ax)
        self._full_niw_range = True
        return self

    def cut_niv(self, niv_cut):
        """Copy with |v| < niv_cut (local_n_point.py:134-154)."""
        if self._num_vn_dimensions == 0:
            raise ValueError("Cannot cut fermionic frequencies if there are none.")
        if niv_cut > self.niv:
            return self
        s = slice(self.niv - niv_cut, self.niv + niv_cut) if self._full_niv_range else slice(0, niv_cut)
        mat = self.mat[..., s, s] if self._num_vn_dimensions == 2 else self.mat[..., s]
        return self._like(np.ascontiguousarray(mat))

    def sum_over_vn(self, beta, axis=(-1,)):
        """(1/beta^len(axis)) sum over fermionic axes (four_point.py:114-144)."""
        if len(axis) > self._num_vn_dimensions:
            raise ValueError(f"Cannot sum over more fermionic axes than available in {self.current_shape}.")
        out = self._like(np.sum(self.mat, axis=axis) / beta ** len(axis))
        out._num_vn_dimensions -= len(axis)
        return out

    def sum_over_all_vn(self, beta):
        if self._num_vn_dimensions == 0:
            return self
        return self.sum_over_vn(beta, axis=(-1,) if self._num_vn_dimensions == 1 else (-2, -1))

    def _like(self, mat):
        return FourPoint(mat, self.channel, self._nq, self._num_wn_dimensions, self._num_vn_dimensions,
                         self._full_niw_range, self._full_niv_range, self._has_compressed_q_dimension,
                         self.frequency_notation)

    def __mul__(self, other):
        if not isinstance(other, (int, float, complex)):
            raise ValueError("Multiplication only supported with numbers or numpy arrays.")
        return self._like(self.mat * other)

    __rmul__ = __mul__

    def __truediv__(self, other):
        return self.__mul__(1.0 / other)

    def __neg__(self):
        return self.__mul__(-1.0)

    # -- sums of objects with the same momentum and frequency structure (four_point.py:49-71, 332-440);
    # the reference also broadcasts a num_vn=1 operand against a num_vn=2 one by its diagonal, which no
    # call of the hot path needs: mismatched frequency structures raise ValueError here.
    def _operand(self, other):
        if isinstance(other, (int, float, complex, np.ndarray)):
            return other
        if not isinstance(other, FourPoint):
            raise ValueError(f"Operations are only supported with FourPoint objects, numbers or arrays, "
                             f"not {type(other)}.")
        if other._num_vn_dimensions != self._num_vn_dimensions or other._num_wn_dimensions != self._num_wn_dimensions:
            raise ValueError("FourPoint objects with different frequency dimensions cannot be added here.")
        if other._full_niw_range != self._full_niw_range:
            other = other._like(other.mat.copy())
            other.to_full_niw_range() if self._full_niw_range else other.to_half_niw_range()
        if other._has_compressed_q_dimension != self._has_compressed_q_dimension:
            other = other._like(other.mat)
            other.compress_q_dimension() if self._has_compressed_q_dimension else other.decompress_q_dimension()
        if other.mat.shape != self.mat.shape:
            raise ValueError(f"Shapes {self.mat.shape} and {other.mat.shape} do not match.")
        return other.mat

    def add(self, other):
        return self._like(self.mat + self._operand(other))

    def sub(self, other):
        return self._like(self.mat - self._operand(other))

    __add__ = add
    __radd__ = add
    __sub__ = sub

    def __rsub__(self, other):
        return self._like(self._operand(other) - self.mat)

    def __getitem__(self, item):
        return self.mat[item]

    def __setitem__(self, key, value):
        self.mat[key] = value

    def permute_orbitals(self, permutation="abcd->abcd", copy=True):
        """Orbital permutation given as 'abcd->...' (four_point.py:303-327)."""
        split = permutation.split("->")
        if len(split) != 2 or len(split[0]) != 4 or len(split[1]) != 4:
            raise ValueError("Invalid permutation.")
        if split[0] == split[1]:
            return self
        lead = "i" if self._has_compressed_q_dimension else "ijk"
        mat = np.einsum(f"{lead}{split[0]}...->{lead}{split[1]}...", self.mat)
        if not copy:
            self.mat = mat
            return self
        return self._like(np.ascontiguousarray(mat))

    def save(self, output_dir="./", name="please_give_me_a_name"):
        """np.save of the half bosonic range as {output_dir}/{name}.npy (local_n_point.py:296-303)."""
        import os

        full = self._full_niw_range
        np.save(os.path.join(output_dir, f"{name}.npy"), self.to_half_niw_range().mat, allow_pickle=False)
        if full:
            self.to_full_niw_range()

    def free(self, trim=False):
        """Drop the array (n_point_base.py:174-185)."""
        self.mat = None

    # -- GPU-backed methods -------------------------------------------------------------------
    def invert(self, copy=True):
        """Inverse in compound indices, returned in the half bosonic range (four_point.py:590-618).

        num_vn = 1: inverse of every [(o1,o2),(o4,o3)] block per (q, w, v);
        num_vn = 0: per (q, w). Dense num_vn = 2 inversion is not on the hot path and raises.
        """
        from dgamore_amd import kernels as K

        torch = _torch()
        obj = self._like(self.mat.copy()) if copy else self
        obj.to_half_niw_range().compress_q_dimension()
        nb = obj.n_bands
        if obj._num_vn_dimensions == 1:
            d = K.ref4_to_qf(K.to_dev(obj.mat))
            niv = obj.niv
            inv, _, _ = K.chi0_prepare(d, nb, niv, 1.0)
            obj.mat = K.qf_to_ref4(inv, nb).cpu().numpy()
            return obj
        if obj._num_vn_dimensions == 0:
            d = K.ref4w_to_qw(K.to_dev(obj.mat))  # [w][nb^4][q]
            nw, _, nq = d.shape
            flat = d.permute(1, 0, 2).reshape(nb**4, nw * nq).contiguous()
            out = torch.empty_like(flat)
            info = torch.zeros(nw * nq, dtype=torch.int32, device=flat.device)
            _lib.call("dgami_block_invert", nw * nq, nb, K._p(flat), K._p(out), K._p(info, torch.int32), K._stream())
            if int(torch.count_nonzero(info)):
                raise np.linalg.LinAlgError("Singular matrix")
            obj.mat = K.qw_to_ref4w(out.reshape(nb**4, nw, nq).permute(1, 0, 2).contiguous(), nb).cpu().numpy()
            return obj
        raise NotImplementedError("dense compound inversion (num_vn = 2) is not part of the GPU hot path")

    def invert_and_sum_over_last_vn(self, beta):
        """(1/beta) sum_v' of the compound inverse, via LU solves with nb^2 indicator RHS (four_point.py:620-642)."""
        if self._num_vn_dimensions != 2:
            raise NotImplementedError("Method only implemented for objects with two fermionic frequency dimensions.")
        from dgamore_amd import config
        from dgamore_amd import kernels as K

        torch = _torch()
        self.to_half_niw_range().compress_q_dimension()
        nq, nb, nw, nv = self.mat.shape[0], self.n_bands, self.mat.shape[5], self.mat.shape[6]
        n = nb * nb * nv
        a = self.mat.transpose(0, 5, 1, 2, 6, 4, 3, 7).reshape(nq * nw, n, n)  # [q,w,(o1,o2,v),(o4,o3,v')]
        d = K.to_dev(a)
        out = torch.empty((nq * nw, n, nb * nb), dtype=K.C128, device=d.device)
        info = torch.zeros(nq * nw, dtype=torch.int32, device=d.device)
        _lib.call("dgami_solve_sum_dense", nq * nw, nb, nv // 2, float(beta), K._p(d), int(config.precision.ladder),
                  K._p(out), K._p(info, torch.int32), K._stream())
        if int(torch.count_nonzero(info)):
            bad = torch.nonzero(info).cpu().numpy().ravel()[:8].tolist()
            raise np.linalg.LinAlgError(f"Singular matrix (q, w) flat indices {bad}")
        x = out.cpu().numpy().reshape(nq, nw, nb, nb, nv, nb, nb)  # [q,w,o1,o2,v,o4,o3]
        self.mat = np.ascontiguousarray(x.transpose(0, 2, 3, 6, 5, 1, 4))
        self._num_vn_dimensions = 1
        return self

    def invert_and_sum_over_last_vn_v2(self, beta):
        """Memory-lean variant of the reference (four_point.py:644-684); identical here."""
        return self.invert_and_sum_over_last_vn(beta)

    @staticmethod
    def load(filename, channel=SpinChannel.NONE, nq=(1, 1, 1), num_wn_dimensions=1, num_vn_dimensions=2,
             full_niw_range=False, full_niv_range=True, has_compressed_q_dimension=True,
             frequency_notation=FrequencyNotation.PH):
        return FourPoint(np.load(filename, allow_pickle=False), channel, nq, num_wn_dimensions, num_vn_dimensions,
                         full_niw_range, full_niv_range, has_compressed_q_dimension, frequency_notation)


class LocalFourPoint(FourPoint):
    """Local four-point object [o1, o2, o3, o4, (w), (v), (v')] (local_four_point.py; the constructor
    arguments of LocalFourPoint.load, local_four_point.py:671-691)."""

    def __init__(self, mat, channel=SpinChannel.NONE, num_wn_dimensions=1, num_vn_dimensions=2,
                 full_niw_range=True, full_niv_range=True, frequency_notation=FrequencyNotation.PH):
        super().__init__(mat, channel, (1, 1, 1), num_wn_dimensions, num_vn_dimensions, full_niw_range,
                         full_niv_range, False, frequency_notation)

    @property
    def n_bands(self):
        return self.mat.shape[0]

    @property
    def nq_tot(self):
        return 1

    def _like(self, mat):
        return LocalFourPoint(mat, self.channel, self._num_wn_dimensions, self._num_vn_dimensions,
                              self._full_niw_range, self._full_niv_range, self.frequency_notation)

    def compress_q_dimension(self):
        return self

    def decompress_q_dimension(self):
        return self

    def permute_orbitals(self, permutation="abcd->abcd", copy=True):
        split = permutation.split("->")
        if len(split) != 2 or len(split[0]) != 4 or len(split[1]) != 4:
            raise ValueError("Invalid permutation.")
        if split[0] == split[1]:
            return self
        mat = np.einsum(f"{split[0]}...->{split[1]}...", self.mat)
        if not copy:
            self.mat = mat
            return self
        return self._like(np.ascontiguousarray(mat))

    @staticmethod
    def load(filename, channel=SpinChannel.NONE, num_wn_dimensions=1, num_vn_dimensions=2, full_niw_range=False,
             full_niv_range=True, frequency_notation=FrequencyNotation.PH):
        return LocalFourPoint(np.load(filename, allow_pickle=False), channel, num_wn_dimensions, num_vn_dimensions,
                              full_niw_range, full_niv_range, frequency_notation)
