"""Module-level configuration singletons used by the hot-path functions.

The reference keeps its run parameters in module-level singletons
(config.py:219-230) that the hot-path functions read (config.sys.beta,
config.box.niv_core, config.lattice.q_grid, ...). This module provides the
subset those functions read, with the same attribute names.
"""

from types import SimpleNamespace

sys = SimpleNamespace(beta=1.0, mu=0.0, n=0.0, n_bands=1, occ=None, occ_k=None, occ_dmft=None)
box = SimpleNamespace(niw_core=0, niv_core=0, niv_shell=0, niv_full=0, niv_dmft=0)
lattice = SimpleNamespace(nk=(1, 1, 1), nq=(1, 1, 1), k_grid=None, q_grid=None)
self_consistency = SimpleNamespace(mixing=1.0, max_iter=1, epsilon=1e-4, mixing_strategy="linear",
                                   mixing_history_length=2, previous_sc_path="", restrict_chi_phys=False)
output = SimpleNamespace(output_path="", eliashberg_path="")
lambda_correction = SimpleNamespace(perform_lambda_correction=False, type="spch")
eliashberg = SimpleNamespace(n_eig=4, epsilon=1e-12, symmetry="random", save_fq=False, construct_fq_cheap=False,
                             matvec_precision="fp64")
memory = SimpleNamespace(save_memory_for_chiq_aux=False, save_memory_for_chi0q=False, save_memory_for_sde=False)
dmft = SimpleNamespace(type="w2dyn", input_path="", fname_1p="1p-data.hdf5", fname_2p="g4iw_sym.hdf5")
precision = SimpleNamespace(ladder=0)  # 0 plain FP64, 1 FP32 + FP64 refinement, 2 split-BF16 + FP64 refinement
