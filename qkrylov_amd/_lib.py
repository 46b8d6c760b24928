"""ctypes binding of libqkrylov.so (the C ABI in include/qkrylov/c_api.h).

The shared library is built in-tree (``qkrylov_amd/libqkrylov.so``) by
``__graft_entry__.build()`` / ``make -C qkrylov_amd/csrc``.  Loading fails
loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("QKRYLOV_LIB", HERE / "libqkrylov.so"))

SUCCESS = 0
ERROR_INVALID_ARG = -1
ERROR_EXCEPTION = -2

c_void_p = C.c_void_p
c_int = C.c_int
c_float = C.c_float
c_double = C.c_double
c_size_t = C.c_size_t
c_u64 = C.c_uint64
c_i64 = C.c_int64
P = C.POINTER


class LanczosResultC(C.Structure):
    _fields_ = [("energy", c_float), ("iterations", c_int), ("converged", c_int)]


class DavidsonResultC(C.Structure):
    _fields_ = [("iterations", c_int), ("converged", c_int)]


class FTLMResultC(C.Structure):
    _fields_ = [("beta", c_float), ("partition_function", c_float),
                ("internal_energy", c_float), ("specific_heat", c_float)]


class LanczosResultF64(C.Structure):
    _fields_ = [("energy", c_double), ("iterations", c_int), ("converged", c_int),
                ("second_passes", c_int)]


class FTLMResultF64(C.Structure):
    _fields_ = [("beta", c_double), ("partition_function", c_double),
                ("internal_energy", c_double), ("specific_heat", c_double)]


class MlpTrainInfo(C.Structure):
    _fields_ = [("pairs", c_int), ("epochs", c_int), ("loss_first", c_double),
                ("loss_last", c_double), ("solve_rel_res", c_double)]


# name -> (restype, argtypes)
SIGNATURES = {
    # reference ABI (include/qkrylov/c_api.h:30-168 of the reference)
    "qkrylov_sector_create": (c_void_p, []),
    "qkrylov_sector_destroy": (None, [c_void_p]),
    "qkrylov_sector_set_sz": (c_int, [c_void_p, c_int]),
    "qkrylov_sector_set_hubbard_particles": (c_int, [c_void_p, c_int, c_int]),
    "qkrylov_sector_set_n": (c_int, [c_void_p, c_int]),
    "qkrylov_sector_set_nb": (c_int, [c_void_p, c_int]),
    "qkrylov_sector_get_sz": (c_int, [c_void_p, P(c_int), P(c_int)]),
    "qkrylov_sector_get_hubbard_particles": (c_int, [c_void_p, P(c_int), P(c_int), P(c_int)]),
    "qkrylov_sector_get_n": (c_int, [c_void_p, P(c_int), P(c_int)]),
    "qkrylov_sector_get_nb": (c_int, [c_void_p, P(c_int), P(c_int)]),
    "qkrylov_spinhalf_basis_create": (c_void_p, [c_int, c_void_p]),
    "qkrylov_fermion_basis_create": (c_void_p, [c_int, c_void_p]),
    "qkrylov_hubbard_basis_create": (c_void_p, [c_int, c_void_p]),
    "qkrylov_tj_basis_create": (c_void_p, [c_int, c_void_p]),
    "qkrylov_basis_destroy": (None, [c_void_p]),
    "qkrylov_basis_dimension": (c_u64, [c_void_p]),
    "qkrylov_basis_nsites": (c_int, [c_void_p]),
    "qkrylov_basis_state": (c_u64, [c_void_p, c_u64]),
    "qkrylov_basis_index": (c_i64, [c_void_p, c_u64]),
    "qkrylov_basis_contains": (c_int, [c_void_p, c_u64]),
    "qkrylov_spinhalf_site_create": (c_void_p, []),
    "qkrylov_fermion_site_create": (c_void_p, []),
    "qkrylov_hubbard_site_create": (c_void_p, []),
    "qkrylov_tj_site_create": (c_void_p, []),
    "qkrylov_site_destroy": (None, [c_void_p]),
    "qkrylov_opsum_create": (c_void_p, []),
    "qkrylov_opsum_destroy": (None, [c_void_p]),
    "qkrylov_opsum_clear": (c_int, [c_void_p]),
    "qkrylov_opsum_add_term_1body": (c_int, [c_void_p, c_float, c_float, C.c_char_p, c_int]),
    "qkrylov_opsum_add_term_2body": (c_int, [c_void_p, c_float, c_float, C.c_char_p, c_int,
                                             C.c_char_p, c_int]),
    "qkrylov_opsum_add_term_nbody": (c_int, [c_void_p, c_float, c_float, c_int,
                                             P(C.c_char_p), P(c_int)]),
    "qkrylov_hamiltonian_create": (c_void_p, [c_void_p, c_void_p, c_void_p]),
    "qkrylov_hamiltonian_destroy": (None, [c_void_p]),
    "qkrylov_hamiltonian_dimension": (c_u64, [c_void_p]),
    "qkrylov_hamiltonian_apply": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "qkrylov_hamiltonian_apply_complex": (c_int, [c_void_p, c_void_p, c_void_p]),
    "qkrylov_hamiltonian_diagonal": (c_int, [c_void_p, c_void_p]),
    "qkrylov_lanczos_ground_state": (c_int, [c_void_p, c_int, c_float, P(LanczosResultC)]),
    "qkrylov_lanczos_ground_state_complex": (c_int, [c_void_p, c_int, c_float, P(LanczosResultC),
                                                     c_void_p]),
    "qkrylov_davidson_lowest_complex": (c_int, [c_void_p, c_int, c_int, c_float, c_void_p,
                                                c_void_p, P(DavidsonResultC)]),
    "qkrylov_continued_fraction_coeffs_complex": (c_int, [c_void_p, c_void_p, c_int, c_void_p,
                                                          c_void_p, P(c_float), P(c_int)]),
    "qkrylov_evaluate_spectral_function": (c_float, [c_void_p, c_void_p, c_size_t, c_float,
                                                     c_float, c_float, c_float]),
    "qkrylov_ftlm": (c_int, [c_void_p, c_float, c_int, c_int, P(FTLMResultC)]),
    # extensions
    "qkrylov_last_error": (C.c_char_p, []),
    "qkrylov_device_info": (c_int, [C.c_char_p, c_size_t]),
    "qkrylov_device_count": (c_int, []),
    "qkrylov_sector_reset": (c_int, [c_void_p]),
    "qkrylov_opsum_add_term_f64": (c_int, [c_void_p, c_double, c_double, c_int, P(C.c_char_p),
                                           P(c_int)]),
    "qkrylov_opsum_size": (c_size_t, [c_void_p]),
    "qkrylov_basis_states": (c_int, [c_void_p, c_u64, c_u64, c_void_p]),
    "qkrylov_basis_indices": (c_int, [c_void_p, c_void_p, c_u64, c_void_p]),
    "qkrylov_hamiltonian_apply_f64": (c_int, [c_void_p, c_void_p, c_void_p, c_int]),
    "qkrylov_hamiltonian_apply_device_f64": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_u64,
                                                     c_void_p]),
    "qkrylov_hamiltonian_diagonal_f64": (c_int, [c_void_p, c_void_p]),
    "qkrylov_hamiltonian_info": (c_int, [c_void_p, C.c_char_p, c_size_t]),
    "qkrylov_hamiltonian_autotune": (c_int, [c_void_p, c_int, C.c_char_p, c_size_t]),
    "qkrylov_lanczos_ground_state_f64": (c_int, [c_void_p, c_int, c_double, c_void_p,
                                                 P(LanczosResultF64), c_void_p]),
    "qkrylov_block_dot_noalloc_f64": (c_int, [c_void_p, C.c_uint64, c_int, c_void_p, C.c_uint64, c_int,
                                              C.c_uint64, c_void_p, c_void_p]),
    "qkrylov_lanczos_checkpointed_f64": (c_int, [c_void_p, c_int, c_double, c_void_p, C.c_char_p, c_int,
                                                 c_int, C.c_char_p, P(LanczosResultF64), c_void_p]),
    "qkrylov_eigenpairs_save_f64": (c_int, [C.c_char_p, C.c_uint64, c_int, c_void_p, c_void_p,
                                            C.c_uint64]),
    "qkrylov_eigenpairs_info": (c_int, [C.c_char_p, P(C.c_uint64), P(c_int), P(c_int), P(C.c_uint64)]),
    "qkrylov_eigenpairs_load_f64": (c_int, [C.c_char_p, c_void_p, c_void_p]),
    "qkrylov_hamiltonian_fingerprint": (C.c_uint64, [c_void_p]),
    "qkrylov_block_lanczos_f64": (c_int, [c_void_p, c_int, c_int, c_int, c_double, c_void_p, c_void_p,
                                          P(c_int), P(c_int)]),
    "qkrylov_davidson_lowest_f64": (c_int, [c_void_p, c_int, c_int, c_double, c_void_p, c_void_p,
                                            P(DavidsonResultC)]),
    "qkrylov_continued_fraction_coeffs_f64": (c_int, [c_void_p, c_void_p, c_int, c_void_p,
                                                      c_void_p, P(c_double), P(c_int)]),
    "qkrylov_continued_fraction_block_f64": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p,
                                                     c_void_p, c_void_p, c_void_p]),
    "qkrylov_continued_fraction_block_device_f64": (c_int, [c_void_p, c_void_p, c_int, c_int,
                                                            c_void_p, c_void_p, c_void_p, c_void_p]),
    "qkrylov_evaluate_spectral_function_f64": (c_double, [c_void_p, c_void_p, c_size_t, c_double,
                                                          c_double, c_double, c_double]),
    "qkrylov_ftlm_f64": (c_int, [c_void_p, c_double, c_int, c_int, P(FTLMResultF64)]),
    "qkrylov_ftlm_sweep_f64": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_int,
                                       c_void_p, c_void_p, c_void_p]),
    "qkrylov_sz_q_device_f64": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_u64, c_void_p]),
    "qkrylov_hamiltonian_sync": (c_int, [c_void_p]),
    "qkrylov_block_dot_f64": (c_int, [c_void_p, c_int, c_void_p, c_int, c_u64, c_void_p]),
    "qkrylov_reference_start_f64": (c_int, [c_u64, C.c_uint, c_void_p]),
    "qkrylov_reference_normals_f64": (c_int, [c_u64, C.c_uint, c_void_p]),
    # learned Davidson preconditioner
    "qkrylov_mlp_create": (c_void_p, [C.c_uint]),
    "qkrylov_mlp_destroy": (None, [c_void_p]),
    "qkrylov_mlp_num_params": (c_int, []),
    "qkrylov_mlp_get_params": (c_int, [c_void_p, c_void_p]),
    "qkrylov_mlp_set_params": (c_int, [c_void_p, c_void_p]),
    "qkrylov_mlp_eval": (c_int, [c_void_p, c_void_p, c_u64, c_void_p, c_void_p, c_void_p]),
    "qkrylov_mlp_features": (c_int, [c_void_p, c_double, c_void_p]),
    "qkrylov_operator_create": (c_void_p, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "qkrylov_operator_destroy": (None, [c_void_p]),
    "qkrylov_operator_dimension_in": (c_u64, [c_void_p]),
    "qkrylov_operator_dimension_out": (c_u64, [c_void_p]),
    "qkrylov_operator_apply_f64": (c_int, [c_void_p, c_void_p, c_void_p, c_int]),
    "qkrylov_operator_apply_device_f64": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_u64, c_u64,
                                                  c_void_p]),
    "qkrylov_hermitian_eigh_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p]),
    "qkrylov_mlp_correction_f64": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_void_p,
                                           c_void_p, c_void_p, c_void_p]),
    "qkrylov_mlp_pair_loss_f64": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                          c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "qkrylov_jd_correction_f64": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_int, c_double,
                                          c_void_p, P(c_double)]),
    "qkrylov_mlp_train_f64": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_int,
                                      c_double, c_void_p, P(MlpTrainInfo)]),
    "qkrylov_davidson_lowest_ex_f64": (c_int, [c_void_p, c_int, c_int, c_double, c_int, c_void_p,
                                               c_void_p, c_void_p, P(DavidsonResultC)]),
}

_lib = None


def load() -> C.CDLL:
    """Load libqkrylov.so (RTLD_GLOBAL) and declare every signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing; build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " or `make -C qkrylov_amd/csrc`")
    lib = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    return load().qkrylov_last_error().decode(errors="replace")


def check(code: int, what: str = "") -> None:
    if code == SUCCESS:
        return
    if code == ERROR_INVALID_ARG:
        raise ValueError(f"{what}: invalid argument")
    raise RuntimeError(f"{what}: {last_error()}")
