"""ctypes binding of libinq_amd.so (the C ABI declared in include/inq_amd.h).

torch is imported first so that the library's libamdhip64.so.7 dependency
resolves to the HIP runtime torch already loaded (one runtime per process).
If the shared library is missing this module raises: there is no fallback.
"""
import ctypes
import os

import torch  # noqa: F401  (must precede loading the HIP library)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "lib", "libinq_amd.so"))

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)
_c_flt_p = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double

OK = 0
ERR_ARGUMENT = 1
ERR_HIP = 2
ERR_MEMORY = 3
ERR_UNSUPPORTED = 4

# INQAMD_COMMUTATOR_*: the [V_nl, r] form of the current (inq_amd.h)
COMMUTATORS = {"inq": 0, "corrected": 1}


def commutator_code(name):
    if name not in COMMUTATORS:
        raise ValueError(f"commutator must be one of {sorted(COMMUTATORS)}, not {name!r}")
    return COMMUTATORS[name]


class SystemDesc(ctypes.Structure):
    """inqamd_system_desc."""

    _fields_ = [
        ("n", ctypes.c_int * 3),
        ("lattice", ctypes.c_double * 9),
        ("vloc", _c_dbl_p),
        ("nproj", ctypes.c_int),
        ("max_sphere", ctypes.c_int),
        ("max_nlm", ctypes.c_int),
        ("points", _c_int_p),
        ("positions", _c_flt_p),
        ("matrices", _c_dbl_p),
        ("coeff", _c_dbl_p),
        ("nlm", _c_int_p),
        ("sphere_size", _c_int_p),
        ("max_bands", ctypes.c_int),
    ]


class VectorPotentialSeries(ctypes.Structure):
    """inqamd_vector_potential_series."""

    _fields_ = [
        ("n", ctypes.c_int),
        ("coeff", _c_dbl_p),
        ("omega", ctypes.c_double),
        ("direction", ctypes.c_double * 3),
    ]


class Tuning(ctypes.Structure):
    """inqamd_tuning."""

    _fields_ = [
        ("z_tile_y", ctypes.c_int),
        ("xy_threads", ctypes.c_int),
        ("band_chunk", ctypes.c_int),
        ("projector_bands", ctypes.c_int),
    ]


class TDDesc(ctypes.Structure):
    """inqamd_td_desc."""

    _fields_ = [
        ("nst", ctypes.c_int),
        ("nsteps", ctypes.c_int),
        ("dt", ctypes.c_double),
        ("order", ctypes.c_int),
        ("occ", ctypes.c_double),
        ("ncoef", ctypes.c_int),
        ("omega", ctypes.c_double),
        ("direction", ctypes.c_double * 3),
        ("checkpoint_every", ctypes.c_int),
        ("band_slice", ctypes.c_int),
        ("store_powers", ctypes.c_int),
        ("commutator", ctypes.c_int),
    ]


# (name, restype, argtypes) for every symbol of include/inq_amd.h
SIGNATURES = [
    ("inqamd_hamiltonian_create", _i, [ctypes.POINTER(SystemDesc), _vp, ctypes.POINTER(_vp)]),
    ("inqamd_hamiltonian_destroy", None, [_vp]),
    ("inqamd_set_stream", _i, [_vp, _vp]),
    ("inqamd_set_tuning", _i, [_vp, ctypes.POINTER(Tuning)]),
    ("inqamd_get_tuning", _i, [_vp, ctypes.POINTER(Tuning)]),
    ("inqamd_last_error", ctypes.c_char_p, [_vp]),
    ("inqamd_version", ctypes.c_char_p, []),
    ("inqamd_selftest_mfma", _i, [_vp, _vp]),
    ("inqamd_selftest_mfma_4x4", _i, [_vp, _vp]),
    ("inqamd_stream_probe", _i, [_i, _vp, _vp, ctypes.c_long, _i, _i, _i, _i, _i, _vp,
                                 ctypes.POINTER(ctypes.c_float)]),
    ("inqamd_set_kpoint", _i, [_vp, _c_dbl_p]),
    ("inqamd_set_uniform_vector_potential", _i, [_vp, _c_dbl_p]),
    ("inqamd_set_scalar_potential", _i, [_vp, _vp]),
    ("inqamd_hamiltonian_apply", _i, [_vp, _vp, _vp, _i, _i]),
    ("inqamd_non_local", _i, [_vp, _vp, _vp, _i, _i]),
    ("inqamd_hamiltonian_apply_fourier", _i, [_vp, _vp, _vp, _i, _i]),
    ("inqamd_hamiltonian_profile", _i, [_vp, _vp, _vp, _i, _i, _c_dbl_p, _c_int_p]),
    ("inqamd_transform_to_fourier", _i, [_vp, _vp, _vp, _i, _i]),
    ("inqamd_transform_to_real", _i, [_vp, _vp, _vp, _i, _i, _i]),
    ("inqamd_projector_project", _i, [_vp, _vp, _vp, _i, _i]),
    ("inqamd_projector_apply", _i, [_vp, _vp, _vp, _i, _i]),
    ("inqamd_band_energies", _i, [_vp, _vp, _vp, _vp, _i, _i]),
    ("inqamd_band_energies_noalloc", _i, [_vp, _vp, _vp, _vp, _i, _i]),
    ("inqamd_energy", _i, [_vp, _vp, _i, _i, _vp, _vp, _vp]),
    ("inqamd_laplacian_add", _i, [_vp, _vp, _vp, _i, _i, _d, _c_dbl_p]),
    ("inqamd_preconditioner", _i, [_vp, _vp, _i, _i]),
    ("inqamd_orthogonalize", _i, [_vp, _vp, _i, _i, _i]),
    ("inqamd_steepest_descent", _i, [_vp, _vp, _i, _i, _i]),
    ("inqamd_subspace_diagonalization", _i, [_vp, _vp, _i, _i, _c_dbl_p]),
    ("inqamd_density", _i, [_vp, _vp, _i, _i, _c_dbl_p, _vp]),
    ("inqamd_save_states", _i, [_vp, _vp, _i, _i, ctypes.c_long, _i, ctypes.c_char_p]),
    ("inqamd_load_states", _i, [_vp, _vp, _i, _i, ctypes.c_long, _i, ctypes.c_char_p]),
    ("inqamd_poisson", _i, [_vp, _vp, _vp]),
    ("inqamd_dipole", _i, [_vp, _vp, _c_dbl_p]),
    ("inqamd_lda_xc", _i, [_vp, _vp, _vp, _c_dbl_p, _c_dbl_p]),
    ("inqamd_xc", _i, [_vp, _vp, ctypes.c_int, _vp, _c_dbl_p, _c_dbl_p]),
    ("inqamd_ks_potential", _i, [_vp, _vp, _vp, _i, _i, _c_dbl_p]),
    ("inqamd_etrs_scf_step", _i, [_vp, _d, _c_dbl_p, _c_dbl_p, _i, _vp, _i, _i, _vp, _c_dbl_p,
                                  _i, _i, ctypes.POINTER(ctypes.c_int)]),
    ("inqamd_shift", _i, [_vp, _d, _d, _vp, _vp, _i, _i]),
    ("inqamd_exponential_in_place", _i, [_vp, _d, _d, _i, _vp, _i, _i]),
    ("inqamd_exponential_2_for_1", _i, [_vp, _d, _d, _d, _d, _i, _vp, _vp, _i, _i]),
    ("inqamd_etrs_step", _i, [_vp, _d, _c_dbl_p, _c_dbl_p, _i, _vp, _i, _i]),
    ("inqamd_etrs_graph_create", _i, [_vp, _vp, _i, _i, _d, _i, ctypes.POINTER(VectorPotentialSeries),
                                      ctypes.POINTER(_vp)]),
    ("inqamd_etrs_graph_launch", _i, [_vp, _c_dbl_p, _c_dbl_p]),
    ("inqamd_etrs_graph_set_step", _i, [_vp, ctypes.c_long]),
    ("inqamd_graph_destroy", None, [_vp]),
    ("inqamd_set_position_commutator", _i, [_vp, _i]),
    ("inqamd_current", _i, [_vp, _vp, _i, _i, _d, _c_dbl_p, _c_dbl_p]),
    ("inqamd_td_create", _i, [_vp, ctypes.POINTER(TDDesc), ctypes.POINTER(_vp)]),
    ("inqamd_td_forward", _i, [_vp, _c_dbl_p, _vp, _i, _vp, _c_dbl_p]),
    ("inqamd_td_backward", _i, [_vp, _d, _c_dbl_p, _vp, _vp]),
    ("inqamd_td_memory", ctypes.c_long, [_vp]),
    ("inqamd_td_forward_count", ctypes.c_long, [_vp]),
    ("inqamd_td_destroy", None, [_vp]),
    ("inqamd_batch_create", _i, [_vp, _i, _vp, ctypes.POINTER(_vp)]),
    ("inqamd_batch_apply", _i, [_vp, _vp, _vp, _vp, _vp]),
    ("inqamd_batch_set_tuning", _i, [_vp, ctypes.c_long, _i, _i, _i, _i]),
    ("inqamd_batch_set_stream", _i, [_vp, _vp]),
    ("inqamd_batch_info", _i, [_vp, _c_int_p, _c_int_p, _c_int_p, _c_int_p]),
    ("inqamd_batch_last_error", ctypes.c_char_p, [_vp]),
    ("inqamd_batch_destroy", None, [_vp]),
    ("inqamd_containing_cube", _i, [_c_dbl_p, _c_int_p, _c_dbl_p, _d, _c_int_p, _c_int_p]),
    ("inqamd_sphere_build", _i, [_c_dbl_p, _c_int_p, _c_dbl_p, _d, _i, _vp, _vp, _vp, _c_int_p]),
    ("inqamd_gaussian_projector_matrix", _i, [_c_dbl_p, _i, _vp, _vp, _d, _i, _vp, _i]),
]

_lib = None


def lib():
    """Load libinq_amd.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libinq_amd.so not found at {LIB_PATH}: build it with __graft_entry__.build() "
                "(make -C inq-wannier-fork_amd); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class InqAmdError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"inq_amd error {code}: {message}")
        self.code = code


def check(rc, handle=None):
    if rc != OK:
        msg = lib().inqamd_last_error(handle)
        raise InqAmdError(rc, msg.decode() if msg else "")


def dvec(v):
    return (ctypes.c_double * len(v))(*[float(x) for x in v])


def ivec(v):
    return (ctypes.c_int * len(v))(*[int(x) for x in v])
