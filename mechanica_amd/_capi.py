"""ctypes declarations of the C ABI in include/mxg.h.

Loading fails loudly when the built library is missing: there is no CPU
fallback for any entry point.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# MXG_LIBRARY names another build of the same library (for timing a
# variant side by side); by default the in-tree build is loaded.
LIB_PATH = os.environ.get("MXG_LIBRARY") or os.path.join(_HERE, "libmxg.so")

c_double_p = C.POINTER(C.c_double)
c_float_p = C.POINTER(C.c_float)
c_int_p = C.POINTER(C.c_int)
c_u8_p = C.POINTER(C.c_uint8)
c_u16_p = C.POINTER(C.c_uint16)
c_u64_p = C.POINTER(C.c_uint64)


class PotentialInfo(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("flags", C.c_uint32),
        ("n", C.c_int32),
        ("alpha", C.c_float * 4),
        ("a", C.c_float),
        ("b", C.c_float),
        ("r0_plusone", C.c_float),
        ("dpd_alpha", C.c_float),
        ("dpd_gamma", C.c_float),
        ("dpd_sigma", C.c_float),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "mxg_potential_create_lj126": (C.c_void_p, [C.c_double] * 5),
    "mxg_potential_create_morse": (C.c_void_p, [C.c_double] * 6),
    "mxg_potential_create_harmonic": (C.c_void_p, [C.c_double] * 5),
    "mxg_potential_create_harmonic_angle": (C.c_void_p, [C.c_double] * 5),
    "mxg_potential_create_harmonic_dihedral": (C.c_void_p, [C.c_double, C.c_int, C.c_double, C.c_double]),
    "mxg_potential_eval_r": (C.c_int, [C.c_void_p, C.c_float, c_float_p, c_float_p]),
    "mxg_potential_create_dpd": (C.c_void_p, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_int]),
    "mxg_potential_free": (None, [C.c_void_p]),
    "mxg_potential_error": (C.c_int, []),
    "mxg_potential_get_info": (C.c_int, [C.c_void_p, C.POINTER(PotentialInfo)]),
    "mxg_potential_coefficients": (C.c_int, [C.c_void_p, c_float_p, C.c_int]),
    "mxg_potential_eval": (C.c_int, [C.c_void_p, C.c_float, c_float_p, c_float_p]),
    "mxg_potential_eval_ex": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float, c_float_p, c_float_p]),
    "mxg_engine_init": (C.c_int, [C.POINTER(C.c_void_p), c_double_p, c_double_p, c_int_p, C.c_double, C.c_double]),
    "mxg_ensemble_init": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, c_double_p, c_double_p, c_int_p,
                                    C.c_double, C.c_double]),
    "mxg_engine_free": (None, [C.c_void_p]),
    "mxg_engine_addtype": (C.c_int, [C.c_void_p, C.c_double, C.c_float, C.c_int]),
    "mxg_engine_addpot": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int]),
    "mxg_engine_add_random_force": (C.c_int, [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_float]),
    "mxg_engine_add_force_berendsen": (C.c_int, [C.c_void_p, C.c_int, C.c_float]),
    "mxg_engine_add_force_constant": (C.c_int, [C.c_void_p, C.c_int, c_float_p]),
    "mxg_engine_add_force_friction": (C.c_int, [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float]),
    "mxg_engine_set_target_energy": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "mxg_engine_set_seed": (C.c_int, [C.c_void_p, C.c_uint64]),
    "mxg_engine_set_dpd_noise": (C.c_int, [C.c_void_p, C.c_int]),
    "mxg_engine_set_max_dist_fraction": (C.c_int, [C.c_void_p, C.c_float]),
    "mxg_engine_set_virial": (C.c_int, [C.c_void_p, C.c_int]),
    "mxg_engine_set_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "mxg_engine_set_integrator": (C.c_int, [C.c_void_p, C.c_int]),
    "mxg_engine_set_brick": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "mxg_engine_get_brick": (C.c_int, [C.c_void_p, c_int_p, c_int_p, c_int_p]),
    "mxg_engine_autotune": (C.c_int, [C.c_void_p, C.c_int, c_int_p]),
    "mxg_engine_tune_log": (C.c_int, [C.c_void_p, C.c_int, c_int_p, c_double_p]),
    "mxg_engine_set_waves": (C.c_int, [C.c_void_p, C.c_int]),
    "mxg_engine_get_waves": (C.c_int, [C.c_void_p]),
    "mxg_engine_load": (C.c_int, [C.c_void_p, C.c_int, c_double_p, c_double_p, c_int_p, c_int_p, c_int_p,
                                  c_u16_p]),
    "mxg_engine_set_replica_seed": (C.c_int, [C.c_void_p, C.c_int, C.c_uint64]),
    "mxg_engine_replica_seed": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_uint64)]),
    "mxg_engine_set_bonds": (C.c_int, [C.c_void_p, C.c_int, c_int_p, c_int_p, C.POINTER(C.c_void_p),
                                       c_double_p]),
    "mxg_engine_set_boundaries": (C.c_int, [C.c_void_p, c_int_p, c_float_p, c_float_p, c_float_p]),
    "mxg_engine_set_boundary_potential": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "mxg_engine_set_species": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "mxg_engine_set_species_values": (C.c_int, [C.c_void_p, c_float_p]),
    "mxg_engine_get_species": (C.c_int, [C.c_void_p, c_float_p, c_float_p]),
    "mxg_engine_add_flux": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                      C.c_float, C.c_float]),
    "mxg_engine_set_angles": (C.c_int, [C.c_void_p, C.c_int, c_int_p, c_int_p, c_int_p, C.POINTER(C.c_void_p)]),
    "mxg_engine_set_dihedrals": (C.c_int, [C.c_void_p, C.c_int, c_int_p, c_int_p, c_int_p, c_int_p,
                                           C.POINTER(C.c_void_p)]),
    "mxg_engine_set_exclusions": (C.c_int, [C.c_void_p, C.c_int, c_int_p, c_int_p]),
    "mxg_engine_term_energies": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    "mxg_engine_step": (C.c_int, [C.c_void_p, C.c_int]),
    "mxg_engine_sync": (C.c_int, [C.c_void_p]),
    "mxg_engine_unload": (C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    "mxg_engine_state": (C.c_int, [C.c_void_p, c_float_p, c_int_p, c_float_p, c_float_p, c_float_p]),
    "mxg_engine_neighbour_counts": (C.c_int, [C.c_void_p, c_int_p]),
    "mxg_engine_number_density": (C.c_int, [C.c_void_p, c_float_p]),
    "mxg_engine_observables": (C.c_int, [C.c_void_p, C.c_int, c_double_p]),
    "mxg_engine_bond_active": (C.c_int, [C.c_void_p, c_u8_p]),
    "mxg_engine_geometry": (C.c_int, [C.c_void_p, C.c_int, c_int_p, c_double_p]),
    "mxg_engine_time": (C.c_int, [C.c_void_p, c_u64_p]),
    "mxg_engine_error": (C.c_int, [C.c_void_p]),
    "mxg_engine_error_message": (C.c_char_p, [C.c_void_p]),
    "mxg_engine_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "mxg_engine_timing": (C.c_int, [C.c_void_p, c_double_p, c_double_p]),
    "mxg_adjoint_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, c_double_p, c_int_p, C.c_double, C.c_double,
                                     C.c_float, C.c_int, c_double_p, c_double_p, c_int_p]),
    "mxg_adjoint_free": (None, [C.c_void_p]),
    "mxg_adjoint_error_message": (C.c_char_p, [C.c_void_p]),
    "mxg_adjoint_forward": (C.c_int, [C.c_void_p, c_double_p, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "mxg_adjoint_backward": (C.c_int, [C.c_void_p, c_double_p, C.c_double, C.c_double, C.c_void_p, C.c_int, C.c_int,
                                       C.c_void_p, C.c_void_p, c_double_p, C.c_void_p]),
    "mxg_adjoint_set_pair_kind": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]),
    "mxg_adjoint_set_bonds": (C.c_int, [C.c_void_p, C.c_int, c_int_p, c_int_p, C.c_double, C.c_double]),
    "mxg_adjoint_forces": (C.c_int, [C.c_void_p, c_double_p, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                     c_double_p, C.c_void_p]),
    "mxg_morse_forces_noalloc": (C.c_int, [C.c_int, c_double_p, c_int_p, C.c_double, C.c_int, c_double_p,
                                           c_double_p, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.c_void_p]),
    "mxg_adjoint_rdf": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_void_p]),
    "mxg_adjoint_rdf_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_void_p,
                                           C.c_void_p]),
    "mxg_device_count": (C.c_int, []),
}

_lib = None


def lib():
    """Returns the loaded libmxg.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libmxg.so is not built (expected at %s); run __graft_entry__.build() "
                "or python mechanica_amd/build.py" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
