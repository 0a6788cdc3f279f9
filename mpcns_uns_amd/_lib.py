"""ctypes binding of libmhdg.so (include/mhdg.h, include/mhdg_mesh.h).

The library is built in-tree by mpcns_uns_amd/build.py.  Loading fails loudly when it is missing: there is
no fallback implementation of any entry point.
"""
import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libmhdg.so")


class Incidence(C.Structure):
    _fields_ = [("ptr", C.POINTER(C.c_int64)), ("idx", C.POINTER(C.c_int32)), ("sgn", C.POINTER(C.c_int8))]


class MeshDesc(C.Structure):
    _fields_ = [
        ("n_edges", C.c_int32), ("n_faces", C.c_int32), ("n_cells", C.c_int32),
        ("f2e", Incidence), ("c2f", Incidence), ("e2f", Incidence), ("f2c", Incidence),
        ("cell_material", C.POINTER(C.c_int32)),
        ("face_kind", C.POINTER(C.c_int8)), ("edge_mode", C.POINTER(C.c_int8)),
        ("face_area", C.POINTER(C.c_double)), ("face_center", C.POINTER(C.c_double)),
        ("face_area_mag", C.POINTER(C.c_double)), ("cell_center", C.POINTER(C.c_double)),
        ("cell_volume", C.POINTER(C.c_double)), ("edge_vec", C.POINTER(C.c_double)),
        ("edge_mid", C.POINTER(C.c_double)), ("edge_len", C.POINTER(C.c_double)),
        ("ec_ptr", C.POINTER(C.c_int64)), ("ec_cell", C.POINTER(C.c_int32)),
        ("ec_weight", C.POINTER(C.c_double)), ("ef_weight", C.POINTER(C.c_double)),
        ("star1", C.POINTER(C.c_double)), ("star2", C.POINTER(C.c_double)),
        ("e2n", Incidence), ("n_nodes", C.c_int32),
    ]


class Config(C.Structure):
    _fields_ = [
        ("gamma", C.c_double), ("inver_MA2", C.c_double), ("M_ref", C.c_double), ("M_H", C.c_double),
        ("M_Na", C.c_double), ("ne_floor", C.c_double), ("hall_coef", C.c_double), ("inver_Rem", C.c_double),
        ("cfl", C.c_double), ("recon_rel_limit", C.c_double), ("rho_floor", C.c_double), ("p_floor", C.c_double),
        ("rho_floor_dt", C.c_double),
        ("second_order", C.c_int32), ("lorentz_force", C.c_int32), ("hall", C.c_int32), ("resistive", C.c_int32),
        ("eta_r_inner", C.c_double), ("eta_r_outer", C.c_double), ("eta_width", C.c_double),
        ("hall_r_min", C.c_double), ("hall_r_max", C.c_double),
        ("ind_rtol", C.c_double), ("ind_max_it", C.c_int32), ("ind_restart", C.c_int32),
        ("cuct", C.c_double), ("fs", (C.c_double * 5) * 2), ("u_inf", C.c_double * 3), ("B_imf", C.c_double * 3),
        ("wall_dirichlet", C.c_int32),
        ("implicit_ideal", C.c_int32),
        ("n_species", C.c_int32),
        ("ion_mass", C.c_double * 10),
        ("ion_fs", (C.c_double * 5) * 10),
        ("src_rate", C.c_double * 10),
        ("src_scale_height", C.c_double), ("src_r0", C.c_double), ("src_T", C.c_double),
        ("induce_coeff", C.c_double),
        ("mg_levels", C.c_int32), ("mg_agglom_passes", C.c_int32), ("mg_pre", C.c_int32), ("mg_post", C.c_int32),
        ("mg_coarse", C.c_int32), ("mg_relax", C.c_double), ("mg_fine_hall", C.c_int32),
    ]


class MarsParams(C.Structure):
    """mhdg_mars_params (include/mhdg.h): scales and particle masses of the Mars point-implicit sources."""
    _fields_ = [
        ("kB", C.c_double), ("q", C.c_double), ("L", C.c_double), ("U", C.c_double), ("B", C.c_double), ("n", C.c_double),
        ("rho_ref", C.c_double), ("p_ref", C.c_double), ("time_ref", C.c_double),
        ("rho_floor", C.c_double), ("p_floor", C.c_double),
        ("ion_mass", C.c_double * 10),
        ("n_neutral", C.c_int32), ("neutral_stride", C.c_int32),
        ("neutral_mass", C.c_double * 10),
    ]


_lib = None


def lib():
    """The loaded libmhdg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        P, D, I32, I64 = C.c_void_p, C.c_double, C.c_int32, C.c_int64
        L.mhdg_last_error.restype = C.c_char_p
        L.mhdg_mesh_error.restype = C.c_char_p
        L.mhdg_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(Config), C.POINTER(P)]
        L.mhdg_destroy.argtypes = [P]
        L.mhdg_last_error.argtypes = [P]
        L.mhdg_config_defaults.argtypes = [C.POINTER(Config)]
        L.mhdg_mesh_shell.argtypes = [C.c_int, C.c_int, C.c_int, D, D, D, D, C.c_uint64, C.POINTER(P)]
        L.mhdg_mesh_from_cells.argtypes = [I32, P, I64, P, P, P, P, I64, P, P, P, C.POINTER(P)]
        L.mhdg_mesh_free.argtypes = [P]
        L.mhdg_stream_calibration.argtypes = [C.c_int, I64, C.c_int, C.POINTER(D), C.POINTER(D)]
        L.mhdg_mesh_hilbert_order.argtypes = [P, P]
        L.mhdg_mesh_renumber_cells.argtypes = [P, P, C.POINTER(P)]
        L.mhdg_mesh_counts.argtypes = [P, P]
        L.mhdg_mesh_array.argtypes = [P, C.c_char_p, C.POINTER(P), C.POINTER(I64), C.POINTER(I32)]
        L.mhdg_mesh_boundary.argtypes = [P, P, I32, I32, P]
        L.mhdg_mesh_desc_fill.argtypes = [P, C.POINTER(MeshDesc)]
        L.mhdg_mesh_background.argtypes = [P, P, C.c_int, P, P, D, P, P]
        L.mhdg_mesh_cell_vector.argtypes = [P, P, P]
        L.mhdg_mesh_seeded_state.argtypes = [P, C.POINTER(Config), D, C.c_uint64, P, P, P, P]
        L.mhdg_mesh_read_unsbin.argtypes = [C.c_char_p, C.POINTER(P)]
        L.mhdg_mesh_read_unsbin_parts.argtypes = [C.POINTER(C.c_char_p), C.c_int32, C.POINTER(P)]
        L.mhdg_mesh_background_dipoles.argtypes = [P, P, I32, P, P, P, D, P, P]
        L.mhdg_dipole_edge_integrals.argtypes = [I32, P, P, P, D, I64, P, P, P]
        L.mhdg_mesh_write_unsbin.argtypes = [P, C.c_char_p]
        L.mhdg_mesh_topology_hash.argtypes = [P, C.POINTER(C.c_uint64)]
        L.mhdg_mesh_tag_names.argtypes = [P, C.c_char_p, C.c_int]
        L.mhdg_mesh_set_tag_names.argtypes = [P, C.c_char_p]
        L.mhdg_fields_write.argtypes = [C.c_char_p, C.c_uint64, P, I32]
        L.mhdg_mesh_write_vtu.argtypes = [P, C.c_char_p, I32, P, I32, P]
        L.mhdg_write_pvtu.argtypes = [C.c_char_p, I32, P, I32, P, I32, P]
        L.mhdg_write_pvd.argtypes = [C.c_char_p, I32, P, P]
        L.mhdg_fields_read.argtypes = [C.c_char_p, C.POINTER(P), C.POINTER(C.c_uint64), C.POINTER(I32)]
        L.mhdg_fields_get.argtypes = [P, I32, P]
        L.mhdg_fields_free.argtypes = [P]
        L.mhdg_mg_build.argtypes = [C.POINTER(MeshDesc), C.POINTER(Config), P, C.POINTER(P)]
        L.mhdg_mg_free.argtypes = [P]
        L.mhdg_mg_levels.argtypes = [P]
        L.mhdg_mg_level_size.argtypes = [P, C.c_int, P, P, P]
        L.mhdg_mg_level_arrays.argtypes = [P, C.c_int, P, P, P, P, P]
        L.mhdg_mg_level_nodes.argtypes = [P, C.c_int, P, P, P]
        for name, args in {
            "mhdg_cell_vector": [P, P, P], "mhdg_cell_current": [P, P, P],
            "mhdg_limited_gradients": [P, P, C.c_int, P, C.c_int, P, P],
            "mhdg_flux_rhs": [P, C.c_int, P, P, P, P, P],
            "mhdg_ideal_emf": [P, P, P, P, P, P, D, P],
            "mhdg_faraday": [P, P, P, D], "mhdg_faraday_edge_dt": [P, P, P, P],
            "mhdg_div_b": [P, P, P, P, P, P], "mhdg_div_b_noalloc": [P, P, P, P], "mhdg_mars_implicit_sources": [P, P, P, D], "mhdg_local_dt": [P, P, P, P, P],
            "mhdg_add_sources": [P, C.c_int, P, P, P, P, D, P],
            "mhdg_ind_set_frozen": [P, P, P], "mhdg_ind_apply_F": [P, P, P],
            "mhdg_ind_solve": [P, P, P, P, D, P, C.c_int, P, P, P],
            "mhdg_ind_flags": [P, P, P, P], "mhdg_mg_apply": [P, P, D, P, P], "mhdg_mg_info": [P, P, I32, P, P], "mhdg_ind_set_ideal_state": [P, P, P], "mhdg_ind_ideal_linear": [P, P, P],
            "mhdg_step_explicit": [P, D, P], "mhdg_set_species_state": [P, C.c_int, P], "mhdg_get_species_state": [P, C.c_int, P],
            "mhdg_get_species_prim": [P, C.c_int, P], "mhdg_ind_apply_G": [P, P, P], "mhdg_ind_jvp": [P, P, P, P],
            "mhdg_ind_newton": [P, P, P, P, D, P, D, C.c_int, D, P, P, P, P],
            "mhdg_shell_energy": [P, D, D, P],
            "mhdg_step_vjp": [P, C.c_int, D, D, D, P, P, P, P, P, P, P, P],
            "mhdg_explicit_tangent": [P, C.c_int, D, P, P, P, P, P, P, P, D, D, P, P, P, P, P], "mhdg_set_recon_rel_limit": [P, D], "mhdg_step_hall": [P, D, C.c_int, P, P, P],
            "mhdg_set_state": [P, P, P, P], "mhdg_get_state": [P, P, P, P],
            "mhdg_set_background": [P, P, P], "mhdg_step": [P, D, C.c_int, P, P, P],
            "mhdg_get_derived": [P, C.c_int, P], "mhdg_prepare": [P], "mhdg_state_save": [P], "mhdg_state_restore": [P], "mhdg_sync": [P], "mhdg_device_info": [C.c_char_p, C.c_int],
            "mhdg_profile": [P, C.c_int], "mhdg_profile_only": [P, C.c_char_p], "mhdg_profile_report": [P, C.c_char_p, C.c_int],
        }.items():
            getattr(L, name).argtypes = args
        _lib = L
    return _lib


def ptr(a):
    """ctypes pointer to a contiguous numpy array's data (None passes NULL).  The pointer object holds a
    reference to the array, so a temporary made for the call (ptr(f64(x))) lives until the call returns."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("arrays crossing the C-ABI must be contiguous")
    return a.ctypes.data_as(C.c_void_p)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
