"""ctypes signatures of the vxg C ABI (include/vxg.h).

The table is prefix-free: ``bind(lib, "vxg_")`` binds the product library.
Test code binds the two parity checkers, which export the same symbol set
under other prefixes, with the same table.
"""
import ctypes as C

_P = C.c_void_p
_D = C.c_double
_F = C.c_float
_I = C.c_int
_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int)
_UP = C.POINTER(C.c_uint32)
_S = C.c_char_p

# name: (restype, argtypes)
SIGNATURES = {
    "create": (_P, [_D]),
    "destroy": (None, [_P]),
    "last_error": (_S, [_P]),
    "voxel_size": (_D, [_P]),
    "set_gravity": (None, [_P, _F]),
    "gravity": (_F, [_P]),
    "enable_floor": (None, [_P, _I]),
    "is_floor_enabled": (_I, [_P]),
    "enable_collisions": (None, [_P, _I]),
    "is_collisions_enabled": (_I, [_P]),
    "set_ambient_temperature": (None, [_P, _F, _I]),
    "ambient_temperature": (_F, [_P]),
    "do_time_step": (_I, [_P, _F]),
    "step": (_I, [_P, _F, _I, _IP]),
    "recommended_time_step": (_F, [_P]),
    "reset_time": (None, [_P]),
    "current_time": (_F, [_P]),
    "state_info": (_F, [_P, _I, _I]),
    "collision_count": (_I, [_P]),
    "set_precision": (_I, [_P, _I]),
    "profile": (_I, [_P, _I]),
    "profile_read": (_I, [_P, _DP, _IP, _I]),
    "material_add": (_I, [_P, _F, _F]),
    "material_count": (_I, [_P]),
    "material_set_model_linear": (_I, [_P, _I, _F, _F]),
    "material_set_model_bilinear": (_I, [_P, _I, _F, _F, _F, _F]),
    "material_set_model": (_I, [_P, _I, _I, _FP, _FP]),
    "material_set": (None, [_P, _I, _I, _F]),
    "material_get": (_F, [_P, _I, _I]),
    "material_set_external_scale": (None, [_P, _I, _D, _D, _D]),
    "material_last_error": (_S, [_P, _I]),
    "material_stress": (_F, [_P, _I, _F, _F, _I]),
    "material_modulus": (_F, [_P, _I, _F]),
    "material_model_data": (_I, [_P, _I, _FP, _FP, _I]),
    "voxel_set": (_I, [_P, _I, _I, _I, _I]),
    "voxel_count": (_I, [_P]),
    "voxel_find": (_I, [_P, _I, _I, _I]),
    "voxel_info": (_I, [_P, _I, _IP]),
    "voxel_set_temperature": (None, [_P, _I, _F]),
    "voxel_temperature": (_F, [_P, _I]),
    "voxel_get_state": (None, [_P, _I, _I, _DP, _DP, _DP, _DP, _UP]),
    "ext_set_fixed": (None, [_P, _I, _I]),
    "ext_set_displacement": (None, [_P, _I, _I, _D]),
    "ext_set_displacement_all": (None, [_P, _I, _DP, _DP]),
    "ext_clear_displacement": (None, [_P, _I, _I]),
    "ext_clear_displacement_all": (None, [_P, _I]),
    "ext_set_force": (None, [_P, _I, _F, _F, _F]),
    "ext_set_moment": (None, [_P, _I, _F, _F, _F]),
    "ext_reset": (None, [_P, _I]),
    "ext_get": (_I, [_P, _I, _IP, _FP, _FP, _DP, _DP]),
    "link_count": (_I, [_P]),
    "link_find": (_I, [_P, _I, _I, _I, _I]),
    "link_info": (_I, [_P, _I, _IP, _IP]),
    "link_get_transverse": (None, [_P, _I, _I, _FP, _FP]),
    "link_get_state": (None, [_P, _I, _I, _FP, _FP, _FP, _FP, _UP, _DP, _DP, _DP]),
}

# extensions of the HIP library with no counterpart in the reference (robots, device-side
# actuation, bulk voxel placement); bound where the library exports them
EXTENSIONS = {
    "voxel_set_many": (_I, [_P, _I, _IP, _IP, _IP]),
    "robot_add": (_I, [_P]),
    "robot_select": (_I, [_P, _I]),
    "robot_count": (_I, [_P]),
    "robot_voxels": (_I, [_P, _I, _IP, _IP]),
    "robot_recommended_time_step": (_F, [_P, _I]),
    "robot_time": (_F, [_P, _I]),
    "robot_step": (_I, [_P, _I, _FP, _IP]),
    "robot_com": (_I, [_P, _DP]),
    "robot_com_noalloc": (_I, [_P, _DP]),
    "voxel_get_materials": (_I, [_P, _I, _I, _IP, _IP]),
    "ext_set_displacement_ramp": (None, [_P, _I, _DP, _DP]),
    "set_actuation": (_I, [_P, _I, _F, _F, _F]),
    "voxel_set_phase": (_I, [_P, _I, _I, _FP]),
    "voxel_set_modulus": (_I, [_P, _I, _I, _FP]),
    "clear_moduli": (None, [_P]),
    "adjoint_run": (_I, [_P, _I, _I, _FP]),
    "adjoint_backward": (_I, [_P, _DP, _DP, _DP]),
    "adjoint_free": (None, [_P]),
    "set_graphs": (_I, [_P, _I]),
    "graph_stats": (_I, [_P, _IP, _IP]),
    "autotune": (_I, [_P, _I]),
    "launch_config": (_I, [_P, _I, _IP, _IP, _IP]),
    "set_launch_config": (_I, [_P, _I, _I, _I]),
    "linear_solve": (_I, [_P, C.c_double, _I, _IP, _DP]),
    "set_fp64_path": (_I, [_P, _I]),
    "load_json": (_I, [_P, C.c_char_p]),
    "robot_state_info": (_I, [_P, _I, _I, _DP]),
    "checkpoint_save": (_I, [_P, _I]),
    "checkpoint_restore": (_I, [_P, _I]),
    "checkpoint_free": (None, [_P, _I]),
    "save_json": (_I, [_P, C.c_char_p]),
    "fp64_path": (_I, [_P]),
    "index_bounds": (None, [_P, _IP, _IP]),
    "material_remove": (_I, [_P, _I]),
    "material_replace": (_I, [_P, _I, _I]),
    "set_voxel_size": (None, [_P, _D]),
}


class ABI:
    """The C ABI of one library, bound under one symbol prefix."""

    def __init__(self, lib, prefix, name=None):
        self.lib = lib
        self.prefix = prefix
        self.name = name or prefix.rstrip("_")
        for key, (res, args) in SIGNATURES.items():
            fn = getattr(lib, prefix + key)
            fn.restype = res
            fn.argtypes = args
            setattr(self, key, fn)
        self.extensions = set()
        for key, (res, args) in EXTENSIONS.items():
            if not hasattr(lib, prefix + key):
                continue
            fn = getattr(lib, prefix + key)
            fn.restype = res
            fn.argtypes = args
            setattr(self, key, fn)
            self.extensions.add(key)

    def __repr__(self):
        return f"ABI({self.name})"


def bind(path_or_lib, prefix, name=None):
    lib = path_or_lib if isinstance(path_or_lib, C.CDLL) else C.CDLL(str(path_or_lib))
    return ABI(lib, prefix, name)
