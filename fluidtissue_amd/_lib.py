"""ctypes binding of libftreinit.so (include/ftreinit.h).

The library is the product: there is no fallback.  If the shared object is
missing, importing this module raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libftreinit.so")

FT_OK = 0
FT_ERR_ARG = -1
FT_ERR_NODEVICE = -2
FT_ERR_HIP = -3
FT_ERR_NOMEM = -4
FT_ERR_T2E = -5
FT_ERR_DEGENERATE = -6
FT_ERR_STATE = -7
FT_ERR_BUDGET = -8
FT_ERR_SOLVER = -9
FT_ERR_INTERNAL = -10

FT_DIST_POINTS = 0
FT_DIST_TRIANGLES = 1
FT_BLEND_INIT = 0
FT_BLEND_UPDATE = 1
FT_SCHED_STATIC = 0
FT_SCHED_QUEUE = 1
FT_SCHED_BINS = 2

FT_PHI_SPHERE_WAVE = 1
FT_PHI_STAR = 2
FT_PHI_SPHERES = 3
FT_PHI_ELLIPSOIDS = 4

FT_NOALLOC_SCRATCH_BYTES = 128 << 20

FT_OBJ_INTERFACE = 0x02
FT_OBJ_BOUNDARY = 0x04


class FtOpts(C.Structure):
    _fields_ = [
        ("mode", C.c_int32),
        ("blend", C.c_int32),
        ("w_phi", C.c_double),
        ("w_ifn", C.c_double),
        ("band", C.c_double),
        ("bin_scale", C.c_float),
        ("want_nearest", C.c_int32),
        ("query_block", C.c_int32),
        ("want_normals", C.c_int32),
        ("waves_per_simd", C.c_int32),
        ("schedule", C.c_int32),
        ("reserved", C.c_int32 * 2),
    ]


class FtCounts(C.Structure):
    _fields_ = [
        ("n_ifp", C.c_int64),
        ("n_tri", C.c_int64),
        ("n_if_elems", C.c_int64),
        ("n_clamped", C.c_int64),
        ("bad_tet", C.c_int64),
    ]


class FtAdvectInfo(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("restarts", C.c_int32),
        ("error", C.c_double),
        ("converged", C.c_int32),
        ("skipped", C.c_int32),
    ]


class FtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ftreinit error {code}: {msg}")
        self.code = code


# PyTorch ships its own HIP runtime under the same SONAME (libamdhip64.so.7)
# as /opt/rocm's.  Loading torch first makes this library bind to that one
# runtime, so device pointers from torch tensors and from this library live
# in a single HIP/HSA context; loading it second would put two runtimes in
# the process and torch then finds no GPU.
try:
    import torch  # noqa: F401
except ImportError:
    pass

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run `python fluidtissue_amd/build.py`")

lib = C.CDLL(LIB_PATH)

_P = C.c_void_p
_H = C.c_void_p
_SIGS = {
    "ft_opts_default": (None, [C.POINTER(FtOpts)]),
    "ft_last_error": (C.c_char_p, []),
    "ft_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
    "ft_create": (C.c_int, [C.c_int32, C.POINTER(_H)]),
    "ft_destroy": (C.c_int, [_H]),
    "ft_set_nodes": (C.c_int, [_H, _P, C.c_int64]),
    "ft_set_mesh": (C.c_int, [_H, _P, C.c_int64, _P, C.c_int64, _P, C.c_int64, _P]),
    "ft_reset_phi_points": (C.c_int, [_H, _P, C.c_int64, _P]),
    "ft_reinit": (C.c_int, [_H, _P, _P, C.POINTER(FtOpts), C.POINTER(FtCounts)]),
    "ft_reinit_device": (C.c_int, [_H, _P, _P, C.POINTER(FtOpts), C.POINTER(FtCounts), _P]),
    "ft_reinit_device_noalloc": (C.c_int, [_H, _P, _P, C.POINTER(FtOpts), C.POINTER(FtCounts), _P]),
    "ft_get_interface": (C.c_int, [_H, _P, _P, _P, _P, _P, _P, _P]),
    "ft_get_nearest": (C.c_int, [_H, _P]),
    "ft_interface_device": (C.c_int, [_H, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
    "ft_set_mesh_batch": (C.c_int, [_H, C.c_int32, _P, _P, _P, _P, _P, _P, _P]),
    "ft_set_kuhn_batch": (C.c_int, [_H, C.c_int32, _P, _P, _P, C.c_double, C.c_uint64]),
    "ft_batch_info": (C.c_int, [_H, C.POINTER(C.c_int32), _P, _P, _P]),
    "ft_get_mesh": (C.c_int, [_H, _P, _P, _P, _P]),
    "ft_get_field_counts": (C.c_int, [_H, _P]),
    "ft_mesh_create_kuhn": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                      C.c_uint64, C.POINTER(C.c_int32)]),
    "ft_mesh_create": (C.c_int, [C.c_int32, _P, C.c_int64, _P, C.c_int64, _P, C.c_int64, _P,
                                 C.POINTER(C.c_int32)]),
    "ft_mesh_info": (C.c_int, [C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                               C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "ft_mesh_release": (C.c_int, [C.c_int32]),
    "ft_set_batch_meshes": (C.c_int, [_H, C.c_int32, _P]),
    "ft_reinit_batch": (C.c_int, [_H, _P, _P, C.POINTER(FtOpts), _P, _P]),
    "ft_reinit_graph": (C.c_int, [_H, _P, _P, C.POINTER(FtOpts), _P, _P]),
    "ft_graph_stats": (C.c_int, [_H, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "ft_get_normals": (C.c_int, [_H, _P]),
    "ft_interface_curvature": (C.c_int, [_H, _P, _P]),
    "ft_surface_tension": (C.c_int, [_H, C.c_float, C.c_int32, _P, _P, _P]),
    "ft_advect": (C.c_int, [_H, _P, _P, C.c_double, C.c_double, C.c_int32, C.POINTER(FtAdvectInfo)]),
    "ft_advect_device": (C.c_int, [_H, _P, _P, C.c_double, C.c_double, C.c_int32, C.POINTER(FtAdvectInfo),
                                   _P]),
    "ft_fem_assemble": (C.c_int, [_H, C.c_int32, _P, _P, _P, _P]),
    "ft_surface_tension_device": (C.c_int, [_H, C.c_float, C.c_int32, _P, _P, _P, _P]),
    "ft_get_counts": (C.c_int, [_H, _P]),
    "ft_write_interface_obj": (C.c_int, [_H, C.c_char_p, C.c_int32]),
    "ft_write_obj": (C.c_int, [_H, C.c_char_p, C.c_int32, C.c_int32, _P, C.c_int64]),
    "ft_normals_device": (C.c_int, [_H, C.POINTER(C.c_void_p)]),
    "ft_set_timing": (C.c_int, [_H, C.c_int32]),
    "ft_get_timing": (C.c_int, [_H, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "ft_kuhn_counts": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "ft_kuhn_mesh": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, _P, _P, C.c_double, C.c_uint64,
                               _P, _P, _P, _P]),
    "ft_phi_field": (C.c_int, [C.c_int32, C.c_int64, _P, C.c_uint64, _P]),
}
OPTIONAL = set()


def _bind():
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


_bind()


def check(rc):
    if rc != FT_OK:
        raise FtError(rc, lib.ft_last_error().decode(errors="replace"))
    return rc


def exported_symbols():
    return sorted(_SIGS)
