"""ctypes binding of libprandtl_hip.so (the C ABI declared in include/prandtl_hip.h).

The HIP library is the product path: when it is missing this module raises,
there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PRANDTL_HIP_LIB") or os.path.join(_HERE, "libprandtl_hip.so")

PRANDTL_OK = 0
PRANDTL_ERR_ARG = 1
PRANDTL_ERR_UNSUPPORTED = 2
PRANDTL_ERR_HIP = 3
PRANDTL_ERR_STATE = 4
PRANDTL_ERR_NOMEM = 5
PRANDTL_ERR_IO = 6
PRANDTL_FP64 = 0
PRANDTL_MIXED = 1
PRANDTL_SHAPES = 4
PRANDTL_KERNEL_RHS = 0
PRANDTL_KERNEL_VISC = 1
PRANDTL_ALGO_TILE = 0
PRANDTL_ALGO_LINE = 1
PRANDTL_ENS_MAX_PER_CU = 4

EXPORTED = (
    "prandtl_create", "prandtl_destroy", "prandtl_sizes", "prandtl_rhs", "prandtl_rk3_step",
    "prandtl_set_graph", "prandtl_set_shape", "prandtl_autotune", "prandtl_run",
    "prandtl_max_wave_speed", "prandtl_alpha", "prandtl_check_state", "prandtl_lgl",
    "prandtl_debug_gradients", "prandtl_primitives", "prandtl_profile_begin", "prandtl_profile_end", "prandtl_last_error",
    "prandtl_ensemble_create", "prandtl_ensemble_destroy", "prandtl_ensemble_case",
    "prandtl_ensemble_rk3_step", "prandtl_ensemble_last_error",
    "prandtl_checkpoint_write", "prandtl_checkpoint_read", "prandtl_checkpoint_save", "prandtl_checkpoint_load",
    "prandtl_create_quad", "prandtl_hmin", "prandtl_mesh_read", "prandtl_mesh_refine", "prandtl_mesh_view",
    "prandtl_mesh_name", "prandtl_mesh_free", "prandtl_mesh_last_error", "prandtl_axis_recover",
    "prandtl_create_hex", "prandtl_mesh_dim", "prandtl_mesh_view_hex",
    "prandtl_set_algorithm", "prandtl_get_algorithm", "prandtl_rhs_noalloc",
    "prandtl_set_time", "prandtl_get_time", "prandtl_set_boundary_split",
    "prandtl_ensemble_set_mode", "prandtl_ensemble_autotune",
    "prandtl_peak_fma", "prandtl_peak_copy",
)


class MeshDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("order", ctypes.c_int), ("nelem", ctypes.c_int * 3),
                ("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3),
                ("bc", (ctypes.c_int * 2) * 3), ("bc_state", ((ctypes.c_double * 5) * 2) * 3)]


class QuadMesh(ctypes.Structure):
    _fields_ = [("nvert", ctypes.c_int), ("vert", ctypes.POINTER(ctypes.c_double)),
                ("nelem", ctypes.c_int), ("elem", ctypes.POINTER(ctypes.c_int)),
                ("elem_coords", ctypes.POINTER(ctypes.c_double)),
                ("nbdr", ctypes.c_int), ("bdr", ctypes.POINTER(ctypes.c_int)),
                ("bdr_attr", ctypes.POINTER(ctypes.c_int)),
                ("nattr", ctypes.c_int), ("attr_bc", ctypes.POINTER(ctypes.c_int)),
                ("attr_state", ctypes.POINTER(ctypes.c_double))]


class HexMesh(QuadMesh):
    """prandtl_hex_mesh: the same fields (vert [nv][3], elem [ne][8], elem_coords
    [ne][8][3], bdr [nb][4], attr_state [nattr][5])."""


# boundary types (PRANDTL_BC_*) by the reference's configuration names
BC_TYPES = {"periodic": 0, "slip": 1, "supersonic-inflow": 2, "supersonic-outflow": 3,
            "symmetry": 4, "specified-state": 5, "no-slip-adiabatic": 6, "axis": 7}


class Params(ctypes.Structure):
    _fields_ = [("gamma", ctypes.c_double), ("Pr", ctypes.c_double), ("mu", ctypes.c_double),
                ("alpha_max", ctypes.c_double), ("alpha_min", ctypes.c_double),
                ("parabolic", ctypes.c_int), ("blending", ctypes.c_int),
                ("ec_surface", ctypes.c_int), ("quirks", ctypes.c_int), ("precision", ctypes.c_int),
                ("sutherland", ctypes.c_int), ("R_gas", ctypes.c_double),
                ("axisymmetric", ctypes.c_int), ("axis_rho_floor", ctypes.c_double)]


class PrandtlError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("prandtl error %d: %s" % (code, msg))
        self.code = code


_lib = None


def lib():
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libprandtl_hip.so is not built; run __graft_entry__.build()")
    # torch ships its own libamdhip64.so.7; importing it first makes the
    # dynamic loader resolve this library's libamdhip64.so.7 to that same
    # runtime, so device pointers and streams are shared with torch.
    import torch  # noqa: F401
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    L.prandtl_create.argtypes = [ctypes.POINTER(MeshDesc), ctypes.POINTER(Params), ctypes.POINTER(vp)]
    L.prandtl_destroy.argtypes = [vp]
    L.prandtl_destroy.restype = None
    L.prandtl_sizes.argtypes = [vp, ip, ip, ip]
    L.prandtl_rhs.argtypes = [vp, vp, vp, vp]
    L.prandtl_rhs_noalloc.argtypes = [ctypes.POINTER(MeshDesc), ctypes.POINTER(Params), vp, vp, vp]
    L.prandtl_rk3_step.argtypes = [vp, vp, ctypes.c_double, vp]
    L.prandtl_set_graph.argtypes = [vp, ctypes.c_int]
    L.prandtl_set_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.prandtl_set_algorithm.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.prandtl_set_time.argtypes = [vp, ctypes.c_double, vp]
    L.prandtl_ensemble_set_mode.argtypes = [vp, ctypes.c_int]
    L.prandtl_peak_fma.argtypes = [ctypes.c_int, vp, dp, vp]
    L.prandtl_peak_copy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int, dp, vp]
    L.prandtl_ensemble_autotune.argtypes = [vp, vp, dp, ip, dp, vp]
    L.prandtl_get_time.argtypes = [vp, dp]
    L.prandtl_set_boundary_split.argtypes = [vp, ctypes.c_int, dp, ctypes.c_double, ctypes.c_double, dp, dp]
    L.prandtl_get_algorithm.argtypes = [vp, ctypes.c_int, ip]
    L.prandtl_autotune.argtypes = [vp, vp, ip, dp, vp]
    L.prandtl_run.argtypes = [vp, vp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                              ctypes.c_int, ip, vp]
    L.prandtl_max_wave_speed.argtypes = [vp, dp, ctypes.c_int, vp]
    L.prandtl_alpha.argtypes = [vp, dp, dp, vp]
    L.prandtl_check_state.argtypes = [vp, vp, vp]
    L.prandtl_lgl.argtypes = [ctypes.c_int, dp, dp]
    L.prandtl_debug_gradients.argtypes = [vp, vp, vp, vp]
    L.prandtl_primitives.argtypes = [vp, vp, vp, vp]
    L.prandtl_axis_recover.argtypes = [vp, vp, vp, vp]
    L.prandtl_profile_begin.argtypes = [vp]
    L.prandtl_profile_end.argtypes = [vp, dp, ip]
    L.prandtl_last_error.argtypes = [vp]
    L.prandtl_last_error.restype = ctypes.c_char_p
    cp = ctypes.c_char_p
    L.prandtl_checkpoint_write.argtypes = [cp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(MeshDesc), ctypes.c_int, dp]
    L.prandtl_checkpoint_read.argtypes = [cp, ctypes.c_int, ctypes.POINTER(MeshDesc), ctypes.c_int, dp, dp]
    L.prandtl_checkpoint_save.argtypes = [vp, vp, cp, ctypes.c_int, ctypes.c_double, vp]
    L.prandtl_checkpoint_load.argtypes = [vp, vp, cp, ctypes.c_int, dp, vp]
    L.prandtl_ensemble_create.argtypes = [ctypes.c_int, ctypes.POINTER(MeshDesc), ctypes.POINTER(Params),
                                          ctypes.POINTER(vp)]
    L.prandtl_ensemble_destroy.argtypes = [vp]
    L.prandtl_ensemble_destroy.restype = None
    L.prandtl_ensemble_case.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
    L.prandtl_ensemble_rk3_step.argtypes = [vp, ctypes.POINTER(vp), dp, vp]
    L.prandtl_ensemble_last_error.argtypes = [vp]
    L.prandtl_ensemble_last_error.restype = ctypes.c_char_p
    L.prandtl_create_quad.argtypes = [ctypes.POINTER(QuadMesh), ctypes.c_int, ctypes.POINTER(Params),
                                      ctypes.POINTER(vp)]
    L.prandtl_create_hex.argtypes = [ctypes.POINTER(HexMesh), ctypes.c_int, ctypes.POINTER(Params),
                                     ctypes.POINTER(vp)]
    L.prandtl_hmin.argtypes = [vp, dp]
    L.prandtl_mesh_dim.argtypes = [vp]
    L.prandtl_mesh_dim.restype = ctypes.c_int
    L.prandtl_mesh_view_hex.argtypes = [vp, ctypes.POINTER(HexMesh)]
    L.prandtl_mesh_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
    L.prandtl_mesh_refine.argtypes = [vp]
    L.prandtl_mesh_view.argtypes = [vp, ctypes.POINTER(QuadMesh)]
    L.prandtl_mesh_name.argtypes = [vp, ctypes.c_int, ip]
    L.prandtl_mesh_name.restype = ctypes.c_char_p
    L.prandtl_mesh_free.argtypes = [vp]
    L.prandtl_mesh_last_error.argtypes = []
    L.prandtl_mesh_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def check(rc, ctx=None):
    if rc != PRANDTL_OK:
        msg = lib().prandtl_last_error(ctx)
        raise PrandtlError(rc, msg.decode() if msg else "")
    return rc


def check_ensemble(rc, ens=None):
    if rc != PRANDTL_OK:
        msg = lib().prandtl_ensemble_last_error(ens)
        raise PrandtlError(rc, msg.decode() if msg else "")
    return rc
