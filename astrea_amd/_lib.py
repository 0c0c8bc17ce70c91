"""ctypes binding of libastrea_hip.so (C ABI declared in include/astrea_hip.h).

The library is built in-tree by astrea_amd._build; importing this module
never builds anything.  ``load()`` raises if the library is missing: there is
no CPU fallback on the product path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libastrea_hip.so")

AH_OK, AH_ERR_INVALID, AH_ERR_UNSUPPORTED, AH_ERR_HIP, AH_ERR_NOMEM, AH_ERR_STATE, AH_ERR_TOO_LARGE = 0, -1, -2, -3, -4, -5, -6
AH_STATIC_SCRATCH_BYTES = 64 << 20
AH_EDGE, AH_WRAP = 0, 1
AH_PCM, AH_PLM, AH_PPM, AH_WENO3, AH_WENO5, AH_WENO7 = range(6)
AH_MINMOD, AH_VANLEER, AH_OSPRE, AH_VANALBADA1, AH_KOREN, AH_SUPERBEE = range(6)
AH_FLAT_OFF, AH_FLAT_ASTREA, AH_FLAT_COLELLA90 = range(3)
AH_PPM_MC, AH_PPM_C, AH_PPM_PH = range(3)
AH_LLF, AH_HLLC, AH_HLLD, AH_DOTS, AH_ES, AH_LW = range(6)
AH_EULER, AH_RK4, AH_SSPRK22, AH_SSPRK33, AH_SSPRK43, AH_SSPRK53, AH_SSPRK54, AH_SSPRK104 = range(8)
AH_FP64, AH_SPLIT = 0, 1
AH_COMPAT_ASTREA, AH_COMPAT_PHYSICAL = 0, 1
AH_VARS_H1, AH_VARS_H2, AH_VARS_M6, AH_VARS_ALL = 0x13, 0x17, 0x77, 0xFF

# every extern "C" symbol of include/astrea_hip.h
EXPORTS = ["ah_version", "ah_status_str", "ah_plan_create", "ah_plan_destroy", "ah_plan_layout",
           "ah_plan_workspace_size", "ah_plan_bind", "ah_plan_reset", "ah_plan_step", "ah_plan_status",
           "ah_evolve_space", "ah_compute_L", "ah_plan_profile", "ah_plan_profile_read", "ah_plan_step_graph",
           "ah_plan_compact", "ah_plan_set_blocks", "ah_plan_time_export", "ah_plan_time_import",
           "ah_plan_init_ic", "ah_plan_snapshot", "ah_plan_set_capture", "ah_plan_status_async",
           "ah_evolve_space_ct", "ah_compute_L_ct", "ah_evolve_time", "ah_evolve_space_static"]
AH_IC_RIEMANN_1D, AH_IC_QUADRANTS_2D = 0, 1
AH_CAPTURE_OFF, AH_CAPTURE_EVERY_STEP, AH_CAPTURE_CHECKPOINTS = 0, 1, 2
KERNEL_KINDS = ["sweep1d", "recon_ax0", "flux_ax0", "recon_ax1", "flux_ax1", "combine", "dt", "advance"]


class AhScheme(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int), ("recon", ctypes.c_int), ("limiter", ctypes.c_int),
                ("flattener", ctypes.c_int), ("solver", ctypes.c_int), ("rk", ctypes.c_int),
                ("precision", ctypes.c_int), ("varmask", ctypes.c_int), ("magnetic_2d", ctypes.c_int),
                ("tile_s", ctypes.c_int), ("tile_t", ctypes.c_int), ("compat", ctypes.c_int),
                ("av", ctypes.c_int), ("ppm_author", ctypes.c_int), ("fuse2d", ctypes.c_int),
                ("reserved", ctypes.c_int * 1)]


class AhProblem(ctypes.Structure):
    _fields_ = [("n0", ctypes.c_int), ("n1", ctypes.c_int), ("boundary", ctypes.c_int),
                ("checkpoints", ctypes.c_int), ("dx", ctypes.c_double), ("gamma", ctypes.c_double),
                ("cfl", ctypes.c_double), ("t_end", ctypes.c_double), ("max_steps", ctypes.c_longlong)]


class AhProblemStatus(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("dt", ctypes.c_double), ("steps", ctypes.c_longlong),
                ("active", ctypes.c_int), ("failed", ctypes.c_int)]


class AhTimeState(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("dt", ctypes.c_double), ("steps", ctypes.c_longlong),
                ("idx", ctypes.c_int), ("active", ctypes.c_int), ("failed", ctypes.c_int), ("parity", ctypes.c_int)]


class AhIc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("high_order", ctypes.c_int), ("x0", ctypes.c_double),
                ("y0", ctypes.c_double), ("w", (ctypes.c_double * 8) * 4)]


class AstreaHipError(RuntimeError):
    pass


_lib = None


def load():
    """Load the in-tree library (once) and declare the signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AstreaHipError(f"{LIB_PATH} is missing: build it with `python -m astrea_amd._build` "
                             "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    vp, i, ll, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_double
    P = ctypes.POINTER
    lib.ah_version.argtypes = [P(i), P(i)]
    lib.ah_status_str.argtypes = [i]
    lib.ah_status_str.restype = ctypes.c_char_p
    lib.ah_plan_create.argtypes = [P(AhScheme), i, P(AhProblem), P(vp)]
    lib.ah_plan_destroy.argtypes = [vp]
    lib.ah_plan_layout.argtypes = [vp, P(ll), P(i), P(i), P(ll)]
    lib.ah_plan_workspace_size.argtypes = [vp, P(ctypes.c_size_t)]
    lib.ah_plan_bind.argtypes = [vp, vp, vp, ctypes.c_size_t]
    lib.ah_plan_reset.argtypes = [vp, vp]
    lib.ah_plan_step.argtypes = [vp, i, vp]
    lib.ah_plan_status.argtypes = [vp, P(AhProblemStatus), vp]
    lib.ah_plan_status_async.argtypes = [vp, vp, vp]
    lib.ah_evolve_space.argtypes = [vp, i, i, vp, vp, vp, P(d), vp]
    lib.ah_compute_L.argtypes = [vp, i, i, vp, vp, vp, vp]
    lib.ah_evolve_space_ct.argtypes = [vp, i, i, vp, vp, vp, P(d), vp, vp, vp, vp]
    lib.ah_compute_L_ct.argtypes = [vp, i, i, vp, vp, vp, vp, vp]
    lib.ah_evolve_time.argtypes = [vp, i, i, vp, vp, vp, vp, vp, vp, d, vp]
    lib.ah_evolve_space_static.argtypes = [P(AhScheme), P(AhProblem), i, vp, vp, vp, P(d), vp]
    lib.ah_plan_profile.argtypes = [vp, i]
    lib.ah_plan_step_graph.argtypes = [vp, i, i, vp]
    lib.ah_plan_compact.argtypes = [vp, P(i), vp]
    lib.ah_plan_set_blocks.argtypes = [vp, i]
    lib.ah_plan_time_export.argtypes = [vp, P(AhTimeState), vp]
    lib.ah_plan_time_import.argtypes = [vp, P(AhTimeState), vp]
    lib.ah_plan_init_ic.argtypes = [vp, i, P(AhIc), vp]
    lib.ah_plan_snapshot.argtypes = [vp, i, vp, vp]
    lib.ah_plan_set_capture.argtypes = [vp, i, i, vp, i, vp, vp, vp]
    lib.ah_plan_profile_read.argtypes = [vp, P(d), P(i)]
    for name in EXPORTS:
        if name != "ah_status_str":
            getattr(lib, name).restype = i
    _lib = lib
    return lib


def check(rc, what=""):
    if rc != AH_OK:
        msg = load().ah_status_str(rc).decode()
        raise AstreaHipError(f"{what}: {msg} ({rc})")
    return rc
