"""ctypes binding of ``libdiffskill_mpm.so`` (the C ABI in include/diffskill_mpm.h).

The shared library is built in-tree by ``diffskill_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdiffskill_mpm.so")

MPM_SPHERE, MPM_CAPSULE, MPM_ROLLINGPIN, MPM_ROLLINGPIN_EXT, MPM_BOX, MPM_GRIPPER, \
    MPM_PRISM, MPM_KNIFE, MPM_CHOPSTICKS, MPM_GRIPPER2, MPM_CYLINDER, MPM_TORUS = range(12)

SHAPE_CODES = {
    "Sphere": MPM_SPHERE, "Capsule": MPM_CAPSULE, "RollingPin": MPM_ROLLINGPIN,
    "RollingPinExt": MPM_ROLLINGPIN_EXT, "Box": MPM_BOX, "Gripper": MPM_GRIPPER,
    "Prism": MPM_PRISM, "Knife": MPM_KNIFE, "Chopsticks": MPM_CHOPSTICKS, "Gripper2": MPM_GRIPPER2,
    "Cylinder": MPM_CYLINDER, "Torus": MPM_TORUS,
}


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("n_grid", ctypes.c_int), ("n_particles", ctypes.c_int), ("n_prims", ctypes.c_int), ("pad0", ctypes.c_int),
        ("dx", ctypes.c_double), ("inv_dx", ctypes.c_double), ("dt", ctypes.c_double),
        ("p_vol", ctypes.c_double), ("p_mass", ctypes.c_double),
        ("mu", ctypes.c_double), ("lam", ctypes.c_double), ("yield_stress", ctypes.c_double),
        ("gravity", ctypes.c_double * 3), ("ground_friction", ctypes.c_double), ("lower_bound", ctypes.c_double),
    ]


class PrimDesc(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int), ("action_dim", ctypes.c_int),
        ("friction", ctypes.c_double), ("softness", ctypes.c_double),
        ("size", ctypes.c_double * 3), ("h", ctypes.c_double), ("r", ctypes.c_double),
        ("ph", ctypes.c_double * 2), ("prot", ctypes.c_double * 4),
        ("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3),
        ("min_gap", ctypes.c_double), ("max_gap", ctypes.c_double),
        ("action_scale", ctypes.c_double * 7),
    ]


EXPORTS = [
    "mpm_workspace_bytes", "mpm_create", "mpm_destroy", "mpm_frame", "mpm_stride", "mpm_particle_offset",
    "mpm_adjoint", "mpm_prim_buffers", "mpm_set_frame", "mpm_get_frame", "mpm_reset_frame0",
    "mpm_set_prim_frame", "mpm_get_prim_frame", "mpm_set_actions", "mpm_substep_fwd", "mpm_substep_bwd",
    "mpm_grid_mass", "mpm_grid_mass_bwd", "mpm_density_loss_bwd", "mpm_add_obs_grad", "mpm_action_grad", "mpm_clear_grads",
    "mpm_update_prims", "mpm_copy_frame", "mpm_set_profiling", "mpm_profile_read", "mpm_last_error",
    "mpm_set_launch", "mpm_get_launch", "mpm_min_dist_cols", "mpm_min_dist", "mpm_min_dist_bwd",
    "mpm_collision_bytes", "mpm_set_collisions", "mpm_get_collisions", "mpm_set_deterministic", "mpm_get_deterministic",
    "mpm_resort_bytes", "mpm_set_resort", "mpm_get_resort", "mpm_slot_order",
    "pc_nearest", "pc_nearest_noalloc", "pc_last_error", "pc_sinkhorn_work_bytes", "pc_softmin", "pc_plan_grad",
    "pc_sinkhorn_step_work_bytes", "pc_sinkhorn_step",
]

PHASES = ["bin", "compact", "blocks", "scatter", "stress", "p2g", "grid_op", "g2p", "p2g_grid", "grid_op_keep",
          "g2p_bwd", "grid_op_bwd", "p2g_bwd", "stress_bwd", "collide", "collide_bwd",
          "prim_reduce"]

_LIB = None


def load():
    """Load the HIP library (raises when it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libdiffskill_mpm.so is not built; run diffskill_amd.build.build() "
                           "(hipcc --offload-arch=gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, i32, i64 = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int64
    fpp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "mpm_workspace_bytes": (i64, [vp, i32, i32, i32, i32]),
        "mpm_create": (i32, [vp, i32, vp, i32, i32, i32, vp, i64, vp, ctypes.POINTER(vp)]),
        "mpm_destroy": (i32, [vp]),
        "mpm_frame": (i32, [vp, i32, vp, fpp]),
        "mpm_stride": (i32, [vp]),
        "mpm_particle_offset": (i32, [vp, i32]),
        "mpm_adjoint": (i32, [vp, i32, fpp]),
        "mpm_prim_buffers": (i32, [vp, fpp, fpp, fpp, fpp]),
        "mpm_set_frame": (i32, [vp, i32, vp, vp]),
        "mpm_get_frame": (i32, [vp, i32, vp, vp]),
        "mpm_reset_frame0": (i32, [vp, vp, vp]),
        "mpm_set_prim_frame": (i32, [vp, i32, vp, vp]),
        "mpm_get_prim_frame": (i32, [vp, i32, vp, vp]),
        "mpm_set_actions": (i32, [vp, i32, i32, vp, vp]),
        "mpm_substep_fwd": (i32, [vp, i32, i32, vp]),
        "mpm_substep_bwd": (i32, [vp, i32, i32, vp]),
        "mpm_grid_mass": (i32, [vp, i32, vp, vp]),
        "mpm_grid_mass_bwd": (i32, [vp, i32, vp, vp]),
        "mpm_density_loss_bwd": (i32, [vp, i32, vp, vp, ctypes.c_float, vp]),
        "mpm_add_obs_grad": (i32, [vp, i32, vp, vp, vp]),
        "mpm_action_grad": (i32, [vp, i32, i32, vp, vp]),
        "mpm_clear_grads": (i32, [vp, i32, vp]),
        "mpm_update_prims": (i32, [vp, vp, vp]),
        "mpm_copy_frame": (i32, [vp, i32, i32, vp]),
        "mpm_set_profiling": (i32, [vp, i32]),
        "mpm_set_launch": (i32, [vp, i32, i32, i32]),
        "mpm_min_dist_cols": (i32, [vp, i32]),
        "mpm_min_dist": (i32, [vp, i32, vp, i32, vp]),
        "mpm_min_dist_bwd": (i32, [vp, i32, vp, i32, vp]),
        "mpm_get_launch": (i32, [vp, vp, vp, vp]),
        "mpm_collision_bytes": (i64, [vp, i32, i32]),
        "mpm_set_collisions": (i32, [vp, i32, vp, i32, vp, vp, i64, vp]),
        "mpm_get_collisions": (i32, [vp, i32, vp, vp, vp, vp]),
        "mpm_profile_read": (i32, [vp, vp, vp, i32]),
        "mpm_set_deterministic": (i32, [vp, i32]),
        "mpm_get_deterministic": (i32, [vp]),
        "mpm_resort_bytes": (i64, [vp, i32]),
        "mpm_set_resort": (i32, [vp, i32, vp, i64]),
        "mpm_get_resort": (i32, [vp]),
        "mpm_slot_order": (i32, [vp, i32, vp, vp]),
        "mpm_last_error": (ctypes.c_char_p, []),
        "pc_nearest": (i32, [vp, i32, vp, i32, vp, vp, vp, vp]),
        "pc_nearest_noalloc": (i32, [vp, i32, vp, i32, vp, vp, vp]),
        "pc_sinkhorn_work_bytes": (ctypes.c_longlong, [i32, i32]),
        "pc_softmin": (i32, [vp, i32, vp, i32, vp, ctypes.c_float, i32, vp, vp, vp]),
        "pc_plan_grad": (i32, [vp, i32, vp, i32, vp, vp, ctypes.c_float, i32, vp, vp, vp]),
        "pc_sinkhorn_step_work_bytes": (ctypes.c_longlong, [i32, i32]),
        "pc_sinkhorn_step": (i32, [vp, i32, vp, i32, vp, ctypes.c_float, i32, i32, vp, vp]),
        "pc_last_error": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def check(rc, what=""):
    """Raise RuntimeError with the library's message when rc != 0."""
    if rc != 0:
        msg = load().mpm_last_error().decode(errors="replace")
        raise RuntimeError("%s failed: %s" % (what or "mpm call", msg))
    return rc
