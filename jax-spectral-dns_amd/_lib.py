"""ctypes binding of libjsd.so (declarations in include/jsd.h).

The library is built in-tree (``lib/libjsd.so``) by ``build.py``; nothing
here compiles or falls back to another implementation: if the shared library
is missing or no GPU is visible the calls raise.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional

_HERE = os.path.dirname(os.path.realpath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libjsd.so")

KIND_NS, KIND_PERTURBATION, KIND_DUAL = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("nxm", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nzm", ctypes.c_int),
        ("px", ctypes.c_int),
        ("pz", ctypes.c_int),
        ("ppx", ctypes.c_int),
        ("ppz", ctypes.c_int),
        ("lx", ctypes.c_double),
        ("lz", ctypes.c_double),
        ("re_tau", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("kind", ctypes.c_int),
        ("linearise", ctypes.c_int),
        ("D", _dp),
        ("kx", _dp),
        ("kz", _dp),
        ("lhs_inv", _dp),
        ("poisson", _dp),
        ("v_a", _dp),
        ("infl_inv", _dp),
        ("rhs_base", _dp),
        ("rhs_ns", _dp),
        ("lhs_inv_ns", _dp),
        ("base_u", _dp),
        ("base_du", _dp),
        ("y_weights", _dp),
        ("adjoint", ctypes.c_int),
        ("y_eig", _dp),
        ("no_coupling_term", ctypes.c_int),
    ]


class JsdError(RuntimeError):
    pass


_lib: Optional[ctypes.CDLL] = None

# every symbol include/jsd.h declares
EXPORTS = (
    "jsd_plan_create",
    "jsd_plan_destroy",
    "jsd_nonlinear",
    "jsd_step",
    "jsd_dual_step",
    "jsd_energy",
    "jsd_ydiff",
    "jsd_step_vjp",
    "jsd_step_vjp_state",
    "jsd_step_record_bytes",
    "jsd_step_record",
    "jsd_step_vjp_recorded",
    "jsd_energy_grad",
    "jsd_profile",
    "jsd_set_base_flow",
    "jsd_set_base_flow_device",
    "jsd_step_source",
    "jsd_substep_terms",
    "jsd_laplacian",
    "jsd_dissipation_accumulate",
    "jsd_volume_integral",
    "jsd_energy_noalloc",
    "jsd_batch_create",
    "jsd_batch_destroy",
    "jsd_batch_step",
    "jsd_batch_step_vjp",
    "jsd_batch_step_vjp_state",
    "jsd_batch_info",
    "jsd_batch_tiles",
    "jsd_last_error",
    "jsd_version",
)


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise JsdError("libjsd.so not built (expected %s); run __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    lib.jsd_plan_create.argtypes = [ctypes.POINTER(PlanDesc), ctypes.POINTER(vp)]
    lib.jsd_plan_destroy.argtypes = [vp]
    lib.jsd_nonlinear.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    lib.jsd_step.argtypes = [vp, vp, ctypes.c_double, ctypes.c_int, vp]
    lib.jsd_dual_step.argtypes = [vp, vp, vp, vp]
    lib.jsd_energy.argtypes = [vp, vp, _dp, vp]
    lib.jsd_ydiff.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_int, vp]
    lib.jsd_step_vjp.argtypes = [vp, vp, ctypes.c_int, vp]
    lib.jsd_step_vjp_state.argtypes = [vp, vp, vp, ctypes.c_double, vp]
    lib.jsd_step_record_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
    lib.jsd_step_record.argtypes = [vp, vp, vp, ctypes.c_double, vp]
    lib.jsd_step_vjp_recorded.argtypes = [vp, vp, vp, vp, ctypes.c_double, vp]
    lib.jsd_energy_grad.argtypes = [vp, vp, vp, ctypes.c_double, vp]
    lib.jsd_profile.argtypes = [vp, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_long)]
    lib.jsd_set_base_flow.argtypes = [vp, _dp, _dp, vp]
    lib.jsd_set_base_flow_device.argtypes = [vp, vp, vp, vp]
    lib.jsd_step_source.argtypes = [vp, vp, vp, ctypes.c_double, vp, ctypes.c_int, vp]
    lib.jsd_substep_terms.argtypes = [vp, vp, ctypes.c_int, ctypes.c_double, vp, vp, vp, vp, ctypes.c_int, vp]
    lib.jsd_laplacian.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_double, vp]
    lib.jsd_dissipation_accumulate.argtypes = [vp, vp, vp, ctypes.c_double, vp]
    lib.jsd_volume_integral.argtypes = [vp, vp, _dp, vp]
    lib.jsd_energy_noalloc.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_double, vp, vp, vp]
    lib.jsd_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.c_int, vp,
                                     ctypes.POINTER(vp)]
    lib.jsd_batch_destroy.argtypes = [vp]
    lib.jsd_batch_step.argtypes = [vp, ctypes.c_int, vp]
    lib.jsd_batch_step_vjp.argtypes = [vp, ctypes.c_int, vp]
    lib.jsd_batch_step_vjp_state.argtypes = [vp, ctypes.POINTER(vp), vp]
    lib.jsd_batch_info.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.jsd_batch_tiles.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.jsd_last_error.restype = ctypes.c_char_p
    lib.jsd_version.restype = ctypes.c_int
    for name in EXPORTS:
        getattr(lib, name).restype = getattr(lib, name).restype or ctypes.c_int
    lib.jsd_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().jsd_last_error().decode(errors="replace")
        raise JsdError("libjsd error %d: %s" % (rc, msg))
