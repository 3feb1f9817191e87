"""MI355X-native explicit finite-volume update of ForestFV.

The product is C host code (octree, AMR, packing: ``host/``) and gfx950 HIP
kernels behind the C ABI of ``include/forestfv_dev.h`` (``csrc/``), built
into ``libforestfv.so``.  This module is a thin ctypes layer over that
library for tests and benchmarks; it adds no computation of its own.

The entry points mirror the reference's per-RUN driver
(src/c/advancetimeexplicit.c:3-29):

    mesh  = Mesh.uniform((1, 1, 1), (0, 0, 0), 1.0, level=6, bc=[4]*6)
    dev   = Device(params, mesh)          # fv_dev_attach
    dev.set_state(S)                      # LEAF.S, [n][NEQ]
    dev.advance(100)                      # advancetimeexplicit x 100
    S     = dev.get_state()               # fv_dev_get_state
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libforestfv.so")

FV_MAXMAT = 4
FV_MAXLEVEL = 10

FV_OK, FV_ERR_NAN, FV_ERR_ARG, FV_ERR_HIP, FV_ERR_MESH, FV_ERR_NOMEM = 0, -1, -2, -3, -4, -5
FK_BOUNDARY, FK_SAME, FK_MIRROR, FK_COARSE, FK_FINE = 0, 1, 2, 3, 4
RK_REF2, RK_SSP3, RK_SSP2 = 2, 3, 22
PREC_FP64, PREC_MIXED, PREC_EXACT = 0, 1, 2
FIX_NONCONS, FIX_ENERGY, FIX_ALL = 1, 2, 3


class FVError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"forestfv status {code}: {msg}")
        self.code = code


class fv_params(C.Structure):
    _fields_ = [
        ("nmat", C.c_int),
        ("gamma", C.c_double * FV_MAXMAT),
        ("pinf", C.c_double * FV_MAXMAT),
        ("amin", C.c_double),
        ("avf_lim", C.c_double),
        ("prelax", C.c_int),
        ("rk", C.c_int),
        ("limiter", C.c_int),
        ("dt", C.c_double),
        ("bc", C.c_int * 6),
        ("fix", C.c_int),
        ("primtv", C.c_int),
        ("bc_fixed_p", C.c_int),
        ("p_inlet", C.c_double),
        ("p_outlet", C.c_double),
    ]


class fv_crit(C.Structure):
    _fields_ = [
        ("uniform_on", C.c_int), ("uniform_lvl", C.c_int),
        ("prs_on", C.c_int), ("prs_lvl", C.c_int), ("prs_thr", C.c_double),
        ("rho_on", C.c_int), ("rho_lvl", C.c_int), ("rho_thr", C.c_double),
        ("vel_on", C.c_int), ("vel_lvl", C.c_int), ("vel_thr", C.c_double),
        ("inter0_on", C.c_int), ("inter0_lvl", C.c_int), ("inter0_lrs", C.c_int),
        ("inter1_on", C.c_int), ("inter1_lvl", C.c_int), ("inter1_lrs", C.c_int),
    ]


class fv_mesh(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("nroot", C.c_int * 3),
        ("origin", C.c_double * 3),
        ("rootsize", C.c_double),
        ("maxlevel", C.c_int),
        ("root", C.POINTER(C.c_int)),
        ("level", C.POINTER(C.c_int)),
        ("ci", C.POINTER(C.c_int)),
        ("tag", C.POINTER(C.c_int)),
        ("lfc_neigs", C.POINTER(C.c_int)),
        ("neigtr", C.POINTER(C.c_int)),
        ("bc", C.POINTER(C.c_int)),
        ("SI", C.POINTER(C.c_int)),
        ("bcside", C.c_int * 6),
        ("hkey", C.c_void_p),
        ("hval", C.c_void_p),
        ("hcap", C.c_size_t),
    ]


_L2 = FV_MAXLEVEL + 2


class fv_geo(C.Structure):
    _fields_ = [
        ("frame", (C.c_double * 9) * 6),
        ("h", C.c_double * _L2),
        ("area", C.c_double * _L2),
        ("vol", C.c_double * _L2),
        ("dcf", (C.c_double * _L2) * 3),
        ("dcc", (C.c_double * _L2) * 3),
    ]


class fv_pack(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("nbr", C.POINTER(C.c_int)),
        ("info", C.POINTER(C.c_uint32)),
        ("nfine", C.c_int),
        ("fine4", C.POINTER(C.c_int)),
        ("geo", fv_geo),
    ]


class fv_m5_params(C.Structure):
    _fields_ = [
        ("nmat", C.c_int),
        ("gamma", C.c_double * FV_MAXMAT),
        ("pinf", C.c_double * FV_MAXMAT),
        ("mush", C.c_double * FV_MAXMAT),
        ("amin", C.c_double),
        ("avf_lim", C.c_double),
        ("dt", C.c_double),
        ("zerou", C.c_int * 3),
        ("prelax", C.c_int),
        ("order", C.c_int),
        ("rk_steps", C.c_int),
        ("yield_", C.c_double),
        ("fix", C.c_int),
        ("limiter", C.c_int),
    ]


_lib = None

# Device contexts still open (Device, M5Device, MLP).  They are closed by an
# atexit hook registered when the library is loaded: Python runs it before the
# interpreter and the C runtime tear down, so every fv_*_detach / hipFree runs
# while the HIP runtime (and a profiler's exit handlers) are still alive.  After
# the hook has run, __del__ does nothing.
_live = weakref.WeakSet()
_shutdown = False
_atexit_registered = False


def _close_all():
    """Close every open device context; registered with atexit by lib()."""
    global _shutdown
    for obj in list(_live):
        try:
            obj.close()
        except Exception:
            pass
    _shutdown = True


def _track(obj):
    _live.add(obj)


def lib():
    """The product library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FVError(FV_ERR_ARG, f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        sig = {
            "fv_mesh_uniform": (P(fv_mesh), [P(C.c_int), P(C.c_double), C.c_double, C.c_int, P(C.c_int)]),
            "fv_mesh_from_leaves": (P(fv_mesh), [P(C.c_int), P(C.c_double), C.c_double, P(C.c_int), C.c_int,
                                                 P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int)]),
            "fv_mesh_free": (None, [P(fv_mesh)]),
            "fv_mesh_find": (C.c_int, [P(fv_mesh), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
            "fv_mesh_centre": (None, [P(fv_mesh), C.c_int, P(C.c_double)]),
            "fv_tag_from_address": (C.c_int, [C.c_int, P(C.c_int)]),
            "fv_address_from_tag": (C.c_int, [C.c_int, P(C.c_int), P(C.c_int)]),
            "fv_mesh_pack": (C.c_int, [P(fv_mesh), P(fv_pack)]),
            "fv_pack_free": (None, [P(fv_pack)]),
            "fv_geo_build": (None, [C.c_double, P(fv_geo)]),
            "fv_criterionsmth": (C.c_int, [P(fv_mesh), C.c_int, P(C.c_int)]),
            "fv_mesh_adapt": (C.c_int, [P(P(fv_mesh)), P(fv_params), P(C.c_int), P(P(C.c_double))]),
            "fv_neq": (C.c_int, [C.c_int]),
            "fv_relax_host": (None, [P(fv_params), P(C.c_double)]),
            "fv_init_pure": (None, [P(fv_params), P(C.c_double), C.c_int, C.c_double, P(C.c_double), P(C.c_double)]),
            "fv_init_split_x": (C.c_int, [P(fv_mesh), P(fv_params), P(C.c_double), P(C.c_double), C.c_double,
                                          P(C.c_double)]),
            "fv_init_bubble": (C.c_int, [P(fv_mesh), P(fv_params), P(C.c_double), P(C.c_double), P(C.c_double),
                                         C.c_double, C.c_double, C.c_uint, C.c_double, P(C.c_double)]),
            "fv_dev_attach": (C.c_int, [P(C.c_void_p), P(fv_params), P(fv_pack)]),
            "fv_dev_reattach": (C.c_int, [C.c_void_p, P(fv_pack)]),
            "fv_dev_detach": (C.c_int, [C.c_void_p]),
            "fv_dev_set_state": (C.c_int, [C.c_void_p, P(C.c_double)]),
            "fv_dev_get_state": (C.c_int, [C.c_void_p, P(C.c_double)]),
            "fv_dev_get_faces": (C.c_int, [C.c_void_p, P(C.c_double)]),
            "fv_dev_advance": (C.c_int, [C.c_void_p, C.c_int]),
            "fv_dev_begin_step": (C.c_int, [C.c_void_p]),
            "fv_dev_stage": (C.c_int, [C.c_void_p, C.c_int]),
            "fv_dev_set_dt": (C.c_int, [C.c_void_p, C.c_double]),
            "fv_dev_cfl_dt": (C.c_int, [C.c_void_p, C.c_double, P(C.c_double)]),
            "fv_dev_set_precision": (C.c_int, [C.c_void_p, C.c_int]),
            "fv_dev_set_fused": (C.c_int, [C.c_void_p, C.c_int]),
            "fv_dev_indicator": (C.c_int, [C.c_void_p, P(fv_crit), P(C.c_int), P(C.c_int)]),
            "fv_dev_interface_location": (C.c_int, [C.c_void_p, P(fv_crit), P(C.c_int)]),
            "fv_dev_time_stages": (C.c_int, [C.c_void_p, C.c_int, P(C.c_double), P(C.c_double)]),
            "fv_dev_size": (C.c_int, [C.c_void_p, P(C.c_int), P(C.c_int)]),
            "fv_dev_last_error": (C.c_char_p, []),
            "fv_dev_selftest_guard": (C.c_int, []),
            "fv_dev_set_bricks": (C.c_int, [C.c_void_p, C.c_int]),
            "fv_dev_brick_count": (C.c_int, [C.c_void_p, P(C.c_int), P(C.c_int)]),
            "fv_dev_brick_kinds": (C.c_int, [C.c_void_p, P(C.c_int), P(C.c_int)]),
            "fv_cfl_dt_noalloc": (C.c_int, [P(fv_params), P(fv_geo), C.c_int, C.c_void_p, C.c_void_p, C.c_double,
                                            C.c_void_p, C.c_void_p]),
            "fv_dev_stream_copy": (C.c_int, [C.c_size_t, C.c_int, P(C.c_double)]),
            "fv_dev_set_limiter_flags": (C.c_int, [C.c_void_p, P(C.c_uint8)]),
            "fv_dev_get_limiter_flags": (C.c_int, [C.c_void_p, P(C.c_uint8)]),
            "fv_mlp_create": (C.c_int, [P(C.c_void_p), C.c_uint]),
            "fv_mlp_destroy": (C.c_int, [C.c_void_p]),
            "fv_mlp_num_weights": (C.c_int, []),
            "fv_mlp_get_weights": (C.c_int, [C.c_void_p, P(C.c_float)]),
            "fv_mlp_set_weights": (C.c_int, [C.c_void_p, P(C.c_float)]),
            "fv_mlp_set_options": (C.c_int, [C.c_void_p, P(C.c_float), C.c_float]),
            "fv_mlp_sample": (C.c_int, [C.c_void_p, C.c_void_p, P(fv_crit)]),
            "fv_mlp_dataset": (C.c_int, [C.c_void_p, P(C.c_long), P(C.c_float), P(C.c_int), P(C.c_float)]),
            "fv_mlp_clear_dataset": (C.c_int, [C.c_void_p]),
            "fv_mlp_train": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_uint, P(C.c_float)]),
            "fv_mlp_grad": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_int), P(C.c_float), C.c_int, P(C.c_float),
                                      P(C.c_float)]),
            "fv_mlp_forward": (C.c_int, [C.c_void_p, P(C.c_float), C.c_int, P(C.c_float)]),
            "fv_mlp_infer": (C.c_int, [C.c_void_p, P(C.c_float), C.c_int, P(C.c_float)]),
            "fv_dev_indicator_mlp": (C.c_int, [C.c_void_p, C.c_void_p, P(fv_crit), P(C.c_int), P(C.c_int),
                                               P(C.c_uint8)]),
            "fv_dev_totals": (C.c_int, [C.c_void_p, P(C.c_double)]),
            "fv_dev_set_forest": (C.c_int, [C.c_void_p, P(fv_mesh)]),
            "fv_dev_get_forest": (C.c_int, [C.c_void_p, P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int)]),
            "fv_dev_get_pack": (C.c_int, [C.c_void_p, P(C.c_int), P(C.c_uint32), P(C.c_int), P(C.c_int)]),
            "fv_dev_criterionsmth": (C.c_int, [C.c_void_p, C.c_int, P(C.c_int)]),
            "fv_dev_regrid": (C.c_int, [C.c_void_p, P(fv_crit), C.c_int, P(C.c_int)]),
            "fv_restart_write": (C.c_int, [C.c_char_p, P(fv_mesh), P(C.c_double), C.c_int, C.c_int, C.c_double,
                                           C.c_int]),
            "fv_restart_read": (C.c_int, [C.c_char_p, P(C.c_int), P(C.c_double), C.c_double, P(C.c_int),
                                          P(P(fv_mesh)), P(P(C.c_double)), P(C.c_int), P(C.c_int), P(C.c_double)]),
            "fv_dev_get_ntrees": (C.c_int, [C.c_void_p, P(C.c_int)]),
            "fv_dev_write_restart": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int, C.c_double, C.c_int]),
            "fv_m5_neq": (C.c_int, [C.c_int]),
            "fv_m5_gfm_faces": (C.c_int, [P(fv_m5_params), C.c_int, P(C.c_double), P(C.c_double), P(C.c_double),
                                          P(C.c_double), P(C.c_double), P(C.c_double)]),
            "fv_m5_rk": (C.c_int, [P(fv_m5_params), P(fv_pack), C.c_int, P(C.c_double), P(C.c_double),
                                   P(C.c_double), P(C.c_uint8)]),
            "fv_m5_gfm_rel": (C.c_int, [P(fv_m5_params), P(fv_pack), P(C.c_double), P(C.c_double), P(C.c_uint8)]),
            "fv_m5_dev_attach": (C.c_int, [P(fv_m5_params), P(fv_pack), P(C.c_void_p)]),
            "fv_m5_dev_detach": (C.c_int, [C.c_void_p]),
            "fv_m5_dev_set_state": (C.c_int, [C.c_void_p, P(C.c_double)]),
            "fv_m5_dev_get_state": (C.c_int, [C.c_void_p, P(C.c_double)]),
            "fv_m5_dev_begin_step": (C.c_int, [C.c_void_p]),
            "fv_m5_dev_stage": (C.c_int, [C.c_void_p, C.c_int]),
            "fv_m5_dev_advance": (C.c_int, [C.c_void_p, C.c_int]),
            "fv_m5_dev_faceinterp": (C.c_int, [C.c_void_p]),
            "fv_m5_dev_residual": (C.c_int, [C.c_void_p]),
            "fv_m5_dev_get_stage_arrays": (C.c_int, [C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_uint8)]),
            "fv_m5_dev_time_steps": (C.c_int, [C.c_void_p, C.c_int, P(C.c_double), P(C.c_double)]),
            "fv_m5_dev_set_precision": (C.c_int, [C.c_void_p, C.c_int]),
            "fv_m5_dev_set_dt": (C.c_int, [C.c_void_p, C.c_double]),
            "fv_m5_dev_set_limiter_flags": (C.c_int, [C.c_void_p, P(C.c_uint8)]),
            "fv_m5_dev_get_limiter_flags": (C.c_int, [C.c_void_p, P(C.c_uint8)]),
            "fv_m5_dev_set_forest": (C.c_int, [C.c_void_p, P(fv_mesh)]),
            "fv_m5_dev_get_forest": (C.c_int, [C.c_void_p, P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int)]),
            "fv_m5_dev_interface_location": (C.c_int, [C.c_void_p, P(fv_crit), P(C.c_int)]),
            "fv_m5_dev_indicator": (C.c_int, [C.c_void_p, P(fv_crit), P(C.c_int), P(C.c_int)]),
            "fv_m5_dev_regrid": (C.c_int, [C.c_void_p, P(fv_crit), C.c_int, P(C.c_int)]),
            "fv_mlp_sample_m5": (C.c_int, [C.c_void_p, C.c_void_p, P(fv_crit)]),
            "fv_m5_dev_indicator_mlp": (C.c_int, [C.c_void_p, C.c_void_p, P(fv_crit), P(C.c_int), P(C.c_int),
                                                  P(C.c_uint8)]),
            "fv_m5_dev_regrid_mlp": (C.c_int, [C.c_void_p, C.c_void_p, P(fv_crit), C.c_int, P(C.c_int)]),
            "fv_m5_dev_cfl_dt": (C.c_int, [C.c_void_p, C.c_double, P(C.c_double)]),
            "fv_m5_init_slab_bubble": (C.c_int, [P(fv_mesh), P(fv_m5_params), P(C.c_double), C.c_double,
                                                 P(C.c_double), C.c_double, C.c_double, C.c_double, C.c_double,
                                                 P(C.c_double)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
        global _atexit_registered
        if not _atexit_registered:
            atexit.register(_close_all)
            _atexit_registered = True
    return _lib


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(rc, what):
    if rc != FV_OK:
        msg = lib().fv_dev_last_error()
        raise FVError(rc, f"{what}: {msg.decode() if msg else ''}")


def neq(nmat):
    return 3 * nmat + 3


def make_params(nmat=2, gamma=(1.4, 1.4), pinf=(0.0, 0.0), amin=1e-8, avf_lim=0.5, prelax=0,
                rk=RK_SSP3, limiter=0, dt=1e-3, bc=(4, 4, 4, 4, 4, 4), fix=FIX_ALL, primtv=0, fixed_p=None):
    """fixed_p = (p_inlet, p_outlet): fixed-pressure ghost states at inlet (code 3) and outlet
    (code 2) faces, as the reference's CASE 3400/3410 with (150e6, 1e6); None: plain copies."""
    p = fv_params()
    p.nmat = nmat
    for k in range(nmat):
        p.gamma[k] = gamma[k]
        p.pinf[k] = pinf[k]
    p.amin, p.avf_lim, p.prelax, p.rk, p.limiter, p.dt = amin, avf_lim, prelax, rk, limiter, dt
    for k in range(6):
        p.bc[k] = bc[k]
    p.fix = fix
    p.primtv = primtv
    if fixed_p is not None:
        p.bc_fixed_p = 1
        p.p_inlet, p.p_outlet = fixed_p
    return p


def restart_write(path, mesh, S, istep=0, time=0.0, binary=False):
    """restart_save / export_ultra_fast layout (fv_restart_write); binary: lossless."""
    S = np.ascontiguousarray(S, dtype=np.float64)
    assert S.shape[0] == mesh.n
    _check(lib().fv_restart_write(os.fsencode(path), mesh._p, _dp(S), S.shape[1], istep, time, int(binary)),
           "fv_restart_write")


def restart_read(path, nroot, origin, rootsize, bc):
    """restart_load (fv_restart_read), either format: (Mesh, S, istep, time)."""
    pm = C.POINTER(fv_mesh)()
    ps = C.POINTER(C.c_double)()
    nq, st, t = C.c_int(), C.c_int(), C.c_double()
    nr = (C.c_int * 3)(*nroot)
    og = (C.c_double * 3)(*origin)
    b = (C.c_int * 6)(*bc)
    _check(lib().fv_restart_read(os.fsencode(path), nr, og, rootsize, b, C.byref(pm), C.byref(ps), C.byref(nq),
                                 C.byref(st), C.byref(t)), "fv_restart_read")
    m = Mesh(pm)
    S = np.ctypeslib.as_array(ps, shape=(m.n * nq.value,)).reshape(m.n, nq.value).copy()
    libc_free(ps)
    return m, S, st.value, t.value


def m5_neq(nmat):
    return 3 * nmat + 12


# fv_m5_params.fix: the consistent model-5 update (FV_M5_FIX_SIGNS of forestfv_m5.h)
M5_FIX_SIGNS = 1
# fv_m5_params.fix bit: the volume-fraction-weighted elastic mixture (FV_M5_FIX_MIXTURE)
M5_FIX_MIXTURE = 2


def make_m5_params(nmat=3, gamma=(4.4, 6.12, 1.4), pinf=(6e8, 3.43e8, 0.0), mush=(2.6e10, 0.0, 0.0), amin=1e-6,
                   avf_lim=0.5, dt=1e-9, zerou=(0, 0, 0), prelax=1, order=2, rk_steps=3, yield_stress=0.0, fix=0,
                   limiter=0):
    """Model-5 material table (MATERGAMA, MATERPINF, MATERMUSH) and stage constants."""
    p = fv_m5_params()
    p.nmat = nmat
    for k in range(nmat):
        p.gamma[k], p.pinf[k], p.mush[k] = gamma[k], pinf[k], mush[k]
    p.amin, p.avf_lim, p.dt = amin, avf_lim, dt
    for k in range(3):
        p.zerou[k] = zerou[k]
    p.prelax, p.order, p.rk_steps = prelax, order, rk_steps
    p.yield_ = yield_stress
    p.fix = fix
    p.limiter = limiter
    return p


def m5_gfm_faces(params, UL, UR, frame, area):
    """GFM flux of interface sub-faces (fv_m5_gfm_faces): returns flux [nf][NEQ], unc [nf][3]."""
    UL = np.ascontiguousarray(UL, dtype=np.float64)
    UR = np.ascontiguousarray(UR, dtype=np.float64)
    fr = np.ascontiguousarray(frame, dtype=np.float64)
    ar = np.ascontiguousarray(area, dtype=np.float64)
    nf = UL.shape[0]
    flux = np.zeros((nf, m5_neq(params.nmat)))
    unc = np.zeros((nf, 3))
    _check(lib().fv_m5_gfm_faces(C.byref(params), nf, _dp(UL), _dp(UR), _dp(fr), _dp(ar), _dp(flux), _dp(unc)),
           "fv_m5_gfm_faces")
    return flux, unc


def m5_rk(params, pack, irk, SN, S, RHS, gfm_face):
    """Model-5 RK stage (fv_m5_rk): returns (SN, S) after the stage."""
    SN = np.ascontiguousarray(SN, dtype=np.float64).copy()
    S = np.ascontiguousarray(S, dtype=np.float64).copy()
    R = np.ascontiguousarray(RHS, dtype=np.float64)
    gf = np.ascontiguousarray(gfm_face, dtype=np.uint8)
    _check(lib().fv_m5_rk(C.byref(params), C.byref(pack.pk), irk, _dp(SN), _dp(S), _dp(R),
                          gf.ctypes.data_as(C.POINTER(C.c_uint8))), "fv_m5_rk")
    return SN, S


def m5_gfm_rel(params, pack, S_before, S, gfm_face):
    """Re-projection after relaxation (fv_m5_gfm_rel): returns the new S."""
    Sb = np.ascontiguousarray(S_before, dtype=np.float64)
    S = np.ascontiguousarray(S, dtype=np.float64).copy()
    gf = np.ascontiguousarray(gfm_face, dtype=np.uint8)
    _check(lib().fv_m5_gfm_rel(C.byref(params), C.byref(pack.pk), _dp(Sb), _dp(S),
                               gf.ctypes.data_as(C.POINTER(C.c_uint8))), "fv_m5_gfm_rel")
    return S


class M5Device:
    """The whole model-5 stage on the device (fv_m5_dev_*): MUSCL face states, the residual with
    its HLLC and GFM fluxes, m5_RK and the relaxation, on the leaves and face kinds of `mesh`."""

    def __init__(self, params, mesh, prec=None):
        self.params = params
        self.pack = mesh.pack()
        self.n = mesh.n
        self.neq = m5_neq(params.nmat)
        h = C.c_void_p()
        _check(lib().fv_m5_dev_attach(C.byref(params), C.byref(self.pack.pk), C.byref(h)), "fv_m5_dev_attach")
        self.h = h
        _track(self)
        if prec is not None:
            self.set_precision(prec)

    def close(self):
        if getattr(self, "h", None):
            lib().fv_m5_dev_detach(self.h)
            self.h = None

    def __del__(self):
        if _lib is not None and not _shutdown:
            try:
                self.close()
            except Exception:
                pass

    def set_state(self, S):
        S = np.ascontiguousarray(S, dtype=np.float64)
        assert S.shape == (self.n, self.neq)
        _check(lib().fv_m5_dev_set_state(self.h, _dp(S)), "fv_m5_dev_set_state")

    def get_state(self):
        S = np.zeros((self.n, self.neq))
        _check(lib().fv_m5_dev_get_state(self.h, _dp(S)), "fv_m5_dev_get_state")
        return S

    def begin_step(self):
        _check(lib().fv_m5_dev_begin_step(self.h), "fv_m5_dev_begin_step")

    def set_precision(self, prec):
        """PREC_EXACT (the reference's operation order), PREC_FP64 (default: the non-conservative
        sources formed once per leaf) or PREC_MIXED (FP32 slopes, face records, HLLC and GFM flux)."""
        _check(lib().fv_m5_dev_set_precision(self.h, prec), "fv_m5_dev_set_precision")

    def stage(self, irk):
        _check(lib().fv_m5_dev_stage(self.h, irk), "fv_m5_dev_stage")

    def set_dt(self, dt):
        _check(lib().fv_m5_dev_set_dt(self.h, dt), "fv_m5_dev_set_dt")
        self.params.dt = dt

    def set_limiter_flags(self, lim):
        """Per-leaf slope flags of limiter 2 (0 minmod, 1 MC)."""
        a = np.ascontiguousarray(lim, dtype=np.uint8)
        assert a.shape == (self.n,)
        _check(lib().fv_m5_dev_set_limiter_flags(self.h, a.ctypes.data_as(C.POINTER(C.c_uint8))),
               "fv_m5_dev_set_limiter_flags")

    def get_limiter_flags(self):
        a = np.zeros(self.n, np.uint8)
        _check(lib().fv_m5_dev_get_limiter_flags(self.h, a.ctypes.data_as(C.POINTER(C.c_uint8))),
               "fv_m5_dev_get_limiter_flags")
        return a

    def set_forest(self, mesh):
        """Leaf addresses of the attached mesh, for interface_location and regrids on the device."""
        _check(lib().fv_m5_dev_set_forest(self.h, mesh._p), "fv_m5_dev_set_forest")

    def get_forest(self):
        """(root [n], level [n], ci [n][3], SI [n][2]) of the device mesh."""
        root = np.zeros(self.n, np.int32)
        level = np.zeros(self.n, np.int32)
        ci = np.zeros((self.n, 3), np.int32)
        SI = np.zeros((self.n, 2), np.int32)
        _check(lib().fv_m5_dev_get_forest(self.h, _ip(root), _ip(level), _ip(ci), _ip(SI)), "fv_m5_dev_get_forest")
        return root, level, ci, SI

    def interface_location(self, crit, copy=True):
        """interface_location after a step: SI of the device state, face kinds refreshed; SI [n][2]
        (copy=False: SI stays on the device, as the time loop needs it, and None is returned)."""
        SI = np.zeros((self.n, 2), np.int32) if copy else None
        _check(lib().fv_m5_dev_interface_location(self.h, C.byref(crit), _ip(SI) if copy else None),
               "fv_m5_dev_interface_location")
        return SI

    def indicator(self, crit):
        """(SI [n][2], target levels [n]) of the device state; the mesh is unchanged."""
        SI = np.zeros((self.n, 2), np.int32)
        split = np.zeros(self.n, np.int32)
        _check(lib().fv_m5_dev_indicator(self.h, C.byref(crit), _ip(SI), _ip(split)), "fv_m5_dev_indicator")
        return SI, split

    def regrid(self, crit, maxlevel):
        """meshadaptation on the device; returns the new leaf count."""
        n = C.c_int()
        _check(lib().fv_m5_dev_regrid(self.h, C.byref(crit), maxlevel, C.byref(n)), "fv_m5_dev_regrid")
        self.n = n.value
        return self.n

    def sample_mlp(self, mlp, crit):
        """Append this state's stencil features and plain-indicator labels to mlp's training set."""
        _check(lib().fv_mlp_sample_m5(self.h, mlp._h, C.byref(crit)), "fv_mlp_sample_m5")

    def indicator_mlp(self, mlp, crit):
        """(SI, targets, smooth flags) of the learned indicator; the flags become limiter 2's."""
        SI = np.zeros((self.n, 2), np.int32)
        split = np.zeros(self.n, np.int32)
        sm = np.zeros(self.n, np.uint8)
        _check(lib().fv_m5_dev_indicator_mlp(self.h, mlp._h, C.byref(crit), _ip(SI), _ip(split),
                                             sm.ctypes.data_as(C.POINTER(C.c_uint8))), "fv_m5_dev_indicator_mlp")
        return SI, split, sm

    def regrid_mlp(self, mlp, crit, maxlevel):
        """meshadaptation on the learned targets; returns the new leaf count."""
        n = C.c_int()
        _check(lib().fv_m5_dev_regrid_mlp(self.h, mlp._h, C.byref(crit), maxlevel, C.byref(n)), "fv_m5_dev_regrid_mlp")
        self.n = n.value
        return self.n

    def cfl_dt(self, cfl):
        """cfl * min h / (max|u_i| + c) over the device state, c the mixture's longitudinal speed."""
        d = C.c_double()
        _check(lib().fv_m5_dev_cfl_dt(self.h, cfl, C.byref(d)), "fv_m5_dev_cfl_dt")
        return d.value

    def advance(self, nsteps):
        _check(lib().fv_m5_dev_advance(self.h, nsteps), "fv_m5_dev_advance")

    def faceinterp(self):
        _check(lib().fv_m5_dev_faceinterp(self.h), "fv_m5_dev_faceinterp")

    def residual(self):
        _check(lib().fv_m5_dev_residual(self.h), "fv_m5_dev_residual")

    def stage_arrays(self):
        """(SF [n][6][NEQ], RHS [n][NEQ], gfm_face [n][6]) of the last operations."""
        SF = np.zeros((self.n, 6, self.neq))
        RHS = np.zeros((self.n, self.neq))
        gf = np.zeros((self.n, 6), np.uint8)
        _check(lib().fv_m5_dev_get_stage_arrays(self.h, _dp(SF), _dp(RHS), gf.ctypes.data_as(C.POINTER(C.c_uint8))),
               "fv_m5_dev_get_stage_arrays")
        return SF, RHS, gf

    def time_steps(self, nsteps):
        """(ms per stage, ms per residual kernel launch) over nsteps steps."""
        a, b = C.c_double(), C.c_double()
        _check(lib().fv_m5_dev_time_steps(self.h, nsteps, C.byref(a), C.byref(b)), "fv_m5_dev_time_steps")
        return a.value, b.value


def init_pure(params, rho_init, mat, pres, u=(0.0, 0.0, 0.0)):
    """Cell state of pure material `mat` at pressure `pres` (init_S.c:70-107)."""
    S = np.zeros(neq(params.nmat))
    r = np.ascontiguousarray(rho_init, dtype=np.float64)
    uu = np.ascontiguousarray(u, dtype=np.float64)
    lib().fv_init_pure(C.byref(params), _dp(r), mat, pres, _dp(uu), _dp(S))
    return S


def make_crit(**kw):
    c = fv_crit()
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class Mesh:
    """Forest of octrees on the host (fv_mesh); arrays are copied out as numpy."""

    def __init__(self, ptr):
        if not ptr:
            raise FVError(FV_ERR_MESH, "mesh construction failed")
        self._p = ptr

    @classmethod
    def uniform(cls, nroot, origin, rootsize, level, bc):
        nr = (C.c_int * 3)(*nroot)
        og = (C.c_double * 3)(*origin)
        b = (C.c_int * 6)(*bc)
        return cls(lib().fv_mesh_uniform(nr, og, rootsize, level, b))

    @classmethod
    def from_leaves(cls, nroot, origin, rootsize, bc, root, level, ci):
        root = np.ascontiguousarray(root, dtype=np.int32)
        level = np.ascontiguousarray(level, dtype=np.int32)
        ci = np.ascontiguousarray(ci, dtype=np.int32).reshape(-1, 3)
        st = C.c_int(0)
        nr = (C.c_int * 3)(*nroot)
        og = (C.c_double * 3)(*origin)
        b = (C.c_int * 6)(*bc)
        p = lib().fv_mesh_from_leaves(nr, og, rootsize, b, len(root), _ip(root), _ip(level), _ip(ci), C.byref(st))
        if not p:
            raise FVError(st.value, "fv_mesh_from_leaves")
        return cls(p)

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.fv_mesh_free(self._p)
            self._p = None

    @property
    def m(self):
        return self._p.contents

    @property
    def n(self):
        return self.m.n

    def _arr(self, name, shape):
        ptr = getattr(self.m, name)
        size = int(np.prod(shape))
        return np.ctypeslib.as_array(ptr, shape=(size,)).reshape(shape).copy()

    def arrays(self):
        n = self.n
        return dict(
            root=self._arr("root", (n,)), level=self._arr("level", (n,)), ci=self._arr("ci", (n, 3)),
            tag=self._arr("tag", (n,)), lfc_neigs=self._arr("lfc_neigs", (n, 6)),
            neigtr=self._arr("neigtr", (n, 6, 4)), bc=self._arr("bc", (n, 6)), SI=self._arr("SI", (n, 2)),
        )

    def set_SI(self, SI):
        SI = np.ascontiguousarray(SI, dtype=np.int32).reshape(-1)
        C.memmove(self.m.SI, SI.ctypes.data, SI.nbytes)

    def centres(self):
        out = np.zeros((self.n, 3))
        c = (C.c_double * 3)()
        for i in range(self.n):
            lib().fv_mesh_centre(self._p, i, c)
            out[i] = c[:]
        return out

    def find(self, root, level, i, j, k):
        return lib().fv_mesh_find(self._p, root, level, i, j, k)

    def criterionsmth(self, LEVEL, split):
        split = np.ascontiguousarray(split, dtype=np.int32).copy()
        _check(lib().fv_criterionsmth(self._p, LEVEL, _ip(split)), "fv_criterionsmth")
        return split

    def adapt(self, params, split, S):
        """Merge/split by target level (fv_mesh_adapt); the mesh is replaced in place and
        the remapped state [n][NEQ] is returned."""
        nq = neq(params.nmat)
        split = np.ascontiguousarray(split, dtype=np.int32)
        sp = C.cast(C.c_void_p(lib_malloc(S)), C.POINTER(C.c_double))
        pm = C.pointer(self._p)
        ps = C.pointer(sp)
        rc = lib().fv_mesh_adapt(pm, C.byref(params), _ip(split), ps)
        if rc != FV_OK:
            libc_free(ps.contents)
            _check(rc, "fv_mesh_adapt")
        self._p = pm.contents
        out = np.ctypeslib.as_array(ps.contents, shape=(self.n * nq,)).reshape(self.n, nq).copy()
        libc_free(ps.contents)
        return out

    def pack(self):
        return Pack(self)

    def init_split_x(self, params, rho_init, p_init, xs):
        S = np.zeros((self.n, neq(params.nmat)))
        r = np.ascontiguousarray(rho_init, dtype=np.float64)
        p = np.ascontiguousarray(p_init, dtype=np.float64)
        _check(lib().fv_init_split_x(self._p, C.byref(params), _dp(r), _dp(p), xs, _dp(S)), "fv_init_split_x")
        return S

    def init_m5_slab_bubble(self, params, rho_init, z_slab, centre, radius, p_liquid, p_gas, smear=2.0):
        """Model-5 state (materials solid, liquid, gas): a sharp solid slab below z_slab, a tanh
        gas bubble in the liquid, A = I in the solid and 0 in the fluids, at rest."""
        S = np.zeros((self.n, m5_neq(params.nmat)))
        r = np.ascontiguousarray(rho_init, dtype=np.float64)
        c = np.ascontiguousarray(centre, dtype=np.float64)
        _check(lib().fv_m5_init_slab_bubble(self._p, C.byref(params), _dp(r), z_slab, _dp(c), radius, p_liquid,
                                            p_gas, smear, _dp(S)), "fv_m5_init_slab_bubble")
        return S

    def init_bubble(self, params, rho_init, p_init, centre, radius, smear=2.0, seed=1234, perturb=0.0):
        S = np.zeros((self.n, neq(params.nmat)))
        r = np.ascontiguousarray(rho_init, dtype=np.float64)
        p = np.ascontiguousarray(p_init, dtype=np.float64)
        c = np.ascontiguousarray(centre, dtype=np.float64)
        _check(lib().fv_init_bubble(self._p, C.byref(params), _dp(r), _dp(p), _dp(c), radius, smear, seed,
                                    perturb, _dp(S)), "fv_init_bubble")
        return S


_libc = C.CDLL(None)
_libc.malloc.restype = C.c_void_p
_libc.malloc.argtypes = [C.c_size_t]
_libc.free.argtypes = [C.c_void_p]


def lib_malloc(S):
    S = np.ascontiguousarray(S, dtype=np.float64)
    p = _libc.malloc(max(S.nbytes, 8))
    C.memmove(p, S.ctypes.data, S.nbytes)
    return p


def libc_free(ptr):
    _libc.free(C.cast(ptr, C.c_void_p))


class Pack:
    """Device-bound image of a mesh (fv_pack)."""

    def __init__(self, mesh):
        self.pk = fv_pack()
        _check(lib().fv_mesh_pack(mesh._p, C.byref(self.pk)), "fv_mesh_pack")
        self.n = self.pk.n

    def __del__(self):
        if _lib is not None and getattr(self, "pk", None) is not None:
            _lib.fv_pack_free(C.byref(self.pk))
            self.pk = None

    def arrays(self):
        n, nf = self.pk.n, self.pk.nfine
        nbr = np.ctypeslib.as_array(self.pk.nbr, shape=(6 * n,)).reshape(6, n).copy()
        info = np.ctypeslib.as_array(self.pk.info, shape=(n,)).copy()
        fine4 = np.ctypeslib.as_array(self.pk.fine4, shape=(4 * max(nf, 1),)).reshape(-1, 4)[:nf].copy()
        return dict(nbr=nbr, info=info, fine4=fine4)


def stream_copy_gbps(nbytes=1 << 30, reps=20):
    """Measured device copy bandwidth (read + write GB/s)."""
    g = C.c_double()
    _check(lib().fv_dev_stream_copy(nbytes, reps, C.byref(g)), "fv_dev_stream_copy")
    return g.value


class Device:
    """One device context (fv_dev) over a packed mesh."""

    def __init__(self, params, mesh_or_pack, prec=None):
        pk = mesh_or_pack if isinstance(mesh_or_pack, Pack) else Pack(mesh_or_pack)
        self._pack = pk
        self.params = params
        self.nq = neq(params.nmat)
        self.n = pk.n
        h = C.c_void_p()
        _check(lib().fv_dev_attach(C.byref(h), C.byref(params), C.byref(pk.pk)), "fv_dev_attach")
        self._h = h
        _track(self)
        if prec is not None:
            self.set_precision(prec)

    def close(self):
        if getattr(self, "_h", None):
            lib().fv_dev_detach(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None and not _shutdown:
            self.close()

    def reattach(self, mesh):
        pk = Pack(mesh)
        _check(lib().fv_dev_reattach(self._h, C.byref(pk.pk)), "fv_dev_reattach")
        self._pack = pk
        self.n = pk.n

    def totals(self):
        """Volume integral of every state field (fv_dev_totals)."""
        t = np.zeros(self.nq)
        _check(lib().fv_dev_totals(self._h, _dp(t)), "fv_dev_totals")
        return t

    def write_restart(self, path, istep=0, time=0.0, binary=False):
        """restart file of the device mesh and state (fv_dev_write_restart)."""
        _check(lib().fv_dev_write_restart(self._h, os.fsencode(path), istep, time, int(binary)), "fv_dev_write_restart")

    def set_forest(self, mesh):
        """Upload the mesh's leaf addresses and SI for device regrids (fv_dev_set_forest)."""
        _check(lib().fv_dev_set_forest(self._h, mesh._p), "fv_dev_set_forest")

    def regrid(self, crit, maxlevel):
        """Device-resident mesh adaptation (fv_dev_regrid); returns the new leaf count."""
        nn = C.c_int()
        _check(lib().fv_dev_regrid(self._h, C.byref(crit), maxlevel, C.byref(nn)), "fv_dev_regrid")
        self.n = nn.value
        return self.n

    def criterionsmth(self, LEVEL, split):
        out = np.ascontiguousarray(split, dtype=np.int32).copy()
        _check(lib().fv_dev_criterionsmth(self._h, LEVEL, _ip(out)), "fv_dev_criterionsmth")
        return out

    def get_forest(self):
        """(root, level, ci, SI) of the device mesh."""
        n = self.n
        root = np.zeros(n, np.int32)
        level = np.zeros(n, np.int32)
        ci = np.zeros((n, 3), np.int32)
        SI = np.zeros((n, 2), np.int32)
        _check(lib().fv_dev_get_forest(self._h, _ip(root), _ip(level), _ip(ci), _ip(SI)), "fv_dev_get_forest")
        return root, level, ci, SI

    def get_pack(self):
        """(nbr [6][n], info [n], fine4 [nfine][4]) of the device mesh."""
        n = self.n
        nbr = np.zeros((6, n), np.int32)
        info = np.zeros(n, np.uint32)
        nf = C.c_int()
        _check(lib().fv_dev_get_pack(self._h, None, None, C.byref(nf), None), "fv_dev_get_pack")
        fine4 = np.zeros((max(nf.value, 1), 4), np.int32)
        _check(lib().fv_dev_get_pack(self._h, _ip(nbr), info.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(nf),
                                     _ip(fine4)), "fv_dev_get_pack")
        return nbr, info, fine4[:nf.value]

    def set_state(self, S):
        S = np.ascontiguousarray(S, dtype=np.float64)
        assert S.shape == (self.n, self.nq), (S.shape, self.n, self.nq)
        _check(lib().fv_dev_set_state(self._h, _dp(S)), "fv_dev_set_state")

    def get_state(self):
        S = np.empty((self.n, self.nq))
        _check(lib().fv_dev_get_state(self._h, _dp(S)), "fv_dev_get_state")
        return S

    def get_faces(self):
        SF = np.empty((self.n, 6, self.nq))
        _check(lib().fv_dev_get_faces(self._h, _dp(SF)), "fv_dev_get_faces")
        return SF

    def advance(self, nsteps):
        _check(lib().fv_dev_advance(self._h, nsteps), "fv_dev_advance")

    def begin_step(self):
        _check(lib().fv_dev_begin_step(self._h), "fv_dev_begin_step")

    def stage(self, irk):
        _check(lib().fv_dev_stage(self._h, irk), "fv_dev_stage")

    def set_dt(self, dt):
        _check(lib().fv_dev_set_dt(self._h, dt), "fv_dev_set_dt")

    def set_precision(self, prec):
        _check(lib().fv_dev_set_precision(self._h, prec), "fv_dev_set_precision")

    def set_fused(self, on):
        _check(lib().fv_dev_set_fused(self._h, 1 if on else 0), "fv_dev_set_fused")

    def set_bricks(self, on):
        """4x4x4 same-level blocks through the LDS block kernel: True or 2 (the default) plain and
        graded blocks (2:1 or 1:4 faces on the surface), 1 plain blocks only, False or 0 every leaf
        per leaf."""
        mode = (2 if on else 0) if isinstance(on, bool) else int(on)
        _check(lib().fv_dev_set_bricks(self._h, mode), "fv_dev_set_bricks")

    def brick_count(self):
        """(blocks, leaves outside blocks) of the current mesh."""
        a, b = C.c_int(), C.c_int()
        _check(lib().fv_dev_brick_count(self._h, C.byref(a), C.byref(b)), "fv_dev_brick_count")
        return a.value, b.value

    def brick_kinds(self):
        """(plain blocks, graded blocks) of the current mesh."""
        a, b = C.c_int(), C.c_int()
        _check(lib().fv_dev_brick_kinds(self._h, C.byref(a), C.byref(b)), "fv_dev_brick_kinds")
        return a.value, b.value

    def cfl_dt(self, cfl):
        d = C.c_double()
        _check(lib().fv_dev_cfl_dt(self._h, cfl, C.byref(d)), "fv_dev_cfl_dt")
        return d.value

    def indicator(self, crit, to_host=True):
        """(SI, split); to_host=False leaves both on the device, as the device regrid reads them."""
        if not to_host:
            _check(lib().fv_dev_indicator(self._h, C.byref(crit), None, None), "fv_dev_indicator")
            return None, None
        SI = np.zeros((self.n, 2), dtype=np.int32)
        split = np.zeros(self.n, dtype=np.int32)
        _check(lib().fv_dev_indicator(self._h, C.byref(crit), _ip(SI), _ip(split)), "fv_dev_indicator")
        return SI, split

    def interface_location(self, crit, copy=True):
        """SI from the current state; the following stages use it (loop.c:23).  copy=False: SI stays
        on the device and None is returned."""
        SI = np.zeros((self.n, 2), dtype=np.int32) if copy else None
        _check(lib().fv_dev_interface_location(self._h, C.byref(crit), _ip(SI) if copy else None),
               "fv_dev_interface_location")
        return SI

    def set_limiter_flags(self, lim):
        lim = np.ascontiguousarray(lim, dtype=np.uint8)
        assert lim.shape == (self.n,)
        _check(lib().fv_dev_set_limiter_flags(self._h, lim.ctypes.data_as(C.POINTER(C.c_uint8))), "limiter flags")

    def get_limiter_flags(self):
        lim = np.zeros(self.n, dtype=np.uint8)
        _check(lib().fv_dev_get_limiter_flags(self._h, lim.ctypes.data_as(C.POINTER(C.c_uint8))), "limiter flags")
        return lim

    def indicator_mlp(self, mlp, crit, smooth=True, to_host=True):
        """Learned indicator: (SI, split, smooth flags); smooth=False leaves the flags
        on the device only (they are the per-leaf limiter either way), to_host=False
        SI and split too."""
        if not to_host:
            _check(lib().fv_dev_indicator_mlp(self._h, mlp._h, C.byref(crit), None, None, None), "fv_dev_indicator_mlp")
            return None, None, None
        SI = np.zeros((self.n, 2), dtype=np.int32)
        split = np.zeros(self.n, dtype=np.int32)
        sm = np.zeros(self.n, dtype=np.uint8) if smooth else None
        _check(lib().fv_dev_indicator_mlp(self._h, mlp._h, C.byref(crit), _ip(SI), _ip(split),
                                          None if sm is None else sm.ctypes.data_as(C.POINTER(C.c_uint8))),
               "fv_dev_indicator_mlp")
        return SI, split, sm

    def time_stages(self, nrep):
        a, b = C.c_double(), C.c_double()
        _check(lib().fv_dev_time_stages(self._h, nrep, C.byref(a), C.byref(b)), "fv_dev_time_stages")
        return a.value, b.value


MLP_K = 112
MLP_H = 64
MLP_NO = 4


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MLP:
    """Learned refinement / limiter indicator (fv_mlp): weights, dataset and training on the device."""

    def __init__(self, seed=1234):
        h = C.c_void_p()
        _check(lib().fv_mlp_create(C.byref(h), seed), "fv_mlp_create")
        self._h = h
        _track(self)
        self.nw = lib().fv_mlp_num_weights()

    def close(self):
        if getattr(self, "_h", None):
            lib().fv_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None and not _shutdown:
            self.close()

    def weights(self):
        w = np.zeros(self.nw, dtype=np.float32)
        _check(lib().fv_mlp_get_weights(self._h, _fp(w)), "weights")
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        _check(lib().fv_mlp_set_weights(self._h, _fp(w)), "weights")

    def set_options(self, class_weights=None, smooth_thr=0.0):
        cw = None if class_weights is None else np.ascontiguousarray(class_weights, dtype=np.float32)
        _check(lib().fv_mlp_set_options(self._h, None if cw is None else _fp(cw), smooth_thr), "options")

    def sample(self, dev, crit):
        _check(lib().fv_mlp_sample(dev._h, self._h, C.byref(crit)), "fv_mlp_sample")

    def dataset_size(self):
        n = C.c_long()
        _check(lib().fv_mlp_dataset(self._h, C.byref(n), None, None, None), "dataset")
        return n.value

    def dataset(self):
        n = C.c_long()
        _check(lib().fv_mlp_dataset(self._h, C.byref(n), None, None, None), "dataset")
        X = np.zeros((n.value, MLP_K), dtype=np.float32)
        cls = np.zeros(n.value, dtype=np.int32)
        sm = np.zeros(n.value, dtype=np.float32)
        _check(lib().fv_mlp_dataset(self._h, C.byref(n), _fp(X), _ip(cls), _fp(sm)), "dataset")
        return X, cls, sm

    def clear_dataset(self):
        _check(lib().fv_mlp_clear_dataset(self._h), "clear")

    def train(self, epochs, batch=1024, lr=1e-3, seed=7):
        loss = C.c_float()
        _check(lib().fv_mlp_train(self._h, epochs, batch, lr, seed, C.byref(loss)), "fv_mlp_train")
        return loss.value

    def grad(self, X, cls, smooth):
        X = np.ascontiguousarray(X, dtype=np.float32)
        cls = np.ascontiguousarray(cls, dtype=np.int32)
        smooth = np.ascontiguousarray(smooth, dtype=np.float32)
        g = np.zeros(self.nw, dtype=np.float32)
        loss = C.c_float()
        _check(lib().fv_mlp_grad(self._h, _fp(X), _ip(cls), _fp(smooth), len(cls), _fp(g), C.byref(loss)), "grad")
        return g, loss.value

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros((X.shape[0], MLP_NO), dtype=np.float32)
        _check(lib().fv_mlp_forward(self._h, _fp(X), X.shape[0], _fp(out)), "forward")
        return out

    def infer(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros((X.shape[0], MLP_NO), dtype=np.float32)
        _check(lib().fv_mlp_infer(self._h, _fp(X), X.shape[0], _fp(out)), "infer")
        return out
