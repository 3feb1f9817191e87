"""ctypes binding of the CPU restatement in oracle/ (test infrastructure only).

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
oracle the HIP path is tested against; never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "libfvoracle.so")

ORC_MAXMAT = 4


class orc_params(C.Structure):
    _fields_ = [
        ("nmat", C.c_int),
        ("gamma", C.c_double * ORC_MAXMAT),
        ("pinf", C.c_double * ORC_MAXMAT),
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


class orc_mesh(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("nroot", C.c_int * 3),
        ("origin", C.c_double * 3),
        ("rootsize", C.c_double),
        ("root", C.POINTER(C.c_int)),
        ("level", C.POINTER(C.c_int)),
        ("ci", C.POINTER(C.c_int)),
        ("lfc_neigs", C.POINTER(C.c_int)),
        ("neigtr", C.POINTER(C.c_int)),
        ("bc", C.POINTER(C.c_int)),
        ("SI", C.POINTER(C.c_int)),
    ]


class orc_cell(C.Structure):
    _fields_ = [(nm, C.c_double * 6) for nm in ("nx", "ny", "nz", "nxt1", "nyt1", "nzt1", "nxt2", "nyt2", "nzt2")] + [
        ("Area", (C.c_double * 4) * 6),
        ("Vol", C.c_double),
        ("dist_cf", C.c_double * 6),
        ("dist_cc", C.c_double * 6),
        ("xc", C.c_double), ("yc", C.c_double), ("zc", C.c_double),
    ]


class orc_crit(C.Structure):
    _fields_ = [
        ("uniform_on", C.c_int), ("uniform_lvl", C.c_int),
        ("prs_on", C.c_int), ("prs_lvl", C.c_int), ("prs_thr", C.c_double),
        ("rho_on", C.c_int), ("rho_lvl", C.c_int), ("rho_thr", C.c_double),
        ("vel_on", C.c_int), ("vel_lvl", C.c_int), ("vel_thr", C.c_double),
        ("inter0_on", C.c_int), ("inter0_lvl", C.c_int), ("inter0_lrs", C.c_int),
        ("inter1_on", C.c_int), ("inter1_lvl", C.c_int), ("inter1_lrs", C.c_int),
    ]


class orc_m5_params(C.Structure):
    _fields_ = [
        ("nmat", C.c_int),
        ("gamma", C.c_double * ORC_MAXMAT),
        ("pinf", C.c_double * ORC_MAXMAT),
        ("mush", C.c_double * ORC_MAXMAT),
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


M5_REC = 39   # orc_m5_record layout (fv_oracle_m5.h)
M5_REC_FIELDS = {"r": 0, "e": 1, "u": 2, "v": 3, "w": 4, "phi0": 5, "ra": 6, "avf": 9, "are": 12, "c2Y": 15,
                 "p_hydro": 18, "Amat": 21, "st": 30}

_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib(variant="O2"):
    """The oracle library: the -O2 build (the checker), or "O0", the same sources built
    without optimisation as the reference makefile does (CPU baseline only)."""
    if variant not in _libs:
        path = ORACLE_SO if variant == "O2" else os.path.join(ORACLE_DIR, "_build", "libfvoracle_O0.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.POINTER
        dp = P(C.c_double)
        L.orc_geometry.argtypes = [P(orc_mesh), P(orc_cell)]
        L.orc_advance.argtypes = [P(orc_mesh), P(orc_cell), P(orc_params), dp, C.c_int]
        L.orc_stage.argtypes = [P(orc_mesh), P(orc_cell), P(orc_params), C.c_int, dp, dp, dp, dp]
        L.orc_faceinterp.argtypes = [P(orc_mesh), P(orc_cell), P(orc_params), dp, dp]
        L.orc_relax.argtypes = [P(orc_params), dp]
        L.orc_relax.restype = None
        L.orc_hllc_state.argtypes = [P(orc_params), dp, dp, dp, dp, dp, C.c_int, dp, dp]
        L.orc_hllc_state.restype = None
        L.orc_physical_flux.argtypes = [P(orc_params), dp, dp, dp]
        L.orc_physical_flux.restype = None
        L.orc_interface_location.argtypes = [P(orc_mesh), P(orc_params), P(orc_crit), dp, P(C.c_int)]
        L.orc_criterionsplt.argtypes = [P(orc_mesh), P(orc_params), P(orc_crit), dp, P(C.c_int), P(C.c_int)]
        L.orc_cfl_dt.argtypes = [P(orc_mesh), P(orc_cell), P(orc_params), dp, C.c_double]
        L.orc_cfl_dt.restype = C.c_double
        L.orc_set_cell_limiter.argtypes = [C.c_void_p]
        L.orc_set_cell_limiter.restype = None
        L.orc_m5_set_cell_limiter.argtypes = [C.c_void_p]
        L.orc_m5_set_cell_limiter.restype = None
        L.orc_advance_par.argtypes = [P(orc_mesh), P(orc_cell), P(orc_params), dp, C.c_int, C.c_int]
        L.orc_num_threads.argtypes = []
        L.orc_cons2prim.argtypes = [P(orc_params), dp, dp]
        L.orc_cons2prim.restype = None
        L.orc_prim2cons.argtypes = [P(orc_params), dp, dp]
        L.orc_prim2cons.restype = None
        L.orc_m5_record.argtypes = [P(orc_m5_params), dp, dp]
        L.orc_m5_record.restype = None
        L.orc_m5_gfm_face.argtypes = [P(orc_m5_params), dp, dp, dp, C.c_double, dp, dp]
        L.orc_m5_rk.argtypes = [P(orc_mesh), P(orc_cell), P(orc_m5_params), C.c_int, dp, dp, dp, P(C.c_uint8)]
        L.orc_m5_gfm_rel.argtypes = [P(orc_mesh), P(orc_cell), P(orc_m5_params), dp, dp, P(C.c_uint8)]
        L.orc_m5_relax.argtypes = [P(orc_m5_params), dp]
        L.orc_m5_faceinterp.argtypes = [P(orc_mesh), P(orc_cell), P(orc_m5_params), dp, dp]
        L.orc_m5_residual.argtypes = [P(orc_mesh), P(orc_cell), P(orc_m5_params), dp, dp, dp, P(C.c_uint8)]
        L.orc_m5_stage.argtypes = [P(orc_mesh), P(orc_cell), P(orc_m5_params), C.c_int, dp, dp, dp, dp,
                                   P(C.c_uint8)]
        L.orc_m5_advance.argtypes = [P(orc_mesh), P(orc_cell), P(orc_m5_params), dp, C.c_int]
        L.orc_m5_cfl_dt.argtypes = [P(orc_mesh), P(orc_cell), P(orc_m5_params), dp, C.c_double]
        L.orc_m5_cfl_dt.restype = C.c_double
        L.orc_m5_face.argtypes = [P(orc_m5_params), dp, dp, dp, dp, dp, C.c_int, C.c_double, dp, dp]
        L.orc_m5_plastic.argtypes = [P(orc_m5_params), dp]
        ip = P(C.c_int)
        L.orc_criterionsmth.argtypes = [P(orc_mesh), C.c_int, ip]
        L.orc_refineh.argtypes = [P(orc_mesh), P(orc_params), ip, dp, ip, ip, ip, ip, dp, ip, ip]
        L.orc_m5_refineh.argtypes = [P(orc_mesh), P(orc_m5_params), ip, dp, ip, ip, ip, ip, dp, ip, ip]
        L.orc_m5_interface_location.argtypes = [P(orc_mesh), P(orc_m5_params), P(orc_crit), dp, ip]
        L.orc_m5_criterionsplt.argtypes = [P(orc_mesh), P(orc_m5_params), P(orc_crit), dp, ip, ip]
        L.orc_bc_ghost.argtypes = [P(orc_params), dp, dp, dp, dp, C.c_int, dp, dp]
        L.orc_bc_ghost.restype = None
        _libs[variant] = L
    return _libs[variant]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def params_from(p):
    """orc_params with the same field values as a forestfv fv_params."""
    q = orc_params()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(orc_params))
    return q


def crit_from(c):
    q = orc_crit()
    C.memmove(C.byref(q), C.byref(c), C.sizeof(orc_crit))
    return q


class OracleMesh:
    """Oracle view of a product mesh's topology (geometry is recomputed by the oracle)."""

    def __init__(self, mesh, SI=None, variant="O2", geometry_from=None):
        self.L = lib(variant)
        a = mesh.arrays()
        m = mesh.m
        self.arrays = {k: np.ascontiguousarray(v, dtype=np.int32) for k, v in a.items()}
        if SI is not None:
            self.arrays["SI"] = np.ascontiguousarray(SI, dtype=np.int32)
        om = orc_mesh()
        om.n = m.n
        for k in range(3):
            om.nroot[k] = m.nroot[k]
            om.origin[k] = m.origin[k]
        om.rootsize = m.rootsize
        for name in ("root", "level", "ci", "lfc_neigs", "neigtr", "bc", "SI"):
            setattr(om, name, _ip(self.arrays[name]))
        self.om = om
        self.n = m.n
        if geometry_from is not None:
            # the same mesh's geometry, computed by another build of the same sources
            assert geometry_from.n == m.n
            self.cells = geometry_from.cells
        else:
            self.cells = (orc_cell * m.n)()
            rc = self.L.orc_geometry(C.byref(om), self.cells)
            assert rc == 0

    def set_SI(self, SI):
        self.arrays["SI"] = np.ascontiguousarray(SI, dtype=np.int32)
        self.om.SI = _ip(self.arrays["SI"])

    def advance(self, params, S, nsteps):
        p = params_from(params)
        S = np.ascontiguousarray(S, dtype=np.float64).copy()
        rc = self.L.orc_advance(C.byref(self.om), self.cells, C.byref(p), _dp(S), nsteps)
        if rc != 0:
            raise RuntimeError(f"orc_advance status {rc}")
        return S

    def advance_par(self, params, S, nsteps, nthreads=0):
        """orc_advance_par: the same steps with OpenMP over the leaves (same bits)."""
        p = params_from(params)
        S = np.ascontiguousarray(S, dtype=np.float64).copy()
        rc = self.L.orc_advance_par(C.byref(self.om), self.cells, C.byref(p), _dp(S), nsteps, nthreads)
        if rc != 0:
            raise RuntimeError(f"orc_advance_par status {rc}")
        return S

    def stage(self, params, irk, SN, S):
        p = params_from(params)
        nq = 3 * params.nmat + 3
        SN = np.ascontiguousarray(SN, dtype=np.float64)
        S = np.ascontiguousarray(S, dtype=np.float64).copy()
        SF = np.zeros((self.n, 6, nq))
        RHS = np.zeros((self.n, nq))
        rc = self.L.orc_stage(C.byref(self.om), self.cells, C.byref(p), irk, _dp(SN), _dp(S), _dp(SF), _dp(RHS))
        if rc != 0:
            raise RuntimeError(f"orc_stage status {rc}")
        return S, SF, RHS

    def faceinterp(self, params, S):
        p = params_from(params)
        nq = 3 * params.nmat + 3
        S = np.ascontiguousarray(S, dtype=np.float64)
        SF = np.zeros((self.n, 6, nq))
        self.L.orc_faceinterp(C.byref(self.om), self.cells, C.byref(p), _dp(S), _dp(SF))
        return SF

    def m5_rk(self, m5p, irk, SN, S, RHS, gfm_face):
        """orc_m5_rk: returns (SN, S) after model-5 RK stage irk."""
        SN = np.ascontiguousarray(SN, dtype=np.float64).copy()
        S = np.ascontiguousarray(S, dtype=np.float64).copy()
        R = np.ascontiguousarray(RHS, dtype=np.float64)
        gf = np.ascontiguousarray(gfm_face, dtype=np.uint8)
        rc = self.L.orc_m5_rk(C.byref(self.om), self.cells, C.byref(m5_from(m5p)), irk, _dp(SN), _dp(S), _dp(R),
                             gf.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc != 0:
            raise RuntimeError(f"orc_m5_rk status {rc}")
        return SN, S

    def m5_gfm_rel(self, m5p, S_before, S, gfm_face):
        Sb = np.ascontiguousarray(S_before, dtype=np.float64)
        S = np.ascontiguousarray(S, dtype=np.float64).copy()
        gf = np.ascontiguousarray(gfm_face, dtype=np.uint8)
        rc = self.L.orc_m5_gfm_rel(C.byref(self.om), self.cells, C.byref(m5_from(m5p)), _dp(Sb), _dp(S),
                                  gf.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc != 0:
            raise RuntimeError(f"orc_m5_gfm_rel status {rc}")
        return S

    def m5_faceinterp(self, m5p, S):
        """orc_m5_faceinterp: MUSCL face states [n][6][NEQ]."""
        nq = 3 * m5p.nmat + 12
        S = np.ascontiguousarray(S, dtype=np.float64)
        SF = np.zeros((self.n, 6, nq))
        rc = self.L.orc_m5_faceinterp(C.byref(self.om), self.cells, C.byref(m5_from(m5p)), _dp(S), _dp(SF))
        if rc != 0:
            raise RuntimeError(f"orc_m5_faceinterp status {rc}")
        return SF

    def m5_residual(self, m5p, S, SF=None):
        """orc_m5_residual: (RHS [n][NEQ], gfm_face [n][6])."""
        nq = 3 * m5p.nmat + 12
        S = np.ascontiguousarray(S, dtype=np.float64)
        SF = np.zeros((self.n, 6, nq)) if SF is None else np.ascontiguousarray(SF, dtype=np.float64)
        RHS = np.zeros((self.n, nq))
        gf = np.zeros((self.n, 6), np.uint8)
        rc = self.L.orc_m5_residual(C.byref(self.om), self.cells, C.byref(m5_from(m5p)), _dp(S), _dp(SF), _dp(RHS),
                                   gf.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc != 0:
            raise RuntimeError(f"orc_m5_residual status {rc}")
        return RHS, gf

    def m5_stage(self, m5p, irk, SN, S):
        """orc_m5_stage: (S, SF, RHS, gfm_face) after stage irk."""
        nq = 3 * m5p.nmat + 12
        SN = np.ascontiguousarray(SN, dtype=np.float64)
        S = np.ascontiguousarray(S, dtype=np.float64).copy()
        SF = np.zeros((self.n, 6, nq))
        RHS = np.zeros((self.n, nq))
        gf = np.zeros((self.n, 6), np.uint8)
        rc = self.L.orc_m5_stage(C.byref(self.om), self.cells, C.byref(m5_from(m5p)), irk, _dp(SN), _dp(S), _dp(SF),
                                _dp(RHS), gf.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc != 0:
            raise RuntimeError(f"orc_m5_stage status {rc}")
        return S, SF, RHS, gf

    def m5_cfl_dt(self, m5p, S, cfl):
        return self.L.orc_m5_cfl_dt(C.byref(self.om), self.cells, C.byref(m5_from(m5p)),
                                    _dp(np.ascontiguousarray(S, dtype=np.float64)), cfl)

    def m5_advance(self, m5p, S, nsteps):
        S = np.ascontiguousarray(S, dtype=np.float64).copy()
        rc = self.L.orc_m5_advance(C.byref(self.om), self.cells, C.byref(m5_from(m5p)), _dp(S), nsteps)
        if rc != 0:
            raise RuntimeError(f"orc_m5_advance status {rc}")
        return S

    def interface_location(self, params, crit, S):
        SI = np.zeros((self.n, 2), dtype=np.int32)
        self.L.orc_interface_location(C.byref(self.om), C.byref(params_from(params)), C.byref(crit_from(crit)),
                                     _dp(np.ascontiguousarray(S)), _ip(SI))
        return SI

    def criterionsplt(self, params, crit, S, SI):
        split = np.zeros(self.n, dtype=np.int32)
        SI = np.ascontiguousarray(SI, dtype=np.int32)
        self.L.orc_criterionsplt(C.byref(self.om), C.byref(params_from(params)), C.byref(crit_from(crit)),
                                _dp(np.ascontiguousarray(S)), _ip(SI), _ip(split))
        return split

    def criterionsmth(self, LEVEL, split):
        """orc_criterionsmth: the smoothed target levels (a new array)."""
        sp = np.ascontiguousarray(split, dtype=np.int32).copy()
        rc = self.L.orc_criterionsmth(C.byref(self.om), int(LEVEL), _ip(sp))
        if rc != 0:
            raise RuntimeError(f"orc_criterionsmth status {rc}")
        return sp

    def refineh(self, params, split, S):
        """orc_refineh: dict of the new leaves (root, level, ci, S, SI, split) in DFS order."""
        nq = 3 * params.nmat + 3
        cap = 8 * self.n
        sp = np.ascontiguousarray(split, dtype=np.int32)
        S = np.ascontiguousarray(S, dtype=np.float64)
        out = {"root": np.zeros(cap, np.int32), "level": np.zeros(cap, np.int32), "ci": np.zeros((cap, 3), np.int32),
               "S": np.zeros((cap, nq)), "SI": np.zeros((cap, 2), np.int32), "split": np.zeros(cap, np.int32)}
        nout = C.c_int()
        rc = self.L.orc_refineh(C.byref(self.om), C.byref(params_from(params)), _ip(sp), _dp(S), C.byref(nout),
                               _ip(out["root"]), _ip(out["level"]), _ip(out["ci"]), _dp(out["S"]), _ip(out["SI"]),
                               _ip(out["split"]))
        if rc != 0:
            raise RuntimeError(f"orc_refineh status {rc}")
        return {k: v[:nout.value].copy() for k, v in out.items()}

    def m5_interface_location(self, m5p, crit, S):
        SI = np.zeros((self.n, 2), dtype=np.int32)
        self.L.orc_m5_interface_location(C.byref(self.om), C.byref(m5_from(m5p)), C.byref(crit_from(crit)),
                                        _dp(np.ascontiguousarray(S, dtype=np.float64)), _ip(SI))
        return SI

    def m5_criterionsplt(self, m5p, crit, S, SI):
        split = np.zeros(self.n, dtype=np.int32)
        SI = np.ascontiguousarray(SI, dtype=np.int32)
        self.L.orc_m5_criterionsplt(C.byref(self.om), C.byref(m5_from(m5p)), C.byref(crit_from(crit)),
                                   _dp(np.ascontiguousarray(S, dtype=np.float64)), _ip(SI), _ip(split))
        return split

    def m5_refineh(self, m5p, split, S):
        """orc_m5_refineh: dict of the new leaves (root, level, ci, S, SI, split) in DFS order."""
        nq = 3 * m5p.nmat + 12
        cap = 8 * self.n
        sp = np.ascontiguousarray(split, dtype=np.int32)
        S = np.ascontiguousarray(S, dtype=np.float64)
        out = {"root": np.zeros(cap, np.int32), "level": np.zeros(cap, np.int32), "ci": np.zeros((cap, 3), np.int32),
               "S": np.zeros((cap, nq)), "SI": np.zeros((cap, 2), np.int32), "split": np.zeros(cap, np.int32)}
        nout = C.c_int()
        rc = self.L.orc_m5_refineh(C.byref(self.om), C.byref(m5_from(m5p)), _ip(sp), _dp(S), C.byref(nout),
                                  _ip(out["root"]), _ip(out["level"]), _ip(out["ci"]), _dp(out["S"]), _ip(out["SI"]),
                                  _ip(out["split"]))
        if rc != 0:
            raise RuntimeError(f"orc_m5_refineh status {rc}")
        return {k: v[:nout.value].copy() for k, v in out.items()}

    def cfl_dt(self, params, S, cfl):
        return self.L.orc_cfl_dt(C.byref(self.om), self.cells, C.byref(params_from(params)),
                                _dp(np.ascontiguousarray(S)), cfl)

    def geometry(self):
        g = {}
        for name in ("nx", "ny", "nz", "nxt1", "nyt1", "nzt1", "nxt2", "nyt2", "nzt2", "dist_cf", "dist_cc"):
            g[name] = np.array([list(getattr(c, name)) for c in self.cells])
        g["Area"] = np.array([[list(r) for r in c.Area] for c in self.cells])
        g["Vol"] = np.array([c.Vol for c in self.cells])
        g["xc"] = np.array([[c.xc, c.yc, c.zc] for c in self.cells])
        return g


_keep = {}


def set_cell_limiter(lim):
    """Per-leaf limiter flags for params.limiter == 2 (None clears)."""
    if lim is None:
        _keep.pop("lim", None)
        lib().orc_set_cell_limiter(None)
        return
    a = np.ascontiguousarray(lim, dtype=np.uint8)
    _keep["lim"] = a
    lib().orc_set_cell_limiter(a.ctypes.data)


def m5_set_cell_limiter(lim):
    """Per-leaf slope flags of model 5 for limiter == 2 (None clears)."""
    if lim is None:
        _keep.pop("lim5", None)
        lib().orc_m5_set_cell_limiter(None)
        return
    a = np.ascontiguousarray(lim, dtype=np.uint8)
    _keep["lim5"] = a
    lib().orc_m5_set_cell_limiter(a.ctypes.data)


def cons2prim(params, vec):
    v = np.ascontiguousarray(vec, dtype=np.float64)
    out = np.zeros(3 * params.nmat + 2)
    lib().orc_cons2prim(C.byref(params_from(params)), _dp(v), _dp(out))
    return out


def prim2cons(params, prim):
    v = np.ascontiguousarray(prim, dtype=np.float64)
    out = np.zeros(3 * params.nmat + 3)
    lib().orc_prim2cons(C.byref(params_from(params)), _dp(v), _dp(out))
    return out


def relax(params, vec):
    v = np.ascontiguousarray(vec, dtype=np.float64).copy()
    lib().orc_relax(C.byref(params_from(params)), _dp(v))
    return v


def hllc(params, sfL, sfR, n, t1, t2, bc=0):
    nq = 3 * params.nmat + 3
    flux = np.zeros(nq)
    us = np.zeros(3)
    lib().orc_hllc_state(C.byref(params_from(params)), _dp(np.ascontiguousarray(sfL, dtype=np.float64)),
                         _dp(np.ascontiguousarray(sfR, dtype=np.float64)), _dp(np.ascontiguousarray(n, dtype=np.float64)),
                         _dp(np.ascontiguousarray(t1, dtype=np.float64)), _dp(np.ascontiguousarray(t2, dtype=np.float64)), bc,
                         _dp(flux), _dp(us))
    return flux, us


def bc_ghost(params, sf, n, t1, t2, code):
    """orc_bc_ghost: (ghost velocity [3], ghost conservative vector [NEQ])."""
    nq = 3 * params.nmat + 3
    uvw = np.zeros(3)
    vec = np.zeros(nq)
    lib().orc_bc_ghost(C.byref(params_from(params)), _dp(np.ascontiguousarray(sf, dtype=np.float64)),
                       _dp(np.ascontiguousarray(n, dtype=np.float64)), _dp(np.ascontiguousarray(t1, dtype=np.float64)),
                       _dp(np.ascontiguousarray(t2, dtype=np.float64)), int(code), _dp(uvw), _dp(vec))
    return uvw, vec


def physical_flux(params, vec, n):
    nq = 3 * params.nmat + 3
    flux = np.zeros(nq)
    lib().orc_physical_flux(C.byref(params_from(params)), _dp(np.ascontiguousarray(vec, dtype=np.float64)),
                            _dp(np.ascontiguousarray(n, dtype=np.float64)), _dp(flux))
    return flux


def m5_from(p):
    """orc_m5_params with the same field values as a forestfv fv_m5_params."""
    q = orc_m5_params()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(orc_m5_params))
    return q


def m5_record(m5p, vec):
    """orc_m5_record: dict of the face record fields of one model-5 state vector."""
    out = np.zeros(M5_REC)
    v = np.ascontiguousarray(vec, dtype=np.float64)
    lib().orc_m5_record(C.byref(m5_from(m5p)), _dp(v), _dp(out))
    nm = m5p.nmat
    rec = {}
    for k, off in M5_REC_FIELDS.items():
        if k in ("ra", "avf", "are", "c2Y", "p_hydro"):
            rec[k] = out[off:off + nm].copy()
        elif k == "Amat":
            rec[k] = out[off:off + 9].reshape(3, 3).copy()
        elif k == "st":
            rec[k] = out[off:off + 9].reshape(3, 3).copy()
        else:
            rec[k] = out[off]
    return rec


def m5_gfm_face(m5p, UL, UR, frame, area):
    """orc_m5_gfm_face: (flux [NEQ], unc [3]) of one interface sub-face; raises on NaN."""
    nq = 3 * m5p.nmat + 12
    flux = np.zeros(nq)
    unc = np.zeros(3)
    rc = lib().orc_m5_gfm_face(C.byref(m5_from(m5p)), _dp(np.ascontiguousarray(UL, dtype=np.float64)),
                               _dp(np.ascontiguousarray(UR, dtype=np.float64)),
                               _dp(np.ascontiguousarray(frame, dtype=np.float64)), float(area), _dp(flux), _dp(unc))
    if rc != 0:
        raise FloatingPointError("NaN in the GFM face flux")
    return flux, unc


def m5_relax(m5p, vec):
    """orc_m5_relax: the relaxed copy of one model-5 state vector; raises on failure."""
    v = np.ascontiguousarray(vec, dtype=np.float64).copy()
    rc = lib().orc_m5_relax(C.byref(m5_from(m5p)), _dp(v))
    if rc != 0:
        raise FloatingPointError(f"orc_m5_relax status {rc}")
    return v


def m5_face(m5p, XL, CL, XR, CR, frame, bc, area):
    """orc_m5_face: (flux [NEQ], unc [3], gfm) of one face between two face states."""
    nq = 3 * m5p.nmat + 12
    flux = np.zeros(nq)
    unc = np.zeros(3)
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (XL, CL, XR, CR, frame)]
    rc = lib().orc_m5_face(C.byref(m5_from(m5p)), *[_dp(x) for x in a], int(bc), float(area), _dp(flux), _dp(unc))
    if rc < 0:
        raise FloatingPointError("NaN in the face flux")
    return flux, unc, rc


def m5_plastic(m5p, vec):
    """orc_m5_plastic: (the state after the plastic step, whether A changed)."""
    v = np.ascontiguousarray(vec, dtype=np.float64).copy()
    changed = lib().orc_m5_plastic(C.byref(m5_from(m5p)), _dp(v))
    return v, bool(changed)
