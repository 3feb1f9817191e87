"""Shared test configurations: Sod (config C0), water/air bubble, AMR meshes."""
import numpy as np

import forestfv_public_amd as F

SOD_GAMMA = 1.4
SOD_RHO = (1.0, 0.125)
SOD_P = (1.0, 0.1)
WATER_AIR = dict(gamma=(6.12, 1.4), pinf=(3.43e8, 0.0))
WATER_AIR_RHO = (1000.0, 1.2)


def sod_params(rk=F.RK_SSP3, prelax=0, dt=None, level=6, limiter=0, bc=(4, 4, 4, 4, 4, 4), fix=F.FIX_ALL):
    """Config C0: two identical ideal-gas materials, left rho=1 p=1, right rho=0.125 p=0.1."""
    h = 1.0 / (1 << (level - 1))
    if dt is None:
        # CFL 0.5 against the largest signal speed of the Sod solution: |u| + c <= 1.93 < 2
        dt = 0.5 * h / 2.0
    return F.make_params(nmat=2, gamma=(SOD_GAMMA, SOD_GAMMA), pinf=(0.0, 0.0), amin=1e-8, avf_lim=0.5,
                         prelax=prelax, rk=rk, limiter=limiter, dt=dt, bc=bc, fix=fix)


def sod_case(level=6, nroot=(1, 1, 1), **kw):
    p = sod_params(level=level, **kw)
    rootsize = 1.0 / nroot[0]
    m = F.Mesh.uniform(nroot, (0.0, 0.0, 0.0), rootsize, level, list(p.bc))
    S = m.init_split_x(p, SOD_RHO, SOD_P, 0.5)
    return p, m, S


def water_air_params(dt=1e-9, prelax=1, rk=F.RK_SSP3, bc=(2, 2, 2, 2, 2, 2), limiter=0, fix=F.FIX_ALL):
    return F.make_params(nmat=2, gamma=WATER_AIR["gamma"], pinf=WATER_AIR["pinf"], amin=1e-6, avf_lim=0.5,
                         prelax=prelax, rk=rk, limiter=limiter, dt=dt, bc=bc, fix=fix)


def bubble_case(level=5, nroot=(1, 1, 1), L=4e-3, R=0.4e-3, perturb=0.01, prelax=1, dt=None, seed=1234, **kw):
    """Config C1 physics on a uniform grid: air bubble (1 bar) in water (50 bar)."""
    h = L / nroot[0] / (1 << (level - 1))
    if dt is None:
        c_w = np.sqrt(6.12 * (5e6 + 3.43e8) / 1000.0)
        dt = 0.4 * h / c_w
    p = water_air_params(dt=dt, prelax=prelax, **kw)
    m = F.Mesh.uniform(nroot, (0.0, 0.0, 0.0), L / nroot[0], level, list(p.bc))
    c = (0.5 * L, 0.5 * L, 0.5 * L)
    S = m.init_bubble(p, WATER_AIR_RHO, (5e6, 1e5), c, R, 2.0, seed, perturb)
    return p, m, S


def random_amr_mesh(params, base_level=2, max_level=4, nroot=(1, 1, 1), seed=0, frac=0.3, passes=3, L=1.0):
    """A 2:1 face-balanced AMR mesh built by random refinement through the host AMR path."""
    rng = np.random.default_rng(seed)
    m = F.Mesh.uniform(nroot, (0.0, 0.0, 0.0), L / nroot[0], base_level, list(params.bc))
    nq = F.neq(params.nmat)
    S = np.tile(np.array([[1.0, 1e-8, 0, 0, 0, 2.5, 1 - 1e-8, 2.5, 1e-8][:nq]]), (m.n, 1))
    for _ in range(passes):
        a = m.arrays()
        split = a["level"].copy()
        pick = rng.random(m.n) < frac
        split[pick] = np.minimum(a["level"][pick] + 1, max_level)
        split = m.criterionsmth(max_level, split)
        S = m.adapt(params, split, S)
    return m


THREE = dict(gamma=(6.12, 1.4, 2.0), pinf=(3.43e8, 0.0, 1.0e8))
THREE_RHO = (1000.0, 1.2, 2000.0)


def three_material_params(prelax=3, dt=None, level=4, rk=F.RK_SSP3, bc=(2, 2, 2, 2, 2, 2), limiter=0):
    """Water, air and a stiffened third material (gamma 2, p_inf 1e8 Pa) at 10 bar."""
    h = 1.0 / (1 << (level - 1))
    if dt is None:
        dt = 0.2 * h / 1500.0
    return F.make_params(nmat=3, gamma=THREE["gamma"], pinf=THREE["pinf"], amin=1e-6, avf_lim=0.5, prelax=prelax,
                         rk=rk, limiter=limiter, dt=dt, bc=bc)


def three_material_case(level=4, prelax=3, **kw):
    """Three slabs along x (water | air | third material) blended over about two cells,
    in pressure equilibrium at 1e6 Pa, with a smooth velocity field: every cell is a
    convex combination of the pure states at its own velocity."""
    p = three_material_params(prelax=prelax, level=level, **kw)
    m = F.Mesh.uniform((1, 1, 1), (0.0, 0.0, 0.0), 1.0, level, list(p.bc))
    c = m.centres()
    h = 1.0 / (1 << (level - 1))
    s0 = 0.5 * (1 - np.tanh((c[:, 0] - 0.35) / (1.5 * h)))
    s2 = 0.5 * (1 + np.tanh((c[:, 0] - 0.65) / (1.5 * h)))
    w = np.stack([s0, np.clip(1 - s0 - s2, 0, 1), s2], 1)
    w /= w.sum(1, keepdims=True)
    S = np.zeros((m.n, F.neq(3)))
    for i in range(m.n):
        u = (8.0 * np.sin(2 * np.pi * c[i, 1]), 5.0 * np.cos(2 * np.pi * c[i, 2]), 3.0 * np.sin(2 * np.pi * c[i, 0]))
        for k in range(3):
            S[i] += w[i, k] * F.init_pure(p, THREE_RHO, k, 1.0e6, u)
    return p, m, S


def sphere_amr_mesh(params, base_level=4, max_level=6, nroot=(1, 1, 1), radius=0.3, L=1.0, centre=None):
    """An AMR mesh refined towards a sphere: leaves within `radius` of the centre go to
    max_level, the smoothing adds the 2:1 halo, so the mesh has large same-level regions
    (many complete 4x4x4 blocks) and 2:1 faces between them."""
    m = F.Mesh.uniform(nroot, (0.0, 0.0, 0.0), L / nroot[0], base_level, list(params.bc))
    nq = F.neq(params.nmat)
    c0 = np.array(centre if centre is not None else (0.5 * L * nroot[0] / nroot[0],) * 3)
    S = np.tile(np.array([[1.0, 1e-8, 0, 0, 0, 2.5, 1 - 1e-8, 2.5, 1e-8][:nq]]), (m.n, 1))
    for _ in range(max_level - base_level):
        c = m.centres()
        lv = m.arrays()["level"]
        split = np.where(np.linalg.norm(c - c0, axis=1) < radius, max_level, lv).astype(np.int32)
        split = m.criterionsmth(max_level, split)
        S = m.adapt(params, split, S)
    return m
