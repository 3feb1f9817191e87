"""The LDS block stage kernel (k_stage_brick) against the per-leaf kernels.

A brick is 64 consecutive leaves in DFS order forming a 4x4x4 block of one
level whose surface faces have same-level neighbours or boundaries (plain), or
also 2:1 and 1:4 faces whose neighbours are outside bricks (graded).  The
block kernel evaluates every face once and keeps face states in LDS; the
per-leaf kernels evaluate each face from both sides.  The two give the same
state to rounding (the HLLC flux is antisymmetric under exchange of sides
with the normal reversed), here within 1e-11 of each field's largest value per
stage (the largest element difference, not an average) and within the
oracle's 1e-10 over many steps; the brick list equals a brute-force search
over integer leaf boxes."""
import numpy as np
import pytest

import forestfv_public_amd as F
import oracle_lib as O
from cases import (bubble_case, sod_case, sod_params, sphere_amr_mesh, three_material_case, water_air_params)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def expected_bricks(m, params, graded=True):
    """Brick starts by integer geometry: 64 consecutive same-level leaves at global positions
    base + Morton(l).  Plain: every surface face touches one same-level leaf or the domain
    boundary.  Graded (when on): the others, less those with a leaf in any candidate across one
    of their 1:4 faces.  Returns (plain starts, graded starts)."""
    a = m.arrays()
    lvl, ci, root = a["level"], a["ci"], a["root"]
    lfc, ntr = a["lfc_neigs"], a["neigtr"]
    nr = m.m.nroot
    n = m.n
    L = int(lvl.max())
    D = 1 << (L - 1)
    rx = np.stack([root % nr[0], (root // nr[0]) % nr[1], root // (nr[0] * nr[1])], 1)
    size = D >> (lvl - 1)
    g = rx * (D // 1) + ci * size[:, None]          # global position of each leaf at the finest level
    mort = np.array([[(l & 1) | (((l >> 3) & 1) << 1), ((l >> 1) & 1) | (((l >> 4) & 1) << 1),
                      ((l >> 2) & 1) | (((l >> 5) & 1) << 1)] for l in range(64)])
    plain, cand = [], []
    i = 0
    while i + 63 < n:
        if np.all(lvl[i:i + 64] == lvl[i]) and np.array_equal(g[i:i + 64], g[i] + mort * size[i]):
            # surface faces: same-level neighbours or boundary (lfc 0) make a plain brick
            ok = True
            for k in range(64):
                for f in range(6):
                    j = ntr[i + k, f, 0]
                    if lfc[i + k, f] == 4 or (lfc[i + k, f] == 1 and lvl[j] != lvl[i]):
                        ok = False
            (plain if ok else cand).append(i)
            i += 64
            continue
        i += 1
    if not graded:
        return plain, []
    member = np.zeros(n, bool)
    for b in plain + cand:
        member[b:b + 64] = True
    kept = [b for b in cand
            if not any(member[ntr[b + k, f, q]] for k in range(64) for f in range(6) if lfc[b + k, f] == 4
                       for q in range(4))]
    return plain, kept


def compare_paths(p, m, S, nstage=3, prec=F.PREC_FP64, SI=None, tol=1e-11):
    if SI is not None:
        m.set_SI(SI)
    devs = []
    for on in (True, False):
        d = F.Device(p, m, prec=prec)
        d.set_bricks(on)
        d.set_state(S)
        devs.append(d)
    nb, nrest = devs[0].brick_count()
    say(f"  n={m.n} bricks={nb} rest={nrest}")
    worst = 0.0
    for irk in range(1, nstage + 1):
        devs[0].stage(irk)
        devs[1].stage(irk)
        A, B = devs[0].get_state(), devs[1].get_state()
        # per-field scale; the momentum components share the momentum magnitude, so a
        # component that is zero up to rounding (Sod's transverse momentum) is not
        # divided by its own noise
        nm = p.nmat
        scale = np.abs(B).max(0)
        scale[nm:nm + 3] = np.sqrt((B[:, nm:nm + 3] ** 2).sum(1)).max()
        scale[scale == 0] = 1.0
        err = (np.abs(A - B) / scale).max()
        worst = max(worst, err)
        assert np.all(np.isfinite(A))
        assert err <= tol, f"stage {irk}: max relative difference {err:.3e}"
    for d in devs:
        d.close()
    return nb, nrest, worst


def test_brick_list_matches_brute_force_on_amr_mesh():
    p = sod_params(bc=(1, 2, 5, 4, 4, 1))
    for args in ((3, 5, (1, 1, 1), 0.6, 1.0, None), (3, 5, (2, 1, 1), 0.45, 2.0, (1.2, 0.6, 0.6))):
        m = sphere_amr_mesh(p, *args)
        dev = F.Device(p, m)
        nb, nrest = dev.brick_count()
        dev.close()
        plain, graded = expected_bricks(m, p)
        dev = F.Device(p, m)
        npl, ngr = dev.brick_kinds()
        dev.set_bricks(1)
        nb1, nrest1 = dev.brick_count()
        dev.close()
        say(f"[bricks] n={m.n} device {nb} bricks ({npl} plain, {ngr} graded), brute force {len(plain)} + "
            f"{len(graded)}")
        assert (npl, ngr) == (len(plain), len(graded))
        assert nb == npl + ngr and nrest == m.n - 64 * nb
        assert nb1 == npl and nrest1 == m.n - 64 * npl
        assert nb > 0 and nrest > 0


def test_brick_list_on_a_uniform_mesh_and_across_roots():
    p = sod_params()
    m = F.Mesh.uniform((1, 1, 1), (0, 0, 0), 1.0, 5, list(p.bc))   # 16^3: 64 bricks
    dev = F.Device(p, m)
    assert dev.brick_count() == (64, 0)
    dev.close()
    m2 = F.Mesh.uniform((2, 2, 2), (0, 0, 0), 0.5, 2, list(p.bc))  # 8 roots of 2^3 leaves: one brick across them
    dev = F.Device(p, m2)
    assert dev.brick_count() == (1, 0)
    dev.close()


CASES = {
    "bubble_uniform": lambda: bubble_case(level=5),
    "bubble_literal": lambda: bubble_case(level=5, fix=0),
    "sod_bc_mix": lambda: sod_case(level=5, prelax=1, bc=(1, 2, 5, 4, 4, 1)),
    "bubble_prelax3": lambda: bubble_case(level=4, prelax=3),
    "bubble_mc": lambda: bubble_case(level=5, limiter=1),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_brick_path_matches_per_leaf_path(name, prec):
    p, m, S = CASES[name]()
    say(f"[brick vs per-leaf] {name} {prec}")
    tol = 1e-11 if prec == "fp64" else 1e-5   # FP32 faces: both paths round differently
    compare_paths(p, m, S, prec={"fp64": F.PREC_FP64, "mixed": F.PREC_MIXED}[prec], tol=tol)


@pytest.mark.parametrize("fix", [F.FIX_ALL, 0])
def test_brick_path_on_amr_with_interface_flags(fix):
    """Bricks next to per-leaf leaves, 2:1 faces outside bricks, random interface flags (so
    same-level faces are computed once or by both sides), corrected and literal residual."""
    p = water_air_params(prelax=1, fix=fix)
    m = sphere_amr_mesh(p, 3, 5, (1, 1, 1), 0.6 * 4e-3, 4e-3, (1.6e-3, 2e-3, 2.2e-3))
    S = m.init_bubble(p, (1000.0, 1.2), (5e6, 1e5), (2e-3, 2e-3, 2e-3), 0.8e-3, 2.0, 9, 0.01)
    h = 4e-3 / 16
    p.dt = 0.2 * h / 1900.0
    rng = np.random.default_rng(11)
    SI = np.zeros((m.n, 2), np.int32)
    SI[rng.random(m.n) < 0.3, 0] = 1
    nb, nrest, worst = compare_paths(p, m, S, nstage=3, SI=SI)
    assert nb > 0 and nrest > 0


def graded_face_counts(m, graded):
    """(2:1 faces, 1:4 faces) on the surfaces of the graded blocks: the block finer or coarser."""
    a = m.arrays()
    fine_side = coarse_side = 0
    for b in graded:
        for k in range(64):
            for f in range(6):
                if a["lfc_neigs"][b + k, f] == 4:
                    coarse_side += 1
                elif a["lfc_neigs"][b + k, f] == 1 and a["level"][a["neigtr"][b + k, f, 0]] < a["level"][b]:
                    fine_side += 1
    return fine_side, coarse_side


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("prelax", [1, 3])
def test_graded_bricks_match_per_leaf_path(prec, prelax):
    """Blocks with 2:1 faces (level-5 blocks beside level-4 leaves) and 1:4 faces (level-4 blocks
    beside level-5 families) on their surface against the per-leaf path, and the graded mode
    against the plain-only mode, on two AMR meshes of the bubble."""
    L = 4e-3
    meshes = ((3, 5, (1, 1, 1), 0.6 * L, L, (1.6e-3, 2e-3, 2.2e-3)),
              (4, 5, (1, 1, 1), 0.25 * L, L, (1.1e-3, 1.3e-3, 2.9e-3)))
    pr = {"fp64": F.PREC_FP64, "mixed": F.PREC_MIXED}[prec]
    tol = 1e-11 if prec == "fp64" else 1e-5
    kinds = [0, 0]
    for args in meshes:
        p = water_air_params(prelax=prelax)
        m = sphere_amr_mesh(p, *args)
        S = m.init_bubble(p, (1000.0, 1.2), (5e6, 1e5), (2e-3, 2e-3, 2e-3), 0.8e-3, 2.0, 9, 0.01)
        p.dt = 0.2 * (L / 32) / 1900.0
        dev = F.Device(p, m)
        npl, ngr = dev.brick_kinds()
        dev.close()
        plain, graded = expected_bricks(m, p)
        assert (npl, ngr) == (len(plain), len(graded))
        fs, cs = graded_face_counts(m, graded)
        kinds[0] += fs
        kinds[1] += cs
        say(f"[graded] {prec} PRELAX {prelax} n={m.n}: {npl} plain, {ngr} graded blocks; {fs} 2:1 and {cs} 1:4 "
            f"block faces")
        assert ngr > 0
        compare_paths(p, m, S, prec=pr, tol=tol)
        # graded (2) against plain-only (1): the faces the graded blocks take are the per-leaf ones
        devs = []
        for mode in (2, 1):
            d = F.Device(p, m, prec=pr)
            d.set_bricks(mode)
            d.set_state(S)
            devs.append(d)
        for irk in (1, 2, 3):
            for d in devs:
                d.stage(irk)
        A, B = devs[0].get_state(), devs[1].get_state()
        for d in devs:
            d.close()
        scale = np.abs(B).max(0)
        scale[2:5] = np.sqrt((B[:, 2:5] ** 2).sum(1)).max()
        scale[scale == 0] = 1.0
        err = (np.abs(A - B) / scale).max()
        say(f"  graded vs plain-only: {err:.3e}")
        assert np.all(np.isfinite(A)) and err <= tol
    assert kinds[0] > 0 and kinds[1] > 0


def test_brick_path_three_materials():
    p, m, S = three_material_case(level=5, prelax=0)
    compare_paths(p, m, S)


def test_brick_path_fixed_pressure_boundaries():
    p, m, S = sod_case(level=5, prelax=1, bc=(3, 2, 4, 4, 1, 5))
    p.bc_fixed_p = 1
    p.p_inlet, p.p_outlet = 1.5, 0.05
    compare_paths(p, m, S)


@pytest.mark.parametrize("mesh", ["fine_blocks", "coarse_blocks"])
def test_brick_path_100_steps_against_the_oracle(mesh):
    """The default FP64 stage (plain and graded blocks on) over 100 steps of the bubble on an
    AMR mesh: relative L1 <= 1e-10 per conserved field against the CPU restatement.  The
    first mesh's graded blocks have 2:1 faces (finer than their neighbours), the second's
    1:4 faces (coarser than them)."""
    p = water_air_params(prelax=1)
    L = 2.0 ** -8
    if mesh == "fine_blocks":
        m = sphere_amr_mesh(p, 3, 5, (1, 1, 1), 0.6 * L, L, (0.4 * L, 0.5 * L, 0.55 * L))
    else:
        m = sphere_amr_mesh(p, 4, 5, (1, 1, 1), 0.25 * L, L, (0.275 * L, 0.325 * L, 0.725 * L))
    S = m.init_bubble(p, (1000.0, 1.2), (5e6, 1e5), (2.0 ** -9,) * 3, 0.2 * L, 2.0, 5, 0.01)
    p.dt = 0.3 * (L / 16) / 1900.0
    plain, graded = expected_bricks(m, p)
    fs, cs = graded_face_counts(m, graded)
    say(f"[brick 100 steps] {mesh}: {len(graded)} graded blocks, {fs} 2:1 and {cs} 1:4 block faces")
    assert len(graded) > 0 and (fs if mesh == "fine_blocks" else cs) > 0
    dev = F.Device(p, m)
    assert dev.brick_kinds() == (len(plain), len(graded))
    nb, nrest = dev.brick_count()
    dev.set_state(S)
    dev.advance(100)
    Sg = dev.get_state()
    dev.close()
    So = O.OracleMesh(m).advance(p, S, 100)
    num = np.abs(Sg - So).sum(0)
    den = np.abs(So).sum(0)
    mom = np.sqrt((So[:, 2:5] ** 2).sum(1)).sum()
    den[2:5] = np.maximum(den[2:5], mom)
    err = num / den
    say(f"[brick 100 steps] bricks={nb} rest={nrest} rel L1 {err}")
    assert nb > 0
    assert np.all(err <= 1e-10), err
