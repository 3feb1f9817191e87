"""Pins of the oracle's refinement indicator (orc_interface_location,
orc_criterionsplt) that do not restate its sweeps.

interface_location.c:14-47 marks a leaf SI = 1 when a non-boundary face has
phi_L phi_R < 0 (phi = clipped alpha - AVF_LIM of the two cells); its passes
SI_lvl = 2 .. ADAPT_INTER_MAX - 1 (:51-82) give SI_lvl to unmarked leaves next
to a leaf marked SI_lvl - 1.  For one indicator this is breadth-first
layering: SI = 1 + the graph distance (over faces, 2:1 faces included) to the
nearest marked leaf while that is below ADAPT_INTER_MAX, else 0.  The tests
check that against scipy's multi-source shortest paths on random AMR forests
and against the closed form |i - i_plane| for a plane on a uniform grid.

criterionsplt.c:3-140 on a planar jump of a uniform grid: only the two
columns next to the plane see the jump; the jump ratio is |a - b| / max(|a|,
|b|), with >= against the threshold; a later criterion overwrites an earlier
one (velocity over density over pressure) rather than taking the maximum;
the velocity criterion reads the density jump (criterionsplt.c:113); the
interface layers [1, lrs0) and [lrs0, lrs1) give their levels.
"""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import shortest_path

import forestfv_public_amd as F
import oracle_lib as O
from cases import random_amr_mesh, sod_params


def adjacency(m):
    """Leaf graph over the non-boundary faces of a forest (neighbour lists of fv_mesh, whose face
    lists are pinned against brute-force geometry in test_mesh.py)."""
    a = m.arrays()
    rows, cols = [], []
    for i in range(m.n):
        for f in range(6):
            if a["bc"][i, f] != 0:
                continue
            for k in range(max(1, a["lfc_neigs"][i, f])):
                rows.append(i)
                cols.append(a["neigtr"][i, f, k])
    g = coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(m.n, m.n)).tocsr()
    assert (g != g.T).nnz == 0
    return g


def layers_by_bfs(g, marked, inter_max):
    SI = np.zeros(g.shape[0], np.int32)
    if not marked.any():
        return SI
    d = shortest_path(g, unweighted=True, indices=np.flatnonzero(marked)).min(axis=0)
    ok = np.isfinite(d) & (1 + d < inter_max)
    SI[ok] = 1 + d[ok].astype(np.int32)
    return SI


def state_of(p, alpha0, alpha1=None, rho=1.0, pres=1.0, u=(0.0, 0.0, 0.0)):
    """Cells of the given fractions at one density and pressure (ideal gas, gamma 1.4)."""
    n = len(alpha0)
    nm = p.nmat
    al = np.zeros((n, nm))
    al[:, 0] = alpha0
    if nm == 3:
        al[:, 1] = alpha1 if alpha1 is not None else 0.3 * (1 - alpha0)
        al[:, 2] = 1.0 - al[:, 0] - al[:, 1]
    else:
        al[:, 1] = 1.0 - al[:, 0]
    g = np.array(p.gamma[:nm])
    rho = np.broadcast_to(np.asarray(rho, float), (n,))
    pres = np.broadcast_to(np.asarray(pres, float), (n,))
    u = np.broadcast_to(np.asarray(u, float), (n, 3))
    S = np.zeros((n, F.neq(nm)))
    ar = al * rho[:, None]
    are = al * pres[:, None] / (g - 1.0)
    S[:, :nm] = ar
    S[:, nm:nm + 3] = rho[:, None] * u
    S[:, nm + 3] = are.sum(1) + 0.5 * rho * (u * u).sum(1)
    S[:, nm + 4:2 * nm + 3] = al[:, :nm - 1]
    S[:, 2 * nm + 3:] = are
    return S


def params(nmat=2):
    return F.make_params(nmat=nmat, gamma=(1.4,) * nmat, pinf=(0.0,) * nmat, amin=1e-8, avf_lim=0.5)


def crit_layers(lrs0, lrs1=0):
    return F.make_crit(inter0_on=1, inter0_lvl=6, inter0_lrs=lrs0, inter1_on=1 if lrs1 else 0, inter1_lvl=5,
                       inter1_lrs=lrs1)


def inter_max(c):
    return c.inter0_lrs + c.inter1_lrs if c.inter1_on == 1 else c.inter0_lrs


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("lrs", [(4, 0), (3, 5)])
def test_interface_layers_are_breadth_first_layers_on_amr_forests(seed, lrs):
    p = params(2)
    m = random_amr_mesh(p, base_level=3, max_level=5, seed=seed, frac=0.3, passes=2)
    c = m.centres()
    rng = np.random.default_rng(seed)
    centre = rng.uniform(0.3, 0.7, size=3)
    r = np.linalg.norm(c - centre, axis=1)
    alpha = np.where(r < rng.uniform(0.15, 0.3), 0.9, 0.1)
    S = state_of(p, alpha)
    crit = crit_layers(*lrs)
    SI = O.OracleMesh(m).interface_location(p, crit, S)
    g = adjacency(m)
    phi = alpha - 0.5
    gc = g.tocoo()
    marked = np.zeros(m.n, bool)
    marked[gc.row[phi[gc.row] * phi[gc.col] < 0.0]] = True
    assert marked.any()
    want = layers_by_bfs(g, marked, inter_max(crit))
    assert np.array_equal(SI[:, 0], want)
    assert not SI[:, 1].any()


@pytest.mark.parametrize("N,lrs", [(16, 6), (32, 12)])
def test_a_plane_gives_layers_by_cell_distance(N, lrs):
    p = params(2)
    level = int(np.log2(N)) + 1
    m = F.Mesh.uniform((1, 1, 1), (0.0, 0.0, 0.0), 1.0, level, [4] * 6)
    c = m.centres()
    i = np.floor(c[:, 0] * N).astype(int)
    S = state_of(p, np.where(i < N // 2, 0.8, 0.2))
    crit = crit_layers(lrs)
    SI = O.OracleMesh(m).interface_location(p, crit, S)
    d = np.where(i < N // 2, N // 2 - 1 - i, i - N // 2)
    want = np.where(1 + d < lrs, 1 + d, 0)
    assert np.array_equal(SI[:, 0], want)


@pytest.mark.parametrize("which", [0, 1])
def test_three_materials_layer_each_indicator(which):
    """With one fraction crossing AVF_LIM and the other not, the other indicator stays 0 and the
    crossing one is breadth-first layered (SI[1] follows the second fraction)."""
    p = params(3)
    m = random_amr_mesh(p, base_level=3, max_level=5, seed=4 + which, frac=0.3, passes=2)
    c = m.centres()
    cross = np.where(c[:, 1] + 0.3 * c[:, 2] < 0.55, 0.7, 0.2)
    if which == 0:
        a0, a1 = cross, np.full(m.n, 0.1)
    else:
        a0, a1 = np.full(m.n, 0.1), cross
    S = state_of(p, a0, a1)
    crit = crit_layers(5)
    SI = O.OracleMesh(m).interface_location(p, crit, S)
    g = adjacency(m).tocoo()
    phi = cross - 0.5
    marked = np.zeros(m.n, bool)
    marked[g.row[phi[g.row] * phi[g.col] < 0.0]] = True
    want = layers_by_bfs(g.tocsr(), marked, 5)
    assert np.array_equal(SI[:, which], want)
    assert not SI[:, 1 - which].any()


# ---------------------------------------------------------------- criterionsplt

N = 16


def plane_mesh():
    return F.Mesh.uniform((1, 1, 1), (0.0, 0.0, 0.0), 1.0, 5, [4] * 6)


def near_plane(m):
    i = np.floor(m.centres()[:, 0] * N).astype(int)
    return (i == N // 2 - 1) | (i == N // 2), i < N // 2


def split_of(m, S, **kw):
    p = params(2)
    crit = F.make_crit(uniform_on=1, uniform_lvl=2, **kw)
    return O.OracleMesh(m).criterionsplt(p, crit, S, np.zeros((m.n, 2), np.int32))


def test_density_jump_marks_the_two_columns_at_the_plane():
    m = plane_mesh()
    near, left = near_plane(m)
    S = state_of(params(2), np.full(m.n, 0.3), rho=np.where(left, 1.0, 0.125))
    ratio = (1.0 - 0.125) / 1.0
    for thr, fires in ((0.5, True), (ratio, True), (np.nextafter(ratio, 1.0), False)):
        split = split_of(m, S, rho_on=1, rho_lvl=6, rho_thr=thr)
        want = np.where(near & fires, 6, 2)
        assert np.array_equal(split, want), thr


def test_pressure_jump_is_the_mixture_pressure_ratio():
    m = plane_mesh()
    near, left = near_plane(m)
    S = state_of(params(2), np.full(m.n, 0.3), pres=np.where(left, 2.0, 1.5))
    assert np.array_equal(split_of(m, S, prs_on=1, prs_lvl=7, prs_thr=0.25), np.where(near, 7, 2))
    assert np.array_equal(split_of(m, S, prs_on=1, prs_lvl=7, prs_thr=0.2500001), np.full(m.n, 2))


def test_a_later_criterion_overwrites_an_earlier_one():
    m = plane_mesh()
    near, left = near_plane(m)
    S = state_of(params(2), np.full(m.n, 0.3), rho=np.where(left, 1.0, 0.5), pres=np.where(left, 2.0, 1.0))
    split = split_of(m, S, prs_on=1, prs_lvl=7, prs_thr=0.1, rho_on=1, rho_lvl=4, rho_thr=0.1)
    assert np.array_equal(split, np.where(near, 4, 2))   # not max(7, 4)


def test_the_velocity_criterion_reads_the_density_jump():
    m = plane_mesh()
    near, left = near_plane(m)
    p = params(2)
    # a velocity jump at uniform density does not fire it
    S = state_of(p, np.full(m.n, 0.3), u=np.where(left[:, None], [[1.0, 0, 0]], [[3.0, 0, 0]]))
    assert np.array_equal(split_of(m, S, rho_on=1, rho_lvl=3, rho_thr=0.1, vel_on=1, vel_lvl=5, vel_thr=0.1),
                          np.full(m.n, 2))
    # a density jump at uniform velocity does
    S = state_of(p, np.full(m.n, 0.3), rho=np.where(left, 1.0, 0.5))
    assert np.array_equal(split_of(m, S, rho_on=1, rho_lvl=3, rho_thr=0.1, vel_on=1, vel_lvl=5, vel_thr=0.1),
                          np.where(near, 5, 2))


def test_interface_layers_give_their_levels():
    m = plane_mesh()
    p = params(2)
    S = state_of(p, np.full(m.n, 0.3))
    rng = np.random.default_rng(0)
    SI = rng.integers(0, 9, size=(m.n, 2)).astype(np.int32)
    crit = F.make_crit(uniform_on=1, uniform_lvl=2, inter0_on=1, inter0_lvl=6, inter0_lrs=3, inter1_on=1,
                       inter1_lvl=4, inter1_lrs=6)
    split = O.OracleMesh(m).criterionsplt(p, crit, S, SI)
    lvl = np.full((m.n, 2), 2)
    lvl[(SI >= 1) & (SI < 3)] = 6
    lvl[(SI >= 3) & (SI < 6)] = 4
    assert np.array_equal(split, lvl.max(axis=1))
