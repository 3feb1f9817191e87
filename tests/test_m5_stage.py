"""The whole model-5 stage: MUSCL face states (m5_faceinterp.c,
m5_reconctruct.c), the residual (m5_residual.c) with its HLLC (m5_hllc.c,
m5_wave_speeds.c, m5_flux.c) and GFM face fluxes, m5_RK and the relaxation
(m5_relaxationvec.c, m5_relaxation_solid.c).

CPU tests pin the restatement in oracle/fv_oracle_m5.c:
  * the relaxation equals the pinned model-4 relaxation on fluid states for
    PRELAX 1, 2 and 3 (to rounding: the phasic pressure is formed
    differently), carries A by the solid-fraction ratio, and with a
    deformed solid keeps mass, momentum and energy and sets sum(alpha rho e)
    to rho E - rho |u|^2 / 2;
  * the MUSCL face states equal a brute-force minmod on lines of cells;
  * the face flux is the physical flux on supersonic faces, conserves mass and
    momentum between equal states, switches to the GFM flux exactly where phi
    changes sign, carries no mass through a wall and averages the mirrored
    fluxes at a symmetry face;
  * the residual of a resting state has the closed form of the reference's
    star-state energy, a mirrored face takes the negated flux of the earlier
    leaf (not its own two-sided value: the model-5 flux is not antisymmetric), and
    on a uniform mesh the mass residuals sum to the wall fluxes (zero);
  * a stage is the composition of the pinned operators.
GPU tests compare fv_m5_dev_* with the oracle.
"""
import math

import numpy as np
import pytest

import forestfv_public_amd as F
import oracle_lib as O
from test_m5 import FLUIDS, RHO, SOLID_WATER_AIR, fluid_side, frame_of, layout, m5_state, near_identity, solid_side


def m5p(table=SOLID_WATER_AIR, nmat=3, dt=2.0 ** -30, prelax=1, order=2, rk_steps=3):
    return F.make_m5_params(nmat=nmat, gamma=table["gamma"][:nmat], pinf=table["pinf"][:nmat],
                            mush=table["mush"][:nmat], amin=1e-6, avf_lim=0.5, dt=dt, prelax=prelax, order=order,
                            rk_steps=rk_steps)


def m4_twin(p, prelax):
    """The model-4 parameters of the same two fluids (m4_relaxationvec.c is pinned in test_oracle.py)."""
    return F.make_params(nmat=2, gamma=tuple(p.gamma[:2]), pinf=tuple(p.pinf[:2]), amin=p.amin, avf_lim=p.avf_lim,
                         prelax=prelax)


def unrelaxed(p, rng, a0):
    """A two-material state whose phasic pressures differ by up to a factor 20."""
    x = m5_state(p, (a0, 1.0 - a0), RHO[:2], 1e6, rng.normal(scale=20.0, size=3), near_identity(rng))
    i1, i2, i3, i4, i5 = layout(2)
    x[i4] *= rng.uniform(0.9, 1.2)
    x[i4 + 1] *= rng.uniform(0.5, 20.0)
    return x


# ---------------------------------------------------------------- relaxation

@pytest.mark.parametrize("prelax", [1, 2, 3])
def test_m5_relaxation_of_fluids_is_the_model4_relaxation(prelax):
    p = m5p(FLUIDS, nmat=2, prelax=prelax)
    q = m4_twin(p, prelax)
    rng = np.random.default_rng(prelax)
    i1, i2, i3, i4, i5 = layout(2)
    for trial in range(40):
        x = unrelaxed(p, rng, rng.uniform(0.05, 0.95))
        z = O.relax(q, x[:9])
        if np.isnan(z).any():
            # the damped Newton of PRELAX 2 from pe = 0.001 diverged: the reference exits, so does model 5
            with pytest.raises(FloatingPointError):
                O.m5_relax(p, x)
            continue
        y = O.m5_relax(p, x)
        # the same equations; model 5 forms p0_k as (g-1)(are-e_el)/alpha - g pinf where model 4
        # has (are (g-1) - alpha g pinf)/alpha, so the two agree to rounding
        assert np.allclose(y[:9], z, rtol=1e-11, atol=0), f"trial {trial}"
        # A is carried by the ratio of the limited solid fractions (m5_relaxationvec.c:153-155, :555-558)
        assert np.array_equal(y[i5:], x[i5:] * (y[i3] / x[i3]))


@pytest.mark.parametrize("prelax", [1, 2, 3])
def test_m5_relaxation_of_an_undeformed_solid_is_the_fluid_relaxation(prelax):
    """With A = I the elastic energy and deviatoric stress vanish exactly, so MATERMUSH does not matter."""
    ps = m5p(SOLID_WATER_AIR, nmat=2, prelax=prelax)
    pf = m5p(FLUIDS, nmat=2, prelax=prelax)
    rng = np.random.default_rng(10 + prelax)
    i1, i2, i3, i4, i5 = layout(2)
    for _ in range(20):
        x = unrelaxed(pf, rng, rng.uniform(0.05, 0.95))
        x[i5:] = np.eye(3).ravel()
        assert np.array_equal(O.m5_relax(ps, x), O.m5_relax(pf, x))


@pytest.mark.parametrize("prelax", [1, 2, 3])
@pytest.mark.parametrize("nmat", [2, 3])
def test_m5_relaxation_keeps_the_conserved_state_and_the_internal_energy(prelax, nmat):
    if prelax == 1 and nmat == 3:
        pytest.skip("PRELAX 1 is the two-phase closed form (m5_relaxationvec.c:103-110)")
    p = m5p(SOLID_WATER_AIR, nmat=nmat, prelax=prelax)
    rng = np.random.default_rng(20 + prelax + nmat)
    i1, i2, i3, i4, i5 = layout(nmat)
    for _ in range(30):
        x = solid_side(p, rng) if rng.random() < 0.5 else fluid_side(p, rng)
        x[i4:i5] *= rng.uniform(0.95, 1.05, size=nmat)
        y = O.m5_relax(p, x)
        assert np.array_equal(y[:i3], x[:i3])                    # alpha rho_k, rho u, rho E untouched
        r = x[:nmat].sum()
        ke = 0.5 * (x[i1:i1 + 3] @ x[i1:i1 + 3]) / r
        assert math.fsum(y[i4:i5]) == pytest.approx(x[i2] - ke, rel=1e-12)
        assert np.all(y[i3:i4] > 0.0) and y[i3:i4].sum() < 1.0


def test_m5_relaxation_prelax0_only_limits_the_fractions():
    p = m5p(SOLID_WATER_AIR, nmat=3, prelax=0)
    rng = np.random.default_rng(30)
    i1, i2, i3, i4, i5 = layout(3)
    x = solid_side(p, rng)
    y = O.m5_relax(p, x)
    assert np.array_equal(y, x)
    x[i3] = -0.5
    y = O.m5_relax(p, x)
    assert y[i3] == p.amin / 100.0 and np.array_equal(np.delete(y, i3), np.delete(x, i3))


# ---------------------------------------------------------------- MUSCL

def test_m5_face_states_are_minmod_on_lines_of_cells():
    """PRELAX 0 leaves the face state as reconstructed; a field that varies along x only has
    zero deltas across y and z faces, so each x line is a 1-D MUSCL problem."""
    level = 3
    L = 2.0 ** -4
    m = F.Mesh.uniform((1, 1, 1), (0.0, 0.0, 0.0), L, level, [4] * 6)
    om = O.OracleMesh(m)
    p = m5p(SOLID_WATER_AIR, nmat=3, prelax=0)
    rng = np.random.default_rng(31)
    nq = F.m5_neq(3)
    i1, i2, i3, i4, i5 = layout(3)
    nx = 1 << (level - 1)
    h = L / nx
    ci = m.arrays()["ci"].reshape(-1, 3)
    column = [solid_side(p, rng) if rng.random() < 0.5 else fluid_side(p, rng) for _ in range(nx)]
    S = np.array([column[c[0]] for c in ci])
    SF = om.m5_faceinterp(p, S)
    lim = np.ones(nq)
    lim[i4:i5] = 0.0
    index = {tuple(c): i for i, c in enumerate(ci)}
    checked = 0
    for i, c in enumerate(ci):
        for f, d in ((1, 1), (3, -1)):              # x+ and x- (operatetp.c:60-69)
            front, back = (c[0] + d, c[1], c[2]), (c[0] - d, c[1], c[2])
            if front not in index:
                want = S[i]                          # boundary face: no slope
            else:
                df = S[index[front]] - S[i]
                db = (S[index[back]] - S[i]) if back in index else np.zeros(nq)
                a, b = -db / h, df / h
                g = np.where(a * b > 0.0, np.where(np.abs(a) > np.abs(b), lim * b, lim * a), 0.0)
                g = np.where(np.abs(df) > 1e-10, g, 0.0)
                want = S[i] + (h / 2.0) * g
            assert np.array_equal(SF[i, f], want), (i, f)
            checked += 1
        for f in (0, 2, 4, 5):
            assert np.array_equal(SF[i, f], S[i])
    assert checked == 2 * m.n


def test_m5_face_states_average_four_finer_neighbours():
    """A 1:4 face sees the mean of its four finer neighbours (m5_facevalues_MUSCL.c:20-33)."""
    from cases import random_amr_mesh
    base = F.make_params(bc=(4, 4, 4, 4, 4, 4))
    m = random_amr_mesh(base, base_level=2, max_level=4, seed=3, frac=0.3, passes=2, L=2.0 ** -4)
    om = O.OracleMesh(m)
    p = m5p(FLUIDS, nmat=3, prelax=0)
    rng = np.random.default_rng(32)
    nq = F.m5_neq(3)
    S = np.array([fluid_side(p, rng, u=np.zeros(3)) for _ in range(m.n)])
    S[:, 0] = rng.uniform(1.0, 2.0, size=m.n)     # only alpha rho_0 varies
    SF = om.m5_faceinterp(p, S)
    a = m.arrays()
    lfc = a["lfc_neigs"].reshape(-1, 6)
    nbr = a["neigtr"].reshape(-1, 6, 4)
    found = 0
    for i in range(m.n):
        for f in range(6):
            if lfc[i, f] == 4:
                fo = (2, 3, 0, 1, 5, 4)[f]
                X = sum(S[j, 0] / 4.0 for j in nbr[i, f])
                df = X - S[i, 0]
                cl = om.cells[i]
                if lfc[i, fo] == 0:
                    continue
                back = (S[nbr[i, fo, 0], 0] if lfc[i, fo] == 1 else sum(S[j, 0] / 4.0 for j in nbr[i, fo])) - S[i, 0]
                a_, b_ = -back / cl.dist_cc[fo], df / cl.dist_cc[f]
                g = (b_ if abs(a_) > abs(b_) else a_) if a_ * b_ > 0.0 else 0.0
                g = g if abs(df) > 1e-10 else 0.0
                assert SF[i, f, 0] == S[i, 0] + cl.dist_cf[f] * g
                found += 1
    assert found > 10


# ---------------------------------------------------------------- face flux

def m5_physical_flux(p, x, n):
    """F(x).n of m5_flux.c:10-95 (flag 0): +st in the momentum flux, -A carried."""
    nm = p.nmat
    i1, i2, i3, i4, i5 = layout(nm)
    rec = O.m5_record(p, x)
    r, e, st = rec["r"], rec["e"], rec["st"]
    u = np.array([rec["u"], rec["v"], rec["w"]])
    un = u @ n
    f = np.zeros(F.m5_neq(nm))
    f[0:nm] = rec["ra"] * un
    f[i1:i1 + 3] = r * u * un + st @ n
    f[i2] = r * e * un - n @ (st @ u)
    f[i3:i4] = rec["avf"][:nm - 1] * un
    f[i4:i5] = rec["are"] * un
    f[i5:] = -rec["Amat"].ravel() * un
    return f


@pytest.mark.parametrize("nmat", [2, 3])
def test_m5_supersonic_faces_are_the_physical_flux(nmat):
    p = m5p(SOLID_WATER_AIR, nmat=nmat)
    rng = np.random.default_rng(40 + nmat)
    for trial in range(20):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        fr = frame_of(n)
        side = solid_side if trial % 2 else fluid_side
        xL = side(p, rng, u=3000.0 * n)
        xR = side(p, rng, u=3000.0 * n)
        flux, unc, g = O.m5_face(p, xL, xL, xR, xR, fr, 0, 1.0)
        assert g == 0
        want = m5_physical_flux(p, xL, fr[:3])
        assert np.allclose(flux, want, rtol=1e-10, atol=1e-9 * np.abs(want).max())
        rec = O.m5_record(p, xL)
        assert np.array_equal(unc, [rec["u"], rec["v"], rec["w"]])   # flag 0: the left velocity


def test_m5_equal_states_give_the_physical_mass_and_momentum_flux():
    """Subsonic face between equal fluid states: the star state keeps the partial densities and the
    momentum, so the mass and momentum fluxes are consistent (the energy is not, see the next test)."""
    p = m5p(FLUIDS, nmat=3)
    rng = np.random.default_rng(50)
    for _ in range(20):
        n = rng.normal(size=3)
        fr = frame_of(n)
        x = fluid_side(p, rng)
        flux, unc, g = O.m5_face(p, x, x, x, x, fr, 0, 1.0)
        want = m5_physical_flux(p, x, fr[:3])
        assert np.allclose(flux[:6], want[:6], rtol=1e-9, atol=1e-9 * np.abs(want[:6]).max())


def test_m5_resting_face_energy_is_the_reference_star_energy():
    """At rest the star energy of m5_star_region.c:69 divides by r*(u* + S_K): E* = -E, so the
    HLLC energy flux is S_L (E* - E) = -2 S_L rho E with S_L = -c."""
    p = m5p(FLUIDS, nmat=3)
    rng = np.random.default_rng(51)
    i1, i2, i3, i4, i5 = layout(3)
    for _ in range(10):
        fr = frame_of(rng.normal(size=3))
        x = fluid_side(p, rng, u=np.zeros(3))
        flux, unc, g = O.m5_face(p, x, x, x, x, fr, 0, 2.0)
        rec = O.m5_record(p, x)
        c = math.sqrt(sum(rec["c2Y"]))
        assert flux[i2] == pytest.approx(2.0 * (2.0 * c * x[i2]), rel=1e-12)
        assert np.allclose(unc, 0.0, atol=1e-12 * c)


@pytest.mark.parametrize("nmat", [2, 3])
def test_m5_face_switches_to_ghost_fluid_where_phi_changes_sign(nmat):
    p = m5p(SOLID_WATER_AIR, nmat=nmat)
    rng = np.random.default_rng(60 + nmat)
    for trial in range(20):
        fr = frame_of(rng.normal(size=3))
        xL = solid_side(p, rng)
        xR = fluid_side(p, rng)
        if trial % 2:
            xL, xR = xR, xL
        flux, unc, g = O.m5_face(p, xL, xL, xR, xR, fr, 0, 0.5)
        assert g == 1
        want, wunc = O.m5_gfm_face(p, xL, xR, fr, 0.5)
        assert np.array_equal(flux, want) and np.array_equal(unc, wunc)
        # phi is read from the cell states: equal-sign cells give the plain HLLC flux
        flux2, _, g2 = O.m5_face(p, xL, xL, xR, xL, fr, 0, 0.5)
        assert g2 == 0


def test_m5_wall_face_carries_no_mass():
    p = m5p(SOLID_WATER_AIR, nmat=3)
    rng = np.random.default_rng(70)
    for _ in range(20):
        fr = frame_of(rng.normal(size=3))
        x = solid_side(p, rng, u=rng.normal(scale=5.0, size=3))
        flux, unc, g = O.m5_face(p, x, x, x, x, fr, 4, 1.0)
        assert np.abs(flux[:3]).max() <= 1e-12 * np.abs(x[:3]).max() * 1e4
        assert abs(unc @ fr[:3]) <= 1e-12 * np.abs(unc).max() + 1e-300


def test_m5_symmetry_face_is_the_mean_of_the_mirrored_physical_fluxes():
    p = m5p(SOLID_WATER_AIR, nmat=3)
    rng = np.random.default_rng(71)
    i1, i2, i3, i4, i5 = layout(3)
    for _ in range(10):
        fr = frame_of(rng.normal(size=3))
        n = fr[:3]
        C = solid_side(p, rng)
        X = solid_side(p, rng)           # the face state is not used by the symmetry flux
        flux, unc, g = O.m5_face(p, X, C, X, C, fr, 1, 1.0)
        rec = O.m5_record(p, C)
        u = np.array([rec["u"], rec["v"], rec["w"]])
        um = u - 2.0 * (u @ n) * n
        uc = 0.5 * (u + um)
        assert np.allclose(unc, uc, rtol=1e-12, atol=1e-12 * np.abs(u).max())
        r, e, st = rec["r"], rec["e"], rec["st"]
        mom = 0.5 * (r * u * (u @ n) + r * um * (um @ n)) + st @ n
        assert np.abs(flux[:3]).max() <= 1e-9 * np.abs(C[:3]).max()
        assert np.allclose(flux[i1:i1 + 3], mom, rtol=1e-10, atol=1e-10 * np.abs(mom).max())
        assert np.allclose(flux[i5:], rec["Amat"].ravel() * (uc @ n), atol=1e-10)


# ---------------------------------------------------------------- residual and stage

def smooth_field(p, m, rng, L):
    """A solid sphere (solid fraction 0.95) in water with a little air, smeared over a few cells so
    that phi = alpha_solid - 0.5 changes sign on a shell of faces; pressure, velocity and the
    deformation A vary smoothly with +-1 % perturbations."""
    nm = p.nmat
    c = m.centres() / L
    k = rng.uniform(1.0, 3.0, size=(4, 3))
    ph = rng.uniform(0, 2 * np.pi, size=4)
    wave = [np.sin(2 * np.pi * (c @ k[q]) + ph[q]) for q in range(4)]
    r = np.linalg.norm(c - np.array([0.45, 0.52, 0.48]), axis=1)
    a0 = 0.5 + 0.45 * np.tanh((0.3 - r) / 0.06)
    S = np.zeros((m.n, F.m5_neq(nm)))
    for i in range(m.n):
        if nm == 3:
            alpha = (a0[i], 0.9 * (1.0 - a0[i]), 0.1 * (1.0 - a0[i]))
        else:
            alpha = (a0[i], 1.0 - a0[i])
        u = 5.0 * np.array([wave[0][i], wave[1][i], wave[2][i]])
        A = np.eye(3) + 0.01 * np.outer([wave[1][i], wave[2][i], wave[3][i]], [1.0, -0.5, 0.25])
        S[i] = m5_state(p, alpha, RHO[:nm], 1e7 * (1.0 + 0.01 * wave[3][i]), u, A)
    return S


def m5_case(seed, amr=True, nmat=3, table=SOLID_WATER_AIR, bc=(4, 4, 4, 4, 4, 4), prelax=3, order=2, L=2.0 ** -6,
            interface_frac=0.2):
    from cases import random_amr_mesh
    rng = np.random.default_rng(seed)
    base = F.make_params(bc=bc)
    if amr:
        m = random_amr_mesh(base, base_level=3, max_level=5, seed=seed, frac=0.25, passes=2, L=L)
    else:
        m = F.Mesh.uniform((1, 1, 1), (0.0, 0.0, 0.0), L, 4, list(bc))
    p = m5p(table, nmat=nmat, prelax=prelax, order=order)
    S = smooth_field(p, m, rng, L)
    # interface leaves (SI[0] != 0 keeps a leaf from reusing its neighbour's flux)
    SI = np.zeros((m.n, 2), np.int32)
    SI[:, 0] = (rng.random(m.n) < interface_frac).astype(np.int32)
    m.set_SI(SI)
    om = O.OracleMesh(m)
    return p, m, om, S


def test_m5_resting_residual_has_the_closed_form_energy():
    p = m5p(FLUIDS, nmat=3)
    m = F.Mesh.uniform((1, 1, 1), (0.0, 0.0, 0.0), 2.0 ** -4, 3, [4] * 6)
    om = O.OracleMesh(m)
    rng = np.random.default_rng(80)
    x = fluid_side(p, rng, u=np.zeros(3))
    S = np.tile(x, (m.n, 1))
    SF = om.m5_faceinterp(p, S)
    RHS, gf = om.m5_residual(p, S, SF)
    i1, i2, i3, i4, i5 = layout(3)
    rec = O.m5_record(p, x)
    c = math.sqrt(sum(rec["c2Y"]))
    A = om.cells[0].Area[0][0]
    assert not gf.any()
    # every computed face (walls included, u* = 0 there) adds A 2 c rho E; a mirrored face takes the
    # negation of its earlier neighbour's value (m5_residual.c:82-87), which cancels one computed face
    a = m.arrays()
    bc = a["bc"].reshape(-1, 6)
    nb = a["neigtr"].reshape(-1, 6, 4)[:, :, 0]
    units = np.where(bc != 0, 1, np.where(nb < np.arange(m.n)[:, None], -1, 1)).sum(axis=1)
    assert np.allclose(RHS[:, i2], units * (A * (2.0 * c * x[i2])), rtol=1e-12, atol=1e-12 * A * c * x[i2])
    assert (units > 0).any() and (units == 0).any()
    assert np.abs(RHS[:, i1:i1 + 3]).max() <= 1e-12 * A * np.abs(rec["st"]).max() * 10


def cell_frame(cl, f):
    return np.array([cl.nx[f], cl.ny[f], cl.nz[f], cl.nxt1[f], cl.nyt1[f], cl.nzt1[f],
                     cl.nxt2[f], cl.nyt2[f], cl.nzt2[f]])


def test_m5_mirrored_faces_take_the_negated_earlier_flux():
    """Model 5's face flux is not antisymmetric (the star energy, the -A transport), so a mirrored
    face differs from its own two-sided evaluation: on the conservative rows, which carry no source
    terms, the residual with mirrors minus the residual without is the sum over mirrored faces of
    -F(earlier leaf) - F(own evaluation), each from orc_m5_face."""
    p, m, om, S = m5_case(81, table=FLUIDS)
    SF = om.m5_faceinterp(p, S)
    a = m.arrays()
    SI = a["SI"].reshape(-1, 2)
    bc = a["bc"].reshape(-1, 6)
    lfc = a["lfc_neigs"].reshape(-1, 6)
    nb = a["neigtr"].reshape(-1, 6, 4)[:, :, 0]
    RHS_m, _ = om.m5_residual(p, S, SF)
    om1 = O.OracleMesh(m, SI=np.ones((m.n, 2), np.int32))
    RHS_2, _ = om1.m5_residual(p, S, SF)
    nc = p.nmat + 4
    FO = (2, 3, 0, 1, 5, 4)
    want = np.zeros((m.n, nc))
    mirrored = 0
    for i in range(m.n):
        for f in range(6):
            j = nb[i, f]
            if bc[i, f] or lfc[i, f] != 1 or lfc[j, FO[f]] != 1 or SI[i, 0] or SI[j, 0] or not j < i:
                continue
            ci, cj = om.cells[i], om.cells[j]
            Fi, _, _ = O.m5_face(p, SF[i, f], S[i], SF[j, FO[f]], S[j], cell_frame(ci, f), 0, ci.Area[f][0])
            Fj, _, _ = O.m5_face(p, SF[j, FO[f]], S[j], SF[i, f], S[i], cell_frame(cj, FO[f]), 0, cj.Area[FO[f]][0])
            want[i] += (-Fj - Fi)[:nc]
            mirrored += 1
    scale = np.abs(RHS_2[:, :nc]).max(axis=0)
    assert mirrored > 100
    assert np.allclose((RHS_m[:, :nc] - RHS_2[:, :nc]) / scale, want / scale, rtol=0, atol=1e-12)
    assert np.abs(want).max() > 0


def test_m5_mass_residuals_sum_to_the_wall_fluxes():
    """Uniform mesh, no interface leaves: every interior face is computed once and negated, so the
    mass residuals cancel up to the reflective-wall faces, which carry no mass."""
    p = m5p(FLUIDS, nmat=3, prelax=3)
    m = F.Mesh.uniform((1, 1, 1), (0.0, 0.0, 0.0), 2.0 ** -4, 4, [4] * 6)
    om = O.OracleMesh(m)
    rng = np.random.default_rng(82)
    S = smooth_field(p, m, rng, 2.0 ** -4)
    SF = om.m5_faceinterp(p, S)
    RHS, _ = om.m5_residual(p, S, SF)
    for k in range(3):
        tot = math.fsum(RHS[:, k])
        assert abs(tot) <= 1e-12 * np.abs(RHS[:, k]).sum()


@pytest.mark.parametrize("irk", [1, 2, 3])
def test_m5_stage_is_the_composition_of_its_operators(irk):
    p, m, om, S = m5_case(83 + irk)
    rng = np.random.default_rng(90 + irk)
    SN = S.copy()
    Sin = S if irk == 1 else S * (1.0 + 1e-4 * rng.normal(size=S.shape))
    if irk != 1:
        i1, i2, i3, i4, i5 = layout(3)
        Sin = np.array([O.m5_relax(p, x) for x in Sin])
    S1, SF, RHS, gf = om.m5_stage(p, irk, SN, Sin)
    SF2 = om.m5_faceinterp(p, Sin)
    R2, g2 = om.m5_residual(p, Sin, SF2)
    _, Srk = om.m5_rk(p, irk, SN, Sin, R2, g2)
    Srel = np.array([O.m5_relax(p, x) for x in Srk])
    Sfin = om.m5_gfm_rel(p, Srk, Srel, g2)
    assert np.array_equal(SF, SF2) and np.array_equal(RHS, R2) and np.array_equal(gf, g2)
    assert np.array_equal(S1, Sfin)
    assert gf.any()


def test_m5_first_order_residual_reads_cell_states():
    p, m, om, S = m5_case(87, order=1)
    RHS1, g1 = om.m5_residual(p, S, np.zeros((m.n, 6, F.m5_neq(3))))
    # ORDER 1 with face states equal to the cell states is the ORDER 2 residual
    p2 = m5p(SOLID_WATER_AIR, nmat=3, order=2)
    RHS2, g2 = om.m5_residual(p2, S, np.repeat(S[:, None, :], 6, axis=1))
    assert np.array_equal(RHS1, RHS2) and np.array_equal(g1, g2)


def test_m5_advance_runs_steps_of_stages():
    p, m, om, S = m5_case(88, table=FLUIDS)
    p.dt = 1e-12
    S2 = om.m5_advance(p, S, 1)
    SN = S.copy()
    Sx = S.copy()
    for irk in (1, 2, 3):
        Sx, _, _, _ = om.m5_stage(p, irk, SN, Sx)
    assert np.array_equal(S2, Sx)


# ---------------------------------------------------------------- device parity (GPU)

# (table, nmat, prelax, bitwise): without a solid and without PRELAX 3 no pow is evaluated, so the
# device and the restatement agree bit for bit; the elastic EOS and PRELAX 3 go through pow, where
# the device and glibc may differ in the last bit
DEVICE_CASES = {
    "fluids2_prelax1": (FLUIDS, 2, 1, True),
    "fluids3_prelax0": (FLUIDS, 3, 0, True),
    "fluids3_prelax2": (FLUIDS, 3, 2, True),
    "solid3_prelax3": (SOLID_WATER_AIR, 3, 3, False),
    "solid2_prelax0": (SOLID_WATER_AIR, 2, 0, False),
}


def say(*a):
    print(*a, flush=True)


def assert_close(got, want, bitwise, what, rel=1e-11):
    if bitwise:
        assert np.array_equal(got, want), f"{what}: max diff {np.abs(got - want).max()}"
        return
    g = got.reshape(got.shape[0], -1)
    w = want.reshape(want.shape[0], -1)
    scale = np.abs(w).max(axis=0) + 1e-300
    err = (np.abs(g - w).max(axis=0) / scale).max()
    assert err <= rel, f"{what}: relative error {err}"


def rel_l1(got, want):
    num = np.abs(got - want).sum(axis=0)
    den = np.abs(want).sum(axis=0) + 1e-300
    return (num / den).max()


@pytest.mark.gpu
@pytest.mark.parametrize("amr", [False, True])
@pytest.mark.parametrize("name", list(DEVICE_CASES))
def test_m5_device_face_states_and_residual_match_oracle(name, amr):
    table, nmat, prelax, bitwise = DEVICE_CASES[name]
    p, m, om, S = m5_case(100 + nmat + prelax, amr=amr, nmat=nmat, table=table, prelax=prelax)
    say(f"[m5 dev] {name} amr={amr} n={m.n}")
    dev = F.M5Device(p, m, prec=F.PREC_EXACT)
    dev.set_state(S)
    dev.faceinterp()
    dev.residual()
    SF, RHS, gf = dev.stage_arrays()
    SFo = om.m5_faceinterp(p, S)
    assert_close(SF, SFo, bitwise, "face states")
    RHSo, gfo = om.m5_residual(p, S, SFo)
    assert np.array_equal(gf, gfo)
    assert gfo.any()
    assert_close(RHS, RHSo, bitwise, "residual")
    dev.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(DEVICE_CASES))
def test_m5_device_stages_match_oracle(name):
    table, nmat, prelax, bitwise = DEVICE_CASES[name]
    p, m, om, S = m5_case(110 + nmat + prelax, nmat=nmat, table=table, prelax=prelax)
    p.dt = 2e-9
    dev = F.M5Device(p, m, prec=F.PREC_EXACT)
    dev.set_state(S)
    dev.begin_step()
    SN = S.copy()
    So = S.copy()
    for irk in (1, 2, 3):
        say(f"[m5 dev stage] {name} n={m.n} irk={irk}")
        dev.stage(irk)
        So, SFo, RHSo, gfo = om.m5_stage(p, irk, SN, So)
        _, RHS, gf = dev.stage_arrays()
        assert np.array_equal(gf, gfo)
        assert_close(RHS, RHSo, bitwise, f"residual of stage {irk}", rel=1e-10)
        Sd = dev.get_state()
        if bitwise:
            assert np.array_equal(Sd, So), f"stage {irk}"
        else:
            assert rel_l1(Sd, So) <= 1e-12, f"stage {irk}: {rel_l1(Sd, So)}"
            So = Sd.copy()         # continue both from the same state
            dev.set_state(So)
    dev.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["fluids3_prelax0", "fluids3_prelax2", "solid3_prelax3"])
def test_m5_device_steps_match_oracle(name):
    table, nmat, prelax, bitwise = DEVICE_CASES[name]
    p, m, om, S = m5_case(120 + nmat + prelax, nmat=nmat, table=table, prelax=prelax)
    p.dt = 2e-9
    dev = F.M5Device(p, m, prec=F.PREC_EXACT)
    dev.set_state(S)
    dev.advance(3)
    Sd = dev.get_state()
    So = om.m5_advance(p, S, 3)
    say(f"[m5 dev steps] {name} n={m.n} rel L1 {rel_l1(Sd, So):.2e}")
    if bitwise:
        assert np.array_equal(Sd, So)
    else:
        assert rel_l1(Sd, So) <= 1e-10
    dev.close()


@pytest.mark.gpu
@pytest.mark.parametrize("bc", [(1, 2, 5, 4, 1, 5), (2, 1, 1, 5, 4, 2)])
def test_m5_device_boundary_codes_match_oracle(bc):
    """Symmetry faces average the mirrored cell fluxes (m5_hllc.c:66-88), walls force u* = 0 or
    mirror all velocity components, outlets copy the face state."""
    p, m, om, S = m5_case(130 + bc[0], nmat=3, table=FLUIDS, prelax=0, bc=bc)
    dev = F.M5Device(p, m, prec=F.PREC_EXACT)
    dev.set_state(S)
    dev.faceinterp()
    dev.residual()
    SF, RHS, gf = dev.stage_arrays()
    SFo = om.m5_faceinterp(p, S)
    RHSo, gfo = om.m5_residual(p, S, SFo)
    say(f"[m5 dev bc] {bc} n={m.n}")
    assert np.array_equal(SF, SFo) and np.array_equal(RHS, RHSo) and np.array_equal(gf, gfo)
    dev.close()


@pytest.mark.gpu
def test_m5_device_first_order_matches_oracle():
    p, m, om, S = m5_case(140, nmat=3, table=FLUIDS, prelax=0, order=1)
    dev = F.M5Device(p, m, prec=F.PREC_EXACT)
    dev.set_state(S)
    dev.residual()
    _, RHS, gf = dev.stage_arrays()
    RHSo, gfo = om.m5_residual(p, S, np.zeros((m.n, 6, F.m5_neq(3))))
    assert np.array_equal(RHS, RHSo) and np.array_equal(gf, gfo)
    dev.close()


@pytest.mark.gpu
def test_m5_device_and_oracle_fail_on_the_same_run():
    """The literal model-5 update is anti-diffusive (+dt RHS / V with the outward flux) and its star
    energy grows at rest; with the two-phase PRELAX 1 this case reaches NaN within three steps.
    The reference would print and exit; the oracle and the device both report it."""
    table, nmat, prelax, _ = DEVICE_CASES["fluids2_prelax1"]
    p, m, om, S = m5_case(120 + nmat + prelax, nmat=nmat, table=table, prelax=prelax)
    p.dt = 2e-9
    with pytest.raises(RuntimeError):
        om.m5_advance(p, S, 3)
    dev = F.M5Device(p, m, prec=F.PREC_EXACT)
    dev.set_state(S)
    with pytest.raises(F.FVError) as e:
        dev.advance(3)
    assert e.value.code == F.FV_ERR_NAN
    dev.close()


@pytest.mark.gpu
def test_m5_device_reports_nan_as_a_status():
    """A leaf with zero density gives 0/0 velocities; the reference prints and exits, the device
    returns FV_ERR_NAN."""
    p, m, om, S = m5_case(141, nmat=3, table=FLUIDS, prelax=0)
    S[5, :3] = 0.0
    S[5, 3:6] = 0.0
    dev = F.M5Device(p, m, prec=F.PREC_EXACT)
    dev.set_state(S)
    with pytest.raises(F.FVError) as e:
        dev.advance(1)
    assert e.value.code == F.FV_ERR_NAN
    dev.close()


@pytest.mark.gpu
@pytest.mark.parametrize("nroot,level", [((1, 1, 1), 1), ((2, 1, 1), 1), ((2, 2, 1), 3)])
def test_m5_device_single_leaf_and_root_bricks_match_oracle(nroot, level):
    """One leaf with six boundary faces, two single-leaf roots side by side, and a brick of level-3
    roots: the root-crossing faces and a stencil made only of ghost values."""
    p = m5p(FLUIDS, nmat=3, prelax=0)
    bc = (1, 2, 5, 4, 4, 1)
    m = F.Mesh.uniform(nroot, (0.0, 0.0, 0.0), 2.0 ** -6, level, list(bc))
    rng = np.random.default_rng(150 + level)
    S = smooth_field(p, m, rng, 2.0 ** -6 * nroot[0])
    om = O.OracleMesh(m)
    dev = F.M5Device(p, m, prec=F.PREC_EXACT)
    dev.set_state(S)
    dev.faceinterp()
    dev.residual()
    SF, RHS, gf = dev.stage_arrays()
    SFo = om.m5_faceinterp(p, S)
    RHSo, gfo = om.m5_residual(p, S, SFo)
    say(f"[m5 dev small] nroot={nroot} level={level} n={m.n}")
    assert np.array_equal(SF, SFo) and np.array_equal(RHS, RHSo) and np.array_equal(gf, gfo)
    p.dt = 1e-10
    dev2 = F.M5Device(p, m, prec=F.PREC_EXACT)
    dev2.set_state(S)
    dev2.advance(1)
    assert np.array_equal(dev2.get_state(), om.m5_advance(p, S, 1))
    dev.close()
    dev2.close()


@pytest.mark.gpu
@pytest.mark.parametrize("table,prelax,amr", [("fluids", 0, True), ("solid", 1, False), ("solid", 3, True)])
def test_m5_device_fp64_sources_per_leaf_equal_exact_to_rounding(table, prelax, amr):
    """The default FP64 residual forms the non-conservative terms once per leaf from the sum of
    the face velocities (linear in each face's u* n^T); the exact path adds them face by face as
    m5_residual.c:176-214 does.  The two agree to rounding per stage and over three steps."""
    tab = SOLID_WATER_AIR if table == "solid" else FLUIDS
    p, m, om, S = m5_case(170 + prelax, amr=amr, nmat=3, table=tab, prelax=prelax)
    p.fix = F.M5_FIX_SIGNS | F.M5_FIX_MIXTURE   # a runnable update for the three steps
    p.dt = 1e-9
    out = {}
    for prec in (F.PREC_EXACT, F.PREC_FP64):
        dev = F.M5Device(p, m, prec=prec)
        dev.set_state(S)
        dev.faceinterp()
        dev.residual()
        _, RHS, gf = dev.stage_arrays()
        dev.set_state(S)
        dev.advance(3)
        out[prec] = (RHS, gf, dev.get_state())
        dev.close()
    (R0, g0, S0), (R1, g1, S1) = out[F.PREC_EXACT], out[F.PREC_FP64]
    assert np.array_equal(g0, g1)
    scale = np.abs(R0).max(axis=0) + 1e-300
    assert (np.abs(R1 - R0).max(axis=0) / scale).max() <= 1e-12
    scale = np.abs(S0).max(axis=0) + 1e-300
    err = (np.abs(S1 - S0).max(axis=0) / scale).max()
    say(f"[m5 fp64 vs exact] {table} prelax={prelax}: three steps, largest difference {err:.2e}")
    assert err <= 1e-12
