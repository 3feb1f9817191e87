"""Model-5 ghost-fluid operators: the GFM face flux (m5_hllc_gfm.c and the
functions it calls), the stage-1 interface re-projection inside m5_RK
(m5_RK.c, m5_gfm.c) and the re-projection after relaxation (m5_gfm_rel.c).

The CPU tests pin the restatement in oracle/fv_oracle_m5.c against closed
forms: the elastic EOS against its matrix formula, supersonic faces against
the physical flux, the literal star-state energy at rest, rotation
covariance, the m5_RK stage tables, positivity fallback and freezes, and the
re-projection geometry (last-face overwrite in m5_gfm, summed normal in
m5_gfm_rel, leaf-order updates).  The GPU tests compare fv_m5_* with the
oracle on random inputs.
"""
import numpy as np
import pytest

import forestfv_public_amd as F
import oracle_lib as O

SOLID_WATER_AIR = dict(gamma=(4.4, 6.12, 1.4), pinf=(6e8, 3.43e8, 0.0), mush=(2.6e10, 0.0, 0.0))
FLUIDS = dict(gamma=(4.4, 6.12, 1.4), pinf=(6e8, 3.43e8, 0.0), mush=(0.0, 0.0, 0.0))
RHO = (2700.0, 1000.0, 1.2)


def m5p(table=SOLID_WATER_AIR, nmat=3, dt=2.0 ** -30, zerou=(0, 0, 0)):
    return F.make_m5_params(nmat=nmat, gamma=table["gamma"][:nmat], pinf=table["pinf"][:nmat],
                            mush=table["mush"][:nmat], amin=1e-6, avf_lim=0.5, dt=dt, zerou=zerou)


def elastic_energy(A, mu):
    """mu / (4 |G|^(1/6)) tr((G - |G|^(1/3) I)^2), G = A A^T (state_equation.c:269-316)."""
    G = A @ A.T
    d = abs(np.linalg.det(G))
    M = G - d ** (1.0 / 3.0) * np.eye(3)
    return mu / (4.0 * d ** (1.0 / 6.0)) * np.trace(M @ M)


def deviatoric_stress(A, mu):
    """-mu/|G|^(1/6) (G^2 - tr(G^2)/3 I - |G|^(1/3) (G - tr(G)/3 I)) (state_equation.c:143-204)."""
    G = A @ A.T
    d = abs(np.linalg.det(G))
    G2 = G @ G
    I = np.eye(3)
    return -mu / d ** (1.0 / 6.0) * (G2 - np.trace(G2) / 3.0 * I - d ** (1.0 / 3.0) * (G - np.trace(G) / 3.0 * I))


def m5_state(p, alpha, rho, pres, u, A):
    """State vector with every material at pressure pres; the elastic energy of a solid is in its
    alpha rho e, so that the record recovers pres (m5_facercnstr.c:143-150)."""
    nm = p.nmat
    alpha = np.asarray(alpha, float)
    rho = np.asarray(rho, float)
    u = np.asarray(u, float)
    A = np.asarray(A, float)
    x = np.zeros(F.m5_neq(nm))
    ra = alpha * rho
    x[0:nm] = ra
    r = ra.sum()
    x[nm:nm + 3] = r * u
    are = np.zeros(nm)
    for k in range(nm):
        el = elastic_energy(A, p.mush[k]) if p.mush[k] != 0.0 else 0.0
        are[k] = alpha[k] * (pres + p.gamma[k] * p.pinf[k]) / (p.gamma[k] - 1.0) + el
    x[nm + 3] = are.sum() + 0.5 * r * (u @ u)
    x[nm + 4:2 * nm + 3] = alpha[:nm - 1]
    x[2 * nm + 3:3 * nm + 3] = are
    x[3 * nm + 3:] = A.ravel()
    return x


def random_rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def near_identity(rng, eps=0.02):
    return np.eye(3) + eps * rng.normal(size=(3, 3))


def _side(p, rng, a0, u):
    if p.nmat == 3:
        a1 = rng.uniform(0.2, 0.8) * (1.0 - a0)
        alpha = (a0, a1, 1.0 - a0 - a1)
    else:
        alpha = (a0, 1.0 - a0)
    u = rng.normal(scale=30.0, size=3) if u is None else u
    return m5_state(p, alpha, RHO[:p.nmat], rng.uniform(1e5, 5e7), u, near_identity(rng))


def solid_side(p, rng, u=None):
    return _side(p, rng, rng.uniform(0.9, 0.98), u)


def fluid_side(p, rng, u=None):
    return _side(p, rng, rng.uniform(0.02, 0.1), u)


def frame_of(n):
    """An orthonormal (n, t1, t2) frame."""
    n = np.asarray(n, float) / np.linalg.norm(n)
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    t1 = np.cross(n, a)
    t1 /= np.linalg.norm(t1)
    return np.concatenate([n, t1, np.cross(n, t1)])


def layout(nm):
    i1 = nm
    i2 = nm + 3
    i3 = nm + 4
    i4 = i3 + nm - 1
    i5 = i4 + nm
    return i1, i2, i3, i4, i5


def physical_flux(p, x, u, n):
    """F(x).n with velocity u (m5_flux_gfm.c:10-87 with the side's own face state)."""
    nm = p.nmat
    i1, i2, i3, i4, i5 = layout(nm)
    rec = O.m5_record(p, x)
    r, e, st = rec["r"], rec["e"], rec["st"]
    un = u @ n
    f = np.zeros(F.m5_neq(nm))
    f[0:nm] = rec["ra"] * un
    f[i1:i1 + 3] = r * u * un - st @ n
    f[i2] = r * e * un - n @ (st @ u)
    f[i3:i4] = rec["avf"][:nm - 1] * un
    f[i4:i5] = rec["are"] * un
    f[i5:] = rec["Amat"].ravel() * un
    return f


# ---------------------------------------------------------------- oracle pins (CPU)

def test_m5_record_undeformed_and_volumetric_states():
    p = m5p()
    rng = np.random.default_rng(0)
    for A in (np.eye(3), 1.07 * np.eye(3)):
        alpha = (0.6, 0.3, 0.1)
        x = m5_state(p, alpha, RHO, 2e6, (3.0, -1.0, 2.0), A)
        rec = O.m5_record(p, x)
        # no shear: zero elastic energy and an isotropic stress -sum alpha_k p_k
        assert np.allclose(rec["p_hydro"], 2e6, rtol=1e-12, atol=1e-3)
        assert np.allclose(rec["st"], -2e6 * np.eye(3), rtol=1e-12, atol=1e-3)
        r = np.dot(alpha, RHO)
        c2 = np.array(alpha) / r * np.array(p.gamma[:3]) * (2e6 + np.array(p.pinf[:3]))
        assert np.allclose(rec["c2Y"], c2, rtol=1e-12)
        assert rec["phi0"] == pytest.approx(0.6 - 0.5, abs=1e-15)
    del rng


def test_m5_record_sheared_solid_matches_the_matrix_formulas():
    p = m5p()
    rng = np.random.default_rng(1)
    for _ in range(5):
        A = near_identity(rng, 0.05)
        alpha = (0.95, 0.03, 0.02)
        pres = rng.uniform(1e5, 1e8)
        x = m5_state(p, alpha, RHO, pres, rng.normal(size=3), A)
        rec = O.m5_record(p, x)
        # the elastic energy of the record equals the matrix formula: pressure recovered
        assert np.allclose(rec["p_hydro"], pres, rtol=1e-9)
        # stress: isotropic part plus the deviatoric elastic stress of the solid
        want = -pres * np.eye(3) + deviatoric_stress(A, p.mush[0])
        assert np.allclose(rec["st"], want, rtol=1e-9, atol=1e-6 * abs(want).max())


def test_m5_record_stress_rotates_as_a_tensor():
    p = m5p()
    rng = np.random.default_rng(2)
    A = near_identity(rng, 0.05)
    Q = random_rotation(rng)
    x = m5_state(p, (0.95, 0.03, 0.02), RHO, 3e7, (1.0, 2.0, 3.0), A)
    xq = m5_state(p, (0.95, 0.03, 0.02), RHO, 3e7, Q @ np.array([1.0, 2.0, 3.0]), Q @ A @ Q.T)
    st, stq = O.m5_record(p, x)["st"], O.m5_record(p, xq)["st"]
    assert np.allclose(stq, Q @ st @ Q.T, rtol=1e-9, atol=1e-7 * abs(st).max())


def test_m5_gfm_supersonic_faces_are_the_physical_flux():
    p = m5p(FLUIDS)
    rng = np.random.default_rng(3)
    area = 2.0 ** -12
    for _ in range(4):
        n = rng.normal(size=3)
        fr = frame_of(n)
        nv = fr[0:3]
        # region 0: everything moves right faster than sound (air-rich mixture, c about 400 m/s)
        uL = 3000.0 * nv + rng.normal(scale=50, size=3)
        uR = 3000.0 * nv + rng.normal(scale=50, size=3)
        alpha_air = (0.001, 0.001, 0.998)
        xL = m5_state(p, alpha_air, RHO, 1e5, uL, near_identity(rng))
        xR = m5_state(p, alpha_air, RHO, 2e5, uR, near_identity(rng))
        flux, unc = O.m5_gfm_face(p, xL, xR, fr, area)
        assert np.allclose(flux, area * physical_flux(p, xL, uL, nv), rtol=1e-12, atol=1e-15)
        assert np.allclose(unc, uL, rtol=1e-14)
        # region 3: supersonic to the left; the right side's ghost carries the left tangential velocity
        uL, uR = -uL, -uR
        xL = m5_state(p, alpha_air, RHO, 1e5, uL, near_identity(rng))
        xR = m5_state(p, alpha_air, RHO, 2e5, uR, near_identity(rng))
        flux, unc = O.m5_gfm_face(p, xL, xR, fr, area)
        ug = (uR @ nv) * nv + (uL - (uL @ nv) * nv)
        assert np.allclose(unc, ug, rtol=1e-12)
        assert np.allclose(flux, area * physical_flux(p, xR, ug, nv), rtol=1e-11, atol=1e-15)


def test_m5_gfm_rest_state_keeps_the_reference_star_energy():
    """At rest on both sides u* = 0 and the left star state is taken.  The reference
    divides by r*(u* + S_K) (m5_star_region.c:73), so E* = -E and the energy
    flux is -2 S_L rho E; the mass flux vanishes and the momentum flux is -sigma.n."""
    p = m5p(FLUIDS)
    rng = np.random.default_rng(4)
    x = m5_state(p, (0.2, 0.7, 0.1), RHO, 1e6, (0.0, 0.0, 0.0), np.eye(3))
    for _ in range(3):
        fr = frame_of(rng.normal(size=3))
        nv = fr[0:3]
        flux, unc = O.m5_gfm_face(p, x, x, fr, 1.0)
        rec = O.m5_record(p, x)
        c = np.sqrt(rec["c2Y"].sum())
        SL = -c
        rhoE = x[3 + 3]
        i1, i2, i3, i4, i5 = layout(3)
        assert np.allclose(flux[0:3], 0.0, atol=1e-12 * rec["r"] * c)
        assert np.allclose(flux[i1:i1 + 3], -rec["st"] @ nv, rtol=1e-12)
        assert flux[i2] == pytest.approx(-2.0 * SL * rhoE, rel=1e-12)
        assert np.allclose(unc, 0.0, atol=1e-12)


def test_m5_gfm_flux_is_rotation_covariant():
    p = m5p()
    rng = np.random.default_rng(5)
    nm = 3
    i1, i2, i3, i4, i5 = layout(nm)
    for trial in range(6):
        Q = random_rotation(rng)
        alphaL, alphaR = (0.95, 0.03, 0.02), (0.04, 0.9, 0.06)
        uL, uR = rng.normal(scale=40, size=3), rng.normal(scale=40, size=3)
        AL, AR = near_identity(rng), near_identity(rng)
        pL, pR = rng.uniform(1e5, 5e7), rng.uniform(1e5, 5e7)
        fr = frame_of(rng.normal(size=3))
        xL = m5_state(p, alphaL, RHO, pL, uL, AL)
        xR = m5_state(p, alphaR, RHO, pR, uR, AR)
        frq = np.concatenate([Q @ fr[0:3], Q @ fr[3:6], Q @ fr[6:9]])
        xLq = m5_state(p, alphaL, RHO, pL, Q @ uL, Q @ AL @ Q.T)
        xRq = m5_state(p, alphaR, RHO, pR, Q @ uR, Q @ AR @ Q.T)
        f, u = O.m5_gfm_face(p, xL, xR, fr, 1.0)
        fq, uq = O.m5_gfm_face(p, xLq, xRq, frq, 1.0)
        scale = np.abs(f).max()
        tol = dict(rtol=1e-8, atol=1e-9 * scale)
        assert np.allclose(fq[0:nm], f[0:nm], **tol)
        assert np.allclose(fq[i1:i1 + 3], Q @ f[i1:i1 + 3], **tol)
        assert np.allclose(fq[i2:i5], f[i2:i5], **tol)
        assert np.allclose(fq[i5:].reshape(3, 3), Q @ f[i5:].reshape(3, 3) @ Q.T, **tol)
        assert np.allclose(uq, Q @ u, rtol=1e-8, atol=1e-9 * np.abs(u).max())


# m5_RK on a small uniform forest

def m5_mesh(level=3, L=2.0 ** -6, bc=(4, 4, 4, 4, 4, 4)):
    m = F.Mesh.uniform((1, 1, 1), (0.0, 0.0, 0.0), L, level, list(bc))
    return m, O.OracleMesh(m)


def random_m5_field(p, n, rng, solid_frac=0.5):
    S = np.zeros((n, F.m5_neq(p.nmat)))
    for i in range(n):
        S[i] = solid_side(p, rng) if rng.random() < solid_frac else fluid_side(p, rng)
    return S


def test_m5_rk_stage_tables_and_the_plus_dt_sign():
    p = m5p(FLUIDS, dt=2.0 ** -8)
    m, om = m5_mesh()
    rng = np.random.default_rng(6)
    n = m.n
    SN = random_m5_field(p, n, rng)
    vol = np.array([om.cells[i].Vol for i in range(n)])
    k = np.ldexp(rng.integers(1, 64, size=SN.shape).astype(float), -4) * rng.choice([-1.0, 1.0], size=SN.shape)
    RHS = vol[:, None] * k
    gf = np.zeros((n, 6), np.uint8)
    _, S1 = om.m5_rk(p, 1, SN, SN, RHS, gf)
    # m5_RK.c:55-60: S = SN + DT * RHS / Vol, a plus where RK.c:25 has a minus
    assert np.all(np.sign(S1 - SN) == np.sign(k))
    assert np.allclose(S1 - SN, p.dt * k, rtol=1e-6, atol=4 * np.spacing(np.abs(SN)).max())
    assert np.array_equal(S1, SN + p.dt * ((1.0 / vol)[:, None] * RHS))
    zero = np.zeros_like(RHS)
    SNc, S0 = om.m5_rk(p, 0, np.zeros_like(SN), SN, zero, gf)
    assert np.array_equal(SNc, SN) and np.array_equal(S0, SN)
    for irk in (2, 4):
        _, S = om.m5_rk(p, irk, SN, SN, zero, gf)
        assert np.array_equal(S, SN)       # 3/4 + 1/4 and 1/2 + 1/2 of the same value are exact
    _, S3 = om.m5_rk(p, 3, SN, SN, zero, gf)
    assert np.allclose(S3, SN, rtol=2 ** -51)


def test_m5_rk_positivity_fallback_and_velocity_freezes():
    m, om = m5_mesh()
    rng = np.random.default_rng(7)
    n = m.n
    for zerou in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
        p = m5p(FLUIDS, dt=2.0 ** -20, zerou=zerou)
        SN = random_m5_field(p, n, rng)
        vol = np.array([om.cells[i].Vol for i in range(n)])
        RHS = vol[:, None] * rng.normal(size=SN.shape) * 1e3
        RHS[5, 1] = -vol[5] * SN[5, 1] / p.dt * 4.0                 # drives alpha rho_1 of leaf 5 negative
        gf = np.zeros((n, 6), np.uint8)
        _, S = om.m5_rk(p, 1, SN, SN, RHS, gf)
        assert S[5, 1] == SN[5, 1]
        c = zerou.index(1)
        frozen_A = {0: (1, 2, 3, 6), 1: (1, 3, 5, 7), 2: (2, 5, 6, 7)}[c]
        i5 = layout(3)[4]
        assert np.array_equal(S[:, 3 + c], SN[:, 3 + c])
        for q in range(9):
            same = np.array_equal(S[:, i5 + q], SN[:, i5 + q])
            assert same == (q in frozen_A), (zerou, q)


def interface_case(p, om, m, rng, leaves, faces, crossing=True):
    """SN, RHS, gfm_face with the given leaves' solid fraction crossing AVF_LIM in stage 1."""
    n = m.n
    SN = random_m5_field(p, n, rng, solid_frac=0.0)
    vol = np.array([om.cells[i].Vol for i in range(n)])
    RHS = vol[:, None] * rng.normal(size=SN.shape) * 1e2
    i3 = layout(3)[2]
    gf = np.zeros((n, 6), np.uint8)
    for c in leaves:
        SN[c] = solid_side(p, rng)
        SN[c, i3] = 0.52
        if crossing:
            RHS[c, i3] = -vol[c] * 0.04 / p.dt                       # alpha_solid 0.52 -> 0.48
    for c, f in faces:
        gf[c, f] = 1
    return SN, RHS, gf, vol


def neighbour(m, c, f):
    a = m.arrays()
    return int(a["neigtr"][c, f, 0]) if a["neigtr"].ndim == 3 else int(a["neigtr"].reshape(m.n, 6, 4)[c, f, 0])


def test_m5_gfm_single_face_projects_on_the_face_normal():
    p = m5p(SOLID_WATER_AIR, dt=2.0 ** -30)
    m, om = m5_mesh()
    rng = np.random.default_rng(8)
    c, f = 21, 1                                  # an interior leaf of the 4^3 forest and its +x face
    SN, RHS, gf, vol = interface_case(p, om, m, rng, [c], [(c, f)])
    _, S = om.m5_rk(p, 1, SN, SN, RHS, gf)
    j = neighbour(m, c, f)
    dV = p.dt / vol[c]
    rho_new = (SN[c, 0:3] + dV * RHS[c, 0:3]).sum()
    ut = (SN[c, 3:6] + dV * RHS[c, 3:6]) / rho_new
    # the solid fraction of c after the update and the neighbour's, averaged, signs the normal
    i3 = layout(3)[2]
    phi = 0.5 * ((SN[c, i3] + p.dt * (RHS[c, i3] / vol[c]) - 0.5) + (SN[j, i3] - 0.5))
    nf = np.array([om.cells[c].nx[f], om.cells[c].ny[f], om.cells[c].nz[f]])
    nn = -np.sign(phi) * nf
    uj = SN[j, 3:6] / SN[j, 0:3].sum()
    unew = (ut @ nn) * nn + (uj - (uj @ nn) * nn)
    assert np.allclose(S[c, 3:6], rho_new * unew, rtol=1e-12, atol=1e-12 * rho_new * np.abs(unew).max())
    assert np.allclose(S[c, 0:3], SN[c, 0:3] + dV * RHS[c, 0:3], rtol=1e-15)
    E = SN[c, 6] + p.dt * (RHS[c, 6] / vol[c])
    E = E - rho_new * 0.5 * (ut @ ut) + rho_new * 0.5 * (unew @ unew)
    assert S[c, 6] == pytest.approx(E, rel=1e-12)
    # every other leaf took the plain update
    others = np.ones(m.n, bool)
    others[c] = False
    plain = SN + p.dt * (RHS / vol[:, None])
    assert np.array_equal(S[others], plain[others])


def test_m5_gfm_keeps_the_last_interface_face_and_gfm_rel_sums_the_normals():
    """m5_gfm.c:92-94 overwrites the interface normal per face and both
    re-projections keep the last face's tangential velocity times its area
    share (m5_gfm.c:160-171); m5_gfm_rel.c:88-90 sums the normals."""
    p = m5p(SOLID_WATER_AIR, dt=2.0 ** -30)
    m, om = m5_mesh()
    rng = np.random.default_rng(9)
    c = 21
    faces = [(c, 3), (c, 2)]                      # the -x and +y faces of leaf c
    SN, RHS, gf, vol = interface_case(p, om, m, rng, [c], faces)
    _, S = om.m5_rk(p, 1, SN, SN, RHS, gf)
    i3 = layout(3)[2]
    dV = p.dt / vol[c]
    rho_new = (SN[c, 0:3] + dV * RHS[c, 0:3]).sum()
    ut = (SN[c, 3:6] + dV * RHS[c, 3:6]) / rho_new
    a_new = SN[c, i3] + p.dt * (RHS[c, i3] / vol[c])

    def face_normal(f):
        return np.array([om.cells[c].nx[f], om.cells[c].ny[f], om.cells[c].nz[f]])

    # neighbours before c in leaf order already hold their stage-1 state
    upd = SN + p.dt * (RHS / vol[:, None])
    f_last = 3                                    # face order 0..5: -x (3) comes after +y (2)
    j_last = neighbour(m, c, f_last)
    nb = upd if j_last < c else SN
    phi = 0.5 * ((a_new - 0.5) + (nb[j_last, i3] - 0.5))
    nn = -np.sign(phi) * face_normal(f_last)
    uj = nb[j_last, 3:6] / nb[j_last, 0:3].sum()
    unew = (ut @ nn) * nn + 0.5 * (uj - (uj @ nn) * nn)
    assert np.allclose(S[c, 3:6], rho_new * unew, rtol=1e-12, atol=1e-12 * rho_new * np.abs(unew).max())

    # after relaxation: same faces, the normal is the phi- and area-weighted sum
    Sb = SN.copy()
    Srel = SN.copy()
    Srel[c, i3] = 0.47                            # relaxation moved leaf c across AVF_LIM
    out = om.m5_gfm_rel(p, Sb, Srel, gf)
    rho = Srel[c, 0:3].sum()
    u0 = Srel[c, 3:6] / rho
    acc = np.zeros(3)
    for f in (2, 3):                              # face order
        j = neighbour(m, c, f)
        phi = 0.5 * ((Srel[c, i3] - 0.5) + (Sb[j, i3] - 0.5))
        acc += phi * om.cells[c].Area[f][0] * face_normal(f)
    nn = -acc / np.linalg.norm(acc)
    uj = Sb[j_last, 3:6] / Sb[j_last, 0:3].sum()
    unew = (u0 @ nn) * nn + 0.5 * (uj - (uj @ nn) * nn)
    assert np.allclose(out[c, 3:6], rho * unew, rtol=1e-12, atol=1e-12 * rho * np.abs(unew).max())
    E = Srel[c, 6] - rho * 0.5 * (u0 @ u0) + rho * 0.5 * (unew @ unew)
    assert out[c, 6] == pytest.approx(E, rel=1e-12)
    others = np.ones(m.n, bool)
    others[c] = False
    assert np.array_equal(out[others], Srel[others])


def test_m5_gfm_later_leaf_sees_the_earlier_leafs_new_velocity():
    """m5_RK updates leaves in place in leaf order: of two neighbouring marked
    leaves, the later one reads the earlier one's re-projected velocity, the
    earlier one reads the later one's old state."""
    p = m5p(SOLID_WATER_AIR, dt=2.0 ** -30)
    m, om = m5_mesh()
    rng = np.random.default_rng(10)
    a, fa = 21, 1
    b = neighbour(m, a, fa)
    fb = 3 if fa == 1 else 1
    assert neighbour(m, b, fb) == a
    lo, hi = min(a, b), max(a, b)
    SN, RHS, gf, vol = interface_case(p, om, m, rng, [a, b], [(a, fa), (b, fb)])
    _, S = om.m5_rk(p, 1, SN, SN, RHS, gf)
    i3 = layout(3)[2]

    def expect(c, f, nb_state):
        dV = p.dt / vol[c]
        rho_new = (SN[c, 0:3] + dV * RHS[c, 0:3]).sum()
        ut = (SN[c, 3:6] + dV * RHS[c, 3:6]) / rho_new
        j = neighbour(m, c, f)
        a_new = SN[c, i3] + p.dt * (RHS[c, i3] / vol[c])
        phi = 0.5 * ((a_new - 0.5) + (nb_state[j, i3] - 0.5))
        nf = np.array([om.cells[c].nx[f], om.cells[c].ny[f], om.cells[c].nz[f]])
        nn = -np.sign(phi) * nf
        uj = nb_state[j, 3:6] / nb_state[j, 0:3].sum()
        return rho_new * ((ut @ nn) * nn + (uj - (uj @ nn) * nn))

    f_lo = fa if lo == a else fb
    f_hi = fb if hi == b else fa
    assert np.allclose(S[lo, 3:6], expect(lo, f_lo, SN), rtol=1e-12, atol=1e-9)
    assert np.allclose(S[hi, 3:6], expect(hi, f_hi, S), rtol=1e-12, atol=1e-9)
    assert not np.allclose(S[hi, 3:6], expect(hi, f_hi, SN), rtol=1e-9, atol=1e-9)


# ---------------------------------------------------------------- device parity (GPU)

def random_faces(p, rng, nf, solid=True):
    nq = F.m5_neq(p.nmat)
    UL = np.zeros((nf, nq))
    UR = np.zeros((nf, nq))
    FR = np.zeros((nf, 9))
    area = np.ldexp(1.0, -rng.integers(8, 14, size=nf))
    for q in range(nf):
        sideA = solid_side(p, rng) if solid else fluid_side(p, rng)
        sideB = fluid_side(p, rng)
        UL[q], UR[q] = (sideA, sideB) if rng.random() < 0.5 else (sideB, sideA)
        if rng.random() < 0.3:
            FR[q] = frame_of(rng.normal(size=3))
        else:
            ax = rng.integers(0, 3)
            nvec = np.zeros(3)
            nvec[ax] = 1.0 if rng.random() < 0.5 else -1.0
            FR[q] = frame_of(nvec)
    # a share of near-sonic and supersonic faces so that every HLLC region occurs
    for q in range(0, nf, 7):
        nvec = FR[q, 0:3]
        sgn = 1.0 if (q // 7) % 2 == 0 else -1.0
        for U in (UL, UR):
            r = U[q, 0:p.nmat].sum()
            u_old = U[q, p.nmat:p.nmat + 3] / r
            u_new = u_old + sgn * 3000.0 * nvec
            U[q, p.nmat + 3] += 0.5 * r * (u_new @ u_new - u_old @ u_old)
            U[q, p.nmat:p.nmat + 3] = r * u_new
    return UL, UR, FR, area


def hllc_region(p, xL, xR, fr):
    """0 left, 1 left star, 2 right star, 3 right state (m5_hllc_gfm.c:56-82) from the records."""
    rL, rR = O.m5_record(p, xL), O.m5_record(p, xR)
    n, t1, t2 = fr[0:3], fr[3:6], fr[6:9]
    V = np.stack([n, t1, t2])
    uL = np.array([rL["u"], rL["v"], rL["w"]]) @ n
    uR = np.array([rR["u"], rR["v"], rR["w"]]) @ n
    cL, cR = np.sqrt(rL["c2Y"].sum()), np.sqrt(rR["c2Y"].sum())
    SL, SR = min(uL - cL, uR - cR), max(uL + cL, uR + cR)
    s11L, s11R = (V @ rL["st"] @ V.T)[0, 0], (V @ rR["st"] @ V.T)[0, 0]
    u_S = ((rL["r"] * uL ** 2 - s11L) - (rR["r"] * uR ** 2 - s11R) - SL * rL["r"] * uL + SR * rR["r"] * uR) / (
        rL["r"] * uL - rR["r"] * uR - SL * rL["r"] + SR * rR["r"])
    if SL >= 0:
        return 0
    if u_S >= 0:
        return 1
    return 2 if SR > 0 else 3


@pytest.mark.parametrize("nmat", [2, 3])
def test_random_faces_cover_every_hllc_region(nmat):
    p = m5p(SOLID_WATER_AIR, nmat=nmat)
    UL, UR, FR, area = random_faces(p, np.random.default_rng(11 + nmat), 700)
    regions = np.bincount([hllc_region(p, UL[q], UR[q], FR[q]) for q in range(len(UL))], minlength=4)
    assert np.all(regions >= 20), regions


@pytest.mark.gpu
@pytest.mark.parametrize("table,nmat", [("fluids", 3), ("fluids", 2), ("solid", 3), ("solid", 2)])
def test_m5_gfm_faces_match_oracle(table, nmat):
    tab = FLUIDS if table == "fluids" else SOLID_WATER_AIR
    p = m5p(tab, nmat=nmat)
    rng = np.random.default_rng(11 + nmat)
    nf = 700

    UL, UR, FR, area = random_faces(p, rng, nf, solid=(table == "solid"))
    flux, unc = F.m5_gfm_faces(p, UL, UR, FR, area)
    want = np.zeros_like(flux)
    wunc = np.zeros_like(unc)
    for q in range(nf):
        want[q], wunc[q] = O.m5_gfm_face(p, UL[q], UR[q], FR[q], area[q])
    print(f"[m5 gfm faces] {table} nmat={nmat} nf={nf}", flush=True)
    if table == "fluids":
        assert np.array_equal(flux, want)
        assert np.array_equal(unc, wunc)
    else:
        # the elastic EOS goes through pow(|G|, 1/3) and pow(|G|, 1/6): device and glibc pow differ in the last bit
        scale = np.abs(want).max(axis=1, keepdims=True)
        assert np.all(np.abs(flux - want) <= 1e-11 * scale)
        assert np.allclose(unc, wunc, rtol=1e-11, atol=1e-11 * np.abs(wunc).max())


@pytest.mark.gpu
def test_m5_gfm_faces_report_nan():
    p = m5p(FLUIDS)
    rng = np.random.default_rng(12)
    UL, UR, FR, area = random_faces(p, rng, 4, solid=False)
    UL[2, 0:3] = 0.0                                # zero density: u = 0/0
    with pytest.raises(F.FVError) as e:
        F.m5_gfm_faces(p, UL, UR, FR, area)
    assert e.value.code == F.FV_ERR_NAN


def rk_case(seed, level=3, amr=False):
    rng = np.random.default_rng(seed)
    p = m5p(SOLID_WATER_AIR, dt=2.0 ** -30)
    if amr:
        from cases import random_amr_mesh
        base = F.make_params(bc=(4, 4, 4, 4, 4, 4))
        m = random_amr_mesh(base, base_level=3, max_level=5, seed=seed, frac=0.25, passes=2, L=2.0 ** -6)
    else:
        m = F.Mesh.uniform((1, 1, 1), (0.0, 0.0, 0.0), 2.0 ** -6, level, [4] * 6)
    om = O.OracleMesh(m)
    n = m.n
    i3 = layout(3)[2]
    SN = random_m5_field(p, n, rng, solid_frac=0.4)
    vol = np.array([om.cells[i].Vol for i in range(n)])
    RHS = vol[:, None] * rng.normal(size=SN.shape) * 1e2
    a = m.arrays()
    bc = a["bc"].reshape(n, 6)
    # a quarter of the leaves sit at the solid fraction 0.52 and cross AVF_LIM this stage
    marked = rng.random(n) < 0.25
    SN[marked, i3] = 0.52
    RHS[marked, i3] = -vol[marked] * 0.04 / p.dt
    gf = ((rng.random((n, 6)) < 0.4) & (bc == 0)).astype(np.uint8)
    S = SN + vol[:, None] * 0.0
    S[~marked] += rng.normal(size=S[~marked].shape) * 1e-3 * np.abs(S[~marked]).max(axis=0)
    return p, m, om, SN, S, RHS, gf


@pytest.mark.gpu
@pytest.mark.parametrize("amr", [False, True])
def test_m5_rk_matches_oracle_bitwise(amr):
    p, m, om, SN, S, RHS, gf = rk_case(13, amr=amr)
    pk = m.pack()
    for irk in (0, 1, 2, 3, 4):
        Sin = SN if irk == 1 else S
        SNg, Sg = F.m5_rk(p, pk, irk, SN, Sin, RHS, gf)
        SNo, So = om.m5_rk(p, irk, SN, Sin, RHS, gf)
        print(f"[m5 rk] amr={amr} n={m.n} irk={irk}", flush=True)
        assert np.array_equal(Sg, So), f"irk {irk}: max diff {np.abs(Sg - So).max()}"
        if irk == 0:
            assert np.array_equal(SNg, SNo)
    # the stage-1 case did re-project leaves, with chains of neighbouring marked leaves
    _, S1 = om.m5_rk(p, 1, SN, SN, RHS, gf)
    plain = SN + p.dt * (RHS / np.array([om.cells[i].Vol for i in range(m.n)])[:, None])
    assert (np.abs(S1[:, 3:6] - plain[:, 3:6]).max(axis=1) > 0).sum() > 3


@pytest.mark.gpu
@pytest.mark.parametrize("amr", [False, True])
def test_m5_gfm_rel_matches_oracle_bitwise(amr):
    p, m, om, SN, S, RHS, gf = rk_case(14, amr=amr)
    i3 = layout(3)[2]
    rng = np.random.default_rng(15)
    Srel = S.copy()
    cross = rng.random(m.n) < 0.3
    Srel[cross, i3] = np.where(S[cross, i3] > 0.5, 0.45, 0.55)
    Srel[~cross, 3:6] *= 1.0 + 1e-3 * rng.normal(size=(int((~cross).sum()), 3))
    got = F.m5_gfm_rel(p, m.pack(), S, Srel, gf)
    want = om.m5_gfm_rel(p, S, Srel, gf)
    print(f"[m5 gfm_rel] amr={amr} n={m.n} crossing={int(cross.sum())}", flush=True)
    assert np.array_equal(got, want)
    assert (np.abs(want - Srel).max(axis=1) > 0).sum() > 3


@pytest.mark.gpu
def test_m5_rk_rejects_an_interface_flag_on_a_boundary_face():
    p, m, om, SN, S, RHS, gf = rk_case(16)
    i3 = layout(3)[2]
    gf[:] = 0
    SN[0, i3] = 0.52
    RHS[0, i3] = -om.cells[0].Vol * 0.04 / p.dt
    gf[0, :] = 1                                   # leaf 0 is a corner leaf: three of its faces are walls
    with pytest.raises(F.FVError) as e:
        F.m5_rk(p, m.pack(), 1, SN, SN, RHS, gf)
    assert e.value.code == F.FV_ERR_ARG
