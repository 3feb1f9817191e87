"""Perfect plasticity of the model-5 solid (fv_m5_params.yield).

The reference has no plasticity to match: the plastic_rlx call in
m5_relaxation_solid.c:14 is commented out and m5_plastic_rlx
(strdata.h:809) is never defined.  The return map is specified by what it
must satisfy, and the CPU tests pin the oracle (orc_m5_plastic) on exactly
that, through the stress of the pinned elastic EOS record (orc_m5_record,
state_equation.c:143-204) rather than through the map's own formula:
  * below the yield stress nothing changes, bit for bit;
  * above it the von Mises stress of the deviatoric part of the record's
    stress equals the yield stress afterwards;
  * det(A) is kept (no volume change), G' = A' A'^T has the eigenvectors of
    G = A A^T, and its log strains ln(l_i) - ln(det G)/3 are those of G
    scaled by one common factor in (0, 1);
  * only A changes: the partial densities, momentum, total energy, fractions
    and phasic energies stay, so the released elastic energy becomes heat;
  * the map is frame indifferent: A -> R A gives R A', A -> A R gives A' R;
  * a uniaxial stretch diag(a, 1, 1) returns to diag(b, c, c) with b c^2 = a;
  * in a stage the map runs on each leaf before its relaxation (where
    m5_relaxation_solid.c:14 would call plastic_rlx), so a stage with a
    yield stress equals the stage whose relaxation sees the plastic states.
GPU tests compare fv_m5_dev_* with yield > 0 against the oracle.
"""
import numpy as np
import pytest

import forestfv_public_amd as F
import oracle_lib as O
from test_m5 import RHO, SOLID_WATER_AIR, layout, m5_state
from test_m5_stage import assert_close, m5_case, m5p

YIELD = 3e8


def plastic_params(nmat=3, prelax=3, yield_stress=YIELD, **kw):
    p = m5p(SOLID_WATER_AIR, nmat=nmat, prelax=prelax, **kw)
    p.yield_ = yield_stress
    return p


def von_mises(p, x):
    st = O.m5_record(p, x)["st"]
    dev = st - np.trace(st) / 3.0 * np.eye(3)
    return np.sqrt(1.5 * np.sum(dev * dev))


def solid_state(p, A, pres=1e7, u=(3.0, -1.0, 2.0)):
    nm = p.nmat
    alpha = (0.97, 0.02, 0.01)[:nm] if nm == 3 else (0.97, 0.03)
    return m5_state(p, alpha, RHO[:nm], pres, u, A)


def log_strains(A):
    G = A @ A.T
    l, Q = np.linalg.eigh(G)
    return l, Q, np.log(l) - np.log(l).sum() / 3.0


def random_rotation(rng):
    Q, R = np.linalg.qr(rng.normal(size=(3, 3)))
    Q = Q @ np.diag(np.sign(np.diag(R)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    return Q


def deformed(rng, eps):
    """A deformation with principal stretches exp(eps * k) (random k), random rotations and volume change."""
    k = rng.normal(size=3)
    k -= k.mean()
    return random_rotation(rng) @ np.diag(np.exp(eps * k) * rng.uniform(0.95, 1.05)) @ random_rotation(rng)


@pytest.mark.parametrize("nmat", [2, 3])
def test_below_yield_nothing_changes(nmat):
    rng = np.random.default_rng(1)
    p = plastic_params(nmat)
    for trial in range(20):
        x = solid_state(p, deformed(rng, 1e-3))
        assert von_mises(p, x) < YIELD
        y, changed = O.m5_plastic(p, x)
        assert not changed
        assert np.array_equal(y, x)


@pytest.mark.parametrize("nmat", [2, 3])
@pytest.mark.parametrize("eps", [0.02, 0.1, 0.5])
def test_above_yield_the_state_returns_to_the_yield_surface(nmat, eps):
    rng = np.random.default_rng(int(eps * 100) + nmat)
    p = plastic_params(nmat)
    i1, i2, i3, i4, i5 = layout(nmat)
    for trial in range(20):
        A = deformed(rng, eps)
        x = solid_state(p, A)
        assert von_mises(p, x) > YIELD
        y, changed = O.m5_plastic(p, x)
        assert changed
        An = y[i5:].reshape(3, 3)
        # on the yield surface
        assert abs(von_mises(p, y) - YIELD) <= 1e-9 * YIELD
        # volume kept
        assert abs(np.linalg.det(An) - np.linalg.det(A)) <= 1e-12 * abs(np.linalg.det(A))
        # coaxial with G, log strains scaled by one factor in (0, 1)
        G, Gn = A @ A.T, An @ An.T
        assert np.abs(G @ Gn - Gn @ G).max() <= 1e-12 * np.abs(G).max() ** 2
        l, Q, k = log_strains(A)
        ln, Qn, kn = log_strains(An)
        t = kn @ k / (k @ k)
        assert 0.0 < t < 1.0
        assert np.abs(kn - t * k).max() <= 1e-9 * np.abs(k).max()
        # only A changes
        assert np.array_equal(y[:i5], x[:i5])


def test_released_elastic_energy_becomes_heat():
    """rho E and the phasic energies stay, so the hydrostatic pressure of the solid rises
    by (gamma - 1) times the released elastic energy over its fraction."""
    rng = np.random.default_rng(7)
    p = plastic_params(3)
    for trial in range(5):
        x = solid_state(p, deformed(rng, 0.1))
        y, _ = O.m5_plastic(p, x)
        r0, r1 = O.m5_record(p, x), O.m5_record(p, y)
        assert r1["p_hydro"][0] > r0["p_hydro"][0]
        assert np.array_equal(r1["p_hydro"][1:], r0["p_hydro"][1:])


def test_the_map_is_frame_indifferent():
    rng = np.random.default_rng(3)
    p = plastic_params(3)
    i5 = layout(3)[4]
    for trial in range(10):
        A = deformed(rng, 0.2)
        R = random_rotation(rng)
        a = O.m5_plastic(p, solid_state(p, A))[0][i5:].reshape(3, 3)
        left = O.m5_plastic(p, solid_state(p, R @ A))[0][i5:].reshape(3, 3)
        right = O.m5_plastic(p, solid_state(p, A @ R))[0][i5:].reshape(3, 3)
        assert np.abs(left - R @ a).max() <= 1e-12
        assert np.abs(right - a @ R).max() <= 1e-12


@pytest.mark.parametrize("a", [1.05, 0.9, 1.5])
def test_uniaxial_stretch_returns_along_its_axes(a):
    p = plastic_params(3)
    i5 = layout(3)[4]
    y, changed = O.m5_plastic(p, solid_state(p, np.diag([a, 1.0, 1.0])))
    assert changed
    An = y[i5:].reshape(3, 3)
    assert np.abs(An - np.diag(np.diag(An))).max() <= 1e-15
    b, c, c2 = np.diag(An)
    assert abs(c - c2) <= 1e-15
    assert abs(b * c * c - a) <= 1e-13
    assert (b - 1.0) * (a - 1.0) > 0 and abs(b - 1.0) < abs(a - 1.0)
    assert abs(von_mises(p, y) - YIELD) <= 1e-9 * YIELD


def test_fluids_and_zero_yield_never_change():
    rng = np.random.default_rng(4)
    fl = F.make_m5_params(nmat=3, gamma=(4.4, 6.12, 1.4), pinf=(6e8, 3.43e8, 0.0), mush=(0.0, 0.0, 0.0),
                          amin=1e-6, avf_lim=0.5, dt=1e-9, yield_stress=1.0)
    el = plastic_params(3, yield_stress=0.0)
    for p in (fl, el):
        x = solid_state(p, deformed(rng, 0.3))
        y, changed = O.m5_plastic(p, x)
        assert not changed and np.array_equal(y, x)


def test_a_stage_applies_the_map_before_the_relaxation():
    """The stage with a yield stress is the composition of the pinned operators with the
    map applied to each leaf between m5_RK and its relaxation."""
    p, m, om, S = m5_case(21, nmat=3, prelax=3)
    i5 = layout(3)[4]
    rng = np.random.default_rng(21)
    for i in range(m.n):   # strain every leaf past yield
        A = S[i, i5:].reshape(3, 3)
        S[i, i5:] = (A @ deformed(rng, 0.05)).ravel()
    p.dt = 1e-9
    p.yield_ = YIELD
    got = om.m5_stage(p, 1, S.copy(), S.copy())[0]
    SF = om.m5_faceinterp(p, S)
    R, g = om.m5_residual(p, S, SF)
    _, Srk = om.m5_rk(p, 1, S.copy(), S.copy(), R, g)
    nchanged = 0
    Srel = Srk.copy()
    for i in range(m.n):
        y, changed = O.m5_plastic(p, Srk[i])
        nchanged += changed
        Srel[i] = O.m5_relax(p, y)
    want = om.m5_gfm_rel(p, Srk, Srel, g)
    assert nchanged > m.n // 2
    assert np.array_equal(got, want)
    # and the map made a difference
    p.yield_ = 0.0
    assert not np.array_equal(om.m5_stage(p, 1, S.copy(), S.copy())[0], got)


@pytest.mark.parametrize("bad", [-1.0, float("nan")])
def test_a_negative_yield_stress_is_rejected(bad):
    p, m, om, S = m5_case(22, nmat=3, prelax=3)
    p.yield_ = bad
    with pytest.raises(Exception):
        om.m5_stage(p, 1, S.copy(), S.copy())


# ---------------------------------------------------------------- device

def strained_case(seed, prelax, eps=0.005):
    """The smooth solid/water/air case with every leaf's A strained past yield and its phasic
    energies re-derived for that A at the leaf's fractions, densities and velocity, 100 bar."""
    p, m, om, S = m5_case(seed, nmat=3, prelax=prelax)
    i1, i2, i3, i4, i5 = layout(3)
    rng = np.random.default_rng(seed)
    # a smooth volume-preserving strain field exp(eps M(x)), M traceless symmetric
    c = m.centres() / m.m.rootsize
    B = rng.normal(size=(3, 3, 3))
    B = B + B.transpose(0, 2, 1)
    B -= np.trace(B, axis1=1, axis2=2)[:, None, None] / 3.0 * np.eye(3)
    k = rng.uniform(0.5, 1.5, size=(3, 3))
    ph = rng.uniform(0, 2 * np.pi, size=3)
    for i in range(m.n):
        M = sum(np.sin(2 * np.pi * c[i] @ k[q] + ph[q]) * B[q] for q in range(3))
        w, V = np.linalg.eigh(M)
        alpha = np.append(S[i, i3:i4], 1.0 - S[i, i3:i4].sum())
        solid = np.clip((alpha[0] - 0.05) / 0.9, 0.0, 1.0) ** 2   # strained inside the solid only
        D = V @ np.diag(np.exp(eps * solid * w)) @ V.T
        rho = S[i, :3] / alpha
        u = S[i, i1:i2] / S[i, :3].sum()
        S[i] = m5_state(p, alpha, rho, 1e7, u, S[i, i5:].reshape(3, 3) @ D)
    p.dt = 2e-9
    p.yield_ = YIELD
    return p, m, om, S


@pytest.mark.gpu
@pytest.mark.parametrize("prelax", [0, 1, 3])
def test_m5_device_stages_with_plasticity_match_oracle(prelax):
    from test_m5_stage import say
    p, m, om, S = strained_case(40 + prelax, prelax)
    say(f"[m5 plastic dev] prelax={prelax} n={m.n}")
    dev = F.M5Device(p, m)
    dev.set_state(S)
    dev.begin_step()
    SN, So = S.copy(), S.copy()
    for irk in (1, 2, 3):
        dev.stage(irk)
        So = om.m5_stage(p, irk, SN, So)[0]
        assert_close(dev.get_state(), So, False, f"stage {irk}", rel=1e-10)
    # the map acted: the elastic stages differ
    q = plastic_params(3, prelax=prelax, yield_stress=0.0)
    q.dt = p.dt
    El = om.m5_stage(q, 1, S.copy(), S.copy())[0]
    Pl = om.m5_stage(p, 1, S.copy(), S.copy())[0]
    assert not np.allclose(El, Pl, rtol=1e-6, atol=0)
    dev.close()


@pytest.mark.gpu
def test_m5_device_steps_with_plasticity_match_oracle():
    from test_m5_stage import say
    p, m, om, S = strained_case(47, 3)
    say(f"[m5 plastic dev steps] n={m.n}")
    dev = F.M5Device(p, m)
    dev.set_state(S)
    dev.advance(2)
    want = om.m5_advance(p, S.copy(), 2)
    assert_close(dev.get_state(), want, False, "two steps", rel=1e-10)
    dev.close()


@pytest.mark.gpu
def test_m5_device_rejects_a_negative_yield_stress():
    p, m, om, S = m5_case(48, nmat=3, prelax=3)
    p.yield_ = -1.0
    with pytest.raises(Exception):
        F.M5Device(p, m)
