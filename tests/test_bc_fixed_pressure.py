"""Fixed-pressure inlet and outlet faces (fv_params.bc_fixed_p).

m4_bc_inlet and m4_bc_outlet (m4_bc.c:44-113, 121-186) act only for the
reference's CASE 3400 and 3410: the ghost state of an inlet (code 3) or
outlet (code 2) face keeps the own face state's densities, fractions and
velocity and puts every material at a fixed pressure (150 MPa in, 1 MPa out
there); for every other CASE those faces copy the own face state.

CPU tests pin the oracle:
  * with bc_fixed_p = 0, an inlet or outlet face is the HLLC flux between the
    face state and itself, bit for bit;
  * with bc_fixed_p = 1 the face flux equals the interior HLLC flux against an
    explicit state vector built from the definition (alpha_k rho_k, rho u,
    alpha_k, alpha_k (P + gamma_k pinf_k) / (gamma_k - 1), rho E with that
    energy), for two and three materials;
  * a resting uniform state at the outlet pressure stays at rest with every
    side an outlet; an inlet at a higher pressure drives inflow through its
    side and nowhere else in the first stage.
GPU tests: the exact path bit-identical to the oracle per stage, and the
plain FP64, mixed-precision and fused paths against it over steps.
"""
import numpy as np
import pytest

import forestfv_public_amd as F
import oracle_lib as O
from cases import sod_case

WATER_AIR = dict(gamma=(6.12, 1.4, 1.4), pinf=(3.43e8, 0.0, 0.0))


def params(nmat, fixed_p=None):
    return F.make_params(nmat=nmat, gamma=WATER_AIR["gamma"][:nmat], pinf=WATER_AIR["pinf"][:nmat], amin=1e-8,
                         avf_lim=0.5, prelax=0, fixed_p=fixed_p)


def state(p, alpha, rho, pres, u):
    """Conservative vector [alpha_k rho_k | rho u | rho E | alpha_k (nmat-1) | alpha_k rho_k e_k]."""
    nm = p.nmat
    alpha, rho, u = (np.asarray(x, float) for x in (alpha, rho, u))
    g, pi = np.array(p.gamma[:nm]), np.array(p.pinf[:nm])
    ar = alpha * rho
    are = alpha * (pres + g * pi) / (g - 1.0)
    r = ar.sum()
    return np.concatenate([ar, r * u, [are.sum() + 0.5 * r * u @ u], alpha[:-1], are])


def rotated_frame(rng):
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return Q[:, 0], Q[:, 1], Q[:, 2]


@pytest.mark.parametrize("nmat", [2, 3])
@pytest.mark.parametrize("bc", [2, 3])
def test_without_the_flag_inlet_and_outlet_copy_the_face_state(nmat, bc):
    rng = np.random.default_rng(bc + nmat)
    p = params(nmat)
    a = rng.dirichlet(np.ones(nmat))
    x = state(p, a, (1000.0, 1.2, 0.9)[:nmat], 2e5, rng.normal(size=3))
    n, t1, t2 = rotated_frame(rng)
    f_bc, us_bc = O.hllc(p, x, x, n, t1, t2, bc=bc)
    f_in, us_in = O.hllc(p, x, x, n, t1, t2, bc=0)
    assert np.array_equal(f_bc, f_in) and np.array_equal(us_bc, us_in)


@pytest.mark.parametrize("nmat", [2, 3])
@pytest.mark.parametrize("bc", [2, 3])
def test_the_ghost_is_the_state_at_the_fixed_pressure(nmat, bc):
    rng = np.random.default_rng(10 * bc + nmat)
    P_in, P_out = 150e6, 1e6
    p = params(nmat, fixed_p=(P_in, P_out))
    q = params(nmat)
    for trial in range(10):
        a = rng.dirichlet(np.ones(nmat))
        rho = (1000.0, 1.2, 0.9)[:nmat] * rng.uniform(0.9, 1.1, size=nmat)
        u = 20.0 * rng.normal(size=3)
        x = state(p, a, rho, rng.uniform(1e5, 5e7), u)
        ghost = state(p, a, rho, P_in if bc == 3 else P_out, u)
        n, t1, t2 = rotated_frame(rng)
        f_bc, us_bc = O.hllc(p, x, x, n, t1, t2, bc=bc)
        f_in, us_in = O.hllc(q, x, ghost, n, t1, t2, bc=0)
        scale = np.abs(f_in).max()
        assert np.abs(f_bc - f_in).max() <= 1e-13 * scale
        assert np.abs(us_bc - us_in).max() <= 1e-12 * (np.abs(us_in).max() + 1.0)


def test_rest_at_the_outlet_pressure_stays_at_rest():
    P_out = 0.1
    p, m, S = sod_case(level=4, prelax=1, bc=(2, 2, 2, 2, 2, 2))
    p.bc_fixed_p, p.p_inlet, p.p_outlet = 1, 1.0, P_out
    x = state(p, (0.5, 0.5), (0.125, 0.125), P_out, (0.0, 0.0, 0.0))
    S[:] = x
    om = O.OracleMesh(m)
    S1 = om.stage(p, 1, S.copy(), S.copy())[0]
    assert np.abs(S1[:, 2:5]).max() <= 1e-14
    assert np.abs(S1 - S).max() <= 1e-14 * np.abs(S).max()


def test_an_inlet_at_higher_pressure_drives_inflow_through_its_side():
    p, m, S = sod_case(level=4, prelax=1, bc=(4, 3, 4, 4, 4, 4))   # inlet on +x
    p.bc_fixed_p, p.p_inlet, p.p_outlet = 1, 1.0, 0.1
    S[:] = state(p, (0.5, 0.5), (0.125, 0.125), 0.1, (0.0, 0.0, 0.0))
    om = O.OracleMesh(m)
    S1 = om.stage(p, 1, S.copy(), S.copy())[0]
    c = m.centres()
    h = c[:, 0].max()
    plus_x = np.isclose(c[:, 0], h)
    assert np.all(S1[plus_x, 2] < 0.0)                           # rho u points into the domain
    assert np.abs(S1[~plus_x, 2:5]).max() <= 1e-14               # the other leaves see no face change yet
    assert np.all(S1[plus_x, 0] + S1[plus_x, 1] > S[plus_x, 0] + S[plus_x, 1])   # mass enters


# ---------------------------------------------------------------- device

def inlet_outlet_case(level=4, prelax=1, **kw):
    p, m, S = sod_case(level=level, prelax=prelax, bc=(3, 2, 4, 1, 2, 5), **kw)
    c = m.centres()
    S[:, 2] = 0.05 * np.sin(6.0 * c[:, 1]) * (S[:, 0] + S[:, 1])
    r = S[:, 0] + S[:, 1]
    S[:, 5] = S[:, 7] + S[:, 8] + 0.5 * S[:, 2] ** 2 / r
    p.bc_fixed_p, p.p_inlet, p.p_outlet = 1, 1.2, 0.08
    return p, m, S


def say(*a):
    print(*a, flush=True)


def rel_l1(a, b):
    return np.abs(a - b).sum(axis=0) / (np.abs(b).sum(axis=0) + 1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("prelax", [0, 1])
def test_device_exact_path_matches_oracle_bitwise(prelax):
    p, m, S = inlet_outlet_case(prelax=prelax)
    say(f"[fixed-p exact] prelax={prelax} n={m.n}")
    dev = F.Device(p, m, prec=F.PREC_EXACT)
    dev.set_state(S)
    om = O.OracleMesh(m)
    SN, Sc = S.copy(), S.copy()
    for irk in (1, 2, 3):
        Sc, SF, _ = om.stage(p, irk, SN, Sc)
        dev.stage(irk)
        assert np.array_equal(dev.get_faces(), SF), f"faces of stage {irk}"
        assert np.array_equal(dev.get_state(), Sc), f"state of stage {irk}"
    dev.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_device_fast_paths_match_oracle_over_steps(prec):
    p, m, S = inlet_outlet_case(level=5)
    say(f"[fixed-p {prec}] n={m.n}")
    So = O.OracleMesh(m).advance(p, S, 20)
    out = []
    for fused in (False, True):
        dev = F.Device(p, m, prec=F.PREC_FP64 if prec == "fp64" else F.PREC_MIXED)
        dev.set_fused(fused)
        dev.set_state(S)
        dev.advance(20)
        out.append(dev.get_state())
        dev.close()
    for Sg in out:
        err = rel_l1(Sg, So)
        say(f"[fixed-p {prec}] rel L1 per field {err}")
        assert np.all(err <= (1e-10 if prec == "fp64" else 1e-4))
    # the fixed pressures act: the plain copies give another state
    q, _, _ = inlet_outlet_case(level=5)
    q.bc_fixed_p = 0
    assert rel_l1(O.OracleMesh(m).advance(q, S, 20), So).max() > 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("bad", ["flag", "pressure"])
def test_device_rejects_bad_fixed_pressure_settings(bad):
    p, m, S = inlet_outlet_case(level=3)
    if bad == "flag":
        p.bc_fixed_p = 2
    else:
        p.p_inlet = float("nan")
    with pytest.raises(F.FVError):
        F.Device(p, m)
