"""Throughput of the explicit finite-volume update on one MI355X.

A step is one RK stage (face reconstruction + HLLC residual + RK update +
pressure relaxation) over every leaf of the workload mesh, with the state
resident in HBM.  The default workload is the 3-D Rayleigh collapse of an air
bubble in water (4 mm cube, R = 0.4 mm, water stiffened gas at 50 bar, air at
1 bar, tanh-smeared interface) on an octree adapted to reference levels 5-9
(16^3 to 256^3 equivalent) by the interface and pressure-jump indicator, the
plain FP64 path.  Synthetic initial data, no I/O.

Prints one JSON line (metric of the contract) and, with --profile-only,
nothing but runs the timed loop (for rocprofv3).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import forestfv_public_amd as F  # noqa: E402

HBM_PEAK_GBPS = 8000.0
METRIC = "Mcell-updates/s per RK stage (plain FP64 vs FP32+MLP) @ L1 err; HBM GB/s % peak"

L_DOMAIN = 4e-3
R_BUBBLE = 0.4e-3
WATER_AIR = dict(gamma=(6.12, 1.4), pinf=(3.43e8, 0.0))
RHO0 = (1000.0, 1.2)
P0 = (5e6, 1e5)


def algorithmic_bytes(neq):
    """Per leaf per RK stage: read S^k, read S^n, write S^k+1 in FP64, six int32
    neighbour ids and an 8-byte volume: B = 24 NEQ + 32."""
    return 24 * neq + 32


def traffic_bytes(workload, prec, kernel):
    """HBM bytes per launch of `kernel` measured by rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes (calibrated against a copy kernel of known size), from the
    summaries committed under profiles/; None where no such measurement exists."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as fh:
            t = json.load(fh)
        return t[workload][prec][kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def dominant_kernel(prec, fused, nbricks, n):
    """(kernel name, leaves one launch updates) of the kernel that dominates a stage."""
    if fused and prec != "exact":
        return "k_stage_fused", n
    if prec != "exact" and nbricks > 0:
        return "k_stage_brick", 64 * nbricks
    return {"exact": "k_flux_rk", "fp64": "k_flux_fast", "mixed": "k_flux_pipe"}[prec], n


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def c1_params(dt=None, prelax=1):
    return F.make_params(nmat=2, gamma=WATER_AIR["gamma"], pinf=WATER_AIR["pinf"], amin=1e-6, avf_lim=0.5,
                         prelax=prelax, rk=F.RK_SSP3, limiter=0, dt=dt if dt else 1e-9, bc=(2, 2, 2, 2, 2, 2))


def c1_crit(maxlevel, layers):
    return F.make_crit(uniform_on=1, uniform_lvl=5, prs_on=1, prs_lvl=maxlevel, prs_thr=0.2,
                       inter0_on=1, inter0_lvl=maxlevel, inter0_lrs=layers, inter1_on=0)


def build_c1(maxlevel=9, base=5, layers=64, log_fn=log):
    """Geometric pre-adaptation (main.c:107-134): init, indicator, smoothing, refine, repeated."""
    p = c1_params()
    m = F.Mesh.uniform((1, 1, 1), (0, 0, 0), L_DOMAIN, base, list(p.bc))
    crit = c1_crit(maxlevel, layers)
    c = (0.5 * L_DOMAIN,) * 3
    for ipass in range(maxlevel - base + 1):
        S = m.init_bubble(p, RHO0, P0, c, R_BUBBLE, 2.0, 1234, 0.01)
        dev = F.Device(p, m)
        dev.set_state(S)
        SI, split = dev.indicator(crit)
        dev.close()
        m.set_SI(SI)
        split = m.criterionsmth(maxlevel, split)
        S = m.adapt(p, split, S)
        log_fn(f"  geo pass {ipass}: {m.n} leaves, levels {np.bincount(m.arrays()['level'])[base:]}")
    S = m.init_bubble(p, RHO0, P0, c, R_BUBBLE, 2.0, 1234, 0.01)
    dev = F.Device(p, m)
    dev.set_state(S)
    SI, _ = dev.indicator(crit)
    dev.close()
    m.set_SI(SI)
    hmin = L_DOMAIN / (1 << (maxlevel - 1))
    cw = np.sqrt(6.12 * (P0[0] + 3.43e8) / 1000.0)
    p.dt = 0.4 * hmin / cw
    return p, m, S


def build_uniform(level):
    p = c1_params()
    m = F.Mesh.uniform((1, 1, 1), (0, 0, 0), L_DOMAIN, level, list(p.bc))
    S = m.init_bubble(p, RHO0, P0, (0.5 * L_DOMAIN,) * 3, R_BUBBLE, 2.0, 1234, 0.01)
    hmin = L_DOMAIN / (1 << (level - 1))
    p.dt = 0.4 * hmin / np.sqrt(6.12 * (P0[0] + 3.43e8) / 1000.0)
    return p, m, S


def cpu_baseline(p, m, S, workload, seconds=6.0):
    """The CPU restatement (oracle/, gcc -ffp-contract=off) timed on the host on the
    headline workload itself (the same mesh, parameters and state as the GPU line),
    on whole RK stages: the -O2 build with OpenMP over all the threads the process
    may use (mirrored faces recomputed, bit-identical to the sequential run) for
    about `seconds`, the -O2 build on one core for one stage, and the -O0 build (the
    reference makefile's flags, src/makefile:39) on one core for one stage.  Also
    checks one device stage against the CPU stage on that workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    t0 = time.perf_counter()
    om = O.OracleMesh(m)
    t_geo = time.perf_counter() - t0
    nthr = int(O.lib().orc_num_threads())
    # OpenMP, whole SSP-RK3 steps
    t0 = time.perf_counter()
    om.advance_par(p, S, 1, nthr)
    one = time.perf_counter() - t0
    nsteps = max(1, int(seconds / max(one, 1e-3)))
    t0 = time.perf_counter()
    om.advance_par(p, S, nsteps, nthr)
    t_par = time.perf_counter() - t0
    stages_par = 3 * nsteps
    # one core, one stage (stage 1 from SN = S), -O2 and -O0
    t0 = time.perf_counter()
    S1, _, _ = om.stage(p, 1, S, S)
    t_1 = time.perf_counter() - t0
    om0 = O.OracleMesh(m, variant="O0", geometry_from=om)
    t0 = time.perf_counter()
    S1_O0, _, _ = om0.stage(p, 1, S, S)
    t_0 = time.perf_counter() - t0
    # the device's stage 1 on the same state against the CPU stage
    dev = F.Device(p, m)
    dev.set_state(S)
    dev.stage(1)
    Sg = dev.get_state()
    dev.close()
    num = np.abs(Sg - S1).sum(0)
    den = np.abs(S1).sum(0)
    mscale = np.sqrt((S1[:, 2:5] ** 2).sum(1)).sum()
    den[2:5] = np.maximum(den[2:5], mscale)
    den[den == 0] = 1
    l1 = float((num / den).max())
    same_O0 = bool(np.array_equal(S1, S1_O0))
    n = m.n
    return {
        "value": n * stages_par / t_par / 1e6, "unit": "Mcell-updates/s per RK stage", "cores": nthr, "kind": "port",
        "sample": f"{workload} ({n} leaves), SSP-RK3, PRELAX {p.prelax}, {stages_par} stages in {t_par:.1f} s, "
                  f"gcc -O2 -ffp-contract=off, OpenMP {nthr} threads (os.cpu_count {os.cpu_count()}); geometry "
                  f"{t_geo:.1f} s not timed",
        "value_1core": n / t_1 / 1e6,
        "sample_1core": f"{workload}, RK stage 1, {t_1:.1f} s, gcc -O2, 1 thread",
        "value_O0": n / t_0 / 1e6,
        "sample_O0": f"{workload}, RK stage 1, {t_0:.1f} s, gcc -O0 (the reference makefile's flags), 1 thread; "
                     f"bit-identical to -O2: {same_O0}",
    }, l1, 1


def mixture_pressure(p, S):
    """Sum_k alpha_k p_k = Sum_k (gamma_k - 1) (alpha rho e)_k - alpha_k gamma_k pinf_k (two materials)."""
    a0 = S[:, 6]
    g, pi = p.gamma, p.pinf
    return ((g[0] - 1.0) * S[:, 7] - a0 * g[0] * pi[0]) + ((g[1] - 1.0) * S[:, 8] - (1.0 - a0) * g[1] * pi[1])


def m5_mixture_pressure(p, S, chunk=1 << 22):
    """sum_k alpha_k p_k of model-5 states in the mixture mode: p_k = (gamma_k - 1) (alpha rho e_k
    - alpha_k E_k(G)) / alpha_k - gamma_k pinf_k with E_k = mu_k / (4 |G|^(1/6)) tr((G - |G|^(1/3)
    I)^2), G = A A^T (state_equation.c:269-316 weighted by alpha_k); fractions as stored (the
    record's clip at AMIN / 100 does not change them in these states)."""
    nm = p.nmat
    out = np.empty(S.shape[0])
    for a0 in range(0, S.shape[0], chunk):
        X = S[a0:a0 + chunk]
        al = np.empty((X.shape[0], nm))
        al[:, :nm - 1] = X[:, nm + 4:2 * nm + 3]
        al[:, nm - 1] = 1.0 - al[:, :nm - 1].sum(1)
        are = X[:, 2 * nm + 3:3 * nm + 3]
        A = X[:, 3 * nm + 3:].reshape(-1, 3, 3)
        G = A @ np.transpose(A, (0, 2, 1))
        d = np.abs(np.linalg.det(G))
        with np.errstate(divide="ignore", invalid="ignore"):
            M = G - np.cbrt(d)[:, None, None] * np.eye(3)
            tr = np.einsum("nij,nji->n", M, M)
            E1 = np.where(d != 0.0, tr / (4.0 * d ** (1.0 / 6.0)), 0.0)
        pm = np.zeros(X.shape[0])
        for k in range(nm):
            g, pi, mu = p.gamma[k], p.pinf[k], p.mush[k]
            pk = (g - 1.0) * (are[:, k] - al[:, k] * mu * E1) / al[:, k] - g * pi
            pm += al[:, k] * pk
        out[a0:a0 + chunk] = pm
    return out


def rel_l1(a, b):
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def gradient_crit(crit):
    """The gradient part of an indicator (criterion_grad thresholds and the uniform
    level, without the interface layers): the labels the network learns.  The
    interface band is not a function of a 3x3x3 stencil; the plain indicator keeps
    forcing it when the network decides."""
    g = F.make_crit()
    C.memmove(C.addressof(g), C.addressof(crit), C.sizeof(crit))
    g.inter0_on = 0
    g.inter1_on = 0
    return g


def build_variant(p_liquid, centre, base=4, maxlevel=8, layers=16, seed=1234, mlp=None):
    """A seeded variant of the bubble case (driving pressure, bubble position) on a
    pre-adapted forest, as build_c1 builds C1; with mlp, every pre-adaptation pass
    is sampled too (those passes carry the refine labels)."""
    p = c1_params()
    m = F.Mesh.uniform((1, 1, 1), (0, 0, 0), L_DOMAIN, base, list(p.bc))
    crit = c1_crit(maxlevel, layers)
    crit.uniform_lvl = base
    for _ in range(maxlevel - base + 1):
        S = m.init_bubble(p, RHO0, (p_liquid, P0[1]), centre, R_BUBBLE, 2.0, seed, 0.01)
        dev = F.Device(p, m)
        dev.set_state(S)
        if mlp is not None:
            mlp.sample(dev, gradient_crit(crit))
        SI, split = dev.indicator(crit)
        dev.close()
        m.set_SI(SI)
        split = m.criterionsmth(maxlevel, split)
        S = m.adapt(p, split, S)
    S = m.init_bubble(p, RHO0, (p_liquid, P0[1]), centre, R_BUBBLE, 2.0, seed, 0.01)
    hmin = L_DOMAIN / (1 << (maxlevel - 1))
    p.dt = 0.4 * hmin / np.sqrt(6.12 * (p_liquid + 3.43e8) / 1000.0)
    return p, m, S, crit


def train_on_variants(mlp, log_fn=log, nvariants=12):
    """Config 4's training set: stencil/flag pairs of the plain indicator on seeded
    variants of the bubble case (driving pressure 20-200 bar, bubble centre moved by
    up to a radius), sampled at every pre-adaptation pass, at the start and after 15
    steps.  Returns the number of pairs."""
    Lc = 0.5 * L_DOMAIN
    variants = [(pl, (Lc + dx, Lc + dy, Lc)) for pl in (2e6, 5e6, 1e7, 2e7)
                for dx, dy in ((0.0, 0.0), (0.8 * R_BUBBLE, -0.5 * R_BUBBLE), (-0.6 * R_BUBBLE, 0.0))]
    variants = variants[::len(variants) // nvariants][:nvariants]
    for k, (pl, c) in enumerate(variants):
        p, m, S, crit = build_variant(pl, c, seed=100 + k, mlp=mlp)
        dev = F.Device(p, m)
        dev.set_state(S)
        g = gradient_crit(crit)
        mlp.sample(dev, g)
        dev.advance(15)
        mlp.sample(dev, g)
        dev.close()
    X, cls, sm = mlp.dataset()
    del X
    freq = np.bincount(cls, minlength=3).astype(float)
    cw = (freq.sum() / (3.0 * np.maximum(freq, 1.0))).astype(np.float32)
    mlp.set_options(class_weights=cw)
    log_fn(f"  training set: {len(cls)} pairs from {len(variants)} variants, class frequencies {freq.astype(int)}")
    return len(cls)


def holdout_accuracy(mlp, p, m, S, crit):
    """Balanced accuracy of the network's refine/keep/coarsen against the plain
    gradient indicator on a state it was not trained on."""
    probe = F.MLP(seed=5)
    dev = F.Device(p, m)
    dev.set_state(S)
    probe.sample(dev, gradient_crit(crit))
    dev.close()
    X, ct, st = probe.dataset()
    probe.close()
    out = mlp.infer(X)
    pc = np.argmax(out[:, :3], 1)
    bal = float(np.mean([(pc[ct == k] == k).mean() for k in range(3) if (ct == k).any()]))
    smooth_acc = float(((out[:, 3] > 0).astype(np.float32) == st).mean())
    return bal, smooth_acc


def fp32_mlp_leg(p, m, S, crit, cycles=4, every=5, log_fn=log):
    """FP32 faces + learned indicator: the network is trained on the device from
    the plain indicator's labels on this state, then every `every` steps it
    proposes targets and the per-leaf limiter (smooth -> MC); the stages run in
    mixed precision.  Returns cell-updates/s with the indicator's cost included,
    and the relative L1 in density and pressure against the FP64 path driven the
    same way."""
    nstage = 3
    t0 = time.perf_counter()
    mlp = F.MLP(seed=7)
    npairs = train_on_variants(mlp, log_fn)
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = mlp.train(2, batch=4096, lr=2e-3)
    t_train = time.perf_counter() - t0
    bal, sacc = holdout_accuracy(mlp, p, m, S, crit)
    log_fn(f"  learned indicator: {npairs} pairs, 2 epochs, loss {loss:.4f}, {t_train:.2f} s; on the held-out C1 "
           f"state balanced accuracy {bal:.3f}, smooth-flag accuracy {sacc:.3f}")
    q = F.make_params()
    C.memmove(C.addressof(q), C.addressof(p), C.sizeof(p))
    q.limiter = 2
    out = {}
    # plain: FP64 stages, plain indicator (minmod); learned: mixed stages, network
    # indicator (per-leaf limiter); "fp64_learned" runs FP64 stages with the
    # network and is the reference for the precision error
    for leg, prec, par, learned in (("plain", F.PREC_FP64, p, False), ("learned", F.PREC_MIXED, q, True),
                                    ("fp64_learned", F.PREC_FP64, q, True)):
        dev = F.Device(par, m, prec=prec)
        for rep in range(2):               # the first pass warms up
            dev.set_state(S)
            t1 = time.perf_counter()
            for _ in range(cycles if rep else 1):
                # SI and the targets stay on the device, as the device regrid reads them
                if learned:
                    dev.indicator_mlp(mlp, crit, smooth=False, to_host=False)
                else:
                    dev.indicator(crit, to_host=False)
                dev.advance(every)
            t = time.perf_counter() - t1
        out[leg] = (dev.get_state(), t)
        dev.close()
    mlp.close()
    stages = cycles * every * nstage
    rate = {k: m.n * stages / v[1] / 1e6 for k, v in out.items()}
    Sm, Sd = out["learned"][0], out["fp64_learned"][0]
    return {
        "value": round(rate["learned"], 2), "unit": "Mcell-updates/s per RK stage",
        "plain_same_driver": round(rate["plain"], 2),
        "speedup_vs_plain_same_driver": round(rate["learned"] / rate["plain"], 3),
        "stages": stages, "indicator_every_steps": every,
        "limiter": "per leaf (learned smooth flag: MC, else minmod)",
        "l1_rel_rho_vs_fp64": rel_l1(Sm[:, 0] + Sm[:, 1], Sd[:, 0] + Sd[:, 1]),
        "l1_rel_p_vs_fp64": rel_l1(mixture_pressure(q, Sm), mixture_pressure(q, Sd)),
        "train_s": round(t_train, 2), "train_pairs": npairs, "dataset_build_s": round(t_data, 1),
        "train_set": "12 seeded bubble variants (20-200 bar, centre moved by up to a radius), levels 4-8, "
                     "sampled at every pre-adaptation pass, at the start and after 15 steps; labels from the "
                     "gradient criteria (interface band left to the plain indicator, which is forced)",
        "holdout_c1_balanced_accuracy": round(bal, 4), "holdout_c1_smooth_flag_accuracy": round(sacc, 4),
    }


def regrid_leg(p, m, S, crit, maxlevel=9, every=5, cycles=2, log_fn=log):
    """Adaptation every `every` steps as loop.c:19-21 runs it, on the C1 mesh:
    fv_dev_regrid (indicator, criterionsmth, merge/split, connectivity on the
    device) against the host path (indicator, then fv_criterionsmth +
    fv_mesh_adapt + fv_mesh_pack on the host and a re-upload).  Reports the
    regrid times and the loop's cell-updates/s with each."""
    nstage = 3
    out = {}
    for leg in ("device", "host"):
        if leg == "device":
            mesh = m
        else:                                                # the host leg adapts its own copy
            a = m.arrays()
            mm = m.m
            mesh = F.Mesh.from_leaves(tuple(mm.nroot), tuple(mm.origin), mm.rootsize, list(p.bc), a["root"],
                                      a["level"], a["ci"])
            mesh.set_SI(a["SI"])
        dev = F.Device(p, mesh)
        dev.set_state(S)
        if leg == "device":
            dev.set_forest(mesh)
        t_reg = t_step = 0.0
        cells = 0
        sizes = []
        for c in range(cycles):
            dev.advance(1)                                   # settle the kernels before timing
            t0 = time.perf_counter()
            if leg == "device":
                dev.regrid(crit, maxlevel)
            else:
                SI, split = dev.indicator(crit)
                mesh.set_SI(SI)
                split = mesh.criterionsmth(maxlevel, split)
                S2 = mesh.adapt(p, split, dev.get_state())
                dev.reattach(mesh)
                dev.set_state(S2)
            t_reg += time.perf_counter() - t0
            sizes.append(dev.n)
            ms, _ = dev.time_stages(every * nstage)
            t_step += ms * 1e-3 * every * nstage
            cells += dev.n * every * nstage
        dev.close()
        out[leg] = dict(ms=1e3 * t_reg / cycles, loop=cells / (t_step + t_reg) / 1e6, sizes=sizes)
        log_fn(f"  regrid {leg}: {out[leg]['ms']:.1f} ms per regrid, leaves {sizes}, "
               f"loop {out[leg]['loop']:.1f} Mcell-updates/s")
    return {
        "device_ms_per_regrid": round(out["device"]["ms"], 2), "host_path_ms_per_regrid": round(out["host"]["ms"], 1),
        "leaves_after": out["device"]["sizes"], "same_leaf_counts": out["device"]["sizes"] == out["host"]["sizes"],
        "every_steps": every,
        "loop_value_device_regrid": round(out["device"]["loop"], 2),
        "loop_value_host_regrid": round(out["host"]["loop"], 2),
        "unit": "Mcell-updates/s per RK stage, regrid time included",
    }


def collapse_leg(p, layers=64, log_fn=log):
    """The bubble to its first collapse with the C driver (forestfv_public_amd/fv_run):
    the AMR loop in FP64 with a device regrid every 5 steps, then the precision report
    at that time: FP64 and mixed stages on the fixed C1 mesh with the same fixed dt,
    per-field relative L1 and Linf of the mixed state against the FP64 one."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "forestfv_public_amd", "fv_run")

    def run(*args):
        r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise SystemExit(f"fv_run {args} failed: {r.stderr[-500:]}")
        return json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][-1])

    amr = run("collapse", 5, 9, layers, 5, F.PREC_FP64, 6000)
    log_fn(f"  collapse (AMR, FP64): {amr['steps']} steps, t = {amr['t_collapse_s']:.3e} s = "
           f"{amr['t_collapse_over_rayleigh']:.3f} Rayleigh, R_min/R0 {amr['R_min_over_R0']:.3f}, "
           f"loop {amr['loop_rate_Mcell_s']:.0f} Mcell-updates/s")
    hmin = L_DOMAIN / 256
    dt = 0.25 * hmin / np.sqrt(6.12 * (5e6 + 3.43e8) / 1000.0)
    nsteps = int(np.ceil(amr["t_collapse_s"] / dt))
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for prec, key in ((F.PREC_FP64, "fp64"), (F.PREC_MIXED, "mixed")):
            path = os.path.join(td, key + ".bin")
            r = run("collapse", 5, 9, layers, 0, prec, nsteps, repr(dt), path)
            n, nq = np.fromfile(path, dtype=np.int32, count=2)
            out[key] = (r, np.fromfile(path, dtype=np.float64, offset=8).reshape(n, nq))
    (r64, S64), (rmx, Smx) = out["fp64"], out["mixed"]
    names = ["alpha_rho_water", "alpha_rho_air", "rho_u", "rho_v", "rho_w", "rho_E", "alpha_water",
             "alpha_rho_e_water", "alpha_rho_e_air"]
    fields = {}
    for k, nm in enumerate(names):
        a, b = Smx[:, k], S64[:, k]
        fields[nm] = {"l1": rel_l1(a, b), "linf": float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))}
    for nm, a, b in (("rho", Smx[:, 0] + Smx[:, 1], S64[:, 0] + S64[:, 1]),
                     ("pressure", mixture_pressure(p, Smx), mixture_pressure(p, S64))):
        fields[nm] = {"l1": rel_l1(a, b), "linf": float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))}
    log_fn(f"  collapse time, fixed mesh, {nsteps} steps: mixed vs FP64 rho L1 {fields['rho']['l1']:.2e} "
           f"Linf {fields['rho']['linf']:.2e}, p L1 {fields['pressure']['l1']:.2e} Linf {fields['pressure']['linf']:.2e}")
    return {
        "amr_fp64": {k: amr[k] for k in ("steps", "t_collapse_s", "rayleigh_time_s", "t_collapse_over_rayleigh",
                                         "R_min_over_R0", "leaves_start", "leaves_min", "leaves_end", "regrids",
                                         "regrid_ms", "loop_rate_Mcell_s", "water_mass_change", "air_mass_change")},
        "fixed_mesh_to_collapse": {"steps": nsteps, "dt_s": dt, "leaves": int(S64.shape[0]),
                                   "stage_rate_fp64": r64["stage_rate_Mcell_s"],
                                   "stage_rate_mixed": rmx["stage_rate_Mcell_s"],
                                   "R_min_over_R0_fp64": r64["R_min_over_R0"], "R_min_over_R0_mixed": rmx["R_min_over_R0"],
                                   "mixed_vs_fp64": fields},
    }


# config 2 materials (solid slab, water, air) for the model-5 stage
C2_TABLE = dict(gamma=(3.0, 6.12, 1.4), pinf=(2.0e10, 3.43e8, 0.0), mush=(26e9, 0.0, 0.0))
C2_RHO = np.array([2700.0, 1000.0, 1.2])
C2_YIELD = 0.3e9      # von Mises yield stress of the slab, perfect plasticity


C2_Z_SLAB = 0.25 * L_DOMAIN                  # the slab fills the bottom quarter of the cube
C2_STANDOFF = 1.2                            # bubble centre to slab face, in bubble radii
C2_P = (1e7, 1e5)                            # water at 100 bar, air at 1 bar
C2_FIX = F.M5_FIX_SIGNS | F.M5_FIX_MIXTURE   # the consistent update with the weighted elastic mixture


def c2_params(prelax=3):
    return F.make_m5_params(nmat=3, gamma=C2_TABLE["gamma"], pinf=C2_TABLE["pinf"], mush=C2_TABLE["mush"],
                            amin=1e-6, avf_lim=0.5, prelax=prelax, order=2, rk_steps=3, fix=C2_FIX,
                            yield_stress=C2_YIELD)


def c2_crit(maxlevel, base, layers):
    return F.make_crit(uniform_on=1, uniform_lvl=base, prs_on=1, prs_lvl=maxlevel, prs_thr=0.2,
                       inter0_on=1, inter0_lvl=maxlevel, inter0_lrs=layers, inter1_on=0)


def c2_state(p, m, dx=0.0, standoff=C2_STANDOFF, p_liquid=C2_P[0]):
    """Config 2 initial data on mesh m: the solid slab below z = 1 mm with a sharp face, the
    0.4 mm air bubble `standoff` radii above it smeared over two finest cells, water at p_liquid
    (100 bar), air at 1 bar, A = I in the solid and 0 in the fluids, at rest
    (fv_m5_init_slab_bubble)."""
    centre = (0.5 * L_DOMAIN + dx, 0.5 * L_DOMAIN, C2_Z_SLAB + standoff * R_BUBBLE)
    return m.init_m5_slab_bubble(p, C2_RHO, C2_Z_SLAB, centre, R_BUBBLE, p_liquid, C2_P[1], 2.0)


def build_c2(maxlevel=9, base=5, layers=64, log_fn=log, mlp=None, **init):
    """Geometric pre-adaptation of config 2 on the device (main.c:107-134 order): initial data on
    the current mesh, then the model-5 indicator, criterionsmth and refinement
    (fv_m5_dev_regrid), repeated; the host mesh is rebuilt from the device's leaves each pass.
    With mlp, every pass is sampled for the learned indicator's training set (gradient labels)."""
    p = c2_params()
    m = F.Mesh.uniform((1, 1, 1), (0, 0, 0), L_DOMAIN, base, [2] * 6)
    crit = c2_crit(maxlevel, base, layers)
    p.dt = 1e-12
    dev = F.M5Device(p, m)
    dev.set_forest(m)
    for ipass in range(maxlevel - base + 1):
        dev.set_state(c2_state(p, m, **init))
        if mlp is not None:
            dev.sample_mlp(mlp, gradient_crit(crit))
        dev.regrid(crit, maxlevel)
        root, level, ci, SI = dev.get_forest()
        m = F.Mesh.from_leaves((1, 1, 1), (0, 0, 0), L_DOMAIN, [2] * 6, root, level, ci)
        m.set_SI(SI)
        log_fn(f"  c2 geo pass {ipass}: {m.n} leaves, levels {np.bincount(level)[base:]}")
    S = c2_state(p, m, **init)
    dev.set_state(S)
    m.set_SI(dev.interface_location(crit))
    return p, m, S, dev, crit


def gas_volume(S, level, L=L_DOMAIN):
    h = L / (1 << (level - 1))
    return float(((1.0 - S[:, 7] - S[:, 8]) * h ** 3).sum())


def c2_loop(dev, crit, maxlevel, nsteps, every=5, cfl=0.4, t_end=None, probe=0, log_fn=log, mlp=None):
    """loop.c:12-30 for model 5: the CFL step, one SSP-RK3 step (timed with hipEvents per stage),
    interface_location after it, and a device regrid every `every` steps (on the learned targets
    with their slope flags when mlp is given); stops after nsteps or at t_end.  probe > 0: the gas
    volume every `probe` steps."""
    t = 0.0
    ms_stage_tot = ms_res_tot = t_other = 0.0
    cells = 0
    leaves = [dev.n]
    vols = []
    step = 0
    t_wall = time.perf_counter()
    while step < nsteps and (t_end is None or t < t_end):
        t0 = time.perf_counter()
        if step > 0 and step % every == 0:
            if mlp is not None:
                dev.regrid_mlp(mlp, crit, maxlevel)
            else:
                dev.regrid(crit, maxlevel)
            leaves.append(dev.n)
        dt = dev.cfl_dt(cfl)
        dev.set_dt(dt)
        t_other += time.perf_counter() - t0
        ms_stage, ms_res = dev.time_steps(1)
        ms_stage_tot += 3 * ms_stage
        ms_res_tot += 3 * ms_res
        cells += 3 * dev.n
        t0 = time.perf_counter()
        dev.interface_location(crit, copy=False)   # SI stays on the device, as loop.c needs it
        t_other += time.perf_counter() - t0
        t += dt
        step += 1
        if probe and (step % probe == 0 or step == nsteps):
            root, level, ci, SI = dev.get_forest()
            vols.append((t, gas_volume(dev.get_state(), level)))
            log_fn(f"  c2 step {step}: t {t:.3e} s, {dev.n} leaves, gas volume {vols[-1][1]:.4e} m^3")
    wall = time.perf_counter() - t_wall
    return {"steps": step, "t": t, "leaves": leaves, "ms_stage": ms_stage_tot / (3 * step), "cells": cells,
            "ms_residual": ms_res_tot / (3 * step), "stage_rate": cells / (ms_stage_tot * 1e-3) / 1e6,
            "loop_rate": cells / wall / 1e6, "other_s": t_other, "wall_s": wall, "vols": vols}


def c2_leg(maxlevel=9, base=5, layers=64, steps=15, every=5, check=True, log_fn=log):
    """Config 2 at its size: the slab, water and bubble pre-adapted on the device to reference
    levels 5-9 (config levels 4-8) with 64 interface layers on both interfaces, then the model-5
    loop (CFL 0.4, SSP-RK3, PRELAX 3, von Mises 0.3 GPa, interface_location every step, a device
    regrid every 5 steps) in FP64 and in mixed precision.  The oracle is timed and checked on the
    32^3 uniform state."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    t0 = time.perf_counter()
    p, m, S, dev, crit = build_c2(maxlevel, base, layers, log_fn)
    setup = time.perf_counter() - t0
    out = {"workload": "c2_bubble_near_elastoplastic_slab_amr_L5-9", "levels": "5-9 (config 4-8)",
           "interface_layers": layers, "leaves_start": m.n, "setup_s": round(setup, 1),
           "model": "model 5, fix = FV_M5_FIX_SIGNS | FV_M5_FIX_MIXTURE, PRELAX 3, von Mises 0.3 GPa, SSP-RK3, CFL 0.4"}
    dev.advance(1)                       # settle the kernels before timing
    dev.set_state(S)
    r = c2_loop(dev, crit, maxlevel, steps, every, log_fn=log_fn)
    Sg = dev.get_state()
    dev.close()
    B = algorithmic_bytes(F.m5_neq(3))
    out.update({"steps": r["steps"], "leaves": r["leaves"], "ms_per_stage": round(r["ms_stage"], 3),
                "ms_residual_kernel": round(r["ms_residual"], 3), "value": round(r["stage_rate"], 2),
                "loop_value": round(r["loop_rate"], 2),
                "loop_ms_per_step": {"wall": round(1e3 * r["wall_s"] / r["steps"], 2),
                                     "stages": round(3 * r["ms_stage"], 2),
                                     "cfl_regrid_interface_location": round(1e3 * r["other_s"] / r["steps"], 2)},
                "finite": bool(np.all(np.isfinite(Sg))),
                "unit": "Mcell-updates/s per RK stage; loop_value with CFL, interface_location and regrids"})
    out["roofline"] = {"bound": "hbm", "algorithmic_bytes_per_leaf": B, "peak": HBM_PEAK_GBPS,
                       "achieved": round(B * r["stage_rate"] * 1e6 / 1e9, 1)}
    out["roofline"]["frac"] = round(out["roofline"]["achieved"] / HBM_PEAK_GBPS, 4)
    log_fn(f"[c2] {r['leaves']} leaves: {r['ms_stage']:.3f} ms/stage, residual {r['ms_residual']:.3f} ms, "
           f"{r['stage_rate']:.1f} Mcell-upd/s, loop {r['loop_rate']:.1f}")
    # mixed precision against FP64 on the fixed start mesh, the same CFL step, 10 steps
    mesh_res = {}
    for prec in (F.PREC_FP64, F.PREC_MIXED):
        d = F.M5Device(p, m)
        d.set_precision(prec)
        d.set_state(S)
        p.dt = d.cfl_dt(0.4)
        d.set_dt(p.dt)
        d.advance(1)
        d.set_state(S)
        ms = [d.time_steps(1)[0] for _ in range(10)]
        mesh_res[prec] = (d.get_state(), float(np.mean(ms)))
        d.close()
    S64, Smx = mesh_res[F.PREC_FP64][0], mesh_res[F.PREC_MIXED][0]
    nm = 3
    names = [f"alpha_rho_{k}" for k in range(nm)] + ["rho_u", "rho_v", "rho_w", "rho_E"] + \
        [f"alpha_{k}" for k in range(nm - 1)] + [f"alpha_rho_e_{k}" for k in range(nm)] + \
        [f"A_{a}{b}" for a in range(3) for b in range(3)]
    scale = np.abs(S64).sum(0)
    scale[nm:nm + 3] = np.sqrt((S64[:, nm:nm + 3] ** 2).sum(1)).sum()
    scale[-9:] = np.sqrt((S64[:, -9:] ** 2).sum(1)).sum()
    l1 = np.abs(Smx - S64).sum(0) / np.maximum(scale, 1e-300)
    rho64, rhomx = S64[:, :nm].sum(1), Smx[:, :nm].sum(1)
    out_l1_rho = rel_l1(rhomx, rho64)
    out_l1_p = rel_l1(m5_mixture_pressure(p, Smx), m5_mixture_pressure(p, S64))
    out["mixed"] = {"rel_l1_rho_vs_fp64_10steps": out_l1_rho, "rel_l1_p_vs_fp64_10steps": out_l1_p,
                    "ms_per_stage": round(mesh_res[F.PREC_MIXED][1], 3),
                    "ms_per_stage_fp64_same_mesh": round(mesh_res[F.PREC_FP64][1], 3),
                    "speedup_vs_fp64_stage": round(mesh_res[F.PREC_FP64][1] / mesh_res[F.PREC_MIXED][1], 3),
                    "rel_l1_vs_fp64_10steps": {k: float(v) for k, v in zip(names, l1)},
                    "finite": bool(np.all(np.isfinite(Smx)))}
    log_fn(f"[c2] mixed: {mesh_res[F.PREC_MIXED][1]:.3f} ms/stage against {mesh_res[F.PREC_FP64][1]:.3f}, "
           f"max field L1 vs FP64 after 10 steps {l1.max():.2e} ({names[int(l1.argmax())]})")
    if check:
        q = c2_params()
        mm = F.Mesh.uniform((1, 1, 1), (0, 0, 0), L_DOMAIN, 6, [2] * 6)
        S0 = c2_state(q, mm)
        om = O.OracleMesh(mm)
        q.dt = om.m5_cfl_dt(q, S0, 0.4)
        d = F.M5Device(q, mm)
        d.set_state(S0)
        d.advance(1)
        Sd = d.get_state()
        d.close()
        t1 = time.perf_counter()
        So = om.m5_advance(q, S0, 1)
        t_cpu = time.perf_counter() - t1
        num = np.abs(Sd - So).sum(axis=0)
        den = np.abs(So).sum(axis=0)
        den[nm:nm + 3] = np.sqrt((So[:, nm:nm + 3] ** 2).sum(1)).sum()
        den[-9:] = np.sqrt((So[:, -9:] ** 2).sum(1)).sum()
        out["rel_l1_vs_oracle_1step_32^3"] = float((num / np.maximum(den, 1e-300)).max())
        out["cpu_value_1core"] = round(mm.n * 3 / t_cpu / 1e6, 4)
        out["cpu_sample"] = f"32^3 config-2 state, 1 step (3 stages), {t_cpu:.1f} s, 1 thread, gcc -O2"
        out["speedup_vs_cpu_1core"] = round(out["value"] / out["cpu_value_1core"], 1)
    return out


def c2_collapse(maxlevel=8, base=5, layers=64, every=5, t_over_rayleigh=1.6, log_fn=log):
    """Config 2 run past the bubble's collapse on reference levels 5-8: the model-5 loop as in
    c2_leg, with the gas volume every 50 steps; the collapse is the first minimum of the gas volume."""
    p, m, S, dev, crit = build_c2(maxlevel, base, layers, log_fn)
    t_r = 0.915 * R_BUBBLE * np.sqrt(1000.0 / (C2_P[0] - C2_P[1]))
    v0 = gas_volume(S, m.arrays()["level"])
    r = c2_loop(dev, crit, maxlevel, 10 ** 7, every, t_end=t_over_rayleigh * t_r, probe=50, log_fn=log_fn)
    dev.close()
    ts = np.array([v[0] for v in r["vols"]])
    vs = np.array([v[1] for v in r["vols"]])
    k = int(np.argmax((vs[1:-1] <= vs[:-2]) & (vs[1:-1] < vs[2:]))) + 1 if len(vs) > 2 else int(np.argmin(vs))
    return {"workload": f"c2_collapse_amr_L{base}-{maxlevel}", "steps": r["steps"], "t_end_s": r["t"],
            "rayleigh_time_s": t_r, "t_collapse_over_rayleigh": float(ts[k] / t_r),
            "gas_volume_min_over_v0": float(vs[k] / v0), "R_min_over_R0": float((vs[k] / v0) ** (1.0 / 3.0)),
            "leaves_start": r["leaves"][0], "leaves_min": int(min(r["leaves"])), "leaves_max": int(max(r["leaves"])),
            "regrids": len(r["leaves"]) - 1, "stage_rate": round(r["stage_rate"], 1),
            "loop_rate": round(r["loop_rate"], 1), "wall_s": round(r["wall_s"], 1),
            "gas_volume_over_v0": [[round(float(t / t_r), 4), round(float(v / v0), 4)] for t, v in zip(ts, vs)]}


def train_on_c2_variants(mlp, log_fn=log):
    """The model-5 half of config 4's training set: stencil/flag pairs of the plain gradient
    indicator on seeded variants of config 2 (stand-off 1.0-2.0 radii, driving pressure 20-200
    bar) pre-adapted to reference levels 4-7, sampled at every pre-adaptation pass, at the start
    and after 15 steps."""
    nv = 0
    for standoff in (1.0, 1.5, 2.0):
        for pl in (2e6, 2e7):
            p, m, S, dev, crit = build_c2(maxlevel=7, base=4, layers=16, log_fn=lambda *a: None, mlp=mlp,
                                          standoff=standoff, p_liquid=pl)
            g = gradient_crit(crit)
            dev.sample_mlp(mlp, g)
            for _ in range(15):
                dev.set_dt(dev.cfl_dt(0.4))
                dev.advance(1)
            dev.sample_mlp(mlp, g)
            dev.close()
            nv += 1
    return nv


def train_c4_network(log_fn=log, seed=7):
    """Config 4's network: trained on the device on about 5 M stencil/flag pairs of the plain
    indicator from seeded variants of configs 1 and 2 (model 4 bubbles and model 5 slab-and-bubble
    cases, 20-200 bar, bubble moved / stand-off 1.0-2.0 radii), two epochs of Adam."""
    mlp = F.MLP(seed=seed)
    t0 = time.perf_counter()
    train_on_variants(mlp, log_fn, nvariants=6)
    n1 = mlp.dataset_size()
    nv = train_on_c2_variants(mlp, log_fn)
    npairs = mlp.dataset_size()
    X, cls, sm = mlp.dataset()
    del X
    freq = np.bincount(cls, minlength=3).astype(float)
    mlp.set_options(class_weights=(freq.sum() / (3.0 * np.maximum(freq, 1.0))).astype(np.float32))
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = mlp.train(2, batch=4096, lr=2e-3)
    t_train = time.perf_counter() - t0
    log_fn(f"  c4 training set: {npairs} pairs ({n1} from 6 config-1 variants, {npairs - n1} from {nv} config-2 "
           f"variants), classes {freq.astype(int)}; 2 epochs, loss {loss:.4f}, {t_train:.2f} s")
    return mlp, {"train_pairs": int(npairs), "pairs_config1": int(n1), "pairs_config2": int(npairs - n1),
                 "class_frequencies": [int(v) for v in freq], "loss": round(float(loss), 4),
                 "dataset_build_s": round(t_data, 1), "train_s": round(t_train, 2)}


def c4_leg(maxlevel=10, base=5, layers=64, steps=10, every=5, log_fn=log):
    """Config 4: config 2 on reference levels 5-10 (config 4-9) with the learned indicator driving
    the regrids (the plain refinement forced, the smooth flag choosing MC per leaf) and the
    mixed-precision stage, against the plain FP64 path (plain indicator, minmod) in the same
    driver from the same mesh and state."""
    mlp, train = train_c4_network(log_fn)
    t0 = time.perf_counter()
    p, m, S, dev, crit = build_c2(maxlevel, base, layers, log_fn)
    setup = time.perf_counter() - t0
    dev.close()
    out = {"workload": "c4_learned_indicator_c2_amr_L5-10", "levels": "5-10 (config 4-9)", "interface_layers": layers,
           "leaves_start": m.n, "setup_s": round(setup, 1), "training": train}
    res = {}
    for leg, prec, limiter, learned in (("plain", F.PREC_FP64, 0, False), ("learned", F.PREC_MIXED, 2, True)):
        q = c2_params()
        q.limiter = limiter
        q.dt = 1e-12
        d = F.M5Device(q, m)
        d.set_precision(prec)
        d.set_forest(m)
        d.set_state(S)
        d.advance(1)                      # settle the kernels
        d.set_state(S)
        if learned:
            _, _, flags = d.indicator_mlp(mlp, crit)    # the slope flags of the start mesh
            res["mlp_flags"] = flags
        r = c2_loop(d, crit, maxlevel, steps, every, log_fn=log_fn, mlp=mlp if learned else None)
        res[leg] = r
        res[leg + "_state"] = d.get_state()
        d.close()
        log_fn(f"[c4] {leg}: leaves {r['leaves']}, {r['ms_stage']:.2f} ms/stage, {r['stage_rate']:.1f} Mcell-upd/s "
               f"per stage, loop {r['loop_rate']:.1f}")
    mlp.close()
    for leg in ("plain", "learned"):
        r = res[leg]
        out[leg] = {"leaves": r["leaves"], "steps": r["steps"], "ms_per_stage": round(r["ms_stage"], 3),
                    "value": round(r["stage_rate"], 2), "loop_value": round(r["loop_rate"], 2),
                    "finite": bool(np.all(np.isfinite(res[leg + "_state"])))}
    # the learned leg's precision: FP64 and mixed stages with the same network's slope flags, ten
    # steps on the start mesh at the CFL step
    mlp2 = res.pop("mlp_flags")
    fin = {}
    for prec in (F.PREC_FP64, F.PREC_MIXED):
        q = c2_params()
        q.limiter = 2
        q.dt = 1e-12
        d = F.M5Device(q, m)
        d.set_precision(prec)
        d.set_state(S)
        d.set_limiter_flags(mlp2)
        d.set_dt(d.cfl_dt(0.4))
        d.advance(10)
        fin[prec] = d.get_state()
        d.close()
    nm = 3
    out["mixed_vs_fp64_10steps"] = {
        "rel_l1_rho": rel_l1(fin[F.PREC_MIXED][:, :nm].sum(1), fin[F.PREC_FP64][:, :nm].sum(1)),
        "rel_l1_p": rel_l1(m5_mixture_pressure(c2_params(), fin[F.PREC_MIXED]),
                           m5_mixture_pressure(c2_params(), fin[F.PREC_FP64])),
        "case": "the start mesh, the learned slope flags, ten steps at the CFL-0.4 step"}
    del fin
    log_fn(f"[c4] mixed vs FP64 after 10 steps: rho L1 {out['mixed_vs_fp64_10steps']['rel_l1_rho']:.2e}, "
           f"p L1 {out['mixed_vs_fp64_10steps']['rel_l1_p']:.2e}")
    out["speedup_stage"] = round(res["learned"]["stage_rate"] / res["plain"]["stage_rate"], 3)
    out["speedup_loop"] = round(res["learned"]["loop_rate"] / res["plain"]["loop_rate"], 3)
    out["unit"] = "Mcell-updates/s per RK stage; loop_value with CFL, interface_location and regrids"
    return out


def c4_ab(maxlevel=7, ref_maxlevel=8, base=5, layers=32, every=5, t_over_rayleigh=1.6, log_fn=log):
    """What the learned indicator gains over the plain one (criterionsplt.c:40-140): config 2 to
    1.6 Rayleigh times on reference levels 5-7 driven by the plain indicator (minmod) and by the
    learned one (its targets with the plain refinement forced, MC where it flags smooth flow), in
    FP64, against the plain run on levels 5-8 as the reference.  The error is the mean absolute
    difference of the gas volume history (in V0) from the reference's; with the mean leaf count."""
    mlp, train = train_c4_network(log_fn)
    t_r = 0.915 * R_BUBBLE * np.sqrt(1000.0 / (C2_P[0] - C2_P[1]))
    runs = {}
    for leg, lvl, learned in (("reference", ref_maxlevel, False), ("plain", maxlevel, False),
                              ("learned", maxlevel, True)):
        p, m, S, dev, crit = build_c2(lvl, base, layers, log_fn=lambda *a: None)
        dev.close()
        q = c2_params()
        q.limiter = 2 if learned else 0
        q.dt = 1e-12
        d = F.M5Device(q, m)
        d.set_forest(m)
        d.set_state(S)
        if learned:
            d.indicator_mlp(mlp, crit)
        v0 = gas_volume(S, m.arrays()["level"])
        r = c2_loop(d, crit, lvl, 10 ** 7, every, t_end=t_over_rayleigh * t_r, probe=25, log_fn=lambda *a: None,
                    mlp=mlp if learned else None)
        d.close()
        ts = np.array([0.0] + [v[0] for v in r["vols"]]) / t_r
        vs = np.array([v0] + [v[1] for v in r["vols"]]) / v0
        runs[leg] = {"t": ts, "v": vs, "leaves_mean": float(np.mean(r["leaves"])), "steps": r["steps"],
                     "wall": r["wall_s"]}
        log_fn(f"[c4 a/b] {leg}: levels {base}-{lvl}, {r['steps']} steps, mean leaves {np.mean(r['leaves']):.0f}, "
               f"min gas volume {vs.min():.4f} at {ts[int(np.argmin(vs))]:.3f} t_R")
    mlp.close()
    ref = runs["reference"]
    grid = np.linspace(0.0, t_over_rayleigh, 161)
    vr = np.interp(grid, ref["t"], ref["v"])
    out = {"workload": f"c2 to {t_over_rayleigh} Rayleigh times, levels {base}-{maxlevel} against a "
                       f"levels {base}-{ref_maxlevel} plain reference, FP64, interface layers {layers}",
           "training": train}
    for leg in ("plain", "learned"):
        v = np.interp(grid, runs[leg]["t"], runs[leg]["v"])
        out[leg] = {"gas_volume_error_mean": float(np.mean(np.abs(v - vr))),
                    "gas_volume_min": float(runs[leg]["v"].min()),
                    "t_min_over_rayleigh": float(runs[leg]["t"][int(np.argmin(runs[leg]["v"]))]),
                    "leaves_mean": round(runs[leg]["leaves_mean"], 1), "steps": runs[leg]["steps"]}
    out["reference"] = {"gas_volume_min": float(ref["v"].min()), "t_min_over_rayleigh": float(ref["t"][int(np.argmin(ref["v"]))]),
                        "leaves_mean": round(ref["leaves_mean"], 1), "steps": ref["steps"]}
    out["error_ratio_learned_over_plain"] = round(out["learned"]["gas_volume_error_mean"] /
                                                  max(out["plain"]["gas_volume_error_mean"], 1e-300), 3)
    out["leaves_ratio_learned_over_plain"] = round(out["learned"]["leaves_mean"] / out["plain"]["leaves_mean"], 3)
    return out


def uniform_leg(level=9, warmup=3, steps=3, log_fn=log):
    """The FP64 stage on the uniform 2^(level-1) cube of the same bubble (256^3 =
    16.8 M leaves at level 9, every leaf in a 4x4x4 block): the block kernel's
    roofline at a size where its launch fills the chip many times over."""
    t = time.perf_counter()
    p, m, S = build_uniform(level)
    dev = F.Device(p, m, prec=F.PREC_FP64)
    dev.set_state(S)
    dev.time_stages(3 * warmup)
    ms_stage, ms_kernel = dev.time_stages(3 * steps)
    nbricks, nrest = dev.brick_count()
    ok = bool(np.all(np.isfinite(dev.get_state())))
    dev.close()
    if not ok:
        raise SystemExit("non-finite state after the uniform-cube stages")
    B = algorithmic_bytes(F.neq(p.nmat))
    achieved = B * 64 * nbricks / (ms_kernel * 1e-3) / 1e9
    side = 1 << (level - 1)
    wl = f"uniform_{side}^3_water_air_bubble"
    log_fn(f"uniform {side}^3: {m.n} leaves, {ms_stage:.3f} ms/stage, block kernel {ms_kernel:.3f} ms "
           f"({time.perf_counter() - t:.1f} s)")
    return {"workload": wl, "leaves": m.n, "ms_per_stage": round(ms_stage, 4),
            "value": round(m.n / (ms_stage * 1e-3) / 1e6, 2), "unit": "Mcell-updates/s per RK stage",
            "roofline": {"bound": "hbm", "kernel": "k_stage_brick", "leaves_per_launch": 64 * nbricks,
                         "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBPS, 4), "ms_kernel": round(ms_kernel, 4),
                         "traffic": traffic_bytes(wl, "fp64", "k_stage_brick"), "algorithmic_bytes_per_leaf": B}}


def c3_run(steps, warmup, layers, nsteps_cmp=300, log_fn=log):
    """Config 3 at its size: the C1 bubble pre-adapted to reference levels 5-10 (config
    levels 4-9); FP64 and mixed-precision stage rates, and the mixed state against the FP64
    one after nsteps_cmp steps on the fixed mesh (per-field relative L1 and Linf)."""
    t0 = time.perf_counter()
    p, m, S = build_c1(maxlevel=10, layers=layers, log_fn=log_fn)
    log_fn(f"setup {time.perf_counter() - t0:.1f} s, {m.n} leaves")
    nq = F.neq(p.nmat)
    B = algorithmic_bytes(nq)
    out = {"metric": "Mcell-updates/s per RK stage (config 3: FP64 and mixed precision)", "unit":
           "Mcell-updates/s per RK stage", "n_gpus": 1, "steps": steps, "warmup": warmup, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "data": "synthetic",
           "config": {"workload": "c3_rayleigh_collapse_amr_L5-10", "leaves": m.n, "levels": "5-10 (config 4-9)",
                      "interface_layers": layers, "prelax": p.prelax, "rk": "SSP-RK3"}}
    states = {}
    for prec, name in ((F.PREC_FP64, "fp64"), (F.PREC_MIXED, "mixed")):
        dev = F.Device(p, m, prec=prec)
        dev.set_state(S)
        dev.time_stages(3 * max(1, warmup))
        ms, msf = dev.time_stages(3 * steps)
        dev.set_state(S)
        t1 = time.perf_counter()
        dev.advance(nsteps_cmp)
        wall = time.perf_counter() - t1
        states[name] = dev.get_state()
        dev.close()
        out[name] = {"value": round(m.n / (ms * 1e-3) / 1e6, 2), "ms_per_stage": round(ms, 4),
                     "ms_dominant_kernel": round(msf, 4),
                     "achieved_GBps": round(B * m.n / (ms * 1e-3) / 1e9, 1),
                     "frac_of_8TBps": round(B * m.n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
                     "run_s": round(wall, 2)}
        log_fn(f"[c3] {name}: {ms:.3f} ms/stage")
    a, b = states["mixed"], states["fp64"]
    names = ["alpha_rho_water", "alpha_rho_air", "rho_u", "rho_v", "rho_w", "rho_E", "alpha_water",
             "alpha_rho_e_water", "alpha_rho_e_air"]
    err = {}
    for k, nm in enumerate(names):
        den = np.abs(b[:, k]).sum() if k not in (2, 3, 4) else np.sqrt((b[:, 2:5] ** 2).sum(axis=1)).sum()
        dmax = np.abs(b[:, k]).max() if k not in (2, 3, 4) else np.sqrt((b[:, 2:5] ** 2).sum(axis=1)).max()
        err[nm] = {"l1": float(np.abs(a[:, k] - b[:, k]).sum() / den),
                   "linf": float(np.abs(a[:, k] - b[:, k]).max() / dmax)}
    pa, pb = mixture_pressure(p, a), mixture_pressure(p, b)
    err["pressure"] = {"l1": float(np.abs(pa - pb).sum() / np.abs(pb).sum()),
                       "linf": float(np.abs(pa - pb).max() / np.abs(pb).max())}
    out["mixed_vs_fp64"] = {"steps": nsteps_cmp, "fields": err}
    out["value"] = out["fp64"]["value"]
    out["ms_per_stage"] = out["fp64"]["ms_per_stage"]
    out["ms_per_step"] = round(3 * out["fp64"]["ms_per_stage"], 4)
    out["dtype"] = "f64"
    out["speedup_mixed_vs_fp64"] = round(out["mixed"]["value"] / out["fp64"]["value"], 3)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed SSP-RK3 steps (3 stages each)")
    ap.add_argument("--warmup", type=int, default=3, help="untimed SSP-RK3 steps before the timed ones")
    ap.add_argument("--workload", default="c1",
                    choices=["c1", "u64", "u128", "u256", "u320", "c2", "c2_collapse", "c3", "c4", "c4_ab"],
                    help="c2: config 2 (model 5, the bubble near an elastoplastic slab, AMR levels 5-9); "
                         "c2_collapse: config 2 on levels 5-8 past the collapse; c3: config 3, the bubble adapted "
                         "to levels 5-10, FP64 against mixed precision")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-only", action="store_true")
    ap.add_argument("--layers", type=int, default=64)
    ap.add_argument("--prec", default="fp64", choices=["fp64", "mixed", "exact"])
    ap.add_argument("--fused", action="store_true", help="one fused stage kernel without stored face states")
    ap.add_argument("--no-mlp", action="store_true", help="skip the FP32 + learned-indicator leg")
    ap.add_argument("--no-regrid", action="store_true", help="skip the device / host regrid leg")
    ap.add_argument("--no-collapse", action="store_true", help="skip the run to the first collapse")
    ap.add_argument("--no-m5", action="store_true", help="skip the config-2 leg (model 5: solid, GFM, AMR)")
    ap.add_argument("--no-u256", action="store_true", help="skip the uniform 256^3 block-kernel roofline leg")
    ap.add_argument("--no-c4", action="store_true", help="skip the config-4 leg (learned indicator on config 2)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU halo exchange is "
              "designed but not built", file=sys.stderr)
        sys.exit(2)

    if a.workload == "c2":
        leg = c2_leg(steps=a.steps, check=not a.no_cpu_baseline)
        print(json.dumps({"metric": "Mcell-updates/s per RK stage (model 5: solid, GFM)", "value": leg["value"],
                          "unit": "Mcell-updates/s per RK stage", "n_gpus": 1, "steps": leg["steps"], "warmup": 1,
                          "ms_per_step": round(3 * leg["ms_per_stage"], 4), "ms_per_stage": leg["ms_per_stage"],
                          "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                          "config": {"workload": leg["workload"], "leaves": leg["leaves"]}, "c2": leg}),
              flush=True)
        return
    if a.workload == "c2_collapse":
        print(json.dumps(c2_collapse()), flush=True)
        return
    if a.workload == "c4":
        leg = c4_leg(steps=a.steps)
        print(json.dumps({"metric": METRIC, "value": leg["learned"]["value"], "unit": "Mcell-updates/s per RK stage",
                          "n_gpus": 1, "steps": leg["learned"]["steps"], "warmup": 1, "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "f32 faces / f64 state", "data": "synthetic",
                          "config": {"workload": leg["workload"], "leaves": leg["leaves_start"]}, "c4": leg}),
              flush=True)
        return
    if a.workload == "c4_ab":
        print(json.dumps(c4_ab()), flush=True)
        return

    if a.workload == "c3":
        print(json.dumps(c3_run(a.steps, a.warmup, a.layers)), flush=True)
        return

    t_setup = time.perf_counter()
    if a.workload == "c1":
        p, m, S = build_c1(layers=a.layers)
        cfg = {"workload": "c1_rayleigh_collapse_amr_L5-9", "leaves": m.n, "levels": "5-9 (config 4-8)",
               "domain_m": L_DOMAIN, "bubble_radius_m": R_BUBBLE, "interface_layers": a.layers}
    else:
        lvl = {"u64": 7, "u128": 8, "u256": 9, "u320": None}[a.workload]
        if lvl is None:
            p = c1_params()
            m = F.Mesh.uniform((5, 5, 5), (0, 0, 0), L_DOMAIN / 5, 7, list(p.bc))
            S = m.init_bubble(p, RHO0, P0, (0.5 * L_DOMAIN,) * 3, R_BUBBLE, 2.0, 1234, 0.01)
            p.dt = 0.4 * (L_DOMAIN / 320) / np.sqrt(6.12 * (P0[0] + 3.43e8) / 1000.0)
        else:
            p, m, S = build_uniform(lvl)
        cfg = {"workload": f"uniform_{a.workload[1:]}^3_water_air_bubble", "leaves": m.n}
    nq = F.neq(p.nmat)
    cfg.update({"prelax": p.prelax, "rk": "SSP-RK3", "limiter": "minmod", "neq": nq, "precision": a.prec})
    log(f"setup {time.perf_counter() - t_setup:.1f} s, {m.n} leaves")

    prec = {"fp64": F.PREC_FP64, "mixed": F.PREC_MIXED, "exact": F.PREC_EXACT}[a.prec]
    dev = F.Device(p, m, prec=prec)
    dev.set_fused(a.fused)
    dev.set_state(S)
    ns = 3   # SSP-RK3: a step is three stages, timed as whole steps
    dev.time_stages(ns * max(1, a.warmup))
    t0 = time.perf_counter()
    ms_stage, ms_flux = dev.time_stages(ns * a.steps)
    wall = time.perf_counter() - t0
    Sg = dev.get_state()
    nbricks, nrest = dev.brick_count() if a.prec != "exact" and not a.fused else (0, m.n)
    nplain, ngraded = dev.brick_kinds() if nbricks > 0 else (0, 0)
    dev.close()
    if not np.all(np.isfinite(Sg)):
        raise SystemExit("non-finite state after the timed stages")
    copy_gbps = F.stream_copy_gbps(1 << 30, 10 if a.profile_only else 20)
    if a.profile_only:
        return

    n = m.n
    fused = a.fused and a.prec != "exact"
    # the dominant launch is the plain blocks' (the graded blocks run in a second, smaller launch)
    kernel, units = dominant_kernel(a.prec, fused, nplain if nplain > 0 else nbricks, n)
    cfg["stage_kernels"] = ("fused" if fused else
                            f"4x4x4 blocks ({nplain} plain + {ngraded} graded blocks, {64 * nbricks} leaves) + "
                            f"recon+flux ({nrest} leaves)"
                            if nbricks > 0 else "recon+flux")
    value = n / (ms_stage * 1e-3) / 1e6
    B = algorithmic_bytes(nq)
    if kernel in ("k_stage_brick", "k_stage_fused"):
        # the kernel does a whole stage for its leaves: B per leaf over its own launch time
        achieved = B * units / (ms_flux * 1e-3) / 1e9
    else:
        # the per-leaf path spreads a stage over two kernels: B per leaf over the stage
        achieved = B * n / (ms_stage * 1e-3) / 1e9
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-updates/s per RK stage",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ns * ms_stage, 4),
        "ms_per_stage": round(ms_stage, 4), "stages_per_step": ns,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "exact": "f64", "mixed": "f32 faces / f64 state"}[a.prec],
        "data": "synthetic", "config": cfg,
        "wall_s_timed": round(wall, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic_bytes(cfg["workload"], a.prec, kernel),
                     "kernel": kernel, "leaves_per_launch": units, "algorithmic_bytes_per_leaf": B,
                     "ms_kernel": round(ms_flux, 4), "stage_achieved_GBps": round(B * n / (ms_stage * 1e-3) / 1e9, 1),
                     "stream_copy_GBps": round(copy_gbps, 1)},
    }
    if a.workload == "c1" and a.prec == "fp64" and not a.fused and not a.no_mlp:
        leg = fp32_mlp_leg(p, m, S, c1_crit(9, a.layers))
        leg["speedup_vs_plain_stage"] = round(leg["value"] / value, 3)
        out["fp32_mlp"] = leg
    if a.workload == "c1" and a.prec == "fp64" and not a.fused and not a.no_regrid:
        out["regrid"] = regrid_leg(p, m, S, c1_crit(9, a.layers))
    if a.workload == "c1" and a.prec == "fp64" and not a.fused and not a.no_collapse:
        out["collapse"] = collapse_leg(p, a.layers)
    if a.workload == "c1" and a.prec == "fp64" and not a.fused and not a.no_m5:
        out["c2"] = c2_leg()
    if a.workload == "c1" and a.prec == "fp64" and not a.fused and not a.no_c4:
        out["c4"] = c4_leg()
    if a.workload == "c1" and a.prec == "fp64" and not a.fused and not a.no_u256:
        out["u256"] = uniform_leg(9)
    if not a.no_cpu_baseline:
        cb, l1, stages = cpu_baseline(p, m, S, cfg["workload"])
        out["cpu_baseline"] = cb
        out["l1_rel_vs_cpu"] = l1
        out["l1_rel_vs_cpu_case"] = "the workload above, RK stage 1, device FP64 path against the CPU restatement"
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
