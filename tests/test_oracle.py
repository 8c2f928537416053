"""Pins of the FP64 CPU oracle against what the mathematics fixes.

The reference ships no numeric golden vectors for this path (its tests cover coupling
arithmetic only), so each oracle function is checked against closed forms, exact discrete
recurrences, conservation laws and symmetries chosen so that a dropped term, a wrong sign or index,
or a transposed operand fails at least one of them.
"""
import math

import numpy as np
import pytest

from cosmolattice_amd.clmi import ModelSpec, lphi4, lphi4U1

FLD_S, PI_S, FLD_CS, PI_CS, FLD_U1, PI_U1 = 0, 1, 2, 3, 6, 7


def free_field(m2: float) -> ModelSpec:
    return ModelSpec(name="free", n_scalars=1, terms=[(0.5 * m2, [2])], fld0=[0.0], pi0=[0.0])


def plane_wave(N, n, amp=1.0, phase=0.0):
    x = np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    return amp * np.cos(2 * np.pi * (n[0] * X + n[1] * Y + n[2] * Z) / N + phase)


def smooth_random(N, rng, kmax=3, amp=0.1):
    f = np.zeros((N, N, N))
    for _ in range(6):
        n = rng.integers(-kmax, kmax + 1, size=3)
        f += plane_wave(N, n, amp * rng.uniform(0.2, 1.0), rng.uniform(0, 2 * np.pi))
    return f


# ------------------------------------------------------------------------------------------------
# RNG
# ------------------------------------------------------------------------------------------------

def test_philox_known_answers(oracle_mod):
    """Philox4x32-10 known-answer vectors published with Random123 (kat_vectors)."""
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in kat:
        got = oracle_mod.philox4x32_10(*ctr, *key)
        assert tuple(int(g) for g in got) == want


# ------------------------------------------------------------------------------------------------
# Laplacian and potential
# ------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("N,n", [(8, (1, 2, 3)), (12, (5, 0, 1)), (10, (0, 3, 4))])
def test_laplacian_plane_wave_eigenvalue(oracle_mod, N, n):
    """L e^{ik.x} = -(4/dx^2) sum_i sin^2(k_i dx/2) e^{ik.x} for the periodic 7-point stencil."""
    dx = 0.37
    f = plane_wave(N, n, 1.3, 0.4)
    lam = -(4 / dx ** 2) * sum(math.sin(math.pi * ni / N) ** 2 for ni in n)
    out = oracle_mod.laplacian(f, dx)
    np.testing.assert_allclose(out, lam * f, rtol=0, atol=1e-12 * abs(lam))


def test_lphi4_potential_matches_model_formulas(oracle_mod):
    """dV/dphi from the polynomial equals the model header's potDeriv (models/lphi4.h:195-206) and a
    central finite difference of V (models/lphi4.h:154-177)."""
    q = 100.0
    spec = lphi4(q)
    lat = oracle_mod.OracleLattice(spec, 4, 1.0, 0.01)
    import ctypes as C

    rng = np.random.default_rng(1)
    for _ in range(20):
        phi, chi = rng.uniform(-2, 2, size=2)
        buf = np.array([phi, chi, 0.0, 0.0])
        p = buf.ctypes.data_as(C.POINTER(C.c_double))
        m = C.byref(lat.m)
        v = lat.lib.orc_scalar_potential(m, p)
        d0 = lat.lib.orc_scalar_potential_deriv(m, p, 0)
        d1 = lat.lib.orc_scalar_potential_deriv(m, p, 1)
        assert v == pytest.approx(0.25 * phi ** 4 + 0.5 * q * (phi * chi) ** 2, rel=1e-14)
        assert d0 == pytest.approx(phi ** 3 + q * phi * chi ** 2, rel=1e-13, abs=1e-13)
        assert d1 == pytest.approx(q * chi * phi ** 2, rel=1e-13, abs=1e-13)
        h = 1e-5
        for k, d in ((0, d0), (1, d1)):
            e = np.zeros(4)
            e[k] = h
            vp = lat.lib.orc_scalar_potential(m, (buf + e).ctypes.data_as(C.POINTER(C.c_double)))
            vm = lat.lib.orc_scalar_potential(m, (buf - e).ctypes.data_as(C.POINTER(C.c_double)))
            assert (vp - vm) / (2 * h) == pytest.approx(d, rel=1e-6, abs=1e-6)
    # masses from potDeriv2 at the homogeneous IC (models/lphi4.h:215-226)
    assert spec.mass2([1.0, 0.0]) == [3.0, q]


# ------------------------------------------------------------------------------------------------
# Evolvers: exact discrete solutions, reversibility, conservation
# ------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("evolver", ["LF", "VV2"])
def test_free_plane_wave_exact_recurrence(oracle_mod, evolver):
    """Free massive field, static background: a plane wave obeys the exact leapfrog recurrence
    phi_n = A cos(n W dt) with cos(W dt) = 1 - w^2 dt^2 / 2 and w^2 = m^2 + (4/dx^2) sum sin^2.
    Synced momenta are pi_n = -A sin(n W dt) sin(W dt) / dt (same for VV2)."""
    N, dx, dt, m2 = 8, 0.5, 0.05, 0.7
    n = (1, 2, 0)
    A = 0.3
    lat = oracle_mod.OracleLattice(free_field(m2), N, dx, dt, expansion=False)
    f0 = plane_wave(N, n, A)
    lat.upload(FLD_S, 0, 0, f0)
    w2 = m2 + (4 / dx ** 2) * sum(math.sin(math.pi * ni / N) ** 2 for ni in n)
    W = math.acos(1 - w2 * dt * dt / 2) / dt
    steps = 37
    if evolver == "LF":
        lat.lf_init()
        lat.lf_evolve(steps)
        lat.lf_sync()
    else:
        lat.vv_evolve([1.0], steps)
    phi = lat.download(FLD_S, 0)
    pi = lat.download(PI_S, 0)
    shape = f0 / A
    np.testing.assert_allclose(phi, A * math.cos(steps * W * dt) * shape, atol=1e-13)
    np.testing.assert_allclose(pi, -A * math.sin(steps * W * dt) * math.sin(W * dt) / dt * shape, atol=1e-12)


def test_composition_weights_satisfy_order_conditions(oracle_mod):
    """Symmetric compositions of a second-order symmetric stage: the stage weights sum to 1, and
    sum(g^3) = 0 from order 4, sum(g^5) = 0 from order 6 (Yoshida 1990).  The product's velocity- and
    position-Verlet tables equal the oracle's.

    The order-4 table keeps the reference's second weight, -1.702414403875838200264
    (positionverletparameters.h:25, velocityverletparameters.h:29), which differs from Yoshida's
    closed form -2^(1/3) / (2 - 2^(1/3)) = -1.70241438391931... by -2.0e-8: a VV4/PV4 step of the
    reference advances the fields by (1 - 2.0e-8) dt.  The drop-in reproduces that value."""
    from cosmolattice_amd.clmi import PV_WEIGHTS, VV_WEIGHTS
    order = {"PV2": 2, "PV4": 4, "PV6": 6, "PV6_2": 6, "PV8": 8, "PV10": 10}
    x1 = 1.0 / (2.0 - 2.0 ** (1.0 / 3.0))
    w4 = oracle_mod.PV_WEIGHTS["PV4"]
    assert abs(w4[0] - x1) < 1e-15
    assert abs(w4[1] - (-(2.0 ** (1.0 / 3.0)) * x1) + 1.995652e-8) < 1e-13
    for name, ws in oracle_mod.PV_WEIGHTS.items():
        n = len(ws)
        g = np.array([ws[s if s < n else 2 * n - 2 - s] for s in range(2 * n - 1)])
        tol = 2.1e-8 if name == "PV4" else 1e-14
        assert abs(g.sum() - 1.0) < tol, name
        if order[name] >= 4:
            assert abs(np.sum(g ** 3)) < (2e-7 if name == "PV4" else 1e-13), name
        if order[name] >= 6:
            assert abs(np.sum(g ** 5)) < 1e-13, name
        assert PV_WEIGHTS[name] == ws
        assert VV_WEIGHTS["VV" + name[2:]] == ws


def test_position_verlet_free_plane_wave_exact_recurrence(oracle_mod):
    """Position Verlet (drift w dt/2, kick w dt, drift w dt/2) on a free plane wave, static
    background: phi_n = A cos(n th) with cos(th) = 1 - w^2 dt^2 / 2 as for leapfrog, but the
    momenta are pi_n = -A w^2 dt sin(n th) / sin(th), which differs from the velocity-Verlet value
    by the factor 1 / (1 - w^2 dt^2 / 4): a PV step run as a VV step fails this."""
    N, dx, dt, m2 = 8, 0.5, 0.05, 0.7
    n = (2, 0, 1)
    A = 0.3
    lat = oracle_mod.OracleLattice(free_field(m2), N, dx, dt, expansion=False)
    f0 = plane_wave(N, n, A)
    lat.upload(FLD_S, 0, 0, f0)
    w2 = m2 + (4 / dx ** 2) * sum(math.sin(math.pi * ni / N) ** 2 for ni in n)
    th = math.acos(1 - w2 * dt * dt / 2)
    steps = 41
    lat.pv_evolve([1.0], steps)
    shape = f0 / A
    np.testing.assert_allclose(lat.download(FLD_S, 0), A * math.cos(steps * th) * shape, atol=1e-13)
    np.testing.assert_allclose(lat.download(PI_S, 0), -A * w2 * dt * math.sin(steps * th) / math.sin(th) * shape,
                               atol=1e-12)


@pytest.mark.parametrize("scheme", ["PV4", "PV6"])
def test_position_verlet_time_reversible(oracle_mod, scheme):
    """A symmetric composition of position-Verlet stages with dt -> -dt retraces its path."""
    ws = oracle_mod.PV_WEIGHTS[scheme]
    N, dx, dt = 8, 0.6, 0.02
    spec = lphi4(100.0)
    rng = np.random.default_rng(4)
    phi0 = 1.0 + smooth_random(N, rng)
    chi0 = smooth_random(N, rng, amp=0.05)
    pi0 = smooth_random(N, rng, amp=0.1) - 0.5
    fw = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=False)
    fw.upload(FLD_S, 0, 0, phi0)
    fw.upload(FLD_S, 1, 0, chi0)
    fw.upload(PI_S, 0, 0, pi0)
    fw.pv_evolve(ws, 20)
    bw = oracle_mod.OracleLattice(spec, N, dx, -dt, expansion=False)
    for k in [(FLD_S, 0), (FLD_S, 1), (PI_S, 0), (PI_S, 1)]:
        bw.upload(*k, 0, fw.download(*k))
    bw.pv_evolve(ws, 20)
    np.testing.assert_allclose(bw.download(FLD_S, 0), phi0, atol=1e-12)
    np.testing.assert_allclose(bw.download(FLD_S, 1), chi0, atol=1e-12)
    np.testing.assert_allclose(bw.download(PI_S, 0), pi0, atol=1e-12)
    assert np.max(np.abs(fw.download(FLD_S, 0) - phi0)) > 1e-3


def test_position_verlet_orders(oracle_mod):
    """Global error after a fixed time shrinks as dt^p: 2^2 for PV2 and 2^4 for PV4 when dt halves
    (error measured against a PV4 run at dt/8).  A wrong stage order or weight drops PV4 to second
    order."""
    N, dx, T = 8, 0.7, 0.4
    spec = lphi4(30.0)
    rng = np.random.default_rng(6)
    init = [(FLD_S, 0, 1.0 + smooth_random(N, rng)), (FLD_S, 1, smooth_random(N, rng, amp=0.1)),
            (PI_S, 0, smooth_random(N, rng, amp=0.2))]

    def run(ws, dt):
        lat = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=False)
        for k, i, v in init:
            lat.upload(k, i, 0, v)
        lat.pv_evolve(ws, int(round(T / dt)))
        return lat.download(FLD_S, 0)

    pv2, pv4 = oracle_mod.PV_WEIGHTS["PV2"], oracle_mod.PV_WEIGHTS["PV4"]
    ref = run(pv4, 0.0025)
    e2 = [np.max(np.abs(run(pv2, dt) - ref)) for dt in (0.04, 0.02)]
    e4 = [np.max(np.abs(run(pv4, dt) - ref)) for dt in (0.04, 0.02)]
    assert 3.5 < e2[0] / e2[1] < 4.5
    assert 13.0 < e4[0] / e4[1] < 19.0
    assert e4[1] < e2[1] / 20


def test_leapfrog_time_reversible(oracle_mod):
    """Synced leapfrog with dt -> -dt retraces its path (static background)."""
    N, dx, dt = 8, 0.6, 0.02
    spec = lphi4(100.0)
    rng = np.random.default_rng(3)
    phi0 = 1.0 + smooth_random(N, rng)
    chi0 = smooth_random(N, rng, amp=0.05)
    pi0 = smooth_random(N, rng, amp=0.1) - 0.5
    fw = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=False)
    fw.upload(FLD_S, 0, 0, phi0)
    fw.upload(FLD_S, 1, 0, chi0)
    fw.upload(PI_S, 0, 0, pi0)
    fw.lf_init()
    fw.lf_evolve(25)
    fw.lf_sync()
    bw = oracle_mod.OracleLattice(spec, N, dx, -dt, expansion=False)
    for k in [(FLD_S, 0), (FLD_S, 1), (PI_S, 0), (PI_S, 1)]:
        bw.upload(*k, 0, fw.download(*k))
    bw.lf_init()
    bw.lf_evolve(25)
    bw.lf_sync()
    np.testing.assert_allclose(bw.download(FLD_S, 0), phi0, atol=1e-12)
    np.testing.assert_allclose(bw.download(FLD_S, 1), chi0, atol=1e-12)
    np.testing.assert_allclose(bw.download(PI_S, 0), pi0, atol=1e-12)
    assert np.max(np.abs(fw.download(FLD_S, 0) - phi0)) > 1e-3  # it did move


def _energy_scalar(lat):
    av = lat.averages()
    return 0.5 * av["pi2S"] + 0.5 * av["grad2S"] + av["pot"]


def _max_energy_drift(oracle_mod, dt, T=1.0):
    N, dx = 8, 0.7
    spec = lphi4(30.0)
    rng = np.random.default_rng(5)
    lat = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=False)
    lat.upload(FLD_S, 0, 0, 1.0 + smooth_random(N, rng))
    lat.upload(FLD_S, 1, 0, smooth_random(N, rng, amp=0.1))
    lat.upload(PI_S, 0, 0, smooth_random(N, rng, amp=0.2))
    lat.lf_init()
    e0 = _energy_scalar(lat)
    steps_per = int(round(0.1 / dt))
    worst = 0.0
    for _ in range(int(round(T / 0.1))):
        lat.lf_evolve(steps_per)
        lat.lf_sync()
        worst = max(worst, abs(_energy_scalar(lat) / e0 - 1))
    return worst


def test_energy_conserved_second_order(oracle_mod):
    """Without expansion the discrete energy <pi^2/2 + (grad+ phi)^2/2 + V> is conserved to O(dt^2):
    halving dt shrinks the worst violation by ~4.  This ties the forward-difference gradient
    energy to the 7-point kick, and V to dV/dphi."""
    e1 = _max_energy_drift(oracle_mod, 0.02)
    e2 = _max_energy_drift(oracle_mod, 0.01)
    assert e1 < 5e-3
    assert 3.5 < e1 / e2 < 4.5


def reference_run(lat, steps, out_every, evolver="LF", ws=(1.0,)):
    """The reference's main loop (source/cosmolattice.cpp:168-201): measure (sync + Friedmann check,
    which runs setAllAverages) every out_every steps, then evolve.  Returns the Friedmann records."""
    recs = []
    if evolver == "LF":
        lat.lf_init()
    for i in range(steps):
        if i % out_every == 0:
            if evolver == "LF":
                lat.lf_sync()
            recs.append(lat.hubble_constraint())
        if evolver == "LF":
            lat.lf_evolve(1)
        elif evolver.startswith("PV"):
            lat.pv_evolve(list(ws), 1)
        else:
            lat.vv_evolve(list(ws), 1)
    if evolver == "LF":
        lat.lf_sync()
    recs.append(lat.hubble_constraint())
    return recs


def _init_expanding(oracle_mod, spec, N, dx, dt, seed=7):
    rng = np.random.default_rng(seed)
    lat = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=True)
    lat.upload(FLD_S, 0, 0, spec.fld0[0] + smooth_random(N, rng, amp=0.05))
    if spec.n_scalars > 1:
        lat.upload(FLD_S, 1, 0, smooth_random(N, rng, amp=0.05))
    lat.upload(PI_S, 0, 0, spec.pi0[0] + smooth_random(N, rng, amp=0.05))
    # ModelInitializer::initialize (initializers/modelinitializer.h:71-80): aDotI = sqrt(RHS)
    lat.set_background(aI=1.0, aSI=1.0, aIM=1.0, aDotI=1.0)
    _, _, rhs = lat.hubble_constraint()
    lat.set_background(aDotI=math.sqrt(rhs))
    return lat


@pytest.mark.parametrize("evolver", ["LF", "VV2", "PV2"])
def test_friedmann_constraint_second_order(oracle_mod, evolver):
    """With self-consistent expansion the first Friedmann equation holds to O(dt^2): the worst
    |rel_diff_friedmann| shrinks ~4x when dt halves.  A wrong sign or factor in the scale-factor
    kernel (ScaleFactorKernels::get), the energies, or the leapfrog's a / aDot staggering breaks
    this."""
    spec = lphi4(50.0)
    worst = []
    for dt in (0.02, 0.01):
        lat = _init_expanding(oracle_mod, spec, 8, 0.8, dt)
        recs = reference_run(lat, int(round(1.0 / dt)), int(round(0.1 / dt)), evolver)
        worst.append(max(abs(r[0]) for r in recs))
        assert lat.background()["aI"] > 1.2  # expansion is not inert
    assert worst[0] < 1e-3
    assert 3.0 < worst[0] / worst[1] < 5.2


def test_setallaverages_overwrites_half_step_cache(oracle_mod):
    """Averages::setAllAverages (definitions/averages.h:86-94) replaces pi2AvSI by the integer-time
    average; the next leapfrog step then averages that value into pi2AvI."""
    spec = lphi4(10.0)
    lat = _init_expanding(oracle_mod, spec, 6, 0.8, 0.02)
    lat.lf_init()
    lat.lf_evolve(3)
    lat.lf_sync()
    before = lat.background()["pi2AvSI"]
    lat.set_all_averages()
    after = lat.background()
    assert after["pi2AvSI"] == after["pi2AvI"] != before
    assert after["potAvSI"] == after["potAvI"]


# ------------------------------------------------------------------------------------------------
# U(1) sector
# ------------------------------------------------------------------------------------------------

def _u1_state(N, rng, amp=0.2):
    f = {}
    f[(FLD_CS, 0)] = 0.7 + smooth_random(N, rng, amp=amp)
    f[(FLD_CS, 1)] = 0.7 + smooth_random(N, rng, amp=amp)
    f[(PI_CS, 0)] = smooth_random(N, rng, amp=amp)
    f[(PI_CS, 1)] = smooth_random(N, rng, amp=amp)
    for d in range(3):
        f[(FLD_U1, d)] = smooth_random(N, rng, amp=amp)
        f[(PI_U1, d)] = smooth_random(N, rng, amp=amp)
    return f


def _load(lat, state):
    for (k, c), a in state.items():
        lat.upload(k, 0, c, a)


def _u1_spec(gQ=1.3):
    spec = lphi4U1()
    spec.fStar, spec.omegaStar = 1.0, 1.0
    spec.gQ = gQ
    return spec


def test_u1_gauge_invariance(oracle_mod):
    """phi -> e^{i chi} phi, A_i -> A_i + (chi(x+i) - chi(x))/(dx gQ) with links exp(-i dx gQ A):
    energies are invariant and one leapfrog step is covariant."""
    N, dx, dt = 6, 0.9, 0.03
    spec = _u1_spec()
    rng = np.random.default_rng(11)
    st = _u1_state(N, rng)
    chi = rng.uniform(-3, 3, size=(N, N, N))
    g = {}
    ph = np.exp(1j * chi)
    for (kf, kp) in ((FLD_CS, FLD_CS), (PI_CS, PI_CS)):
        z = (st[(kf, 0)] + 1j * st[(kf, 1)]) * ph
        g[(kf, 0)], g[(kf, 1)] = z.real, z.imag
    for d in range(3):
        g[(FLD_U1, d)] = st[(FLD_U1, d)] + (np.roll(chi, -1, axis=d) - chi) / (dx * spec.gQ)
        g[(PI_U1, d)] = st[(PI_U1, d)]
    a = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=False)
    b = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=False)
    _load(a, st)
    _load(b, g)
    av, bv = a.averages(), b.averages()
    for k in ("pot", "pi2CS", "grad2CS", "pi2U1", "B2U1"):
        assert bv[k] == pytest.approx(av[k], rel=1e-11)
    a.lf_init(), b.lf_init()
    a.lf_evolve(1), b.lf_evolve(1)
    a.lf_sync(), b.lf_sync()
    za = a.download(PI_CS, 0, 0) + 1j * a.download(PI_CS, 0, 1)
    zb = b.download(PI_CS, 0, 0) + 1j * b.download(PI_CS, 0, 1)
    # the step moved A, so the gauge function is unchanged only up to the drift of A, which is
    # gauge invariant (pi_A is); phi picks up the same phase
    np.testing.assert_allclose(zb, za * ph, atol=1e-11)
    for d in range(3):
        np.testing.assert_allclose(b.download(PI_U1, 0, d), a.download(PI_U1, 0, d), atol=1e-11)


def test_magnetic_field_pinned(oracle_mod):
    """The magnetic field of the measurements: |B|^2 averages to the C oracle's field-strength
    reduction B2U1, it is gauge invariant, and a plane wave A_z = sin(2 pi n x/N) has the closed-form
    B_y = -(sin(2 pi n (x+1)/N) - sin(2 pi n x/N))/dx with B_x = B_z = 0."""
    N, dx = 8, 0.7
    spec = _u1_spec()
    rng = np.random.default_rng(5)
    st = _u1_state(N, rng)
    A = [st[(FLD_U1, d)] for d in range(3)]
    B = oracle_mod.magnetic_field(A, dx)
    a = oracle_mod.OracleLattice(spec, N, dx, 0.01, expansion=False)
    _load(a, st)
    assert np.mean(B[0] ** 2 + B[1] ** 2 + B[2] ** 2) == pytest.approx(a.averages()["B2U1"], rel=1e-13)
    chi = rng.uniform(-3, 3, size=(N, N, N))
    Bg = oracle_mod.magnetic_field([A[d] + (np.roll(chi, -1, axis=d) - chi) for d in range(3)], dx)
    for d in range(3):
        np.testing.assert_allclose(Bg[d], B[d], atol=1e-12)
    n = 3
    x = np.arange(N)[:, None, None] * np.ones((1, N, N))
    Az = np.sin(2 * np.pi * n * x / N)
    Bp = oracle_mod.magnetic_field([np.zeros_like(Az), np.zeros_like(Az), Az], dx)
    np.testing.assert_allclose(Bp[1], -(np.sin(2 * np.pi * n * (x + 1) / N) - Az) / dx, atol=1e-13)
    assert np.abs(Bp[0]).max() == 0 and np.abs(Bp[2]).max() == 0


def test_u1_gauss_law_conserved(oracle_mod):
    """Initial data obeying the discrete Gauss law (pi_A = 0, charge density zero) keep obeying it
    under leapfrog evolution: the kick's current and the charge density satisfy the discrete
    continuity equation."""
    N, dx, dt = 8, 0.8, 0.02
    spec = _u1_spec(gQ=2.0)
    rng = np.random.default_rng(13)
    st = _u1_state(N, rng)
    s = smooth_random(N, rng, amp=0.3)
    st[(PI_CS, 0)] = s * st[(FLD_CS, 0)]
    st[(PI_CS, 1)] = s * st[(FLD_CS, 1)]
    for d in range(3):
        st[(PI_U1, d)] = np.zeros((N, N, N))
    lat = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=False)
    _load(lat, st)
    lat.lf_init()
    g0 = lat.gauss_u1()
    assert g0[1] < 1e-14 and g0[2] < 1e-14
    lat.lf_evolve(40)
    lat.lf_sync()
    diff, lhs, rhs = lat.gauss_u1()
    assert lhs > 1e-3  # electric field and charges did develop
    assert diff < 1e-12


def _gauss_lattice(oracle_mod, N, dx, gQ, piA, phi, pi):
    spec = _u1_spec(gQ=gQ)
    spec.fStar, spec.omegaStar = 1.7, 0.9  # qnorm = 2/dx (fStar/omegaStar)^2 must enter
    lat = oracle_mod.OracleLattice(spec, N, dx, 0.01, expansion=False)
    for d in range(3):
        lat.upload(PI_U1, 0, d, piA[d])
        lat.upload(FLD_U1, 0, d, np.zeros((N, N, N)))
    for c in range(2):
        lat.upload(FLD_CS, 0, c, phi[c])
        lat.upload(PI_CS, 0, c, pi[c])
    qnorm = 2.0 / dx * (spec.fStar / spec.omegaStar) ** 2
    return lat, qnorm


@pytest.mark.parametrize("charged", [False, True])
def test_gauss_ratio_checkerboard_closed_form(oracle_mod, charged):
    """GaussLaws::checkU1 (definitions/gausslaws.h:36-50) on checkerboard data, where every lattice
    average is exact: pi_A1 = A s, pi_A2 = B s, pi_A3 = C t with s = (-1)^(x+y), t = (-1)^z, and a
    constant phi = (v, 0) with pi_phi = (0, p s).  The backward differences give LHS_1 = 2 A s / dx,
    LHS_2 = 2 B s / dx, LHS_3 = 2 C t / dx; the charge density is rho = qnorm gQ Im[conj(pi) phi] =
    -qnorm gQ p v s, so RHS = -dx rho = R0 s with R0 = dx qnorm gQ p v.  Since s and t are
    independent fair signs, <|a s + b t|> = max(|a|, |b|), and LHS_1^2 + LHS_2^2 + LHS_3^2 + RHS^2 is
    the constant 4 (A^2 + B^2 + C^2) / dx^2 + R0^2.  A forward difference, a wrong sign of RHS, a
    dropped dx or qnorm, or any of the denominators <|LHS| + |RHS|>, <sqrt(sum)>, <LHS^2 + RHS^2>
    fails at least one of the three outputs."""
    N, dx, gQ = 8, 0.7, 1.9
    A, B, C = 0.3, 0.45, 0.05
    v, p = 1.3, (0.8 if charged else 0.0)
    x = np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    s = (-1.0) ** (X + Y)
    t = (-1.0) ** Z
    piA = [A * s, B * s, C * t]
    phi = [v * np.ones((N, N, N)), np.zeros((N, N, N))]
    pi = [np.zeros((N, N, N)), p * s]
    lat, qnorm = _gauss_lattice(oracle_mod, N, dx, gQ, piA, phi, pi)
    R0 = dx * qnorm * gQ * p * v
    a_lhs, b_lhs = 2 * (A + B) / dx, 2 * C / dx
    want_diff = max(abs(a_lhs - R0), abs(b_lhs))
    want_sum = 4 * (A * A + B * B + C * C) / dx**2 + R0 * R0
    got = lat.gauss_u1()
    assert got[0] == pytest.approx(want_diff / want_sum, rel=1e-13)
    assert got[1] == pytest.approx(max(abs(a_lhs), abs(b_lhs)), rel=1e-13)
    assert got[2] == pytest.approx(abs(R0), abs=1e-15, rel=1e-13)
    # the wrong candidates named above really differ at these numbers
    wrong = [
        want_diff / (max(abs(a_lhs), abs(b_lhs)) + abs(R0)),  # <|LHS| + |RHS|>
        want_diff / math.sqrt(want_sum),  # <sqrt(sum)>
        want_diff / (4 * ((A + B) ** 2 + C * C) / dx**2 + R0 * R0),  # (sum_i LHS_i)^2 + RHS^2
    ]
    if charged:  # with RHS = 0 these two coincide with the right answer
        wrong += [max(abs(a_lhs + R0), abs(b_lhs)) / want_sum,  # RHS sign
                  max(abs(-a_lhs - R0), abs(b_lhs)) / want_sum]  # forward differences
    for w in wrong:
        assert abs(w - got[0]) > 1e-3 * got[0]


def test_gauss_ratio_plane_wave_closed_form(oracle_mod):
    """Uncharged plane waves pi_A,i = A_i cos(theta + d_i), theta = 2 pi n.x / N, n = (1, 2, 3):
    LHS_i = (2 A_i / dx) sin(pi n_i / N) cos(theta + d_i - pi n_i / N) exactly, so
    <LHS_1^2 + LHS_2^2 + LHS_3^2> = sum_i (2 A_i sin(pi n_i / N) / dx)^2 / 2 (2n.x/N is not a multiple
    of N, so <cos^2> = 1/2 on the lattice).  LHS = |Z| cos(theta + arg Z) with
    Z = sum_i (2 A_i / dx) sin(pi n_i / N) e^{i (d_i - pi n_i / N)}, and n.x mod N takes every
    residue equally often, so <|LHS|> = |Z| (1/N) sum_m |cos(2 pi m / N + arg Z)|."""
    N, dx, gQ = 12, 0.55, 1.1
    n = np.array([1, 2, 3])
    amps = np.array([0.4, -0.25, 0.3])
    dph = np.array([0.3, 1.7, -0.9])
    x = np.arange(N)
    X, Y, Z3 = np.meshgrid(x, x, x, indexing="ij")
    theta = 2 * np.pi * (n[0] * X + n[1] * Y + n[2] * Z3) / N
    piA = [amps[i] * np.cos(theta + dph[i]) for i in range(3)]
    zero = np.zeros((N, N, N))
    lat, _ = _gauss_lattice(oracle_mod, N, dx, gQ, piA, [zero + 0.9, zero], [zero, zero])
    c = 2 * amps / dx * np.sin(np.pi * n / N)
    want_sum = 0.5 * float(np.sum(c * c))
    Zc = complex(np.sum(c * np.exp(1j * (dph - np.pi * n / N))))
    m = np.arange(N)
    mean_abs_cos = float(np.mean(np.abs(np.cos(2 * np.pi * m / N + np.angle(Zc)))))
    want_lhs = abs(Zc) * mean_abs_cos
    got = lat.gauss_u1()
    assert got[1] == pytest.approx(want_lhs, rel=1e-12)
    assert got[2] == 0.0
    assert got[0] == pytest.approx(want_lhs / want_sum, rel=1e-12)
    # (sum_i LHS_i)^2 in place of sum_i LHS_i^2 would give |Z|^2 / 2 instead
    assert abs(abs(Zc) ** 2 / 2 - want_sum) > 0.05 * want_sum


def _u1_energy(lat):
    av = lat.averages()
    w2 = (lat.spec.omegaStar / lat.spec.fStar) ** 2
    return av["pi2CS"] + av["grad2CS"] + av["pot"] + 0.5 * w2 * (av["pi2U1"] + av["B2U1"])


def _u1_energy_drift(oracle_mod, dt):
    N, dx = 6, 0.9
    spec = _u1_spec(gQ=1.1)
    spec.omegaStar = 0.8  # exercises the (omegaStar/fStar)^2 factors of the gauge energies
    rng = np.random.default_rng(17)
    lat = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=False)
    _load(lat, _u1_state(N, rng))
    lat.lf_init()
    e0 = _u1_energy(lat)
    worst = 0.0
    per = int(round(0.1 / dt))
    for _ in range(6):
        lat.lf_evolve(per)
        lat.lf_sync()
        worst = max(worst, abs(_u1_energy(lat) / e0 - 1))
    return worst


def test_u1_energy_conserved_second_order(oracle_mod):
    """Charged scalar + U(1) without expansion: the energy with the reference's normalisations
    (definitions/energies.h:47-85) is conserved to O(dt^2)."""
    e1 = _u1_energy_drift(oracle_mod, 0.02)
    e2 = _u1_energy_drift(oracle_mod, 0.01)
    assert e1 < 2e-3
    assert 3.0 < e1 / e2 < 5.0


def test_charged_scalar_without_gauge_coupling_is_two_real_scalars(oracle_mod):
    """gQ = 0 and V = |phi|^4: the complex scalar evolves exactly like two real scalars with
    V = (phi1^2 + phi2^2)^2 / 2 (kick factor aV/2 of complexscalarkernels.h:33-34)."""
    N, dx, dt = 6, 0.8, 0.02
    spec_c = _u1_spec(gQ=0.0)
    spec_s = ModelSpec(name="two", n_scalars=2, terms=[(0.5, [4, 0]), (1.0, [2, 2]), (0.5, [0, 4])])
    rng = np.random.default_rng(19)
    st = _u1_state(N, rng)
    c = oracle_mod.OracleLattice(spec_c, N, dx, dt, expansion=False)
    s = oracle_mod.OracleLattice(spec_s, N, dx, dt, expansion=False)
    _load(c, st)
    for k in range(2):
        s.upload(FLD_S, k, 0, st[(FLD_CS, k)])
        s.upload(PI_S, k, 0, st[(PI_CS, k)])
    c.lf_init(), s.lf_init()
    c.lf_evolve(15), s.lf_evolve(15)
    c.lf_sync(), s.lf_sync()
    for k in range(2):
        np.testing.assert_allclose(c.download(FLD_CS, 0, k), s.download(FLD_S, k), rtol=0, atol=1e-13)
        np.testing.assert_allclose(c.download(PI_CS, 0, k), s.download(PI_S, k), rtol=0, atol=1e-12)


# ------------------------------------------------------------------------------------------------
# Spectra and Gaussian random fields
# ------------------------------------------------------------------------------------------------

def test_spectrum_parseval_and_counts(oracle_mod):
    """Sum over bins >= 1 of sum|f_n|^2 / N^6 equals the variance of f (Parseval), and bin counts
    match a brute-force enumeration of integer mode vectors."""
    N = 12
    rng = np.random.default_rng(23)
    f = rng.normal(size=(N, N, N)) + 0.3
    s, cnt, mean = oracle_mod.spectrum_bins(f, 1.0)
    assert s[1:].sum() / N ** 6 == pytest.approx(f.var(), rel=1e-12)
    assert cnt[0] == 1 and s[0] == pytest.approx((f.sum()) ** 2, rel=1e-12)
    brute = np.zeros_like(cnt)
    rng_n = [i if i <= N // 2 else i - N for i in range(N)]
    for a in rng_n:
        for b in rng_n:
            for c in rng_n:
                r = math.sqrt(a * a + b * b + c * c)
                brute[min(int(math.floor(r + 0.5)), len(cnt) - 1)] += 1
    np.testing.assert_array_equal(cnt, brute)
    assert cnt.sum() == N ** 3


def test_spectrum_plane_wave_single_bin(oracle_mod):
    """A plane wave of amplitude A puts A^2 N^6 / 2 into the bin of |n| and nothing elsewhere."""
    N = 16
    n = (3, 4, 0)  # |n| = 5
    A = 0.7
    s, _, mean = oracle_mod.spectrum_bins(plane_wave(N, n, A, 0.3), 1.0)
    assert s[5] == pytest.approx(A * A * N ** 6 / 2, rel=1e-12)
    assert np.sum(np.delete(s, 5)) < 1e-12 * s[5]


def test_spectrum_version3_plane_wave_closed_forms(oracle_mod):
    """PS_version 3 (measurements/powerspectrum.h:111-131) bins |n| f for PS_type 1 and |n|^3 f for
    PS_type 2, with prefactors kIR dx / (2 pi N^5) and (kIR dx / N)^3 / (2 pi^2).  For a plane wave
    A cos(2 pi n0.x / N) (power A^2 N^6 / 2 at +-n0) these give |n0| A^2 / 2 and 2 pi |n0|^3 A^2 / count
    in the bin of |n0|, whatever kIR and dk = 1: a dropped |n| weight, a leftover 1/dk or bin |n|
    factor, or a wrong power of N fails.  PS_version 1 of the same wave equals PS_version 3 times
    mean|n| / |n0| (type 1), which ties the two code paths together."""
    from cosmolattice_amd.clmi import power_spectrum

    N, n0, A, kIR = 16, (3, 4, 0), 0.7, 0.37
    f = plane_wave(N, n0, A, 0.3)
    s1, c, mn = oracle_mod.spectrum_bins(f, 1.0, npow=1)
    s3, _, _ = oracle_mod.spectrum_bins(f, 1.0, npow=3)
    s0, _, _ = oracle_mod.spectrum_bins(f, 1.0)
    _, ps1 = power_spectrum(s1, c, mn, N, kIR, 1.0, ps_type=1, ps_version=3)
    _, ps2 = power_spectrum(s3, c, mn, N, kIR, 1.0, ps_type=2, ps_version=3)
    _, ps1v1 = power_spectrum(s0, c, mn, N, kIR, 1.0, ps_type=1, ps_version=1)
    assert ps1[5] == pytest.approx(5 * A * A / 2, rel=1e-12)
    assert ps2[5] == pytest.approx(2 * math.pi * 125 * A * A / c[5], rel=1e-12)
    assert np.abs(np.delete(ps1, 5)).max() < 1e-12 * ps1[5]
    assert ps1v1[5] == pytest.approx(ps1[5] * mn[5] / 5, rel=1e-12)


def test_grf_statistics_and_structure(oracle_mod):
    """Vacuum modes: E|phi_k|^2 = (w*/f*)^2 (L/dx^2)^3 / (2 w_k) below the cut, zero above and at
    k = 0; the real-space field reproduces the half-space modes (exact Hermitian completion);
    pi_k + aDot phi_k is uncorrelated with phi_k and has power w_k^2 E|phi_k|^2."""
    N, kIR, m2, aDot = 16, 0.3, 1.7, 0.4
    dx = 2 * math.pi / kIR / N
    L = N * dx
    spec = lphi4(1.0)
    spec.fStar, spec.omegaStar = 1.0, 1.0
    kcut = 6.5 * kIR
    pref = (L / dx ** 2) ** 1.5 / math.sqrt(2)
    phik, pik = oracle_mod.grf_modes(N, 42, 3, m2, aDot, kIR, kcut, pref)
    Nh = N // 2 + 1
    kx, ky, kz = np.meshgrid(np.arange(N), np.arange(N), np.arange(Nh), indexing="ij")
    nx = np.where(kx <= N // 2, kx, kx - N)
    ny = np.where(ky <= N // 2, ky, ky - N)
    n2 = nx ** 2 + ny ** 2 + kz ** 2
    k = np.sqrt(n2) * kIR
    keep = (n2 > 0) & (k <= kcut)
    assert np.all(phik[~keep] == 0) and np.all(pik[~keep] == 0)
    om = np.sqrt(k ** 2 + m2)
    expect = (L / dx ** 2) ** 3 / (2 * om)
    ratio = (np.abs(phik) ** 2 / expect)[keep & (kz > 0) & (2 * kz < N)]
    assert abs(ratio.mean() - 1) < 5 / math.sqrt(ratio.size)
    d = pik + aDot * phik
    r2 = (np.abs(d) ** 2 / (om ** 2 * expect))[keep & (kz > 0) & (2 * kz < N)]
    assert abs(r2.mean() - 1) < 5 / math.sqrt(r2.size)
    cross = (np.real(np.conj(phik) * d) / (om * expect))[keep & (kz > 0) & (2 * kz < N)]
    assert abs(cross.mean()) < 5 / math.sqrt(cross.size)
    phi, pi = oracle_mod.grf_scalar(N, dx, spec, 42, 3, m2, aDot, kIR, kcut, 0.25, -0.5)
    back = np.fft.rfftn(phi - 0.25)
    np.testing.assert_allclose(back, phik, atol=1e-9 * np.abs(phik).max())
    assert phi.mean() == pytest.approx(0.25, abs=1e-12)
    assert pi.mean() == pytest.approx(-0.5, abs=1e-12)


def test_ps_type2_of_vacuum_grf_is_the_continuum_vacuum_spectrum(oracle_mod):
    """PowerSpectrumMeasurer's PS_type 2 normalisation (measurements/powerspectrum.h:85-110), applied by
    the host's power_spectrum to the bin sums, turns the vacuum initial conditions of
    initializers/fluctuationsgenerator.h:39-71 into the textbook vacuum spectrum of a free field,
    Delta_phi(k) = (omegaStar/fStar)^2 k^3 / (4 pi^2 omega_k) in program units: with
    E|phi_k|^2 = (omegaStar/fStar)^2 (L/dx^2)^3 / (2 omega_k) the prefactor (kIR n dx / N)^3 / (2 pi^2 dk)
    reduces to k^3 / (2 pi^2) because L = N dx.  Checked per bin below the cut, averaged over 6
    realisations, against k_b^3 / (4 pi^2) times the bin's mean of 1/omega_k (PS_version 1: k_b from the
    bin's mean |n|), within 5 standard errors.  A dropped factor 2, a wrong power of N or dx, or a
    missing omegaStar/fStar fails by far more."""
    from cosmolattice_amd.clmi import power_spectrum

    N, kIR, m2 = 64, 0.1, 1.3
    dx = 2 * math.pi / kIR / N
    spec = lphi4(1.0)
    spec.fStar, spec.omegaStar = 1.0, 0.8  # (omegaStar/fStar)^2 = 0.64 must appear
    kcut = 20.5 * kIR
    nseeds = 6
    acc = None
    for seed in range(nseeds):
        phi, _ = oracle_mod.grf_scalar(N, dx, spec, 900 + seed, 0, m2, 0.0, kIR, kcut, 0.0, 0.0)
        s, c, mn = oracle_mod.spectrum_bins(phi, 1.0)
        _, ps = power_spectrum(s, c, mn, N, kIR, 1.0, ps_type=2, ps_version=1)
        acc = ps if acc is None else acc + ps
    ps = acc / nseeds
    # the bin's mean of 1/omega_k over its modes (full mode set, as the bins count them)
    nn = oracle_mod.mode_norms(N).ravel()
    b = np.minimum(np.floor(nn + 0.5).astype(np.int64), len(c) - 1)
    inv_om = np.bincount(b, weights=1.0 / np.sqrt((nn * kIR) ** 2 + m2), minlength=len(c)) / np.maximum(c, 1)
    w2 = (spec.omegaStar / spec.fStar) ** 2
    want = w2 * (mn * kIR) ** 3 / (4 * math.pi ** 2) * inv_om
    checked = 0
    for k in range(2, len(c)):
        if (k + 0.5) * kIR > kcut or c[k] < 200:  # whole bins below the cut with enough modes
            continue
        # |phi_k|^2 is exponentially distributed; the bin holds about c/2 independent complex modes
        se = 1.0 / math.sqrt(nseeds * c[k] / 2.0)
        assert abs(ps[k] / want[k] - 1.0) < 5 * se, (k, ps[k], want[k])
        checked += 1
    assert checked >= 10


# ------------------------------------------------------------------------------------------------
# Complex scalar + U(1) initial conditions
# ------------------------------------------------------------------------------------------------

def u1_spec(gQ=3.0):
    spec = lphi4U1(10.0, phi0_gev=5.6964e18, lam=1.0, pi0_gev2=0.0)
    spec.gQ = gQ
    return spec


def test_u1_ics_satisfy_gauss_law_and_carry_no_charge(oracle_mod):
    """The constrained phases make every mode's charge vanish, and pi_A from the Fourier-space
    Gauss solve satisfies the lattice Gauss law (backward differences) at every site; the law then
    stays satisfied under the leapfrog."""
    N, kIR = 16, 0.5
    dx = 2 * math.pi / kIR / N
    spec = u1_spec()
    m2 = spec.cs_mass2(spec.fld0)
    phi, pi, piA = oracle_mod.grf_u1(N, dx, spec, 11, 1, m2, 0.2, kIR, 1e9, spec.fld0, spec.pi0)
    rho = pi[0] * phi[1] - pi[1] * phi[0]
    assert abs(rho.sum()) < 1e-13 * np.abs(rho).sum()
    lat = oracle_mod.OracleLattice(spec, N, dx, 0.01, expansion=False)
    for c in range(2):
        lat.upload(FLD_CS, 0, c, phi[c])
        lat.upload(PI_CS, 0, c, pi[c])
    for d in range(3):
        lat.upload(PI_U1, 0, d, piA[d])
    g = lat.gauss_u1()
    assert g[1] > 1.0 and g[0] < 1e-14
    # a deliberately wrong solve (forward instead of backward lattice momentum) violates the law
    j0k = np.fft.rfftn(-dx * dx * 2.0 / dx * (spec.fStar / spec.omegaStar) ** 2 * spec.gQ * rho)
    wrong = [np.fft.irfftn(np.conj(p), s=(N, N, N), axes=(0, 1, 2)) for p in oracle_mod.gauss_solve_modes(np.conj(j0k), N)]
    for d in range(3):
        lat.upload(PI_U1, 0, d, wrong[d])
    assert lat.gauss_u1()[0] > 1e-3
    for d in range(3):
        lat.upload(PI_U1, 0, d, piA[d])
    lat.lf_init()
    lat.lf_evolve(20)
    lat.lf_sync()
    assert lat.gauss_u1()[0] < 1e-12


def test_u1_ics_statistics_and_structure(oracle_mod):
    """E|phi_c,k|^2 = sigma_c^2 (half a real scalar's power: phi = (phi_0 + i phi_1)/sqrt2),
    E|pi_c,k + aDot phi_c,k|^2 = omega_0^2 sigma_c^2 with component 0's frequency for both, zero
    mode and modes beyond the cut are zero, and the real fields reproduce the half-space modes."""
    N, kIR, aDot = 16, 0.3, 0.4
    dx = 2 * math.pi / kIR / N
    L = N * dx
    pref = (L / dx ** 2) ** 1.5 / math.sqrt(2)
    m2 = [1.7, 0.6]
    kcut = 6.5 * kIR
    modes = oracle_mod.grf_cs_modes(N, 5, 2, m2, aDot, kIR, kcut, pref)
    Nh = N // 2 + 1
    kx, ky, kz = np.meshgrid(np.arange(N), np.arange(N), np.arange(Nh), indexing="ij")
    nx = np.where(kx <= N // 2, kx, kx - N)
    ny = np.where(ky <= N // 2, ky, ky - N)
    n2 = nx ** 2 + ny ** 2 + kz ** 2
    k = np.sqrt(n2) * kIR
    keep = (n2 > 0) & (k <= kcut)
    inner = keep & (kz > 0) & (2 * kz < N)
    om0 = np.sqrt(k ** 2 + m2[0])
    for c in range(2):
        phik, pik = modes[c]
        assert np.all(phik[~keep] == 0) and np.all(pik[~keep] == 0)
        sig2 = pref ** 2 / (2 * np.sqrt(k ** 2 + m2[c]))
        r = (np.abs(phik) ** 2 / sig2)[inner]
        assert abs(r.mean() - 1) < 5 / math.sqrt(r.size)
        d = pik + aDot * phik
        r2 = (np.abs(d) ** 2 / (om0 ** 2 * sig2))[inner]
        assert abs(r2.mean() - 1) < 5 / math.sqrt(r2.size)
        back = np.fft.rfftn(np.fft.irfftn(phik, s=(N, N, N), axes=(0, 1, 2)), axes=(0, 1, 2))
        np.testing.assert_allclose(back, phik, atol=1e-12 * np.abs(phik).max())


# ------------------------------------------------------------------------------------------------
# Reverse-mode leapfrog
# ------------------------------------------------------------------------------------------------

def _fd_loss(oracle_mod, spec, N, dx, dt, phi0, pi0, nsteps, expansion, weights):
    """A smooth scalar of the final state: a fixed linear functional of phi_T, pi_T plus a(T)."""
    o = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=expansion)
    for f in range(spec.n_scalars):
        o.upload(FLD_S, f, 0, phi0[f])
        o.upload(PI_S, f, 0, pi0[f])
    if expansion:
        _, _, rhs = o.hubble_constraint()
        o.set_background(aDotI=math.sqrt(rhs))
    else:
        o.set_all_averages()
    o.lf_init()
    o.lf_evolve(nsteps)
    o.lf_sync()
    wf, wp, wa = weights
    bg = o.background()
    return (sum(np.sum(wf[f] * o.download(FLD_S, f) ** 2) + np.sum(wp[f] * o.download(PI_S, f))
                for f in range(spec.n_scalars)) + wa[0] * bg["aI"] + wa[1] * bg["aDotI"])


@pytest.mark.parametrize("expansion", [True, False])
def test_adjoint_restatement_matches_central_differences(oracle_mod, expansion):
    """The FP64 reverse-mode leapfrog (including the scale-factor chain and the initial
    background computed from the fields) against central differences of the forward run, for a
    loss quadratic in phi_T and linear in pi_T, a(T), aDot(T); directional derivatives along random
    directions and every potential coefficient."""
    import dataclasses

    spec = lphi4(100.0, n_fields=2, phi0_gev=5.6964e18, lam=1.0, pi0_program=-0.5)
    N, dt, kIR, nsteps = 8, 0.005, 0.8, 12
    dx = 2 * math.pi / kIR / N
    rng = np.random.default_rng(21)
    phi0 = np.stack([spec.fld0[f] + 0.05 * rng.normal(size=(N,) * 3) for f in range(2)])
    pi0 = np.stack([spec.pi0[f] + 0.05 * rng.normal(size=(N,) * 3) for f in range(2)])
    wf = [rng.normal(size=(N,) * 3) / N ** 3 for _ in range(2)]
    wp = [rng.normal(size=(N,) * 3) / N ** 3 for _ in range(2)]
    wa = (0.7, -1.3)
    # seeds of the loss at the final state
    o = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=expansion)
    for f in range(2):
        o.upload(FLD_S, f, 0, phi0[f])
        o.upload(PI_S, f, 0, pi0[f])
    if expansion:
        _, _, rhs = o.hubble_constraint()
        o.set_background(aDotI=math.sqrt(rhs))
    else:
        o.set_all_averages()
    o.lf_init()
    o.lf_evolve(nsteps)
    o.lf_sync()
    seed_f = [2 * wf[f] * o.download(FLD_S, f) for f in range(2)]
    gphi, gpi, gc = oracle_mod.adjoint_lf(spec, N, dx, dt, phi0, pi0, nsteps, seed_f, wp,
                                          wa if expansion else (0.0, 0.0), expansion)
    weights = (wf, wp, wa if expansion else (0.0, 0.0))
    for which in ("phi", "pi"):
        d = rng.normal(size=phi0.shape)
        eps = 1e-5
        if which == "phi":
            lp = _fd_loss(oracle_mod, spec, N, dx, dt, phi0 + eps * d, pi0, nsteps, expansion, weights)
            lm = _fd_loss(oracle_mod, spec, N, dx, dt, phi0 - eps * d, pi0, nsteps, expansion, weights)
            ad = sum(np.sum(gphi[f] * d[f]) for f in range(2))
        else:
            lp = _fd_loss(oracle_mod, spec, N, dx, dt, phi0, pi0 + eps * d, nsteps, expansion, weights)
            lm = _fd_loss(oracle_mod, spec, N, dx, dt, phi0, pi0 - eps * d, nsteps, expansion, weights)
            ad = sum(np.sum(gpi[f] * d[f]) for f in range(2))
        assert ad == pytest.approx((lp - lm) / (2 * eps), rel=1e-6)
    for t in range(len(spec.terms)):
        c0 = spec.terms[t][0]
        h = 1e-5 * abs(c0)
        vals = []
        for sgn in (1, -1):
            terms = [(c + (sgn * h if i == t else 0.0), e) for i, (c, e) in enumerate(spec.terms)]
            vals.append(_fd_loss(oracle_mod, dataclasses.replace(spec, terms=terms), N, dx, dt, phi0, pi0, nsteps,
                                 expansion, weights))
        assert gc[t] == pytest.approx((vals[0] - vals[1]) / (2 * h), rel=1e-6)


def test_adjoint_restatement_long_run_needs_small_difference_steps(oracle_mod):
    """Over 2000 leapfrog steps of the q = 100 model the final state depends on the coupling
    nonlinearly enough that a central difference with relative step 1e-4 misses the derivative by
    a large factor, while steps of 1e-7 and 1e-8 converge to the reverse-mode value (to 1e-4, the
    bar of the sensitivity runs).
    The adjoint is the exact derivative of the discrete map; long-run gradient checks must use
    small steps."""
    import dataclasses

    spec = lphi4(100.0, n_fields=2, phi0_gev=5.6964e18, lam=1.0, pi0_program=-0.5)
    N, dt, kIR, nsteps = 8, 0.005, 0.4, 2000
    dx = 2 * math.pi / kIR / N
    rng = np.random.default_rng(3)
    phi0 = np.stack([spec.fld0[f] + smooth_random(N, rng, amp=0.05) + 0.01 * rng.normal(size=(N,) * 3)
                     for f in range(2)])
    pi0 = np.stack([spec.pi0[f] + smooth_random(N, rng, amp=0.05) + 0.01 * rng.normal(size=(N,) * 3)
                    for f in range(2)])
    wf = [rng.normal(size=(N,) * 3) / N ** 3 for _ in range(2)]
    wp = [rng.normal(size=(N,) * 3) / N ** 3 for _ in range(2)]
    wa = (0.7, -1.3)
    o = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=True)
    for f in range(2):
        o.upload(FLD_S, f, 0, phi0[f])
        o.upload(PI_S, f, 0, pi0[f])
    _, _, rhs = o.hubble_constraint()
    o.set_background(aDotI=math.sqrt(rhs))
    o.lf_init()
    o.lf_evolve(nsteps)
    o.lf_sync()
    seed_f = [2 * wf[f] * o.download(FLD_S, f) for f in range(2)]
    _, _, gc = oracle_mod.adjoint_lf(spec, N, dx, dt, phi0, pi0, nsteps, seed_f, wp, wa, True)

    def fd(rel):
        h = rel * abs(spec.terms[1][0])
        vals = []
        for sgn in (1, -1):
            terms = [(c + (sgn * h if i == 1 else 0.0), e) for i, (c, e) in enumerate(spec.terms)]
            vals.append(_fd_loss(oracle_mod, dataclasses.replace(spec, terms=terms), N, dx, dt, phi0, pi0, nsteps,
                                 True, (wf, wp, wa)))
        return (vals[0] - vals[1]) / (2 * h)

    assert gc[1] == pytest.approx(fd(1e-7), rel=1e-4)
    assert gc[1] == pytest.approx(fd(1e-8), rel=1e-4)
    assert abs(fd(1e-4) - gc[1]) > 1e-2 * abs(gc[1])
