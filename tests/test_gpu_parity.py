"""GPU parity: the HIP path (through the C ABI) against the FP64 CPU oracle on identical inputs.

Tolerances: FP64 mode must match the oracle within 1e-10 relative on fields and volume averages
(the bar stated for the FP64 drop-in); spectra within 1e-8.  FP32-bulk mode is checked against the
same FP64 oracle with the looser bounds written in each test.
"""
import math

import numpy as np
import pytest

from cosmolattice_amd.clmi import (CLMI_F32_BULK, CLMI_F64, FLD_CS, FLD_S, FLD_U1, MAG_U1, PI_CS, PI_S, PI_U1,
                                   Lattice, ModelSpec, lphi4, lphi4U1)

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def say(*a):
    print(*a, flush=True)


def smooth_random(N, rng, kmax=3, amp=0.1, waves=6):
    x = np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    f = np.zeros((N, N, N))
    for _ in range(waves):
        n = rng.integers(-kmax, kmax + 1, size=3)
        f += amp * rng.uniform(0.2, 1.0) * np.cos(2 * np.pi * (n[0] * X + n[1] * Y + n[2] * Z) / N
                                                  + rng.uniform(0, 2 * np.pi))
    return f


def bench_spec(q=100.0, n_fields=2):
    # lambda = 1, phi0 = 1 in program units; fStar / MPl = 2.34 so the expansion is live
    return lphi4(q, n_fields=n_fields, phi0_gev=5.6964e18, lam=1.0, pi0_program=-0.5)


def scalar_ics(spec, N, rng, noise=0.02):
    out = {}
    for f in range(spec.n_scalars):
        out[(FLD_S, f)] = spec.fld0[f] + smooth_random(N, rng, amp=0.08) + noise * rng.normal(size=(N, N, N))
        out[(PI_S, f)] = spec.pi0[f] + smooth_random(N, rng, amp=0.08) + noise * rng.normal(size=(N, N, N))
    return out


def load_both(lat, orc, ics):
    for (k, i), a in ics.items():
        lat.upload(k, i, 0, a)
        orc.upload(k, i, 0, a)


def init_background(lat, orc):
    """aDotI from the Friedmann constraint on the loaded fields (modelinitializer.h:71-80)."""
    orc.set_background(aI=1.0, aSI=1.0, aIM=1.0, aDotI=1.0)
    _, _, rhs = orc.hubble_constraint()
    orc.set_background(aDotI=math.sqrt(rhs))
    lat.set_background(**orc.background())


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def expected_scalar_pass(N, prec, flags=0):
    """The scalar lattice pass a context of this size, precision and flags runs: the full-row pass for
    64, 128, 192, 256 and 512 (FP32-bulk) except FP64 at 256^3 without a memory-mode or shape flag, the
    tiled pass for the other multiples of 64, and the generic pass with CLMI_FLAG_NO_TILED."""
    from cosmolattice_amd.clmi import CLMI_FLAG_NO_TILED

    if flags & CLMI_FLAG_NO_TILED:
        return "k_scalar_pass"
    asked = (flags >> 8) & 15 or flags & 6
    if N in (64, 128, 192) or (N == 256 and (prec == CLMI_F32_BULK or asked)) or (N == 512 and prec == CLMI_F32_BULK):
        return "k_scalar_rows"
    return "k_scalar_tiled" if N % 64 == 0 else "k_scalar_pass"


def assert_plan(lat, kernel):
    """The context runs `kernel` (first word of a line of clmi_plan_report), so a comparison with the
    generic pass is not the generic pass against itself."""
    plan = lat.plan()
    say("plan", plan.replace("\n", "; "))
    assert kernel in [ln.split()[0] for ln in plan.splitlines()], plan
    return plan


def run_reference_loop(obj, steps, out_every, evolver="LF", ws=(1.0,)):
    recs = []
    if evolver == "LF":
        obj.lf_init()
    for i in range(steps):
        if i % out_every == 0:
            if evolver == "LF":
                obj.lf_sync()
            recs.append(obj.hubble_constraint())
        if evolver == "LF":
            obj.lf_evolve(1)
        elif evolver == "PV":
            obj.pv_evolve(list(ws), 1)
        else:
            obj.vv_evolve(list(ws), 1)
    if evolver == "LF":
        obj.lf_sync()
    recs.append(obj.hubble_constraint())
    return np.array(recs)


@pytest.mark.parametrize("N", [16, 32, 20])
def test_lf_lphi4_fp64_parity_100_steps(gpu, oracle_mod, N):
    """2-field lphi4, self-consistent expansion, leapfrog, 100 steps with measurements (sync +
    Friedmann check) every 10 steps: fields, momenta, background and Friedmann records match the
    oracle within 1e-10 relative."""
    say(f"case lf_lphi4_fp64 N={N}")
    spec = bench_spec()
    dx, dt = 0.4909, 0.005
    rng = np.random.default_rng(100 + N)
    ics = scalar_ics(spec, N, rng)
    lat = Lattice(spec, N, dx, dt, CLMI_F64, expansion=True)
    orc = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=True)
    load_both(lat, orc, ics)
    init_background(lat, orc)
    rg = run_reference_loop(lat, 100, 10)
    ro = run_reference_loop(orc, 100, 10)
    for k in [(FLD_S, 0), (FLD_S, 1), (PI_S, 0), (PI_S, 1)]:
        e = relerr(lat.download(*k, 0), orc.download(*k))
        say(k, e)
        assert e < FP64_TOL, k
    bg, bo = lat.background(), orc.background()
    for key in bo:
        if bo[key] != 0:
            assert abs(bg[key] - bo[key]) <= FP64_TOL * abs(bo[key]), key
    np.testing.assert_allclose(rg[:, 1:], ro[:, 1:], rtol=FP64_TOL)
    np.testing.assert_allclose(rg[:, 0], ro[:, 0], rtol=1e-6, atol=1e-13)
    assert bg["aI"] > 1.0


@pytest.mark.parametrize("N,prec,flags", [
    (64, CLMI_F64, 0), (64, CLMI_F32_BULK, 0), (128, CLMI_F32_BULK, 0), (128, CLMI_F64, 0),
    (192, CLMI_F32_BULK, 0), (192, CLMI_F64, 0), (256, CLMI_F32_BULK, 0), (256, CLMI_F64, 0),
    (256, CLMI_F32_BULK, 4), (256, CLMI_F64, 6),
    (256, CLMI_F32_BULK, 2), (256, CLMI_F64, 2), (320, CLMI_F32_BULK, 0), (320, CLMI_F64, 0),
    (256, CLMI_F32_BULK, 1 << 8), (256, CLMI_F64, 1 << 8), (256, CLMI_F32_BULK, 3 << 8),
    (256, CLMI_F64, 3 << 8), (128, CLMI_F32_BULK, 1 << 8), (192, CLMI_F64, 1 << 8),
    (64, CLMI_F32_BULK, 3 << 8), (256, CLMI_F32_BULK, (1 << 8) | 4), (256, CLMI_F64, (1 << 8) | 2),
    (256, CLMI_F32_BULK, 7 << 8), (256, CLMI_F64, 7 << 8), (192, CLMI_F32_BULK, 7 << 8),
    (256, CLMI_F32_BULK, 15 << 8), (256, CLMI_F64, 15 << 8), (128, CLMI_F64, 15 << 8),
    (512, CLMI_F32_BULK, 0)])
def test_fast_pass_bit_identical_to_generic(gpu, N, prec, flags):
    """The full-row pass (N = 64, 128, 192, 256, 512, all workgroup shapes and memory modes) and the
    tiled x-marching pass (other multiples of 64) give fields and momenta bit-identical to the generic
    one-site-per-thread pass after 40 leapfrog steps on a static background (same per-site
    arithmetic); their volume sums differ only in summation order."""
    from cosmolattice_amd.clmi import CLMI_FLAG_NO_TILED

    say(f"case fast_vs_generic N={N} prec={prec} flags={flags}")
    spec = bench_spec()
    rng = np.random.default_rng(N + prec)
    ics = scalar_ics(spec, N, rng)
    res = []
    for flags in (flags, CLMI_FLAG_NO_TILED):
        lat = Lattice(spec, N, 0.4909, 0.005, prec, expansion=False, flags=flags)
        assert_plan(lat, expected_scalar_pass(N, prec, flags))
        for (k, i), a in ics.items():
            lat.upload(k, i, 0, a)
        lat.set_background(aI=1.1, aSI=1.05)
        lat.lf_init()
        lat.lf_evolve(25)
        lat.lf_sync()
        lat.lf_evolve(15)
        out = [lat.download(k, i, 0) for (k, i) in sorted(ics)]
        res.append((out, lat.averages()))
        lat.close()
    for a, b in zip(res[0][0], res[1][0]):
        assert np.array_equal(a, b)
    for k in ("pi2S", "grad2S", "pot"):
        assert res[0][1][k] == pytest.approx(res[1][1][k], rel=1e-13)


def runtime_poly_spec():
    """Two fields with a potential that is not lphi4's (a chi^4 and a phi^2 term added), so the passes
    evaluate it from the runtime exponent table rather than the compile-time lphi4 form."""
    base = bench_spec()
    return ModelSpec(name="lphi4+", n_scalars=2, alpha=base.alpha, fStar=base.fStar, omegaStar=base.omegaStar,
                     terms=[(0.25, [4, 0]), (50.0, [2, 2]), (0.1, [0, 4]), (0.3, [2, 0])],
                     fld0=list(base.fld0), pi0=list(base.pi0))


@pytest.mark.parametrize("N,prec", [(64, CLMI_F64), (128, CLMI_F32_BULK), (192, CLMI_F64), (256, CLMI_F32_BULK),
                                    (256, CLMI_F64), (320, CLMI_F32_BULK), (512, CLMI_F32_BULK)])
def test_fast_pass_runtime_potential_bit_identical_to_generic(gpu, N, prec):
    """With a potential outside the compile-time lphi4 form, the full-row pass (64, 128, 192, 256 FP32-bulk,
    512) and the tiled pass (256 FP64, 320) evaluate the runtime exponent table and still give fields and
    momenta bit-identical to the generic pass after 40 leapfrog steps on a static background."""
    from cosmolattice_amd.clmi import CLMI_FLAG_NO_TILED

    say(f"case fast_vs_generic runtime potential N={N} prec={prec}")
    spec = runtime_poly_spec()
    ics = scalar_ics(spec, N, np.random.default_rng(N + 7 * prec))
    res = []
    for flags in (0, CLMI_FLAG_NO_TILED):
        lat = Lattice(spec, N, 0.4909, 0.005, prec, expansion=False, flags=flags)
        plan = assert_plan(lat, expected_scalar_pass(N, prec, flags))
        assert flags or "PK=0" in plan
        for (k, i), a in ics.items():
            lat.upload(k, i, 0, a)
        lat.set_background(aI=1.1, aSI=1.05)
        lat.lf_init()
        lat.lf_evolve(25)
        lat.lf_sync()
        lat.lf_evolve(15)
        res.append([lat.download(k, i, 0) for (k, i) in sorted(ics)])
        lat.close()
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)


def test_lf_runtime_potential_fp64_parity_64(gpu, oracle_mod):
    """The same non-lphi4 potential at 64^3 (full-row pass, runtime exponent table) against the oracle:
    100 leapfrog steps with expansion and measurements every 25, within 1e-10."""
    say("case lf_runtime_potential_fp64_64")
    spec = runtime_poly_spec()
    N, dx, dt = 64, 0.4909, 0.005
    ics = scalar_ics(spec, N, np.random.default_rng(640))
    lat = Lattice(spec, N, dx, dt, CLMI_F64)
    orc = oracle_mod.OracleLattice(spec, N, dx, dt)
    load_both(lat, orc, ics)
    init_background(lat, orc)
    rg = run_reference_loop(lat, 100, 25)
    ro = run_reference_loop(orc, 100, 25)
    for k in [(FLD_S, 0), (FLD_S, 1), (PI_S, 0), (PI_S, 1)]:
        e = relerr(lat.download(*k, 0), orc.download(*k))
        say(k, e)
        assert e < FP64_TOL, k
    np.testing.assert_allclose(rg[:, 1:], ro[:, 1:], rtol=FP64_TOL)
    lat.close()


@pytest.mark.parametrize("N,prec,mm", [(64, CLMI_F64, 15), (64, CLMI_F32_BULK, 15), (128, CLMI_F32_BULK, 15),
                                       (256, CLMI_F64, 15), (256, CLMI_F64, 0), (320, CLMI_F32_BULK, 0)])
def test_verlet_passes_bit_identical_to_generic(gpu, N, prec, mm):
    """Velocity Verlet (kick-only and drift+kick passes) and position Verlet (drift-only and drift+kick
    passes) give fields and momenta bit-identical to the generic one-site-per-thread pass on a static
    background (fourth-order compositions, 6 steps): through the full-row pass (memory mode 15 asks for
    it) and, without that flag, through the default pass, which is the tiled pass at 256^3 FP64 and
    320^3."""
    from cosmolattice_amd.clmi import CLMI_FLAG_MEMMODE, CLMI_FLAG_NO_TILED, PV_WEIGHTS, VV_WEIGHTS

    say(f"case verlet_vs_generic N={N} prec={prec} mm={mm}")
    spec = bench_spec()
    rng = np.random.default_rng(3 * N + prec)
    ics = scalar_ics(spec, N, rng)
    for scheme in ("vv", "pv"):
        res = []
        for flags in (CLMI_FLAG_MEMMODE(mm) if mm else 0, CLMI_FLAG_NO_TILED):
            lat = Lattice(spec, N, 0.4909, 0.005, prec, expansion=False, flags=flags)
            assert_plan(lat, expected_scalar_pass(N, prec, flags))
            for (k, i), a in ics.items():
                lat.upload(k, i, 0, a)
            lat.set_background(aI=1.1, aSI=1.05, aIM=1.0)
            if scheme == "vv":
                lat.vv_evolve(VV_WEIGHTS["VV4"], 6)
            else:
                lat.pv_evolve(PV_WEIGHTS["PV4"], 6)
            res.append([lat.download(k, i, 0) for (k, i) in sorted(ics)])
            lat.close()
        for a, b in zip(res[0], res[1]):
            assert np.array_equal(a, b), scheme


@pytest.mark.parametrize("N,prec,n_fields,scheme", [(64, CLMI_F64, 1, "VV2"), (64, CLMI_F32_BULK, 2, "VV4"),
                                                    (128, CLMI_F32_BULK, 2, "VV2"), (128, CLMI_F64, 2, "VV6"),
                                                    (256, CLMI_F32_BULK, 2, "VV2"), (256, CLMI_F64, 2, "VV4"),
                                                    (320, CLMI_F32_BULK, 2, "VV2"), (320, CLMI_F64, 1, "VV2")])
def test_vv_double_kick_bit_identical_to_two_passes(gpu, N, prec, n_fields, scheme):
    """Velocity Verlet with the second half-kick of each stage and the first half-kick of the next in one
    double-kick pass (the default on the full-row and the tiled pass) gives the same bits as the two separate passes
    (CLMI_FLAG_NO_DOUBLE_KICK): fields, momenta and the whole background, with expansion, after three
    calls of uneven length (the first and last stage of a call stay single kicks) and a download
    between them."""
    from cosmolattice_amd.clmi import CLMI_FLAG_NO_DOUBLE_KICK, VV_WEIGHTS

    say(f"case vv double kick N={N} prec={prec} nf={n_fields} {scheme}")
    spec = bench_spec(n_fields=n_fields)
    rng = np.random.default_rng(5 * N + prec + n_fields)
    ics = scalar_ics(spec, N, rng)
    res = []
    for flags in (0, CLMI_FLAG_NO_DOUBLE_KICK):
        lat = Lattice(spec, N, 0.4909, 0.005, prec, expansion=True, flags=flags)
        assert_plan(lat, expected_scalar_pass(N, prec))
        for (k, i), a in ics.items():
            lat.upload(k, i, 0, a)
        lat.set_background(aDotI=0.6)
        snaps = []
        for n in (5, 1, 4):
            lat.vv_evolve(VV_WEIGHTS[scheme], n)
            snaps.append(([lat.download(k, i, 0) for (k, i) in sorted(ics)], lat.background()))
        res.append(snaps)
        lat.close()
    for j, ((fa, ba), (fb, bb)) in enumerate(zip(res[0], res[1])):
        for a, b in zip(fa, fb):
            assert np.array_equal(a, b), j
        assert ba == bb, j
    assert res[0][-1][1]["aI"] > 1.0


@pytest.mark.parametrize("prec", [CLMI_F64, CLMI_F32_BULK])
def test_autotune_keeps_state_and_results(gpu, prec):
    """Autotuning runs passes on the live data but leaves fields and background untouched, and the
    plan it picks gives bit-identical results (volume sums included) to the default plan."""
    say(f"case autotune prec={prec}")
    spec = bench_spec()
    N = 128
    rng = np.random.default_rng(77)
    ics = scalar_ics(spec, N, rng)
    outs = []
    for tune in (False, True):
        lat = Lattice(spec, N, 0.4909, 0.005, prec)
        for (k, i), a in ics.items():
            lat.upload(k, i, 0, a)
        lat.set_background(aDotI=0.6)
        lat.lf_init()
        lat.lf_evolve(7)
        if tune:
            before = [lat.download(k, i, 0) for (k, i) in sorted(ics)]
            bg_before = lat.background()
            rep = lat.autotune(2)
            say(rep)
            assert "XC=" in rep
            assert all(np.array_equal(x, lat.download(k, i, 0)) for x, (k, i) in zip(before, sorted(ics)))
            assert lat.background() == bg_before
        lat.lf_evolve(13)
        outs.append(([lat.download(k, i, 0) for (k, i) in sorted(ics)], lat.background()))
        lat.close()
    for a, b in zip(outs[0][0], outs[1][0]):
        assert np.array_equal(a, b)
    assert outs[0][1] == outs[1][1]


def test_lf_lphi4_fp64_parity_64_fast_pass(gpu, oracle_mod):
    """64^3 (full-row pass, then the separate background-update kernel) against the oracle: 100 leapfrog
    steps, measurements every 25."""
    say("case lf_lphi4_fp64_64")
    spec = bench_spec()
    N, dx, dt = 64, 0.4909, 0.005
    rng = np.random.default_rng(64)
    ics = scalar_ics(spec, N, rng)
    lat = Lattice(spec, N, dx, dt, CLMI_F64)
    orc = oracle_mod.OracleLattice(spec, N, dx, dt)
    load_both(lat, orc, ics)
    init_background(lat, orc)
    rg = run_reference_loop(lat, 100, 25)
    ro = run_reference_loop(orc, 100, 25)
    for k in [(FLD_S, 0), (FLD_S, 1), (PI_S, 0), (PI_S, 1)]:
        e = relerr(lat.download(*k, 0), orc.download(*k))
        say(k, e)
        assert e < FP64_TOL, k
    np.testing.assert_allclose(rg[:, 1:], ro[:, 1:], rtol=FP64_TOL)


def test_vv2_single_field_fp64_parity(gpu, oracle_mod):
    """Single-inflaton lphi4 with velocity Verlet (VV2), expansion on: 60 steps."""
    say("case vv2_1field_fp64")
    spec = bench_spec(n_fields=1)
    N, dx, dt = 24, 0.98, 0.01
    rng = np.random.default_rng(7)
    ics = scalar_ics(spec, N, rng)
    lat = Lattice(spec, N, dx, dt, CLMI_F64)
    orc = oracle_mod.OracleLattice(spec, N, dx, dt)
    load_both(lat, orc, ics)
    init_background(lat, orc)
    rg = run_reference_loop(lat, 60, 20, evolver="VV")
    ro = run_reference_loop(orc, 60, 20, evolver="VV")
    for k in [(FLD_S, 0), (PI_S, 0)]:
        assert relerr(lat.download(*k, 0), orc.download(*k)) < FP64_TOL
    np.testing.assert_allclose(rg[:, 1:], ro[:, 1:], rtol=FP64_TOL)


@pytest.mark.parametrize("scheme,N", [("PV2", 16), ("PV4", 20)])
def test_pv_lphi4_fp64_parity(gpu, oracle_mod, scheme, N):
    """2-field lphi4 with position Verlet, expansion on, 40 steps with Friedmann checks every 10:
    fields, momenta, background and Friedmann records match the oracle within 1e-10."""
    say(f"case pv_lphi4_fp64 {scheme} N={N}")
    ws = oracle_mod.PV_WEIGHTS[scheme]
    spec = bench_spec()
    dx, dt = 0.4909, 0.005
    rng = np.random.default_rng(200 + N)
    ics = scalar_ics(spec, N, rng)
    lat = Lattice(spec, N, dx, dt, CLMI_F64, expansion=True)
    orc = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=True)
    load_both(lat, orc, ics)
    init_background(lat, orc)
    rg = run_reference_loop(lat, 40, 10, evolver="PV", ws=ws)
    ro = run_reference_loop(orc, 40, 10, evolver="PV", ws=ws)
    for k in [(FLD_S, 0), (FLD_S, 1), (PI_S, 0), (PI_S, 1)]:
        e = relerr(lat.download(*k, 0), orc.download(*k))
        say(k, e)
        assert e < FP64_TOL, k
    bg, bo = lat.background(), orc.background()
    for key in bo:
        if bo[key] != 0:
            assert abs(bg[key] - bo[key]) <= FP64_TOL * abs(bo[key]), key
    np.testing.assert_allclose(rg[:, 1:], ro[:, 1:], rtol=FP64_TOL)
    assert bg["aI"] > 1.0


def test_pv_static_background_parity(gpu, oracle_mod):
    """PV2 without expansion (drift coefficients from aI, no background updates), FP64 and FP32-bulk."""
    say("case pv_static")
    spec = bench_spec()
    N, dx, dt = 16, 0.4909, 0.01
    ics = scalar_ics(spec, N, np.random.default_rng(5))
    orc = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=False)
    for (k, i), a in ics.items():
        orc.upload(k, i, 0, a)
    orc.pv_evolve([1.0], 30)
    for prec, tol in ((CLMI_F64, FP64_TOL), (CLMI_F32_BULK, 1e-5)):
        lat = Lattice(spec, N, dx, dt, prec, expansion=False)
        for (k, i), a in ics.items():
            lat.upload(k, i, 0, a)
        lat.pv_evolve([1.0], 30)
        for k in [(FLD_S, 0), (FLD_S, 1), (PI_S, 0), (PI_S, 1)]:
            e = relerr(lat.download(*k, 0), orc.download(*k))
            say(prec, k, e)
            assert e < tol, (prec, k)
        lat.close()


def test_primitives_fp64_parity(gpu, oracle_mod):
    """kick, drift and every volume average as single passes (no fusion involved)."""
    say("case primitives_fp64")
    spec = bench_spec()
    N, dx = 16, 0.7
    rng = np.random.default_rng(9)
    ics = scalar_ics(spec, N, rng)
    lat = Lattice(spec, N, dx, 0.01, CLMI_F64)
    orc = oracle_mod.OracleLattice(spec, N, dx, 0.01)
    load_both(lat, orc, ics)
    lat.set_background(aI=1.3, aSI=1.2)
    orc.set_background(aI=1.3, aSI=1.2)
    ag, ao = lat.averages(), orc.averages()
    for k in ("pi2S", "grad2S", "pot"):
        assert ag[k] == pytest.approx(ao[k], rel=1e-13)
    lat.kick(0.013)
    orc.kick(0.013)
    lat.drift(0.021)
    orc.drift(0.021)
    for k in [(FLD_S, 0), (FLD_S, 1), (PI_S, 0), (PI_S, 1)]:
        assert relerr(lat.download(*k, 0), orc.download(*k)) < 1e-14
    assert lat.scale_factor_kernel() == pytest.approx(orc.scale_factor_kernel(), rel=1e-13)


def test_f32_bulk_tracks_fp64_at_256(gpu):
    """The headline workload at its full size: 2-field lphi4 at 256^3 (q = 100, kIR = 0.05, dt = 0.005,
    self-consistent expansion) from the same seeded vacuum initial conditions in FP32-bulk and FP64.
    After 400 leapfrog steps the scale factors agree to 1e-8, both Friedmann residuals stay small and
    agree, and the chi spectrum in the band k in [0.5, 1.5] agrees to 1e-3 on average."""
    say("case f32_bulk vs fp64 at 256^3")
    spec = lphi4(100.0, n_fields=2, phi0_gev=5.6964e18, lam=1.0, pi0_program=-0.5)
    N, kIR, dt = 256, 0.05, 0.005
    dx = 2 * math.pi / kIR / N
    out = {}
    for prec in (CLMI_F64, CLMI_F32_BULK):
        lat = Lattice(spec, N, dx, dt, prec, expansion=True)
        kin = sum(0.5 * p * p for p in spec.pi0)
        aDot0 = spec.fStar / spec.MPl * math.sqrt((kin + spec.potential(spec.fld0)) / 3.0)
        lat.set_background(aI=1.0, aSI=1.0, aIM=1.0, aDotI=aDot0)
        m2 = spec.mass2(spec.fld0)
        for f in range(2):
            lat.grf_scalar(f, 4242, f, m2[f], aDot0, kIR, N * kIR / 4, spec.fld0[f], spec.pi0[f])
        _, _, rhs = lat.hubble_constraint()
        lat.set_background(aDotI=math.sqrt(rhs))
        lat.lf_init()
        lat.lf_evolve(400)
        lat.lf_sync()
        hc = lat.hubble_constraint()[0]
        s, c, _ = lat.spectrum_bins(FLD_S, 1, 0, 1.0)
        out[prec] = (lat.background()["aI"], hc, s / np.maximum(c, 1))
        lat.close()
    a64, h64, p64 = out[CLMI_F64]
    a32, h32, p32 = out[CLMI_F32_BULK]
    assert a64 > 1.0
    assert abs(a32 - a64) / a64 < 1e-8
    assert abs(h64) < 1e-3 and abs(h32) < 1e-3
    assert abs(h32 - h64) < 1e-5
    n = np.arange(len(p64))
    band = (n * kIR >= 0.5) & (n * kIR <= 1.5)
    err = np.abs(p32[band] - p64[band]) / p64[band]
    assert err.mean() < 1e-3, err.mean()


def test_free_plane_wave_closed_form_256(gpu):
    """Full-size check at 256^3: a free massive plane wave follows phi_n = A cos(n W dt) exactly
    (size-independent property of the leapfrog recurrence), in FP64 and FP32-bulk."""
    say("case free_plane_wave_256")
    N, dx, dt, m2 = 256, 0.4909, 0.005, 3.0
    spec = ModelSpec(name="free", n_scalars=1, terms=[(0.5 * m2, [2])])
    n = (5, 17, 3)
    A = 0.5
    x = np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    shape = np.cos(2 * np.pi * (n[0] * X + n[1] * Y + n[2] * Z) / N)
    w2 = m2 + (4 / dx ** 2) * sum(math.sin(math.pi * ni / N) ** 2 for ni in n)
    W = math.acos(1 - w2 * dt * dt / 2) / dt
    steps = 200
    for prec, tol in ((CLMI_F64, 1e-12), (CLMI_F32_BULK, 2e-5)):
        lat = Lattice(spec, N, dx, dt, prec, expansion=False)
        lat.upload(FLD_S, 0, 0, A * shape)
        lat.lf_init()
        lat.lf_evolve(steps)
        lat.lf_sync()
        phi = lat.download(FLD_S, 0, 0)
        err = np.max(np.abs(phi - A * math.cos(steps * W * dt) * shape))
        say("prec", prec, "err", err)
        assert err < tol * A
        lat.close()


def test_f32_bulk_tracks_fp64(gpu, oracle_mod):
    """FP32-bulk (hi/lo fields, FP32 momenta, FP64 reductions) stays close to the FP64 oracle over
    200 leapfrog steps: fields within 1e-5 of their scale, Friedmann violation comparable."""
    say("case f32_bulk_vs_fp64")
    spec = bench_spec()
    N, dx, dt = 32, 0.4909, 0.005
    rng = np.random.default_rng(21)
    ics = scalar_ics(spec, N, rng)
    lat = Lattice(spec, N, dx, dt, CLMI_F32_BULK)
    orc = oracle_mod.OracleLattice(spec, N, dx, dt)
    load_both(lat, orc, ics)
    init_background(lat, orc)
    rg = run_reference_loop(lat, 200, 50)
    ro = run_reference_loop(orc, 200, 50)
    for k in [(FLD_S, 0), (FLD_S, 1), (PI_S, 0), (PI_S, 1)]:
        e = relerr(lat.download(*k, 0), orc.download(*k))
        say(k, e)
        assert e < 1e-5
    assert np.max(np.abs(rg[:, 0] - ro[:, 0])) < 1e-6
    assert lat.background()["aI"] == pytest.approx(orc.background()["aI"], rel=1e-8)


def test_deterministic_reductions(gpu):
    """Two identical runs give bit-identical fields and background (fixed-order reductions)."""
    say("case determinism")
    spec = bench_spec()
    N = 32
    rng = np.random.default_rng(5)
    ics = scalar_ics(spec, N, rng)
    outs = []
    for _ in range(2):
        lat = Lattice(spec, N, 0.4909, 0.005, CLMI_F64)
        for (k, i), a in ics.items():
            lat.upload(k, i, 0, a)
        lat.set_background(aDotI=0.5)
        lat.lf_init()
        lat.lf_evolve(30)
        outs.append((lat.download(FLD_S, 1, 0), lat.background()))
        lat.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1]


def u1_spec():
    spec = lphi4U1(e2_over_lambda=10.0, lam=1.0)
    return spec


def u1_ics(spec, N, rng, gauss_consistent=False):
    f = {}
    for c in range(2):
        f[(FLD_CS, c)] = spec.fld0[c] + smooth_random(N, rng, amp=0.1)
        f[(PI_CS, c)] = spec.pi0[c] + smooth_random(N, rng, amp=0.1)
    for d in range(3):
        f[(FLD_U1, d)] = smooth_random(N, rng, amp=0.05)
        f[(PI_U1, d)] = smooth_random(N, rng, amp=0.05)
    if gauss_consistent:
        s = smooth_random(N, rng, amp=0.2)
        f[(PI_CS, 0)] = s * f[(FLD_CS, 0)]
        f[(PI_CS, 1)] = s * f[(FLD_CS, 1)]
        for d in range(3):
            f[(PI_U1, d)] = np.zeros((N, N, N))
    return f


def load_u1(lat, st):
    for (k, c), a in st.items():
        lat.upload(k, 0, c, a)


@pytest.mark.parametrize("N", [16, 64])
def test_lf_lphi4u1_fp64_parity(gpu, oracle_mod, N):
    """Charged scalar + U(1), expansion on, 40 leapfrog steps with measurements every 10: all ten
    arrays and the background match the oracle within 1e-10; Gauss checks agree (N = 16: the
    one-site-per-thread pass; N = 64: the tiled pass)."""
    say(f"case lf_lphi4U1_fp64 N={N}")
    spec = u1_spec()
    dx, dt = 0.6, 0.01
    rng = np.random.default_rng(31)
    st = u1_ics(spec, N, rng)
    lat = Lattice(spec, N, dx, dt, CLMI_F64)
    orc = oracle_mod.OracleLattice(spec, N, dx, dt)
    load_u1(lat, st)
    load_u1(orc, st)
    init_background(lat, orc)
    rg = run_reference_loop(lat, 40, 10)
    ro = run_reference_loop(orc, 40, 10)
    for (k, c) in st:
        e = relerr(lat.download(k, 0, c), orc.download(k, 0, c))
        say((k, c), e)
        assert e < FP64_TOL, (k, c)
    np.testing.assert_allclose(rg[:, 1:], ro[:, 1:], rtol=FP64_TOL)
    gg, go = lat.gauss_u1(), orc.gauss_u1()
    np.testing.assert_allclose(gg, go, rtol=1e-9)


@pytest.mark.parametrize("N", [16, 64])
def test_vv4_lphi4u1_fp64_parity(gpu, oracle_mod, N):
    """Charged scalar + U(1) with velocity Verlet (VV4: three stages, each a kick pass and a fused
    drift + kick pass), expansion on, 10 steps with measurements every 5: all ten arrays and the
    Friedmann records match the oracle within 1e-10."""
    say(f"case vv4_lphi4U1_fp64 N={N}")
    ws = [1.351207191959657771818, -1.702414403875838200264]
    spec = u1_spec()
    dx, dt = 0.6, 0.01
    st = u1_ics(spec, N, np.random.default_rng(41))
    lat = Lattice(spec, N, dx, dt, CLMI_F64)
    orc = oracle_mod.OracleLattice(spec, N, dx, dt)
    load_u1(lat, st)
    load_u1(orc, st)
    init_background(lat, orc)
    rg = run_reference_loop(lat, 10, 5, evolver="VV", ws=ws)
    ro = run_reference_loop(orc, 10, 5, evolver="VV", ws=ws)
    for (k, c) in st:
        e = relerr(lat.download(k, 0, c), orc.download(k, 0, c))
        say((k, c), e)
        assert e < FP64_TOL, (k, c)
    np.testing.assert_allclose(rg[:, 1:], ro[:, 1:], rtol=FP64_TOL)


@pytest.mark.parametrize("N", [16, 64])
def test_pv_lphi4u1_fp64_parity(gpu, oracle_mod, N):
    """Charged scalar + U(1) with position Verlet (PV2), expansion on, 20 steps with measurements
    every 10: all ten arrays and the Friedmann records match the oracle within 1e-10."""
    say(f"case pv_lphi4U1_fp64 N={N}")
    spec = u1_spec()
    dx, dt = 0.6, 0.01
    st = u1_ics(spec, N, np.random.default_rng(37))
    lat = Lattice(spec, N, dx, dt, CLMI_F64)
    orc = oracle_mod.OracleLattice(spec, N, dx, dt)
    load_u1(lat, st)
    load_u1(orc, st)
    init_background(lat, orc)
    rg = run_reference_loop(lat, 20, 10, evolver="PV")
    ro = run_reference_loop(orc, 20, 10, evolver="PV")
    for (k, c) in st:
        e = relerr(lat.download(k, 0, c), orc.download(k, 0, c))
        say((k, c), e)
        assert e < FP64_TOL, (k, c)
    np.testing.assert_allclose(rg[:, 1:], ro[:, 1:], rtol=FP64_TOL)


@pytest.mark.parametrize("N,prec,ty8", [(64, CLMI_F64, False), (64, CLMI_F32_BULK, False),
                                        (128, CLMI_F32_BULK, False), (128, CLMI_F64, False),
                                        (64, CLMI_F64, True), (128, CLMI_F32_BULK, True),
                                        (192, CLMI_F64, False), (192, CLMI_F32_BULK, False)])
def test_tiled_gauge_pass_bit_identical_to_generic(gpu, N, prec, ty8):
    """The tiled charged-scalar + U(1) pass gives the same fields, bit for bit, as the
    one-site-per-thread pass (flag CLMI_FLAG_NO_TILED); volume averages agree to summation order."""
    from cosmolattice_amd.clmi import CLMI_FLAG_GAUGE_TY8, CLMI_FLAG_NO_TILED

    say(f"case gauge tiled N={N} prec={prec} ty8={ty8}")
    spec = u1_spec()
    dx, dt = 0.6, 0.01
    rng = np.random.default_rng(35)
    st = u1_ics(spec, N, rng)
    res = []
    for flags in (CLMI_FLAG_GAUGE_TY8 if ty8 else 0, CLMI_FLAG_NO_TILED):
        lat = Lattice(spec, N, dx, dt, prec, expansion=False, flags=flags)
        plan = assert_plan(lat, "k_gauge_pass" if flags == CLMI_FLAG_NO_TILED else "k_gauge_tiled")
        assert flags == CLMI_FLAG_NO_TILED or f"TY={8 if ty8 else 4}" in plan
        load_u1(lat, st)
        lat.set_background(aI=1.0, aSI=1.0, aIM=1.0, aDotI=0.0)
        lat.lf_init()
        lat.lf_evolve(7)
        lat.lf_sync()
        res.append(({kc: lat.download(kc[0], 0, kc[1]) for kc in st}, lat.averages(), lat.gauss_u1()))
        lat.close()
    for kc in st:
        assert np.array_equal(res[0][0][kc], res[1][0][kc]), kc
    for key in res[0][1]:
        assert res[0][1][key] == pytest.approx(res[1][1][key], rel=1e-12, abs=1e-300), key
    np.testing.assert_allclose(res[0][2], res[1][2], rtol=1e-9)


@pytest.mark.parametrize("N,prec", [(64, CLMI_F64), (64, CLMI_F32_BULK), (128, CLMI_F32_BULK)])
def test_tiled_gauge_verlet_bit_identical_to_generic(gpu, N, prec):
    """Velocity and position Verlet (fourth order, 4 steps) through the tiled charged-scalar + U(1)
    pass give fields bit-identical to the one-site-per-thread pass on a static background: the
    kick-only, drift-only and drift+kick forms of the tiled pass."""
    from cosmolattice_amd.clmi import CLMI_FLAG_NO_TILED, PV_WEIGHTS, VV_WEIGHTS

    say(f"case gauge tiled verlet N={N} prec={prec}")
    spec = u1_spec()
    rng = np.random.default_rng(36)
    st = u1_ics(spec, N, rng)
    for scheme in ("vv", "pv"):
        res = []
        for flags in (0, CLMI_FLAG_NO_TILED):
            lat = Lattice(spec, N, 0.6, 0.01, prec, expansion=False, flags=flags)
            assert_plan(lat, "k_gauge_tiled" if flags == 0 else "k_gauge_pass")
            load_u1(lat, st)
            lat.set_background(aI=1.1, aSI=1.05, aIM=1.0, aDotI=0.0)
            if scheme == "vv":
                lat.vv_evolve(VV_WEIGHTS["VV4"], 4)
            else:
                lat.pv_evolve(PV_WEIGHTS["PV4"], 4)
            res.append({kc: lat.download(kc[0], 0, kc[1]) for kc in st})
            lat.close()
        for kc in st:
            assert np.array_equal(res[0][kc], res[1][kc]), (scheme, kc)


def test_lphi4u1_gauss_conserved_fp64_and_f32(gpu):
    """Gauss-consistent initial data keep the Gauss violation at rounding level on the GPU."""
    say("case u1_gauss")
    spec = u1_spec()
    N, dx, dt = 32, 0.6, 0.01
    rng = np.random.default_rng(33)
    st = u1_ics(spec, N, rng, gauss_consistent=True)
    for prec, tol in ((CLMI_F64, 1e-11), (CLMI_F32_BULK, 1e-4)):
        lat = Lattice(spec, N, dx, dt, prec, expansion=False)
        load_u1(lat, st)
        lat.lf_init()
        lat.lf_evolve(50)
        lat.lf_sync()
        diff, lhs, rhs = lat.gauss_u1()
        say("prec", prec, diff, lhs, rhs)
        assert lhs > 1e-4
        assert diff < tol
        lat.close()


def test_spectrum_bins_parity(gpu, oracle_mod):
    """hipFFT R2C + deterministic shell binning equals the numpy restatement within 1e-12, and a
    repeated call is bit-identical.  The cases cover the streaming pass at the production size (256^3,
    8.4 M modes, 1024 workgroups) and with four and with one LDS
    histogram per workgroup (nbins <= 2048 and 2048 < nbins <= 8192), small N (lines shorter than a
    wave: up to 5 and 14 lines per 64-mode chunk) and the bin-sorted table used above 8192 bins.  In
    FP64 the R2C transform reads the field array itself, so the field must come back unchanged at every
    size checked, including the ensemble sizes 64^3 and 192^3."""
    say("case spectrum")
    spec = bench_spec()
    cases = ((32, 1.0, (CLMI_F64, CLMI_F32_BULK)), (24, 0.5, (CLMI_F64, CLMI_F32_BULK)), (128, 1.0, (CLMI_F64,)),
             (256, 1.0, (CLMI_F32_BULK, CLMI_F64)),
             (30, 0.7, (CLMI_F64,)), (8, 0.4, (CLMI_F64,)), (24, 0.01, (CLMI_F64,)), (24, 0.002, (CLMI_F64,)),
             (64, 1.0, (CLMI_F64,)), (192, 1.0, (CLMI_F64,)))
    for N, dk, precs in cases:
        rng = np.random.default_rng(N)
        f = 1.0 + smooth_random(N, rng) + 0.1 * rng.normal(size=(N, N, N))
        for prec in precs:
            say("spectrum", N, dk, prec)
            lat = Lattice(spec, N, 0.5, 0.01, prec)
            lat.upload(FLD_S, 1, 0, f)
            s, c, m = lat.spectrum_bins(FLD_S, 1, 0, dk)
            s2, _, _ = lat.spectrum_bins(FLD_S, 1, 0, dk)
            # the transform reads the field array itself in FP64: it must come back unchanged
            assert np.array_equal(lat.download(FLD_S, 1, 0), f) or prec == CLMI_F32_BULK
            so, co, mo = oracle_mod.spectrum_bins(lat.download(FLD_S, 1, 0), dk)
            np.testing.assert_array_equal(s, s2)
            np.testing.assert_array_equal(c, co)
            # the oracle's mean |n| is a sequential double sum over up to ~1e5 modes per bin
            np.testing.assert_allclose(m, mo, rtol=1e-12)
            np.testing.assert_allclose(s, so, rtol=1e-12, atol=1e-12 * so.max())
            lat.close()


@pytest.mark.parametrize("N,dk", [(32, 1.0), (256, 1.0), (24, 0.002)])
def test_spectrum_version3_weighted_bins(gpu, oracle_mod, N, dk):
    """PS_version 3 bins |n|^p |f_n|^2 (p = 1 for PS_type 1, 3 for PS_type 2;
    measurements/powerspectrum.h:111-131): the weighted sums of both binning passes (streaming at
    dk = 1, the bin-sorted table at dk = 0.002) equal the numpy restatement within 1e-12, and a plane
    wave cos(2 pi n0.x / N) of amplitude A gives the closed forms PS_1 = |n0| A^2 / 2 and
    PS_2 = 2 pi |n0|^3 A^2 / count in the bin of |n0| (the host's prefactors, through the C ABI)."""
    say("case spectrum version 3", N, dk)
    spec = bench_spec()
    rng = np.random.default_rng(N + 3)
    f = 1.0 + smooth_random(N, rng) + 0.1 * rng.normal(size=(N, N, N))
    lat = Lattice(spec, N, 0.5, 0.01, CLMI_F64)
    lat.upload(FLD_S, 1, 0, f)
    for p in (1, 3):
        s, c, m = lat.spectrum_bins(FLD_S, 1, 0, dk, npow=p)
        so, co, mo = oracle_mod.spectrum_bins(f, dk, npow=p)
        np.testing.assert_array_equal(c, co)
        np.testing.assert_allclose(s, so, rtol=1e-12, atol=1e-12 * so.max())
    if dk == 1.0:
        n0, A = (3, 4, 0), 0.7  # |n0| = 5
        x = np.arange(N)
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        lat.upload(FLD_S, 1, 0, A * np.cos(2 * np.pi * (n0[0] * X + n0[1] * Y + n0[2] * Z) / N))
        _, ps1 = lat.power_spectrum(FLD_S, 1, 0, 1.0, ps_type=1, ps_version=3)
        _, ps2 = lat.power_spectrum(FLD_S, 1, 0, 1.0, ps_type=2, ps_version=3)
        _, c, _ = lat.spectrum_bins(FLD_S, 1, 0, 1.0)
        assert ps1[5] == pytest.approx(5 * A * A / 2, rel=1e-12)
        assert ps2[5] == pytest.approx(2 * np.pi * 125 * A * A / c[5], rel=1e-12)
        assert np.abs(np.delete(ps1, 5)).max() < 1e-12 * ps1[5]
    lat.close()


def test_spectrum_leaves_fp64_field_unchanged_512(gpu):
    """In FP64 the spectrum's R2C transform reads the field array itself (out of place): at 512^3, where
    hipFFT plans differently from the smaller sizes, the field comes back bit-identical and a repeated
    spectrum is bit-identical."""
    say("case spectrum 512 fp64 input preserved")
    spec = bench_spec()
    N = 512
    f = np.random.default_rng(512).normal(size=(N, N, N))
    lat = Lattice(spec, N, 0.5, 0.01, CLMI_F64)
    lat.upload(FLD_S, 1, 0, f)
    s1, _, _ = lat.spectrum_bins(FLD_S, 1, 0, 1.0)
    s2, _, _ = lat.spectrum_bins(FLD_S, 1, 0, 1.0)
    assert np.array_equal(lat.download(FLD_S, 1, 0), f)
    assert np.array_equal(s1, s2)
    # Parseval: the bins hold the full N^3 mode set
    assert s1.sum() == pytest.approx(N ** 3 * np.sum(f * f), rel=1e-11)
    lat.close()


@pytest.mark.parametrize("prec", [CLMI_F64, CLMI_F32_BULK])
def test_norm_averages_and_magnetic_spectrum(gpu, oracle_mod, prec):
    """<|v|>, <|v|^2> of the complex scalar, its momentum, the U(1) momenta and the magnetic field
    equal numpy on the lattice's own arrays; <|B|^2> equals the evolver's field-strength reduction;
    the spectrum of each B component equals the oracle's binning of numpy's B, and those of the complex
    scalar's and the gauge field's components equal the oracle's binning of the lattice's arrays."""
    say("case norms", prec)
    spec = u1_spec()
    N, dx = 32, 0.7
    rng = np.random.default_rng(17)
    st = u1_ics(spec, N, rng)
    for d in range(3):
        st[(FLD_U1, d)] = st[(FLD_U1, d)] + 0.02 * rng.normal(size=(N, N, N))
    lat = Lattice(spec, N, dx, 0.01, prec)
    load_u1(lat, st)
    got = {(k, c): lat.download(k, 0, c) for (k, c) in st}
    for kind, nc in ((FLD_CS, 2), (PI_CS, 2), (PI_U1, 3)):
        q = sum(got[(kind, c)] ** 2 for c in range(nc))
        m1, m2 = lat.norm_averages(kind)
        assert m1 == pytest.approx(np.mean(np.sqrt(q)), rel=1e-12)
        assert m2 == pytest.approx(np.mean(q), rel=1e-12)
    B = oracle_mod.magnetic_field([got[(FLD_U1, d)] for d in range(3)], dx)
    q = B[0] ** 2 + B[1] ** 2 + B[2] ** 2
    m1, m2 = lat.norm_averages(MAG_U1)
    assert m1 == pytest.approx(np.mean(np.sqrt(q)), rel=1e-12)
    assert m2 == pytest.approx(np.mean(q), rel=1e-12)
    # the evolver's FP32-bulk reduction rounds each field-strength difference to FP32
    assert m2 == pytest.approx(lat.averages()["B2U1"], rel=1e-12 if prec == CLMI_F64 else 1e-6)
    for c in range(3):
        s, cnt, _ = lat.spectrum_bins(MAG_U1, 0, c, 1.0)
        so, co, _ = oracle_mod.spectrum_bins(B[c], 1.0)
        np.testing.assert_array_equal(cnt, co)
        np.testing.assert_allclose(s, so, rtol=1e-11, atol=1e-12 * so.max())
    # spectra of the complex scalar's and the gauge momenta's components
    for kind, c in ((FLD_CS, 0), (FLD_CS, 1), (PI_CS, 1), (FLD_U1, 2), (PI_U1, 0)):
        s, cnt, _ = lat.spectrum_bins(kind, 0, c, 1.0)
        so, co, _ = oracle_mod.spectrum_bins(got[(kind, c)], 1.0)
        np.testing.assert_array_equal(cnt, co)
        np.testing.assert_allclose(s, so, rtol=1e-11, atol=1e-12 * so.max())
    with pytest.raises(Exception):
        lat.norm_averages(FLD_S)
    lat.close()


def test_grf_parity(gpu, oracle_mod):
    """Gaussian-random-field initial conditions on the GPU equal the numpy restatement (same
    Philox stream) to rounding, in both precisions."""
    say("case grf")
    spec = bench_spec()
    N, kIR = 32, 0.4
    dx = 2 * math.pi / kIR / N
    m2 = spec.mass2([1.0, 0.0])
    for prec, tol in ((CLMI_F64, 1e-12), (CLMI_F32_BULK, 1e-5)):
        lat = Lattice(spec, N, dx, 0.005, prec)
        for f in range(2):
            lat.grf_scalar(f, 12345, f, m2[f], 0.3, kIR, N * kIR / 4, spec.fld0[f], spec.pi0[f])
            phi_o, pi_o = oracle_mod.grf_scalar(N, dx, spec, 12345, f, m2[f], 0.3, kIR, N * kIR / 4,
                                                spec.fld0[f], spec.pi0[f])
            ef = np.max(np.abs(lat.download(FLD_S, f, 0) - phi_o)) / np.max(np.abs(phi_o - spec.fld0[f]))
            ep = np.max(np.abs(lat.download(PI_S, f, 0) - pi_o)) / np.max(np.abs(pi_o - spec.pi0[f]))
            say("grf", prec, f, ef, ep)
            assert ef < tol and ep < tol
        lat.close()


def test_grf_u1_parity(gpu, oracle_mod):
    """Complex scalar + U(1) initial conditions on the GPU equal the numpy restatement (same Philox
    stream) to rounding; the GPU's own Gauss check then reads zero violation."""
    say("case grf_u1")
    spec = lphi4U1(10.0, phi0_gev=5.6964e18, lam=1.0, pi0_gev2=0.0)
    N, kIR = 32, 0.4
    dx = 2 * math.pi / kIR / N
    m2 = spec.cs_mass2(spec.fld0)
    for prec, tol in ((CLMI_F64, 1e-12), (CLMI_F32_BULK, 1e-5)):
        lat = Lattice(spec, N, dx, 0.005, prec, expansion=False)
        lat.grf_u1(99, 1, m2, 0.3, kIR, N * kIR / 4, spec.fld0, spec.pi0)
        phi, pi, piA = oracle_mod.grf_u1(N, dx, spec, 99, 1, m2, 0.3, kIR, N * kIR / 4, spec.fld0, spec.pi0)
        for c in range(2):
            ef = np.max(np.abs(lat.download(FLD_CS, 0, c) - phi[c])) / np.max(np.abs(phi[c] - spec.fld0[c]))
            ep = np.max(np.abs(lat.download(PI_CS, 0, c) - pi[c])) / np.max(np.abs(pi[c] - spec.pi0[c]))
            say("grf_u1 cs", prec, c, ef, ep)
            assert ef < tol and ep < tol
        for d in range(3):
            assert np.all(lat.download(FLD_U1, 0, d) == 0)
            ea = np.max(np.abs(lat.download(PI_U1, 0, d) - piA[d])) / np.max(np.abs(piA[d]))
            say("grf_u1 piA", prec, d, ea)
            assert ea < tol
        g = lat.gauss_u1()
        say("gauss", prec, g)
        assert g[0] < (1e-13 if prec == CLMI_F64 else 1e-6)
        lat.close()


@pytest.mark.parametrize("prec,group_bytes", [(CLMI_F32_BULK, 0), (CLMI_F64, 0), (CLMI_F32_BULK, 1),
                                               (CLMI_F32_BULK, 64 << 20)])
def test_ensemble_bit_identical_to_individual_runs(gpu, prec, group_bytes):
    """An ensemble of uneven lattices (64^3, 128^3, 192^3; different seeds, couplings and dt)
    advanced in batched, graph-replayed launches equals each lattice evolved on its own, bit for
    bit, after every call (graph replays interleaved with observations and syncs)."""
    from cosmolattice_amd.clmi import Ensemble

    say(f"case ensemble prec={prec} group_bytes={group_bytes}")
    kIR = 0.2
    members, twins = [], []
    for i, (N, q, dt) in enumerate([(64, 100.0, 0.005), (128, 80.0, 0.004), (192, 100.0, 0.005), (64, 120.0, 0.006)]):
        spec = bench_spec(q=q)
        dx = 2 * math.pi / kIR / N
        pair = []
        for _ in range(2):
            lat = Lattice(spec, N, dx, dt, prec, expansion=True)
            lat.set_background(aI=1.0, aSI=1.0, aIM=1.0, aDotI=0.5)
            m2 = spec.mass2(spec.fld0)
            for f in range(2):
                lat.grf_scalar(f, 1000 + i, f, m2[f], 0.5, kIR, N * kIR / 4, spec.fld0[f], spec.pi0[f])
            _, _, rhs = lat.hubble_constraint()
            lat.set_background(aDotI=math.sqrt(rhs))
            lat.lf_init()
            pair.append(lat)
        members.append(pair[0])
        twins.append(pair[1])
    ens = Ensemble(members, group_bytes=group_bytes)
    say(ens.autotune(1))  # changes launch plans only

    def compare(tag):
        for m, t in zip(members, twins):
            for f in range(2):
                for k in (FLD_S, PI_S):
                    assert np.array_equal(m.download(k, f, 0), t.download(k, f, 0)), (tag, m.N, k, f)
            assert m.background() == t.background(), (tag, m.N)

    for chunk in (1, 5, 4, 3):
        ens.lf_evolve(chunk)
        for t in twins:
            t.lf_evolve(chunk)
        compare(f"after {chunk}")
    for lat in members + twins:
        lat.lf_sync()
    compare("sync")
    ens.lf_evolve(6)
    for t in twins:
        t.lf_evolve(6)
    compare("after sync")
    ens.close()
    for lat in members + twins:
        lat.close()


def test_ensemble_rejects_unsupported_members(gpu):
    from cosmolattice_amd.clmi import ClmiError, Ensemble

    spec = bench_spec()
    a = Lattice(spec, 64, 0.5, 0.005, CLMI_F32_BULK)
    b = Lattice(spec, 64, 0.5, 0.005, CLMI_F64)
    c = Lattice(spec, 256, 0.5, 0.005, CLMI_F32_BULK)
    for group in ([a, b], [a, c], [a, a]):
        with pytest.raises(ClmiError):
            Ensemble(group)
    for lat in (a, b, c):
        lat.close()


def test_profile_timing_samples_one_pass_in_n(gpu):
    """clmi_profile_enable(ctx, n) times every n-th drift + kick pass with HIP events (n = 1: every pass):
    12 fused leapfrog steps give 12 timed launches with n = 1 and 3 with n = 4, and the average duration
    is positive; the fields do not depend on the timing."""
    say("case profile sampling")
    spec = bench_spec()
    N = 64
    ics = scalar_ics(spec, N, np.random.default_rng(12))
    outs = []
    for every in (1, 4):
        lat = Lattice(spec, N, 0.4909, 0.005, CLMI_F64, expansion=False)
        for (k, i), a in ics.items():
            lat.upload(k, i, 0, a)
        lat.set_background(aI=1.1, aSI=1.05)
        lat.lf_init()
        lat.lf_evolve(1)
        lat.profile_enable(True, every=every)
        lat.lf_evolve(12)
        ms, launches, bpl = lat.profile_read()
        lat.profile_enable(False)
        say("every", every, "launches", launches, "ms", ms)
        assert launches == (12 if every == 1 else 3)
        assert ms > 0 and bpl > 0
        lat.lf_sync()
        outs.append([lat.download(k, i, 0) for (k, i) in sorted(ics)])
        lat.close()
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


def test_invalid_arguments_are_rejected_with_a_message(gpu):
    """The C ABI refuses what it cannot run, with a message and without touching the device state:
    lattice sizes outside even 4..1024, a negative step count, the adjoint of a model with a complex scalar,
    unknown memory modes, and a plan-report buffer too short for the text."""
    import ctypes as C

    from cosmolattice_amd.clmi import CLMI_FLAG_MEMMODE, ClmiError

    say("case invalid arguments")
    spec = bench_spec()
    for n in (2, 33, 1026):
        with pytest.raises(ClmiError, match="N must be even"):
            Lattice(spec, n, 0.5, 0.005, CLMI_F64)
    with pytest.raises(ClmiError, match="memory mode"):
        Lattice(spec, 48, 0.5, 0.005, CLMI_F32_BULK, flags=CLMI_FLAG_MEMMODE(15))
    lat = Lattice(spec, 16, 0.5, 0.005, CLMI_F64)
    with pytest.raises(ClmiError, match="nsteps"):
        lat.lf_evolve(-1)
    buf = C.create_string_buffer(8)
    assert lat.lib.clmi_plan_report(lat._h, buf, len(buf)) != 0
    assert lat.plan().startswith("k_scalar_pass")
    lat.close()
    u1 = Lattice(u1_spec(), 16, 0.6, 0.01, CLMI_F64)
    z = [np.zeros((16,) * 3)] * u1.spec.n_scalars
    with pytest.raises((ClmiError, ValueError)):
        u1.adjoint_lf(3, z, z, (0.0, 0.0))
    u1.close()
