"""GPU tests at the full sizes of the five benchmark workloads, through the C ABI.

* lphi4, one field, 64^3, velocity Verlet (VV2), dt = 0.01, 1000 steps, FP64 with expansion: the whole
  run against the FP64 oracle from the same initial fields (evolvers/velocityverlet.h:51-114).
* lphi4U1 at 256^3 (e^2/lambda = 10, kIR = 0.05, dt = 0.005) from the Gauss-consistent vacuum initial
  conditions: the tiled charged-scalar + U(1) pass bit-identical to the generic pass, 5 FP64 leapfrog
  steps against the oracle, and 2000 steps in both precisions with the Gauss and energy constraints
  (evolvers/kernels/u1kernels.h:31-52, definitions/gausslaws.h:31-53, hubbleconstraint.h:27-36).
* 48-member ensemble, 16 lattices each at 64^3, 128^3 and 192^3, 20 steps in both grouping modes: every
  member bit-identical to its own single-lattice run.
* Reverse-mode leapfrog at 128^3 over 500 steps: dL/dq and two initial-mode derivatives of the FP64
  adjoint against FP64 central differences to 1e-6.
* A model with a real singlet next to the charged scalar + U(1) (both sectors in one pass and one
  background update): oracle parity at small N, and at 512^3 FP32-bulk, where the two sectors' partial
  sums need more rows than either alone.
* The largest accepted lattice, 1024^3, in both precisions: the default pass against the generic pass.

The 256^3 FP32-bulk vs FP64 comparison of the two-field lphi4 run over its full 4000 steps lives in
test_gpu_accuracy.py.
"""
import math

import numpy as np
import pytest

from cosmolattice_amd.clmi import (CLMI_F32_BULK, CLMI_F64, CLMI_FLAG_NO_TILED, FLD_CS, FLD_S, FLD_U1, PI_CS, PI_S,
                                   PI_U1, Lattice, ModelSpec, lphi4, lphi4U1)
from test_gpu_parity import FP64_TOL, init_background, relerr, say, smooth_random

pytestmark = pytest.mark.gpu


def vacuum_lattice(spec, N, kIR, dt, prec, seed, expansion=True, flags=0):
    """Seeded vacuum initial conditions as the reference's initializers set them: aDot from the
    homogeneous energy (scalefactorinitializer.h:49-64), Gaussian fluctuations cut at N kIR / 4, then
    aDot from the Friedmann constraint on the fluctuating fields (modelinitializer.h:71-80)."""
    dx = 2 * math.pi / kIR / N
    lat = Lattice(spec, N, dx, dt, prec, expansion=expansion, flags=flags)
    kin = sum(0.5 * p * p for p in spec.pi0)
    aDot0 = spec.fStar / spec.MPl * math.sqrt((kin + spec.potential(spec.fld0)) / 3.0)
    lat.set_background(aI=1.0, aSI=1.0, aIM=1.0, aDotI=aDot0)
    m2 = spec.mass2(spec.fld0)
    for f in range(spec.n_scalars):
        lat.grf_scalar(f, seed, f, m2[f], aDot0, kIR, N * kIR / 4, spec.fld0[f], spec.pi0[f])
    if expansion:
        _, _, rhs = lat.hubble_constraint()
        lat.set_background(aDotI=math.sqrt(rhs))
    lat.lf_init()
    return lat, dx


def test_c1_single_field_vv2_64_1000_steps(gpu, oracle_mod):
    """lphi4 with one field (lambda = 1, phi0 = 1, pi0 = -0.5 in program units, fStar/MPl = 2.34) on
    64^3, kIR = 0.1, VV2, dt = 0.01, 1000 steps with expansion, FP64: the field, its momentum, the
    background and the Friedmann records (every 100 steps) equal the oracle's within 1e-10."""
    say("case c1 64^3 VV2 1000 steps")
    spec = lphi4(0.0, n_fields=1, phi0_gev=5.6964e18, lam=1.0, pi0_program=-0.5)
    N, kIR, dt = 64, 0.1, 0.01
    lat, dx = vacuum_lattice(spec, N, kIR, dt, CLMI_F64, seed=11)
    orc = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=True)
    for k in (FLD_S, PI_S):
        orc.upload(k, 0, 0, lat.download(k, 0, 0))
    orc.set_background(**lat.background())
    rec_g, rec_o = [], []
    for _ in range(10):
        lat.vv_evolve([1.0], 100)
        orc.vv_evolve([1.0], 100)
        rec_g.append(lat.hubble_constraint())
        rec_o.append(orc.hubble_constraint())
    for k in (FLD_S, PI_S):
        e = relerr(lat.download(k, 0, 0), orc.download(k, 0))
        say("c1", k, e)
        assert e < FP64_TOL, k
    bg, bo = lat.background(), orc.background()
    for key in bo:
        if bo[key] != 0:
            assert abs(bg[key] - bo[key]) <= FP64_TOL * abs(bo[key]), key
    rec_g, rec_o = np.array(rec_g), np.array(rec_o)
    np.testing.assert_allclose(rec_g[:, 1:], rec_o[:, 1:], rtol=FP64_TOL)
    # the relative violation is a difference of nearly equal numbers: compare it absolutely
    np.testing.assert_allclose(rec_g[:, 0], rec_o[:, 0], rtol=0, atol=1e-12)
    assert bg["aI"] > 5.0  # the expansion is live over the run
    assert np.max(np.abs(rec_g[:, 0])) < 1e-3
    lat.close()


# ---- lphi4U1 at 256^3 -----------------------------------------------------------------------------

U1_N, U1_KIR, U1_DT = 256, 0.05, 0.005


def u1_vacuum(prec, seed=9000, expansion=True, flags=0):
    spec = lphi4U1(10.0, phi0_gev=5.6964e18, lam=1.0, pi0_gev2=0.0)
    dx = 2 * math.pi / U1_KIR / U1_N
    lat = Lattice(spec, U1_N, dx, U1_DT, prec, expansion=expansion, flags=flags)
    lat.set_background(aI=1.0, aSI=1.0, aIM=1.0, aDotI=0.0)
    lat.grf_u1(seed, 1, spec.cs_mass2(spec.fld0), 0.0, U1_KIR, U1_N * U1_KIR / 4, spec.fld0, spec.pi0)
    if expansion:
        _, _, rhs = lat.hubble_constraint()
        lat.set_background(aDotI=math.sqrt(rhs))
    else:
        lat.set_all_averages()
    return spec, lat, dx


U1_KINDS = [(FLD_CS, 0), (FLD_CS, 1), (PI_CS, 0), (PI_CS, 1)] + [(k, d) for k in (FLD_U1, PI_U1) for d in range(3)]


@pytest.mark.parametrize("prec", [CLMI_F64, CLMI_F32_BULK])
def test_c3_u1_256_tiled_bit_identical_to_generic(gpu, prec):
    """At 256^3 the tiled charged-scalar + U(1) pass gives all ten arrays bit-identical to the generic
    one-site-per-thread pass after 7 leapfrog steps from the vacuum initial conditions (static
    background, so the two passes' different summation orders cannot enter)."""
    say(f"case c3 tiled vs generic prec={prec}")
    res = []
    for flags in (0, CLMI_FLAG_NO_TILED):
        _, lat, _ = u1_vacuum(prec, expansion=False, flags=flags)
        plan = lat.plan()
        say("plan", plan.replace("\n", "; "))
        assert ("k_gauge_tiled" if flags == 0 else "k_gauge_pass") in plan
        lat.lf_init()
        lat.lf_evolve(7)
        lat.lf_sync()
        res.append({kc: lat.download(kc[0], 0, kc[1]) for kc in U1_KINDS})
        lat.close()
    for kc in U1_KINDS:
        assert np.array_equal(res[0][kc], res[1][kc]), kc


@pytest.mark.parametrize("prec", [CLMI_F64, CLMI_F32_BULK])
def test_c3_u1_256_autotune_keeps_state_and_results(gpu, prec):
    """At 256^3 the autotuner times every x-chunk of the tiled charged-scalar + U(1) pass on the live
    lattice without changing it, and the chunk it picks gives fields, momenta and the background (volume
    sums included) bit-identical to the default chunk over 6 leapfrog steps with expansion: the pass
    writes one partial row per group of 32 planes (at 256^3) whatever the chunk."""
    say(f"case c3 autotune prec={prec}")
    outs = []
    for tune in (False, True):
        _, lat, _ = u1_vacuum(prec)
        lat.lf_init()
        lat.lf_evolve(3)
        if tune:
            before = {kc: lat.download(kc[0], 0, kc[1]) for kc in U1_KINDS}
            bg_before = lat.background()
            rep = lat.autotune(2)
            say(rep)
            assert "gauge XC=" in rep
            assert all(np.array_equal(before[kc], lat.download(kc[0], 0, kc[1])) for kc in U1_KINDS)
            assert lat.background() == bg_before
        lat.lf_evolve(3)
        lat.lf_sync()
        outs.append(({kc: lat.download(kc[0], 0, kc[1]) for kc in U1_KINDS}, lat.background()))
        lat.close()
    for kc in U1_KINDS:
        assert np.array_equal(outs[0][0][kc], outs[1][0][kc]), kc
    assert outs[0][1] == outs[1][1]


def test_c3_u1_256_fp64_matches_oracle(gpu, oracle_mod):
    """lphi4U1 at 256^3 with expansion: 5 FP64 leapfrog steps plus a sync on the GPU equal the oracle's
    from the same initial arrays and background, within 1e-10 on all ten arrays, the background and the
    Friedmann and Gauss records."""
    say("case c3 fp64 vs oracle")
    spec, lat, dx = u1_vacuum(CLMI_F64)
    orc = oracle_mod.OracleLattice(spec, U1_N, dx, U1_DT, expansion=True)
    for k, c in U1_KINDS:
        orc.upload(k, 0, c, lat.download(k, 0, c))
    orc.set_background(**lat.background())
    for obj in (lat, orc):
        obj.lf_init()
        obj.lf_evolve(5)
        obj.lf_sync()
    for k, c in U1_KINDS:
        e = relerr(lat.download(k, 0, c), orc.download(k, 0, c))
        say("c3", (k, c), e)
        assert e < FP64_TOL, (k, c)
    bg, bo = lat.background(), orc.background()
    for key in bo:
        if bo[key] != 0:
            assert abs(bg[key] - bo[key]) <= FP64_TOL * abs(bo[key]), key
    hg, ho = lat.hubble_constraint(), orc.hubble_constraint()
    np.testing.assert_allclose(hg[1:], ho[1:], rtol=FP64_TOL)
    gg, go = lat.gauss_u1(), orc.gauss_u1()
    np.testing.assert_allclose(gg[1:], go[1:], rtol=1e-9)
    # the ratio is rounding noise over a sum of squares (<= 1e-15); both must be at that level
    assert gg[0] < 1e-13 and go[0] < 1e-13
    lat.close()


def test_c3_u1_256_long_run_constraints(gpu):
    """lphi4U1 at 256^3 over 2000 leapfrog steps with expansion, FP64 and FP32-bulk from the same
    initial conditions: the Gauss violation stays at rounding level (FP64 < 1e-12, FP32-bulk < 1e-5), the
    Friedmann (energy-constraint) violation of both modes stays below 1e-3 and FP32-bulk's differs from
    FP64's by less than 1e-5 at every check (every 250 steps), and the scale factors agree to 2^-24.
    FP32-bulk forms each site's energy terms in FP32 (summed over sites in FP64); the lattice is close to
    homogeneous, so their rounding is nearly the same at every site and does not average out: the energy
    density, and through the Friedmann equation the scale factor (a ~ 190 here), carries a relative error
    of order 2^-24."""
    say("case c3 long run")
    recs = {}
    for prec in (CLMI_F64, CLMI_F32_BULK):
        _, lat, _ = u1_vacuum(prec)
        lat.lf_init()
        r = []
        for _ in range(8):
            lat.lf_evolve(250)
            lat.lf_sync()
            r.append((lat.hubble_constraint()[0], lat.gauss_u1()[0], lat.background()["aI"]))
        recs[prec] = np.array(r)
        say("c3 long", prec, r[-1])
        lat.close()
    r64, r32 = recs[CLMI_F64], recs[CLMI_F32_BULK]
    assert np.all(r64[:, 1] < 1e-12)
    assert np.all(r32[:, 1] < 1e-5)
    assert np.all(np.abs(r64[:, 0]) < 1e-3) and np.all(np.abs(r32[:, 0]) < 1e-3)
    assert np.max(np.abs(r32[:, 0] - r64[:, 0])) < 1e-5
    assert r64[-1, 2] > 1.0
    assert np.max(np.abs(r32[:, 2] - r64[:, 2]) / r64[:, 2]) < 2.0**-24


# ---- batched ensemble ------------------------------------------------------------------------------

def test_c4_ensemble_48_members_bit_identical(gpu):
    """48 two-field lphi4 realisations (q = 100, kIR = 0.05, dt = 0.005, distinct seeds), 16 each at 64^3,
    128^3 and 192^3, FP32-bulk: 10 steps in the default grouping (one group per size), then 10 in
    cache-sized groups (128 MiB), with a sync between.  Every member equals the same lattice advanced
    on its own, bit for bit, fields, momenta and background."""
    from cosmolattice_amd.clmi import Ensemble

    say("case c4 ensemble 48")
    spec = lphi4(100.0, n_fields=2, phi0_gev=5.6964e18, lam=1.0, pi0_program=-0.5)
    sizes = [64] * 16 + [128] * 16 + [192] * 16
    members = [vacuum_lattice(spec, N, 0.05, 0.005, CLMI_F32_BULK, 7000 + i)[0] for i, N in enumerate(sizes)]
    for gb, steps in ((0, 10), (128 << 20, 10)):
        ens = Ensemble(members, group_bytes=gb)
        ens.lf_evolve(steps)
        say("groups", ens.groups)  # formed at the first step
        assert ens.groups == (3 if gb == 0 else 25)
        ens.close()
        for m in members:
            m.lf_sync()
    for i, (m, N) in enumerate(zip(members, sizes)):
        twin, _ = vacuum_lattice(spec, N, 0.05, 0.005, CLMI_F32_BULK, 7000 + i)
        twin.lf_evolve(10)
        twin.lf_sync()
        twin.lf_evolve(10)
        twin.lf_sync()
        for f in range(2):
            for k in (FLD_S, PI_S):
                assert np.array_equal(m.download(k, f, 0), twin.download(k, f, 0)), (i, N, k, f)
        assert m.background() == twin.background(), (i, N)
        twin.close()
    for m in members:
        m.close()


# ---- adjoint at 128^3 -----------------------------------------------------------------------------

def test_c5_adjoint_128_500_steps_matches_central_differences(gpu):
    """Reverse-mode leapfrog on two-field lphi4 at 128^3 (q = 100, kIR = 0.05, dt = 0.005, expansion) over
    500 steps; loss = final energy fraction of chi + chi power in k in [0.5, 1.5].  dL/dq and the
    derivatives by the real parts of two initial phi modes from the FP64 adjoint equal FP64 central
    differences (relative steps 1e-5 for q, 1e-6 for the modes) within 1e-6, and dL/dq with the
    FP32-bulk forward within 1e-4."""
    import torch

    from cosmolattice_amd.adjoint import chi_energy_fraction_and_band, leapfrog_run, mode_gradients

    say("case c5 adjoint 128^3 500 steps")
    N, kIR, dt, steps = 128, 0.05, 0.005, 500
    spec = lphi4(100.0, n_fields=2, phi0_gev=5.6964e18, lam=1.0, pi0_program=-0.5)
    lat, dx = vacuum_lattice(spec, N, kIR, dt, CLMI_F64, seed=5000)
    phi0 = np.stack([lat.download(FLD_S, f, 0) for f in range(2)])
    pi0 = np.stack([lat.download(PI_S, f, 0) for f in range(2)])
    lat.close()
    coeffs = [c for c, _ in spec.terms]

    def run(cvals, p0, grads):
        c = torch.tensor(cvals, dtype=torch.float64, requires_grad=grads)
        p = torch.from_numpy(p0.copy()).requires_grad_(grads)
        q = torch.from_numpy(pi0.copy()).requires_grad_(grads)
        phiT, piT, bgT = leapfrog_run(c, p, q, spec, N, dx, dt, steps, CLMI_F64, True, 0)
        L = chi_energy_fraction_and_band(phiT, piT, bgT, c, spec, dx, kIR, (0.5, 1.5))
        if not grads:
            return float(L.detach())
        L.backward()
        return float(L.detach()), c.grad.numpy(), p.grad.numpy()

    _, gc, gp = run(coeffs, phi0, True)
    h = 1e-5 * coeffs[1]
    cp, cm = list(coeffs), list(coeffs)
    cp[1] += h
    cm[1] -= h
    # q enters as the coefficient q/2 of phi^2 chi^2
    fd_q = 0.5 * (run(cp, phi0, False) - run(cm, phi0, False)) / (2 * h)
    say("dL/dq adjoint", 0.5 * gc[1], "central difference", fd_q)
    assert abs(0.5 * gc[1] - fd_q) <= 1e-6 * abs(fd_q)
    # the FP32-bulk forward with the FP64 adjoint meets the 1e-4 bar against the same differences
    c32 = torch.tensor(coeffs, dtype=torch.float64, requires_grad=True)
    p32 = torch.from_numpy(phi0.copy()).requires_grad_(True)
    q32 = torch.from_numpy(pi0.copy()).requires_grad_(True)
    phiT, piT, bgT = leapfrog_run(c32, p32, q32, spec, N, dx, dt, steps, CLMI_F32_BULK, True, 0)
    chi_energy_fraction_and_band(phiT, piT, bgT, c32, spec, dx, kIR, (0.5, 1.5)).backward()
    say("dL/dq FP32-bulk adjoint", 0.5 * float(c32.grad[1]))
    assert abs(0.5 * float(c32.grad[1]) - fd_q) <= 1e-4 * abs(fd_q)
    modes = [np.array([-9, 7, 20]), np.array([-4, 14, 5])]
    mg = mode_gradients(gp[0], modes)
    x = np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    for j, n in enumerate(modes):
        shape = 2.0 / N ** 3 * np.cos(2 * np.pi * (n[0] * X + n[1] * Y + n[2] * Z) / N)
        eps = 1e-6 * np.max(np.abs(phi0[0])) / np.max(np.abs(shape))
        pp, pm = phi0.copy(), phi0.copy()
        pp[0] += eps * shape
        pm[0] -= eps * shape
        fd = (run(coeffs, pp, False) - run(coeffs, pm, False)) / (2 * eps)
        say("mode", n, "adjoint", mg[j][0], "central difference", fd)
        assert abs(mg[j][0] - fd) <= 1e-6 * abs(fd)


# ---- singlet next to the charged scalar + U(1) ----------------------------------------------------

def mixed_spec():
    """A real singlet with V = phi^4/4 + m^2 phi^2/2 decoupled from lphi4U1's charged scalar."""
    spec = lphi4U1(10.0, phi0_gev=5.6964e18, lam=1.0, pi0_gev2=0.0)
    spec.name = "singlet+lphi4U1"
    spec.n_scalars = 1
    spec.terms = [(0.25, [4]), (0.5 * 0.7, [2])]
    spec.fld0 = [0.8] + list(spec.fld0)
    spec.pi0 = [-0.3] + list(spec.pi0)
    return spec


def mixed_ics(spec, N, rng):
    st = {}
    st[(FLD_S, 0)] = 0.8 + smooth_random(N, rng, amp=0.1)
    st[(PI_S, 0)] = -0.3 + smooth_random(N, rng, amp=0.1)
    for c in range(2):
        st[(FLD_CS, c)] = spec.fld0[1 + c] + smooth_random(N, rng, amp=0.1)
        st[(PI_CS, c)] = smooth_random(N, rng, amp=0.1)
    for d in range(3):
        st[(FLD_U1, d)] = smooth_random(N, rng, amp=0.05)
        st[(PI_U1, d)] = smooth_random(N, rng, amp=0.05)
    return st


def load_mixed(obj, st):
    for (k, c), a in st.items():
        if k in (FLD_S, PI_S):
            obj.upload(k, c, 0, a)
        else:
            obj.upload(k, 0, c, a)


def download_mixed(obj, k, c, gpu_side):
    if k in (FLD_S, PI_S):
        return obj.download(k, c, 0) if gpu_side else obj.download(k, c)
    return obj.download(k, 0, c)


@pytest.mark.parametrize("N", [16, 64])
def test_singlet_plus_u1_fp64_parity(gpu, oracle_mod, N):
    """A model with a real singlet and the charged scalar + U(1): both sectors run in the same step and
    share one background update.  40 leapfrog steps with expansion and measurements every 10 equal the
    oracle within 1e-10 (N = 16: generic passes; N = 64: full-row scalar pass + tiled U(1) pass)."""
    from test_gpu_parity import run_reference_loop

    say(f"case singlet+u1 N={N}")
    spec = mixed_spec()
    dx, dt = 0.6, 0.01
    st = mixed_ics(spec, N, np.random.default_rng(55))
    lat = Lattice(spec, N, dx, dt, CLMI_F64)
    orc = oracle_mod.OracleLattice(spec, N, dx, dt)
    load_mixed(lat, st)
    load_mixed(orc, st)
    init_background(lat, orc)
    rg = run_reference_loop(lat, 40, 10)
    ro = run_reference_loop(orc, 40, 10)
    for (k, c) in st:
        e = relerr(download_mixed(lat, k, c, True), download_mixed(orc, k, c, False))
        say((k, c), e)
        assert e < FP64_TOL, (k, c)
    np.testing.assert_allclose(rg[:, 1:], ro[:, 1:], rtol=FP64_TOL)
    lat.close()


def test_singlet_plus_u1_512_f32_partial_rows(gpu):
    """At 512^3 FP32-bulk the singlet's full-row pass writes 8192 partial-sum rows and the U(1) pass its
    own after them, more than either pass alone.  Three leapfrog steps with expansion agree with the same
    run through the generic passes (4096 rows each): the scale factor to 1e-12 and the fields to FP32
    rounding, so every partial row was written and summed where the update reads it."""
    say("case singlet+u1 512 f32")
    spec = mixed_spec()
    N, dx, dt = 512, 0.3, 0.005
    x = np.arange(N)
    wave = 0.05 * np.cos(2 * np.pi * (3 * x[:, None, None] + 5 * x[None, :, None] + 2 * x[None, None, :]) / N)
    out = []
    for flags in (0, CLMI_FLAG_NO_TILED):
        lat = Lattice(spec, N, dx, dt, CLMI_F32_BULK, expansion=True, flags=flags)
        lat.upload(FLD_S, 0, 0, 0.8 + wave)
        lat.upload(PI_S, 0, 0, np.full((N, N, N), -0.3))
        for c in range(2):
            lat.upload(FLD_CS, 0, c, spec.fld0[1 + c] + wave)
        lat.upload(PI_U1, 0, 1, wave)
        lat.set_background(aI=1.0, aSI=1.0, aIM=1.0, aDotI=1.0)
        _, _, rhs = lat.hubble_constraint()
        lat.set_background(aDotI=math.sqrt(rhs))
        lat.lf_init()
        lat.lf_evolve(3)
        lat.lf_sync()
        out.append((lat.background(), lat.download(FLD_S, 0, 0), lat.download(PI_CS, 0, 1)))
        lat.close()
    (b0, f0, p0), (b1, f1, p1) = out
    assert b0["aI"] == pytest.approx(b1["aI"], rel=1e-12)
    assert b0["aDotI"] == pytest.approx(b1["aDotI"], rel=1e-10)
    for key in ("pi2AvI", "CSpi2AvI", "U1pi2AvI", "grad2AvI", "CSgrad2AvI", "potAvI"):
        assert b0[key] == pytest.approx(b1[key], rel=1e-9), key
    assert relerr(f0, f1) < 1e-6 and relerr(p0, p1) < 1e-5


@pytest.mark.parametrize("prec", [CLMI_F32_BULK, CLMI_F64])
def test_max_size_1024_default_pass_matches_generic(gpu, prec):
    """The largest lattice the library accepts, N = 1024 (2^30 sites, the limit of 32-bit site indexing;
    51.5 GB of lattice arrays in FP32-bulk, 68.7 GB in FP64).  Two-field lphi4 vacuum initial conditions
    drawn on the device (the same seed twice), then 6 leapfrog steps with expansion through the default
    pass (the tiled pass; 65536 partial-sum rows per update) and through the generic pass.  The fields and
    momenta agree bitwise, compared through their mean, mean square and mean squared forward gradient,
    computed on the device by one kernel; the background agrees to summation-order rounding; the energy
    constraint stays at the leapfrog's own level.  Nothing of N^3 size crosses to the host."""
    say(f"case max size 1024 prec={prec}")
    spec = lphi4(100.0, n_fields=2, phi0_gev=5.6964e18, lam=1.0, pi0_program=-0.5)
    # the 256^3 workload's spacing and fluctuation cut-off (dx = 0.49, kCutOff = N kIR / 4 = 3.2) in a box
    # 4 times longer
    N, kIR, dt = 1024, 0.0125, 0.005
    res = []
    for flags in (0, CLMI_FLAG_NO_TILED):
        lat, _ = vacuum_lattice(spec, N, kIR, dt, prec, seed=77, flags=flags)
        plan = lat.plan()
        say("plan", plan.replace("\n", "; "))
        assert plan.split()[0] == ("k_scalar_tiled" if flags == 0 else "k_scalar_pass"), plan
        lat.lf_evolve(6)
        lat.lf_sync()
        stats = [lat.field_stats(k, f, 0) for k in (FLD_S, PI_S) for f in range(2)]
        res.append((stats, lat.background(), lat.hubble_constraint()))
        say("stats", stats[0], "aI", res[-1][1]["aI"], "constraint", res[-1][2])
        lat.close()
    (s0, b0, h0), (s1, b1, h1) = res
    assert s0 == s1
    for key in ("aI", "aSI", "aDotI", "pi2AvI", "grad2AvI", "potAvI"):
        assert b0[key] == pytest.approx(b1[key], rel=1e-11), key
    assert b0["aI"] > 1.0
    # (the 256^3 run's constraint stays below 1.6e-4 over its 4000 steps; here 7.1e-5 after 6)
    assert abs(h0[0]) < 2e-4 and h0[0] == pytest.approx(h1[0], rel=1e-6, abs=1e-12)
