"""GPU: the reverse-mode leapfrog (clmi_adjoint_lf, and its torch.autograd wrapper) against central
finite differences of the FP64 oracle's forward run, and FP32-bulk against FP64.

Forward map: setAllAverages and aDotI = sqrt(Friedmann RHS) from the initial fields, LeapFrog
constructor, nsteps steps, sync.  Loss: final energy fraction of chi plus chi's variance in a
k band (cosmolattice_amd.adjoint.chi_energy_fraction_and_band).  Tolerance of the gradient checks:
1e-4 relative, as the sensitivity runs require.
"""
import math

import numpy as np
import pytest
import torch

from cosmolattice_amd.adjoint import chi_energy_fraction_and_band, leapfrog_run, mode_gradients
from cosmolattice_amd.clmi import CLMI_F32_BULK, CLMI_F64, FLD_S, PI_S, Lattice
from test_gpu_parity import bench_spec, say, smooth_random

pytestmark = pytest.mark.gpu

N, DT, KIR = 16, 0.005, 0.4
DX = 2 * math.pi / KIR / N
BAND = (0.8, 2.4)


def initial_fields(spec, rng):
    phi = np.stack([spec.fld0[f] + smooth_random(N, rng, amp=0.05) + 0.01 * rng.normal(size=(N,) * 3)
                    for f in range(spec.n_scalars)])
    pi = np.stack([spec.pi0[f] + smooth_random(N, rng, amp=0.05) + 0.01 * rng.normal(size=(N,) * 3)
                   for f in range(spec.n_scalars)])
    return phi, pi


def oracle_loss(oracle_mod, spec, coeffs, phi0, pi0, nsteps, expansion=True):
    """The forward map on the FP64 oracle, then the loss (in torch, FP64, CPU)."""
    from cosmolattice_amd.adjoint import _spec_with

    s = _spec_with(spec, coeffs)
    orc = oracle_mod.OracleLattice(s, N, DX, DT, expansion=expansion)
    for f in range(2):
        orc.upload(FLD_S, f, 0, phi0[f])
        orc.upload(PI_S, f, 0, pi0[f])
    if expansion:
        _, _, rhs = orc.hubble_constraint()
        orc.set_background(aDotI=math.sqrt(rhs))
    else:
        orc.set_all_averages()
    orc.lf_init()
    orc.lf_evolve(nsteps)
    orc.lf_sync()
    phiT = torch.from_numpy(np.stack([orc.download(FLD_S, f) for f in range(2)]))
    piT = torch.from_numpy(np.stack([orc.download(PI_S, f) for f in range(2)]))
    bg = orc.background()
    bgT = torch.tensor([bg["aI"], bg["aDotI"]], dtype=torch.float64)
    return float(chi_energy_fraction_and_band(phiT, piT, bgT, torch.tensor(coeffs, dtype=torch.float64), s, DX,
                                              KIR, BAND))


def adjoint_grads(spec, coeffs, phi0, pi0, nsteps, precision=CLMI_F64, expansion=True, checkpoint_every=0):
    c = torch.tensor(coeffs, dtype=torch.float64, requires_grad=True)
    p = torch.from_numpy(phi0.copy()).requires_grad_(True)
    q = torch.from_numpy(pi0.copy()).requires_grad_(True)
    phiT, piT, bgT = leapfrog_run(c, p, q, spec, N, DX, DT, nsteps, precision, expansion, checkpoint_every)
    L = chi_energy_fraction_and_band(phiT, piT, bgT, c, spec, DX, KIR, BAND)
    L.backward()
    return float(L.detach()), c.grad.numpy(), p.grad.numpy(), q.grad.numpy()


@pytest.mark.parametrize("expansion", [True, False])
def test_adjoint_matches_central_differences(gpu, oracle_mod, expansion):
    say(f"case adjoint FD expansion={expansion}")
    spec = bench_spec()
    coeffs = [c for c, _ in spec.terms]
    rng = np.random.default_rng(3)
    phi0, pi0 = initial_fields(spec, rng)
    nsteps = 60
    L, gc, gphi, gpi = adjoint_grads(spec, coeffs, phi0, pi0, nsteps, expansion=expansion, checkpoint_every=8)
    L0 = oracle_loss(oracle_mod, spec, coeffs, phi0, pi0, nsteps, expansion)
    say("loss", L, L0)
    assert L == pytest.approx(L0, rel=1e-10)
    # directional derivatives along smooth random directions of phi0 and pi0
    for which in ("phi", "pi"):
        d = np.stack([smooth_random(N, rng, amp=1.0) for _ in range(2)])
        eps = 1e-4 * np.max(np.abs(phi0 if which == "phi" else pi0)) / np.max(np.abs(d))
        if which == "phi":
            lp = oracle_loss(oracle_mod, spec, coeffs, phi0 + eps * d, pi0, nsteps, expansion)
            lm = oracle_loss(oracle_mod, spec, coeffs, phi0 - eps * d, pi0, nsteps, expansion)
            ad = float(np.sum(gphi * d))
        else:
            lp = oracle_loss(oracle_mod, spec, coeffs, phi0, pi0 + eps * d, nsteps, expansion)
            lm = oracle_loss(oracle_mod, spec, coeffs, phi0, pi0 - eps * d, nsteps, expansion)
            ad = float(np.sum(gpi * d))
        fd = (lp - lm) / (2 * eps)
        say(which, ad, fd)
        assert ad == pytest.approx(fd, rel=1e-4)
    # couplings: every potential monomial coefficient (q enters through the q/2 phi^2 chi^2 term)
    for t in range(len(coeffs)):
        h = 1e-4 * abs(coeffs[t])
        cp, cm = list(coeffs), list(coeffs)
        cp[t] += h
        cm[t] -= h
        fd = (oracle_loss(oracle_mod, spec, cp, phi0, pi0, nsteps, expansion)
              - oracle_loss(oracle_mod, spec, cm, phi0, pi0, nsteps, expansion)) / (2 * h)
        say("coeff", t, gc[t], fd)
        assert gc[t] == pytest.approx(fd, rel=1e-4)
    # individual Fourier modes of the initial inflaton
    modes = np.array([[1, 0, 1], [2, -1, 3], [0, 3, 2]])
    mg = mode_gradients(gphi[0], modes)
    for (n, (gre, gim)) in zip(modes, mg):
        x = np.arange(N)
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        arg = 2 * np.pi * (n[0] * X + n[1] * Y + n[2] * Z) / N
        for g, shape in ((gre, 2.0 / N ** 3 * np.cos(arg)), (gim, -2.0 / N ** 3 * np.sin(arg))):
            eps = 1e-4 * np.max(np.abs(phi0[0])) / np.max(np.abs(shape))
            p_ = phi0.copy()
            p_[0] += eps * shape
            m_ = phi0.copy()
            m_[0] -= eps * shape
            fd = (oracle_loss(oracle_mod, spec, coeffs, p_, pi0, nsteps, expansion)
                  - oracle_loss(oracle_mod, spec, coeffs, m_, pi0, nsteps, expansion)) / (2 * eps)
            say("mode", n, g, fd)
            assert g == pytest.approx(fd, rel=1e-4, abs=1e-9 * abs(L))


def test_adjoint_checkpoint_spacing_does_not_change_gradients(gpu):
    say("case adjoint checkpoints")
    spec = bench_spec()
    coeffs = [c for c, _ in spec.terms]
    phi0, pi0 = initial_fields(spec, np.random.default_rng(5))
    res = [adjoint_grads(spec, coeffs, phi0, pi0, 23, checkpoint_every=m) for m in (1, 5, 0)]
    for r in res[1:]:
        assert r[0] == res[0][0]
        for a, b in zip(r[1:], res[0][1:]):
            assert np.array_equal(a, b)


def test_adjoint_leaves_final_state(gpu):
    say("case adjoint final state")
    spec = bench_spec()
    phi0, pi0 = initial_fields(spec, np.random.default_rng(9))
    lats = []
    for run_adjoint in (False, True):
        lat = Lattice(spec, N, DX, DT, CLMI_F64)
        for f in range(2):
            lat.upload(FLD_S, f, 0, phi0[f])
            lat.upload(PI_S, f, 0, pi0[f])
        if run_adjoint:
            z = [np.zeros((N,) * 3)] * 2
            lat.adjoint_lf(17, z, z, (0.0, 0.0), checkpoint_every=4)
        else:
            _, _, rhs = lat.hubble_constraint()
            lat.set_background(aDotI=math.sqrt(rhs))
            lat.lf_init()
            lat.lf_evolve(17)
            lat.lf_sync()
        lats.append(lat)
    for f in range(2):
        for k in (FLD_S, PI_S):
            assert np.array_equal(lats[0].download(k, f, 0), lats[1].download(k, f, 0))
    assert lats[0].background()["aI"] == lats[1].background()["aI"]
    for lat in lats:
        lat.close()


def test_adjoint_f32_bulk_tracks_fp64(gpu):
    say("case adjoint f32 vs f64")
    spec = bench_spec()
    coeffs = [c for c, _ in spec.terms]
    phi0, pi0 = initial_fields(spec, np.random.default_rng(11))
    L64, gc64, gp64, gq64 = adjoint_grads(spec, coeffs, phi0, pi0, 80, CLMI_F64)
    L32, gc32, gp32, gq32 = adjoint_grads(spec, coeffs, phi0, pi0, 80, CLMI_F32_BULK)
    say("loss", L64, L32, "coeff grads", gc64, gc32)
    assert L32 == pytest.approx(L64, rel=1e-5)
    np.testing.assert_allclose(gc32, gc64, rtol=1e-4)
    for a, b in ((gp32, gp64), (gq32, gq64)):
        assert np.max(np.abs(a - b)) < 1e-4 * np.max(np.abs(b))


@pytest.mark.parametrize("expansion,model", [(True, "lphi4"), (False, "lphi4"), (True, "runtime"), (True, "lphi4_1f")])
def test_adjoint_matches_fp64_restatement(gpu, oracle_mod, expansion, model):
    """The HIP adjoint (checkpointed) equals the numpy FP64 reverse-mode restatement on the same
    seeds: adjoint fields and coefficient gradients to 1e-9 relative.  The models cover every form of
    the adjoint kernels: two-field and one-field lphi4 (compile-time exponent table) and a two-field
    potential that is not lphi4's (runtime table, four terms)."""
    from test_gpu_parity import runtime_poly_spec

    say(f"case adjoint vs restatement expansion={expansion} model={model}")
    spec = {"lphi4": bench_spec, "runtime": runtime_poly_spec, "lphi4_1f": lambda: bench_spec(0.0, n_fields=1)}[model]()
    nf = spec.n_scalars
    rng = np.random.default_rng(13)
    phi0, pi0 = initial_fields(spec, rng)
    nsteps = 30
    sf = [rng.normal(size=(N,) * 3) for _ in range(nf)]
    sp = [rng.normal(size=(N,) * 3) for _ in range(nf)]
    sb = (0.3, -0.2) if expansion else (0.0, 0.0)
    lat = Lattice(spec, N, DX, DT, CLMI_F64, expansion=expansion)
    for f in range(nf):
        lat.upload(FLD_S, f, 0, phi0[f])
        lat.upload(PI_S, f, 0, pi0[f])
    gf, gp, gc, _ = lat.adjoint_lf(nsteps, sf, sp, sb, checkpoint_every=7)
    lat.close()
    rf, rp, rc = oracle_mod.adjoint_lf(spec, N, DX, DT, phi0, pi0, nsteps, sf, sp, sb, expansion)
    for a, b in list(zip(gf, rf)) + list(zip(gp, rp)):
        e = np.max(np.abs(a - b)) / np.max(np.abs(b))
        say("field", e)
        assert e < 1e-9
    np.testing.assert_allclose(gc, rc, rtol=1e-9)


@pytest.mark.parametrize("prec", [CLMI_F64, CLMI_F32_BULK])
def test_tiled_kick_transpose_matches_one_site_kernel(gpu, prec):
    """At 64^3 (lphi4, two fields) the reverse sweep's kick transpose runs tiled (x-march, LDS tile);
    with CLMI_FLAG_NO_TILED it runs one site per thread, and the forward runs the generic pass.  The
    per-site arithmetic is the same, so the gradients agree up to the summation order of the volume
    sums: 1e-10 relative in FP64; 1e-5 with the FP32-bulk forward, whose stored fields follow the
    background and can differ by a rounding."""
    from cosmolattice_amd.clmi import CLMI_FLAG_NO_TILED

    say(f"case adjoint tiled kick prec={prec}")
    spec = bench_spec()
    n, kir = 64, 0.1
    dx = 2 * math.pi / kir / n
    rng = np.random.default_rng(41)
    phi0 = [spec.fld0[f] + smooth_random(n, rng, amp=0.05) + 0.01 * rng.normal(size=(n,) * 3) for f in range(2)]
    pi0 = [spec.pi0[f] + smooth_random(n, rng, amp=0.05) + 0.01 * rng.normal(size=(n,) * 3) for f in range(2)]
    seed_f = [rng.normal(size=(n,) * 3) for _ in range(2)]
    seed_p = [rng.normal(size=(n,) * 3) for _ in range(2)]
    res = []
    for flags in (0, CLMI_FLAG_NO_TILED):
        lat = Lattice(spec, n, dx, DT, prec, flags=flags)
        plan = lat.plan()
        say("plan", plan.replace("\n", "; "))
        assert ("k_adj_kick_tiled" in plan) == (flags == 0), plan
        for f in range(2):
            lat.upload(FLD_S, f, 0, phi0[f])
            lat.upload(PI_S, f, 0, pi0[f])
        res.append(lat.adjoint_lf(30, seed_f, seed_p, (0.3, 0.1), checkpoint_every=6))
        lat.close()
    tol = 1e-10 if prec == CLMI_F64 else 1e-5
    for key in (0, 1):
        for a, b in zip(res[0][key], res[1][key]):
            assert np.max(np.abs(a - b)) <= tol * np.max(np.abs(b)), key
    np.testing.assert_allclose(res[0][2], res[1][2], rtol=tol)
    for k, v in res[1][3].items():
        assert res[0][3][k] == pytest.approx(v, rel=tol, abs=1e-300), k
