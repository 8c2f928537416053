"""Throughput benchmark: 2-field lphi4 preheating on a 256^3 lattice, leapfrog, FP32-bulk vs FP64.

Workload (one MI355X): inflaton phi (V = phi^4/4) plus daughter chi coupled by q/2 phi^2 chi^2 with
q = 100, program units lambda = 1, phi0 = 1, pi0 = -0.5, fStar/MPl = 2.34 (so the self-consistent
expansion is live), periodic 256^3, kIR = 0.05, dt = 0.005, seeded vacuum Gaussian fluctuations
on both fields (cut at k = N kIR / 4).  warmup + steps = 4000 leapfrog steps by default (t = 20).

A step is one leapfrog step of the whole lattice including the volume reductions and the
self-consistent scale-factor update (one fused lattice pass + one background-update launch).
site-updates/s = N^3 * steps / seconds of the timed loop (inputs resident in HBM).

Printed: ONE JSON line (rank 0).  value = FP32-bulk site-updates/s; the FP64 drop-in mode, the
accuracy of FP32-bulk against FP64 (Friedmann-constraint violation, band-averaged spectrum error),
the roofline of the dominant kernel (HIP-event timed, algorithmic bytes) and the CPU baseline
(the FP64 oracle, reference-structured, on this host's cores) are reported beside it.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "site-updates/s + %HBM peak, lphi4 256^3: FP32-bulk vs FP64 at stated energy error"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_lattice(clmi, spec, N, kIR, dt, precision, seed, flags=0, round_pi_f32=False):
    dx = 2 * math.pi / kIR / N
    lat = clmi.Lattice(spec, N, dx, dt, precision, expansion=True, flags=flags)
    # ScaleFactorInitializer (initializers/scalefactorinitializer.h:49-64): aDot from the
    # homogeneous energy, then the fluctuations are added with that aDot
    phi0 = spec.fld0
    pi0 = spec.pi0
    kin = sum(0.5 * p * p for p in pi0)
    rho_hom = kin + spec.potential(phi0)
    aDot0 = spec.fStar / spec.MPl * math.sqrt(rho_hom / 3.0)
    lat.set_background(aI=1.0, aSI=1.0, aIM=1.0, aDotI=aDot0)
    m2 = spec.mass2(phi0)
    kcut = N * kIR / 4
    for f in range(spec.n_scalars):
        lat.grf_scalar(f, seed, f, m2[f], aDot0, kIR, kcut, phi0[f], pi0[f])
        if round_pi_f32:  # the initial state as FP32-bulk stores it (momenta in FP32)
            lat.upload(clmi.PI_S, f, 0, lat.download(clmi.PI_S, f, 0).astype(np.float32).astype(np.float64))
    # ModelInitializer (initializers/modelinitializer.h:71-80): aDotI = sqrt(Friedmann RHS)
    _, _, rhs = lat.hubble_constraint()
    lat.set_background(aDotI=math.sqrt(rhs))
    lat.lf_init()
    return lat, dx


# HIP events time one drift + kick pass in prof_every(steps): timing every pass adds two event records per
# step, about 7 us at 256^3, to the throughput being measured.  Short runs time every pass, so that the
# kernel's average still rests on at least 64 launches where the run has them.
PROF_EVERY = 16


def prof_every(steps):
    return max(1, min(PROF_EVERY, steps // 64))


def timed_run(lat, warmup, steps):
    lat.lf_evolve(warmup)
    lat.synchronize()
    lat.profile_enable(True, every=prof_every(steps))
    t0 = time.perf_counter()
    lat.lf_evolve(steps)
    lat.synchronize()
    t1 = time.perf_counter()
    ms, launches, bytes_per = lat.profile_read()
    lat.profile_enable(False)
    return t1 - t0, ms, launches, bytes_per


def horizon_run(clmi, spec, N, kIR, dt, seed, precision, steps, every, ic=None):
    """One run of the workload over `steps` leapfrog steps from its seeded initial state, with a sync and a
    Friedmann check every `every` steps.  ic: None; "pi_f32" rounds the initial momenta to FP32 (what
    FP32-bulk storage does to the initial state); "rel_2^-24" multiplies every initial field and momentum
    value by (1 + 2^-24 u), u uniform in [-1, 1] (seeded).  Returns the Friedmann records, the PS_type 1
    spectra of both fields at the end and the final background."""
    from cosmolattice_amd.clmi import FLD_S, PI_S, power_spectrum

    dx = 2 * math.pi / kIR / N
    lat = clmi.Lattice(spec, N, dx, dt, precision, expansion=True)
    kin = sum(0.5 * p * p for p in spec.pi0)
    aDot0 = spec.fStar / spec.MPl * math.sqrt((kin + spec.potential(spec.fld0)) / 3.0)
    lat.set_background(aI=1.0, aSI=1.0, aIM=1.0, aDotI=aDot0)
    m2 = spec.mass2(spec.fld0)
    rng = np.random.default_rng(seed + 17)
    for f in range(spec.n_scalars):
        lat.grf_scalar(f, seed, f, m2[f], aDot0, kIR, N * kIR / 4, spec.fld0[f], spec.pi0[f])
        if ic == "pi_f32":
            lat.upload(PI_S, f, 0, lat.download(PI_S, f, 0).astype(np.float32).astype(np.float64))
        elif ic == "rel_2^-24":
            for k in (FLD_S, PI_S):
                a = lat.download(k, f, 0)
                lat.upload(k, f, 0, a * (1.0 + 2.0 ** -24 * rng.uniform(-1.0, 1.0, size=a.shape)))
    _, _, rhs = lat.hubble_constraint()
    lat.set_background(aDotI=math.sqrt(rhs))
    lat.lf_init()
    fried = []
    for _ in range(steps // every):
        lat.lf_evolve(every)
        lat.lf_sync()
        fried.append(lat.hubble_constraint()[0])
    spectra = []
    for f in range(spec.n_scalars):
        s_, c_, m_ = lat.spectrum_bins(FLD_S, f, 0, 1.0)
        spectra.append(power_spectrum(s_, c_, m_, N, kIR, 1.0, 1, 1)[1])
    bg = lat.background()
    lat.close()
    return np.array(fried), spectra, bg


def band_errors(pa, pb, nband=8):
    """Relative errors of a spectrum summed over `nband` equal bands of bins (bin 0 left out)."""
    pa, pb = pa[1:], pb[1:]
    edges = np.linspace(0, len(pb), nband + 1).astype(int)
    return np.array([abs(pa[lo:hi].sum() - pb[lo:hi].sum()) / pb[lo:hi].sum()
                     for lo, hi in zip(edges[:-1], edges[1:]) if pb[lo:hi].sum() > 0])


def accuracy_horizon(clmi, spec, N, kIR, dt, seed, steps=4000, every=100):
    """FP32-bulk against FP64 over the workload's full horizon (independent of the timed step count), next
    to the chaotic floor: FP64 runs whose initial state differs at FP32 rounding level."""
    log(f"[bench] accuracy: {N}^3, {steps} steps in FP64, FP32-bulk and two perturbed FP64 runs")
    runs = {"f64": (clmi.CLMI_F64, None), "f32_bulk": (clmi.CLMI_F32_BULK, None),
            "f64_pi_rounded_to_f32": (clmi.CLMI_F64, "pi_f32"), "f64_perturbed_rel_2^-24": (clmi.CLMI_F64, "rel_2^-24")}
    res = {k: horizon_run(clmi, spec, N, kIR, dt, seed, p, steps, every, ic) for k, (p, ic) in runs.items()}
    f64 = res["f64"]
    out = {"steps": steps, "t_end": steps * dt, "friedmann_check_every": every,
           "max_abs_friedmann": {k: float(np.max(np.abs(r[0]))) for k, r in res.items()},
           "end_friedmann": {k: float(r[0][-1]) for k, r in res.items()},
           "band_spectrum_rel_err_vs_f64": {}, "a_rel_diff_vs_f64": {}}
    for k, r in res.items():
        if k == "f64":
            continue
        d = {}
        for f, name in ((0, "phi"), (1, "chi")):
            e = band_errors(r[1][f], f64[1][f])
            d[name] = {"mean": float(e.mean()), "max": float(e.max())}
        out["band_spectrum_rel_err_vs_f64"][k] = d
        out["a_rel_diff_vs_f64"][k] = abs(r[2]["aI"] - f64[2]["aI"]) / f64[2]["aI"]
    e32 = out["band_spectrum_rel_err_vs_f64"]["f32_bulk"]["chi"]["mean"]
    efl = out["band_spectrum_rel_err_vs_f64"]["f64_pi_rounded_to_f32"]["chi"]["mean"]
    out["chi_band_err_f32_bulk_over_floor"] = e32 / efl if efl > 0 else None
    log(f"[bench] accuracy: chi band error f32_bulk {e32:.3e}, floor (pi rounded to f32) {efl:.3e}")
    return out


def cpu_baseline(lat64, spec, N, dx, dt, budget_s):
    """FP64 oracle (reference-structured: one lattice sweep per reference expression) timed on this
    host's cores on a bounded sample of the same workload, starting from the GPU run's state."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    from cosmolattice_amd.clmi import FLD_S, PI_S

    ref = orc.OracleLattice(spec, N, dx, dt, expansion=True)
    lat64.lf_sync()
    for f in range(spec.n_scalars):
        ref.upload(FLD_S, f, 0, lat64.download(FLD_S, f, 0))
        ref.upload(PI_S, f, 0, lat64.download(PI_S, f, 0))
    ref.set_background(**lat64.background())
    ref.lf_init()
    ref.lf_evolve(1)  # first (half-kick) step outside the timing
    # median over timed blocks of 20 steps (at least 5 blocks), within the time budget
    block, rates = 20, []
    t_start = time.perf_counter()
    while len(rates) < 5 or time.perf_counter() - t_start < budget_s:
        t0 = time.perf_counter()
        ref.lf_evolve(block)
        rates.append(N ** 3 * block / (time.perf_counter() - t0))
        if len(rates) >= 100:
            break
    el = time.perf_counter() - t_start
    return {
        "value": float(np.median(rates)),
        "unit": "site-updates/s",
        "cores": orc.num_threads(),
        "kind": "port",
        "sample": f"median of {len(rates)} blocks of {block} leapfrog steps of the same {N}^3 2-field lphi4 lattice "
                  f"(FP64, reference-structured C/OpenMP oracle: 7 lattice sweeps per step), {el:.1f} s in all; "
                  f"block range {min(rates) / 1e9:.3f}-{max(rates) / 1e9:.3f} G site-updates/s",
    }


def c1_bench(clmi, steps=1000, N=64, kIR=0.1, dt=0.01, seed=11):
    """Single-field lphi4 64^3 with velocity Verlet (VV2), FP64, self-consistent expansion, seeded vacuum
    fluctuations: the whole run on the GPU, and the same run on the FP64 oracle on this host's cores."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    from cosmolattice_amd.clmi import FLD_S, PI_S

    spec = clmi.lphi4(0.0, n_fields=1)
    log(f"[bench] c1: lphi4 1 field {N}^3 VV2, {steps} steps, GPU and oracle")
    lat, dx = make_lattice(clmi, spec, N, kIR, dt, clmi.CLMI_F64, seed)
    ref = orc.OracleLattice(spec, N, dx, dt, expansion=True)
    ref.upload(FLD_S, 0, 0, lat.download(FLD_S, 0, 0))
    ref.upload(PI_S, 0, 0, lat.download(PI_S, 0, 0))
    ref.set_background(**lat.background())
    lat.synchronize()
    t0 = time.perf_counter()
    lat.vv_evolve([1.0], steps)
    lat.synchronize()
    gpu_s = time.perf_counter() - t0
    fr_gpu = lat.hubble_constraint()[0]
    t0 = time.perf_counter()
    ref.vv_evolve([1.0], steps)
    cpu_s = time.perf_counter() - t0
    fr_cpu = ref.hubble_constraint()[0]
    err = max(float(np.max(np.abs(lat.download(k, 0, 0) - ref.download(k, 0))) / np.max(np.abs(ref.download(k, 0))))
              for k in (FLD_S, PI_S))
    out = {"config": f"lphi4 1 field {N}^3, kIR={kIR}, VV2, dt={dt}, {steps} steps, FP64, expansion",
           "gpu_site_updates_per_s": N ** 3 * steps / gpu_s, "gpu_seconds": gpu_s,
           "cpu_site_updates_per_s": N ** 3 * steps / cpu_s, "cpu_seconds": cpu_s, "cpu_cores": orc.num_threads(),
           "max_rel_diff_gpu_vs_oracle": err, "friedmann_gpu": fr_gpu, "friedmann_oracle": fr_cpu,
           "aI_end": lat.background()["aI"]}
    lat.close()
    return out


def adjoint_bench(clmi, N=128, kIR=0.05, dt=0.005, steps=2000, check_steps=(500, 1000), seed=5000, n_modes=256):
    """Adjoint sensitivities of the 2-field lphi4 128^3 run: loss = final energy fraction of chi + chi
    power in k in [0.5, 1.5]; dL/dq and dL/d(initial phi mode) for 256 seeded modes from the
    reverse-mode leapfrog (sqrt(T) checkpoints), FP64 and FP32-bulk forward.

    Timing is the full 2000-step run.  The gradients are checked against central differences of FP64
    forward runs at 500 and 1000 steps, where the differences converge as the step shrinks (reported for
    relative steps 1e-4 ... 1e-7).  At 2000 steps (t = 10) the parametric resonance amplifies every
    perturbation so strongly that no step size resolves the derivative: the differences at relative
    steps 1e-4, 1e-6 and 1e-8 are reported to show that, not as a check."""
    import torch
    from cosmolattice_amd.adjoint import chi_energy_fraction_and_band, leapfrog_run, mode_gradients
    from cosmolattice_amd.clmi import FLD_S, PI_S

    spec = clmi.lphi4(100.0, n_fields=2, phi0_gev=5.6964e18, lam=1.0, pi0_program=-0.5)
    log(f"[bench] adjoint: {N}^3, {steps} steps")
    lat, dx = make_lattice(clmi, spec, N, kIR, dt, clmi.CLMI_F64, seed)
    phi0 = np.stack([lat.download(FLD_S, f, 0) for f in range(2)])
    pi0 = np.stack([lat.download(PI_S, f, 0) for f in range(2)])
    lat.close()
    coeffs = [c for c, _ in spec.terms]
    band = (0.5, 1.5)

    def run(prec, cvals, nsteps, p0=phi0, grads=True, q0=pi0):
        c = torch.tensor(cvals, dtype=torch.float64, requires_grad=grads)
        p = torch.from_numpy(p0.copy()).requires_grad_(grads)
        q = torch.from_numpy(q0.copy()).requires_grad_(grads)
        t0 = time.perf_counter()
        phiT, piT, bgT = leapfrog_run(c, p, q, spec, N, dx, dt, nsteps, prec, True, 0)
        L = chi_energy_fraction_and_band(phiT, piT, bgT, c, spec, dx, kIR, band)
        if not grads:
            return float(L.detach()), time.perf_counter() - t0
        L.backward()
        return float(L.detach()), time.perf_counter() - t0, c.grad.numpy(), p.grad.numpy()

    def fd_q(nsteps, rel):
        # q enters as the coefficient q/2 of phi^2 chi^2
        h = rel * abs(coeffs[1])
        cp, cm = list(coeffs), list(coeffs)
        cp[1] += h
        cm[1] -= h
        return 0.5 * (run(clmi.CLMI_F64, cp, nsteps, grads=False)[0] -
                      run(clmi.CLMI_F64, cm, nsteps, grads=False)[0]) / (2 * h)

    L32, t32, gc32, gp32 = run(clmi.CLMI_F32_BULK, coeffs, steps)
    L64, t64, gc64, gp64 = run(clmi.CLMI_F64, coeffs, steps)
    rng = np.random.default_rng(seed)
    kc = N / 4  # modes below the initial-condition cutoff |n| kIR < N kIR / 4
    modes = []
    while len(modes) < n_modes:
        n = rng.integers(-int(kc), int(kc) + 1, size=3)
        if 0 < n[2] < N // 2 and np.linalg.norm(n) < kc and not any(np.array_equal(n, m) for m in modes):
            modes.append(n)

    checks = []
    x = np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    prng = np.random.default_rng(seed + 17)
    phi0_p = phi0 * (1.0 + 2.0 ** -24 * prng.uniform(-1.0, 1.0, size=phi0.shape))
    pi0_p = pi0 * (1.0 + 2.0 ** -24 * prng.uniform(-1.0, 1.0, size=pi0.shape))
    for ns in check_steps:
        _, _, gc, gp = run(clmi.CLMI_F64, coeffs, ns)
        _, _, gc_32, gp_32 = run(clmi.CLMI_F32_BULK, coeffs, ns)
        # chaotic floor: the FP64 adjoint from initial fields perturbed at FP32 rounding level
        _, _, gc_p, gp_p = run(clmi.CLMI_F64, coeffs, ns, p0=phi0_p, q0=pi0_p)
        # A central difference errs by truncation at large steps and by rounding at small ones; the
        # adjoint is checked against the best of the sequence (the step is reported with it).
        def best(adj, fd):
            errs = {k: abs(adj - v) / abs(v) for k, v in fd.items()}
            k = min(errs, key=errs.get)
            return errs[k], float(k)

        fd = {f"{rel:g}": fd_q(ns, rel) for rel in (1e-4, 1e-5, 1e-6, 1e-7)}
        e64, s64 = best(0.5 * gc[1], fd)
        e32, s32 = best(0.5 * gc_32[1], fd)
        mg64, mg32 = mode_gradients(gp[0], modes), mode_gradients(gp_32[0], modes)
        fd_modes = []
        for j in range(2):  # real parts of two modes
            shape = 2.0 / N ** 3 * np.cos(2 * np.pi * (modes[j][0] * X + modes[j][1] * Y + modes[j][2] * Z) / N)
            fdm = {}
            for rel in (1e-6, 1e-8):
                eps = rel * np.max(np.abs(phi0[0])) / np.max(np.abs(shape))
                pp, pm = phi0.copy(), phi0.copy()
                pp[0] += eps * shape
                pm[0] -= eps * shape
                fdm[f"{rel:g}"] = (run(clmi.CLMI_F64, coeffs, ns, pp, grads=False)[0] -
                                   run(clmi.CLMI_F64, coeffs, ns, pm, grads=False)[0]) / (2 * eps)
            em, sm = best(float(mg64[j][0]), fdm)
            fd_modes.append({"mode": [int(v) for v in modes[j]], "adjoint_f64": float(mg64[j][0]),
                             "central_difference_f64_by_rel_step": fdm, "best_rel_err": em, "at_rel_step": sm})
        mgp = mode_gradients(gp_p[0], modes)
        checks.append({"steps": ns,
                       "dL_dq_rel_gap_f32_bulk_vs_f64": abs(gc_32[1] - gc[1]) / abs(gc[1]),
                       "dL_dq_rel_gap_floor_f64_perturbed_rel_2^-24": abs(gc_p[1] - gc[1]) / abs(gc[1]),
                       "modes_max_rel_diff_floor_f64_perturbed": float(np.max(np.abs(mgp - mg64)) / np.max(np.abs(mg64))),
                       "dL_dq": {"adjoint_f64": 0.5 * gc[1], "adjoint_f32_bulk": 0.5 * gc_32[1],
                                 "central_difference_f64_by_rel_step": fd,
                                 "best_rel_err_f64": e64, "at_rel_step_f64": s64,
                                 "best_rel_err_f32_bulk": e32, "at_rel_step_f32_bulk": s32},
                       "mode_checks": fd_modes,
                       "modes_max_rel_diff_f32_vs_f64": float(np.max(np.abs(mg32 - mg64)) / np.max(np.abs(mg64)))})
        log(f"[bench] adjoint check {ns} steps: dL/dq best rel err {e64:.2e} at rel step {s64:g}")
    long_fd = {f"{rel:g}": fd_q(steps, rel) for rel in (1e-4, 1e-6, 1e-8)}
    # the library call alone (no host-side loss, transfers or torch), on a prepared lattice
    call = {}
    for name, prec in (("f32_bulk", clmi.CLMI_F32_BULK), ("f64", clmi.CLMI_F64)):
        lat = clmi.Lattice(spec, N, dx, dt, prec, expansion=True)
        for f in range(2):
            lat.upload(FLD_S, f, 0, phi0[f])
            lat.upload(PI_S, f, 0, pi0[f])
        seeds = [np.zeros((N, N, N)), np.full((N, N, N), 1e-3)]
        lat.synchronize()
        t0 = time.perf_counter()
        lat.adjoint_lf(steps, seeds, [np.zeros((N, N, N))] * 2, (0.0, 0.0))
        call[name] = time.perf_counter() - t0
        lat.close()
    log(f"[bench] adjoint call alone: {call}")
    return {"config": f"lphi4 2 fields {N}^3, q=100, kIR={kIR}, dt={dt}, {steps} leapfrog steps, expansion; "
                      f"loss = E_chi/E_tot + chi power in k in [{band[0]}, {band[1]}]",
            "forward_plus_adjoint_seconds": {"f32_bulk": t32, "f64": t64},
            "forward_plus_adjoint_seconds_note": "torch round trip: two lattice set-ups, host transfers, the loss "
                                                 "(CPU FFT) and its backward, plus clmi_adjoint_lf",
            "adjoint_call_seconds": call,
            "adjoint_call_note": "clmi_adjoint_lf alone on a prepared lattice: forward with checkpoints, per-segment "
                                 "recomputation, reverse sweep (host FP64 seeds in, gradients out)",
            "loss": {"f32_bulk": L32, "f64": L64},
            "dL_dq": {"f32_bulk": 0.5 * gc32[1], "f64": 0.5 * gc64[1]},
            "gradient_checks": checks,
            "long_horizon_central_differences_f64": {
                "steps": steps, "dL_dq_by_rel_step": long_fd,
                "note": "differences do not converge with the step at this horizon (exponential instability); "
                        "the gradient is checked at the shorter horizons above"}}


TRAFFIC_FILE = "profiles/k_scalar_rows_lphi4_2f_256_traffic.md"


def measured_traffic(kernel: str, f32: bool):
    """HBM bytes per launch of the fused pass with the chosen plan, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of the same workload (calibrated against a copy kernel), or None
    when that summary has no row for this kernel instantiation."""
    import re
    path = Path(__file__).resolve().parent / TRAFFIC_FILE
    m = re.search(r"VZ=(\d+) WPR=(\d+) TY=(\d+) XC=\d+ MM=(\d+) PK=(\d+)", kernel)
    t = re.search(r"k_scalar_tiled XC=\d+ NT=(\d+) PK=(\d+)", kernel)
    if not (m or t) or not path.exists():
        return None
    fl = "true" if f32 else "false"
    if m:
        vz, wpr, ty, mm, pk = m.groups()
        name = f"`k_scalar_rows<2, {fl}, {vz}, {wpr}, {ty}, true, true, true, {mm}, {pk}>`"
    else:
        nt, pk = t.groups()
        name = f"`k_scalar_tiled<2, {fl}, true, true, true, 4, {'true' if nt == '1' else 'false'}, {pk}>`"
    for ln in path.read_text().splitlines():
        cells = [c.strip() for c in ln.strip().strip("|").split("|")]
        if len(cells) == 4 and cells[0] == name:
            return (float(cells[2]) + float(cells[3])) * 2 ** 20
    return None


def u1_bench(clmi, steps, N=256, kIR=0.05, dt=0.005, seed=9000, every=250):
    """Charged scalar + U(1) (lphi4U1, e^2/lambda = 10, program units lambda = 1) on 256^3 with the
    Gauss-consistent vacuum initial conditions and self-consistent expansion, `steps` leapfrog steps in
    both precision modes from the same initial state: site-updates/s (the evolve calls only; the checks
    every `every` steps are outside the timed region) and HBM fraction of the fused pass, and at every
    check the Friedmann (energy-constraint) violation (hubbleconstraint.h:27-36) and the Gauss violation
    (gausslaws.h:31-53), with FP32-bulk's differences from FP64."""
    from cosmolattice_amd.clmi import lphi4U1

    spec = lphi4U1(10.0, phi0_gev=5.6964e18, lam=1.0, pi0_gev2=0.0)
    dx = 2 * math.pi / kIR / N
    out = {"config": f"lphi4U1 {N}^3 leapfrog, e^2/lambda = 10, kIR = {kIR}, dt = {dt}, self-consistent expansion",
           "steps": steps, "check_every": every}
    recs = {}
    for name, prec in (("f64", clmi.CLMI_F64), ("f32_bulk", clmi.CLMI_F32_BULK)):
        lat = clmi.Lattice(spec, N, dx, dt, prec, expansion=True)
        lat.set_background(aI=1.0, aSI=1.0, aIM=1.0, aDotI=0.0)
        lat.grf_u1(seed, 1, spec.cs_mass2(spec.fld0), 0.0, kIR, N * kIR / 4, spec.fld0, spec.pi0)
        _, _, rhs = lat.hubble_constraint()
        lat.set_background(aDotI=math.sqrt(rhs))
        lat.lf_init()
        tune = lat.autotune(3)  # x-chunk of the charged-scalar + U(1) pass; results do not depend on it
        sec, kms, launches = 0.0, 0.0, 0
        fried, gauss = [], []
        done = 0
        while done < steps:
            n = min(every, steps - done)
            s_, k_, l_, bpl = timed_run(lat, 0, n)
            sec, kms, launches, done = sec + s_, kms + k_, launches + l_, done + n
            lat.lf_sync()
            fried.append(lat.hubble_constraint()[0])
            gauss.append(lat.gauss_u1()[0])
        avg_s = kms / 1e3 / max(launches, 1)
        recs[name] = (np.array(fried), np.array(gauss), lat.background()["aI"])
        out[name] = {"site_updates_per_s": N ** 3 * steps / sec, "ms_per_step": 1e3 * sec / steps,
                     "avg_launch_us": avg_s * 1e6, "bytes_per_launch": bpl,
                     "hbm_frac": bpl / avg_s / 1e9 / HBM_PEAK_GBS,
                     "max_abs_friedmann": float(np.max(np.abs(fried))), "end_friedmann": float(fried[-1]),
                     "max_gauss_violation": float(np.max(gauss)), "end_gauss_violation": float(gauss[-1]),
                     "aI_end": recs[name][2],
                     "kernel": next((ln[len("# kernel "):] for ln in tune.splitlines() if ln.startswith("# kernel ")),
                                    "generic")}
        log(f"[bench] u1 {name}: {out[name]['ms_per_step']:.4f} ms/step, frac {out[name]['hbm_frac']:.3f}, "
            f"max |Friedmann| {out[name]['max_abs_friedmann']:.3e}, max Gauss {out[name]['max_gauss_violation']:.3e}")
        lat.close()
    (f64, g64, a64), (f32, g32, a32) = recs["f64"], recs["f32_bulk"]
    out["f32_bulk_vs_f64"] = {"max_abs_friedmann_diff": float(np.max(np.abs(f32 - f64))),
                              "aI_rel_diff": abs(a32 - a64) / a64}
    return out


def ensemble_bench(clmi, spec, steps, kIR=0.05, dt=0.005, seed=7000):
    """Batched ensemble: 16 realisations each at 64^3, 128^3 and 192^3 (distinct seeds), FP32-bulk,
    advanced together (one batched launch per step, graph-replayed) versus the same lattices advanced
    one at a time; plus the ensemble-mean chi spectrum of the 64^3 members against the same
    realisations in FP64."""
    from cosmolattice_amd.clmi import FLD_S, Ensemble, power_spectrum

    sizes = [64] * 16 + [128] * 16 + [192] * 16
    log(f"[bench] ensemble: {len(sizes)} lattices, {steps} steps")
    members = []
    for i, N in enumerate(sizes):
        lat, _ = make_lattice(clmi, spec, N, kIR, dt, clmi.CLMI_F32_BULK, seed + i)
        members.append(lat)
    sites = sum(N ** 3 for N in sizes)
    # default: one group per size (three parallel graph branches); split: groups capped at 128 MiB
    timings = {}
    tune = {}
    for label, gb in (("split_groups", 128 << 20), ("default", 0)):
        ens = Ensemble(members, group_bytes=gb)
        if label == "default":
            tune["ensemble"] = ens.autotune(3)
            log(f"[bench] ensemble autotune\n{tune['ensemble']}")
        ens.lf_evolve(2)  # first step of each member alone, then the graphs are captured
        t0 = time.perf_counter()
        ens.lf_evolve(steps)
        timings[label] = (time.perf_counter() - t0, ens.groups)
        ens.close()
    t_ens, ngroups = timings["default"]
    # one lattice at a time, each with its own single-lattice autotuned plan
    for N in (64, 128, 192):
        m = next(m for m in members if m.N == N)
        tune[f"single_{N}"] = m.autotune(2).splitlines()[0]
    for m in members:
        m.autotune(2)
        m.synchronize()
    t0 = time.perf_counter()
    for m in members:
        m.lf_evolve(steps)
    for m in members:
        m.synchronize()
    t_one = time.perf_counter() - t0
    # ensemble-mean chi spectrum of the 64^3 members vs FP64 twins, next to the chaotic floor: FP64 twins
    # whose initial momenta were rounded to FP32
    nspec = 16
    twins, floor = [], []
    for i in range(nspec):
        for lst, rnd in ((twins, False), (floor, True)):
            lat, _ = make_lattice(clmi, spec, 64, kIR, dt, clmi.CLMI_F64, seed + i, round_pi_f32=rnd)
            lat.lf_evolve(4 + 3 * steps)
            lst.append(lat)
    def mean_ps(lats):
        acc = None
        for lat in lats:
            lat.lf_sync()
            s_, c_, m_ = lat.spectrum_bins(FLD_S, 1, 0, 1.0)
            _, ps = power_spectrum(s_, c_, m_, 64, kIR, 1.0, 1, 1)
            acc = ps if acc is None else acc + ps
        return acc / len(lats)
    p32, p64, pfl = mean_ps(members[:nspec]), mean_ps(twins), mean_ps(floor)
    band = np.abs(p32[1:] - p64[1:]) / np.maximum(np.abs(p64[1:]), 1e-300)
    band_fl = np.abs(pfl[1:] - p64[1:]) / np.maximum(np.abs(p64[1:]), 1e-300)
    b8, b8fl = band_errors(p32, p64), band_errors(pfl, p64)
    for lat in members + twins + floor:
        lat.close()
    return {
        "config": "48 lphi4 2-field realisations: 16 x 64^3, 16 x 128^3, 16 x 192^3, FP32-bulk, leapfrog",
        "steps": steps,
        "site_updates": sites,
        "ensemble_site_updates_per_s": sites * steps / t_ens,
        "groups": ngroups,
        "split_groups_site_updates_per_s": sites * steps / timings["split_groups"][0],
        "split_groups": timings["split_groups"][1],
        "one_at_a_time_site_updates_per_s": sites * steps / t_one,
        "autotune": tune,
        "speedup": t_one / t_ens,
        "ms_per_step": 1e3 * t_ens / steps,
        "spectrum_steps": 4 + 3 * steps,
        "mean_chi_spectrum_rel_err_vs_fp64": {"per_bin_mean": float(band.mean()), "per_bin_max": float(band.max()),
                                              "band8_mean": float(b8.mean()), "band8_max": float(b8.max())},
        "floor_fp64_pi_rounded_to_f32": {"per_bin_mean": float(band_fl.mean()), "per_bin_max": float(band_fl.max()),
                                         "band8_mean": float(b8fl.mean()), "band8_max": float(b8fl.max())},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3900)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--seed", type=int, default=20260101)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--flags", type=int, default=0, help="clmi_config.flags for both lattices")
    ap.add_argument("--accuracy-steps", type=int, default=4000,
                    help="leapfrog steps of the FP32-bulk vs FP64 accuracy runs (4000: t = 20)")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=2000)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--u1-steps", type=int, default=2000)
    ap.add_argument("--no-u1", action="store_true")
    ap.add_argument("--no-c1", action="store_true", help="skip the 64^3 VV2 GPU-vs-oracle run")
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB device copy (torch) so HBM counters can be scaled")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library runs each lattice on a single MI355X "
              "(sharding a lattice over several GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)

    from cosmolattice_amd import clmi

    N, kIR, dt, q = args.N, 0.05, 0.005, 100.0
    spec = clmi.lphi4(q, n_fields=2, phi0_gev=5.6964e18, lam=1.0, pi0_program=-0.5)

    res = {}
    lats = {}
    for name, prec in (("f64", clmi.CLMI_F64), ("f32_bulk", clmi.CLMI_F32_BULK)):
        log(f"[bench] {name}: init {N}^3")
        lat, dx = make_lattice(clmi, spec, N, kIR, dt, prec, args.seed, args.flags)
        tune = lat.autotune(5)
        log(f"[bench] {name}: autotune\n{tune}")
        log(f"[bench] {name}: warmup {args.warmup}, timed {args.steps} steps")
        sec, kms, launches, bpl = timed_run(lat, args.warmup, args.steps)
        res[name] = {"seconds": sec, "kernel_ms": kms, "launches": launches, "bytes_per_launch": bpl,
                     "site_updates_per_s": N ** 3 * args.steps / sec, "ms_per_step": 1e3 * sec / args.steps,
                     "kernel": next((ln[len("# kernel "):] for ln in tune.splitlines()
                                     if ln.startswith("# kernel ")), "generic"),
                     "autotune_us": {ln.rsplit(" ", 1)[0]: float(ln.rsplit(" ", 1)[1])
                                     for ln in tune.strip().splitlines() if not ln.startswith("#")}}
        log(f"[bench] {name}: {res[name]['ms_per_step']:.4f} ms/step, "
            f"{res[name]['site_updates_per_s'] / 1e9:.2f} G site-updates/s")
        lats[name] = lat

    from cosmolattice_amd.clmi import FLD_S

    for lat in lats.values():
        lat.lf_sync()

    # spectra (A14) and Gaussian-random-field initial conditions (A15) at 256^3: time per call
    from cosmolattice_amd.clmi import PI_S
    lat = lats["f64"]
    lat.spectrum_bins(FLD_S, 1, 0, 1.0)  # plans and bin tables built on first use
    lat.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        for k in (FLD_S, PI_S):
            lat.spectrum_bins(k, 1, 0, 1.0)
    spec_ms = (time.perf_counter() - t0) / 10 * 1e3
    t0 = time.perf_counter()
    tmp, _ = make_lattice(clmi, spec, N, kIR, dt, clmi.CLMI_F64, args.seed + 1)
    tmp.synchronize()
    init_s = time.perf_counter() - t0
    tmp.close()
    extras = {"spectrum_ms_per_field": spec_ms, "spectrum_modes": N * N * (N // 2 + 1),
              "initial_conditions_s": init_s,
              "note": "spectrum: R2C FFT + deterministic |k|-shell bin sums, host-synchronous call; initial "
                      "conditions: lattice allocation, Philox GRF on both fields, two C2R FFTs each, Friedmann aDot"}

    r32, r64 = res["f32_bulk"], res["f64"]
    def roof(r, f32):
        avg_s = r["kernel_ms"] / 1e3 / max(r["launches"], 1)
        achieved = r["bytes_per_launch"] / avg_s / 1e9
        traffic = measured_traffic(r["kernel"], f32)
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "traffic_source": TRAFFIC_FILE if traffic is not None else None,
                "kernel": r["kernel"] + " (fused drift+kick+reductions)",
                "avg_launch_us": avg_s * 1e6, "timed_launches": r["launches"], "timed_every": prof_every(args.steps),
                "bytes_per_launch": r["bytes_per_launch"], "autotune_us": r["autotune_us"]}

    if args.calibrate_copy:
        # A plain copy kernel of known size (k_copy): its FETCH_SIZE / WRITE_SIZE counters
        # calibrate the counter units against 1 GiB read + 1 GiB written per launch.
        lats["f64"].copy_calibration(1 << 30, 3)

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle)")
        cpu = cpu_baseline(lats["f64"], spec, N, 2 * math.pi / kIR / N, dt, args.cpu_seconds)
    for lat in lats.values():
        lat.close()
    # accuracy of FP32-bulk vs FP64 at the workload's horizon (t = 20), whatever --steps is
    acc = None if args.no_accuracy else accuracy_horizon(clmi, spec, N, kIR, dt, args.seed, args.accuracy_steps)
    ens = None if args.no_ensemble else ensemble_bench(clmi, spec, args.ensemble_steps)
    u1 = None if args.no_u1 else u1_bench(clmi, args.u1_steps)
    c1 = None if args.no_c1 else c1_bench(clmi)
    adj = None if args.no_adjoint else adjoint_bench(clmi, steps=args.adjoint_steps)

    line = {
        "metric": METRIC,
        "value": r32["site_updates_per_s"],
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded vacuum Gaussian random fields (Philox4x32-10) on both fields",
        "config": {"workload": "lphi4 2-field preheating 256^3 leapfrog (q=100, kIR=0.05, dt=0.005, "
                               "self-consistent expansion)", "N": N, "kIR": kIR, "dt": dt, "q": q,
                   "mode": "FP32-bulk: fields f32 hi/lo, momenta f32, reductions + scale factor f64",
                   "fStar_over_MPl": spec.fStar / spec.MPl},
        "roofline": roof(r32, True),
        "cpu_baseline": cpu,
        "fp64": {"value": r64["site_updates_per_s"], "unit": "site-updates/s", "ms_per_step": r64["ms_per_step"],
                 "roofline": roof(r64, False)},
        "f32_over_f64": r32["site_updates_per_s"] / r64["site_updates_per_s"],
        "hbm_frac_f32": roof(r32, True)["frac"],
        "accuracy": acc,
        "spectra_and_ics": extras,
        "ensemble": ens,
        "u1": u1,
        "c1_vv2_64": c1,
        "adjoint": adj,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
