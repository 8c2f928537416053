"""GPU: the C++ host mirror's main programs (cosmolattice_amd/host/_bin/<model>, built from
host/cosmolattice_main.cpp) run the reference's time loop on the GPU and match the FP64 oracle.

Each case writes the initial state as raw fields, runs the model program with a parameter file
(measurement cadence, evolver), reads back its dump and its average_energy_conservation.txt, and
replays the same loop on the oracle: sync + Friedmann check at every output step, then evolve.
"""
import json
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from cosmolattice_amd.clmi import CLMI_F64, FLD_S, PI_S, PV_WEIGHTS, Lattice, VV_WEIGHTS
from test_gpu_parity import bench_spec, say, scalar_ics

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "cosmolattice_amd" / "host" / "_bin"
FP64_TOL = 1e-10
BG_NAMES = ["aI", "aSI", "aIM", "aDotI", "aDotSI", "aDotSIM", "pi2AvI", "pi2AvSI", "pi2AvSIM", "grad2AvI",
            "CSpi2AvI", "CSpi2AvSI", "CSpi2AvSIM", "CSgrad2AvI", "U1pi2AvI", "U1pi2AvSI", "U1pi2AvSIM",
            "U1Mag2AvI", "potAvI", "potAvSI"]


def lphi4_params(N, dx, dt, steps, out_every, evolver="LF", extra="", geometry=None):
    # lambda = 1, phi0 = 5.6964e18 GeV (fStar/MPl = 2.34), pi0 = -0.5 in program units
    phi0 = 5.6964e18
    geometry = geometry or f"lSide = {N * dx!r}"
    return (f"N = {N}\n{geometry}\ndt = {dt!r}\ntMax = {steps * dt!r}\n"
            f"tOutputFreq = {out_every * dt!r}\ntOutputInfreq = {out_every * dt!r}\n"
            f"tOutputRareFreq = {out_every * dt!r}\nevolver = {evolver}\nexpansion = true\n"
            f"print_headers = true\nlambda = 1\nq = 100\ninitial_amplitudes = {phi0!r} 0\n"
            f"initial_momenta = {-0.5 * phi0 * phi0!r} 0\n{extra}")


KIND_NAMES = {0: "fldS", 1: "piS", 2: "fldCS", 3: "piCS", 6: "fldU1", 7: "piU1"}


def write_state(d: Path, arrays: dict, bg: dict, N: int, t: float = 0.0):
    """The saved-state directory the host programs read (raw components + state.json)."""
    d.mkdir(parents=True, exist_ok=True)
    for (k, i, c), a in arrays.items():
        np.ascontiguousarray(a, dtype="<f8").tofile(d / f"{KIND_NAMES[k]}_{i}_{c}.f64")
    state = {"model": "test", "N": N, "tSave": t}
    for n in ("aSI", "aI", "aIM", "aDotSI", "aDotI", "aDotSIM"):
        state[n] = float(bg.get(n, 0.0))
    state["background"] = {f"bg_{n}": float(bg.get(n, 0.0)) for n in BG_NAMES}
    (d / "state.json").write_text(json.dumps(state))


def read_state(d: Path, keys, N):
    arrays = {key: np.fromfile(d / f"{KIND_NAMES[key[0]]}_{key[1]}_{key[2]}.f64", dtype="<f8").reshape(N, N, N)
              for key in keys}
    state = json.loads((d / "state.json").read_text())
    return arrays, {n: state["background"][f"bg_{n}"] for n in BG_NAMES}


def run_program(model, tmp_path, params, *args):
    exe = BIN / model
    if not exe.exists():
        pytest.fail(f"{exe} is not built (run the build first)")
    pf = tmp_path / "run.in"
    pf.write_text(params)
    r = subprocess.run([str(exe), f"input={pf}", f"outputfile={tmp_path}/", *args], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r


def observables(orc, spec, N, dx):
    """Per-field means and energies of the oracle's current state (numpy), in the layout of
    average_scalar_<i>.txt and average_energies.txt."""
    bg = orc.background()
    a, al = bg["aI"], spec.alpha
    means, erow = [], []
    for f in range(spec.n_scalars):
        phi, pi = orc.download(FLD_S, f), orc.download(PI_S, f) * a ** (al - 3)
        means.append([phi.mean(), pi.mean(), (phi ** 2).mean(), (pi ** 2).mean(),
                      math.sqrt(max((phi ** 2).mean() - phi.mean() ** 2, 0)),
                      math.sqrt(max((pi ** 2).mean() - pi.mean() ** 2, 0))])
        g2 = sum((np.roll(phi, -1, d) - phi) ** 2 for d in range(3)).mean() / dx ** 2
        erow += [0.5 * (orc.download(PI_S, f) ** 2).mean() * a ** -6, 0.5 * g2 * a ** -2]
    phis = [orc.download(FLD_S, f) for f in range(spec.n_scalars)]
    for c, e in spec.terms:
        m = np.ones_like(phis[0])
        for f, ef in enumerate(e):
            m = m * phis[f] ** ef
        erow.append(c * m.mean())
    erow.append(sum(erow))
    return means, erow


def oracle_loop(orc, steps, out_every, evolver="LF", obs=None):
    recs = []
    if evolver == "LF":
        orc.lf_init()
    for i in range(steps):
        if i % out_every == 0:
            if evolver == "LF":
                orc.lf_sync()
            if obs is not None:
                obs.append(observables(orc, orc.spec, orc.N, orc.dx))
            recs.append(orc.hubble_constraint())
        if evolver == "LF":
            orc.lf_evolve(1)
        elif evolver.startswith("PV"):
            orc.pv_evolve(PV_WEIGHTS[evolver], 1)
        else:
            orc.vv_evolve(VV_WEIGHTS[evolver], 1)
    return np.array(recs)


@pytest.mark.parametrize("evolver", ["LF", "VV4", "PV4"])
def test_host_program_matches_oracle(gpu, oracle_mod, tmp_path, evolver):
    say(f"case host lphi4 {evolver}")
    spec = bench_spec()
    N, dx, dt, steps, every = 16, 0.4909, 0.005, 60, 10
    ics = scalar_ics(spec, N, np.random.default_rng(7))
    orc = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=True)
    for (k, i), a in ics.items():
        orc.upload(k, i, 0, a)
    orc.set_background(aI=1.0, aSI=1.0, aIM=1.0, aDotI=1.0)
    _, _, rhs = orc.hubble_constraint()
    orc.set_background(aDotI=math.sqrt(rhs))
    write_state(tmp_path / "in", {(k, i, 0): a for (k, i), a in ics.items()}, orc.background(), N)
    run_program("lphi4", tmp_path, lphi4_params(N, dx, dt, steps, every, evolver),
                f"clmi_load_dir={tmp_path / 'in'}", f"clmi_dump_dir={tmp_path}")
    obs = []
    recs = oracle_loop(orc, steps, every, evolver, obs)
    keys = [(k, i, 0) for (k, i) in ics]
    got, bg = read_state(tmp_path, keys, N)
    for (k, i, c) in keys:
        ref = orc.download(k, i)
        e = float(np.max(np.abs(got[(k, i, c)] - ref)) / np.max(np.abs(ref)))
        say((k, i), e)
        assert e < FP64_TOL
    bo = orc.background()
    for n in BG_NAMES:
        if bo[n] != 0:
            assert bg[n] == pytest.approx(bo[n], rel=FP64_TOL), n
    cons = np.loadtxt(tmp_path / "average_energy_conservation.txt", ndmin=2)
    assert cons.shape == (len(recs), 4)
    np.testing.assert_allclose(cons[:, 0], np.arange(len(recs)) * every * dt, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(cons[:, 2:], recs[:, 1:], rtol=FP64_TOL)
    np.testing.assert_allclose(cons[:, 1], recs[:, 0], rtol=1e-6, atol=1e-13)
    assert (tmp_path / "average_energy_conservation.txt").read_text().startswith(
        "#t    rel_diff_friedmann    LHS_friedmann    RHS_friedmann\n")
    # scalar means and energies at every output time
    en = np.loadtxt(tmp_path / "average_energies.txt", ndmin=2)
    np.testing.assert_allclose(en[:, 1:], np.array([o[1] for o in obs]), rtol=FP64_TOL, atol=1e-14)
    assert (tmp_path / "average_energies.txt").read_text().startswith(
        "#t    E^kin_scal0    E^grad_scal0    E^kin_scal1    E^grad_scal1    Vpot_term_0    Vpot_term_1    E_tot\n")
    for f in range(2):
        me = np.loadtxt(tmp_path / f"average_scalar_{f}.txt", ndmin=2)
        np.testing.assert_allclose(me[:, 1:], np.array([o[0][f] for o in obs]), rtol=1e-9, atol=1e-13)


def fnv1a64(s: str) -> int:
    h = 1469598103934665603
    for ch in s.encode():
        h ^= ch
        h = (h * 1099511628211) % (1 << 64)
    return h


def test_host_initializer_matches_python_path(gpu, oracle_mod, tmp_path):
    """ModelInitializer through the model program equals the same steps through the Python binding
    (scale factor from the homogeneous energy, GRF per field, aDot from the Friedmann constraint),
    and its fields equal the numpy GRF restatement."""
    say("case host initializer")
    spec = bench_spec()
    N, kIR, dt = 32, 0.4, 0.005
    dx = 2 * math.pi / kIR / N
    kcut = N * kIR / 4
    run_program("lphi4", tmp_path, lphi4_params(N, dx, dt, 0, 1, extra=f"kCutOff = {kcut!r}\nbaseSeed = 42\n",
                                                geometry=f"kIR = {kIR!r}"),
                f"clmi_dump_dir={tmp_path}")
    keys = [(FLD_S, 0, 0), (PI_S, 0, 0), (FLD_S, 1, 0), (PI_S, 1, 0)]
    got, bg = read_state(tmp_path, keys, N)
    lat = Lattice(spec, N, dx, dt, CLMI_F64, expansion=True)
    aDot0 = spec.fStar / spec.MPl * math.sqrt((0.5 * spec.pi0[0] ** 2 + spec.potential(spec.fld0)) / 3.0)
    lat.set_background(aI=1.0, aSI=1.0, aIM=1.0, aDotI=aDot0)
    m2 = spec.mass2(spec.fld0)
    seed = fnv1a64("42")
    for f in range(2):
        lat.grf_scalar(f, seed, f, m2[f], aDot0, kIR, kcut, spec.fld0[f], spec.pi0[f])
        phi_o, pi_o = oracle_mod.grf_scalar(N, dx, spec, seed, f, m2[f], aDot0, kIR, kcut, spec.fld0[f], spec.pi0[f])
        # kIR and dx pass through the run parameters' geometry rule, so allow last-bit differences
        for k, ref in ((FLD_S, lat.download(FLD_S, f, 0)), (PI_S, lat.download(PI_S, f, 0))):
            assert np.max(np.abs(got[(k, f, 0)] - ref)) <= 1e-13 * np.max(np.abs(ref))
        assert np.max(np.abs(got[(FLD_S, f, 0)] - phi_o)) < 1e-12 * np.max(np.abs(phi_o - spec.fld0[f]))
    _, _, rhs = lat.hubble_constraint()
    assert bg["aDotI"] == pytest.approx(math.sqrt(rhs), rel=1e-14)
    assert bg["aI"] == 1.0
    lat.close()


def test_host_lphi4u1_initializer(gpu, oracle_mod, tmp_path):
    """The lphi4U1 model program's initializer draws the complex scalar with constrained phases and
    the gauge momenta from the Gauss constraint: the fields equal the numpy restatement (stream 1,
    key FNV-1a-64 of the seed string) and the Gauss law holds."""
    say("case host lphi4U1 initializer")
    from cosmolattice_amd.clmi import FLD_CS, PI_CS, PI_U1, lphi4U1

    N, kIR, dt = 32, 0.4, 0.005
    dx = 2 * math.pi / kIR / N
    phi0 = 5.6964e18
    g = math.sqrt(10.0)
    params = (f"N = {N}\nkIR = {kIR!r}\ndt = {dt!r}\ntMax = 0\ntOutputFreq = {dt!r}\ntOutputInfreq = {dt!r}\n"
              f"tOutputRareFreq = {dt!r}\nexpansion = false\nlambda = 1\ngU1s = {g!r}\nCSU1_charges = 1\n"
              f"cmplx_field_initial_norm = {phi0!r}\ncmplx_momentum_initial_norm = 0\nbaseSeed = 3\n"
              f"kCutOff = {N * kIR / 4!r}\n")
    run_program("lphi4U1", tmp_path, params, f"clmi_dump_dir={tmp_path}")
    keys = [(FLD_CS, 0, 0), (FLD_CS, 0, 1), (PI_CS, 0, 0), (PI_CS, 0, 1)] + [(PI_U1, 0, d) for d in range(3)]
    got, _ = read_state(tmp_path, keys, N)
    spec = lphi4U1(10.0, phi0_gev=phi0, lam=1.0, pi0_gev2=0.0)
    m2 = spec.cs_mass2(spec.fld0)
    phi, pi, piA = oracle_mod.grf_u1(N, dx, spec, fnv1a64("3"), 1, m2, 0.0, kIR, N * kIR / 4, spec.fld0, spec.pi0)
    for c in range(2):
        assert np.max(np.abs(got[(FLD_CS, 0, c)] - phi[c])) < 1e-12 * np.max(np.abs(phi[c] - spec.fld0[c]))
        assert np.max(np.abs(got[(PI_CS, 0, c)] - pi[c])) < 1e-12 * np.max(np.abs(pi[c]))
    for d in range(3):
        assert np.max(np.abs(got[(PI_U1, 0, d)] - piA[d])) < 1e-12 * np.max(np.abs(piA[d]))
    orc = oracle_mod.OracleLattice(spec, N, dx, dt, expansion=False)
    for (k, i, c), a in got.items():
        orc.upload(k, i, c, a)
    assert orc.gauss_u1()[0] < 1e-13


def test_host_lphi4u1_measurement_files(gpu, tmp_path):
    """The lphi4U1 program writes the complex-scalar and U(1) files of complexscalarmeasurer.h and
    u1measurer.h at the measurement cadence: component means, norms, Gauss law and norm spectra.  Norms
    and component means come from separate reductions and must agree; <|B|^2> must agree with the
    magnetic energy of average_energies.txt; the Gauss law holds for the program's initial conditions."""
    say("case host lphi4U1 measurement files")
    from cosmolattice_amd.clmi import lphi4U1

    N, kIR, dt = 32, 0.4, 0.005
    phi0 = 5.6964e18
    g = math.sqrt(10.0)
    params = (f"N = {N}\nkIR = {kIR!r}\ndt = {dt!r}\ntMax = {4 * dt!r}\ntOutputFreq = {2 * dt!r}\n"
              f"tOutputInfreq = {2 * dt!r}\ntOutputRareFreq = {2 * dt!r}\nexpansion = false\nlambda = 1\n"
              f"gU1s = {g!r}\nCSU1_charges = 1\ncmplx_field_initial_norm = {phi0!r}\n"
              f"cmplx_momentum_initial_norm = 0\nbaseSeed = 3\nkCutOff = {N * kIR / 4!r}\nprint_headers = true\n")
    run_program("lphi4U1", tmp_path, params)
    spec = lphi4U1(10.0, phi0_gev=phi0, lam=1.0, pi0_gev2=0.0)
    w2 = (spec.omegaStar / spec.fStar) ** 2
    load = lambda n: np.atleast_2d(np.loadtxt(tmp_path / n))
    re, im = load("average_cmplx_scalar_0_0.txt"), load("average_cmplx_scalar_0_1.txt")
    nrm = load("average_norm_cmplx_scalar_0.txt")
    u1 = load("average_norm_U1_0.txt")
    gs = load("average_gauss_U1_0.txt")
    en = load("average_energies.txt")
    for a in (re, im, nrm, u1, gs, en):
        assert a.shape[0] == 2 and np.allclose(a[:, 0], [0.0, 2 * dt])
    # components carry sqrt2: <2 phi_0^2> + <2 phi_1^2> = 2 <|phi|^2>, the same for the momenta
    np.testing.assert_allclose(re[:, 3] + im[:, 3], 2 * nrm[:, 3], rtol=1e-12)
    np.testing.assert_allclose(re[:, 4] + im[:, 4], 2 * nrm[:, 4], rtol=1e-12)
    assert np.all(nrm[:, 1] <= np.sqrt(nrm[:, 3]) * (1 + 1e-14))  # <|phi|> <= sqrt(<|phi|^2>)
    # E_B = w^2 <|B|^2> / 2 (a = 1); A starts at 0 and grows
    assert u1[0, 4] == 0.0 and u1[1, 4] > 0.0
    header = (tmp_path / "average_energies.txt").read_text().splitlines()[0].lstrip("#").split()
    np.testing.assert_allclose(0.5 * w2 * u1[:, 4], en[:, header.index("E^grad_U10")], rtol=1e-12)
    np.testing.assert_allclose(0.5 * w2 * u1[:, 3], en[:, header.index("E^kin_U10")], rtol=1e-12)
    assert np.all(gs[:, 1] < 1e-10)
    for name in ("spectra_norm_cmplx_scalar_0.txt", "spectra_norm_U1_0.txt"):
        text = (tmp_path / name).read_text()
        blocks = [b for b in text.split("\n\n") if b.strip()]
        assert len(blocks) == 2
        rows = np.loadtxt((tmp_path / name).open(), comments="#")
        assert rows.shape == (2 * (int(math.floor(math.sqrt(3) / 2 * N)) + 1), 3)
        assert np.all(np.isfinite(rows)) and rows[:, 1].max() > 0


@pytest.mark.parametrize("version", [1, 2])
def test_host_spectrum_normalisation_parseval(gpu, tmp_path, version):
    """PS_type 1 spectra written by the lphi4 program obey Parseval: with n_b the bin's |n| used in the
    prefactor (b dk for PS_version 2, the mean |n| of the bin's modes for PS_version 1),
    sum_{b >= 1} PS_b dk / n_b = <f^2> - <f>^2, which the same program writes as rms^2 in
    average_scalar_<i>.txt (and a^(2 alpha - 6) times the momentum variance for PS(pi)).  This pins the
    spectrum prefactor of measurements/powerspectrum.h:85-110 for both versions."""
    say(f"case host spectrum parseval v{version}")
    N, dx, dt, steps, every = 32, 0.6, 0.005, 4, 2
    extra = f"PS_type = 1\nPS_version = {version}\nbaseSeed = 5\nkCutOff = 2.0\n"
    run_program("lphi4", tmp_path, lphi4_params(N, dx, dt, steps, every, extra=extra))
    text = (tmp_path / "spectra_scalar_1.txt").read_text()
    blocks = [np.loadtxt(b.splitlines()) for b in
              ["\n".join(ln for ln in blk.splitlines() if not ln.startswith("#")) for blk in text.split("\n\n")]
              if b.strip()]
    means = np.atleast_2d(np.loadtxt(tmp_path / "average_scalar_1.txt"))
    assert len(blocks) == means.shape[0] >= 2
    nb = blocks[0].shape[0]
    kIR = 2 * math.pi / (N * dx)
    if version == 2:
        nbar = np.arange(nb) * 1.0
    else:
        n = np.fft.fftfreq(N, d=1.0 / N)
        n = np.where(n == -N // 2, N // 2, n)
        nn = np.sqrt(n[:, None, None] ** 2 + n[None, :, None] ** 2 + n[None, None, :] ** 2).ravel()
        b = np.minimum(np.floor(nn + 0.5).astype(int), nb - 1)
        nbar = np.bincount(b, weights=nn, minlength=nb) / np.maximum(np.bincount(b, minlength=nb), 1)
    for blk, row in zip(blocks, means):
        np.testing.assert_allclose(blk[:, 0], np.arange(nb) * kIR, rtol=1e-12)
        for col, rms in ((1, row[5]), (2, row[6])):
            total = np.sum(blk[1:, col] / nbar[1:])
            assert total == pytest.approx(rms ** 2, rel=1e-7), (col, total, rms ** 2)


def _spectrum_blocks(path):
    text = path.read_text()
    return [np.loadtxt(b.splitlines()) for b in
            ["\n".join(ln for ln in blk.splitlines() if not ln.startswith("#")) for blk in text.split("\n\n")]
            if b.strip()]


@pytest.mark.parametrize("ps_type", [1, 2])
def test_host_spectrum_version3_against_version1(gpu, tmp_path, ps_type):
    """PS_version 3 (measurements/powerspectrum.h:111-131) puts |n| (PS_type 1) or |n|^3 (PS_type 2)
    inside the bin sum and drops the bin's |n| and 1/dk from the prefactor.  Two runs of the lphi4
    program from the same seed, PS_version 1 and 3 with bin width dk = 0.5, see the same fields, so in
    bin b (|n| in [(b - 1/2) dk, (b + 1/2) dk)) the ratio PS3 / PS1 lies within
    dk [(b - 1/2)/(b + 1/2), (b + 1/2)/(b - 1/2)]^p: a kept 1/dk doubles it, a missing |n| weight
    divides it by about b dk."""
    say(f"case host spectrum v3 type {ps_type}")
    N, dx, dt, steps, every, dk = 32, 0.6, 0.005, 4, 2, 0.5
    out = {}
    for version in (1, 3):
        d = tmp_path / f"v{version}"
        d.mkdir()
        extra = f"PS_type = {ps_type}\nPS_version = {version}\nbaseSeed = 5\nkCutOff = 2.0\ndeltaKBin = {dk}\n"
        run_program("lphi4", d, lphi4_params(N, dx, dt, steps, every, extra=extra))
        out[version] = _spectrum_blocks(d / "spectra_scalar_1.txt")
    p = 1 if ps_type == 1 else 3
    assert len(out[1]) == len(out[3]) >= 2
    for b1, b3 in zip(out[1], out[3]):
        np.testing.assert_array_equal(b1[:, 0], b3[:, 0])
        for col in (1, 2):
            b = np.arange(b1.shape[0])
            # the last bin also takes the corner modes past its upper edge
            ok = (b >= 2) & (b < b1.shape[0] - 1) & (b1[:, col] > 1e-30 * b1[:, col].max())
            ratio = b3[ok, col] / b1[ok, col]
            lo = dk * ((b[ok] - 0.5) / (b[ok] + 0.5)) ** p
            hi = dk * ((b[ok] + 0.5) / (b[ok] - 0.5)) ** p
            assert ok.sum() > 10
            assert np.all(ratio >= lo * (1 - 1e-12)) and np.all(ratio <= hi * (1 + 1e-12)), (col, ratio, lo, hi)


def test_host_backup_and_restart(gpu, tmp_path):
    """A run restarted from its own backup (load_file, as SimulationManager::loadSim) ends where the
    uninterrupted run ends, to rounding: the restart recomputes the cached averages
    (setAllAverages) and skips the LeapFrog start-up half step, as the reference does.

    The restarted loop counts steps from 0 again (source/cosmolattice.cpp:168), so it measures at
    its first step, and a measurement re-syncs the momenta and resets the half-step averages.  The
    two runs therefore agree to rounding only when the uninterrupted run also measures at the
    backup time: backups every 41 steps (t = dt, 42 dt) and measurements every 21 steps (t = 0,
    21 dt, 42 dt)."""
    say("case host restart")
    N, dx, dt = 16, 0.4909, 0.005
    spec = bench_spec()
    # the backup interval in steps is tBackupFreq / dt truncated (runparameters.h:216): 41.5 dt -> 41
    params = lphi4_params(N, dx, dt, 60, 21, extra="kCutOff = 3.0\nbaseSeed = 5\n") + \
        f"tBackupFreq = {41.5 * dt!r}\n"
    full = tmp_path / "full"
    full.mkdir()
    run_program("lphi4", full, params, f"backup_dir={full}/", f"clmi_dump_dir={full}/end")
    backup = full / "lphi4_backup"
    assert (backup / "state.json").exists() and (full / "lphi4_backup~").exists()
    st = json.loads((backup / "state.json").read_text())
    assert st["tSave"] == pytest.approx(42 * dt)
    rest = tmp_path / "rest"
    rest.mkdir()
    run_program("lphi4", rest, params, f"load_file={backup}", f"clmi_dump_dir={rest}/end")
    keys = [(FLD_S, 0, 0), (PI_S, 0, 0), (FLD_S, 1, 0), (PI_S, 1, 0)]
    a, bga = read_state(full / "end", keys, N)
    b, bgb = read_state(rest / "end", keys, N)
    for k in keys:
        assert np.max(np.abs(a[k] - b[k])) <= 1e-11 * np.max(np.abs(a[k])), k
    assert bga["aI"] == pytest.approx(bgb["aI"], rel=1e-12)
    assert json.loads((rest / "end" / "state.json").read_text())["tSave"] == pytest.approx(60 * dt)


def test_host_spectra_verbosity_columns(gpu, tmp_path):
    """spectraVerbosity (runparameters.h:96,233-235; spectrumsaverstd.h:56-82): three runs of the lphi4
    program from the same seed write the same spectra with 0 -> (k, PS...), 1 -> (mean |k| of the bin,
    PS...), 2 -> (k, mean |k|, mode count, PS...).  The mean |k| and the counts equal a brute-force
    enumeration of the lattice modes by |n| shell (the counts add up to N^3), the spectrum columns are the
    same numbers in all three files, and spectraVerbosity = 3 is refused with the reference's message."""
    say("case host spectra verbosity")
    N, dx, dt, steps, every = 32, 0.6, 0.005, 4, 2
    out = {}
    for v in (0, 1, 2):
        d = tmp_path / f"v{v}"
        d.mkdir()
        extra = f"PS_type = 1\nPS_version = 1\nbaseSeed = 5\nkCutOff = 2.0\nspectraVerbosity = {v}\n"
        run_program("lphi4", d, lphi4_params(N, dx, dt, steps, every, extra=extra))
        out[v] = _spectrum_blocks(d / "spectra_scalar_0.txt")
        header = (d / "spectra_scalar_0.txt").read_text().splitlines()[0]
        assert header.startswith({0: "#k ", 1: "#k_avg", 2: "#k    k_avg    multiplicity"}[v]), header
    nb = out[0][0].shape[0]
    kIR = 2 * math.pi / (N * dx)
    n = np.fft.fftfreq(N, d=1.0 / N)
    n = np.where(n == -N // 2, N // 2, n)
    nn = np.sqrt(n[:, None, None] ** 2 + n[None, :, None] ** 2 + n[None, None, :] ** 2).ravel()
    b = np.minimum(np.floor(nn + 0.5).astype(int), nb - 1)
    count = np.bincount(b, minlength=nb).astype(float)
    kavg = np.bincount(b, weights=nn, minlength=nb) / np.maximum(count, 1) * kIR
    assert count.sum() == N ** 3
    assert len(out[0]) == len(out[1]) == len(out[2]) >= 2
    for b0, b1, b2 in zip(out[0], out[1], out[2]):
        assert b0.shape == (nb, 3) and b1.shape == (nb, 3) and b2.shape == (nb, 5)
        np.testing.assert_allclose(b0[:, 0], np.arange(nb) * kIR, rtol=1e-12)
        np.testing.assert_array_equal(b2[:, 0], b0[:, 0])
        np.testing.assert_allclose(b1[:, 0], kavg, rtol=1e-12)
        np.testing.assert_array_equal(b2[:, 1], b1[:, 0])
        np.testing.assert_array_equal(b2[:, 2], count)
        for c in (1, 2):
            np.testing.assert_array_equal(b1[:, c], b0[:, c])
            np.testing.assert_array_equal(b2[:, 2 + c], b0[:, c])
    d = tmp_path / "bad"
    d.mkdir()
    pf = d / "run.in"
    pf.write_text(lphi4_params(N, dx, dt, steps, every, extra="spectraVerbosity = 3\n"))
    r = subprocess.run([str(BIN / "lphi4"), f"input={pf}", f"outputfile={d}/"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and "spectraVerbosity 3 is not a valid spectraVerbosity." in (r.stderr + r.stdout)
