"""C++ host mirror of the CosmoInterface model/evolver API (cosmolattice_amd/host), on the CPU.

A model header is compiled against the mirror and the library configuration it produces is
printed by tests/cpp/model_config.cpp (no GPU involved): the monomials of the potential, the
program units, the initial potential and masses.  They must equal what the model defines in the
reference's equations (models/lphi4.h:125-226, models/lphi4U1.h:96-137).
"""
import math
import subprocess
from pathlib import Path

import pytest

from cosmolattice_amd._build import HOST, host_compile_cmd

ROOT = Path(__file__).resolve().parent.parent
TOOL_SRC = ROOT / "tests" / "cpp" / "model_config.cpp"
REFERENCE_MODELS = Path("/root/reference/models")

LPHI4_IN = """
expansion = true
evolver = LF
N = 64
dt = 0.01
kIR = .7
tOutputFreq = 0.2
tMax = 10
initial_amplitudes = 5.6964e18 0  # GeV
initial_momenta = -4.86735e30 0   # GeV^2
lambda = 9e-14
q = 300
"""

LPHI4U1_IN = """
N = 32
dt = 0.005
kIR = 0.5
lambda = 9e-14
gU1s = 9.486832980505138e-07
CSU1_charges = 1
cmplx_field_initial_norm = 5.6964e18
cmplx_momentum_initial_norm = 0
"""


def config_of(tmp_path, header, cls, params, *extra):
    exe = tmp_path / f"cfg_{cls}"
    r = subprocess.run(host_compile_cmd(TOOL_SRC, exe, header, cls), capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    pf = tmp_path / "params.in"
    pf.write_text(params)
    r = subprocess.run([str(exe), f"input={pf}", *extra], capture_output=True, text=True)
    out = {"rc": r.returncode, "stderr": r.stderr, "term": [], "cs_term": [], "mass2": [], "fld0": [], "pi0": []}
    for line in r.stdout.splitlines():
        k, *v = line.split()
        if k in ("term", "cs_term", "mass2", "fld0", "pi0"):
            out[k].append([float(x) for x in v])
        else:
            out[k] = float(v[0])
    return out


def terms_set(c):
    return sorted((round(t[0], 12), tuple(int(e) for e in t[1:])) for t in c["term"])


def test_lphi4_header_gives_lphi4_monomials(tmp_path):
    c = config_of(tmp_path, "models/lphi4.h", "lphi4", LPHI4_IN)
    assert c["rc"] == 0, c["stderr"]
    q = 300.0
    # V = phi^4/4 + q/2 phi^2 chi^2 in program units (models/lphi4.h:154-177)
    assert terms_set(c) == sorted([(0.25, (4, 0, 0, 0)), (0.5 * q, (2, 2, 0, 0))])
    assert c["n_scalars"] == 2 and c["n_cscalars"] == 0 and c["n_u1"] == 0
    assert c["fStar"] == 5.6964e18
    assert c["omegaStar"] == pytest.approx(math.sqrt(9e-14) * 5.6964e18, rel=1e-15)
    assert c["alpha"] == 1.0
    assert c["dx"] == pytest.approx(2 * math.pi / 0.7 / 64, rel=1e-15)
    assert c["pot0"] == pytest.approx(0.25, rel=1e-15)  # phi = 1, chi = 0
    # masses from potDeriv2 at the homogeneous point: 3 phi^2 + q chi^2, q phi^2
    assert [m[0] for m in c["mass2"]] == pytest.approx([3.0, q], rel=1e-15)
    assert c["MPl"] == 2.435e18


def test_lphi4_header_matches_python_binding(tmp_path):
    from cosmolattice_amd import clmi

    c = config_of(tmp_path, "models/lphi4.h", "lphi4", LPHI4_IN)
    spec = clmi.lphi4(300.0, n_fields=2, phi0_gev=5.6964e18, lam=9e-14, pi0_gev2=-4.86735e30)
    py = sorted((round(cf, 12), tuple((e + [0, 0, 0, 0])[:4])) for cf, e in spec.terms)
    assert terms_set(c) == py
    assert c["omegaStar"] == pytest.approx(spec.omegaStar, rel=1e-15)
    assert c["pi0"][0][0] / c["fStar"] / c["omegaStar"] == pytest.approx(spec.pi0[0], rel=1e-14)


def test_single_field_header(tmp_path):
    params = LPHI4_IN.replace("initial_amplitudes = 5.6964e18 0", "initial_amplitudes = 5.6964e18").replace(
        "initial_momenta = -4.86735e30 0", "initial_momenta = -4.86735e30")
    c = config_of(tmp_path, "models/lphi4_1f.h", "lphi4_1f", params)
    assert c["rc"] == 0, c["stderr"]
    assert terms_set(c) == [(0.25, (4, 0, 0, 0))]
    assert [m[0] for m in c["mass2"]] == pytest.approx([3.0])


def test_lphi4u1_header(tmp_path):
    c = config_of(tmp_path, "models/lphi4U1.h", "lphi4U1", LPHI4U1_IN)
    assert c["rc"] == 0, c["stderr"]
    assert c["n_scalars"] == 0 and c["n_cscalars"] == 1 and c["n_u1"] == 1
    assert c["cs_term"] == [[1.0, 4.0]]  # V = |phi|^4 (models/lphi4U1.h:117-120)
    assert c["gQ"] == pytest.approx(9.486832980505138e-07, rel=1e-15)  # charge 1 x gU1s
    assert c["pot0"] == pytest.approx(1.0, rel=1e-14)  # |phi| = 1 in program units
    # deriv2CS per component at phi_0 = phi_1 = 1/sqrt2: V'' phi_c^2/|phi|^2 + V'/|phi| (1 - phi_c^2/|phi|^2)
    assert c["mass2cs"] == pytest.approx(12.0 * 0.5 + 4.0 * 0.5, rel=1e-14)


@pytest.mark.skipif(not (REFERENCE_MODELS / "lphi4.h").exists(), reason="reference tree not present")
@pytest.mark.parametrize("name,params", [("lphi4", LPHI4_IN), ("lphi4U1", LPHI4U1_IN)])
def test_reference_model_headers_compile_unchanged(tmp_path, name, params):
    """The reference's own model headers compile against the mirror as they are and give the same
    configuration as this package's headers."""
    ref = config_of(tmp_path, str(REFERENCE_MODELS / f"{name}.h"), name, params)
    ours = config_of(tmp_path, f"models/{name}.h", name, params)
    assert ref["rc"] == 0, ref["stderr"]
    assert terms_set(ref) == terms_set(ours)
    assert ref["cs_term"] == ours["cs_term"]
    for k in ("fStar", "omegaStar", "alpha", "gQ", "pot0", "dx"):
        assert ref[k] == pytest.approx(ours[k], rel=1e-15), k
    assert ref["mass2"] == ours["mass2"]


BAD_HEADER = """
#include "CosmoInterface/cosmointerface.h"
namespace TempLat {
  struct BadPars : public DefaultModelPars { static constexpr size_t NScalars = 1; using FloatType = double; };
  class bad : public MakeModel(bad, BadPars) {
  public:
    bad(ParameterParser &p, RunParameters<double> &r, auto tb)
        : AbstractModel(p, r.getLatParams(), tb, r.dt, "bad")
    { fldS0[0] = 1; fStar = 1; omegaStar = 1; setInitialPotentialAndMassesFromPotential(); }
    auto potentialTerms(Tag<0>) const { return FloatType(0.25) * pow<4>(fldS(0_c)); }
    auto potDeriv(Tag<0>) const { return 4 * pow<3>(fldS(0_c)); }  // should be phi^3
    auto potDeriv2(Tag<0>) const { return 3 * pow<2>(fldS(0_c)); }
  };
}
"""


def test_inconsistent_potential_derivative_is_rejected(tmp_path):
    h = tmp_path / "bad.h"
    h.write_text(BAD_HEADER)
    c = config_of(tmp_path, str(h), "bad", "N = 8\ndt = 0.01\nkIR = 1\n")
    assert c["rc"] == 1
    assert "potDeriv(Tag<0>) is not the derivative" in c["stderr"]


def test_run_parameters_geometry_consistency(tmp_path):
    c = config_of(tmp_path, "models/lphi4.h", "lphi4", LPHI4_IN, "lSide=3.0")
    assert c["rc"] == 1 and "not consistent" in c["stderr"]
    c = config_of(tmp_path, "models/lphi4.h", "lphi4", LPHI4_IN.replace("kIR = .7", "lSide = 10"))
    assert c["rc"] == 0 and c["dx"] == pytest.approx(10 / 64, rel=1e-15)


def test_missing_parameter_is_reported(tmp_path):
    c = config_of(tmp_path, "models/lphi4.h", "lphi4", LPHI4_IN.replace("q = 300", ""))
    assert c["rc"] == 1 and "'q' is missing" in c["stderr"]


def test_model_programs_built():
    for name in ("lphi4", "lphi4_1f", "lphi4U1"):
        assert (HOST / "_bin" / name).exists(), f"run the build first: {name} main program missing"


def test_couplings_manager_matches_reference_kat(tmp_path):
    """The mirror's CouplingsManager against the known answers of the reference's own test
    (tests/CosmoInterface/couplingsmanager.cpp:21-58, fixture tests/golden/couplingsmanager_kat.json):
    the compile-time coupling pattern, g*Q consumed in pattern order over coupled pairs only, the
    couplings g, and the two count errors."""
    import json

    from cosmolattice_amd._build import HOST as host_dir

    kat = json.loads((ROOT / "tests" / "golden" / "couplingsmanager_kat.json").read_text())
    exe = tmp_path / "couplings_kat"
    r = subprocess.run(["g++", "-std=c++20", "-O1", "-Wall", f"-I{host_dir}", f"-I{ROOT / 'include'}",
                        str(ROOT / "tests" / "cpp" / "couplings_kat.cpp"), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    for case in kat["cases"]:
        args = [",".join(str(x) for x in case["charges"]), ",".join(str(x) for x in case["couplings"])]
        out = subprocess.run([str(exe), *args], capture_output=True, text=True, check=True).stdout.splitlines()
        rows = [ln.split() for ln in out]
        assert int(next(r_[1] for r_ in rows if r_[0] == "how_many_couples")) == kat["how_many_couples"]
        if "throws" in case:
            assert ["throws", case["throws"]] in rows
            continue
        pattern = {(int(r_[1]), int(r_[2])): int(r_[3]) for r_ in rows if r_[0] == "couples"}
        for m in range(kat["n_matter"]):
            for g in range(kat["n_gauge"]):
                assert pattern[(m, g)] == kat["pattern"][m * kat["n_gauge"] + g]
        charge = {f"{r_[1]},{r_[2]}": float(r_[3]) for r_ in rows if r_[0] == "charge"}
        for k, v in case["effective_charge"].items():
            assert charge[k] == v, k
        coupling = {r_[1]: float(r_[2]) for r_ in rows if r_[0] == "coupling"}
        for k, v in case["coupling"].items():
            assert coupling[k] == v, k


def test_run_parameters_match_reference_parameter_lists(tmp_path):
    """The mirror's RunParameters on the parameter lists of the reference's runparameters test
    (tests/CosmoInterface/runparameters.cpp:23-37, lists in tests/golden/runparameters/): kIR = 1
    gives lSide = 2 pi; no geometry, or kIR and lSide that disagree, are rejected; lSide = 2 gives
    kIR = pi."""
    import json

    from cosmolattice_amd._build import HOST as host_dir

    gold = ROOT / "tests" / "golden" / "runparameters"
    expected = json.loads((gold / "expected.json").read_text())
    exe = tmp_path / "runparams_probe"
    r = subprocess.run(["g++", "-std=c++20", "-O1", "-Wall", f"-I{host_dir}", f"-I{ROOT / 'include'}",
                        str(ROOT / "tests" / "cpp" / "runparams_probe.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    for name, want in expected.items():
        if name == "source":
            continue
        out = subprocess.run([str(exe), f"input={gold / name}"], capture_output=True, text=True, check=True).stdout
        rows = dict(ln.split(None, 1) for ln in out.splitlines())
        if "throws" in want:
            assert rows.get("throws") == want["throws"], name
        else:
            for k, v in want.items():
                assert float(rows[k]) == pytest.approx(v, rel=1e-15), (name, k)
