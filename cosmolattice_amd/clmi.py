"""ctypes binding of libclmi.so (include/clmi.h) and a thin model/evolver layer on top of it.

The classes here mirror the reference's CosmoInterface objects used on the lattice-evolution path
(CosmoLattice source tree):

* ``Lattice`` owns one ``clmi_ctx``: the device-resident field collections and background of an
  ``AbstractModel`` (include/CosmoInterface/abstractmodel.h:147-186).
* ``LeapFrog`` / ``VelocityVerlet`` expose ``evolve`` and ``sync`` with the semantics of
  include/CosmoInterface/evolvers/leapfrog.h:42-116 and velocityverlet.h:51-122.

Everything runs in the HIP library; there is no CPU fallback.  Loading fails loudly when the
library is missing or no GPU is present.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from pathlib import Path
from typing import Sequence

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libclmi.so"

CLMI_OK = 0
CLMI_F64 = 0
CLMI_F32_BULK = 1
CLMI_MAX_SCALARS = 4
CLMI_MAX_TERMS = 8
CLMI_FLAG_NO_TILED = 1
CLMI_FLAG_SHAPE_ALT = 2
CLMI_FLAG_SHAPE_TY4 = 4
CLMI_FLAG_GAUGE_TY8 = 16
CLMI_FLAG_NO_DOUBLE_KICK = 32


def CLMI_FLAG_MEMMODE(m: int) -> int:
    return m << 8

FLD_S, PI_S, FLD_CS, PI_CS, FLD_U1, PI_U1 = 0, 1, 2, 3, 6, 7
MAG_U1 = 64  # derived kind: B = forward-difference curl of the U(1) field / dx

# Reduced Planck mass in GeV, the reference's Constants::reducedMPlanck
# (include/CosmoInterface/abstractmodel/scalefactorbase.h:18).
REDUCED_MPL = 2.435e18


class ClmiConfig(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("n_scalars", C.c_int), ("n_cscalars", C.c_int), ("n_u1", C.c_int),
        ("precision", C.c_int), ("expansion", C.c_int), ("device", C.c_int), ("flags", C.c_int),
        ("dx", C.c_double), ("dt", C.c_double),
        ("alpha", C.c_double), ("fStar", C.c_double), ("omegaStar", C.c_double), ("MPl", C.c_double),
        ("gQ", C.c_double),
        ("n_terms", C.c_int), ("n_cs_terms", C.c_int),
        ("term_coeff", C.c_double * CLMI_MAX_TERMS),
        ("term_exp", (C.c_int * CLMI_MAX_SCALARS) * CLMI_MAX_TERMS),
        ("cs_coeff", C.c_double * CLMI_MAX_TERMS),
        ("cs_pow", C.c_int * CLMI_MAX_TERMS),
    ]


BACKGROUND_FIELDS = [
    "aI", "aSI", "aIM", "aDotI", "aDotSI", "aDotSIM",
    "pi2AvI", "pi2AvSI", "pi2AvSIM", "grad2AvI",
    "CSpi2AvI", "CSpi2AvSI", "CSpi2AvSIM", "CSgrad2AvI",
    "U1pi2AvI", "U1pi2AvSI", "U1pi2AvSIM", "U1Mag2AvI",
    "potAvI", "potAvSI",
]


class Background(C.Structure):
    _fields_ = [(n, C.c_double) for n in BACKGROUND_FIELDS]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n in BACKGROUND_FIELDS}


class Averages(C.Structure):
    _fields_ = [(n, C.c_double) for n in ["pi2S", "grad2S", "pot", "pi2CS", "grad2CS", "pi2U1", "B2U1", "reserved"]]


_lib = None


def load_library(path: Path | None = None) -> C.CDLL:
    """Load libclmi.so once.  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with __graft_entry__.build() (no CPU fallback exists)")
    lib = C.CDLL(str(p))
    P = C.c_void_p
    sig = {
        "clmi_create": (C.c_int, [C.POINTER(ClmiConfig), C.POINTER(P)]),
        "clmi_destroy": (None, [P]),
        "clmi_last_error": (C.c_char_p, [P]),
        "clmi_synchronize": (C.c_int, [P]),
        "clmi_upload_field": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]),
        "clmi_download_field": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]),
        "clmi_set_background": (C.c_int, [P, C.POINTER(Background)]),
        "clmi_get_background": (C.c_int, [P, C.POINTER(Background)]),
        "clmi_lf_init": (C.c_int, [P, C.c_int]),
        "clmi_lf_evolve": (C.c_int, [P, C.c_int]),
        "clmi_lf_sync": (C.c_int, [P]),
        "clmi_lf_is_synced": (C.c_int, [P]),
        "clmi_vv_evolve": (C.c_int, [P, C.POINTER(C.c_double), C.c_int, C.c_int]),
        "clmi_pv_evolve": (C.c_int, [P, C.POINTER(C.c_double), C.c_int, C.c_int]),
        "clmi_kick": (C.c_int, [P, C.c_double]),
        "clmi_drift": (C.c_int, [P, C.c_double]),
        "clmi_set_all_averages": (C.c_int, [P]),
        "clmi_volume_averages": (C.c_int, [P, C.POINTER(Averages)]),
        "clmi_scale_factor_kernel": (C.c_int, [P, C.POINTER(C.c_double)]),
        "clmi_hubble_constraint": (C.c_int, [P, C.POINTER(C.c_double)]),
        "clmi_gauss_u1": (C.c_int, [P, C.POINTER(C.c_double)]),
        "clmi_field_means": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_double)]),
        "clmi_field_stats": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]),
        "clmi_norm_averages": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_double)]),
        "clmi_monomial_averages": (C.c_int, [P, C.POINTER(C.c_double)]),
        "clmi_spectrum_nbins": (C.c_int, [P, C.c_double]),
        "clmi_spectrum_bins": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                         C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "clmi_spectrum_bins_npow": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                              C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "clmi_grf_scalar": (C.c_int, [P, C.c_int, C.c_uint64, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_double]),
        "clmi_grf_u1": (C.c_int, [P, C.c_uint64, C.c_uint32, C.POINTER(C.c_double), C.c_double, C.c_double,
                                  C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "clmi_adjoint_lf": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                      C.POINTER(C.c_double), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                      C.POINTER(C.c_double), C.POINTER(Background)]),
        "clmi_ensemble_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_size_t, C.POINTER(C.c_void_p)]),
        "clmi_ensemble_groups": (C.c_int, [P]),
        "clmi_ensemble_autotune": (C.c_int, [P, C.c_int, C.c_char_p, C.c_size_t]),
        "clmi_ensemble_destroy": (None, [P]),
        "clmi_ensemble_last_error": (C.c_char_p, [P]),
        "clmi_ensemble_lf_evolve": (C.c_int, [P, C.c_int]),
        "clmi_profile_enable": (C.c_int, [P, C.c_int]),
        "clmi_profile_read": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
        "clmi_copy_calibration": (C.c_int, [P, C.c_size_t, C.c_int]),
        "clmi_autotune": (C.c_int, [P, C.c_int, C.c_char_p, C.c_size_t]),
        "clmi_plan_report": (C.c_int, [P, C.c_char_p, C.c_size_t]),
        "clmi_field_stats_noalloc": (C.c_int, [P, C.c_int, C.c_double, P, P]),
        "clmi_noalloc_last_error": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class ClmiError(RuntimeError):
    pass


def field_stats_noalloc(d_field: int, N: int, dx: float, d_out: int, stream: int = 0) -> None:
    """clmi_field_stats_noalloc: {<f>, <f^2>, <sum_i ((f(x+i) - f(x))/dx)^2>} of a caller-owned FP64
    device array of N^3 doubles into 3 caller-owned device doubles, enqueued on `stream` (device
    addresses and the stream handle as integers, e.g. a torch tensor's data_ptr() and a torch stream's
    cuda_stream).  Allocates nothing; the caller synchronises the stream before reading d_out."""
    lib = load_library()
    rc = lib.clmi_field_stats_noalloc(C.c_void_p(d_field or None), N, dx, C.c_void_p(d_out or None),
                                      C.c_void_p(stream or None))
    if rc != CLMI_OK:
        raise ClmiError(f"field_stats_noalloc: {lib.clmi_noalloc_last_error().decode()} (code {rc})")


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------------------------------------
# Models (program units).  Potentials follow models/lphi4.h:154-206 and models/lphi4U1.h:117-129.
# ---------------------------------------------------------------------------------------------

@dataclass
class ModelSpec:
    """Field content, units and potential of a model in program units."""
    name: str
    n_scalars: int = 0
    n_cscalars: int = 0
    n_u1: int = 0
    alpha: float = 1.0
    fStar: float = 1.0
    omegaStar: float = 1.0
    MPl: float = REDUCED_MPL
    gQ: float = 0.0
    terms: list = field(default_factory=list)      # [(coeff, [e0, e1, ...])]
    cs_terms: list = field(default_factory=list)   # [(coeff, power of |phi|)]
    fld0: list = field(default_factory=list)       # homogeneous field values, program units
    pi0: list = field(default_factory=list)        # homogeneous momenta, program units

    def potential(self, phi: Sequence[float]) -> float:
        v = 0.0
        for c, e in self.terms:
            p = c
            for f, ef in enumerate(e):
                p *= phi[f] ** ef
            v += p
        return v

    def cs_mass2(self, comps: Sequence[float]) -> list:
        """Effective masses^2 of the two complex-scalar components at the homogeneous point
        (Potential::deriv2CS, definitions/potential.h:36-43,83-88):
        m_c^2 = V''(|phi|) phi_c^2/|phi|^2 + V'(|phi|)/|phi| (1 - phi_c^2/|phi|^2), with the
        reference's 1e-14 floor on |phi| (abstractmodel.h:255-274)."""
        r = math.sqrt(comps[0] ** 2 + comps[1] ** 2)
        if r == 0.0:
            comps = [1e-14 * 0.70710678118654752, 1e-14 * 0.70710678118654752]
            r = math.sqrt(comps[0] ** 2 + comps[1] ** 2)
        d1 = sum(c * p * r ** (p - 1) for c, p in self.cs_terms if p >= 1)
        d2 = sum(c * p * (p - 1) * r ** (p - 2) for c, p in self.cs_terms if p >= 2)
        return [d2 * x * x / (r * r) + d1 / r * (1 - x * x / (r * r)) for x in comps]

    def mass2(self, phi: Sequence[float]) -> list:
        """Diagonal second derivatives of V at phi (the reference's potDeriv2 at the homogeneous
        initial condition, abstractmodel.h:220-248)."""
        out = []
        for n in range(self.n_scalars):
            d2 = 0.0
            for c, e in self.terms:
                en = e[n] if n < len(e) else 0
                if en < 2:
                    continue
                p = c * en * (en - 1)
                for f, ef in enumerate(e):
                    p *= phi[f] ** (ef - 2 if f == n else ef)
                d2 += p
            out.append(d2)
        return out


def lphi4(q: float, n_fields: int = 2, phi0_gev: float = 5.6964e18, lam: float = 9e-14,
          pi0_gev2: float = -4.86735e30, pi0_program: float | None = None) -> ModelSpec:
    """lphi4 (models/lphi4.h): V = phi^4/4 + q/2 phi^2 chi^2 in program units, alpha = 1,
    fStar = phi0, omegaStar = sqrt(lambda) phi0 (models/lphi4.h:125-127).  n_fields = 1 drops chi
    (the single-inflaton variant written in the same model API).  pi0_program, when given, sets the
    homogeneous momentum directly in program units."""
    fStar = phi0_gev
    omegaStar = math.sqrt(lam) * phi0_gev
    terms = [(0.25, [4, 0][:n_fields])]
    if n_fields == 2:
        terms.append((0.5 * q, [2, 2]))
    return ModelSpec(name="lphi4", n_scalars=n_fields, alpha=1.0, fStar=fStar, omegaStar=omegaStar,
                     terms=terms, fld0=[phi0_gev / fStar] + [0.0] * (n_fields - 1),
                     pi0=[pi0_gev2 / fStar / omegaStar if pi0_program is None else pi0_program]
                     + [0.0] * (n_fields - 1))


def lphi4U1(e2_over_lambda: float = 10.0, phi0_gev: float = 5.6964e18, lam: float = 9e-14,
            pi0_gev2: float = -6.88343e30, charge: float = 1.0) -> ModelSpec:
    """lphi4U1 (models/lphi4U1.h): V = |phi|^4 in program units, fStar = |phi|_0,
    omegaStar = sqrt(lambda) |phi|_0; gauge coupling g = sqrt(e2_over_lambda * lambda), effective
    charge gQ = charge * g (abstractmodel/complexscalarbase.h:71-73)."""
    fStar = phi0_gev
    omegaStar = math.sqrt(lam) * phi0_gev
    g = math.sqrt(e2_over_lambda * lam)
    c0 = phi0_gev / math.sqrt(2.0) / fStar
    p0 = pi0_gev2 / math.sqrt(2.0) / fStar / omegaStar
    return ModelSpec(name="lphi4U1", n_cscalars=1, n_u1=1, alpha=1.0, fStar=fStar, omegaStar=omegaStar,
                     gQ=charge * g, cs_terms=[(1.0, 4)], fld0=[c0, c0], pi0=[p0, p0])


def make_config(spec: ModelSpec, N: int, dx: float, dt: float, precision: int = CLMI_F64,
                expansion: bool = True, device: int = 0, flags: int = 0) -> ClmiConfig:
    cfg = ClmiConfig()
    cfg.flags = flags
    cfg.N = N
    cfg.n_scalars = spec.n_scalars
    cfg.n_cscalars = spec.n_cscalars
    cfg.n_u1 = spec.n_u1
    cfg.precision = precision
    cfg.expansion = 1 if expansion else 0
    cfg.device = device
    cfg.dx = dx
    cfg.dt = dt
    cfg.alpha = spec.alpha
    cfg.fStar = spec.fStar
    cfg.omegaStar = spec.omegaStar
    cfg.MPl = spec.MPl
    cfg.gQ = spec.gQ
    cfg.n_terms = len(spec.terms)
    for t, (c, e) in enumerate(spec.terms):
        cfg.term_coeff[t] = c
        for f, ef in enumerate(e):
            cfg.term_exp[t][f] = ef
    cfg.n_cs_terms = len(spec.cs_terms)
    for k, (c, p) in enumerate(spec.cs_terms):
        cfg.cs_coeff[k] = c
        cfg.cs_pow[k] = p
    return cfg


class Lattice:
    """One device context: the fields, momenta and background of a model on an N^3 lattice."""

    def __init__(self, spec: ModelSpec, N: int, dx: float, dt: float, precision: int = CLMI_F64,
                 expansion: bool = True, device: int = 0, flags: int = 0):
        self.lib = load_library()
        self.spec = spec
        self.N, self.dx, self.dt = N, dx, dt
        self.precision = precision
        self.expansion = expansion
        self.cfg = make_config(spec, N, dx, dt, precision, expansion, device, flags)
        h = C.c_void_p()
        rc = self.lib.clmi_create(C.byref(self.cfg), C.byref(h))
        if rc != CLMI_OK:
            raise ClmiError(f"clmi_create failed ({rc}): {self.lib.clmi_last_error(None).decode()}")
        self._h = h

    # -- plumbing -------------------------------------------------------------------------------
    def _check(self, rc: int, what: str):
        if rc != CLMI_OK:
            raise ClmiError(f"{what} failed ({rc}): {self.lib.clmi_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self.lib.clmi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    # -- data -------------------------------------------------------------------------------------
    def upload(self, kind: int, index: int, comp: int, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.size != self.N ** 3:
            raise ValueError("array must hold N^3 values")
        self._check(self.lib.clmi_upload_field(self._h, kind, index, comp, _dptr(a)), "upload")

    def download(self, kind: int, index: int, comp: int) -> np.ndarray:
        out = np.empty((self.N, self.N, self.N), dtype=np.float64)
        self._check(self.lib.clmi_download_field(self._h, kind, index, comp, _dptr(out)), "download")
        return out

    def set_background(self, **kw):
        b = self.background_struct()
        for k, v in kw.items():
            setattr(b, k, float(v))
        self._check(self.lib.clmi_set_background(self._h, C.byref(b)), "set_background")

    def background_struct(self) -> Background:
        b = Background()
        self._check(self.lib.clmi_get_background(self._h, C.byref(b)), "get_background")
        return b

    def background(self) -> dict:
        return self.background_struct().as_dict()

    # -- evolution ---------------------------------------------------------------------------------
    def lf_init(self, restart: bool = False):
        self._check(self.lib.clmi_lf_init(self._h, 1 if restart else 0), "lf_init")

    def lf_evolve(self, nsteps: int = 1):
        self._check(self.lib.clmi_lf_evolve(self._h, nsteps), "lf_evolve")

    def lf_sync(self):
        self._check(self.lib.clmi_lf_sync(self._h), "lf_sync")

    def is_synced(self) -> bool:
        return bool(self.lib.clmi_lf_is_synced(self._h))

    def vv_evolve(self, ws: Sequence[float], nsteps: int = 1):
        w = np.ascontiguousarray(ws, dtype=np.float64)
        self._check(self.lib.clmi_vv_evolve(self._h, _dptr(w), len(w), nsteps), "vv_evolve")

    def pv_evolve(self, ws: Sequence[float], nsteps: int = 1):
        w = np.ascontiguousarray(ws, dtype=np.float64)
        self._check(self.lib.clmi_pv_evolve(self._h, _dptr(w), len(w), nsteps), "pv_evolve")

    def kick(self, c: float):
        self._check(self.lib.clmi_kick(self._h, c), "kick")

    def drift(self, cdt: float):
        self._check(self.lib.clmi_drift(self._h, cdt), "drift")

    def synchronize(self):
        self._check(self.lib.clmi_synchronize(self._h), "synchronize")

    # -- observables -------------------------------------------------------------------------------
    def set_all_averages(self):
        self._check(self.lib.clmi_set_all_averages(self._h), "set_all_averages")

    def averages(self) -> dict:
        a = Averages()
        self._check(self.lib.clmi_volume_averages(self._h, C.byref(a)), "volume_averages")
        return {n: getattr(a, n) for n, _ in Averages._fields_ if n != "reserved"}

    def scale_factor_kernel(self) -> float:
        k = C.c_double()
        self._check(self.lib.clmi_scale_factor_kernel(self._h, C.byref(k)), "scale_factor_kernel")
        return k.value

    def hubble_constraint(self) -> tuple:
        out = (C.c_double * 3)()
        self._check(self.lib.clmi_hubble_constraint(self._h, out), "hubble_constraint")
        return tuple(out)

    def gauss_u1(self) -> tuple:
        out = (C.c_double * 3)()
        self._check(self.lib.clmi_gauss_u1(self._h, out), "gauss_u1")
        return tuple(out)

    def field_means(self, kind: int, index: int, comp: int = 0, scale: float = 1.0) -> tuple:
        out = (C.c_double * 2)()
        self._check(self.lib.clmi_field_means(self._h, kind, index, comp, scale, out), "field_means")
        return tuple(out)

    def field_stats(self, kind: int, index: int, comp: int = 0) -> tuple:
        out = (C.c_double * 3)()
        self._check(self.lib.clmi_field_stats(self._h, kind, index, comp, out), "field_stats")
        return tuple(out)

    def norm_averages(self, kind: int, index: int = 0) -> tuple:
        """(<|v|>, <|v|^2>) of the complex scalar (FLD_CS, PI_CS), the U(1) momenta (PI_U1) or the
        magnetic field (MAG_U1)."""
        out = (C.c_double * 2)()
        self._check(self.lib.clmi_norm_averages(self._h, kind, index, out), "norm_averages")
        return tuple(out)

    def monomial_averages(self) -> list:
        out = (C.c_double * max(1, len(self.spec.terms)))()
        self._check(self.lib.clmi_monomial_averages(self._h, out), "monomial_averages")
        return list(out)[:len(self.spec.terms)]

    def spectrum_bins(self, kind: int, index: int, comp: int = 0, dk: float = 1.0, npow: int = 0):
        """(sum over the bin's modes of |f_n|^2 |n|^npow, mode count, mean |n|) per bin."""
        nb = self.lib.clmi_spectrum_nbins(self._h, dk)
        s = np.zeros(nb)
        cnt = np.zeros(nb)
        mn = np.zeros(nb)
        if npow == 0:
            rc = self.lib.clmi_spectrum_bins(self._h, kind, index, comp, dk, nb, _dptr(s), _dptr(cnt), _dptr(mn))
        else:
            rc = self.lib.clmi_spectrum_bins_npow(self._h, kind, index, comp, dk, nb, npow, _dptr(s), _dptr(cnt),
                                                  _dptr(mn))
        self._check(rc, "spectrum_bins")
        return s, cnt, mn

    def power_spectrum(self, kind: int, index: int, comp: int = 0, dk: float = 1.0, ps_type: int = 1,
                       ps_version: int = 1) -> tuple:
        """PowerSpectrumMeasurer::powerSpectrum of one component (measurements/powerspectrum.h:85-134):
        PS_version 3 bins |n| f (PS_type 1) or |n|^3 f (PS_type 2) instead of f."""
        npow = (1 if ps_type == 1 else 3) if ps_version == 3 else 0
        s, c, m = self.spectrum_bins(kind, index, comp, dk, npow)
        return power_spectrum(s, c, m, self.N, 2 * math.pi / (self.N * self.dx), dk, ps_type, ps_version)

    def grf_scalar(self, index: int, seed: int, stream: int, mass2: float, aDot: float, kIR: float,
                   kCutOff: float, fld0: float, pi0: float):
        self._check(self.lib.clmi_grf_scalar(self._h, index, seed, stream, mass2, aDot, kIR, kCutOff, fld0, pi0),
                    "grf_scalar")

    def grf_u1(self, seed: int, stream: int, mass2: Sequence[float], aDot: float, kIR: float, kCutOff: float,
               fld0: Sequence[float], pi0: Sequence[float]):
        """Complex scalar + U(1) vacuum initial conditions with the Gauss constraint (clmi_grf_u1)."""
        m = (C.c_double * 2)(*mass2)
        f = (C.c_double * 2)(*fld0)
        p = (C.c_double * 2)(*pi0)
        self._check(self.lib.clmi_grf_u1(self._h, seed, stream, m, aDot, kIR, kCutOff, f, p), "grf_u1")

    def adjoint_lf(self, nsteps: int, seed_fld: Sequence[np.ndarray], seed_pi: Sequence[np.ndarray],
                   seed_bg: Sequence[float] = (0.0, 0.0), checkpoint_every: int = 0, bg_from_fields: bool = True):
        """Reverse-mode derivative of [initial background] -> lf_init -> nsteps steps -> sync from the
        current state (clmi_adjoint_lf).  Returns (dL/dphi_0 per field, dL/dpi_0 per field,
        dL/dterm_coeff, dL/d initial background dict); the lattice ends at the forward's final state."""
        nf = self.spec.n_scalars
        V = self.N ** 3
        seeds_f = [np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1)) for a in seed_fld]
        seeds_p = [np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1)) for a in seed_pi]
        if len(seeds_f) != nf or len(seeds_p) != nf or any(a.size != V for a in seeds_f + seeds_p):
            raise ValueError("one N^3 seed array per scalar field is required")
        gf = [np.empty(V) for _ in range(nf)]
        gp = [np.empty(V) for _ in range(nf)]
        ptrs = lambda arrs: (C.c_void_p * nf)(*[a.ctypes.data for a in arrs])  # noqa: E731
        sb = (C.c_double * 2)(*seed_bg)
        gc = (C.c_double * max(1, len(self.spec.terms)))()
        gb = Background()
        self._check(self.lib.clmi_adjoint_lf(self._h, nsteps, checkpoint_every, 1 if bg_from_fields else 0,
                                             ptrs(seeds_f), ptrs(seeds_p), sb, ptrs(gf), ptrs(gp), gc, C.byref(gb)),
                    "adjoint_lf")
        shape = (self.N,) * 3
        return ([a.reshape(shape) for a in gf], [a.reshape(shape) for a in gp], list(gc)[:len(self.spec.terms)],
                gb.as_dict())

    # -- timing ------------------------------------------------------------------------------------
    def profile_enable(self, on: bool = True, every: int = 1):
        """HIP-event timing of every `every`-th drift + kick pass (clmi_profile_enable); off with on=False."""
        self._check(self.lib.clmi_profile_enable(self._h, max(1, int(every)) if on else 0), "profile_enable")

    def profile_read(self) -> tuple:
        ms = C.c_double()
        n = C.c_int64()
        b = C.c_double()
        self._check(self.lib.clmi_profile_read(self._h, C.byref(ms), C.byref(n), C.byref(b)), "profile_read")
        return ms.value, n.value, b.value

    def autotune(self, reps: int = 5) -> str:
        buf = C.create_string_buffer(4096)
        self._check(self.lib.clmi_autotune(self._h, reps, buf, len(buf)), "autotune")
        return buf.value.decode()

    def plan(self) -> str:
        """The lattice passes this context runs, one line per sector (clmi_plan_report)."""
        buf = C.create_string_buffer(1024)
        self._check(self.lib.clmi_plan_report(self._h, buf, len(buf)), "plan_report")
        return buf.value.decode()

    def copy_calibration(self, nbytes: int, iters: int = 3):
        self._check(self.lib.clmi_copy_calibration(self._h, nbytes, iters), "copy_calibration")


def power_spectrum(sum_f2, count, mean_n, N: int, kIR: float, dk: float = 1.0, ps_type: int = 1,
                   ps_version: int = 1) -> tuple:
    """PowerSpectrumMeasurer::powerSpectrum normalisation (measurements/powerspectrum.h:85-134) applied
    to the binned sums: returns (k values = n_b * kIR, spectrum).  PS_version 1 uses each bin's
    mean |n|, PS_version 2 the bin's nominal centre b * dk.  PS_version 3 takes the |n|^p-weighted sums
    (p = 1 for PS_type 1, 3 for PS_type 2, from clmi_spectrum_bins_npow): the |n| factor is then inside
    the bin sum, and the prefactor has neither the bin's |n| nor 1/dk (powerspectrum.h:111-131)."""
    sum_f2 = np.asarray(sum_f2)
    count = np.asarray(count)
    nb = len(sum_f2)
    nominal = np.arange(nb) * dk
    dx = 2 * math.pi / kIR / N
    if ps_version == 3:
        if ps_type == 1:
            ps = kIR * dx / N ** 5 / 2 / math.pi * sum_f2
        elif ps_type == 2:
            mean = np.where(count > 0, sum_f2 / np.maximum(count, 1), 0.0)
            ps = (kIR * dx / N) ** 3 / 2 / math.pi ** 2 * mean
        else:
            raise ValueError("ps_type must be 1 or 2")
        return nominal * kIR, ps
    if ps_version not in (1, 2):
        raise ValueError("ps_version must be 1, 2 or 3")
    nt = np.asarray(mean_n) if ps_version == 1 else nominal
    if ps_type == 1:
        ps = kIR * nt * dx / dk / N ** 5 / 2 / math.pi * sum_f2
    elif ps_type == 2:
        mean = np.where(count > 0, sum_f2 / np.maximum(count, 1), 0.0)
        ps = (kIR * nt * dx / N) ** 3 / dk / 2 / math.pi ** 2 * mean
    else:
        raise ValueError("ps_type must be 1 or 2")
    return nt * kIR, ps


class Ensemble:
    """Independent lattices (N = 64, 128, 192; same precision, field count and expansion) advanced
    together: one batched launch per leapfrog step, replayed from a captured HIP graph
    (clmi_ensemble_*).  The member Lattices stay usable for initial conditions and observations."""

    def __init__(self, members: Sequence[Lattice], group_bytes: int = 0):
        self.lib = load_library()
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.handle.value for m in self.members])
        h = C.c_void_p()
        rc = self.lib.clmi_ensemble_create(arr, len(self.members), group_bytes, C.byref(h))
        if rc != CLMI_OK:
            raise ClmiError(f"clmi_ensemble_create failed ({rc}): {self.lib.clmi_ensemble_last_error(None).decode()}")
        self._h = h

    def lf_evolve(self, nsteps: int = 1):
        rc = self.lib.clmi_ensemble_lf_evolve(self._h, nsteps)
        if rc != CLMI_OK:
            raise ClmiError(f"clmi_ensemble_lf_evolve failed ({rc}): "
                            f"{self.lib.clmi_ensemble_last_error(self._h).decode()}")

    @property
    def groups(self) -> int:
        return self.lib.clmi_ensemble_groups(self._h)

    def autotune(self, reps: int = 3) -> str:
        """Pick each group's x-chunk by timing its batched pass on the live data (no state change)."""
        buf = C.create_string_buffer(4096)
        rc = self.lib.clmi_ensemble_autotune(self._h, reps, buf, len(buf))
        if rc != CLMI_OK:
            raise ClmiError(f"clmi_ensemble_autotune failed ({rc}): "
                            f"{self.lib.clmi_ensemble_last_error(self._h).decode()}")
        return buf.value.decode()

    def close(self):
        if getattr(self, "_h", None):
            self.lib.clmi_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LeapFrog:
    """LeapFrog evolver (include/CosmoInterface/evolvers/leapfrog.h:23-288) over a Lattice."""

    def __init__(self, lat: Lattice, restart: bool = False):
        self.lat = lat
        lat.lf_init(restart)

    def evolve(self, nsteps: int = 1):
        self.lat.lf_evolve(nsteps)

    def sync(self):
        self.lat.lf_sync()

    @property
    def synced(self) -> bool:
        return self.lat.is_synced()


# Symmetric composition weights w[0..n-1] of the Verlet schemes; a step runs the stages w[0], ...,
# w[n-1], ..., w[0].  Velocity Verlet (evolvers/velocityverletparameters.h:23-47) and position Verlet
# (evolvers/positionverletparameters.h:22-42) use the same published coefficients: Yoshida (1990) for
# orders 4 and 6, Kahan and Li (1997) for 6 (alternative), 8 and 10.
_COMPOSITION = {
    "2": [1.0],
    "4": [1.351207191959657771818, -1.702414403875838200264],
    "6": [0.78451361047755726382, 0.23557321335935813368, -1.1776799841788710069, 1.3151863206839112189],
    "6_2": [0.39216144400731413928, 0.33259913678935943860, -0.70624617255763935981, 0.082213596293550800230,
            0.79854399093482996340],
    "8": [0.74167036435061295345, -0.40910082580003159400, 0.19075471029623837995, -0.57386247111608226666,
          0.29906418130365592384, 0.33462491824529818378, 0.31529309239676659663, -0.79688793935291635402],
    "10": [-0.48159895600253002870, 0.0036303931544595926879, 0.50180317558723140279, 0.28298402624506254868,
           0.80702967895372223806, -0.026090580538592205447, -0.87286590146318071547, -0.52373568062510581643,
           0.44521844299952789252, 0.18612289547097907887, 0.23137327866438360633, -0.52191036590418628905,
           0.74866113714499296793, 0.066736511890604057532, -0.80360324375670830316, 0.91249037635867994571],
}
VV_WEIGHTS = {"VV" + k: v for k, v in _COMPOSITION.items()}
PV_WEIGHTS = {"PV" + k: v for k, v in _COMPOSITION.items()}


class VelocityVerlet:
    """VelocityVerlet evolver (include/CosmoInterface/evolvers/velocityverlet.h:33-278)."""

    def __init__(self, lat: Lattice, kind: str = "VV2"):
        self.lat = lat
        self.ws = VV_WEIGHTS[kind]

    def evolve(self, nsteps: int = 1):
        self.lat.vv_evolve(self.ws, nsteps)

    def sync(self):
        pass  # fields and momenta live at integer times (velocityverlet.h:119-122)


class PositionVerlet:
    """PositionVerlet evolver (include/CosmoInterface/evolvers/positionverlet.h:29-296)."""

    def __init__(self, lat: Lattice, kind: str = "PV2"):
        self.lat = lat
        self.ws = PV_WEIGHTS[kind]

    def evolve(self, nsteps: int = 1):
        self.lat.pv_evolve(self.ws, nsteps)

    def sync(self):
        pass  # only a fixed background needs it (positionverlet.h:115-119); not supported here
