"""FP64 CPU oracle for the lattice-evolution hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module, and
only as the checker (or the timed CPU baseline), never as the thing measured or shipped.

Two parts:
* ctypes wrapper of oracle/_build/liboracle.so (clmi_oracle.c): the reference-structured C
  restatement of the leapfrog / velocity-Verlet evolution, volume averages, scale factor, Friedmann
  and Gauss checks (one lattice sweep per reference expression).
* numpy restatements of the FFT-side functions: |n|-shell binning of power spectra
  (measurements/powerspectrum.h:85-134) and the Gaussian-random-field initial conditions
  (initializers/fluctuationsgenerator.h:39-71,103-150) with the Philox4x32-10 generator.

Pinning (tests/test_oracle.py): the bins by Parseval, mode counts and a single plane wave; the GRF
by its shell statistics, Hermitian structure, cutoff and homogeneous values; Philox by the published
known-answer vectors.  Parity unpinned against the reference (TempLat could not be built or
imported): the random stream itself (TempLat's generator is replaced by Philox4x32-10 with a
documented key/counter).

Reference files are cited relative to the CosmoLattice source tree.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "_build" / "liboracle.so"
MAX_SCALARS = 4
MAX_TERMS = 8


# Symmetric composition weights of the position-Verlet schemes (evolvers/positionverletparameters.h:22-42;
# Yoshida 1990 for orders 4 and 6, Kahan and Li 1997 for 6_2, 8 and 10).  A step runs w[0..n-1] then
# w[n-2..0].
PV_WEIGHTS = {
    "PV2": [1.0],
    "PV4": [1.351207191959657771818, -1.702414403875838200264],
    "PV6": [0.78451361047755726382, 0.23557321335935813368, -1.1776799841788710069, 1.3151863206839112189],
    "PV6_2": [0.39216144400731413928, 0.33259913678935943860, -0.70624617255763935981, 0.082213596293550800230,
              0.79854399093482996340],
    "PV8": [0.74167036435061295345, -0.40910082580003159400, 0.19075471029623837995, -0.57386247111608226666,
            0.29906418130365592384, 0.33462491824529818378, 0.31529309239676659663, -0.79688793935291635402],
    "PV10": [-0.48159895600253002870, 0.0036303931544595926879, 0.50180317558723140279, 0.28298402624506254868,
             0.80702967895372223806, -0.026090580538592205447, -0.87286590146318071547, -0.52373568062510581643,
             0.44521844299952789252, 0.18612289547097907887, 0.23137327866438360633, -0.52191036590418628905,
             0.74866113714499296793, 0.066736511890604057532, -0.80360324375670830316, 0.91249037635867994571],
}


class OrcBackground(C.Structure):
    _fields_ = [(n, C.c_double) for n in [
        "aI", "aSI", "aIM", "aDotI", "aDotSI", "aDotSIM",
        "pi2AvI", "pi2AvSI", "pi2AvSIM", "grad2AvI",
        "CSpi2AvI", "CSpi2AvSI", "CSpi2AvSIM", "CSgrad2AvI",
        "U1pi2AvI", "U1pi2AvSI", "U1pi2AvSIM", "U1Mag2AvI",
        "potAvI", "potAvSI"]]


DP = C.POINTER(C.c_double)


class OrcModel(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("ns", C.c_int), ("ncs", C.c_int), ("nu1", C.c_int),
        ("expansion", C.c_int), ("synced", C.c_int),
        ("dx", C.c_double), ("dt", C.c_double),
        ("alpha", C.c_double), ("fStar", C.c_double), ("omegaStar", C.c_double), ("MPl", C.c_double),
        ("nterms", C.c_int),
        ("coeff", C.c_double * MAX_TERMS),
        ("expo", (C.c_int * MAX_SCALARS) * MAX_TERMS),
        ("ncsterms", C.c_int),
        ("cscoeff", C.c_double * MAX_TERMS),
        ("cspow", C.c_int * MAX_TERMS),
        ("gQ", C.c_double),
        ("fldS", DP * MAX_SCALARS), ("piS", DP * MAX_SCALARS),
        ("fldCS", DP * 2), ("piCS", DP * 2),
        ("fldU1", DP * 3), ("piU1", DP * 3),
        ("bg", OrcBackground),
    ]


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
    lib = C.CDLL(str(LIB_PATH))
    M = C.POINTER(OrcModel)
    for name, res, args in [
        ("orc_lf_init", None, [M]), ("orc_lf_evolve", None, [M]), ("orc_lf_sync", None, [M]),
        ("orc_vv_evolve", None, [M, DP, C.c_int]), ("orc_pv_evolve", None, [M, DP, C.c_int]),
        ("orc_kick", None, [M, C.c_double]), ("orc_drift", None, [M, C.c_double]),
        ("orc_set_all_averages", None, [M]),
        ("orc_avg_pi2S", C.c_double, [M]), ("orc_avg_grad2S", C.c_double, [M]),
        ("orc_avg_potential", C.c_double, [M]), ("orc_avg_pi2CS", C.c_double, [M]),
        ("orc_avg_grad2CS", C.c_double, [M]), ("orc_avg_pi2U1", C.c_double, [M]),
        ("orc_avg_B2U1", C.c_double, [M]),
        ("orc_scale_factor_kernel", C.c_double, [M]), ("orc_rho", C.c_double, [M]),
        ("orc_hubble_constraint", None, [M, DP]), ("orc_gauss_u1", None, [M, DP]),
        ("orc_laplacian", None, [C.c_int, C.c_double, DP, DP]),
        ("orc_scalar_potential", C.c_double, [M, DP]),
        ("orc_scalar_potential_deriv", C.c_double, [M, DP, C.c_int]),
        ("orc_num_threads", C.c_int, []), ("orc_set_num_threads", None, [C.c_int]),
        ("orc_sizeof_model", C.c_size_t, []),
    ]:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.orc_sizeof_model() != C.sizeof(OrcModel):
        raise RuntimeError("orc_model layout mismatch between clmi_oracle.c and oracle.py")
    _lib = lib
    return lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(DP)


class OracleLattice:
    """CPU FP64 lattice with the same model description as cosmolattice_amd.clmi.ModelSpec."""

    def __init__(self, spec, N: int, dx: float, dt: float, expansion: bool = True):
        self.lib = load()
        self.spec = spec
        self.N, self.dx, self.dt = N, dx, dt
        m = OrcModel()
        m.N = N
        m.ns, m.ncs, m.nu1 = spec.n_scalars, spec.n_cscalars, spec.n_u1
        m.expansion = 1 if expansion else 0
        m.synced = 1
        m.dx, m.dt = dx, dt
        m.alpha, m.fStar, m.omegaStar, m.MPl = spec.alpha, spec.fStar, spec.omegaStar, spec.MPl
        m.nterms = len(spec.terms)
        for t, (c, e) in enumerate(spec.terms):
            m.coeff[t] = c
            for f, ef in enumerate(e):
                m.expo[t][f] = ef
        m.ncsterms = len(spec.cs_terms)
        for k, (c, p) in enumerate(spec.cs_terms):
            m.cscoeff[k] = c
            m.cspow[k] = p
        m.gQ = spec.gQ
        shape = (N, N, N)
        self.arrays = {}
        def mk(key):
            a = np.zeros(shape)
            self.arrays[key] = a
            return a
        for f in range(spec.n_scalars):
            m.fldS[f] = _p(mk(("fldS", f, 0)))
            m.piS[f] = _p(mk(("piS", f, 0)))
        for c in range(2 * spec.n_cscalars):
            m.fldCS[c] = _p(mk(("fldCS", 0, c)))
            m.piCS[c] = _p(mk(("piCS", 0, c)))
        for d in range(3 * spec.n_u1):
            m.fldU1[d] = _p(mk(("fldU1", 0, d)))
            m.piU1[d] = _p(mk(("piU1", 0, d)))
        m.bg.aI = m.bg.aSI = m.bg.aIM = 1.0
        self.m = m

    _KIND = {0: "fldS", 1: "piS", 2: "fldCS", 3: "piCS", 6: "fldU1", 7: "piU1"}

    def arr(self, kind: int, index: int, comp: int = 0) -> np.ndarray:
        return self.arrays[(self._KIND[kind], index, comp)]

    def upload(self, kind: int, index: int, comp: int, a: np.ndarray):
        self.arr(kind, index, comp)[...] = np.asarray(a, dtype=np.float64).reshape(self.N, self.N, self.N)

    def download(self, kind: int, index: int, comp: int = 0) -> np.ndarray:
        return self.arr(kind, index, comp).copy()

    def background(self) -> dict:
        return {n: getattr(self.m.bg, n) for n, _ in OrcBackground._fields_}

    def set_background(self, **kw):
        for k, v in kw.items():
            setattr(self.m.bg, k, float(v))

    @property
    def synced(self) -> bool:
        return bool(self.m.synced)

    def lf_init(self):
        self.lib.orc_lf_init(C.byref(self.m))

    def lf_evolve(self, nsteps: int = 1):
        for _ in range(nsteps):
            self.lib.orc_lf_evolve(C.byref(self.m))

    def lf_sync(self):
        self.lib.orc_lf_sync(C.byref(self.m))

    def vv_evolve(self, ws, nsteps: int = 1):
        w = np.ascontiguousarray(ws, dtype=np.float64)
        for _ in range(nsteps):
            self.lib.orc_vv_evolve(C.byref(self.m), _p(w), len(w))

    def pv_evolve(self, ws, nsteps: int = 1):
        w = np.ascontiguousarray(ws, dtype=np.float64)
        for _ in range(nsteps):
            self.lib.orc_pv_evolve(C.byref(self.m), _p(w), len(w))

    def kick(self, c: float):
        self.lib.orc_kick(C.byref(self.m), c)

    def drift(self, cdt: float):
        self.lib.orc_drift(C.byref(self.m), cdt)

    def set_all_averages(self):
        self.lib.orc_set_all_averages(C.byref(self.m))

    def averages(self) -> dict:
        L, m = self.lib, C.byref(self.m)
        out = {"pot": L.orc_avg_potential(m)}
        out["pi2S"] = L.orc_avg_pi2S(m) if self.m.ns else 0.0
        out["grad2S"] = L.orc_avg_grad2S(m) if self.m.ns else 0.0
        out["pi2CS"] = L.orc_avg_pi2CS(m) if self.m.ncs else 0.0
        out["grad2CS"] = L.orc_avg_grad2CS(m) if self.m.ncs else 0.0
        out["pi2U1"] = L.orc_avg_pi2U1(m) if self.m.nu1 else 0.0
        out["B2U1"] = L.orc_avg_B2U1(m) if self.m.nu1 else 0.0
        return out

    def scale_factor_kernel(self) -> float:
        return self.lib.orc_scale_factor_kernel(C.byref(self.m))

    def hubble_constraint(self) -> tuple:
        out = (C.c_double * 3)()
        self.lib.orc_hubble_constraint(C.byref(self.m), out)
        return tuple(out)

    def gauss_u1(self) -> tuple:
        out = (C.c_double * 3)()
        self.lib.orc_gauss_u1(C.byref(self.m), out)
        return tuple(out)


def laplacian(f: np.ndarray, dx: float) -> np.ndarray:
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.empty_like(f)
    load().orc_laplacian(f.shape[0], dx, _p(f), _p(out))
    return out


def set_threads(n: int):
    load().orc_set_num_threads(n)


def num_threads() -> int:
    return load().orc_num_threads()


# ---------------------------------------------------------------------------------------------
# Power spectrum binning (numpy)
# ---------------------------------------------------------------------------------------------

def spectrum_nbins(N: int, dk: float) -> int:
    """floor(sqrt(3)/2 N / dk) bins of runparameters.h:97-98, plus bin 0 for the zero mode."""
    return int(math.floor(math.sqrt(3.0) / 2.0 * N / dk)) + 1


def mode_norms(N: int) -> np.ndarray:
    """|n| of every full-space FFT mode, n components in (-N/2, N/2]."""
    n = np.fft.fftfreq(N, d=1.0 / N)
    n = np.where(n == -N // 2, N // 2, n)
    nx, ny, nz = np.meshgrid(n, n, n, indexing="ij")
    return np.sqrt(nx * nx + ny * ny + nz * nz)


def magnetic_field(A, dx: float):
    """B_i = (d+_j A_k - d+_k A_j)/dx with (i, j, k) cyclic and forward differences, so that
    |B|^2 = F_12^2 + F_13^2 + F_23^2 with the field strength F_ij of definitions/fieldfunctionals.h:140-153
    (TempLat's magneticField as measurements/u1measurer.h:72,95-97 uses it)."""
    def fd(f, ax):
        return np.roll(f, -1, axis=ax) - f
    return [(fd(A[(i + 2) % 3], (i + 1) % 3) - fd(A[(i + 1) % 3], (i + 2) % 3)) / dx for i in range(3)]


def spectrum_bins(f: np.ndarray, dk: float = 1.0, npow: int = 0):
    """|f_n|^2 summed over |n|-shells of the unnormalised forward DFT (full N^3 mode set); returns
    (sum |f_n|^2, number of modes, mean |n|) per bin, bin b = min(floor(|n|/dk + 1/2), nbins-1).
    npow > 0 weights each mode by |n|^npow inside the sum: the projections of |n| f and |n|^3 f that
    PS_version 3 bins (measurements/powerspectrum.h:111-131)."""
    N = f.shape[0]
    nb = spectrum_nbins(N, dk)
    fk = np.fft.fftn(f)
    p = (fk.real ** 2 + fk.imag ** 2).ravel()
    nn = mode_norms(N).ravel()
    if npow:
        p = p * nn ** npow
    b = np.minimum(np.floor(nn / dk + 0.5).astype(np.int64), nb - 1)
    s = np.bincount(b, weights=p, minlength=nb)
    cnt = np.bincount(b, minlength=nb).astype(np.float64)
    sn = np.bincount(b, weights=nn, minlength=nb)
    mean = np.where(cnt > 0, sn / np.maximum(cnt, 1), 0.0)
    return s, cnt, mean


# ---------------------------------------------------------------------------------------------
# Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and Gaussian random fields (numpy)
# ---------------------------------------------------------------------------------------------

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0: int, k1: int):
    """Vectorised Philox4x32-10; counters are uint32-valued arrays, key two uint32 ints."""
    c0, c1, c2, c3 = (np.asarray(x, dtype=np.uint64) & _MASK for x in (c0, c1, c2, c3))
    k0 &= 0xFFFFFFFF
    k1 &= 0xFFFFFFFF
    for r in range(10):
        if r > 0:
            k0 = (k0 + _W0) & 0xFFFFFFFF
            k1 = (k1 + _W1) & 0xFFFFFFFF
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c0, c1, c2, c3 = hi1 ^ c1 ^ np.uint64(k0), lo1, hi0 ^ c3 ^ np.uint64(k1), lo0
    return c0, c1, c2, c3


def _u53(a, b):
    return ((a >> np.uint64(5)).astype(np.float64) * 67108864.0 + (b >> np.uint64(6)).astype(np.float64)) \
        * (1.0 / 9007199254740992.0)


def _normals2(seed: int, stream: int, m, draw: int):
    z = np.zeros_like(m)
    r0, r1, r2, r3 = philox4x32_10(m, z + draw, z + stream, z, seed & 0xFFFFFFFF, seed >> 32)
    u1, u2 = _u53(r0, r1), _u53(r2, r3)
    rad = np.sqrt(-2.0 * np.log(1.0 - u1))
    th = 6.283185307179586477 * u2
    return rad * np.cos(th), rad * np.sin(th)


def n2max_for_cut(kIR: float, kCut: float, N: int) -> int:
    """Largest |n|^2 with sqrt(|n|^2) kIR <= kCut (modes beyond are cut; H(0) = 1)."""
    r = kCut / kIR
    if not r >= 0:
        return -1
    if r > 4.0 * N:
        return 3 * N * N
    n2 = int(math.floor(r * r))
    while math.sqrt(n2 + 1) * kIR <= kCut:
        n2 += 1
    while n2 >= 0 and math.sqrt(n2) * kIR > kCut:
        n2 -= 1
    return n2


def grf_modes(N: int, seed: int, stream: int, mass2: float, aDot: float, kIR: float, kCut: float,
              sigma_pref: float):
    """Half-complex (N, N, N/2+1) Fourier modes of phi and pi for one scalar field
    (fluctuationsgenerator.h:129-150) with the Hermitian completion described in the HIP kernel."""
    Nh = N // 2 + 1
    kx, ky, kz = np.meshgrid(np.arange(N), np.arange(N), np.arange(Nh), indexing="ij")
    m = ((kx * N + ky) * Nh + kz).astype(np.uint64)
    nx = np.where(kx <= N // 2, kx, kx - N)
    ny = np.where(ky <= N // 2, ky, ky - N)
    n2 = nx * nx + ny * ny + kz * kz
    k = np.sqrt(n2.astype(np.float64)) * kIR
    omega = np.sqrt(k * k + max(mass2, 0.0))
    with np.errstate(divide="ignore"):
        sigma = sigma_pref * np.power(2 * omega, -0.5)
    edge = (kz == 0) | (2 * kz == N)
    pm = ((((N - kx) % N) * N + (N - ky) % N) * Nh + kz).astype(np.uint64)
    self_conj = edge & (pm == m)
    partner = edge & (pm < m)
    src = np.where(partner, pm, m)
    gLr, gLi = _normals2(seed, stream, src, 0)
    gRr, gRi = _normals2(seed, stream, src, 1)
    fL = sigma * (gLr + 1j * gLi)
    fR = sigma * (gRr + 1j * gRi)
    fL2 = np.where(partner, np.conj(fR), fL)
    fR2 = np.where(partner, np.conj(fL), fR)
    fR2 = np.where(self_conj, np.conj(fL2), fR2)
    is2 = 0.70710678118654752440
    phik = (fL2 + fR2) * is2
    pik = 1j * omega * (fL2 - fR2) * is2 - aDot * phik
    keep = (m != 0) & (n2 <= n2max_for_cut(kIR, kCut, N))
    phik = np.where(keep, phik, 0)
    pik = np.where(keep, pik, 0)
    return phik, pik


def grf_scalar(N: int, dx: float, spec, seed: int, stream: int, mass2: float, aDot: float, kIR: float,
               kCut: float, fld0: float, pi0: float):
    """Real-space phi, pi of one scalar: inverse FFT of grf_modes (1/N^3 normalisation) plus the
    homogeneous values (scalarsingletinitializer.h:283-284)."""
    L = N * dx
    sigma_pref = spec.omegaStar / spec.fStar * (L / dx ** 2) ** 1.5 / math.sqrt(2.0)
    phik, pik = grf_modes(N, seed, stream, mass2, aDot, kIR, kCut, sigma_pref)
    phi = np.fft.irfftn(phik, s=(N, N, N), axes=(0, 1, 2)) + fld0
    pi = np.fft.irfftn(pik, s=(N, N, N), axes=(0, 1, 2)) + pi0
    return phi, pi


# ---------------------------------------------------------------------------------------------
# Complex scalar + U(1) initial conditions (initializers/u1initializer.h:77-117, 157-230)
# ---------------------------------------------------------------------------------------------

def _uniforms6(seed: int, stream: int, m):
    z = np.zeros_like(m)
    out = []
    for d in range(3):
        r0, r1, r2, r3 = philox4x32_10(m, z + d, z + stream, z, seed & 0xFFFFFFFF, seed >> 32)
        out += [_u53(r0, r1), _u53(r2, r3)]
    return out


def grf_cs_modes(N: int, seed: int, stream: int, mass2, aDot: float, kIR: float, kCut: float, sigma_pref: float):
    """Half-complex modes (phi_0, phi_1, pi_0, pi_1) of the complex scalar: Rayleigh amplitudes,
    phases e00, e01, e10 and the constrained e11 = e01 e10 conj(e00) (u1initializer.h:157-230);
    the mode -n on the kz = 0, N/2 planes uses the draws of n with (e_c0, e_c1) -> (conj e_c1,
    conj e_c0); self-conjugate modes use e01 = conj e00, e10 = e00."""
    Nh = N // 2 + 1
    kx, ky, kz = np.meshgrid(np.arange(N), np.arange(N), np.arange(Nh), indexing="ij")
    m = ((kx * N + ky) * Nh + kz).astype(np.uint64)
    nx = np.where(kx <= N // 2, kx, kx - N)
    ny = np.where(ky <= N // 2, ky, ky - N)
    n2 = nx * nx + ny * ny + kz * kz
    k = np.sqrt(n2.astype(np.float64)) * kIR
    edge = (kz == 0) | (2 * kz == N)
    pm = ((((N - kx) % N) * N + (N - ky) % N) * Nh + kz).astype(np.uint64)
    self_conj = edge & (pm == m)
    partner = edge & (pm < m)
    src = np.where(partner, pm, m)
    u = _uniforms6(seed, stream, src)
    e00 = np.exp(2j * np.pi * u[2])
    e01 = np.exp(2j * np.pi * u[3])
    e10 = np.exp(2j * np.pi * u[4])
    e01 = np.where(self_conj, np.conj(e00), e01)
    e10 = np.where(self_conj, e00, e10)
    e11 = e01 * e10 * np.conj(e00)
    omega0 = np.sqrt(k * k + max(mass2[0], 0.0))
    keep = (m != 0) & (n2 <= n2max_for_cut(kIR, kCut, N))
    out = []
    for comp, (ea, eb) in enumerate(((e00, e01), (e10, e11))):
        with np.errstate(divide="ignore"):
            sigma = sigma_pref * np.power(2 * np.sqrt(k * k + max(mass2[comp], 0.0)), -0.5)
        amp = sigma * np.sqrt(-2.0 * np.log(1.0 - u[comp]))
        c0 = np.where(partner, np.conj(eb), ea)
        c1 = np.where(partner, np.conj(ea), eb)
        phik = 0.5 * amp * (c0 + c1)
        pik = -1j * omega0 * 0.5 * amp * (c0 - c1) - aDot * phik
        out.append((np.where(keep, phik, 0), np.where(keep, pik, 0)))
    return out


def gauss_solve_modes(j0k, N: int):
    """pi_A,i modes from the Gauss constraint: conj(k-_i)/|k-|^2 j0, k-_i = 1 - exp(-2 pi i n_i/N)
    (u1initializer.h:86-104); zero mode zero."""
    Nh = N // 2 + 1
    n = np.meshgrid(np.arange(N), np.arange(N), np.arange(Nh), indexing="ij")
    km = [1.0 - np.exp(-2j * np.pi * ni / N) for ni in n]
    k2 = sum(np.abs(x) ** 2 for x in km)
    k2[0, 0, 0] = 1.0
    out = []
    for i in range(3):
        p = np.conj(km[i]) / k2 * j0k
        p[0, 0, 0] = 0
        out.append(p)
    return out


def grf_u1(N: int, dx: float, spec, seed: int, stream: int, mass2, aDot: float, kIR: float, kCut: float, fld0,
           pi0):
    """Real-space complex-scalar components (phi_0, phi_1, pi_0, pi_1) and pi_A (3 components) of
    the U(1) initial conditions; A = 0."""
    L = N * dx
    sigma_pref = spec.omegaStar / spec.fStar * (L / dx ** 2) ** 1.5 / math.sqrt(2.0)
    modes = grf_cs_modes(N, seed, stream, mass2, aDot, kIR, kCut, sigma_pref)
    s = (N, N, N)
    phi = [np.fft.irfftn(modes[c][0], s=s, axes=(0, 1, 2)) + fld0[c] for c in range(2)]
    pi = [np.fft.irfftn(modes[c][1], s=s, axes=(0, 1, 2)) + pi0[c] for c in range(2)]
    fw = spec.fStar / spec.omegaStar
    rho = 2.0 / dx * fw * fw * spec.gQ * (pi[0] * phi[1] - pi[1] * phi[0])
    j0k = np.fft.rfftn(-dx * dx * rho, axes=(0, 1, 2))
    piA = [np.fft.irfftn(p, s=s, axes=(0, 1, 2)) for p in gauss_solve_modes(j0k, N)]
    return phi, pi, piA


# ---------------------------------------------------------------------------------------------
# Reverse-mode derivative of the leapfrog run (scalar singlets), FP64 numpy
# ---------------------------------------------------------------------------------------------

def _lap_np(f, dx):
    return sum(np.roll(f, 1, a) + np.roll(f, -1, a) - 2 * f for a in range(3)) / (dx * dx)


def _monomials(spec, phi):
    """Per term t: value m_t, first derivatives dm_t/dphi_f, second derivatives (lists of arrays)."""
    nf = spec.n_scalars
    out = []
    for _, e in spec.terms:
        e = list(e) + [0] * (nf - len(e))
        m = np.ones_like(phi[0])
        for f in range(nf):
            m = m * phi[f] ** e[f]
        d1 = []
        for n in range(nf):
            if e[n] == 0:
                d1.append(np.zeros_like(phi[0]))
                continue
            p = np.full_like(phi[0], float(e[n]))
            for f in range(nf):
                p = p * phi[f] ** (e[f] - 1 if f == n else e[f])
            d1.append(p)
        d2 = [[None] * nf for _ in range(nf)]
        for n in range(nf):
            for k in range(nf):
                if n == k:
                    if e[n] < 2:
                        d2[n][k] = np.zeros_like(phi[0])
                        continue
                    p = np.full_like(phi[0], float(e[n] * (e[n] - 1)))
                    for f in range(nf):
                        p = p * phi[f] ** (e[f] - 2 if f == n else e[f])
                else:
                    if e[n] == 0 or e[k] == 0:
                        d2[n][k] = np.zeros_like(phi[0])
                        continue
                    p = np.full_like(phi[0], float(e[n] * e[k]))
                    for f in range(nf):
                        p = p * phi[f] ** (e[f] - 1 if f in (n, k) else e[f])
                d2[n][k] = p
        out.append((m, d1, d2))
    return out


def _sf_partials(spec, a, pi2, G, U):
    al, fm = spec.alpha, spec.fStar / spec.MPl
    A = a ** (2 * al + 1) / 3.0 * fm * fm
    B = (al - 2) * 0.5 * pi2 * a ** -6 + al * 0.5 * G * a ** -2 + (al + 1) * U
    return dict(K=A * B, dpi2=A * (al - 2) * 0.5 * a ** -6, dG=A * al * 0.5 * a ** -2, dU=A * (al + 1),
                da=(2 * al + 1) / a * A * B + A * ((al - 2) * 0.5 * pi2 * -6 * a ** -7 + al * 0.5 * G * -2 * a ** -3))


def adjoint_lf(spec, N, dx, dt, phi0, pi0, nsteps, seed_fld, seed_pi, seed_bg=(0.0, 0.0), expansion=True):
    """FP64 restatement of the reverse-mode leapfrog: forward map = setAllAverages + aDotI from the
    Friedmann constraint, LeapFrog constructor, nsteps steps, sync (on the C oracle, every state
    kept); returns (dL/dphi_0, dL/dpi_0, dL/dcoeff) for the given output seeds."""
    nf = spec.n_scalars
    V = float(N ** 3)
    al = spec.alpha
    orc = OracleLattice(spec, N, dx, dt, expansion=expansion)
    for f in range(nf):
        orc.upload(0, f, 0, phi0[f])
        orc.upload(1, f, 0, pi0[f])
    if expansion:
        _, _, rhs = orc.hubble_constraint()
        orc.set_background(aDotI=math.sqrt(rhs))
    else:
        orc.set_all_averages()
    b_init = orc.background()
    orc.lf_init()
    states = [([orc.download(0, f) for f in range(nf)], [orc.download(1, f) for f in range(nf)])]
    tape = [orc.background()]
    for _ in range(nsteps):
        orc.lf_evolve(1)
        states.append(([orc.download(0, f) for f in range(nf)], [orc.download(1, f) for f in range(nf)]))
        tape.append(orc.background())
    orc.lf_sync()
    b_fin = orc.background()
    phiT = [orc.download(0, f) for f in range(nf)]
    piT = [orc.download(1, f) for f in range(nf)]
    coeff = [c for c, _ in spec.terms]

    lphi = [np.array(s, dtype=np.float64) for s in seed_fld]
    lpi = [np.array(s, dtype=np.float64) for s in seed_pi]
    gc = np.zeros(len(coeff))

    def kick_T(phi, c, aK, aV):
        mons = _monomials(spec, phi)
        RK = sum(np.sum(lpi[f] * _lap_np(phi[f], dx)) for f in range(nf))
        RV = 0.0
        for f in range(nf):
            h = sum(coeff[t] * sum(mons[t][2][f][g] * lpi[g] for g in range(nf)) for t in range(len(coeff)))
            lphi[f] += c * (aK * _lap_np(lpi[f], dx) - aV * h)
        for t in range(len(coeff)):
            st = sum(np.sum(lpi[f] * mons[t][1][f]) for f in range(nf))
            gc[t] += -c * aV * st
            RV += coeff[t] * st
        return RK, RV

    lam_aI, lam_aDotSI, lam_pi2SI, lam_G, lam_U = seed_bg[0], 0.0, 0.0, 0.0, 0.0
    bT = tape[nsteps]
    lam_pi2T = 0.0
    if expansion:
        lam_aDotSI = seed_bg[1]
        lamK = 0.5 * dt * seed_bg[1]
        d = _sf_partials(spec, bT["aI"], b_fin["pi2AvI"], bT["grad2AvI"], bT["potAvI"])
        lam_pi2T = lamK * d["dpi2"]
        lam_G, lam_U = lamK * d["dG"], lamK * d["dU"]
        lam_aI += lamK * d["da"]
    for f in range(nf):
        lpi[f] += lam_pi2T * 2.0 / V * piT[f]
    aK, aV = bT["aI"] ** (1 + al), bT["aI"] ** (3 + al)
    RK, RV = kick_T(phiT, 0.5 * dt, aK, aV)
    if expansion:
        lam_aI += 0.5 * dt * (RK * (1 + al) * bT["aI"] ** al - RV * (3 + al) * bT["aI"] ** (2 + al))

    for n in range(nsteps - 1, -1, -1):
        bn, bn1 = tape[n], tape[n + 1]
        phi1, pih = states[n + 1]
        mons = _monomials(spec, phi1)
        for f in range(nf):
            dv = sum(coeff[t] * mons[t][1][f] for t in range(len(coeff)))
            lphi[f] += lam_G * (-2.0 / V) * _lap_np(phi1[f], dx) + lam_U * dv / V
        for t in range(len(coeff)):
            gc[t] += lam_U * np.sum(mons[t][0]) / V
        Ra = sum(np.sum(lphi[f] * pih[f]) for f in range(nf))
        F = bn1["aSI"] ** (al - 3)
        lamP = 0.0
        if expansion:
            lam_aSI1 = dt * Ra * (al - 3) * bn1["aSI"] ** (al - 4)
            lam_aI1 = lam_aI + 0.5 * lam_aSI1
            lam_aI_n = 0.5 * lam_aSI1 + lam_aI1
            lam_aDotSI1 = lam_aDotSI + dt * lam_aI1
            lam_aDotSI_n = lam_aDotSI1
            lamK = dt * lam_aDotSI1
            d = _sf_partials(spec, bn["aI"], bn1["pi2AvI"], bn["grad2AvI"], bn["potAvI"])
            lam_pi2I = lamK * d["dpi2"]
            lam_G_n, lam_U_n = lamK * d["dG"], lamK * d["dU"]
            lam_aI_n += lamK * d["da"]
            lam_pi2SI_n = 0.5 * lam_pi2I
            lamP = 0.5 * lam_pi2I + lam_pi2SI
        for f in range(nf):
            lpi[f] += dt * F * lphi[f] + lamP * 2.0 / V * pih[f]
        w = 0.5 if n == 0 else 1.0
        aK, aV = bn["aI"] ** (1 + al), bn["aI"] ** (3 + al)
        RK, RV = kick_T(states[n][0], w * dt, aK, aV)
        if expansion:
            lam_aI_n += w * dt * (RK * (1 + al) * bn["aI"] ** al - RV * (3 + al) * bn["aI"] ** (2 + al))
            lam_aI, lam_aDotSI, lam_pi2SI, lam_G, lam_U = lam_aI_n, lam_aDotSI_n, lam_pi2SI_n, lam_G_n, lam_U_n

    if expansion:
        lam_aDotI0 = lam_aDotSI
        lamK0 = -0.5 * dt * lam_aDotSI
        d = _sf_partials(spec, b_init["aI"], b_init["pi2AvI"], b_init["grad2AvI"], b_init["potAvI"])
        lam_pi2I0 = lamK0 * d["dpi2"]
        lam_G += lamK0 * d["dG"]
        lam_U += lamK0 * d["dU"]
        a0, fm = b_init["aI"], spec.fStar / spec.MPl
        lamRho = lam_aDotI0 / (2.0 * b_init["aDotI"]) * a0 ** (2 * al + 2) / 3.0 * fm * fm
        lamP0 = lam_pi2I0 + lam_pi2SI + lamRho * 0.5 * a0 ** -6
        lamG0 = lam_G + lamRho * 0.5 * a0 ** -2
        lamU0 = lam_U + lamRho
        phi_0, pi_0 = states[0]
        mons = _monomials(spec, phi_0)
        for f in range(nf):
            dv = sum(coeff[t] * mons[t][1][f] for t in range(len(coeff)))
            lphi[f] += lamG0 * (-2.0 / V) * _lap_np(phi_0[f], dx) + lamU0 * dv / V
            lpi[f] += lamP0 * 2.0 / V * pi_0[f]
        for t in range(len(coeff)):
            gc[t] += lamU0 * np.sum(mons[t][0]) / V
    return lphi, lpi, gc
