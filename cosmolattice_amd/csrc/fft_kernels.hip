// Kernels around the FFTs: |n|-shell binning of power spectra and Gaussian-random-field modes.
//
// Binning (measurements/powerspectrum.h:85-134 with TempLat's projectRadially): a spectrum is one
// streaming pass over the half-complex R2C array, N*N lines of N/2+1 modes, z fastest.  A mode's bin
// is looked up from |n|^2 in a table built on the host with the host's formula, so bins, counts and
// mean |n| agree by construction.  Along a line |n| grows with kz, so each wave reduces a chunk of 64
// consecutive modes with a segmented scan keyed by (line, bin) and adds the segment sums into its own
// LDS histogram; the workgroup's histograms are added in wave order into one row per workgroup, and a
// second kernel adds the rows in row order.  Every sum has a fixed order, so spectra are
// bit-reproducible (no floating-point atomics).  w = 1 on the kz = 0 and kz = N/2 planes and 2
// elsewhere, so sums run over the full N^3 mode set.  When the histograms do not fit in LDS (more than
// kMaxStreamBins bins) the modes are grouped by bin once per (N, dk) into a bin-sorted index table
// (CSR) and each bin is summed by one workgroup instead.
//
// Gaussian random field (initializers/fluctuationsgenerator.h:39-71,103-150): per mode two complex
// Gaussian amplitudes fL, fR of rms sigma_k per real part; phi_k = (fL+fR)/sqrt2 and
// pi_k = i w_k (fL-fR)/sqrt2 - aDot phi_k.  Random numbers come from Philox4x32-10 (Salmon et al.,
// SC'11) with key = seed and counter = (mode, draw, stream, 0), so values depend only on
// (seed, stream, mode).  On the kz = 0 and kz = N/2 planes the array must be Hermitian: the mode
// with the larger linear index of a conjugate pair takes its partner's draws with L and R swapped
// and conjugated, and a self-conjugate mode uses fR = conj(fL); phi and pi are then exactly
// Hermitian and real after the C2R transform.
#include <algorithm>

#include "clmi_internal.hpp"

namespace clmi {

namespace {
hipError_t fill_bin_table(BinTable &t, int N, double dk, int nbins);
}

// Builds the table for (N, dk, nbins); on failure the table is left empty (N = 0), so the next call
// builds it again.
hipError_t build_bin_table(BinTable &t, int N, double dk, int nbins)
{
  free_bin_table(t);
  const hipError_t e = fill_bin_table(t, N, dk, nbins);
  if (e != hipSuccess) free_bin_table(t);
  return e;
}

namespace {
hipError_t fill_bin_table(BinTable &t, int N, double dk, int nbins)
{
  const uint32_t Nh = N / 2 + 1;
  const uint32_t M = (uint32_t)N * N * Nh;
  const int half = N / 2;
  const uint32_t n2max = 3u * (uint32_t)half * (uint32_t)half;
  // bin of every |n|^2 on the lattice (the host formula, applied once per |n|^2)
  std::vector<uint32_t> bin_of(n2max + 1);
  for (uint32_t n2 = 0; n2 <= n2max; ++n2) {
    const double nn = std::sqrt((double)n2);
    int b = (int)std::floor(nn / dk + 0.5);
    if (b > nbins - 1) b = nbins - 1;
    bin_of[n2] = (uint32_t)b;
  }
  // weighted multiplicity of each |n|^2: the (nx, ny) pairs first, then the kz planes
  std::vector<double> mult_r2(2u * (uint32_t)half * (uint32_t)half + 1, 0.0);
  for (int kx = 0; kx < N; ++kx) {
    const int nx = kx <= half ? kx : kx - N;
    for (int ky = 0; ky < N; ++ky) {
      const int ny = ky <= half ? ky : ky - N;
      mult_r2[nx * nx + ny * ny] += 1.0;
    }
  }
  std::vector<double> mult(n2max + 1, 0.0);
  for (uint32_t r2 = 0; r2 < mult_r2.size(); ++r2) {
    if (mult_r2[r2] == 0.0) continue;
    for (uint32_t kz = 0; kz < Nh; ++kz) {
      const double w = (kz == 0 || 2 * kz == (uint32_t)N) ? 1.0 : 2.0;
      mult[r2 + kz * kz] += w * mult_r2[r2];
    }
  }
  t.count.assign(nbins, 0.0);
  t.mean_n.assign(nbins, 0.0);
  std::vector<long double> sumn(nbins, 0.0L);  // extended precision: mean |n| to about 1e-16
  for (uint32_t n2 = 0; n2 <= n2max; ++n2) {
    if (mult[n2] == 0.0) continue;
    t.count[bin_of[n2]] += mult[n2];
    sumn[bin_of[n2]] += (long double)mult[n2] * std::sqrt((long double)n2);
  }
  for (int b = 0; b < nbins; ++b) t.mean_n[b] = t.count[b] > 0 ? (double)(sumn[b] / t.count[b]) : 0.0;
  t.N = N;
  t.dk = dk;
  t.nbins = nbins;
  hipError_t e;
  if (nbins <= kMaxStreamBins) {
    // streaming pass: 4 waves per workgroup while four histograms fit in 64 KiB of LDS
    t.wpg = (size_t)4 * nbins * sizeof(double) <= 65536 ? 4 : 1;
    const uint64_t chunks = ((uint64_t)M + 63) / 64;
    uint64_t waves = std::min<uint64_t>(4096, (chunks + 7) / 8);
    uint64_t groups = (waves + t.wpg - 1) / t.wpg;
    groups = std::min<uint64_t>(groups, std::max<uint64_t>(1, (4u << 20) / (uint64_t)nbins));
    t.groups = (int)std::max<uint64_t>(1, groups);
    t.chunks_per_wave = (uint32_t)((chunks + (uint64_t)t.groups * t.wpg - 1) / ((uint64_t)t.groups * t.wpg));
    const std::vector<uint16_t> bin16(bin_of.begin(), bin_of.end());
    e = hipMalloc(&t.d_bin, sizeof(uint16_t) * bin16.size());
    if (e != hipSuccess) return e;
    e = hipMemcpy(t.d_bin, bin16.data(), sizeof(uint16_t) * bin16.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    return hipMalloc(&t.d_rows, sizeof(double) * (size_t)t.groups * nbins);
  }
  // bin-sorted index table
  std::vector<uint32_t> start(nbins + 1, 0);
  for (uint32_t kx = 0; kx < (uint32_t)N; ++kx) {
    const int nx = kx <= (uint32_t)half ? (int)kx : (int)kx - N;
    for (uint32_t ky = 0; ky < (uint32_t)N; ++ky) {
      const int ny = ky <= (uint32_t)half ? (int)ky : (int)ky - N;
      for (uint32_t kz = 0; kz < Nh; ++kz) start[bin_of[nx * nx + ny * ny + kz * kz] + 1]++;
    }
  }
  for (int b = 0; b < nbins; ++b) start[b + 1] += start[b];
  std::vector<uint32_t> order(M);
  std::vector<uint32_t> fill(start.begin(), start.end() - 1);
  for (uint32_t kx = 0; kx < (uint32_t)N; ++kx) {
    const int nx = kx <= (uint32_t)half ? (int)kx : (int)kx - N;
    for (uint32_t ky = 0; ky < (uint32_t)N; ++ky) {
      const int ny = ky <= (uint32_t)half ? (int)ky : (int)ky - N;
      for (uint32_t kz = 0; kz < Nh; ++kz)
        order[fill[bin_of[nx * nx + ny * ny + kz * kz]]++] = (kx * N + ky) * Nh + kz;
    }
  }
  e = hipMalloc(&t.d_modes, sizeof(uint32_t) * (size_t)M);
  if (e != hipSuccess) return e;
  e = hipMalloc(&t.d_start, sizeof(uint32_t) * (size_t)(nbins + 1));
  if (e != hipSuccess) return e;
  e = hipMemcpy(t.d_modes, order.data(), sizeof(uint32_t) * (size_t)M, hipMemcpyHostToDevice);
  if (e != hipSuccess) return e;
  return hipMemcpy(t.d_start, start.data(), sizeof(uint32_t) * (size_t)(nbins + 1), hipMemcpyHostToDevice);
}
}  // namespace

void free_bin_table(BinTable &t)
{
  if (t.d_modes) (void)hipFree(t.d_modes);
  if (t.d_start) (void)hipFree(t.d_start);
  if (t.d_bin) (void)hipFree(t.d_bin);
  if (t.d_rows) (void)hipFree(t.d_rows);
  t.d_modes = nullptr;
  t.d_start = nullptr;
  t.d_bin = nullptr;
  t.d_rows = nullptr;
  t.N = 0;
  t.nbins = 0;
}

namespace {

// |n|^NPOW for a mode with |n|^2 = n2 (NPOW = 0, 1 or 3): the weight PS_version 3 puts inside the bin sum
// (measurements/powerspectrum.h:111-131).
template <int NPOW>
__device__ __forceinline__ double npow_weight(uint32_t n2)
{
  if constexpr (NPOW == 0) {
    return 1.0;
  } else {
    const double r = sqrt((double)n2);
    if constexpr (NPOW == 1) return r;
    else return r * r * r;
  }
}

template <int NPOW>
__global__ __launch_bounds__(256) void k_bin_sums(const uint32_t *modes, const uint32_t *start, const double2 *fk,
                                                  uint32_t Nh, uint32_t Nz2, double *out)
{
  __shared__ double sh[256];
  const uint32_t b = blockIdx.x;
  const uint32_t lo = start[b], hi = start[b + 1];
  double s = 0.0;
  for (uint32_t i = lo + threadIdx.x; i < hi; i += 256) {
    const uint32_t m = modes[i];
    const uint32_t kz = m % Nh;
    const double2 v = fk[m];
    double w = (kz == 0 || 2 * kz == Nz2) ? 1.0 : 2.0;
    if constexpr (NPOW != 0) {
      const uint32_t line = m / Nh, kx = line / Nz2, ky = line - kx * Nz2;
      const int nx = 2 * kx <= Nz2 ? (int)kx : (int)kx - (int)Nz2;
      const int ny = 2 * ky <= Nz2 ? (int)ky : (int)ky - (int)Nz2;
      w *= npow_weight<NPOW>((uint32_t)(nx * nx + ny * ny) + kz * kz);
    }
    s += w * (v.x * v.x + v.y * v.y);
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int stride = 128; stride > 0; stride >>= 1) {
    if ((int)threadIdx.x < stride) sh[threadIdx.x] += sh[threadIdx.x + stride];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[b] = sh[0];
}

// Lane-shifted copy of a double by a DPP row pattern (lanes outside the pattern or the row mask
// read 0).
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double x)
{
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, ROWMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, ROWMASK, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Streaming binning.  Wave w of workgroup g reduces the 64-mode chunks [c0, c0 + cpw) of the
// half-space array (modes m = (kx N + ky) Nh + kz) into its LDS histogram h; the workgroup then writes
// the sum of its WPG histograms, in wave order, to rows[g].
template <int WPG, int NPOW>
__global__ __launch_bounds__(64 * WPG) void k_bin_stream(const double2 *__restrict__ fk,
                                                         const uint16_t *__restrict__ bin_of, uint32_t N,
                                                         uint32_t Nh, uint32_t M, uint32_t nbins, uint32_t cpw,
                                                         double *__restrict__ rows)
{
  constexpr int U = 4;  // chunks loaded ahead
  extern __shared__ double hist[];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double *h = hist + wave * nbins;
  for (uint32_t b = lane; b < nbins; b += 64) h[b] = 0.0;
  const uint32_t half = N / 2;
  const uint64_t nchunks = ((uint64_t)M + 63) / 64;
  const uint64_t c0 = ((uint64_t)blockIdx.x * WPG + wave) * cpw;
  const uint64_t cend = c0 + cpw < nchunks ? c0 + cpw : nchunks;
  if (c0 < cend) {
    // line (kx0, ky0) and kz0 of the current chunk's first mode
    const uint64_t m0 = c0 * 64;
    const uint32_t line = (uint32_t)(m0 / Nh);
    uint32_t kz0 = (uint32_t)(m0 - (uint64_t)line * Nh);
    uint32_t kx0 = line / N, ky0 = line - kx0 * N;
    for (uint64_t c = c0; c < cend; c += U) {
      double2 v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint64_t m = (c + u) * 64 + lane;
        v[u] = fk[(c + u < cend && m < M) ? m : 0];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (c + u >= cend) break;
        const bool valid = (c + u) * 64 + lane < M;
        // this lane's line offset dl from the chunk's first line, and its kz
        uint32_t kz = kz0 + lane, dl = 0;
        while (kz >= Nh) {
          kz -= Nh;
          ++dl;
        }
        uint32_t ky = ky0 + dl, kx = kx0;
        while (ky >= N) {
          ky -= N;
          ++kx;
        }
        const int nx = kx <= half ? (int)kx : (int)kx - (int)N;
        const int ny = ky <= half ? (int)ky : (int)ky - (int)N;
        const uint32_t n2 = valid ? (uint32_t)(nx * nx + ny * ny) + kz * kz : 0u;
        const uint32_t b = bin_of[n2];
        const double w = ((kz == 0 || 2 * kz == N) ? 1.0 : 2.0) * npow_weight<NPOW>(n2);
        double x = valid ? w * (v[u].x * v[u].x + v[u].y * v[u].y) : 0.0;
        // keys (line, bin) do not decrease along the chunk; invalid lanes form its tail
        const uint32_t key = valid ? (dl << 16 | b) : 0xFFFFFFFFu;
        const uint32_t kprev = __shfl_up(key, 1, 64);
        const uint64_t heads = __ballot(lane == 0 || kprev != key);
        const uint32_t start = 63 - __clzll((long long)(heads & ((2ull << lane) - 1)));
        // segmented inclusive scan, x = sum over lanes [start, lane]: within rows of 16 lanes, then
        // the last lane of rows 0 and 2 into rows 1 and 3, then lane 31 into rows 2 and 3
        const uint32_t r = lane & 15;
        double o = dpp_f64<0x111, 0xf>(x);  // row_shr:1
        if (r >= 1 && lane >= start + 1) x += o;
        o = dpp_f64<0x112, 0xf>(x);  // row_shr:2
        if (r >= 2 && lane >= start + 2) x += o;
        o = dpp_f64<0x114, 0xf>(x);  // row_shr:4
        if (r >= 4 && lane >= start + 4) x += o;
        o = dpp_f64<0x118, 0xf>(x);  // row_shr:8
        if (r >= 8 && lane >= start + 8) x += o;
        o = dpp_f64<0x142, 0xa>(x);  // row_bcast:15
        if ((lane & 16) && (lane | 15) - 16 >= start) x += o;
        o = dpp_f64<0x143, 0xc>(x);  // row_bcast:31
        if (lane >= 32 && 31 >= start) x += o;
        // a segment ends where the next lane starts one (invalid lanes start their own)
        const bool tail = valid && (lane == 63 || ((heads >> (lane + 1)) & 1));
        // segments of one line have distinct bins; lines are added one after another
        const uint32_t dlast = __builtin_amdgcn_readlane(dl, 63);
        for (uint32_t j = 0; j <= dlast; ++j)
          if (tail && dl == j) h[b] += x;
        kz0 += 64;
        while (kz0 >= Nh) {
          kz0 -= Nh;
          if (++ky0 == N) {
            ky0 = 0;
            ++kx0;
          }
        }
      }
    }
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < nbins; b += 64 * WPG) {
    double s = hist[b];
#pragma unroll
    for (int w = 1; w < WPG; ++w) s += hist[w * nbins + b];
    rows[(size_t)blockIdx.x * nbins + b] = s;
  }
}

// out[b] = sum over rows g of rows[g][b] in a fixed order: wave q of 16 adds rows q, q + 16, ...,
// then the 16 wave sums are added in wave order.
__global__ __launch_bounds__(1024) void k_bin_rows(const double *__restrict__ rows, uint32_t groups, uint32_t nbins,
                                                   double *__restrict__ out)
{
  __shared__ double sh[16][64];
  const uint32_t l = threadIdx.x & 63, q = threadIdx.x >> 6;
  const uint32_t b = blockIdx.x * 64 + l;
  double s = 0.0;
  if (b < nbins) {
#pragma unroll 8
    for (uint32_t g = q; g < groups; g += 16) s += rows[(size_t)g * nbins + b];
  }
  sh[q][l] = s;
  __syncthreads();
  if (q == 0 && b < nbins) {
    double t = sh[0][l];
#pragma unroll
    for (int k = 1; k < 16; ++k) t += sh[k][l];
    out[b] = t;
  }
}

// ---- Philox4x32-10 ---------------------------------------------------------------------------
struct U4 {
  uint32_t v[4];
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1)
{
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    const uint64_t p0 = (uint64_t)M0 * c.v[0];
    const uint64_t p1 = (uint64_t)M1 * c.v[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    U4 n;
    n.v[0] = hi1 ^ c.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ c.v[3] ^ k1;
    n.v[3] = lo0;
    c = n;
  }
  return c;
}

// 53-bit uniform in [0, 1) from two 32-bit words.
__device__ __forceinline__ double u53(uint32_t a, uint32_t b)
{
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// Two standard normals (Box-Muller on one uniform pair) for (mode, draw).
__device__ __forceinline__ void normals2(uint64_t seed, uint32_t stream, uint32_t m, uint32_t draw, double g[2])
{
  U4 c;
  c.v[0] = m;
  c.v[1] = draw;
  c.v[2] = stream;
  c.v[3] = 0u;
  const U4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double u1 = u53(r.v[0], r.v[1]), u2 = u53(r.v[2], r.v[3]);
  const double rad = sqrt(-2.0 * log(1.0 - u1));
  double sn, cs;
  sincos(6.283185307179586477 * u2, &sn, &cs);
  g[0] = rad * cs;
  g[1] = rad * sn;
}

__global__ __launch_bounds__(256) void k_grf_modes(const GrfArgs a, int64_t n2max)
{
  const uint32_t N = a.N, Nh = N / 2 + 1;
  for (uint32_t m = blockIdx.x * 256 + threadIdx.x; m < a.nmodes; m += gridDim.x * 256) {
    const uint32_t kz = m % Nh;
    const uint32_t kxy = m / Nh;
    const uint32_t kx = kxy / N, ky = kxy - kx * N;
    const int nx = kx <= N / 2 ? (int)kx : (int)kx - (int)N;
    const int ny = ky <= N / 2 ? (int)ky : (int)ky - (int)N;
    const int64_t n2 = (int64_t)nx * nx + (int64_t)ny * ny + (int64_t)kz * kz;
    double2 ph = make_double2(0.0, 0.0), pp = make_double2(0.0, 0.0);
    if (m != 0 && n2 <= n2max) {
      const double k = sqrt((double)n2) * a.kIR;
      const double meff2 = a.mass2 > 0 ? a.mass2 : 0.0;
      const double omega = sqrt(k * k + meff2);
      const double sigma = a.sigma_pref * pow(2 * omega, -0.5);
      // canonical representative of the conjugate pair on the kz = 0 and kz = N/2 planes
      uint32_t src = m;
      int mode = 0;  // 0: own draws, 1: partner's draws swapped+conjugated, 2: self-conjugate
      if (kz == 0 || 2 * kz == N) {
        const uint32_t px = (N - kx) % N, py = (N - ky) % N;
        const uint32_t pm = (px * N + py) * Nh + kz;
        if (pm == m)
          mode = 2;
        else if (pm < m) {
          mode = 1;
          src = pm;
        }
      }
      double gL[2], gR[2];
      normals2(a.seed, a.stream, src, 0u, gL);
      normals2(a.seed, a.stream, src, 1u, gR);
      double2 fL = make_double2(sigma * gL[0], sigma * gL[1]);
      double2 fR = make_double2(sigma * gR[0], sigma * gR[1]);
      if (mode == 1) {
        const double2 tL = fL;
        fL = make_double2(fR.x, -fR.y);
        fR = make_double2(tL.x, -tL.y);
      } else if (mode == 2) {
        fR = make_double2(fL.x, -fL.y);
      }
      const double is2 = 0.70710678118654752440;
      ph = make_double2((fL.x + fR.x) * is2, (fL.y + fR.y) * is2);
      // i * omega * (fL - fR) / sqrt2 - aDot * phi
      const double dr = (fL.x - fR.x) * is2, di = (fL.y - fR.y) * is2;
      pp = make_double2(-omega * di - a.aDot * ph.x, omega * dr - a.aDot * ph.y);
    }
    ((double2 *)a.phik)[m] = ph;
    ((double2 *)a.pik)[m] = pp;
  }
}

// Complex-scalar fluctuations from constrained phases (u1initializer.h:157-230) for component
// `comp` (0 or 1): per mode, Rayleigh amplitudes a_c and unit phases e00, e01, e10 drawn from
// Philox (stream, counter (mode, draw 0..2)); e11 = e01 e10 conj(e00) makes every mode's
// contribution to the total charge vanish.
//   phi_c = a_c (e_c0 + e_c1) / 2,  pi_c = -i omega_0 a_c (e_c0 - e_c1) / 2 - aDot phi_c.
// On the kz = 0 and kz = N/2 planes the mode -n takes the draws of n with
// (e_c0, e_c1) -> (conj e_c1, conj e_c0), which keeps the fields real and the phase constraint;
// self-conjugate modes use e01 = conj e00 and e10 = e00.
__global__ __launch_bounds__(256) void k_grf_cs_modes(const GrfCsArgs a, int64_t n2max)
{
  const uint32_t N = a.N, Nh = N / 2 + 1;
  for (uint32_t m = blockIdx.x * 256 + threadIdx.x; m < a.nmodes; m += gridDim.x * 256) {
    const uint32_t kz = m % Nh;
    const uint32_t kxy = m / Nh;
    const uint32_t kx = kxy / N, ky = kxy - kx * N;
    const int nx = kx <= N / 2 ? (int)kx : (int)kx - (int)N;
    const int ny = ky <= N / 2 ? (int)ky : (int)ky - (int)N;
    const int64_t n2 = (int64_t)nx * nx + (int64_t)ny * ny + (int64_t)kz * kz;
    double2 ph = make_double2(0.0, 0.0), pp = make_double2(0.0, 0.0);
    if (m != 0 && n2 <= n2max) {
      const double k = sqrt((double)n2) * a.kIR;
      const double m2c = a.mass2[a.comp] > 0 ? a.mass2[a.comp] : 0.0;
      const double m20 = a.mass2[0] > 0 ? a.mass2[0] : 0.0;
      const double sigma = a.sigma_pref * pow(2 * sqrt(k * k + m2c), -0.5);
      const double omega0 = sqrt(k * k + m20);
      uint32_t src = m;
      int mode = 0;
      if (kz == 0 || 2 * kz == N) {
        const uint32_t px = (N - kx) % N, py = (N - ky) % N;
        const uint32_t pm = (px * N + py) * Nh + kz;
        if (pm == m)
          mode = 2;
        else if (pm < m) {
          mode = 1;
          src = pm;
        }
      }
      double u[6];
#pragma unroll
      for (uint32_t d = 0; d < 3; ++d) {
        U4 c;
        c.v[0] = src;
        c.v[1] = d;
        c.v[2] = a.stream;
        c.v[3] = 0u;
        const U4 r = philox4x32_10(c, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
        u[2 * d] = u53(r.v[0], r.v[1]);
        u[2 * d + 1] = u53(r.v[2], r.v[3]);
      }
      // Rayleigh amplitude with <R^2> = 2 (the |g|^2 of the complex normals of k_grf_modes)
      const double amp = sigma * sqrt(-2.0 * log(1.0 - u[a.comp]));
      const double twopi = 6.283185307179586477;
      double s00, c00, s01, c01, s10, c10;
      sincos(twopi * u[2], &s00, &c00);
      sincos(twopi * u[3], &s01, &c01);
      sincos(twopi * u[4], &s10, &c10);
      double2 e00 = make_double2(c00, s00), e01 = make_double2(c01, s01), e10 = make_double2(c10, s10);
      if (mode == 2) {
        e01 = make_double2(e00.x, -e00.y);
        e10 = e00;
      }
      // e11 = e01 * e10 * conj(e00)
      const double2 t = make_double2(e01.x * e10.x - e01.y * e10.y, e01.x * e10.y + e01.y * e10.x);
      double2 e11 = make_double2(t.x * e00.x + t.y * e00.y, t.y * e00.x - t.x * e00.y);
      double2 ec0 = a.comp == 0 ? e00 : e10, ec1 = a.comp == 0 ? e01 : e11;
      if (mode == 1) {
        const double2 t0 = ec0;
        ec0 = make_double2(ec1.x, -ec1.y);
        ec1 = make_double2(t0.x, -t0.y);
      }
      ph = make_double2(0.5 * amp * (ec0.x + ec1.x), 0.5 * amp * (ec0.y + ec1.y));
      // -i omega0 * d,  d = amp (ec0 - ec1) / 2
      const double dr = 0.5 * amp * (ec0.x - ec1.x), di = 0.5 * amp * (ec0.y - ec1.y);
      pp = make_double2(omega0 * di - a.aDot * ph.x, -omega0 * dr - a.aDot * ph.y);
    }
    ((double2 *)a.phik)[m] = ph;
    ((double2 *)a.pik)[m] = pp;
  }
}

// j0 = -dx^2 rho_Q with rho_Q = (2/dx) (fStar/omegaStar)^2 gQ Im[conj(pi) phi]
// (mattercurrents.h:27-48), from the current complex-scalar arrays.
__global__ void k_u1_source(const void *f0, const float *f0lo, const void *f1, const float *f1lo, const void *p0,
                            const void *p1, int f32, double coef, double *out, uint32_t V)
{
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
    double a0, a1, q0, q1;
    if (f32) {
      a0 = (double)((const float *)f0)[i] + (double)f0lo[i];
      a1 = (double)((const float *)f1)[i] + (double)f1lo[i];
      q0 = (double)((const float *)p0)[i];
      q1 = (double)((const float *)p1)[i];
    } else {
      a0 = ((const double *)f0)[i];
      a1 = ((const double *)f1)[i];
      q0 = ((const double *)p0)[i];
      q1 = ((const double *)p1)[i];
    }
    out[i] = coef * (q0 * a1 - q1 * a0);
  }
}

// Gauss constraint in Fourier space (u1initializer.h:86-104): pi_A,i = conj(k-_i) / |k-|^2 * j0,
// k-_i = 1 - exp(-2 pi i n_i / N), zero mode zero.
__global__ void k_gauss_solve(const double2 *j0, double2 *out, int dir, uint32_t N, uint32_t nmodes)
{
  const uint32_t Nh = N / 2 + 1;
  const double w = 6.283185307179586477 / (double)N;
  for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < nmodes; m += gridDim.x * blockDim.x) {
    const uint32_t kz = m % Nh;
    const uint32_t kxy = m / Nh;
    const uint32_t kx = kxy / N, ky = kxy - kx * N;
    const uint32_t n[3] = {kx, ky, kz};
    double k2 = 0.0, cr = 0.0, ci = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double sn, cs;
      sincos(w * (double)n[i], &sn, &cs);
      const double re = 1.0 - cs, im = sn;  // k-_i = 1 - cos + i sin
      k2 += re * re + im * im;
      if (i == dir) {
        cr = re;
        ci = -im;  // conjugate
      }
    }
    double2 r = make_double2(0.0, 0.0);
    if (m != 0) {
      const double2 j = j0[m];
      r = make_double2((cr * j.x - ci * j.y) / k2, (cr * j.y + ci * j.x) / k2);
    }
    out[m] = r;
  }
}

__global__ void k_grf_finish(const double *src, double scale, double add, void *dst, float *dstlo, int f32,
                             uint32_t V)
{
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
    const double v = src[i] * scale + add;
    if (f32) {
      const float h = (float)v;
      ((float *)dst)[i] = h;
      if (dstlo) dstlo[i] = (float)(v - (double)h);
    } else {
      ((double *)dst)[i] = v;
    }
  }
}

}  // namespace

namespace {
template <int NPOW>
hipError_t bin_sums_pow(const BinTable &t, const void *fk, double *out, hipStream_t s)
{
  if (t.d_bin) {
    const uint32_t N = (uint32_t)t.N, Nh = N / 2 + 1, M = N * N * Nh;
    const size_t lds = (size_t)t.wpg * t.nbins * sizeof(double);
    if (t.wpg == 4)
      hipLaunchKernelGGL((k_bin_stream<4, NPOW>), dim3(t.groups), dim3(256), lds, s, (const double2 *)fk, t.d_bin, N, Nh,
                         M, (uint32_t)t.nbins, t.chunks_per_wave, t.d_rows);
    else
      hipLaunchKernelGGL((k_bin_stream<1, NPOW>), dim3(t.groups), dim3(64), lds, s, (const double2 *)fk, t.d_bin, N, Nh,
                         M, (uint32_t)t.nbins, t.chunks_per_wave, t.d_rows);
    hipLaunchKernelGGL(k_bin_rows, dim3((t.nbins + 63) / 64), dim3(1024), 0, s, t.d_rows, (uint32_t)t.groups,
                       (uint32_t)t.nbins, out);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_bin_sums<NPOW>, dim3(t.nbins), dim3(256), 0, s, t.d_modes, t.d_start, (const double2 *)fk,
                     (uint32_t)(t.N / 2 + 1), (uint32_t)t.N, out);
  return hipGetLastError();
}
}  // namespace

hipError_t launch_bin_sums(const BinTable &t, const void *fk, double *out, hipStream_t s, int npow)
{
  switch (npow) {
    case 0: return bin_sums_pow<0>(t, fk, out, s);
    case 1: return bin_sums_pow<1>(t, fk, out, s);
    case 3: return bin_sums_pow<3>(t, fk, out, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_grf_modes(const GrfArgs &a, int64_t n2max, hipStream_t s)
{
  uint32_t blocks = (a.nmodes + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_grf_modes, dim3(blocks), dim3(256), 0, s, a, n2max);
  return hipGetLastError();
}

hipError_t launch_grf_cs_modes(const GrfCsArgs &a, int64_t n2max, hipStream_t s)
{
  uint32_t blocks = (a.nmodes + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_grf_cs_modes, dim3(blocks), dim3(256), 0, s, a, n2max);
  return hipGetLastError();
}

hipError_t launch_u1_source(const void *f0, const float *f0lo, const void *f1, const float *f1lo, const void *p0,
                            const void *p1, int f32, double coef, double *out, uint32_t V, hipStream_t s)
{
  uint32_t blocks = (V + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_u1_source, dim3(blocks), dim3(256), 0, s, f0, f0lo, f1, f1lo, p0, p1, f32, coef, out, V);
  return hipGetLastError();
}

hipError_t launch_gauss_solve(const void *j0, void *out, int dir, int N, hipStream_t s)
{
  const uint32_t nmodes = (uint32_t)N * N * (N / 2 + 1);
  uint32_t blocks = (nmodes + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(k_gauss_solve, dim3(blocks), dim3(256), 0, s, (const double2 *)j0, (double2 *)out, dir,
                     (uint32_t)N, nmodes);
  return hipGetLastError();
}

hipError_t launch_grf_finish(const double *src, double scale, double add, void *dst, float *dstlo, int f32,
                             uint32_t V, hipStream_t s)
{
  uint32_t blocks = (V + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_grf_finish, dim3(blocks), dim3(256), 0, s, src, scale, add, dst, dstlo, f32, V);
  return hipGetLastError();
}

}  // namespace clmi
