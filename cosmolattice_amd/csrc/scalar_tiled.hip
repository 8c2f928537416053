// Tiled scalar-singlet lattice pass for N a multiple of 64: same arithmetic as scalar_pass.hip
// (bit-identical results), organised for HBM streaming on CDNA4.
//
// Each workgroup (4 wavefronts) owns a 64 (z) x TY (y) column of sites and marches along x over
// XC planes.  Along x the stencil neighbours live in registers (prev / cur / next plane); within a
// plane the y and z neighbours come from a double-buffered LDS tile with a one-site halo.  The halo
// sites are loaded and drifted by dedicated lanes, so every HBM array is read once per pass plus
// the halo (2 rows + 2 columns per plane, and 2 extra planes per x-chunk).  Loads for plane x+2 are
// issued before plane x is computed to hide HBM latency.  One barrier per plane.
//
// Workgroups are dispatched round-robin over the 8 XCDs; the block index is remapped so that the
// workgroups of one XCD own neighbouring tiles and x-chunks, which keeps halo re-reads in that
// XCD's L2.
#include <algorithm>

#include "clmi_internal.hpp"
#include "scalar_poly.hpp"

namespace clmi {
namespace {

constexpr int TZ = 64;

__device__ __forceinline__ void pair_add(float &hi, float &lo, float d)
{
  const float s = lo + d;
  const float h = hi + s;
  const float bb = h - hi;
  const float e = (hi - (h - bb)) + (s - bb);
  hi = h;
  lo = e;
}

// Storage traits: a drifted field value (FP64: double; FP32 bulk: hi/lo pair) and a momentum.
template <bool F32>
struct Tr;
template <>
struct Tr<false> {
  using V = double;
  using P = double;
  struct Raw {
    double f;
    double p;
  };
  __device__ static __forceinline__ Raw load(const ScalarPassArgs &a, int f, uint32_t i, bool need_p)
  {
    Raw r;
    r.f = ((const double *)a.phi_in[f])[i];
    r.p = need_p ? ((const double *)a.pi_in[f])[i] : 0.0;
    return r;
  }
  template <bool DRIFT>
  __device__ static __forceinline__ V value(const Raw &r, double fs, double cdt, float)
  {
    double v = r.f;
    if (DRIFT) v = __dadd_rn(v, __dmul_rn(fs, __dmul_rn(cdt, r.p)));
    return v;
  }
};
template <>
struct Tr<true> {
  using V = float2;
  using P = float;
  struct Raw {
    float h, l, p;
  };
  __device__ static __forceinline__ Raw load(const ScalarPassArgs &a, int f, uint32_t i, bool need_p)
  {
    Raw r;
    r.h = ((const float *)a.phi_in[f])[i];
    r.l = a.philo_in[f][i];
    r.p = need_p ? ((const float *)a.pi_in[f])[i] : 0.0f;
    return r;
  }
  template <bool DRIFT>
  __device__ static __forceinline__ V value(const Raw &r, double, double, float cf)
  {
    float h = r.h, l = r.l;
    if (DRIFT) pair_add(h, l, cf * r.p);
    return make_float2(h, l);
  }
};

template <bool NT, typename S>
__device__ __forceinline__ void store_out(S *p, uint32_t i, S v)
{
  if constexpr (NT)
    __builtin_nontemporal_store(v, p + i);
  else
    p[i] = v;
}

// Partial sums are written per (group of G planes, y-tile, z-tile), independent of the x-chunk
// length XC (a multiple of G), so every XC gives the same summation order.
// PK: the potential's compile-time exponent table (0: runtime table).
// KICK2 (with DRIFT and KICK): kick a second time, by a.ckick2, from the same drifted fields, summing the
// momenta after each kick (slots RED_PI2S and RED_PI2S_KICK2) and storing only the twice-kicked ones; each
// product and sum is formed as a following kick-only pass would form it.
template <int NF, bool F32, bool DRIFT, bool KICK, bool RED, int RY, bool NT, int PK, bool KICK2 = false>
__global__ __launch_bounds__(kBlock) void k_scalar_tiled(const ScalarPassArgs a, const int XC, const int G)
{
  static_assert(!KICK2 || (DRIFT && KICK), "a double kick follows a drift");
  using T = Tr<F32>;
  using V = typename T::V;
  using Raw = typename T::Raw;
  constexpr int TY = 4 * RY;
  __shared__ V lds[2][NF][TY + 2][TZ + 2];

  const int N = a.N;
  const uint32_t NN = (uint32_t)N * N;
  const int nTz = N / TZ, nTy = N / TY;
  const int nwg = gridDim.x;
  const int b = blockIdx.x;
  const int L = (nwg % 8 == 0) ? (b % 8) * (nwg / 8) + b / 8 : b;
  const int tz = L % nTz, ty = (L / nTz) % nTy, cx = L / (nTz * nTy);
  const int z0 = tz * TZ, y0 = ty * TY, x0 = cx * XC;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int z = z0 + lane;

  const double fs = DRIFT ? a.st->fdrift_s : 0.0;
  const double aK = KICK ? a.st->aK : 0.0;
  const double aV = KICK ? a.st->aV : 0.0;
  const float cf = DRIFT ? (float)(a.st->fdrift_s * a.cdt) : 0.0f;
  const float kK = KICK ? (float)(a.ckick * a.st->aK * a.inv_dx2) : 0.0f;
  const float kV = KICK ? (float)(a.ckick * a.st->aV) : 0.0f;
  const float kK2 = KICK2 ? (float)(a.ckick2 * a.st->aK * a.inv_dx2) : 0.0f;
  const float kV2 = KICK2 ? (float)(a.ckick2 * a.st->aV) : 0.0f;
  const bool need_p_nb = DRIFT;  // momenta of neighbouring planes are read only to drift them

  auto site = [&](int x, int y, int zz) -> uint32_t {
    x = x < 0 ? x + N : (x >= N ? x - N : x);
    y = y < 0 ? y + N : (y >= N ? y - N : y);
    zz = zz < 0 ? zz + N : (zz >= N ? zz - N : zz);
    return (uint32_t)x * NN + (uint32_t)y * N + zz;
  };

  // halo lanes: wave 0 -> row y0-1, wave 1 -> row y0+TY, waves 2/3 lanes < TY -> columns z0-1 / z0+TZ
  int hy = 0, hz = 0, hly = 0, hlz = 0;
  bool has_halo = false;
  if (wave == 0) { hy = y0 - 1; hz = z; hly = 0; hlz = lane + 1; has_halo = true; }
  else if (wave == 1) { hy = y0 + TY; hz = z; hly = TY + 1; hlz = lane + 1; has_halo = true; }
  else if (lane < TY) {
    hy = y0 + lane;
    hly = lane + 1;
    if (wave == 2) { hz = z0 - 1; hlz = 0; } else { hz = z0 + TZ; hlz = TZ + 1; }
    has_halo = true;
  }

  V prev[RY][NF], cur[RY][NF], nxt[RY][NF];
  typename T::P pic[RY][NF], pin[RY][NF];
  Raw pf[RY][NF];
  double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;

  // prologue: planes x0-1, x0, x0+1 and the LDS tile of plane x0
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int y = y0 + wave * RY + r;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      Raw q = T::load(a, f, site(x0 - 1, y, z), need_p_nb);
      prev[r][f] = T::template value<DRIFT>(q, fs, a.cdt, cf);
      q = T::load(a, f, site(x0, y, z), true);
      cur[r][f] = T::template value<DRIFT>(q, fs, a.cdt, cf);
      pic[r][f] = q.p;
      q = T::load(a, f, site(x0 + 1, y, z), true);
      nxt[r][f] = T::template value<DRIFT>(q, fs, a.cdt, cf);
      pin[r][f] = q.p;
      lds[0][f][wave * RY + r + 1][lane + 1] = cur[r][f];
    }
  }
  if (has_halo) {
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const Raw q = T::load(a, f, site(x0, hy, hz), need_p_nb);
      lds[0][f][hly][hlz] = T::template value<DRIFT>(q, fs, a.cdt, cf);
    }
  }
  __syncthreads();

  for (int i = 0; i < XC; ++i) {
    const int x = x0 + i;
    const int buf = i & 1;
    // prefetch plane x+2 (own sites) and the halo of plane x+1
    const bool more = i + 1 < XC;
    if (more) {
#pragma unroll
      for (int r = 0; r < RY; ++r)
#pragma unroll
        for (int f = 0; f < NF; ++f) pf[r][f] = T::load(a, f, site(x + 2, y0 + wave * RY + r, z), true);
    }
    Raw hq[NF];
    if (more && has_halo) {
#pragma unroll
      for (int f = 0; f < NF; ++f) hq[f] = T::load(a, f, site(x + 1, hy, hz), need_p_nb);
    }

    // compute plane x
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int ly = wave * RY + r + 1;
      const uint32_t s = site(x, y0 + ly - 1, z);
      double g2 = 0.0;
      float g2f = 0.0f, p2f = 0.0f, p2fb = 0.0f;  // FP32-bulk: per-site terms in FP32, accumulated in FP64
      double phid[NF];
      float phif[NF];
      double lapd[NF];
      float lapf[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const V c = cur[r][f];
        const V vyp = lds[buf][f][ly + 1][lane + 1], vym = lds[buf][f][ly - 1][lane + 1];
        const V vzp = lds[buf][f][ly][lane + 2], vzm = lds[buf][f][ly][lane];
        const V vxp = nxt[r][f], vxm = prev[r][f];
        if constexpr (!F32) {
          phid[f] = c;
          if (DRIFT) store_out<NT>((double *)a.phi_out[f], s, c);
          if (KICK) {
            double l = 0.0;
            l += vxp + vxm - 2 * c;
            l += vyp + vym - 2 * c;
            l += vzp + vzm - 2 * c;
            lapd[f] = l * a.inv_dx2;
          }
          if (RED) {
            const double gx = (vxp - c) * a.inv_dx, gy = (vyp - c) * a.inv_dx, gz = (vzp - c) * a.inv_dx;
            g2 += gx * gx;
            g2 += gy * gy;
            g2 += gz * gz;
          }
        } else {
          phif[f] = c.x + c.y;
          if (DRIFT) {
            store_out<NT>((float *)a.phi_out[f], s, c.x);
            store_out<NT>(a.philo_out[f], s, c.y);
          }
          const float dxp = (vxp.x - c.x) + (vxp.y - c.y);
          const float dyp = (vyp.x - c.x) + (vyp.y - c.y);
          const float dzp = (vzp.x - c.x) + (vzp.y - c.y);
          if (KICK) {
            const float dxm = (vxm.x - c.x) + (vxm.y - c.y);
            const float dym = (vym.x - c.x) + (vym.y - c.y);
            const float dzm = (vzm.x - c.x) + (vzm.y - c.y);
            lapf[f] = ((dxp + dxm) + (dyp + dym)) + (dzp + dzm);
          }
          if (RED) g2f += (dxp * dxp + dyp * dyp) + dzp * dzp;
        }
      }
      double p2 = 0.0, p2b = 0.0;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        if constexpr (!F32) {
          double p = pic[r][f];
          if constexpr (KICK2) {
            const double F = aK * lapd[f] - aV * dpot<NF, PK, double>(a.poly, phid, f);
            p += a.ckick * F;
            p2 += p * p;
            p += a.ckick2 * F;
            p2b += p * p;
            store_out<NT>((double *)a.pi_out[f], s, p);
          } else {
            if (KICK) {
              p += a.ckick * (aK * lapd[f] - aV * dpot<NF, PK, double>(a.poly, phid, f));
              store_out<NT>((double *)a.pi_out[f], s, p);
            }
            p2 += p * p;
          }
        } else {
          float p = pic[r][f];
          if constexpr (KICK2) {
            const float d = dpot<NF, PK, float>(a.poly, phif, f);
            p += kK * lapf[f] - kV * d;
            p2f += p * p;
            p += kK2 * lapf[f] - kV2 * d;
            p2fb += p * p;
            store_out<NT>((float *)a.pi_out[f], s, p);
          } else {
            if (KICK) {
              p += kK * lapf[f] - kV * dpot<NF, PK, float>(a.poly, phif, f);
              store_out<NT>((float *)a.pi_out[f], s, p);
            }
            p2f += p * p;
          }
        }
      }
      if (RED) {
        if constexpr (KICK2) acc3 += F32 ? (double)p2fb : p2b;
        if constexpr (!F32) {
          acc0 += p2;
          acc1 += g2;
          acc2 += pot<NF, PK, double>(a.poly, phid);
        } else {
          acc0 += (double)p2f;
          acc1 += (double)g2f * a.inv_dx2;
          acc2 += (double)pot<NF, PK, float>(a.poly, phif);
        }
      }
    }

    if (RED && (i + 1) % G == 0) {
      // one partial row per (plane group, y-tile, z-tile); the loop barrier below separates
      // consecutive uses of the reduction's shared scratch
      double v[kNumRed] = {acc0, acc1, acc2, 0, 0, 0, 0, 0};
      v[RED_PI2S_KICK2] = acc3;
      const size_t row = ((size_t)(x / G) * nTy + ty) * nTz + tz;
      block_reduce_store<kNumRed>(v, a.partials + row * kNumRed);
      acc0 = acc1 = acc2 = acc3 = 0.0;
    }

    if (more) {
      // LDS tile of plane x+1
#pragma unroll
      for (int r = 0; r < RY; ++r)
#pragma unroll
        for (int f = 0; f < NF; ++f) lds[buf ^ 1][f][wave * RY + r + 1][lane + 1] = nxt[r][f];
      if (has_halo) {
#pragma unroll
        for (int f = 0; f < NF; ++f) lds[buf ^ 1][f][hly][hlz] = T::template value<DRIFT>(hq[f], fs, a.cdt, cf);
      }
      // rotate the register queue
#pragma unroll
      for (int r = 0; r < RY; ++r)
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          prev[r][f] = cur[r][f];
          cur[r][f] = nxt[r][f];
          pic[r][f] = pin[r][f];
          nxt[r][f] = T::template value<DRIFT>(pf[r][f], fs, a.cdt, cf);
          pin[r][f] = pf[r][f].p;
        }
      __syncthreads();
    }
  }
}

template <int NF, bool F32, bool NT, int PK>
hipError_t dispatch_tiled(bool drift, bool kick, bool red, const ScalarPassArgs &a, int blocks, int XC, int G,
                          hipStream_t s, bool kick2)
{
  constexpr int RY = 4;
  const int code = (drift ? 4 : 0) | (kick ? 2 : 0) | (red ? 1 : 0);
  if (kick2) {
    if (!drift || !kick) return hipErrorInvalidValue;
    if (red)
      hipLaunchKernelGGL((k_scalar_tiled<NF, F32, true, true, true, RY, NT, PK, true>), dim3(blocks), dim3(kBlock), 0, s,
                         a, XC, G);
    else
      hipLaunchKernelGGL((k_scalar_tiled<NF, F32, true, true, false, RY, NT, PK, true>), dim3(blocks), dim3(kBlock), 0,
                         s, a, XC, G);
    return hipGetLastError();
  }
#define CLMI_ST(D, K, R)                                                                                        \
  case ((D ? 4 : 0) | (K ? 2 : 0) | (R ? 1 : 0)):                                                              \
    hipLaunchKernelGGL((k_scalar_tiled<NF, F32, D, K, R, RY, NT, PK>), dim3(blocks), dim3(kBlock), 0, s, a, XC, G); \
    break;
  switch (code) {
    CLMI_ST(true, true, true)
    CLMI_ST(true, true, false)
    CLMI_ST(true, false, true)
    CLMI_ST(true, false, false)
    CLMI_ST(false, true, true)
    CLMI_ST(false, true, false)
    CLMI_ST(false, false, true)
    default:
      return hipErrorInvalidValue;
  }
#undef CLMI_ST
  return hipGetLastError();
}

}  // namespace

// Tile plan for N with x-chunk XC (false: the tiled pass does not apply).  G, the plane group of
// the partial sums, depends on N only.
bool scalar_tiled_plan(int N, int nf, int XC, bool nt, TiledPlan *p)
{
  constexpr int TY = 16;
  if (nf < 1 || nf > 2 || N % TZ != 0 || N % TY != 0) return false;
  // planes per partial-sum row: a power of two (so that it divides N, a multiple of 64), the largest
  // not above N/16 (at most 16) up to 256^3 and not above N/64 beyond, at least 4: few rows for the
  // background update to sum (1024 at 256^3), and at most 64 row groups per tile column
  const int target = N <= 256 ? std::min(16, N / 16) : N / 64;
  int G = 4;
  while (2 * G <= target) G *= 2;
  if (XC < G || XC % G != 0 || N % XC != 0) return false;
  const long long tiles = (long long)(N / TZ) * (N / TY);
  p->XC = XC;
  p->G = G;
  p->NT = nt;
  p->blocks = (int)(tiles * (N / XC));
  p->rows = (int)(tiles * (N / G));
  return true;
}

hipError_t launch_scalar_tiled(int nf, bool f32, bool drift, bool kick, bool red, const ScalarPassArgs &a,
                               const TiledPlan &p, hipStream_t s, bool kick2)
{
  // the compile-time lphi4 potential when the run's exponent table is lphi4's
  const bool pk = poly_kind(a.poly, nf) == 1;
  const int code = (nf == 2 ? 4 : 0) | (f32 ? 2 : 0) | (p.NT ? 1 : 0) | (pk ? 8 : 0);
  switch (code) {
    case 0: return dispatch_tiled<1, false, false, 0>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);
    case 1: return dispatch_tiled<1, false, true, 0>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);
    case 2: return dispatch_tiled<1, true, false, 0>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);
    case 3: return dispatch_tiled<1, true, true, 0>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);
    case 4: return dispatch_tiled<2, false, false, 0>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);
    case 5: return dispatch_tiled<2, false, true, 0>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);
    case 6: return dispatch_tiled<2, true, false, 0>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);
    case 7: return dispatch_tiled<2, true, true, 0>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);
    case 10: return dispatch_tiled<1, true, false, 1>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);
    case 11: return dispatch_tiled<1, true, true, 1>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);
    case 14: return dispatch_tiled<2, true, false, 1>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);
    case 15: return dispatch_tiled<2, true, true, 1>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);
    case 12: return dispatch_tiled<2, false, false, 1>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);
    case 13: return dispatch_tiled<2, false, true, 1>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);
    case 8: return dispatch_tiled<1, false, false, 1>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);
    case 9: return dispatch_tiled<1, false, true, 1>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace clmi
