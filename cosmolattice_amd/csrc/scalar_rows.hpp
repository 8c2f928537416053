// Full-row scalar-singlet lattice pass (kernel template; instantiated per precision in
// scalar_rows_f32.hip / scalar_rows_f64.hip) for N = 64, 128, 192, 256: same per-site arithmetic as
// scalar_pass.hip (bit-identical fields), laid out so that no load is partial.
//
// A workgroup owns TY complete z-rows (WPR wavefronts per row, each lane VZ consecutive sites, read
// and written with one vector access per array) and marches along x over XC planes.  x-neighbours
// are kept in registers (prev / cur / next plane), y-neighbours come from an LDS tile holding the
// TY rows of the current plane plus one halo row above and below, and z-neighbours are the lane's
// own registers except at its two edges, which are read from the same LDS row.  Because rows are
// complete there is no z-halo; the y-halo is two full rows per plane, the x-halo two planes per
// chunk.  Memory mode MM: bit 0 non-temporal stores, bit 1 non-temporal loads of the momenta (read
// exactly once), bit 2 non-temporal loads of the workgroup's own field rows (the y-halo rows, which
// the neighbouring tile also reads, always use ordinary loads), bit 3 (with bit 2) ordinary loads for
// the tile's first and last rows, which are the neighbouring tiles' halo rows.
//
// Volume sums are written per (group of G planes, TY-row y-tile), independent of XC.
#pragma once

#include "clmi_internal.hpp"
#include "scalar_poly.hpp"

namespace clmi {
namespace {

__device__ __forceinline__ void pair_add(float &hi, float &lo, float d)
{
  const float s = lo + d;
  const float h = hi + s;
  const float bb = h - hi;
  const float e = (hi - (h - bb)) + (s - bb);
  hi = h;
  lo = e;
}

// Lattice arrays live in device (global) memory.  Accesses go through global-address-space
// pointers even when the pointer itself was read from memory (the batched pass's member table), so
// the compiler emits global rather than flat instructions; flat loads would share the wait counter
// of the LDS tile reads and serialise them behind the prefetch.
template <typename S> using gptr = __attribute__((address_space(1))) S *;
template <typename S> using cgptr = const __attribute__((address_space(1))) S *;

// VZ consecutive elements with vector accesses (the site index of element 0 is a multiple of VZ);
// NTL: non-temporal loads.
template <bool NTL, typename Q>
__device__ __forceinline__ Q ldq(const __attribute__((address_space(1))) Q *p)
{
  if constexpr (NTL) return __builtin_nontemporal_load(p);
  else return *p;
}
template <int VZ, bool NTL = false, typename S>
__device__ __forceinline__ void vload(const S *pg, uint32_t s, S (&o)[VZ])
{
  const cgptr<S> p = (cgptr<S>)pg;
  if constexpr (VZ == 4 && sizeof(S) == 4) {
    typedef S v4 __attribute__((ext_vector_type(4)));
    const v4 v = ldq<NTL, v4>((cgptr<v4>)(p + s));
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
  } else if constexpr (VZ % 2 == 0) {
    typedef S v2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int k = 0; k < VZ; k += 2) {
      const v2 v = ldq<NTL, v2>((cgptr<v2>)(p + s + k));
      o[k] = v.x;
      o[k + 1] = v.y;
    }
  } else if constexpr (VZ == 3 && sizeof(S) == 4) {
    // one 12-byte access (global_load_dwordx3), aligned to the element only (s is a multiple of 3).
    // Only for 4-byte elements: a 3-element vector of doubles is accessed as 32 bytes.
    // (loaded here, not through ldq, so that the pointer keeps the element alignment)
    typedef S v3 __attribute__((ext_vector_type(3), aligned(sizeof(S))));
    const cgptr<v3> q = (cgptr<v3>)(p + s);
    v3 v;
    if constexpr (NTL)
      v = __builtin_nontemporal_load(q);
    else
      v = *q;
    o[0] = v.x; o[1] = v.y; o[2] = v.z;
  } else {
#pragma unroll
    for (int k = 0; k < VZ; ++k) o[k] = ldq<NTL, S>(p + s + k);
  }
}
template <int VZ, bool NT, typename S>
__device__ __forceinline__ void vstore(S *pg, uint32_t s, const S (&o)[VZ])
{
  const gptr<S> p = (gptr<S>)pg;
  if constexpr (VZ == 4 && sizeof(S) == 4) {
    typedef S v4 __attribute__((ext_vector_type(4)));
    v4 v;
    v.x = o[0]; v.y = o[1]; v.z = o[2]; v.w = o[3];
    if constexpr (NT) __builtin_nontemporal_store(v, (gptr<v4>)(p + s)); else *(gptr<v4>)(p + s) = v;
  } else if constexpr (VZ % 2 == 0) {
    typedef S v2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int k = 0; k < VZ; k += 2) {
      v2 v;
      v.x = o[k];
      v.y = o[k + 1];
      if constexpr (NT) __builtin_nontemporal_store(v, (gptr<v2>)(p + s + k)); else *(gptr<v2>)(p + s + k) = v;
    }
  } else if constexpr (VZ == 3 && sizeof(S) == 4) {
    // one 12-byte access (global_store_dwordx3); see vload
    typedef S v3 __attribute__((ext_vector_type(3), aligned(sizeof(S))));
    v3 v;
    v.x = o[0]; v.y = o[1]; v.z = o[2];
    if constexpr (NT) __builtin_nontemporal_store(v, (gptr<v3>)(p + s)); else *(gptr<v3>)(p + s) = v;
  } else {
#pragma unroll
    for (int k = 0; k < VZ; ++k) {
      if constexpr (NT) __builtin_nontemporal_store(o[k], p + s + k); else p[s + k] = o[k];
    }
  }
}

// Drifted field value and raw inputs per storage mode.
template <bool F32>
struct Tr;
template <>
struct Tr<false> {
  using V = double;
  using P = double;
  template <int VZ>
  struct Raw {
    double f[VZ], p[VZ];
  };
  template <int VZ, bool NTF = false, bool NTP = false>
  __device__ static __forceinline__ void load(const ScalarPassArgs &a, int f, uint32_t s, bool need_p, Raw<VZ> &r)
  {
    vload<VZ, NTF>((const double *)a.phi_in[f], s, r.f);
    if (need_p) vload<VZ, NTP>((const double *)a.pi_in[f], s, r.p);
    else
#pragma unroll
      for (int k = 0; k < VZ; ++k) r.p[k] = 0.0;
  }
  template <bool DRIFT, int VZ>
  __device__ static __forceinline__ V value(const Raw<VZ> &r, int k, double fs, double cdt, float)
  {
    double v = r.f[k];
    if (DRIFT) v = __dadd_rn(v, __dmul_rn(fs, __dmul_rn(cdt, r.p[k])));
    return v;
  }
};
template <>
struct Tr<true> {
  using V = float2;
  using P = float;
  template <int VZ>
  struct Raw {
    float h[VZ], l[VZ], p[VZ];
  };
  template <int VZ, bool NTF = false, bool NTP = false>
  __device__ static __forceinline__ void load(const ScalarPassArgs &a, int f, uint32_t s, bool need_p, Raw<VZ> &r)
  {
    vload<VZ, NTF>((const float *)a.phi_in[f], s, r.h);
    vload<VZ, NTF>(a.philo_in[f], s, r.l);
    if (need_p) vload<VZ, NTP>((const float *)a.pi_in[f], s, r.p);
    else
#pragma unroll
      for (int k = 0; k < VZ; ++k) r.p[k] = 0.0f;
  }
  template <bool DRIFT, int VZ>
  __device__ static __forceinline__ V value(const Raw<VZ> &r, int k, double, double, float cf)
  {
    float h = r.h[k], l = r.l[k];
    if (DRIFT) pair_add(h, l, cf * r.p[k]);
    return make_float2(h, l);
  }
};

template <int NW, int K>
__device__ __forceinline__ void block_reduce_rows(double (&v)[K], double *row)
{
  __shared__ double sh[NW][K];
#pragma unroll
  for (int k = 0; k < K; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) sh[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x < K) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += sh[w][threadIdx.x];
    row[threadIdx.x] = s;
  }
}

// LDS tile of one workgroup: plane x in buffer (x - x0) & 1.
template <int NF, bool F32, int VZ, int WPR, int TY>
using RowsTile = typename Tr<F32>::V[2][NF][TY + 2][64 * VZ * WPR];

// One workgroup's work: workgroup b of the nwg that cover this lattice.  PK: the potential's
// compile-time exponent table (0: runtime table).  KICK2 (with DRIFT and KICK): after the kick by
// a.ckick, kick again by a.ckick2 from the same drifted fields -- the kick a following kick-only pass
// would apply, with the same bits -- and sum the momenta after each kick (slots RED_PI2S and
// RED_PI2S_KICK2); only the twice-kicked momenta are stored.
template <int NF, bool F32, int VZ, int WPR, int TY, bool DRIFT, bool KICK, bool RED, int MM, int PK = 0,
          bool KICK2 = false>
__device__ __forceinline__ void rows_body(const ScalarPassArgs &a, const int XC, const int G, const int b,
                                          const int nwg, RowsTile<NF, F32, VZ, WPR, TY> &ldsbuf)
{
  using T = Tr<F32>;
  using V = typename T::V;
  using Raw = typename T::template Raw<VZ>;
  constexpr bool NT = (MM & 1) != 0, NTP = (MM & 2) != 0, NTF = (MM & 4) != 0, EDGE = (MM & 8) != 0;
  constexpr int NW = WPR * TY;  // wavefronts per workgroup
  constexpr int N = 64 * VZ * WPR;

  const uint32_t NN = (uint32_t)N * N;
  const int nTy = N / TY;
  const int L = (nwg % 8 == 0) ? (b % 8) * (nwg / 8) + b / 8 : b;
  const int ty = L % nTy, cx = L / nTy;
  const int y0 = ty * TY, x0 = cx * XC;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = wave / WPR;                    // 0..TY-1
  const int zb = (wave % WPR) * 64 * VZ + lane * VZ;  // first z of this lane
  const int y = y0 + row;

  const double fs = DRIFT ? a.st->fdrift_s : 0.0;
  const double aK = KICK ? a.st->aK : 0.0;
  const double aV = KICK ? a.st->aV : 0.0;
  const float cf = DRIFT ? (float)(a.st->fdrift_s * a.cdt) : 0.0f;
  const float kK = KICK ? (float)(a.ckick * a.st->aK * a.inv_dx2) : 0.0f;
  const float kV = KICK ? (float)(a.ckick * a.st->aV) : 0.0f;
  const float kK2 = KICK2 ? (float)(a.ckick2 * a.st->aK * a.inv_dx2) : 0.0f;
  const float kV2 = KICK2 ? (float)(a.ckick2 * a.st->aV) : 0.0f;
  static_assert(!KICK2 || (DRIFT && KICK), "a double kick follows a drift");

  auto wrapx = [&](int x) -> uint32_t { return (uint32_t)(x < 0 ? x + N : (x >= N ? x - N : x)); };
  auto rowbase = [&](int x, int yy) -> uint32_t {
    yy = yy < 0 ? yy + N : (yy >= N ? yy - N : yy);
    return wrapx(x) * NN + (uint32_t)yy * N;
  };
  // halo rows: the waves of the first tile row also load row y0-1, those of the last row y0+TY
  const bool halo = (row == 0) || (row == TY - 1);
  const int hy = row == 0 ? y0 - 1 : y0 + TY;
  const int hl = row == 0 ? 0 : TY + 1;

  // own-row loads; with EDGE the first and last tile rows (wave-uniform) use ordinary field loads
  const bool edge_row = EDGE && (row == 0 || row == TY - 1);
  auto own_load = [&](int f, uint32_t s, bool need_p, Raw &q) {
    if (edge_row)
      T::template load<VZ, false, NTP>(a, f, s, need_p, q);
    else
      T::template load<VZ, NTF, NTP>(a, f, s, need_p, q);
  };

  V prev[NF][VZ], cur[NF][VZ], nxt[NF][VZ];
  typename T::P pic[NF][VZ], pin[NF][VZ];
  Raw pf[NF], hq[NF];
  double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;

  {
  auto &lds = ldsbuf[0];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    Raw q;
    own_load(f, rowbase(x0 - 1, y) + zb, DRIFT, q);
#pragma unroll
    for (int k = 0; k < VZ; ++k) prev[f][k] = T::template value<DRIFT>(q, k, fs, a.cdt, cf);
    own_load(f, rowbase(x0, y) + zb, true, q);
#pragma unroll
    for (int k = 0; k < VZ; ++k) {
      cur[f][k] = T::template value<DRIFT>(q, k, fs, a.cdt, cf);
      pic[f][k] = q.p[k];
      lds[f][row + 1][zb + k] = cur[f][k];
    }
    own_load(f, rowbase(x0 + 1, y) + zb, true, q);
#pragma unroll
    for (int k = 0; k < VZ; ++k) {
      nxt[f][k] = T::template value<DRIFT>(q, k, fs, a.cdt, cf);
      pin[f][k] = q.p[k];
    }
    if (halo) {
      T::load(a, f, rowbase(x0, hy) + zb, DRIFT, q);
#pragma unroll
      for (int k = 0; k < VZ; ++k) lds[f][hl][zb + k] = T::template value<DRIFT>(q, k, fs, a.cdt, cf);
    }
  }
  }
  __syncthreads();

  for (int i = 0; i < XC; ++i) {
    const int x = x0 + i;
    const bool more = i + 1 < XC;
    auto &lds = ldsbuf[i & 1];
    if (more) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        own_load(f, rowbase(x + 2, y) + zb, true, pf[f]);
        if (halo) T::load(a, f, rowbase(x + 1, hy) + zb, DRIFT, hq[f]);
      }
    }

    const uint32_t s0 = rowbase(x, y) + zb;
    // z-edge neighbours of this lane from the LDS copy of its own row
    V zl[NF], zr[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      zl[f] = lds[f][row + 1][zb == 0 ? N - 1 : zb - 1];
      zr[f] = lds[f][row + 1][zb + VZ == N ? 0 : zb + VZ];
    }

    double p2 = 0.0, g2 = 0.0, vsum = 0.0, p2b = 0.0;
    typename T::P pout[NF][VZ];
#pragma unroll
    for (int k = 0; k < VZ; ++k) {
      double phid[NF];
      float phif[NF];
      double lapd[NF];
      float lapf[NF];
      float g2f = 0.0f, p2f = 0.0f, p2fb = 0.0f;  // FP32-bulk: per-site terms in FP32, accumulated in FP64
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const V c = cur[f][k];
        const V vyp = lds[f][row + 2][zb + k], vym = lds[f][row][zb + k];
        const V vzp = k + 1 < VZ ? cur[f][k + 1 < VZ ? k + 1 : k] : zr[f];
        const V vzm = k > 0 ? cur[f][k > 0 ? k - 1 : k] : zl[f];
        const V vxp = nxt[f][k], vxm = prev[f][k];
        if constexpr (!F32) {
          phid[f] = c;
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
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        if constexpr (!F32) {
          double p = pic[f][k];
          if constexpr (KICK2) {
            const double F = aK * lapd[f] - aV * dpot<NF, PK, double>(a.poly, phid, f);
            p += a.ckick * F;
            p2 += p * p;
            p += a.ckick2 * F;
            p2b += p * p;
          } else {
            if (KICK) p += a.ckick * (aK * lapd[f] - aV * dpot<NF, PK, double>(a.poly, phid, f));
            p2 += p * p;
          }
          pout[f][k] = p;
        } else {
          float p = pic[f][k];
          if constexpr (KICK2) {
            const float d = dpot<NF, PK, float>(a.poly, phif, f);
            p += kK * lapf[f] - kV * d;
            p2f += p * p;
            p += kK2 * lapf[f] - kV2 * d;
            p2fb += p * p;
          } else {
            if (KICK) p += kK * lapf[f] - kV * dpot<NF, PK, float>(a.poly, phif, f);
            p2f += p * p;
          }
          pout[f][k] = p;
        }
      }
      if (RED) {
        if constexpr (!F32) {
          vsum += pot<NF, PK, double>(a.poly, phid);
        } else {
          p2 += (double)p2f;
          if constexpr (KICK2) p2b += (double)p2fb;
          g2 += (double)g2f * a.inv_dx2;
          vsum += (double)pot<NF, PK, float>(a.poly, phif);
        }
      }
    }
    // outputs
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      if (KICK) vstore<VZ, NT>((typename T::P *)a.pi_out[f], s0, pout[f]);
      if (DRIFT) {
        if constexpr (!F32) {
          vstore<VZ, NT>((double *)a.phi_out[f], s0, cur[f]);
        } else {
          float h[VZ], l[VZ];
#pragma unroll
          for (int k = 0; k < VZ; ++k) {
            h[k] = cur[f][k].x;
            l[k] = cur[f][k].y;
          }
          vstore<VZ, NT>((float *)a.phi_out[f], s0, h);
          vstore<VZ, NT>(a.philo_out[f], s0, l);
        }
      }
    }
    if (RED) {
      acc0 += p2;
      acc1 += g2;
      acc2 += vsum;
      if constexpr (KICK2) acc3 += p2b;
    }
    if (RED && (i + 1) % G == 0) {
      double v[kNumRed] = {acc0, acc1, acc2, 0, 0, 0, 0, 0};
      v[RED_PI2S_KICK2] = acc3;
      const size_t prow = (size_t)(x / G) * nTy + ty;
      block_reduce_rows<NW, kNumRed>(v, a.partials + prow * kNumRed);
      acc0 = acc1 = acc2 = acc3 = 0.0;
    }
    if (more) {
      // the other buffer was last read in the previous iteration, which every wavefront finished
      // before the barrier that ended it
      auto &ldn = ldsbuf[(i + 1) & 1];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
#pragma unroll
        for (int k = 0; k < VZ; ++k) {
          prev[f][k] = cur[f][k];
          cur[f][k] = nxt[f][k];
          pic[f][k] = pin[f][k];
          nxt[f][k] = T::template value<DRIFT>(pf[f], k, fs, a.cdt, cf);
          pin[f][k] = pf[f].p[k];
          ldn[f][row + 1][zb + k] = cur[f][k];
        }
        if (halo)
#pragma unroll
          for (int k = 0; k < VZ; ++k) ldn[f][hl][zb + k] = T::template value<DRIFT>(hq[f], k, fs, a.cdt, cf);
      }
      __syncthreads();  // LDS tile of plane x+1 ready
    }
  }
}

template <int NF, bool F32, int VZ, int WPR, int TY, bool DRIFT, bool KICK, bool RED, int MM, int PK,
          bool KICK2 = false>
__global__ __launch_bounds__(64 * WPR * TY) void k_scalar_rows(const ScalarPassArgs a, const int XC, const int G)
{
  __shared__ RowsTile<NF, F32, VZ, WPR, TY> ldsbuf;
  rows_body<NF, F32, VZ, WPR, TY, DRIFT, KICK, RED, MM, PK, KICK2>(a, XC, G, blockIdx.x, gridDim.x, ldsbuf);
}

// kick2: the double-kick form of the drift + kick pass.
template <int NF, bool F32, int VZ, int WPR, int TY, int MM, int PK = 0>
hipError_t dispatch_rows(bool drift, bool kick, bool red, const ScalarPassArgs &a, int blocks, int XC, int G,
                         hipStream_t s, bool kick2 = false)
{
  const int code = (drift ? 4 : 0) | (kick ? 2 : 0) | (red ? 1 : 0);
  const dim3 blk(64 * WPR * TY);
  if (kick2) {
    if (!drift || !kick) return hipErrorInvalidValue;
    if (red)
      hipLaunchKernelGGL((k_scalar_rows<NF, F32, VZ, WPR, TY, true, true, true, MM, PK, true>), dim3(blocks), blk, 0, s,
                         a, XC, G);
    else
      hipLaunchKernelGGL((k_scalar_rows<NF, F32, VZ, WPR, TY, true, true, false, MM, PK, true>), dim3(blocks), blk, 0,
                         s, a, XC, G);
    return hipGetLastError();
  }
#define CLMI_SR(D, K, R)                                                                                       \
  case ((D ? 4 : 0) | (K ? 2 : 0) | (R ? 1 : 0)):                                                             \
    hipLaunchKernelGGL((k_scalar_rows<NF, F32, VZ, WPR, TY, D, K, R, MM, PK>), dim3(blocks), blk, 0, s, a, XC, G); \
    break;
  switch (code) {
    CLMI_SR(true, true, true)
    CLMI_SR(true, true, false)
    CLMI_SR(true, false, true)
    CLMI_SR(true, false, false)
    CLMI_SR(false, true, true)
    CLMI_SR(false, true, false)
    CLMI_SR(false, false, true)
    default:
      return hipErrorInvalidValue;
  }
#undef CLMI_SR
  return hipGetLastError();
}

}  // namespace
}  // namespace clmi
