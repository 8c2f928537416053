// Charged scalar + U(1) lattice pass, x-marching with LDS plane tiles (N a multiple of 64).
//
// A workgroup of TY wavefronts (TY = 4 or 8) owns a tile of TY y-rows x 64 z-sites (one site per
// lane) and marches along x over XC planes.  The drifted values of all five field components (phi_0, phi_1, A_x, A_y,
// A_z) of planes x-1, x and x+1, each with a one-site y/z halo including corners, sit in a
// three-slot LDS ring, so every neighbour the covariant Laplacian, the gauge Laplacian, the mixed
// derivatives d-_j d+_i A_j (the diagonal sites x+i-j) and the links need is read from LDS; each
// field value is loaded from memory and drifted once per tile instead of once per use.  The raw
// values of plane x+2 are loaded while plane x is computed; the momenta of a thread's own site stay
// in registers from that load until the kick uses them.  The per-site arithmetic is
// gauge_site.hpp's, so fields are bit-identical to the one-site-per-thread pass.  Volume sums are
// written per (group of G planes, tile), so they do not depend on the x-chunk XC: every XC gives the
// same partial rows in the same order, and the autotuner may pick any of them.
#include <algorithm>

#include "gauge_site.hpp"

namespace clmi {
namespace {

constexpr int GTZ = 64;  // z sites per tile (one per lane)
// planes per partial-sum row: the largest power of two (so that it divides N, a multiple of 64) not
// above N/8, between 8 and 32 (2048 rows at 256^3 for the background update)
inline int gauge_rows_g(int N)
{
  const int target = std::min(32, N / 8);
  int G = 8;
  while (2 * G <= target) G *= 2;
  return G;
}
constexpr int GNC = 5;   // components: phi_0, phi_1, A_x, A_y, A_z

// Raw (not yet drifted) values of one site's five components.
template <bool F32> struct GRaw;
template <> struct GRaw<false> {
  double f[GNC], p[GNC];
  __device__ __forceinline__ void load(const GaugePassArgs &a, uint32_t s, bool u1, bool drift)
  {
    for (int q = 0; q < 2; ++q) {
      f[q] = ((const double *)a.phi_in[q])[s];
      p[q] = drift ? ((const double *)a.pi_in[q])[s] : 0.0;
    }
    for (int d = 0; d < 3; ++d) {
      f[2 + d] = u1 ? ((const double *)a.A_in[d])[s] : 0.0;
      p[2 + d] = (u1 && drift) ? ((const double *)a.piA_in[d])[s] : 0.0;
    }
  }
  // drifted values exactly as GVal<false>::fld computes them
  __device__ __forceinline__ double value(int q, bool drift, double cs, double cu) const
  {
    double v = f[q];
    if (drift) v += (q < 2 ? cs : cu) * p[q];
    return v;
  }
};
template <> struct GRaw<true> {
  float h[GNC], l[GNC], p[GNC];
  __device__ __forceinline__ void load(const GaugePassArgs &a, uint32_t s, bool u1, bool drift)
  {
    for (int q = 0; q < 2; ++q) {
      h[q] = ((const float *)a.phi_in[q])[s];
      l[q] = a.philo_in[q][s];
      p[q] = drift ? ((const float *)a.pi_in[q])[s] : 0.0f;
    }
    for (int d = 0; d < 3; ++d) {
      h[2 + d] = u1 ? ((const float *)a.A_in[d])[s] : 0.0f;
      l[2 + d] = u1 ? a.Alo_in[d][s] : 0.0f;
      p[2 + d] = (u1 && drift) ? ((const float *)a.piA_in[d])[s] : 0.0f;
    }
  }
  __device__ __forceinline__ P2 value(int q, bool drift, float cs, float cu) const
  {
    P2 v{h[q], l[q]};
    if (drift) pair_add(v.h, v.l, (q < 2 ? cs : cu) * p[q]);
    return v;
  }
};

// Outputs are written with non-temporal stores: nothing in the pass reads them back, and keeping them
// out of L2 leaves room for the halo rows that neighbouring tiles re-read (256^3: 663 -> 643 us FP64,
// 474 -> 461 us FP32-bulk).
template <bool F32> __device__ __forceinline__ void nt_store_v(void *f, float *lo, uint32_t i, typename GVal<F32>::V v)
{
  if constexpr (F32) {
    __builtin_nontemporal_store(v.h, (float *)f + i);
    __builtin_nontemporal_store(v.l, lo + i);
  } else {
    __builtin_nontemporal_store(v, (double *)f + i);
  }
}

// U1: the lattice has a U(1) field, known at compile time (1) or read from the arguments (0).  Fixing it
// removes the per-site selects and branches on it; that pays for FP32-bulk (256^3 pass 479 -> 462 us)
// and not for FP64 (670 -> 683 us), so only FP32-bulk lattices with a U(1) field use U1 = 1.
template <bool F32, bool DRIFT, bool KICK, bool RED, int GTY, int U1 = 0>
__global__ __launch_bounds__(64 * GTY) void k_gauge_tiled(const GaugePassArgs a, const int XC, const int RG)
{
  using G = GVal<F32>;
  using V = typename G::V;
  using P = typename G::P;
  using DT = typename std::conditional<F32, float, double>::type;  // drift factor type
  __shared__ V lds[3][GNC][GTY + 2][GTZ + 2];

  const int N = (int)a.N;
  const uint32_t NN = (uint32_t)N * N;
  const int nTz = N / GTZ, nTy = N / GTY;
  const int nwg = gridDim.x, b = blockIdx.x;
  const int L = (nwg % 8 == 0) ? (b % 8) * (nwg / 8) + b / 8 : b;
  const int tz = L % nTz, ty = (L / nTz) % nTy, cx = L / (nTz * nTy);
  const int z0 = tz * GTZ, y0 = ty * GTY, x0 = cx * XC;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int y = y0 + wave, z = z0 + lane;
  GaugeK k = gauge_consts(a, KICK);
  if constexpr (U1 == 1) k.u1 = true;
  const bool u1 = k.u1;
  DT cs = 0, cu = 0;
  if (DRIFT) {
    if constexpr (F32) {
      cs = (float)(a.st->fdrift_s * a.cdt);
      cu = (float)(a.st->fdrift_u * a.cdt);
    } else {
      cs = a.st->fdrift_s * a.cdt;
      cu = a.st->fdrift_u * a.cdt;
    }
  }

  auto wrap = [&](int v) -> int { return v < 0 ? v + N : (v >= N ? v - N : v); };
  auto site = [&](int xx, int yy, int zz) -> uint32_t {
    return (uint32_t)wrap(xx) * NN + (uint32_t)wrap(yy) * N + (uint32_t)wrap(zz);
  };

  // halo cell of this thread (2 * 66 + 2 * GTY of 64 * GTY threads): rows y0-1 and y0+GTY with corners,
  // columns z0-1, z0+64
  int hly = -1, hlz = 0, hy = 0, hz = 0;
  if (t < GTZ + 2) {
    hly = 0; hlz = t; hy = y0 - 1; hz = z0 - 1 + t;
  } else if (t < 2 * (GTZ + 2)) {
    hly = GTY + 1; hlz = t - (GTZ + 2); hy = y0 + GTY; hz = z0 - 1 + hlz;
  } else if (t < 2 * (GTZ + 2) + GTY) {
    const int r = t - 2 * (GTZ + 2);
    hly = r + 1; hlz = 0; hy = y0 + r; hz = z0 - 1;
  } else if (t < 2 * (GTZ + 2) + 2 * GTY) {
    const int r = t - 2 * (GTZ + 2) - GTY;
    hly = r + 1; hlz = GTZ + 1; hy = y0 + r; hz = z0 + GTZ;
  }
  const bool halo = hly >= 0;

  auto put = [&](int slot, const GRaw<F32> &own, const GRaw<F32> &hal, bool has_halo) {
    for (int q = 0; q < GNC; ++q) {
      lds[slot][q][wave + 1][lane + 1] = own.value(q, DRIFT, cs, cu);
      if (has_halo) lds[slot][q][hly][hlz] = hal.value(q, DRIFT, cs, cu);
    }
  };
  auto slot_of = [&](int xx) -> int { return (xx - x0 + 1) % 3; };

  // own-site momenta of planes x and x + 1 (read with the fields when drifting, else here)
  P pc[GNC], pn[GNC];
  auto keep_p = [&](const GRaw<F32> &r, P (&dst)[GNC]) {
    for (int q = 0; q < GNC; ++q) dst[q] = r.p[q];
  };
  auto load_p = [&](uint32_t s, P (&dst)[GNC]) {
    for (int q = 0; q < 2; ++q) dst[q] = ((const P *)a.pi_in[q])[s];
    for (int d = 0; d < 3; ++d) dst[2 + d] = u1 ? ((const P *)a.piA_in[d])[s] : P(0);
  };

  // prologue: planes x0-1, x0, x0+1
  for (int xx = x0 - 1; xx <= x0 + 1; ++xx) {
    GRaw<F32> own, hal;
    own.load(a, site(xx, y, z), u1, DRIFT);
    if (halo) hal.load(a, site(xx, hy, hz), u1, DRIFT);
    put(slot_of(xx), own, hal, halo);
    if (KICK && xx == x0) { if (DRIFT) keep_p(own, pc); else load_p(site(xx, y, z), pc); }
    if (KICK && xx == x0 + 1) { if (DRIFT) keep_p(own, pn); else load_p(site(xx, y, z), pn); }
  }
  __syncthreads();

  double acc[5] = {0, 0, 0, 0, 0};
  const int ly = wave + 1, lz = lane + 1;
  for (int i = 0; i < XC; ++i) {
    const int x = x0 + i;
    const bool more = i + 1 < XC;
    GRaw<F32> pown, phal;
    if (more) {
      pown.load(a, site(x + 2, y, z), u1, DRIFT);
      if (halo) phal.load(a, site(x + 2, hy, hz), u1, DRIFT);
    }
    const int sc = slot_of(x), sp = slot_of(x + 1), sm = slot_of(x - 1);
    const auto &cur = lds[sc];
    const auto &nx = lds[sp];
    const auto &pv = lds[sm];
    GaugeIn<V> g;
    for (int q = 0; q < 2; ++q) {
      g.fc[q] = cur[q][ly][lz];
      g.fp[0][q] = nx[q][ly][lz];
      g.fm[0][q] = pv[q][ly][lz];
      g.fp[1][q] = cur[q][ly + 1][lz];
      g.fm[1][q] = cur[q][ly - 1][lz];
      g.fp[2][q] = cur[q][ly][lz + 1];
      g.fm[2][q] = cur[q][ly][lz - 1];
    }
    for (int j = 0; j < 3; ++j) {
      const int c = 2 + j;
      g.Ac[j] = cur[c][ly][lz];
      g.Ap[0][j] = nx[c][ly][lz];
      g.Ap[1][j] = cur[c][ly + 1][lz];
      g.Ap[2][j] = cur[c][ly][lz + 1];
      g.Am[0][j] = pv[c][ly][lz];
      g.Am[1][j] = cur[c][ly - 1][lz];
      g.Am[2][j] = cur[c][ly][lz - 1];
    }
    // A_j(x + i - j), i != j
    g.Ad[0][0] = g.Ad[1][1] = g.Ad[2][2] = V{};
    g.Ad[0][1] = nx[3][ly - 1][lz];
    g.Ad[0][2] = nx[4][ly][lz - 1];
    g.Ad[1][0] = pv[2][ly + 1][lz];
    g.Ad[1][2] = cur[4][ly + 1][lz - 1];
    g.Ad[2][0] = pv[2][ly][lz + 1];
    g.Ad[2][1] = cur[3][ly - 1][lz + 1];

    const uint32_t s = site(x, y, z);
    if (DRIFT) {
      for (int q = 0; q < 2; ++q) nt_store_v<F32>(a.phi_out[q], a.philo_out[q], s, g.fc[q]);
      if (u1)
        for (int d = 0; d < 3; ++d) nt_store_v<F32>(a.A_out[d], a.Alo_out[d], s, g.Ac[d]);
    }
    P pin[2], pAin[3] = {0, 0, 0}, pout[2], pAout[3];
    if (KICK) {
      for (int q = 0; q < 2; ++q) pin[q] = pc[q];
      if (u1)
        for (int d = 0; d < 3; ++d) pAin[d] = pc[2 + d];
    } else {
      for (int q = 0; q < 2; ++q) pin[q] = ((const P *)a.pi_in[q])[s];
      if (u1)
        for (int d = 0; d < 3; ++d) pAin[d] = ((const P *)a.piA_in[d])[s];
    }
    if constexpr (F32)
      gauge_site_f32<KICK, RED>(k, a.poly, g, pin, pAin, pout, pAout, acc);
    else
      gauge_site_f64<KICK, RED>(k, a.poly, g, pin, pAin, pout, pAout, acc);
    if (KICK) {
      for (int q = 0; q < 2; ++q) __builtin_nontemporal_store(pout[q], (P *)a.pi_out[q] + s);
      if (u1)
        for (int d = 0; d < 3; ++d) __builtin_nontemporal_store(pAout[d], (P *)a.piA_out[d] + s);
    }
    if (RED && (i + 1) % RG == 0) {
      double v[kNumRed] = {0, 0, acc[2], acc[0], acc[1], acc[3], acc[4], 0};
      const size_t row = ((size_t)(x / RG) * nTy + ty) * nTz + tz;
      block_reduce_store_nt<kNumRed, 64 * GTY>(v, a.partials + row * kNumRed);
      for (int k2 = 0; k2 < 5; ++k2) acc[k2] = 0.0;
    }
    if (more) {
      if (KICK) {
        for (int q = 0; q < GNC; ++q) pc[q] = pn[q];
        if (DRIFT) keep_p(pown, pn);
        else if (i + 2 < XC) load_p(site(x + 2, y, z), pn);
      }
      __syncthreads();  // every wavefront is done with slot sm (plane x - 1)
      put(sm, pown, phal, halo);  // plane x + 2 takes its slot
      __syncthreads();
    }
  }
}

template <bool F32, int GTY, int U1 = 0>
hipError_t dispatch_tiled(bool drift, bool kick, bool red, const GaugePassArgs &a, int blocks, int XC, hipStream_t s)
{
  const int G = gauge_rows_g((int)a.N);
  const int code = (drift ? 4 : 0) | (kick ? 2 : 0) | (red ? 1 : 0);
#define CLMI_GT(D, K, R)                                                                                  \
  case ((D ? 4 : 0) | (K ? 2 : 0) | (R ? 1 : 0)):                                                        \
    hipLaunchKernelGGL((k_gauge_tiled<F32, D, K, R, GTY, U1>), dim3(blocks), dim3(64 * GTY), 0, s, a, XC, G); \
    break;
  switch (code) {
    CLMI_GT(true, true, true)
    CLMI_GT(true, true, false)
    CLMI_GT(true, false, true)
    CLMI_GT(true, false, false)
    CLMI_GT(false, true, true)
    CLMI_GT(false, true, false)
    CLMI_GT(false, false, true)
    default:
      return hipErrorInvalidValue;
  }
#undef CLMI_GT
  return hipGetLastError();
}

}  // namespace

// Shape of the tiled pass with TY-row tiles (4 or 8) and x-chunk XC (a power of two dividing N, a multiple
// of the row group G): workgroups and partial-sum rows; false when N does not fit the tile.
bool gauge_tiled_shape(int N, int ty, int XC, int *blocks, int *rows)
{
  if ((ty != 4 && ty != 8) || N % GTZ != 0 || N % ty != 0) return false;
  const int G = gauge_rows_g(N);
  if (XC < G || XC > N || N % XC != 0 || XC % G != 0 || (XC & (XC - 1)) != 0) return false;
  const long tiles = (long)(N / ty) * (N / GTZ);
  *blocks = (int)(tiles * (N / XC));
  *rows = (int)(tiles * (N / G));
  return true;
}

// Default plan: the longest x-chunk that still gives every CU two workgroups (two 4-wave workgroups
// fill a CU at the pass's two waves per SIMD), else the shortest; longer chunks re-read fewer x-halo
// planes.  Measured at 256^3: XC 16 -> 128 takes the pass from 706 to 652 us (FP64) and 546 to 475 us
// (FP32-bulk).
bool gauge_tiled_plan(int N, int ty, int ncu, int *XC, int *blocks, int *rows)
{
  int best = 0;
  for (int xc = 8; xc <= N; xc *= 2) {
    int b = 0, r = 0;
    if (!gauge_tiled_shape(N, ty, xc, &b, &r)) continue;
    if (best == 0 || b >= 2 * ncu) best = xc;
  }
  if (best == 0) return false;
  *XC = best;
  return gauge_tiled_shape(N, ty, best, blocks, rows);
}

hipError_t launch_gauge_tiled(bool f32, int ty, bool drift, bool kick, bool red, const GaugePassArgs &a, int blocks,
                              int XC, hipStream_t s)
{
  const bool u1 = a.has_u1 != 0;
  if (ty == 8)
    return f32 ? (u1 ? dispatch_tiled<true, 8, 1>(drift, kick, red, a, blocks, XC, s)
                     : dispatch_tiled<true, 8>(drift, kick, red, a, blocks, XC, s))
               : dispatch_tiled<false, 8>(drift, kick, red, a, blocks, XC, s);
  if (ty == 4)
    return f32 ? (u1 ? dispatch_tiled<true, 4, 1>(drift, kick, red, a, blocks, XC, s)
                     : dispatch_tiled<true, 4>(drift, kick, red, a, blocks, XC, s))
               : dispatch_tiled<false, 4>(drift, kick, red, a, blocks, XC, s);
  return hipErrorInvalidValue;
}

}  // namespace clmi
