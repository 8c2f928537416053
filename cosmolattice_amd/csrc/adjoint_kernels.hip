// Reverse-mode (adjoint) lattice kernels of the scalar-singlet leapfrog (clmi_adjoint_lf).
//
// Forward step n of the leapfrog (evolvers/leapfrog.h:42-89):
//   pi+  = pi- + c [aK L(phi) - aV V'(phi)]           (kick)
//   P    = <sum_f pi+_f^2>                           (storeMomentaAverages)
//   background update (scale factor)
//   phi' = phi + dt F pi+,  F = aSI^(alpha-3)          (drift)
//   G    = <sum_f |grad+ phi'_f|^2>,  U = <V(phi')>  (storeFieldsAverages)
// The kernels below apply the transposes of these maps to the adjoint fields lam_phi, lam_pi
// (FP64), reading the forward fields in their storage precision (FP64, or FP32 hi/lo phi and FP32
// pi).  With L the 7-point Laplacian (self-adjoint) and H the Hessian of V:
//   d<|grad+ phi|^2>/dphi = -(2/N^3) L(phi),  d<V>/dphi = V'(phi)/N^3,  d<pi^2>/dpi = (2/N^3) pi.
// Global sums that feed the background adjoint are written as per-block partial rows of kAdjRed
// doubles and combined in a fixed order (k_adj_finish), so results are reproducible.
#include "clmi_internal.hpp"
#include "scalar_poly.hpp"

namespace clmi {
namespace {

template <bool F32> __device__ __forceinline__ double ld_phi(const AdjFields &s, int f, uint32_t i)
{
  if (F32) return (double)((const float *)s.phi[f])[i] + (double)s.philo[f][i];
  return ((const double *)s.phi[f])[i];
}
template <bool F32> __device__ __forceinline__ double ld_pi(const AdjFields &s, int f, uint32_t i)
{
  if (F32) return (double)((const float *)s.pi[f])[i];
  return ((const double *)s.pi[f])[i];
}

struct Nbr {
  uint32_t xp, xm, yp, ym, zp, zm;
};
__device__ __forceinline__ Nbr neighbours(uint32_t s, uint32_t N)
{
  const uint32_t z = s % N, y = (s / N) % N, x = s / (N * N);
  const uint32_t NN = N * N;
  Nbr n;
  n.xp = s - x * NN + ((x + 1 == N) ? 0 : x + 1) * NN;
  n.xm = s - x * NN + ((x == 0) ? N - 1 : x - 1) * NN;
  n.yp = s - y * N + ((y + 1 == N) ? 0 : y + 1) * N;
  n.ym = s - y * N + ((y == 0) ? N - 1 : y - 1) * N;
  n.zp = s - z + ((z + 1 == N) ? 0 : z + 1);
  n.zm = s - z + ((z == 0) ? N - 1 : z - 1);
  return n;
}

// Potential monomials m_t = prod_f phi_f^e_tf and their derivatives.  The field count NF is a template
// parameter so that per-site arrays stay in registers; with PK > 0 the exponent table is
// scalar_poly.hpp's compile-time PolyTable and every loop over terms and powers unrolls.
template <int NF, int PK> __device__ __forceinline__ int n_terms(const ScalarPoly &P)
{
  if constexpr (PK == 0) return P.nterms;
  else return PolyTable<NF, PK>::nt;
}
// Static bound of the term loops: unrolled with a guard t < n_terms, so that the per-term sums index
// their accumulators with constants and stay in registers.
template <int NF, int PK> constexpr int max_terms()
{
  if constexpr (PK == 0) return CLMI_MAX_TERMS;
  else return PolyTable<NF, PK>::nt;
}
template <int NF, int PK> __device__ __forceinline__ int expo(const ScalarPoly &P, int t, int f)
{
  if constexpr (PK == 0) return P.e[t][f];
  else return PolyTable<NF, PK>::e[t][f];
}
__device__ __forceinline__ double ipow(double v, int e)
{
  double r = 1.0;
  for (int i = 0; i < e; ++i) r *= v;
  return r;
}
// prod_f phi_f^(e_tf - [f == n] - [f == m]) (n, m = -1: no derivative)
template <int NF, int PK> __device__ __forceinline__ double term(const ScalarPoly &P, int t, const double *phi, int n, int m)
{
  double p = 1.0;
#pragma unroll
  for (int f = 0; f < NF; ++f) p *= ipow(phi[f], expo<NF, PK>(P, t, f) - (f == n) - (f == m));
  return p;
}
template <int NF, int PK> __device__ __forceinline__ double mono(const ScalarPoly &P, int t, const double *phi)
{
  return term<NF, PK>(P, t, phi, -1, -1);
}
template <int NF, int PK> __device__ __forceinline__ double dmono(const ScalarPoly &P, int t, const double *phi, int n)
{
  const int e = expo<NF, PK>(P, t, n);
  return e == 0 ? 0.0 : (double)e * term<NF, PK>(P, t, phi, n, -1);
}
template <int NF, int PK>
__device__ __forceinline__ double d2mono(const ScalarPoly &P, int t, const double *phi, int n, int m)
{
  const int en = expo<NF, PK>(P, t, n), em = expo<NF, PK>(P, t, m) - (n == m);
  return (en <= 0 || em <= 0) ? 0.0 : (double)(en * em) * term<NF, PK>(P, t, phi, n, m);
}

// (A) fields adjoint of the storeFieldsAverages reductions, optionally with the momentum
// reduction, and the sums that the background adjoint needs:
//   lam_phi_f += lamG (-2/N^3) L(phi_f) + lamU V'_f(phi) / N^3
//   lam_pi_f  += lamP (2/N^3) pi_f                        (when lamP != 0 or with_pi)
//   red[0] = sum lam_phi . pi,  red[1 + t] = sum m_t(phi)
// (pi from a.fwd_red when it is set, else from a.fwd)
template <int NF, bool F32, int PK>
__global__ __launch_bounds__(kBlock) void k_adj_fields(const AdjArgs a, double lamG, double lamU, double lamP,
                                                       int with_pi)
{
  if (a.sc) {
    lamG = a.sc->lam_G;
    lamU = a.sc->lam_U;
  }
  const int nt = n_terms<NF, PK>(a.poly);
  double acc[kAdjRed];
#pragma unroll
  for (int k = 0; k < kAdjRed; ++k) acc[k] = 0.0;
  const double invV = 1.0 / (double)a.V;
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < a.V; s += gridDim.x * kBlock) {
    const Nbr nb = neighbours(s, a.N);
    double phi[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) phi[f] = ld_phi<F32>(a.fwd, f, s);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const double c = phi[f];
      const double lap = (ld_phi<F32>(a.fwd, f, nb.xp) + ld_phi<F32>(a.fwd, f, nb.xm) - 2 * c +
                          ld_phi<F32>(a.fwd, f, nb.yp) + ld_phi<F32>(a.fwd, f, nb.ym) - 2 * c +
                          ld_phi<F32>(a.fwd, f, nb.zp) + ld_phi<F32>(a.fwd, f, nb.zm) - 2 * c) *
                         a.inv_dx2;
      double dv = 0.0;
#pragma unroll
      for (int t = 0; t < max_terms<NF, PK>(); ++t)
        if (t < nt) dv += a.poly.c[t] * dmono<NF, PK>(a.poly, t, phi, f);
      double l = a.lam_phi[f][s];
      l += lamG * (-2.0 * invV) * lap + lamU * dv * invV;
      a.lam_phi[f][s] = l;
      const double p = ld_pi<F32>(a.fwd_red.pi[0] ? a.fwd_red : a.fwd, f, s);
      acc[0] += l * p;
      if (with_pi) a.lam_pi[f][s] += lamP * 2.0 * invV * p;
    }
#pragma unroll
    for (int t = 0; t < max_terms<NF, PK>(); ++t)
      if (t < nt && t + 1 < kAdjRed) acc[1 + t] += mono<NF, PK>(a.poly, t, phi);
  }
  block_reduce_store<kAdjRed>(acc, a.partials + (size_t)blockIdx.x * kAdjRed);
}

// (B) drift and <pi^2> transposes (pointwise): lam_pi_f += cF lam_phi_f + lamP (2/N^3) pi_f
template <int NF, bool F32> __global__ __launch_bounds__(kBlock) void k_adj_pi(const AdjArgs a, double cF, double lamP)
{
  if (a.sc) {
    cF = a.sc->cF;
    lamP = a.sc->lamP;
  }
  const double c2 = lamP * 2.0 / (double)a.V;
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < a.V; s += gridDim.x * kBlock)
#pragma unroll
    for (int f = 0; f < NF; ++f) a.lam_pi[f][s] += cF * a.lam_phi[f][s] + c2 * ld_pi<F32>(a.fwd, f, s);
}

// (C) kick transpose on the fields of step n:
//   lam_phi_f += c [aK L(lam_pi_f) - aV sum_g H_fg(phi) lam_pi_g]
//   red[0] = sum lam_pi . L(phi),  red[1] = sum lam_pi . V'(phi),
//   red[2 + t] = sum_f lam_pi_f d m_t / d phi_f
// FIELDS: then the fields transpose (A) of the same state, with the Laplacian and V' the kick already
// evaluated, in the order (A) applies them:
//   lam_phi_f += lamG (-2/N^3) L(phi_f) + lamU V'_f(phi) / N^3
//   red[F] = sum lam_phi . pi,  red[F + 1 + t] = sum m_t(phi),  F = kAdjFieldSlots  (pi as in (A))
template <int NF, bool F32, bool FIELDS, int PK>
__global__ __launch_bounds__(kBlock) void k_adj_kick(const AdjArgs a, double c, double aK, double aV)
{
  double lamG = 0.0, lamU = 0.0;
  if (a.sc) {
    c = a.sc->ck;
    aK = a.sc->aK;
    aV = a.sc->aV;
    lamG = a.sc->fG;
    lamU = a.sc->fU;
  }
  const int nt = n_terms<NF, PK>(a.poly);
  const double invV = 1.0 / (double)a.V;
  double acc[kAdjRed];
#pragma unroll
  for (int k = 0; k < kAdjRed; ++k) acc[k] = 0.0;
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < a.V; s += gridDim.x * kBlock) {
    const Nbr nb = neighbours(s, a.N);
    double phi[NF], lp[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      phi[f] = ld_phi<F32>(a.fwd, f, s);
      lp[f] = a.lam_pi[f][s];
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const double *L = a.lam_pi[f];
      const double lapl = (L[nb.xp] + L[nb.xm] - 2 * lp[f] + L[nb.yp] + L[nb.ym] - 2 * lp[f] + L[nb.zp] + L[nb.zm] -
                           2 * lp[f]) *
                          a.inv_dx2;
      double hl = 0.0;
#pragma unroll
      for (int g = 0; g < NF; ++g) {
        double h = 0.0;
#pragma unroll
        for (int t = 0; t < max_terms<NF, PK>(); ++t)
          if (t < nt) h += a.poly.c[t] * d2mono<NF, PK>(a.poly, t, phi, f, g);
        hl += h * lp[g];
      }
      double l = a.lam_phi[f][s] + c * (aK * lapl - aV * hl);
      const double c0 = phi[f];
      const double lapf = (ld_phi<F32>(a.fwd, f, nb.xp) + ld_phi<F32>(a.fwd, f, nb.xm) - 2 * c0 +
                           ld_phi<F32>(a.fwd, f, nb.yp) + ld_phi<F32>(a.fwd, f, nb.ym) - 2 * c0 +
                           ld_phi<F32>(a.fwd, f, nb.zp) + ld_phi<F32>(a.fwd, f, nb.zm) - 2 * c0) *
                          a.inv_dx2;
      double dv = 0.0;
#pragma unroll
      for (int t = 0; t < max_terms<NF, PK>(); ++t) {
        if (t >= nt) break;
        const double dm = dmono<NF, PK>(a.poly, t, phi, f);
        dv += a.poly.c[t] * dm;
        if (t + 2 < kAdjFieldSlots) acc[2 + t] += lp[f] * dm;
      }
      acc[0] += lp[f] * lapf;
      acc[1] += lp[f] * dv;
      if (FIELDS) {
        l += lamG * (-2.0 * invV) * lapf + lamU * dv * invV;
        acc[kAdjFieldSlots] += l * ld_pi<F32>(a.fwd_red.pi[0] ? a.fwd_red : a.fwd, f, s);
      }
      a.lam_phi[f][s] = l;
    }
    if (FIELDS)
#pragma unroll
      for (int t = 0; t < max_terms<NF, PK>(); ++t)
        if (t < nt && kAdjFieldSlots + 1 + t < kAdjRed) acc[kAdjFieldSlots + 1 + t] += mono<NF, PK>(a.poly, t, phi);
  }
  block_reduce_store<kAdjRed>(acc, a.partials + (size_t)blockIdx.x * kAdjRed);
}

// Tiled form of (C), for N a multiple of 64, one or two fields and the compile-time lphi4 potential
// (with a runtime exponent table every term's accumulator stays live and the kernel runs out of
// registers, so those runs keep k_adj_kick): a workgroup owns a tile of kAdjTY
// y-rows x 64 z-sites (each wavefront kAdjRY rows, one z-site per lane) and marches along x over XC
// planes.  lam_pi and phi of the x-neighbours come from a register queue (planes x-1, x, x+1; plane
// x+2 is loaded while x is computed), those of the y- and z-neighbours from a double-buffered LDS
// tile of the plane with a one-site halo.  Every value is read from memory once per tile instead of
// seven times; the per-site arithmetic, term order included, is k_adj_kick's, so lam_phi is
// bit-identical to it.  The pointwise inputs of a plane (lam_phi, the momenta of the fields sum) are
// loaded one plane ahead.  One partial row per workgroup.
constexpr int kAdjRY = 1, kAdjTY = 4 * kAdjRY, kAdjTZ = 64;

template <int NF, bool F32, bool FIELDS, int PK>
__global__ __launch_bounds__(kBlock) void k_adj_kick_tiled(const AdjArgs a, double c, double aK, double aV,
                                                           const int XC)
{
  static_assert(kBlock == 4 * 64, "four wavefronts per tile");
  __shared__ double lamL[2][NF][kAdjTY + 2][kAdjTZ + 2];
  __shared__ double phiL[2][NF][kAdjTY + 2][kAdjTZ + 2];
  double lamG = 0.0, lamU = 0.0;
  if (a.sc) {
    c = a.sc->ck;
    aK = a.sc->aK;
    aV = a.sc->aV;
    lamG = a.sc->fG;
    lamU = a.sc->fU;
  }
  const int nt = n_terms<NF, PK>(a.poly);
  const double invV = 1.0 / (double)a.V;
  const int N = (int)a.N;
  const uint32_t NN = (uint32_t)N * N;
  const int nTz = N / kAdjTZ, nTy = N / kAdjTY;
  const int b = blockIdx.x;
  const int tz = b % nTz, ty = (b / nTz) % nTy, cx = b / (nTz * nTy);
  const int z0 = tz * kAdjTZ, y0 = ty * kAdjTY, x0 = cx * XC;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int z = z0 + lane;
  auto site = [&](int x, int y, int zz) -> uint32_t {
    x = x < 0 ? x + N : (x >= N ? x - N : x);
    y = y < 0 ? y + N : (y >= N ? y - N : y);
    zz = zz < 0 ? zz + N : (zz >= N ? zz - N : zz);
    return (uint32_t)x * NN + (uint32_t)y * N + zz;
  };
  const AdjFields &fr = a.fwd_red.pi[0] ? a.fwd_red : a.fwd;

  // halo lanes: wave 0 -> row y0-1, wave 1 -> row y0+TY, waves 2/3 lanes < TY -> columns z0-1 / z0+TZ
  int hy = 0, hz = 0, hly = 0, hlz = 0;
  bool has_halo = false;
  if (wave == 0) { hy = y0 - 1; hz = z; hly = 0; hlz = lane + 1; has_halo = true; }
  else if (wave == 1) { hy = y0 + kAdjTY; hz = z; hly = kAdjTY + 1; hlz = lane + 1; has_halo = true; }
  else if (lane < kAdjTY) {
    hy = y0 + lane;
    hly = lane + 1;
    if (wave == 2) { hz = z0 - 1; hlz = 0; } else { hz = z0 + kAdjTZ; hlz = kAdjTZ + 1; }
    has_halo = true;
  }

  double lprev[kAdjRY][NF], lcur[kAdjRY][NF], lnxt[kAdjRY][NF];
  double pprev[kAdjRY][NF], pcur[kAdjRY][NF], pnxt[kAdjRY][NF];
  double lphi[kAdjRY][NF], pr[kAdjRY][NF];  // pointwise inputs of the plane being computed
  double acc[kAdjRed];
#pragma unroll
  for (int k = 0; k < kAdjRed; ++k) acc[k] = 0.0;

  // prologue: planes x0-1, x0, x0+1 of the own rows, the pointwise inputs of x0, the tile of x0
#pragma unroll
  for (int r = 0; r < kAdjRY; ++r) {
    const int y = y0 + wave * kAdjRY + r;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const uint32_t sm = site(x0 - 1, y, z), s0 = site(x0, y, z), sp = site(x0 + 1, y, z);
      lprev[r][f] = a.lam_pi[f][sm];
      pprev[r][f] = ld_phi<F32>(a.fwd, f, sm);
      lcur[r][f] = a.lam_pi[f][s0];
      pcur[r][f] = ld_phi<F32>(a.fwd, f, s0);
      lnxt[r][f] = a.lam_pi[f][sp];
      pnxt[r][f] = ld_phi<F32>(a.fwd, f, sp);
      lphi[r][f] = a.lam_phi[f][s0];
      pr[r][f] = FIELDS ? ld_pi<F32>(fr, f, s0) : 0.0;
      lamL[0][f][wave * kAdjRY + r + 1][lane + 1] = lcur[r][f];
      phiL[0][f][wave * kAdjRY + r + 1][lane + 1] = pcur[r][f];
    }
  }
  if (has_halo) {
    const uint32_t sh = site(x0, hy, hz);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      lamL[0][f][hly][hlz] = a.lam_pi[f][sh];
      phiL[0][f][hly][hlz] = ld_phi<F32>(a.fwd, f, sh);
    }
  }
  __syncthreads();

  for (int i = 0; i < XC; ++i) {
    const int x = x0 + i;
    const int buf = i & 1;
    const bool more = i + 1 < XC;
    // loads for later planes: own rows at x+2, pointwise inputs at x+1, halo at x+1
    double lf[kAdjRY][NF], pf[kAdjRY][NF], lphin[kAdjRY][NF], prn[kAdjRY][NF], hl[NF], hp[NF];
    if (more) {
#pragma unroll
      for (int r = 0; r < kAdjRY; ++r) {
        const int y = y0 + wave * kAdjRY + r;
        const uint32_t s2 = site(x + 2, y, z), s1 = site(x + 1, y, z);
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          lf[r][f] = a.lam_pi[f][s2];
          pf[r][f] = ld_phi<F32>(a.fwd, f, s2);
          lphin[r][f] = a.lam_phi[f][s1];
          prn[r][f] = FIELDS ? ld_pi<F32>(fr, f, s1) : 0.0;
        }
      }
      if (has_halo) {
        const uint32_t sh = site(x + 1, hy, hz);
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          hl[f] = a.lam_pi[f][sh];
          hp[f] = ld_phi<F32>(a.fwd, f, sh);
        }
      }
    }

#pragma unroll
    for (int r = 0; r < kAdjRY; ++r) {
      const int ly = wave * kAdjRY + r + 1;
      const uint32_t s = site(x, y0 + ly - 1, z);
      double phi[NF], lp[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        phi[f] = pcur[r][f];
        lp[f] = lcur[r][f];
      }
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const double lapl = (lnxt[r][f] + lprev[r][f] - 2 * lp[f] + lamL[buf][f][ly + 1][lane + 1] +
                             lamL[buf][f][ly - 1][lane + 1] - 2 * lp[f] + lamL[buf][f][ly][lane + 2] +
                             lamL[buf][f][ly][lane] - 2 * lp[f]) *
                            a.inv_dx2;
        double hsum = 0.0;
#pragma unroll
        for (int g = 0; g < NF; ++g) {
          double h = 0.0;
#pragma unroll
          for (int t = 0; t < max_terms<NF, PK>(); ++t)
            if (t < nt) h += a.poly.c[t] * d2mono<NF, PK>(a.poly, t, phi, f, g);
          hsum += h * lp[g];
        }
        double l = lphi[r][f] + c * (aK * lapl - aV * hsum);
        const double c0 = phi[f];
        const double lapf = (pnxt[r][f] + pprev[r][f] - 2 * c0 + phiL[buf][f][ly + 1][lane + 1] +
                             phiL[buf][f][ly - 1][lane + 1] - 2 * c0 + phiL[buf][f][ly][lane + 2] +
                             phiL[buf][f][ly][lane] - 2 * c0) *
                            a.inv_dx2;
        double dv = 0.0;
#pragma unroll
        for (int t = 0; t < max_terms<NF, PK>(); ++t) {
          if (t >= nt) break;
          const double dm = dmono<NF, PK>(a.poly, t, phi, f);
          dv += a.poly.c[t] * dm;
          if (t + 2 < kAdjFieldSlots) acc[2 + t] += lp[f] * dm;
        }
        acc[0] += lp[f] * lapf;
        acc[1] += lp[f] * dv;
        if (FIELDS) {
          l += lamG * (-2.0 * invV) * lapf + lamU * dv * invV;
          acc[kAdjFieldSlots] += l * pr[r][f];
        }
        a.lam_phi[f][s] = l;
      }
      if (FIELDS)
#pragma unroll
        for (int t = 0; t < max_terms<NF, PK>(); ++t)
          if (t < nt && kAdjFieldSlots + 1 + t < kAdjRed) acc[kAdjFieldSlots + 1 + t] += mono<NF, PK>(a.poly, t, phi);
    }

    if (more) {
      // tile of plane x+1, then advance the register queues
#pragma unroll
      for (int r = 0; r < kAdjRY; ++r)
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          lamL[buf ^ 1][f][wave * kAdjRY + r + 1][lane + 1] = lnxt[r][f];
          phiL[buf ^ 1][f][wave * kAdjRY + r + 1][lane + 1] = pnxt[r][f];
        }
      if (has_halo)
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          lamL[buf ^ 1][f][hly][hlz] = hl[f];
          phiL[buf ^ 1][f][hly][hlz] = hp[f];
        }
#pragma unroll
      for (int r = 0; r < kAdjRY; ++r)
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          lprev[r][f] = lcur[r][f];
          lcur[r][f] = lnxt[r][f];
          lnxt[r][f] = lf[r][f];
          pprev[r][f] = pcur[r][f];
          pcur[r][f] = pnxt[r][f];
          pnxt[r][f] = pf[r][f];
          lphi[r][f] = lphin[r][f];
          pr[r][f] = prn[r][f];
        }
    }
    __syncthreads();
  }
  block_reduce_store<kAdjRed>(acc, a.partials + (size_t)b * kAdjRed);
}

// Fixed-order sum of `rows` partial rows into red[0..kAdjRed) (shared memory) by kFinThreads threads:
// thread (g, k) adds slot k of rows g, g + NG, ... in that order (coalesced across a wavefront, the
// loads of kFinU rows in flight at once), then the NG group sums are added in group order.
constexpr int kFinThreads = 1024, kFinU = 16;
__device__ __forceinline__ void finish_rows(const double *partials, int rows, double *red)
{
  static_assert(kFinThreads % kAdjRed == 0, "row groups");
  constexpr int NG = kFinThreads / kAdjRed;
  __shared__ double sh[NG][kAdjRed];
  const int k = threadIdx.x % kAdjRed, g = threadIdx.x / kAdjRed;
  double s = 0.0;
  for (int r0 = g; r0 < rows; r0 += kFinU * NG) {
    double q[kFinU];
#pragma unroll
    for (int u = 0; u < kFinU; ++u) {
      const int r = r0 + u * NG;
      q[u] = r < rows ? partials[(size_t)r * kAdjRed + k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kFinU; ++u)
      if (r0 + u * NG < rows) s += q[u];
  }
  sh[g][k] = s;
  __syncthreads();
  if (threadIdx.x < kAdjRed) {
    double t = 0.0;
    for (int q = 0; q < NG; ++q) t += sh[q][threadIdx.x];
    red[threadIdx.x] = t;
  }
  __syncthreads();
}

__global__ __launch_bounds__(kFinThreads) void k_adj_finish(const double *partials, int rows, double *out)
{
  __shared__ double red[kAdjRed];
  finish_rows(partials, rows, red);
  if (threadIdx.x < kAdjRed) out[threadIdx.x] = red[threadIdx.x];
}

// The host arithmetic of the reverse sweep's background chain, moved to the device so that the sweep
// needs no host round trip (launch_adj_chain_fields, launch_adj_finish_chains).
__device__ void chain_fields(AdjScalars *S, const double *red, int fo, const clmi_background *tape, int n, double w,
                             double V, const ModelConsts &mc, int nterms, int ex)
{
  const double dt = mc.dt, al = mc.alpha;
  const clmi_background &bn = tape[n], &bn1 = tape[n + 1];
  for (int t = 0; t < nterms; ++t) S->gc[t] += S->lam_U * red[fo + 1 + t] / V;
  const double F = rpow(bn1.aSI, al - 3);
  double lamP = 0.0;
  S->fG = S->lam_G;
  S->fU = S->lam_U;
  if (ex) {
    // drift: aSI' = (aI + aI')/2 enters F; aI' = aI + dt aDotSI'; aDotSI' = aDotSI + dt K
    const double lam_aSI1 = dt * red[fo] * (al - 3) * rpow(bn1.aSI, al - 4);
    const double lam_aI1 = S->lam_aI + 0.5 * lam_aSI1;
    double n_aI = 0.5 * lam_aSI1 + lam_aI1;
    const double lam_aDotSI1 = S->lam_aDotSI + dt * lam_aI1;
    const double lamK = dt * lam_aDotSI1;
    const SFd d = sf_partials(mc, bn.aI, bn1.pi2AvI, bn.grad2AvI, bn.potAvI);
    const double lam_pi2I = lamK * d.dpi2;
    n_aI += lamK * d.da;
    S->n_aI = n_aI;
    S->n_aDotSI = lam_aDotSI1;
    S->n_G = lamK * d.dG;
    S->n_U = lamK * d.dU;
    S->fG = S->n_G;
    S->fU = S->n_U;
    // pi2AvI = (pi2AvSI_old + P)/2, pi2AvSI' = P
    S->n_pi2SI = 0.5 * lam_pi2I;
    lamP = 0.5 * lam_pi2I + S->lam_pi2SI;
  }
  S->cF = dt * F;
  S->lamP = lamP;
  S->ck = w * dt;
  S->aK = rpow(bn.aI, 1 + al);
  S->aV = rpow(bn.aI, 3 + al);
}

__device__ void chain_kick(AdjScalars *S, const double *red, const clmi_background *tape, int n, const ModelConsts &mc,
                           int nterms, int ex)
{
  const double al = mc.alpha, aI = tape[n].aI, ck = S->ck, aV = S->aV;
  for (int t = 0; t < nterms; ++t) S->gc[t] += -ck * aV * red[2 + t];
  if (ex) {
    S->lam_aI = S->n_aI + ck * (red[0] * (1 + al) * rpow(aI, al) - red[1] * (3 + al) * rpow(aI, 2 + al));
    S->lam_aDotSI = S->n_aDotSI;
    S->lam_pi2SI = S->n_pi2SI;
    S->lam_G = S->n_G;
    S->lam_U = S->n_U;
  }
}

__global__ void k_adj_chain_fields(AdjScalars *S, const double *red, int fo, const clmi_background *tape, int n,
                                   double w, double V, const ModelConsts mc, int nterms, int ex)
{
  if (threadIdx.x == 0) chain_fields(S, red, fo, tape, n, w, V, mc, nterms, ex);
}

// One reverse step's scalar work in one launch: the fixed-order sum of the fused kick + fields pass's
// partial rows (as k_adj_finish, also written to out), then chain_kick of step n and, for n > 0,
// chain_fields of step n - 1 on the fields sums.
__global__ __launch_bounds__(kFinThreads) void k_adj_finish_chains(const double *partials, int rows, double *out,
                                                                   AdjScalars *S, const clmi_background *tape, int n,
                                                                   double w, double V, const ModelConsts mc,
                                                                   int nterms, int ex)
{
  __shared__ double red[kAdjRed];
  finish_rows(partials, rows, red);
  if (threadIdx.x < kAdjRed) out[threadIdx.x] = red[threadIdx.x];
  if (threadIdx.x != 0) return;
  chain_kick(S, red, tape, n, mc, nterms, ex);
  if (n > 0) chain_fields(S, red, kAdjFieldSlots, tape, n - 1, w, V, mc, nterms, ex);
}

}  // namespace

hipError_t launch_adj_chain_fields(AdjScalars *S, const double *red, int fo, const clmi_background *tape, int n,
                                   double w, double V, const ModelConsts &mc, int nterms, bool ex, hipStream_t s)
{
  hipLaunchKernelGGL(k_adj_chain_fields, dim3(1), dim3(64), 0, s, S, red, fo, tape, n, w, V, mc, nterms, (int)ex);
  return hipGetLastError();
}

// Plan of the tiled kick transpose: x-chunk XC (the shortest from 4 up that leaves at most 512
// workgroups) and workgroups, one partial row each; false where it does not apply (N not a multiple of
// 64, more than two fields, a runtime potential table, more rows than the partial buffer holds, or the
// caller asks for the one-site-per-thread kernel).
bool adj_kick_tiled_plan(const AdjArgs &a, int *XC, int *blocks)
{
  const int N = (int)a.N;
  if (a.no_tiled || a.nf < 1 || a.nf > 2 || N % kAdjTZ != 0 || N % kAdjTY != 0) return false;
  if (poly_kind(a.poly, a.nf) != 1) return false;
  const long tiles = (long)(N / kAdjTZ) * (N / kAdjTY);
  int xc = 4;
  while (tiles * (N / xc) > 512 && 2 * xc <= N && N % (2 * xc) == 0) xc *= 2;
  const long b = tiles * (N / xc);
  if (b > kAdjMaxRows) return false;
  *XC = xc;
  *blocks = (int)b;
  return true;
}

int adj_blocks(uint32_t V)
{
  const uint32_t b = (V + kBlock - 1) / kBlock;
  return (int)(b < (uint32_t)kAdjMaxRows ? b : (uint32_t)kAdjMaxRows);
}


namespace {

// Dispatch over the field count (1..4), the storage precision and the potential form (compile-time
// lphi4 table for 1 or 2 fields when the run's table is lphi4's).
template <template <int, bool, int> class K, class... Args>
hipError_t adj_dispatch(const AdjArgs &a, bool f32, hipStream_t s, Args... args)
{
  const int blocks = adj_blocks(a.V);
  const int pk = (a.nf <= 2) ? poly_kind(a.poly, a.nf) : 0;
#define CLMI_ADJ(NF, F, PK)                                                                                     \
  if (a.nf == NF && f32 == F && pk == PK) {                                                                     \
    hipLaunchKernelGGL((K<NF, F, PK>::kernel), dim3(blocks), dim3(kBlock), 0, s, a, args...);                   \
    return hipGetLastError();                                                                                   \
  }
  CLMI_ADJ(1, false, 0) CLMI_ADJ(1, true, 0) CLMI_ADJ(1, false, 1) CLMI_ADJ(1, true, 1)
  CLMI_ADJ(2, false, 0) CLMI_ADJ(2, true, 0) CLMI_ADJ(2, false, 1) CLMI_ADJ(2, true, 1)
  CLMI_ADJ(3, false, 0) CLMI_ADJ(3, true, 0) CLMI_ADJ(4, false, 0) CLMI_ADJ(4, true, 0)
#undef CLMI_ADJ
  return hipErrorInvalidValue;
}

template <int NF, bool F32, int PK> struct FieldsK {
  static constexpr auto kernel = k_adj_fields<NF, F32, PK>;
};
template <int NF, bool F32, int PK> struct PiK {
  static constexpr auto kernel = k_adj_pi<NF, F32>;
};
template <int NF, bool F32, int PK> struct KickK {
  static constexpr auto kernel = k_adj_kick<NF, F32, false, PK>;
};
template <int NF, bool F32, int PK> struct KickFieldsK {
  static constexpr auto kernel = k_adj_kick<NF, F32, true, PK>;
};

hipError_t adj_finish(const AdjArgs &a, int rows, hipStream_t s)
{
  hipLaunchKernelGGL(k_adj_finish, dim3(1), dim3(kFinThreads), 0, s, a.partials, rows, a.red_out);
  return hipGetLastError();
}

// Partial rows the kick transpose writes (the tiled or the one-site-per-thread kernel).
int adj_kick_rows(const AdjArgs &a)
{
  int xc = 0, b = 0;
  return adj_kick_tiled_plan(a, &xc, &b) ? b : adj_blocks(a.V);
}

// The kick transpose with coefficients (c, aK, aV) or, with a.sc set, from a.sc.
template <bool FIELDS>
hipError_t adj_kick_dispatch(const AdjArgs &a, bool f32, double c, double aK, double aV, hipStream_t s)
{
  int xc = 0, b = 0;
  if (!adj_kick_tiled_plan(a, &xc, &b))
    return FIELDS ? adj_dispatch<KickFieldsK>(a, f32, s, c, aK, aV) : adj_dispatch<KickK>(a, f32, s, c, aK, aV);
#define CLMI_ADJT(NF, F)                                                                                      \
  if (a.nf == NF && f32 == F) {                                                                                \
    hipLaunchKernelGGL((k_adj_kick_tiled<NF, F, FIELDS, 1>), dim3(b), dim3(kBlock), 0, s, a, c, aK, aV, xc);  \
    return hipGetLastError();                                                                                  \
  }
  CLMI_ADJT(1, false) CLMI_ADJT(1, true) CLMI_ADJT(2, false) CLMI_ADJT(2, true)
#undef CLMI_ADJT
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_adj_fields(const AdjArgs &a, bool f32, double lamG, double lamU, double lamP, bool with_pi,
                             hipStream_t s)
{
  hipError_t e = adj_dispatch<FieldsK>(a, f32, s, lamG, lamU, lamP, (int)with_pi);
  return e == hipSuccess ? adj_finish(a, adj_blocks(a.V), s) : e;
}

hipError_t launch_adj_pi(const AdjArgs &a, bool f32, double cF, double lamP, hipStream_t s)
{
  return adj_dispatch<PiK>(a, f32, s, cF, lamP);
}

hipError_t launch_adj_kick(const AdjArgs &a, bool f32, double c, double aK, double aV, hipStream_t s)
{
  hipError_t e = adj_kick_dispatch<false>(a, f32, c, aK, aV, s);
  return e == hipSuccess ? adj_finish(a, adj_kick_rows(a), s) : e;
}

hipError_t launch_adj_kick_fields(const AdjArgs &a, bool f32, bool with_fields, hipStream_t s)
{
  if (!a.sc) return hipErrorInvalidValue;
  return with_fields ? adj_kick_dispatch<true>(a, f32, 0.0, 0.0, 0.0, s) : adj_kick_dispatch<false>(a, f32, 0.0, 0.0, 0.0, s);
}

hipError_t launch_adj_finish_chains(const AdjArgs &a, AdjScalars *S, const clmi_background *tape, int n, double w,
                                    double V, const ModelConsts &mc, int nterms, bool ex, hipStream_t s)
{
  hipLaunchKernelGGL(k_adj_finish_chains, dim3(1), dim3(kFinThreads), 0, s, (const double *)a.partials, adj_kick_rows(a),
                     a.red_out, S, tape, n, w, V, mc, nterms, (int)ex);
  return hipGetLastError();
}

}  // namespace clmi
