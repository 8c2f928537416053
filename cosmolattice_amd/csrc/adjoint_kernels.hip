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

__device__ __forceinline__ double ipow(double v, int e)
{
  double r = 1.0;
  for (int i = 0; i < e; ++i) r *= v;
  return r;
}
// monomial t and its first and second derivatives
__device__ __forceinline__ double mono(const ScalarPoly &P, int t, const double *phi, int nf)
{
  double p = 1.0;
  for (int f = 0; f < nf; ++f) p *= ipow(phi[f], P.e[t][f]);
  return p;
}
__device__ __forceinline__ double dmono(const ScalarPoly &P, int t, const double *phi, int nf, int n)
{
  const int e = P.e[t][n];
  if (e == 0) return 0.0;
  double p = (double)e;
  for (int f = 0; f < nf; ++f) p *= ipow(phi[f], f == n ? e - 1 : P.e[t][f]);
  return p;
}
__device__ __forceinline__ double d2mono(const ScalarPoly &P, int t, const double *phi, int nf, int n, int m)
{
  if (n == m) {
    const int e = P.e[t][n];
    if (e < 2) return 0.0;
    double p = (double)(e * (e - 1));
    for (int f = 0; f < nf; ++f) p *= ipow(phi[f], f == n ? e - 2 : P.e[t][f]);
    return p;
  }
  const int en = P.e[t][n], em = P.e[t][m];
  if (en == 0 || em == 0) return 0.0;
  double p = (double)(en * em);
  for (int f = 0; f < nf; ++f) p *= ipow(phi[f], f == n ? en - 1 : (f == m ? em - 1 : P.e[t][f]));
  return p;
}

// (A) fields adjoint of the storeFieldsAverages reductions, optionally with the momentum
// reduction, and the sums that the background adjoint needs:
//   lam_phi_f += lamG (-2/N^3) L(phi_f) + lamU V'_f(phi) / N^3
//   lam_pi_f  += lamP (2/N^3) pi_f                        (when lamP != 0 or with_pi)
//   red[0] = sum lam_phi . pi,  red[1 + t] = sum m_t(phi)
template <bool F32>
__global__ __launch_bounds__(kBlock) void k_adj_fields(const AdjArgs a, double lamG, double lamU, double lamP,
                                                       int with_pi)
{
  const int nf = a.nf;
  double acc[kAdjRed];
#pragma unroll
  for (int k = 0; k < kAdjRed; ++k) acc[k] = 0.0;
  const double invV = 1.0 / (double)a.V;
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < a.V; s += gridDim.x * kBlock) {
    const Nbr nb = neighbours(s, a.N);
    double phi[CLMI_MAX_SCALARS];
    for (int f = 0; f < nf; ++f) phi[f] = ld_phi<F32>(a.fwd, f, s);
    for (int f = 0; f < nf; ++f) {
      const double c = phi[f];
      const double lap = (ld_phi<F32>(a.fwd, f, nb.xp) + ld_phi<F32>(a.fwd, f, nb.xm) - 2 * c +
                          ld_phi<F32>(a.fwd, f, nb.yp) + ld_phi<F32>(a.fwd, f, nb.ym) - 2 * c +
                          ld_phi<F32>(a.fwd, f, nb.zp) + ld_phi<F32>(a.fwd, f, nb.zm) - 2 * c) *
                         a.inv_dx2;
      double dv = 0.0;
      for (int t = 0; t < a.poly.nterms; ++t) dv += a.poly.c[t] * dmono(a.poly, t, phi, nf, f);
      double l = a.lam_phi[f][s];
      l += lamG * (-2.0 * invV) * lap + lamU * dv * invV;
      a.lam_phi[f][s] = l;
      const double p = ld_pi<F32>(a.fwd, f, s);
      acc[0] += l * p;
      if (with_pi) a.lam_pi[f][s] += lamP * 2.0 * invV * p;
    }
    for (int t = 0; t < a.poly.nterms && t + 1 < kAdjRed; ++t) acc[1 + t] += mono(a.poly, t, phi, nf);
  }
  block_reduce_store<kAdjRed>(acc, a.partials + (size_t)blockIdx.x * kAdjRed);
}

// (B) drift and <pi^2> transposes (pointwise): lam_pi_f += cF lam_phi_f + lamP (2/N^3) pi_f
template <bool F32> __global__ __launch_bounds__(kBlock) void k_adj_pi(const AdjArgs a, double cF, double lamP)
{
  const double c2 = lamP * 2.0 / (double)a.V;
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < a.V; s += gridDim.x * kBlock)
    for (int f = 0; f < a.nf; ++f) a.lam_pi[f][s] += cF * a.lam_phi[f][s] + c2 * ld_pi<F32>(a.fwd, f, s);
}

// (C) kick transpose on the fields of step n:
//   lam_phi_f += c [aK L(lam_pi_f) - aV sum_g H_fg(phi) lam_pi_g]
//   red[0] = sum lam_pi . L(phi),  red[1] = sum lam_pi . V'(phi),
//   red[2 + t] = sum_f lam_pi_f d m_t / d phi_f
template <bool F32>
__global__ __launch_bounds__(kBlock) void k_adj_kick(const AdjArgs a, double c, double aK, double aV)
{
  const int nf = a.nf;
  double acc[kAdjRed];
#pragma unroll
  for (int k = 0; k < kAdjRed; ++k) acc[k] = 0.0;
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < a.V; s += gridDim.x * kBlock) {
    const Nbr nb = neighbours(s, a.N);
    double phi[CLMI_MAX_SCALARS], lp[CLMI_MAX_SCALARS];
    for (int f = 0; f < nf; ++f) {
      phi[f] = ld_phi<F32>(a.fwd, f, s);
      lp[f] = a.lam_pi[f][s];
    }
    for (int f = 0; f < nf; ++f) {
      const double *L = a.lam_pi[f];
      const double lapl = (L[nb.xp] + L[nb.xm] - 2 * lp[f] + L[nb.yp] + L[nb.ym] - 2 * lp[f] + L[nb.zp] + L[nb.zm] -
                           2 * lp[f]) *
                          a.inv_dx2;
      double hl = 0.0;
      for (int g = 0; g < nf; ++g) {
        double h = 0.0;
        for (int t = 0; t < a.poly.nterms; ++t) h += a.poly.c[t] * d2mono(a.poly, t, phi, nf, f, g);
        hl += h * lp[g];
      }
      a.lam_phi[f][s] += c * (aK * lapl - aV * hl);
      const double c0 = phi[f];
      const double lapf = (ld_phi<F32>(a.fwd, f, nb.xp) + ld_phi<F32>(a.fwd, f, nb.xm) - 2 * c0 +
                           ld_phi<F32>(a.fwd, f, nb.yp) + ld_phi<F32>(a.fwd, f, nb.ym) - 2 * c0 +
                           ld_phi<F32>(a.fwd, f, nb.zp) + ld_phi<F32>(a.fwd, f, nb.zm) - 2 * c0) *
                          a.inv_dx2;
      double dv = 0.0;
      for (int t = 0; t < a.poly.nterms; ++t) {
        const double dm = dmono(a.poly, t, phi, nf, f);
        dv += a.poly.c[t] * dm;
        if (t + 2 < kAdjRed) acc[2 + t] += lp[f] * dm;
      }
      acc[0] += lp[f] * lapf;
      acc[1] += lp[f] * dv;
    }
  }
  block_reduce_store<kAdjRed>(acc, a.partials + (size_t)blockIdx.x * kAdjRed);
}

// Fixed-order sum of `rows` partial rows into out[0..kAdjRed): thread (g, k) adds slot k of rows g,
// g + NG, ... (coalesced across a wavefront, eight loads in flight), then the NG group sums are added
// in group order.
__global__ __launch_bounds__(kBlock) void k_adj_finish(const double *partials, int rows, double *out)
{
  static_assert(kBlock % kAdjRed == 0, "row groups");
  constexpr int NG = kBlock / kAdjRed;
  __shared__ double sh[NG][kAdjRed];
  const int k = threadIdx.x % kAdjRed, g = threadIdx.x / kAdjRed;
  double s = 0.0;
#pragma unroll 8
  for (int r = g; r < rows; r += NG) s += partials[(size_t)r * kAdjRed + k];
  sh[g][k] = s;
  __syncthreads();
  if (threadIdx.x < kAdjRed) {
    double t = 0.0;
    for (int q = 0; q < NG; ++q) t += sh[q][threadIdx.x];
    out[threadIdx.x] = t;
  }
}

}  // namespace

int adj_blocks(uint32_t V)
{
  const uint32_t b = (V + kBlock - 1) / kBlock;
  return (int)(b < 1024 ? b : 1024);
}

hipError_t launch_adj_fields(const AdjArgs &a, bool f32, double lamG, double lamU, double lamP, bool with_pi,
                             hipStream_t s)
{
  const int blocks = adj_blocks(a.V);
  if (f32)
    hipLaunchKernelGGL(k_adj_fields<true>, dim3(blocks), dim3(kBlock), 0, s, a, lamG, lamU, lamP, (int)with_pi);
  else
    hipLaunchKernelGGL(k_adj_fields<false>, dim3(blocks), dim3(kBlock), 0, s, a, lamG, lamU, lamP, (int)with_pi);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_adj_finish, dim3(1), dim3(kBlock), 0, s, a.partials, blocks, a.red_out);
  return hipGetLastError();
}

hipError_t launch_adj_pi(const AdjArgs &a, bool f32, double cF, double lamP, hipStream_t s)
{
  const int blocks = adj_blocks(a.V);
  if (f32)
    hipLaunchKernelGGL(k_adj_pi<true>, dim3(blocks), dim3(kBlock), 0, s, a, cF, lamP);
  else
    hipLaunchKernelGGL(k_adj_pi<false>, dim3(blocks), dim3(kBlock), 0, s, a, cF, lamP);
  return hipGetLastError();
}

hipError_t launch_adj_kick(const AdjArgs &a, bool f32, double c, double aK, double aV, hipStream_t s)
{
  const int blocks = adj_blocks(a.V);
  if (f32)
    hipLaunchKernelGGL(k_adj_kick<true>, dim3(blocks), dim3(kBlock), 0, s, a, c, aK, aV);
  else
    hipLaunchKernelGGL(k_adj_kick<false>, dim3(blocks), dim3(kBlock), 0, s, a, c, aK, aV);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_adj_finish, dim3(1), dim3(kBlock), 0, s, a.partials, blocks, a.red_out);
  return hipGetLastError();
}

}  // namespace clmi
