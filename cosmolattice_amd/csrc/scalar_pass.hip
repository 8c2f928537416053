// Scalar-singlet lattice pass: [deferred drift] + [kick] + [volume reductions] in one sweep.
//
// Reference expressions fused here (one TempLat sweep each in the reference):
//   drift  fldS += aSI^(alpha-3) * (dt * piS)                   evolvers/leapfrog.h:185-196
//   kick   piS(n) += (w dt) * (aI^(1+alpha) LatLapl(phi_n) - aI^(3+alpha) dV/dphi_n)
//                                                             evolvers/leapfrog.h:131-134,
//                                                             evolvers/kernels/scalarsingletkernels.h:30-40
//   <sum_n pi_n^2>, <sum_n (grad+ phi_n)^2>, <V>               definitions/averages.h:27-35, leapfrog.h:276
//
// A drift pass reads (phi, pi) of the previous step and writes phi into the other buffer of a
// double-buffered pair, because neighbouring sites need the pre-drift values; a fused
// drift+kick pass double-buffers pi as well.  The drifted value of a neighbour is recomputed from
// the same inputs with the same un-contracted operations, so it is bit-identical to the value the
// owning thread stores.
//
// FP32-bulk storage: phi = hi + lo (two floats), pi one float.  The drift increment is added to the
// pair with an error-free TwoSum, so increments far below the ulp of phi are kept in lo.  Stencil
// differences are taken on hi and lo separately.  All reductions accumulate in FP64.
#include "clmi_internal.hpp"
#include "scalar_poly.hpp"

namespace clmi {

namespace {

// Error-free sum of a float pair and a float increment: returns (hi', lo') with hi'+lo' = hi+lo+d
// up to the rounding of lo + d (Knuth TwoSum on hi and lo + d).
__device__ __forceinline__ void pair_add(float &hi, float &lo, float d)
{
  const float s = lo + d;
  const float h = hi + s;
  const float bb = h - hi;
  const float e = (hi - (h - bb)) + (s - bb);
  hi = h;
  lo = e;
}

}  // namespace

namespace {

struct Nbr {
  uint32_t c, xp, xm, yp, ym, zp, zm;
};

__device__ __forceinline__ Nbr neighbours(uint32_t s, uint32_t N, uint32_t NN)
{
  const uint32_t x = s / NN;
  const uint32_t r = s - x * NN;
  const uint32_t y = r / N;
  const uint32_t z = r - y * N;
  const uint32_t base = s - z;  // (x, y, 0)
  Nbr n;
  n.c = s;
  n.zp = base + (z + 1 == N ? 0 : z + 1);
  n.zm = base + (z == 0 ? N - 1 : z - 1);
  n.yp = s - y * N + (y + 1 == N ? 0 : y + 1) * N;
  n.ym = s - y * N + (y == 0 ? N - 1 : y - 1) * N;
  n.xp = s - x * NN + (x + 1 == N ? 0 : x + 1) * NN;
  n.xm = s - x * NN + (x == 0 ? N - 1 : x - 1) * NN;
  return n;
}

template <int NF, bool DRIFT, bool KICK, bool RED>
__global__ __launch_bounds__(kBlock) void k_scalar_pass_f64(const ScalarPassArgs a)
{
  const uint32_t N = a.N, NN = N * N, V = a.V;
  const double fs = DRIFT ? a.st->fdrift_s : 0.0;
  const double aK = KICK ? a.st->aK : 0.0;
  const double aV = KICK ? a.st->aV : 0.0;
  double acc[3] = {0.0, 0.0, 0.0};  // pi2, grad2, pot

  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < V; s += gridDim.x * kBlock) {
    const Nbr nb = neighbours(s, N, NN);
    double phic[NF], lap[NF];
    double g2 = 0.0;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const double *P = (const double *)a.phi_in[f];
      const double *Q = (const double *)a.pi_in[f];
      auto val = [&](uint32_t i) -> double {
        double v = P[i];
        if (DRIFT) v = __dadd_rn(v, __dmul_rn(fs, __dmul_rn(a.cdt, Q[i])));
        return v;
      };
      const double c = val(nb.c);
      phic[f] = c;
      if (DRIFT) ((double *)a.phi_out[f])[s] = c;
      const double vxp = val(nb.xp), vyp = val(nb.yp), vzp = val(nb.zp);
      if (KICK) {
        const double vxm = val(nb.xm), vym = val(nb.ym), vzm = val(nb.zm);
        double l = 0.0;
        l += vxp + vxm - 2 * c;
        l += vyp + vym - 2 * c;
        l += vzp + vzm - 2 * c;
        lap[f] = l * a.inv_dx2;
      }
      if (RED) {
        const double gx = (vxp - c) * a.inv_dx, gy = (vyp - c) * a.inv_dx, gz = (vzp - c) * a.inv_dx;
        g2 += gx * gx;
        g2 += gy * gy;
        g2 += gz * gz;
      }
    }
    double p2 = 0.0;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      double p = ((const double *)a.pi_in[f])[s];
      if (KICK) {
        p += a.ckick * (aK * lap[f] - aV * dpot_d<NF>(a.poly, phic, f));
        ((double *)a.pi_out[f])[s] = p;
      }
      p2 += p * p;
    }
    if (RED) {
      acc[0] += p2;
      acc[1] += g2;
      acc[2] += pot_d<NF>(a.poly, phic);
    }
  }
  if (RED) {
    double v[kNumRed] = {acc[0], acc[1], acc[2], 0, 0, 0, 0, 0};
    block_reduce_store<kNumRed>(v, a.partials + (size_t)blockIdx.x * kNumRed);
  }
}

template <int NF, bool DRIFT, bool KICK, bool RED>
__global__ __launch_bounds__(kBlock) void k_scalar_pass_f32(const ScalarPassArgs a)
{
  const uint32_t N = a.N, NN = N * N, V = a.V;
  const float cf = DRIFT ? (float)(a.st->fdrift_s * a.cdt) : 0.0f;
  const float kK = KICK ? (float)(a.ckick * a.st->aK * a.inv_dx2) : 0.0f;
  const float kV = KICK ? (float)(a.ckick * a.st->aV) : 0.0f;
  double acc[3] = {0.0, 0.0, 0.0};

  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < V; s += gridDim.x * kBlock) {
    const Nbr nb = neighbours(s, N, NN);
    float phic[NF], lap[NF];
    float g2f = 0.0f;  // per-site terms in FP32, accumulated in FP64
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const float *H = (const float *)a.phi_in[f];
      const float *L = a.philo_in[f];
      const float *Q = (const float *)a.pi_in[f];
      auto val = [&](uint32_t i, float &h, float &l) {
        h = H[i];
        l = L[i];
        if (DRIFT) pair_add(h, l, cf * Q[i]);
      };
      float ch, cl;
      val(nb.c, ch, cl);
      if (DRIFT) {
        ((float *)a.phi_out[f])[s] = ch;
        a.philo_out[f][s] = cl;
      }
      phic[f] = ch + cl;
      float xph, xpl, yph, ypl, zph, zpl;
      val(nb.xp, xph, xpl);
      val(nb.yp, yph, ypl);
      val(nb.zp, zph, zpl);
      const float dxp = (xph - ch) + (xpl - cl);
      const float dyp = (yph - ch) + (ypl - cl);
      const float dzp = (zph - ch) + (zpl - cl);
      if (KICK) {
        float xmh, xml, ymh, yml, zmh, zml;
        val(nb.xm, xmh, xml);
        val(nb.ym, ymh, yml);
        val(nb.zm, zmh, zml);
        const float dxm = (xmh - ch) + (xml - cl);
        const float dym = (ymh - ch) + (yml - cl);
        const float dzm = (zmh - ch) + (zml - cl);
        lap[f] = ((dxp + dxm) + (dyp + dym)) + (dzp + dzm);
      }
      if (RED) g2f += (dxp * dxp + dyp * dyp) + dzp * dzp;
    }
    float p2f = 0.0f;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      float p = ((const float *)a.pi_in[f])[s];
      if (KICK) {
        p += kK * lap[f] - kV * dpot_f<NF>(a.poly, phic, f);
        ((float *)a.pi_out[f])[s] = p;
      }
      p2f += p * p;
    }
    if (RED) {
      acc[0] += (double)p2f;
      acc[1] += (double)g2f * a.inv_dx2;
      acc[2] += (double)pot_f<NF>(a.poly, phic);
    }
  }
  if (RED) {
    double v[kNumRed] = {acc[0], acc[1], acc[2], 0, 0, 0, 0, 0};
    block_reduce_store<kNumRed>(v, a.partials + (size_t)blockIdx.x * kNumRed);
  }
}

template <int NF, bool F32>
hipError_t dispatch_modes(bool drift, bool kick, bool red, const ScalarPassArgs &a, int blocks, hipStream_t s)
{
  const int code = (drift ? 4 : 0) | (kick ? 2 : 0) | (red ? 1 : 0);
#define CLMI_SP(D, K, R)                                                                                        \
  case ((D ? 4 : 0) | (K ? 2 : 0) | (R ? 1 : 0)):                                                              \
    if constexpr (F32)                                                                                          \
      hipLaunchKernelGGL((k_scalar_pass_f32<NF, D, K, R>), dim3(blocks), dim3(kBlock), 0, s, a);                \
    else                                                                                                        \
      hipLaunchKernelGGL((k_scalar_pass_f64<NF, D, K, R>), dim3(blocks), dim3(kBlock), 0, s, a);                \
    break;
  switch (code) {
    CLMI_SP(true, true, true)
    CLMI_SP(true, true, false)
    CLMI_SP(true, false, true)
    CLMI_SP(true, false, false)
    CLMI_SP(false, true, true)
    CLMI_SP(false, true, false)
    CLMI_SP(false, false, true)
    default:
      return hipErrorInvalidValue;
  }
#undef CLMI_SP
  return hipGetLastError();
}

}  // namespace

int scalar_pass_blocks(uint32_t V)
{
  uint32_t b = (V + kBlock - 1) / kBlock;
  return (int)(b < (uint32_t)kMaxPassBlocks ? b : (uint32_t)kMaxPassBlocks);
}

hipError_t launch_scalar_pass(int nf, bool f32, bool drift, bool kick, bool red, const ScalarPassArgs &a,
                              int blocks, hipStream_t s)
{
  switch (nf * 2 + (f32 ? 1 : 0)) {
    case 2: return dispatch_modes<1, false>(drift, kick, red, a, blocks, s);
    case 3: return dispatch_modes<1, true>(drift, kick, red, a, blocks, s);
    case 4: return dispatch_modes<2, false>(drift, kick, red, a, blocks, s);
    case 5: return dispatch_modes<2, true>(drift, kick, red, a, blocks, s);
    case 6: return dispatch_modes<3, false>(drift, kick, red, a, blocks, s);
    case 7: return dispatch_modes<3, true>(drift, kick, red, a, blocks, s);
    case 8: return dispatch_modes<4, false>(drift, kick, red, a, blocks, s);
    case 9: return dispatch_modes<4, true>(drift, kick, red, a, blocks, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace clmi
