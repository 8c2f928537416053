// Device-side background update (used by the update kernel).
//
// sum_partials: fixed-order sum of the per-workgroup partial rows (bit-reproducible for a given
// block size).  apply_update: one thread advances the scale factor exactly as the reference's host
// code does and refreshes the kick/drift coefficients the next pass reads.
#pragma once

#include "clmi_internal.hpp"

namespace clmi {

// The scale-factor kernel of ScaleFactorKernels::get/getSI (evolvers/kernels/scalefactorkernels.h:27-124)
// with the energy normalisations of definitions/energies.h:37-85, from the given volume averages.
__device__ inline double sf_kernel_of(double a, double pi2S, double grad2S, double pi2CS, double grad2CS,
                                      double pi2U1, double mag2U1, double pot, const ModelConsts &mc)
{
  // a^-2, a^-4, a^-6 from one division (a serial chain on the updating thread)
  const double ia = 1.0 / a, ia2 = ia * ia, ia4 = ia2 * ia2, ia6 = ia4 * ia2;
  double Eks = 0, Egs = 0, Ekcs = 0, Egcs = 0, Eel = 0, Emag = 0;
  if (mc.ns > 0) {
    Eks = 0.5 * pi2S * ia6;
    Egs = 0.5 * grad2S * ia2;
  }
  if (mc.ncs > 0) {
    Ekcs = pi2CS * ia6;
    Egcs = grad2CS * ia2;
  }
  if (mc.nu1 > 0) {
    const double w2 = (mc.omegaStar / mc.fStar) * (mc.omegaStar / mc.fStar);
    Eel = 0.5 * w2 * ia4 * pi2U1;
    Emag = 0.5 * w2 * ia4 * mag2U1;
  }
  const double fm = mc.fStar / mc.MPl;
  return rpow(a, 2 * mc.alpha + 1) / 3.0 * (fm * fm) *
         ((mc.alpha - 2) * (Eks + Ekcs) + mc.alpha * (Egs + Egcs) + (mc.alpha - 1) * (Eel + Emag) +
          (mc.alpha + 1) * pot);
}

// ScaleFactorKernels::get: integer-time averages.
__device__ inline double sf_kernel(const clmi_background &b, const ModelConsts &mc)
{
  return sf_kernel_of(b.aI, b.pi2AvI, b.grad2AvI, b.CSpi2AvI, b.CSgrad2AvI, b.U1pi2AvI, b.U1Mag2AvI, b.potAvI, mc);
}

// Fixed-order sum of `rows` partial rows by a block of NT threads: thread t adds rows t, t + NT, ... in
// that order, four rows in flight per batch (each row one 64-byte line, read as four 16-byte loads),
// then wavefront shuffles, then the wavefront sums in order.  Result in out[0..kNumRed)
// (shared memory); ends with a barrier.  The rows are 64-byte aligned (kNumRed = 8 doubles).
template <int NT>
__device__ inline void sum_partials(const double *partials, int rows, double *out)
{
  static_assert(kNumRed == 8, "a partial row is four 16-byte vectors");
  typedef double d2 __attribute__((ext_vector_type(2)));
  __shared__ double sh[NT / 64][kNumRed];
  const int t = threadIdx.x;
  const d2 *P = (const d2 *)partials;
  double v[kNumRed];
#pragma unroll
  for (int k = 0; k < kNumRed; ++k) v[k] = 0.0;
  // rows t, t + NT, t + 2 NT, ... in that order, U rows per batch with all their loads in flight (a
  // thread has at most U rows below U * NT rows, so those take one memory round trip)
  constexpr int U = 4;
  for (int r0 = t; r0 < rows; r0 += U * NT) {
    d2 q[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (r0 + u * NT < rows)
#pragma unroll
        for (int k = 0; k < 4; ++k) q[u][k] = P[(size_t)(r0 + u * NT) * 4 + k];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (r0 + u * NT < rows)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          v[2 * k] += q[u][k].x;
          v[2 * k + 1] += q[u][k].y;
        }
  }
#pragma unroll
  for (int k = 0; k < kNumRed; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  if ((t & 63) == 0)
#pragma unroll
    for (int k = 0; k < kNumRed; ++k) sh[t >> 6][k] = v[k];
  __syncthreads();
  if (t < kNumRed) {
    double s = 0.0;
    for (int w = 0; w < NT / 64; ++w) s += sh[w][t];
    out[t] = s;
  }
  __syncthreads();
}

// Executed by one thread.  red: finished volume averages (sums / N^3) of the last pass.
__device__ inline void apply_update(DevState *st, const double *red, const ModelConsts &mc, const UpdateOps &ops)
{
  clmi_background &b = st->bg;
  const double dt = mc.dt;
  double a_drift = -1.0;  // > 0: the drift coefficients use this scale factor instead of aSI
  for (int o = 0; o < ops.n; ++o) {
    const double w = ops.w[o];
    switch (ops.op[o]) {
      case OP_STORE_FIELDS:  // leapfrog.h:269-277
        if (mc.ns > 0) b.grad2AvI = red[RED_GRAD2S];
        if (mc.ncs > 0) b.CSgrad2AvI = red[RED_GRAD2CS];
        if (mc.nu1 > 0) b.U1Mag2AvI = red[RED_B2U1];
        b.potAvI = red[RED_POT];
        break;
      case OP_LF_MOMENTA:  // leapfrog.h:239-267 then 124-129 and 168-184
        if (mc.ns > 0) {
          b.pi2AvSIM = b.pi2AvSI;
          b.pi2AvSI = red[RED_PI2S];
          b.pi2AvI = 0.5 * (b.pi2AvSIM + b.pi2AvSI);
        }
        if (mc.ncs > 0) {
          b.CSpi2AvSIM = b.CSpi2AvSI;
          b.CSpi2AvSI = red[RED_PI2CS];
          b.CSpi2AvI = 0.5 * (b.CSpi2AvSIM + b.CSpi2AvSI);
        }
        if (mc.nu1 > 0) {
          b.U1pi2AvSIM = b.U1pi2AvSI;
          b.U1pi2AvSI = red[RED_PI2U1];
          b.U1pi2AvI = 0.5 * (b.U1pi2AvSIM + b.U1pi2AvSI);
        }
        st->K = sf_kernel(b, mc);
        b.aDotSIM = b.aDotSI;
        b.aDotSI += dt * st->K;
        b.aDotI = (b.aDotSIM + b.aDotSI) / 2.0;
        b.aIM = b.aI;
        b.aI += dt * b.aDotSI;
        b.aSI = (b.aIM + b.aI) / 2.0;
        break;
      case OP_LF_SYNC:  // leapfrog.h:105-112
        if (mc.ns > 0) b.pi2AvI = red[RED_PI2S];
        if (mc.ncs > 0) b.CSpi2AvI = red[RED_PI2CS];
        if (mc.nu1 > 0) b.U1pi2AvI = red[RED_PI2U1];
        st->K = sf_kernel(b, mc);
        b.aDotI = b.aDotSI + dt / 2.0 * st->K;
        break;
      case OP_LF_INIT:  // leapfrog.h:35
        st->K = sf_kernel(b, mc);
        b.aDotSI = b.aDotI - dt / 2.0 * st->K;
        break;
      case OP_VV_SF_HALF:  // velocityverlet.h:165-168
        st->K = sf_kernel(b, mc);
        b.aDotSI = b.aDotI + (dt / 2) * st->K * w;
        break;
      case OP_VV_MOMENTA:  // velocityverlet.h:249-256
        if (mc.ns > 0) b.pi2AvI = red[RED_PI2S];
        if (mc.ncs > 0) b.CSpi2AvI = red[RED_PI2CS];
        if (mc.nu1 > 0) b.U1pi2AvI = red[RED_PI2U1];
        break;
      case OP_VV_MOMENTA_KICK2:  // velocityverlet.h:249-256, scalar singlets only
        b.pi2AvI = red[RED_PI2S_KICK2];
        break;
      case OP_VV_DRIFT_SF:  // velocityverlet.h:239-242
        b.aIM = b.aI;
        b.aI += dt * b.aDotSI * w;
        b.aSI = (b.aIM + b.aI) / 2.0;
        break;
      case OP_VV_SF_ONE:  // velocityverlet.h:171-174
        st->K = sf_kernel(b, mc);
        b.aDotI = b.aDotSI + (dt / 2) * st->K * w;
        break;
      case OP_SET_ALL:  // averages.h:86-116
        b.potAvI = red[RED_POT];
        b.potAvSI = b.potAvI;
        if (mc.ns > 0) {
          b.pi2AvI = red[RED_PI2S];
          b.pi2AvSI = b.pi2AvI;
          b.grad2AvI = red[RED_GRAD2S];
        }
        if (mc.nu1 > 0) {
          b.U1pi2AvI = red[RED_PI2U1];
          b.U1pi2AvSI = b.U1pi2AvI;
          b.U1Mag2AvI = red[RED_B2U1];
        }
        if (mc.ncs > 0) {
          b.CSpi2AvI = red[RED_PI2CS];
          b.CSpi2AvSI = b.CSpi2AvI;
          b.CSgrad2AvI = red[RED_GRAD2CS];
        }
        break;
      case OP_KERNEL_ONLY:
        st->K = sf_kernel(b, mc);
        break;
      case OP_PV_SF_FIRST:  // positionverlet.h:196-209; the first drift of the stage uses aI before it
        a_drift = b.aI;
        b.aIM = b.aI;
        b.aI += dt / 2.0 * b.aDotI * w;
        b.aSI = b.aI;
        break;
      case OP_PV_MID: {  // positionverlet.h:73-94
        // storeFieldsAveragesHalf (positionverlet.h:258-266): the half-step field averages feed getSI
        const double g2SI = red[RED_GRAD2S], cg2SI = red[RED_GRAD2CS], b2SI = red[RED_B2U1];
        b.potAvSI = red[RED_POT];
        // storeMomentaAverages (positionverlet.h:238-256)
        const double pIM = b.pi2AvI, cIM = b.CSpi2AvI, uIM = b.U1pi2AvI;
        if (mc.ns > 0) b.pi2AvI = red[RED_PI2S];
        if (mc.ncs > 0) b.CSpi2AvI = red[RED_PI2CS];
        if (mc.nu1 > 0) b.U1pi2AvI = red[RED_PI2U1];
        b.pi2AvSI = 0.5 * (b.pi2AvI + pIM);
        b.CSpi2AvSI = 0.5 * (b.CSpi2AvI + cIM);
        b.U1pi2AvSI = 0.5 * (b.U1pi2AvI + uIM);
        // kickScaleFactor with ScaleFactorKernels::getSI (positionverlet.h:151-154)
        st->K = sf_kernel_of(b.aI, b.pi2AvSI, g2SI, b.CSpi2AvSI, cg2SI, b.U1pi2AvSI, b2SI, b.potAvSI, mc);
        b.aDotI += dt * st->K * w;
        // driftScaleFactorSecondHalf (positionverlet.h:212-223); the second drift uses the new aI
        b.aI += dt / 2.0 * b.aDotI * w;
        b.aSI = (b.aIM + b.aI) / 2.0;
        a_drift = b.aI;
        break;
      }
      case OP_PV_DRIFT_AI:
        a_drift = b.aI;
        break;
      default:
        break;
    }
    if (o == 0 && ops.snap) *ops.snap = b;
  }
  // a^(1+alpha) and a^(3+alpha) = a^(1+alpha) a^2, a^(alpha-1); the drift factors likewise share one power
  const double aK = rpow(b.aI, 1 + mc.alpha);
  st->aK = aK;
  st->aV = aK * (b.aI * b.aI);
  st->aG = rpow(b.aI, mc.alpha - 1);
  st->aJ = aK;
  const double ad = a_drift > 0.0 ? a_drift : b.aSI;
  const double fu = rpow(ad, mc.alpha - 1);
  st->fdrift_u = fu;
  st->fdrift_s = fu / (ad * ad);
}

}  // namespace clmi
