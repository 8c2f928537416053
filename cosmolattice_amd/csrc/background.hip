// Background (homogeneous) update on the device, Gauss-law and mean reductions, array conversions.
//
// k_update finishes the per-block partial sums of the last lattice pass in a fixed order (so
// results are bit-reproducible run to run), then one thread advances the scale factor exactly as
// the reference's host code does, and refreshes the kick/drift coefficients the next pass reads.
#include "clmi_internal.hpp"

namespace clmi {
namespace {

// ScaleFactorKernels::get (evolvers/kernels/scalefactorkernels.h:27-81) with the energy
// normalisations of definitions/energies.h:37-85.
__device__ double sf_kernel(const clmi_background &b, const ModelConsts &mc)
{
  const double a = b.aI;
  double Eks = 0, Egs = 0, Ekcs = 0, Egcs = 0, Eel = 0, Emag = 0;
  if (mc.ns > 0) {
    Eks = 0.5 * b.pi2AvI * pow(a, -6.0);
    Egs = 0.5 * b.grad2AvI * pow(a, -2.0);
  }
  if (mc.ncs > 0) {
    Ekcs = b.CSpi2AvI * pow(a, -6.0);
    Egcs = b.CSgrad2AvI * pow(a, -2.0);
  }
  if (mc.nu1 > 0) {
    const double w2 = (mc.omegaStar / mc.fStar) * (mc.omegaStar / mc.fStar);
    Eel = 0.5 * w2 * pow(a, -4.0) * b.U1pi2AvI;
    Emag = 0.5 * w2 * pow(a, -4.0) * b.U1Mag2AvI;
  }
  const double fm = mc.fStar / mc.MPl;
  return pow(a, 2 * mc.alpha + 1) / 3.0 * (fm * fm) *
         ((mc.alpha - 2) * (Eks + Ekcs) + mc.alpha * (Egs + Egcs) + (mc.alpha - 1) * (Eel + Emag) +
          (mc.alpha + 1) * b.potAvI);
}

constexpr int kUpdThreads = 1024;

// Fixed-order sum of `rows` partial rows: thread t takes rows t, t + 1024, ... (four loads in
// flight per slot), then wavefront shuffles, then a sum over the 16 wavefronts in order.
__device__ void sum_partials(const double *partials, int rows, double *out /* kNumRed, in LDS */)
{
  __shared__ double sh[kUpdThreads / 64][kNumRed];
  const int t = threadIdx.x;
  double v[kNumRed];
#pragma unroll
  for (int k = 0; k < kNumRed; ++k) v[k] = 0.0;
  int r = t;
  for (; r + 3 * kUpdThreads < rows; r += 4 * kUpdThreads) {
    double q[4][kNumRed];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < kNumRed; ++k) q[u][k] = partials[(size_t)(r + u * kUpdThreads) * kNumRed + k];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < kNumRed; ++k) v[k] += q[u][k];
  }
  for (; r < rows; r += kUpdThreads)
#pragma unroll
    for (int k = 0; k < kNumRed; ++k) v[k] += partials[(size_t)r * kNumRed + k];
#pragma unroll
  for (int k = 0; k < kNumRed; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  if ((t & 63) == 0)
#pragma unroll
    for (int k = 0; k < kNumRed; ++k) sh[t >> 6][k] = v[k];
  __syncthreads();
  if (t < kNumRed) {
    double s = 0.0;
    for (int w = 0; w < kUpdThreads / 64; ++w) s += sh[w][t];
    out[t] = s;
  }
  __syncthreads();
}

__global__ __launch_bounds__(kUpdThreads) void k_update(DevState *st, const double *partials, int rows,
                                                        ModelConsts mc, UpdateOps ops)
{
  __shared__ double tot[kNumRed];
  const int t = threadIdx.x;
  if (rows > 0) {
    sum_partials(partials, rows, tot);
    if (t < kNumRed) st->red[t] = tot[t] * mc.invV;
  }
  if (t != 0) return;

  clmi_background &b = st->bg;
  double red[kNumRed];
  for (int k = 0; k < kNumRed; ++k) red[k] = rows > 0 ? tot[k] * mc.invV : st->red[k];
  const double dt = mc.dt;
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
      default:
        break;
    }
  }
  st->aK = pow(b.aI, 1 + mc.alpha);
  st->aV = pow(b.aI, 3 + mc.alpha);
  st->aG = pow(b.aI, mc.alpha - 1);
  st->aJ = pow(b.aI, 1 + mc.alpha);
  st->fdrift_s = pow(b.aSI, mc.alpha - 3);
  st->fdrift_u = pow(b.aSI, mc.alpha - 1);
}

// GaussLaws::checkU1 (definitions/gausslaws.h:31-53) with the charge density of
// definitions/mattercurrents.h:27-48.  Slots: |LHS-RHS|, |LHS1^2+LHS2^2+LHS3^2+RHS^2|, |RHS|, |LHS|.
template <bool F32>
__global__ __launch_bounds__(kBlock) void k_gauss(const GaussArgs a)
{
  const uint32_t N = a.N, NN = N * N, V = a.V;
  double acc[4] = {0, 0, 0, 0};
  auto ld = [&](const void *p, const float *lo, uint32_t i) -> double {
    if (F32) return (double)((const float *)p)[i] + (lo ? (double)lo[i] : 0.0);
    return ((const double *)p)[i];
  };
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < V; s += gridDim.x * kBlock) {
    const uint32_t x = s / NN, y = (s - x * NN) / N, z = s - x * NN - y * N;
    const uint32_t m[3] = {s - x * NN + (x == 0 ? N - 1 : x - 1) * NN, s - y * N + (y == 0 ? N - 1 : y - 1) * N,
                           s - z + (z == 0 ? N - 1 : z - 1)};
    double l[3], lsum = 0.0;
    for (int i = 0; i < 3; ++i) {
      const double pc = ld(a.piA[i], nullptr, s), pm = ld(a.piA[i], nullptr, m[i]);
      l[i] = (pc - pm) * a.inv_dx;
      lsum += pc - pm;
    }
    const double lhs = lsum * a.inv_dx;
    double rho = 0.0;
    if (a.has_cs) {
      const double pr = ld(a.pi[0], nullptr, s), pi_ = ld(a.pi[1], nullptr, s);
      const double fr = ld(a.phi[0], a.philo[0], s), fi = ld(a.phi[1], a.philo[1], s);
      rho = a.qnorm_gQ * (pr * fi - pi_ * fr);
    }
    const double rhs = -a.dx * rho;
    const double sum = l[0] * l[0] + l[1] * l[1] + l[2] * l[2] + rhs * rhs;
    acc[0] += fabs(lhs - rhs);
    acc[1] += fabs(sum);
    acc[2] += fabs(rhs);
    acc[3] += fabs(lhs);
  }
  double v[kNumRed] = {acc[0], acc[1], acc[2], acc[3], 0, 0, 0, 0};
  block_reduce_store<kNumRed>(v, a.partials + (size_t)blockIdx.x * kNumRed);
}

template <bool F32>
__global__ __launch_bounds__(kBlock) void k_means(const MeansArgs a)
{
  double acc[2] = {0, 0};
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < a.V; s += gridDim.x * kBlock) {
    double v;
    if (F32)
      v = (double)((const float *)a.p)[s] + (a.lo ? (double)a.lo[s] : 0.0);
    else
      v = ((const double *)a.p)[s];
    v *= a.scale;
    acc[0] += v;
    acc[1] += v * v;
  }
  double v[kNumRed] = {acc[0], acc[1], 0, 0, 0, 0, 0, 0};
  block_reduce_store<kNumRed>(v, a.partials + (size_t)blockIdx.x * kNumRed);
}

__global__ __launch_bounds__(kUpdThreads) void k_sum_rows(const double *partials, int rows, double *out)
{
  __shared__ double tot[kNumRed];
  sum_partials(partials, rows, tot);
  if (threadIdx.x < kNumRed) out[threadIdx.x] = tot[threadIdx.x];
}

__global__ void k_split(const double *src, float *hi, float *lo, uint32_t V)
{
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) {
    const double v = src[i];
    const float h = (float)v;
    hi[i] = h;
    lo[i] = (float)(v - (double)h);
  }
}
__global__ void k_to_f32(const double *src, float *dst, uint32_t V)
{
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x) dst[i] = (float)src[i];
}
__global__ void k_join(const float *hi, const float *lo, double *dst, uint32_t V)
{
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < V; i += gridDim.x * blockDim.x)
    dst[i] = (double)hi[i] + (lo ? (double)lo[i] : 0.0);
}

// Plain streaming copy (16-byte vectors), used only to calibrate hardware byte counters.
__global__ __launch_bounds__(256) void k_copy(const float4 *src, float4 *dst, size_t n)
{
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

int conv_blocks(uint32_t V)
{
  uint32_t b = (V + 255) / 256;
  return (int)(b < 8192u ? b : 8192u);
}

}  // namespace

hipError_t launch_update(DevState *st, const double *partials, int rows, const ModelConsts &mc,
                         const UpdateOps &ops, hipStream_t s)
{
  hipLaunchKernelGGL(k_update, dim3(1), dim3(kUpdThreads), 0, s, st, partials, rows, mc, ops);
  return hipGetLastError();
}

hipError_t launch_gauss(const GaussArgs &a, int blocks, hipStream_t s)
{
  if (a.f32)
    hipLaunchKernelGGL(k_gauss<true>, dim3(blocks), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_gauss<false>, dim3(blocks), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_means(const MeansArgs &a, int blocks, hipStream_t s)
{
  if (a.f32)
    hipLaunchKernelGGL(k_means<true>, dim3(blocks), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_means<false>, dim3(blocks), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_sum_rows(const double *partials, int rows, double *out, hipStream_t s)
{
  hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(kUpdThreads), 0, s, partials, rows, out);
  return hipGetLastError();
}

hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t s)
{
  const size_t n = bytes / sizeof(float4);
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, s, (const float4 *)src, (float4 *)dst, n);
  return hipGetLastError();
}

hipError_t launch_split_f64(const double *src, float *hi, float *lo, uint32_t V, hipStream_t s)
{
  hipLaunchKernelGGL(k_split, dim3(conv_blocks(V)), dim3(256), 0, s, src, hi, lo, V);
  return hipGetLastError();
}
hipError_t launch_to_f32(const double *src, float *dst, uint32_t V, hipStream_t s)
{
  hipLaunchKernelGGL(k_to_f32, dim3(conv_blocks(V)), dim3(256), 0, s, src, dst, V);
  return hipGetLastError();
}
hipError_t launch_join_f64(const float *hi, const float *lo, double *dst, uint32_t V, hipStream_t s)
{
  hipLaunchKernelGGL(k_join, dim3(conv_blocks(V)), dim3(256), 0, s, hi, lo, dst, V);
  return hipGetLastError();
}

}  // namespace clmi
