// Background (homogeneous) update on the device, Gauss-law and mean reductions, array conversions.
//
// k_update finishes the per-block partial sums of the last lattice pass in a fixed order (so
// results are bit-reproducible run to run), then one thread advances the scale factor exactly as
// the reference's host code does, and refreshes the kick/drift coefficients the next pass reads.
#include <algorithm>

#include "clmi_internal.hpp"
#include "update_device.hpp"

namespace clmi {
namespace {

constexpr int kUpdThreads = 512;  // 8 waves: two per SIMD, so the updating thread keeps its state in registers

__global__ __launch_bounds__(kUpdThreads) void k_update(DevState *st, const double *partials, int rows,
                                                        ModelConsts mc, UpdateOps ops)
{
  __shared__ double tot[kNumRed];
  const int t = threadIdx.x;
  // Thread 0 reads the state before the partial sums, so that its loads overlap theirs, and writes it
  // back once at the end.
  DevState s;
  if (t == 0) s = *st;
  if (rows > 0) sum_partials<kUpdThreads>(partials, rows, tot);
  if (t != 0) return;
  double red[kNumRed];
  for (int k = 0; k < kNumRed; ++k) red[k] = rows > 0 ? tot[k] * mc.invV : s.red[k];
  for (int k = 0; k < kNumRed; ++k) s.red[k] = red[k];
  apply_update(&s, red, mc, ops);
  *st = s;
}

// The same update for several lattices, one workgroup each.
__global__ __launch_bounds__(kUpdThreads) void k_update_batch(const UpdMember *m, UpdateOps ops)
{
  const UpdMember &u = m[blockIdx.x];
  __shared__ double tot[kNumRed];
  const int t = threadIdx.x;
  // the updating thread copies the member's state and constants into registers at entry, so that their
  // loads overlap the partial-row loads instead of sitting in the scalar chain
  DevState s;
  ModelConsts mc;
  if (t == 0) {
    s = *u.st;
    mc = u.mc;
  }
  if (u.rows > 0) sum_partials<kUpdThreads>(u.partials, u.rows, tot);
  if (t != 0) return;
  double red[kNumRed];
  for (int k = 0; k < kNumRed; ++k) red[k] = u.rows > 0 ? tot[k] * mc.invV : s.red[k];
  for (int k = 0; k < kNumRed; ++k) s.red[k] = red[k];
  apply_update(&s, red, mc, ops);
  *u.st = s;
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
  double acc[3] = {0, 0, 0};
  auto ld = [&](uint32_t i) -> double {
    if (F32) return (double)((const float *)a.p)[i] + (a.lo ? (double)a.lo[i] : 0.0);
    return ((const double *)a.p)[i];
  };
  const uint32_t N = a.N, NN = N * N;
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < a.V; s += gridDim.x * kBlock) {
    const double c = ld(s);
    const double v = c * a.scale;
    acc[0] += v;
    acc[1] += v * v;
    if (a.grad) {
      const uint32_t x = s / NN, y = (s - x * NN) / N, z = s - x * NN - y * N;
      const double dx = ld(s - x * NN + (x + 1 == N ? 0 : x + 1) * NN) - c;
      const double dy = ld(s - y * N + (y + 1 == N ? 0 : y + 1) * N) - c;
      const double dz = ld(s - z + (z + 1 == N ? 0 : z + 1)) - c;
      acc[2] += dx * dx + dy * dy + dz * dz;
    }
  }
  double v[kNumRed] = {acc[0], acc[1], acc[2], 0, 0, 0, 0, 0};
  block_reduce_store<kNumRed>(v, a.partials + (size_t)blockIdx.x * kNumRed);
}

// Component i of B = curl A at site s with forward differences: (d+_j A_k - d+_k A_j) / dx, (i, j, k)
// cyclic.  ld(c, site) loads component c; nb[d] is the site one step forward in direction d.
template <class LD>
__device__ __forceinline__ double curl_component(const LD &ld, int i, uint32_t s, const uint32_t nb[3], double inv_dx)
{
  const int j = (i + 1) % 3, k = (i + 2) % 3;
  return ((ld(k, nb[j]) - ld(k, s)) - (ld(j, nb[k]) - ld(j, s))) * inv_dx;
}

template <bool F32>
__global__ __launch_bounds__(kBlock) void k_norms(const NormArgs a)
{
  double acc[2] = {0, 0};
  auto ld = [&](int c, uint32_t i) -> double {
    if (F32) return (double)((const float *)a.p[c])[i] + (a.lo[c] ? (double)a.lo[c][i] : 0.0);
    return ((const double *)a.p[c])[i];
  };
  const uint32_t N = a.N, NN = N * N;
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < a.V; s += gridDim.x * kBlock) {
    double q = 0.0;
    if (!a.curl) {
      for (int c = 0; c < a.ncomp; ++c) {
        const double v = ld(c, s);
        q += v * v;
      }
    } else {
      const uint32_t x = s / NN, y = (s - x * NN) / N, z = s - x * NN - y * N;
      const uint32_t nb[3] = {s - x * NN + (x + 1 == N ? 0 : x + 1) * NN, s - y * N + (y + 1 == N ? 0 : y + 1) * N,
                              s - z + (z + 1 == N ? 0 : z + 1)};
      for (int i = 0; i < 3; ++i) {
        const double b = curl_component(ld, i, s, nb, a.inv_dx);
        q += b * b;
      }
    }
    acc[0] += sqrt(q);
    acc[1] += q;
  }
  double v[kNumRed] = {acc[0], acc[1], 0, 0, 0, 0, 0, 0};
  block_reduce_store<kNumRed>(v, a.partials + (size_t)blockIdx.x * kNumRed);
}

template <bool F32>
__global__ __launch_bounds__(kBlock) void k_curl_component(const NormArgs a, int comp, double *dst)
{
  auto ld = [&](int c, uint32_t i) -> double {
    if (F32) return (double)((const float *)a.p[c])[i] + (a.lo[c] ? (double)a.lo[c][i] : 0.0);
    return ((const double *)a.p[c])[i];
  };
  const uint32_t N = a.N, NN = N * N;
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < a.V; s += gridDim.x * kBlock) {
    const uint32_t x = s / NN, y = (s - x * NN) / N, z = s - x * NN - y * N;
    const uint32_t nb[3] = {s - x * NN + (x + 1 == N ? 0 : x + 1) * NN, s - y * N + (y + 1 == N ? 0 : y + 1) * N,
                            s - z + (z + 1 == N ? 0 : z + 1)};
    dst[s] = curl_component(ld, comp, s, nb, a.inv_dx);
  }
}

template <bool F32>
__global__ __launch_bounds__(kBlock) void k_monomials(const MonoArgs a)
{
  double acc[kNumRed];
#pragma unroll
  for (int k = 0; k < kNumRed; ++k) acc[k] = 0.0;
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < a.V; s += gridDim.x * kBlock) {
    double phi[CLMI_MAX_SCALARS];
    for (int f = 0; f < a.nf; ++f)
      phi[f] = F32 ? (double)((const float *)a.phi[f])[s] + (double)a.philo[f][s] : ((const double *)a.phi[f])[s];
    for (int t = 0; t < a.poly.nterms && t < kNumRed; ++t) {
      double m = 1.0;
      for (int f = 0; f < a.nf; ++f)
        for (int k = 0; k < a.poly.e[t][f]; ++k) m *= phi[f];
      acc[t] += m;
    }
  }
  block_reduce_store<kNumRed>(acc, a.partials + (size_t)blockIdx.x * kNumRed);
}

__global__ __launch_bounds__(kUpdThreads) void k_sum_rows(const double *partials, int rows, double *out)
{
  __shared__ double tot[kNumRed];
  sum_partials<kUpdThreads>(partials, rows, tot);
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

// Several arrays copied by one launch: each workgroup strides over every array in turn.
__global__ __launch_bounds__(256) void k_copy_set(const CopySet cs)
{
  const size_t stride = (size_t)gridDim.x * 256;
  typedef float v4 __attribute__((ext_vector_type(4)));
  for (int a = 0; a < cs.count; ++a) {
    const v4 *src = (const v4 *)cs.src[a];
    v4 *dst = (v4 *)cs.dst[a];
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < cs.n16[a]; i += stride)
      __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
  }
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

hipError_t launch_monomials(const MonoArgs &a, int blocks, hipStream_t s)
{
  if (a.f32)
    hipLaunchKernelGGL(k_monomials<true>, dim3(blocks), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_monomials<false>, dim3(blocks), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_norms(const NormArgs &a, int blocks, hipStream_t s)
{
  if (a.f32)
    hipLaunchKernelGGL(k_norms<true>, dim3(blocks), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL(k_norms<false>, dim3(blocks), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_curl_component(const NormArgs &a, int comp, double *dst, hipStream_t s)
{
  const uint32_t blocks = std::min<uint32_t>((a.V + kBlock - 1) / kBlock, 16384u);
  if (a.f32)
    hipLaunchKernelGGL(k_curl_component<true>, dim3(blocks), dim3(kBlock), 0, s, a, comp, dst);
  else
    hipLaunchKernelGGL(k_curl_component<false>, dim3(blocks), dim3(kBlock), 0, s, a, comp, dst);
  return hipGetLastError();
}

hipError_t launch_update_batch(const UpdMember *d_members, int n, const UpdateOps &ops, hipStream_t s)
{
  hipLaunchKernelGGL(k_update_batch, dim3(n), dim3(kUpdThreads), 0, s, d_members, ops);
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

hipError_t launch_copy_set(const CopySet &cs, hipStream_t s)
{
  if (cs.count < 0 || cs.count > kCopySetMax) return hipErrorInvalidValue;
  if (cs.count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_set, dim3(2048), dim3(256), 0, s, cs);
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
