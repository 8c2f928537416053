// Charged scalar + U(1) gauge lattice pass, one site per thread: [deferred drift] + [kicks] +
// [volume reductions].  The per-site arithmetic is gauge_site.hpp's; this kernel gathers every
// neighbour from memory (drifting it on the fly) and applies to any N.  Links are recomputed from A
// (nothing but A is stored).  Drifts: fldCS += aSI^(alpha-3) dt piCS, fldU1 += aSI^(alpha-1) dt piU1
// (evolvers/leapfrog.h:203-217).
#include "gauge_site.hpp"

namespace clmi {
namespace {

struct Coord {
  uint32_t x, y, z, N, NN;
  __device__ __forceinline__ uint32_t at(int dx, int dy, int dz) const
  {
    uint32_t xx = (uint32_t)((int)x + dx + (int)N) % N;
    uint32_t yy = (uint32_t)((int)y + dy + (int)N) % N;
    uint32_t zz = (uint32_t)((int)z + dz + (int)N) % N;
    return xx * NN + yy * N + zz;
  }
  // site shifted by +1 along direction i and -1 along direction j (i, j in 0..2; either may be -1)
  __device__ __forceinline__ uint32_t shifted(int i, int si, int j, int sj) const
  {
    int d[3] = {0, 0, 0};
    if (i >= 0) d[i] += si;
    if (j >= 0) d[j] += sj;
    return at(d[0], d[1], d[2]);
  }
};

template <bool F32, bool DRIFT, bool KICK, bool RED>
__global__ __launch_bounds__(kBlock) void k_gauge_pass(const GaugePassArgs a)
{
  using G = GVal<F32>;
  using V = typename G::V;
  using P = typename G::P;
  const uint32_t N = a.N, NN = N * N, Vn = a.V;
  const GaugeK k = gauge_consts(a, KICK);
  // drift factors: FP64 (fdrift * cdt), FP32-bulk (float)(fdrift * cdt)
  const double cs64 = DRIFT ? a.st->fdrift_s * a.cdt : 0.0, cu64 = DRIFT ? a.st->fdrift_u * a.cdt : 0.0;
  const float cs32 = DRIFT ? (float)(a.st->fdrift_s * a.cdt) : 0.0f;
  const float cu32 = DRIFT ? (float)(a.st->fdrift_u * a.cdt) : 0.0f;
  double acc[5] = {0, 0, 0, 0, 0};  // pi2CS, grad2CS, pot, pi2U1, B2U1

  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < Vn; s += gridDim.x * kBlock) {
    Coord c;
    c.N = N;
    c.NN = NN;
    c.x = s / NN;
    c.y = (s - c.x * NN) / N;
    c.z = s - c.x * NN - c.y * N;
    auto phi = [&](int comp, uint32_t i) -> V {
      if constexpr (F32)
        return G::fld(a.phi_in[comp], a.philo_in[comp], a.pi_in[comp], i, DRIFT, cs32);
      else
        return G::fld(a.phi_in[comp], nullptr, a.pi_in[comp], i, DRIFT, cs64);
    };
    auto A = [&](int d, uint32_t i) -> V {
      if (!k.u1) return V{};
      if constexpr (F32)
        return G::fld(a.A_in[d], a.Alo_in[d], a.piA_in[d], i, DRIFT, cu32);
      else
        return G::fld(a.A_in[d], nullptr, a.piA_in[d], i, DRIFT, cu64);
    };
    GaugeIn<V> g;
    for (int q = 0; q < 2; ++q) g.fc[q] = phi(q, s);
    for (int d = 0; d < 3; ++d) {
      const uint32_t jp = c.shifted(d, 1, -1, 0), jm = c.shifted(d, -1, -1, 0);
      for (int q = 0; q < 2; ++q) {
        g.fp[d][q] = phi(q, jp);
        g.fm[d][q] = KICK ? phi(q, jm) : V{};
      }
    }
    for (int j = 0; j < 3; ++j) g.Ac[j] = A(j, s);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        g.Ap[i][j] = A(j, c.shifted(i, 1, -1, 0));
        g.Am[i][j] = (KICK && (i == j || k.u1)) ? A(j, c.shifted(i, -1, -1, 0)) : V{};
        g.Ad[i][j] = (KICK && i != j) ? A(j, c.shifted(i, 1, j, -1)) : V{};
      }
    if (DRIFT) {
      for (int q = 0; q < 2; ++q) G::store(a.phi_out[q], a.philo_out[q], s, g.fc[q]);
      if (k.u1)
        for (int d = 0; d < 3; ++d) G::store(a.A_out[d], a.Alo_out[d], s, g.Ac[d]);
    }
    P pin[2], pAin[3] = {0, 0, 0}, pout[2], pAout[3];
    for (int q = 0; q < 2; ++q) pin[q] = ((const P *)a.pi_in[q])[s];
    if (k.u1)
      for (int d = 0; d < 3; ++d) pAin[d] = ((const P *)a.piA_in[d])[s];
    if constexpr (F32)
      gauge_site_f32<KICK, RED>(k, a.poly, g, pin, pAin, pout, pAout, acc);
    else
      gauge_site_f64<KICK, RED>(k, a.poly, g, pin, pAin, pout, pAout, acc);
    if (KICK) {
      for (int q = 0; q < 2; ++q) ((P *)a.pi_out[q])[s] = pout[q];
      if (k.u1)
        for (int d = 0; d < 3; ++d) ((P *)a.piA_out[d])[s] = pAout[d];
    }
  }
  if (RED) {
    double v[kNumRed] = {0, 0, acc[2], acc[0], acc[1], acc[3], acc[4], 0};
    block_reduce_store<kNumRed>(v, a.partials + (size_t)blockIdx.x * kNumRed);
  }
}

template <bool F32>
hipError_t dispatch(bool drift, bool kick, bool red, const GaugePassArgs &a, int blocks, hipStream_t s)
{
  const int code = (drift ? 4 : 0) | (kick ? 2 : 0) | (red ? 1 : 0);
#define CLMI_GP(D, K, R)                                                                                        \
  case ((D ? 4 : 0) | (K ? 2 : 0) | (R ? 1 : 0)):                                                              \
    hipLaunchKernelGGL((k_gauge_pass<F32, D, K, R>), dim3(blocks), dim3(kBlock), 0, s, a);                      \
    break;
  switch (code) {
    CLMI_GP(true, true, true)
    CLMI_GP(true, true, false)
    CLMI_GP(true, false, true)
    CLMI_GP(true, false, false)
    CLMI_GP(false, true, true)
    CLMI_GP(false, true, false)
    CLMI_GP(false, false, true)
    default:
      return hipErrorInvalidValue;
  }
#undef CLMI_GP
  return hipGetLastError();
}

}  // namespace

int gauge_pass_blocks(uint32_t V)
{
  uint32_t b = (V + kBlock - 1) / kBlock;
  return (int)(b < (uint32_t)kMaxPassBlocks ? b : (uint32_t)kMaxPassBlocks);
}

hipError_t launch_gauge_pass(bool f32, bool drift, bool kick, bool red, const GaugePassArgs &a, int blocks,
                             hipStream_t s)
{
  return f32 ? dispatch<true>(drift, kick, red, a, blocks, s) : dispatch<false>(drift, kick, red, a, blocks, s);
}

}  // namespace clmi
