// Charged scalar + U(1) gauge lattice pass: [deferred drift] + [kicks] + [volume reductions].
//
// Reference expressions fused here (one TempLat sweep each in the reference):
//   drifts   fldCS += aSI^(alpha-3) dt piCS ; fldU1 += aSI^(alpha-1) dt piU1   evolvers/leapfrog.h:203-217
//   CS kick  piCS += (w dt)(aI^(1+alpha) D^2 phi - aI^(3+alpha)/2 dV/dphi)     kernels/complexscalarkernels.h:28-35
//            D^2 phi = dx^-2 [sum_i U_i(x) phi(x+i) + conj(U_i(x-i)) phi(x-i) - 6 phi]
//                                                                         definitions/gaugederivatives.h:48-53
//            U_i(x) = exp(-i dx gQ A_i(x))                                definitions/gaugederivatives.h:92-102
//   U1 kick  piU1_i += (w dt)(aI^(alpha-1)(Lapl A_i - sum_j d-_j d+_i A_j) - aI^(1+alpha) J_i)
//            J_i = 2/dx (fStar/omegaStar)^2 gQ Im[conj(U_i(x) phi(x+i)) phi(x)]
//                                                  kernels/u1kernels.h:31-52, definitions/mattercurrents.h:52-76
//   <|piCS|^2>, <sum_j |D+_j phi|^2>, <V>, <sum_i piU1_i^2>, <F21^2+F31^2+F32^2>
//                                                  definitions/fieldfunctionals.h:80-153, averages.h:39-71
// Links are recomputed from A on the fly (6 sincos per site); nothing but A is stored.
#include "clmi_internal.hpp"

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

__device__ __forceinline__ double cs_dnorm_over_norm_d(const CsPoly &P, double r2)
{
  // (dV/d|phi|)/|phi| = sum_k c_k (2 h_k) |phi|^(2 h_k - 2)
  double d = 0.0;
  for (int k = 0; k < P.nterms; ++k) {
    const int h = P.h[k];
    if (h == 0) continue;
    double p = P.c[k] * (2 * h);
    for (int i = 0; i < h - 1; ++i) p *= r2;
    d += p;
  }
  return d;
}
__device__ __forceinline__ double cs_pot_d(const CsPoly &P, double r2)
{
  double v = 0.0;
  for (int k = 0; k < P.nterms; ++k) {
    double p = P.c[k];
    for (int i = 0; i < P.h[k]; ++i) p *= r2;
    v += p;
  }
  return v;
}
__device__ __forceinline__ float cs_dnorm_over_norm_f(const CsPoly &P, float r2)
{
  float d = 0.0f;
  for (int k = 0; k < P.nterms; ++k) {
    const int h = P.h[k];
    if (h == 0) continue;
    float p = (float)P.c[k] * (float)(2 * h);
    for (int i = 0; i < h - 1; ++i) p *= r2;
    d += p;
  }
  return d;
}

__device__ __forceinline__ void pair_add(float &hi, float &lo, float d)
{
  const float s = lo + d;
  const float h = hi + s;
  const float bb = h - hi;
  const float e = (hi - (h - bb)) + (s - bb);
  hi = h;
  lo = e;
}

// ---------------------------------------------------------------------------------------------
// FP64
// ---------------------------------------------------------------------------------------------
template <bool DRIFT, bool KICK, bool RED>
__global__ __launch_bounds__(kBlock) void k_gauge_pass_f64(const GaugePassArgs a)
{
  const uint32_t N = a.N, NN = N * N, V = a.V;
  const double fs = DRIFT ? a.st->fdrift_s : 0.0;
  const double fu = DRIFT ? a.st->fdrift_u : 0.0;
  const double aK = KICK ? a.st->aK : 0.0, aV = KICK ? a.st->aV : 0.0;
  const double aG = KICK ? a.st->aG : 0.0, aJ = KICK ? a.st->aJ : 0.0;
  const bool u1 = a.has_u1 != 0;
  double acc[5] = {0, 0, 0, 0, 0};  // pi2CS, grad2CS, pot, pi2U1, B2U1

  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < V; s += gridDim.x * kBlock) {
    Coord c;
    c.N = N;
    c.NN = NN;
    c.x = s / NN;
    c.y = (s - c.x * NN) / N;
    c.z = s - c.x * NN - c.y * N;

    auto phi = [&](int comp, uint32_t i) -> double {
      double v = ((const double *)a.phi_in[comp])[i];
      if (DRIFT) v += fs * a.cdt * ((const double *)a.pi_in[comp])[i];
      return v;
    };
    auto A = [&](int d, uint32_t i) -> double {
      if (!u1) return 0.0;
      double v = ((const double *)a.A_in[d])[i];
      if (DRIFT) v += fu * a.cdt * ((const double *)a.piA_in[d])[i];
      return v;
    };
    auto link = [&](int d, uint32_t i, double &cs, double &sn) {
      if (!u1) {
        cs = 1.0;
        sn = 0.0;
        return;
      }
      sincos(-a.dx * a.gQ * A(d, i), &sn, &cs);
    };

    const double pr = phi(0, s), pim = phi(1, s);
    if (DRIFT) {
      ((double *)a.phi_out[0])[s] = pr;
      ((double *)a.phi_out[1])[s] = pim;
      if (u1)
        for (int d = 0; d < 3; ++d) ((double *)a.A_out[d])[s] = A(d, s);
    }

    // ---- complex scalar: covariant Laplacian, forward covariant gradient -----------------
    double plus_r = 0, plus_i = 0, minus_r = 0, minus_i = 0, g2 = 0;
    double cl[3], sl[3];  // links U_d(x)
    for (int d = 0; d < 3; ++d) {
      const uint32_t jp = c.shifted(d, 1, -1, 0), jm = c.shifted(d, -1, -1, 0);
      link(d, s, cl[d], sl[d]);
      const double fpr = phi(0, jp), fpi = phi(1, jp);
      const double ur = cl[d] * fpr - sl[d] * fpi;
      const double ui = cl[d] * fpi + sl[d] * fpr;
      plus_r += ur;
      plus_i += ui;
      if (KICK) {
        double c2, s2;
        link(d, jm, c2, s2);
        const double fmr = phi(0, jm), fmi = phi(1, jm);
        minus_r += c2 * fmr + s2 * fmi;
        minus_i += c2 * fmi - s2 * fmr;
      }
      if (RED) {
        const double gr = (ur - pr) * a.inv_dx, gi = (ui - pim) * a.inv_dx;
        g2 += gr * gr + gi * gi;
      }
    }
    const double r2 = pr * pr + pim * pim;
    double p_r = ((const double *)a.pi_in[0])[s], p_i = ((const double *)a.pi_in[1])[s];
    if (KICK) {
      const double lap_r = (plus_r + minus_r - 6 * pr) * a.inv_dx2;
      const double lap_i = (plus_i + minus_i - 6 * pim) * a.inv_dx2;
      const double nd = cs_dnorm_over_norm_d(a.poly, r2);
      p_r += a.ckick * (aK * lap_r - aV / 2 * (nd * pr));
      p_i += a.ckick * (aK * lap_i - aV / 2 * (nd * pim));
      ((double *)a.pi_out[0])[s] = p_r;
      ((double *)a.pi_out[1])[s] = p_i;
    }
    if (RED) {
      acc[0] += p_r * p_r + p_i * p_i;
      acc[1] += g2;
      acc[2] += cs_pot_d(a.poly, r2);
    }

    // ---- U(1) gauge field ------------------------------------------------------------------
    if (u1) {
      double Ac[3], Ap[3][3];  // Ap[i][j] = A_j(x + i)
      for (int j = 0; j < 3; ++j) Ac[j] = A(j, s);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Ap[i][j] = A(j, c.shifted(i, 1, -1, 0));
      double e2 = 0.0;
      for (int i = 0; i < 3; ++i) {
        double pa = ((const double *)a.piA_in[i])[s];
        if (KICK) {
          double lap = 0.0;
          for (int d = 0; d < 3; ++d) lap += Ap[d][i] + A(i, c.shifted(d, -1, -1, 0)) - 2 * Ac[i];
          lap *= a.inv_dx2;
          double gd = 0.0;
          for (int j = 0; j < 3; ++j) {
            const double a_pimj = A(j, c.shifted(i, 1, j, -1));
            const double a_mj = A(j, c.shifted(j, -1, -1, 0));
            gd += (Ap[i][j] - Ac[j]) - (a_pimj - a_mj);
          }
          gd *= a.inv_dx2;
          // J_i with U_i(x) phi(x+i)
          const uint32_t sp = c.shifted(i, 1, -1, 0);
          const double fpr = phi(0, sp), fpi = phi(1, sp);
          const double ur = cl[i] * fpr - sl[i] * fpi;
          const double ui = cl[i] * fpi + sl[i] * fpr;
          const double J = a.jnorm * a.gQ * (ur * pim - ui * pr);
          pa += a.ckick * (aG * (lap - gd) - aJ * J);
          ((double *)a.piA_out[i])[s] = pa;
        }
        e2 += pa * pa;
      }
      if (RED) {
        double b2 = 0.0;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < i; ++j) {
            const double f = ((Ap[i][j] - Ac[j]) - (Ap[j][i] - Ac[i])) * a.inv_dx;
            b2 += f * f;
          }
        acc[3] += e2;
        acc[4] += b2;
      }
    }
  }
  if (RED) {
    double v[kNumRed] = {0, 0, acc[2], acc[0], acc[1], acc[3], acc[4], 0};
    block_reduce_store<kNumRed>(v, a.partials + (size_t)blockIdx.x * kNumRed);
  }
}

// ---------------------------------------------------------------------------------------------
// FP32 bulk (fields as hi/lo float pairs, momenta float, reductions FP64)
// ---------------------------------------------------------------------------------------------
struct P2 {
  float h, l;
};

template <bool DRIFT, bool KICK, bool RED>
__global__ __launch_bounds__(kBlock) void k_gauge_pass_f32(const GaugePassArgs a)
{
  const uint32_t N = a.N, NN = N * N, V = a.V;
  const float cfs = DRIFT ? (float)(a.st->fdrift_s * a.cdt) : 0.0f;
  const float cfu = DRIFT ? (float)(a.st->fdrift_u * a.cdt) : 0.0f;
  const float kK = KICK ? (float)(a.ckick * a.st->aK * a.inv_dx2) : 0.0f;
  const float kV = KICK ? (float)(a.ckick * a.st->aV * 0.5) : 0.0f;
  const float kG = KICK ? (float)(a.ckick * a.st->aG * a.inv_dx2) : 0.0f;
  const float kJ = KICK ? (float)(a.ckick * a.st->aJ * a.jnorm * a.gQ) : 0.0f;
  const float th_scale = (float)(-a.dx * a.gQ);
  const bool u1 = a.has_u1 != 0;
  double acc[5] = {0, 0, 0, 0, 0};

  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < V; s += gridDim.x * kBlock) {
    Coord c;
    c.N = N;
    c.NN = NN;
    c.x = s / NN;
    c.y = (s - c.x * NN) / N;
    c.z = s - c.x * NN - c.y * N;

    auto phi = [&](int comp, uint32_t i) -> P2 {
      P2 v{((const float *)a.phi_in[comp])[i], a.philo_in[comp][i]};
      if (DRIFT) pair_add(v.h, v.l, cfs * ((const float *)a.pi_in[comp])[i]);
      return v;
    };
    auto A = [&](int d, uint32_t i) -> P2 {
      if (!u1) return P2{0.0f, 0.0f};
      P2 v{((const float *)a.A_in[d])[i], a.Alo_in[d][i]};
      if (DRIFT) pair_add(v.h, v.l, cfu * ((const float *)a.piA_in[d])[i]);
      return v;
    };
    auto diff = [](P2 p, P2 q) -> float { return (p.h - q.h) + (p.l - q.l); };
    // (U - 1) as (cos - 1, sin) with cos - 1 = -2 sin^2(theta/2) to keep small phases accurate
    auto link_m1 = [&](P2 Ad, float &cm1, float &sn) {
      if (!u1) {
        cm1 = 0.0f;
        sn = 0.0f;
        return;
      }
      const float th = th_scale * (Ad.h + Ad.l);
      float sh, ch;
      sincosf(0.5f * th, &sh, &ch);
      sn = 2.0f * sh * ch;
      cm1 = -2.0f * sh * sh;
    };

    const P2 fr = phi(0, s), fi = phi(1, s);
    if (DRIFT) {
      ((float *)a.phi_out[0])[s] = fr.h;
      a.philo_out[0][s] = fr.l;
      ((float *)a.phi_out[1])[s] = fi.h;
      a.philo_out[1][s] = fi.l;
      if (u1)
        for (int d = 0; d < 3; ++d) {
          const P2 v = A(d, s);
          ((float *)a.A_out[d])[s] = v.h;
          a.Alo_out[d][s] = v.l;
        }
    }
    const float pr = fr.h + fr.l, pim = fi.h + fi.l;

    float lap_r = 0.0f, lap_i = 0.0f;
    double g2 = 0.0;
    float cm[3], sl[3];
    P2 Ac[3];
    for (int d = 0; d < 3; ++d) Ac[d] = A(d, s);
    for (int d = 0; d < 3; ++d) {
      const uint32_t jp = c.shifted(d, 1, -1, 0), jm = c.shifted(d, -1, -1, 0);
      link_m1(Ac[d], cm[d], sl[d]);
      const P2 qr = phi(0, jp), qi = phi(1, jp);
      const float qrf = qr.h + qr.l, qif = qi.h + qi.l;
      // U phi(x+d) - phi(x) = (U - 1) phi(x+d) + (phi(x+d) - phi(x))
      const float dr = (cm[d] * qrf - sl[d] * qif) + diff(qr, fr);
      const float di = (cm[d] * qif + sl[d] * qrf) + diff(qi, fi);
      if (KICK) {
        float cm2, s2;
        link_m1(A(d, jm), cm2, s2);
        const P2 mr = phi(0, jm), mi = phi(1, jm);
        const float mrf = mr.h + mr.l, mif = mi.h + mi.l;
        // conj(U(x-d)) phi(x-d) - phi(x)
        const float er = (cm2 * mrf + s2 * mif) + diff(mr, fr);
        const float ei = (cm2 * mif - s2 * mrf) + diff(mi, fi);
        lap_r += dr + er;
        lap_i += di + ei;
      }
      if (RED) {
        const double gr = (double)dr * a.inv_dx, gi = (double)di * a.inv_dx;
        g2 += gr * gr + gi * gi;
      }
    }
    const double r2d = ((double)fr.h + fr.l) * ((double)fr.h + fr.l) + ((double)fi.h + fi.l) * ((double)fi.h + fi.l);
    float p_r = ((const float *)a.pi_in[0])[s], p_i = ((const float *)a.pi_in[1])[s];
    if (KICK) {
      const float nd = cs_dnorm_over_norm_f(a.poly, pr * pr + pim * pim);
      p_r += kK * lap_r - kV * (nd * pr);
      p_i += kK * lap_i - kV * (nd * pim);
      ((float *)a.pi_out[0])[s] = p_r;
      ((float *)a.pi_out[1])[s] = p_i;
    }
    if (RED) {
      acc[0] += (double)p_r * p_r + (double)p_i * p_i;
      acc[1] += g2;
      acc[2] += cs_pot_d(a.poly, r2d);
    }

    if (u1) {
      P2 Ap[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Ap[i][j] = A(j, c.shifted(i, 1, -1, 0));
      double e2 = 0.0;
      for (int i = 0; i < 3; ++i) {
        float pa = ((const float *)a.piA_in[i])[s];
        if (KICK) {
          float lap = 0.0f;
          for (int d = 0; d < 3; ++d) lap += diff(Ap[d][i], Ac[i]) + diff(A(i, c.shifted(d, -1, -1, 0)), Ac[i]);
          float gd = 0.0f;
          for (int j = 0; j < 3; ++j) {
            const P2 a_pimj = A(j, c.shifted(i, 1, j, -1));
            const P2 a_mj = A(j, c.shifted(j, -1, -1, 0));
            gd += diff(Ap[i][j], Ac[j]) - diff(a_pimj, a_mj);
          }
          const uint32_t sp = c.shifted(i, 1, -1, 0);
          const P2 qr = phi(0, sp), qi = phi(1, sp);
          const float qrf = qr.h + qr.l, qif = qi.h + qi.l;
          const float ur = (cm[i] + 1.0f) * qrf - sl[i] * qif;
          const float ui = (cm[i] + 1.0f) * qif + sl[i] * qrf;
          const float J = ur * pim - ui * pr;
          pa += kG * (lap - gd) - kJ * J;
          ((float *)a.piA_out[i])[s] = pa;
        }
        e2 += (double)pa * pa;
      }
      if (RED) {
        double b2 = 0.0;
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < i; ++j) {
            const double f = ((double)diff(Ap[i][j], Ac[j]) - (double)diff(Ap[j][i], Ac[i])) * a.inv_dx;
            b2 += f * f;
          }
        acc[3] += e2;
        acc[4] += b2;
      }
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
    if constexpr (F32)                                                                                          \
      hipLaunchKernelGGL((k_gauge_pass_f32<D, K, R>), dim3(blocks), dim3(kBlock), 0, s, a);                     \
    else                                                                                                        \
      hipLaunchKernelGGL((k_gauge_pass_f64<D, K, R>), dim3(blocks), dim3(kBlock), 0, s, a);                     \
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
