// Per-site arithmetic of the charged scalar + U(1) lattice pass, shared by the one-site-per-thread
// pass (gauge_pass.hip) and the x-marching tiled pass (gauge_tiled.hip), which differ only in how
// they gather the (drifted) neighbour values -- so both give bit-identical fields.
//
// Reference expressions (one TempLat sweep each in the reference):
//   CS kick  piCS += (w dt)(aI^(1+alpha) D^2 phi - aI^(3+alpha)/2 dV/dphi)     kernels/complexscalarkernels.h:28-35
//            D^2 phi = dx^-2 [sum_i U_i(x) phi(x+i) + conj(U_i(x-i)) phi(x-i) - 6 phi]
//                                                                         definitions/gaugederivatives.h:48-53
//            U_i(x) = exp(-i dx gQ A_i(x))                                definitions/gaugederivatives.h:92-102
//   U1 kick  piU1_i += (w dt)(aI^(alpha-1)(Lapl A_i - sum_j d-_j d+_i A_j) - aI^(1+alpha) J_i)
//            J_i = 2/dx (fStar/omegaStar)^2 gQ Im[conj(U_i(x) phi(x+i)) phi(x)]
//                                                  kernels/u1kernels.h:31-52, definitions/mattercurrents.h:52-76
//   sums     <|piCS|^2>, <sum_j |D+_j phi|^2>, <V>, <sum_i piU1_i^2>, <F21^2+F31^2+F32^2>
//                                                  definitions/fieldfunctionals.h:80-153, averages.h:39-71
#pragma once

#include "clmi_internal.hpp"

namespace clmi {
namespace {

// FP32-bulk field value: hi + lo.
struct P2 {
  float h, l;
};

// Drifted field values around one site.  A[j] components j = 0..2 are directions x, y, z.
template <typename V>
struct GaugeIn {
  V fc[2];        // phi components at x
  V fp[3][2];     // phi at x + d
  V fm[3][2];     // phi at x - d
  V Ac[3];        // A_j(x)
  V Ap[3][3];     // [i][j]: A_j(x + i)
  V Am[3][3];     // [i][j]: A_j(x - i)
  V Ad[3][3];     // [i][j]: A_j(x + i - j), i != j
};

// Kick and reduction constants of one pass.
struct GaugeK {
  // FP64
  double ckick, aK, aV, aG, aJ, inv_dx, inv_dx2, dx, gQ, jnorm;
  // FP32-bulk
  float kK, kV, kG, kJ, th_scale;
  bool u1;
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
__device__ __forceinline__ float cs_pot_f(const CsPoly &P, float r2)
{
  float v = 0.0f;
  for (int k = 0; k < P.nterms; ++k) {
    float p = (float)P.c[k];
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

// FP64 site update.  pin, pAin: momenta at x (read); pout, pAout: kicked momenta; acc: the five sums.
template <bool KICK, bool RED>
__device__ __forceinline__ void gauge_site_f64(const GaugeK &k, const CsPoly &poly, const GaugeIn<double> &g,
                                               const double pin[2], const double pAin[3], double pout[2],
                                               double pAout[3], double acc[5])
{
  const double pr = g.fc[0], pim = g.fc[1];
  double plus_r = 0, plus_i = 0, minus_r = 0, minus_i = 0, g2 = 0;
  double cl[3], sl[3];  // links U_d(x)
  for (int d = 0; d < 3; ++d) {
    if (k.u1)
      sincos(-k.dx * k.gQ * g.Ac[d], &sl[d], &cl[d]);
    else {
      cl[d] = 1.0;
      sl[d] = 0.0;
    }
    const double fpr = g.fp[d][0], fpi = g.fp[d][1];
    const double ur = cl[d] * fpr - sl[d] * fpi;
    const double ui = cl[d] * fpi + sl[d] * fpr;
    plus_r += ur;
    plus_i += ui;
    if (KICK) {
      double c2 = 1.0, s2 = 0.0;
      if (k.u1) sincos(-k.dx * k.gQ * g.Am[d][d], &s2, &c2);
      const double fmr = g.fm[d][0], fmi = g.fm[d][1];
      minus_r += c2 * fmr + s2 * fmi;
      minus_i += c2 * fmi - s2 * fmr;
    }
    if (RED) {
      const double gr = (ur - pr) * k.inv_dx, gi = (ui - pim) * k.inv_dx;
      g2 += gr * gr + gi * gi;
    }
  }
  const double r2 = pr * pr + pim * pim;
  double p_r = pin[0], p_i = pin[1];
  if (KICK) {
    const double lap_r = (plus_r + minus_r - 6 * pr) * k.inv_dx2;
    const double lap_i = (plus_i + minus_i - 6 * pim) * k.inv_dx2;
    const double nd = cs_dnorm_over_norm_d(poly, r2);
    p_r += k.ckick * (k.aK * lap_r - k.aV / 2 * (nd * pr));
    p_i += k.ckick * (k.aK * lap_i - k.aV / 2 * (nd * pim));
  }
  pout[0] = p_r;
  pout[1] = p_i;
  if (RED) {
    acc[0] += p_r * p_r + p_i * p_i;
    acc[1] += g2;
    acc[2] += cs_pot_d(poly, r2);
  }
  if (k.u1) {
    double e2 = 0.0;
    for (int i = 0; i < 3; ++i) {
      double pa = pAin[i];
      if (KICK) {
        double lap = 0.0;
        for (int d = 0; d < 3; ++d) lap += g.Ap[d][i] + g.Am[d][i] - 2 * g.Ac[i];
        lap *= k.inv_dx2;
        double gd = 0.0;
        for (int j = 0; j < 3; ++j) {
          const double a_pimj = j == i ? g.Ac[i] : g.Ad[i][j];
          const double a_mj = g.Am[j][j];
          gd += (g.Ap[i][j] - g.Ac[j]) - (a_pimj - a_mj);
        }
        gd *= k.inv_dx2;
        const double fpr = g.fp[i][0], fpi = g.fp[i][1];
        const double ur = cl[i] * fpr - sl[i] * fpi;
        const double ui = cl[i] * fpi + sl[i] * fpr;
        const double J = k.jnorm * k.gQ * (ur * pim - ui * pr);
        pa += k.ckick * (k.aG * (lap - gd) - k.aJ * J);
      }
      pAout[i] = pa;
      e2 += pa * pa;
    }
    if (RED) {
      double b2 = 0.0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < i; ++j) {
          const double f = ((g.Ap[i][j] - g.Ac[j]) - (g.Ap[j][i] - g.Ac[i])) * k.inv_dx;
          b2 += f * f;
        }
      acc[3] += e2;
      acc[4] += b2;
    }
  }
}

__device__ __forceinline__ float p2diff(P2 p, P2 q) { return (p.h - q.h) + (p.l - q.l); }

// (U - 1) as (cos - 1, sin) with cos - 1 = -2 sin^2(theta/2) to keep small phases accurate
__device__ __forceinline__ void link_m1(const GaugeK &k, P2 Ad, float &cm1, float &sn)
{
  if (!k.u1) {
    cm1 = 0.0f;
    sn = 0.0f;
    return;
  }
  const float th = k.th_scale * (Ad.h + Ad.l);
  float sh, ch;
  sincosf(0.5f * th, &sh, &ch);
  sn = 2.0f * sh * ch;
  cm1 = -2.0f * sh * sh;
}

// FP32-bulk site update (stencil differences on hi and lo separately, kick and per-site energy terms in
// FP32, accumulated over sites in FP64, as the scalar FP32-bulk pass does).
template <bool KICK, bool RED>
__device__ __forceinline__ void gauge_site_f32(const GaugeK &k, const CsPoly &poly, const GaugeIn<P2> &g,
                                               const float pin[2], const float pAin[3], float pout[2],
                                               float pAout[3], double acc[5])
{
  const P2 fr = g.fc[0], fi = g.fc[1];
  const float pr = fr.h + fr.l, pim = fi.h + fi.l;
  float lap_r = 0.0f, lap_i = 0.0f;
  float g2 = 0.0f;  // per-site terms in FP32, accumulated over sites in FP64
  float cm[3], sl[3];
  for (int d = 0; d < 3; ++d) {
    link_m1(k, g.Ac[d], cm[d], sl[d]);
    const P2 qr = g.fp[d][0], qi = g.fp[d][1];
    const float qrf = qr.h + qr.l, qif = qi.h + qi.l;
    // U phi(x+d) - phi(x) = (U - 1) phi(x+d) + (phi(x+d) - phi(x))
    const float dr = (cm[d] * qrf - sl[d] * qif) + p2diff(qr, fr);
    const float di = (cm[d] * qif + sl[d] * qrf) + p2diff(qi, fi);
    if (KICK) {
      float cm2, s2;
      link_m1(k, g.Am[d][d], cm2, s2);
      const P2 mr = g.fm[d][0], mi = g.fm[d][1];
      const float mrf = mr.h + mr.l, mif = mi.h + mi.l;
      // conj(U(x-d)) phi(x-d) - phi(x)
      const float er = (cm2 * mrf + s2 * mif) + p2diff(mr, fr);
      const float ei = (cm2 * mif - s2 * mrf) + p2diff(mi, fi);
      lap_r += dr + er;
      lap_i += di + ei;
    }
    if (RED) g2 += dr * dr + di * di;
  }
  float p_r = pin[0], p_i = pin[1];
  if (KICK) {
    const float nd = cs_dnorm_over_norm_f(poly, pr * pr + pim * pim);
    p_r += k.kK * lap_r - k.kV * (nd * pr);
    p_i += k.kK * lap_i - k.kV * (nd * pim);
  }
  pout[0] = p_r;
  pout[1] = p_i;
  if (RED) {
    acc[0] += (double)(p_r * p_r + p_i * p_i);
    acc[1] += (double)g2 * k.inv_dx2;
    acc[2] += (double)cs_pot_f(poly, pr * pr + pim * pim);
  }
  if (k.u1) {
    float e2 = 0.0f;
    for (int i = 0; i < 3; ++i) {
      float pa = pAin[i];
      if (KICK) {
        float lap = 0.0f;
        for (int d = 0; d < 3; ++d) lap += p2diff(g.Ap[d][i], g.Ac[i]) + p2diff(g.Am[d][i], g.Ac[i]);
        float gd = 0.0f;
        for (int j = 0; j < 3; ++j) {
          const P2 a_pimj = j == i ? g.Ac[i] : g.Ad[i][j];
          const P2 a_mj = g.Am[j][j];
          gd += p2diff(g.Ap[i][j], g.Ac[j]) - p2diff(a_pimj, a_mj);
        }
        const P2 qr = g.fp[i][0], qi = g.fp[i][1];
        const float qrf = qr.h + qr.l, qif = qi.h + qi.l;
        const float ur = (cm[i] + 1.0f) * qrf - sl[i] * qif;
        const float ui = (cm[i] + 1.0f) * qif + sl[i] * qrf;
        const float J = ur * pim - ui * pr;
        pa += k.kG * (lap - gd) - k.kJ * J;
      }
      pAout[i] = pa;
      e2 += pa * pa;
    }
    if (RED) {
      float b2 = 0.0f;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < i; ++j) {
          const float f = p2diff(g.Ap[i][j], g.Ac[j]) - p2diff(g.Ap[j][i], g.Ac[i]);
          b2 += f * f;
        }
      acc[3] += (double)e2;
      acc[4] += (double)b2 * k.inv_dx2;
    }
  }
}

// Field value type, drifted load and store per storage mode.
template <bool F32> struct GVal;
template <> struct GVal<false> {
  using V = double;
  using P = double;
  __device__ static __forceinline__ V fld(const void *f, const float *, const void *p, uint32_t i, bool drift,
                                          double c)
  {
    double v = ((const double *)f)[i];
    if (drift) v += c * ((const double *)p)[i];
    return v;
  }
  __device__ static __forceinline__ void store(void *f, float *, uint32_t i, V v) { ((double *)f)[i] = v; }
};
template <> struct GVal<true> {
  using V = P2;
  using P = float;
  __device__ static __forceinline__ V fld(const void *f, const float *lo, const void *p, uint32_t i, bool drift,
                                          float c)
  {
    P2 v{((const float *)f)[i], lo[i]};
    if (drift) pair_add(v.h, v.l, c * ((const float *)p)[i]);
    return v;
  }
  __device__ static __forceinline__ void store(void *f, float *lo, uint32_t i, V v)
  {
    ((float *)f)[i] = v.h;
    lo[i] = v.l;
  }
};

// Constants of one pass from the device background state.
__device__ __forceinline__ GaugeK gauge_consts(const GaugePassArgs &a, bool kick)
{
  GaugeK k;
  k.ckick = a.ckick;
  k.aK = kick ? a.st->aK : 0.0;
  k.aV = kick ? a.st->aV : 0.0;
  k.aG = kick ? a.st->aG : 0.0;
  k.aJ = kick ? a.st->aJ : 0.0;
  k.inv_dx = a.inv_dx;
  k.inv_dx2 = a.inv_dx2;
  k.dx = a.dx;
  k.gQ = a.gQ;
  k.jnorm = a.jnorm;
  k.kK = kick ? (float)(a.ckick * a.st->aK * a.inv_dx2) : 0.0f;
  k.kV = kick ? (float)(a.ckick * a.st->aV * 0.5) : 0.0f;
  k.kG = kick ? (float)(a.ckick * a.st->aG * a.inv_dx2) : 0.0f;
  k.kJ = kick ? (float)(a.ckick * a.st->aJ * a.jnorm * a.gQ) : 0.0f;
  k.th_scale = (float)(-a.dx * a.gQ);
  k.u1 = a.has_u1 != 0;
  return k;
}

}  // namespace
}  // namespace clmi
