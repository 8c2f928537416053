// The scalar-singlet potential V = sum_t c_t prod_f phi_f^e_tf (models/lphi4.h:154-206 for lphi4) and
// its derivatives, as the lattice passes evaluate them per site.
//
// Two forms with the same operation order, so they give the same bits:
//  - runtime: the exponent table comes from the ScalarPoly kernel argument; powers are loops;
//  - compile time (PK > 0): the exponent table is a type and every power is unrolled, so no scalar
//    loop runs per site.  The host selects it with poly_kind() when the run's table equals one of
//    the tables below (the coefficients stay runtime values).
// dV/dphi_n term by term: c_t * e_tn * prod_f phi_f^(e_tf - [f == n]), the factors multiplied in field
// order into the product, terms with e_tn = 0 skipped, the terms summed in order from 0; V likewise
// with c_t * prod_f phi_f^e_tf.  This is the order of the FP64 oracle.
#pragma once

#include "clmi_internal.hpp"

namespace clmi {

// Compile-time exponent tables.  PK 1: lphi4, V = c_0 phi^4 + c_1 phi^2 chi^2 (2 fields) or c_0 phi^4
// (1 field).
template <int NF, int PK> struct PolyTable;
template <> struct PolyTable<2, 1> {
  static constexpr int nt = 2;
  static constexpr int e[2][2] = {{4, 0}, {2, 2}};
};
template <> struct PolyTable<1, 1> {
  static constexpr int nt = 1;
  static constexpr int e[1][1] = {{4}};
};

// Host: the compile-time table equal to P's (exponents and term order), else 0.
inline int poly_kind(const ScalarPoly &P, int nf)
{
  if (nf == 2 && P.nterms == 2 && P.e[0][0] == 4 && P.e[0][1] == 0 && P.e[1][0] == 2 && P.e[1][1] == 2) return 1;
  if (nf == 1 && P.nterms == 1 && P.e[0][0] == 4) return 1;
  return 0;
}

namespace {

template <class T> __device__ __forceinline__ T ipow_rt(T v, int e)
{
  T r = T(1);
  for (int i = 0; i < e; ++i) r *= v;
  return r;
}
template <int E, class T> __device__ __forceinline__ T ipow_ct(T v)
{
  if constexpr (E == 0)
    return T(1);
  else
    return ipow_ct<E - 1, T>(v) * v;
}

// product p * prod_{f >= F} phi_f^(e_tf - [f == n]) (n = -1: no derivative)
template <class Tab, int NF, int t, int n, int F, class T>
__device__ __forceinline__ T poly_mul_ct(T p, const T *phi)
{
  if constexpr (F == NF) {
    return p;
  } else {
    constexpr int e = F == n ? Tab::e[t][F] - 1 : Tab::e[t][F];
    return poly_mul_ct<Tab, NF, t, n, F + 1, T>(p * ipow_ct<e, T>(phi[F]), phi);
  }
}
template <class Tab, int NF, int t, int n, class T>
__device__ __forceinline__ T dpot_ct(const ScalarPoly &P, const T *phi, T d)
{
  if constexpr (t == Tab::nt) {
    return d;
  } else {
    if constexpr (Tab::e[t][n] != 0) d += poly_mul_ct<Tab, NF, t, n, 0, T>((T)P.c[t] * (T)Tab::e[t][n], phi);
    return dpot_ct<Tab, NF, t + 1, n, T>(P, phi, d);
  }
}
template <class Tab, int NF, int t, class T>
__device__ __forceinline__ T pot_ct(const ScalarPoly &P, const T *phi, T v)
{
  if constexpr (t == Tab::nt)
    return v;
  else
    return pot_ct<Tab, NF, t + 1, T>(P, phi, v + poly_mul_ct<Tab, NF, t, -1, 0, T>((T)P.c[t], phi));
}

// dV/dphi_n (n a compile-time constant once the caller's field loop is unrolled)
template <int NF, int PK, class T> __device__ __forceinline__ T dpot(const ScalarPoly &P, const T *phi, int n)
{
  if constexpr (PK == 0) {
    T d = T(0);
    for (int t = 0; t < P.nterms; ++t) {
      const int e = P.e[t][n];
      if (e == 0) continue;
      T p = (T)P.c[t] * (T)e;
#pragma unroll
      for (int f = 0; f < NF; ++f) p *= ipow_rt<T>(phi[f], f == n ? e - 1 : P.e[t][f]);
      d += p;
    }
    return d;
  } else {
    using Tab = PolyTable<NF, PK>;
    if constexpr (NF == 1)
      return dpot_ct<Tab, NF, 0, 0, T>(P, phi, T(0));
    else
      return n == 0 ? dpot_ct<Tab, NF, 0, 0, T>(P, phi, T(0)) : dpot_ct<Tab, NF, 0, 1, T>(P, phi, T(0));
  }
}
template <int NF, int PK, class T> __device__ __forceinline__ T pot(const ScalarPoly &P, const T *phi)
{
  if constexpr (PK == 0) {
    T v = T(0);
    for (int t = 0; t < P.nterms; ++t) {
      T p = (T)P.c[t];
#pragma unroll
      for (int f = 0; f < NF; ++f) p *= ipow_rt<T>(phi[f], P.e[t][f]);
      v += p;
    }
    return v;
  } else {
    return pot_ct<PolyTable<NF, PK>, NF, 0, T>(P, phi, T(0));
  }
}

// The runtime forms under the names the passes use.
template <int NF> __device__ __forceinline__ double dpot_d(const ScalarPoly &P, const double *phi, int n)
{
  return dpot<NF, 0, double>(P, phi, n);
}
template <int NF> __device__ __forceinline__ float dpot_f(const ScalarPoly &P, const float *phi, int n)
{
  return dpot<NF, 0, float>(P, phi, n);
}
template <int NF> __device__ __forceinline__ double pot_d(const ScalarPoly &P, const double *phi)
{
  return pot<NF, 0, double>(P, phi);
}
template <int NF> __device__ __forceinline__ float pot_f(const ScalarPoly &P, const float *phi)
{
  return pot<NF, 0, float>(P, phi);
}

}  // namespace
}  // namespace clmi
