// Symbolic vocabulary for model headers: Tag<N>, the _c literal, field symbols and polynomial
// expressions.
//
// A reference model header (models/lphi4.h:154-226, models/lphi4U1.h:117-137) writes its potential
// as TempLat expressions such as `FloatType(0.25) * pow<4>(fldS(0_c))`.  Here the same source text
// builds an exact polynomial in the program-unit fields instead of a lattice expression; the
// library evaluates the polynomial on the device.  Variables: the real scalar singlets phi_0..phi_3
// and |phi| of the complex scalar.
#pragma once

#include <array>
#include <cmath>
#include <cstddef>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "clmi.h"

namespace TempLat {

// Plain-number maths keeps resolving to the standard functions next to the polynomial overloads.
using std::pow;
using std::sqrt;

template <int N> struct Tag {
  static constexpr int value = N;
  constexpr operator int() const { return N; }
};

namespace detail {
template <char... Cs> constexpr int parse_tag()
{
  int v = 0;
  for (char c : {Cs...}) v = 10 * v + (c - '0');
  return v;
}
}  // namespace detail

// 0_c, 1_c, ... are Tag<0>, Tag<1>, ... (the reference's tag literal).
template <char... Cs> constexpr auto operator""_c() { return Tag<detail::parse_tag<Cs...>()>{}; }

// Polynomial variables.
constexpr int kNumScalarVars = CLMI_MAX_SCALARS;
constexpr int kNormCSVar = CLMI_MAX_SCALARS;  // |phi| of complex scalar 0
constexpr int kNumVars = CLMI_MAX_SCALARS + 1;

using Monomial = std::array<int, kNumVars>;

// Exact polynomial: monomial exponents -> coefficient.
class Poly
{
public:
  Poly() = default;
  Poly(double c)  // NOLINT: constants enter expressions implicitly, as in the reference's algebra
  {
    if (c != 0.0) terms_[Monomial{}] = c;
  }
  static Poly var(int v)
  {
    if (v < 0 || v >= kNumVars) throw std::out_of_range("polynomial variable out of range");
    Monomial m{};
    m[v] = 1;
    Poly p;
    p.terms_[m] = 1.0;
    return p;
  }
  const std::map<Monomial, double> &terms() const { return terms_; }

  Poly &operator+=(const Poly &o)
  {
    for (const auto &[m, c] : o.terms_) add(m, c);
    return *this;
  }
  Poly &operator-=(const Poly &o)
  {
    for (const auto &[m, c] : o.terms_) add(m, -c);
    return *this;
  }
  Poly &operator*=(const Poly &o)
  {
    Poly r;
    for (const auto &[ma, ca] : terms_)
      for (const auto &[mb, cb] : o.terms_) {
        Monomial m{};
        for (int v = 0; v < kNumVars; ++v) m[v] = ma[v] + mb[v];
        r.add(m, ca * cb);
      }
    *this = r;
    return *this;
  }
  Poly operator-() const
  {
    Poly r;
    for (const auto &[m, c] : terms_) r.terms_[m] = -c;
    return r;
  }

  // d/d(variable v)
  Poly derivative(int v) const
  {
    Poly r;
    for (const auto &[m, c] : terms_) {
      if (m[v] == 0) continue;
      Monomial d = m;
      d[v] -= 1;
      r.add(d, c * m[v]);
    }
    return r;
  }
  double eval(const std::array<double, kNumVars> &x) const
  {
    double s = 0.0;
    for (const auto &[m, c] : terms_) {
      double p = c;
      for (int v = 0; v < kNumVars; ++v)
        for (int k = 0; k < m[v]; ++k) p *= x[v];
      s += p;
    }
    return s;
  }
  // True when both polynomials agree term by term to a relative tolerance.
  bool approxEqual(const Poly &o, double rel = 1e-12) const
  {
    Poly d = *this;
    d -= o;
    double scale = 0.0;
    for (const auto &[m, c] : terms_) scale = std::fmax(scale, std::fabs(c));
    for (const auto &[m, c] : o.terms_) scale = std::fmax(scale, std::fabs(c));
    for (const auto &[m, c] : d.terms_)
      if (std::fabs(c) > rel * scale) return false;
    return true;
  }

private:
  void add(const Monomial &m, double c)
  {
    auto it = terms_.find(m);
    if (it == terms_.end()) {
      if (c != 0.0) terms_[m] = c;
      return;
    }
    it->second += c;
    if (it->second == 0.0) terms_.erase(it);
  }
  std::map<Monomial, double> terms_;
};

inline Poly operator+(Poly a, const Poly &b) { return a += b; }
inline Poly operator-(Poly a, const Poly &b) { return a -= b; }
inline Poly operator*(Poly a, const Poly &b) { return a *= b; }
inline Poly operator/(Poly a, double d) { return a *= Poly(1.0 / d); }

template <int N> Poly pow(const Poly &x)
{
  static_assert(N >= 0, "negative powers are not polynomial");
  Poly r(1.0);
  for (int k = 0; k < N; ++k) r *= x;
  return r;
}
inline Poly pow(const Poly &x, int n)
{
  if (n < 0) throw std::invalid_argument("pow: negative powers are not polynomial");
  Poly r(1.0);
  for (int k = 0; k < n; ++k) r *= x;
  return r;
}

// Symbol of complex scalar n; only |phi| (norm) enters a U(1)-invariant potential.
struct ComplexScalarSymbol {
  int index = 0;
};
inline Poly norm(const ComplexScalarSymbol &s)
{
  if (s.index != 0) throw std::out_of_range("only one complex scalar is supported");
  return Poly::var(kNormCSVar);
}
inline Poly norm2(const ComplexScalarSymbol &s) { return pow<2>(norm(s)); }

// fldS: scalar-singlet symbols, fldS(0_c) / fldS(n).
template <size_t NS> struct ScalarSymbols {
  template <int N> Poly operator()(Tag<N>) const
  {
    static_assert(N >= 0 && (size_t)N < NS, "scalar field index out of range");
    return Poly::var(N);
  }
  Poly operator()(int n) const
  {
    if (n < 0 || (size_t)n >= NS) throw std::out_of_range("scalar field index out of range");
    return Poly::var(n);
  }
};
template <size_t NC> struct ComplexScalarSymbols {
  template <int N> ComplexScalarSymbol operator()(Tag<N>) const
  {
    static_assert(N >= 0 && (size_t)N < NC, "complex scalar index out of range");
    return ComplexScalarSymbol{N};
  }
};

// Homogeneous complex value (fldCS0, piCS0): two real components.
template <typename T> struct ComplexValue {
  T re = 0, im = 0;
  T &operator()(Tag<0>) { return re; }
  T &operator()(Tag<1>) { return im; }
  const T &operator()(Tag<0>) const { return re; }
  const T &operator()(Tag<1>) const { return im; }
  T &operator[](int i) { return i == 0 ? re : im; }
  const T &operator[](int i) const { return i == 0 ? re : im; }
};
template <typename T> ComplexValue<T> Complexify(T re, T im) { return ComplexValue<T>{re, im}; }

// Fixed-size collection indexed by Tag or int (fldCS0(0_c), masses2S(j)).
template <typename V, size_t N> struct TaggedArray {
  std::array<V, N> v{};
  template <int I> V &operator()(Tag<I>)
  {
    static_assert(I >= 0 && (size_t)I < N, "index out of range");
    return v[I];
  }
  template <int I> const V &operator()(Tag<I>) const
  {
    static_assert(I >= 0 && (size_t)I < N, "index out of range");
    return v[I];
  }
  V &operator()(int i) { return v.at(i); }
  const V &operator()(int i) const { return v.at(i); }
  V &operator[](int i) { return v.at(i); }
  const V &operator[](int i) const { return v.at(i); }
  TaggedArray &operator=(const std::array<V, N> &a)
  {
    v = a;
    return *this;
  }
  static constexpr size_t size() { return N; }
};

}  // namespace TempLat
