// Model base class with the reference's model-header API (abstractmodel.h:35-80 DefaultModelPars and
// MakeModel, abstractmodel.h:88-300 AbstractModel, abstractmodel/{scalarbase,complexscalarbase,u1base,
// scalefactorbase,modelparametersbase}.h).  A model header written for the reference
// (`struct ModelPars : DefaultModelPars {...}; class M : public MakeModel(M, ModelPars) {...}`) compiles
// against this class unchanged for the sectors this library evolves: real scalar singlets, one
// complex scalar and one U(1) gauge field, with polynomial potentials.
//
// The lattice fields live on the GPU, in the clmi context this class owns.  The potential terms and
// derivatives the header defines are collected symbolically (symbolic.h) into the monomial list
// the library evaluates, and the derivatives are checked against the potential once, when the
// context is created.  Host members aI, aDotI, pi2AvI, ... mirror the device background; they are
// refreshed by the evolvers' sync and by the initializer (the points where the reference's
// measurements read them).
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "CosmoInterface/couplingsmanager.h"
#include "CosmoInterface/runparameters.h"
#include "CosmoInterface/symbolic.h"
#include "clmi.h"

#define STRINGIFY_CLMI_INNER(x) #x
#ifndef STRINGIFY
#define STRINGIFY(x) STRINGIFY_CLMI_INNER(x)
#endif
#ifndef MODELLABEL
#define MODELLABEL model
#endif

namespace TempLat {

class ClmiError : public std::runtime_error
{
public:
  ClmiError(int code, const std::string &msg) : std::runtime_error(msg), code(code) {}
  int code;
};
class PotentialDerivativeNotDefined : public std::runtime_error
{
public:
  using std::runtime_error::runtime_error;
};
class ModelNotSupported : public std::runtime_error
{
public:
  using std::runtime_error::runtime_error;
};

namespace Constants {
template <typename T> constexpr T reducedMPlanck = T(2.435e18);
}


struct DefaultModelPars {
  static constexpr size_t NScalars = 0;
  static constexpr size_t NCScalars = 0;
  static constexpr size_t NU1Flds = 0;
  static constexpr size_t NSU2Doublet = 0;
  static constexpr size_t NSU2Flds = 0;
  static constexpr bool DefectsModel = false;
  static constexpr size_t NPotTerms = 1;
  using FloatType = double;
  typedef CouplingsManager<0, 0> CsU1Couplings;
};

// Lattice size holder passed to model constructors (the reference's MemoryToolBox carries the
// decomposition; here the lattice is one GPU's).
template <size_t NDIM> class MemoryToolBox
{
public:
  MemoryToolBox(ptrdiff_t N, int nGhost) : N(N), nGhost(nGhost) {}
  static std::shared_ptr<MemoryToolBox> makeShared(ptrdiff_t N, int nGhost)
  {
    return std::make_shared<MemoryToolBox>(N, nGhost);
  }
  bool amIRoot() const { return true; }
  int getNProcesses() const { return 1; }
  ptrdiff_t N;
  int nGhost;
};
namespace device::memory {
template <class X> using host_ptr = std::shared_ptr<X>;
}

#define MakeModel(_ModelName, _ModelParsType) ::TempLat::AbstractModel<_ModelName, _ModelParsType>

template <class R, class MP> class AbstractModel
{
public:
  using FloatType = typename MP::FloatType;
  using T = FloatType;
  using CsU1Couplings = typename MP::CsU1Couplings;
  static constexpr size_t NDim = 3;
  static constexpr size_t Ns = MP::NScalars;
  static constexpr size_t NCs = MP::NCScalars;
  static constexpr size_t NU1 = MP::NU1Flds;
  static constexpr size_t NSU2Doublet = MP::NSU2Doublet;
  static constexpr size_t NSU2 = MP::NSU2Flds;
  static constexpr size_t NPotTerms = MP::NPotTerms;
  static constexpr bool DefectsModel = MP::DefectsModel;
  static constexpr T MPl = Constants::reducedMPlanck<T>;
  static_assert(Ns <= CLMI_MAX_SCALARS, "at most CLMI_MAX_SCALARS real scalar singlets");
  static_assert(NCs <= 1 && NU1 <= 1, "at most one complex scalar and one U(1) field");
  static_assert(NU1 == 0 || NCs == 1, "a U(1) field needs the complex scalar it couples to");
  static_assert(NSU2Doublet == 0 && NSU2 == 0, "SU(2) sectors are not part of this library's lattice path");
  static_assert(!DefectsModel, "defect models are not part of this library's lattice path");

  // symbols (potential terms) and homogeneous initial values (abstractmodel/scalarbase.h,
  // complexscalarbase.h)
  ScalarSymbols<Ns> fldS;
  ComplexScalarSymbols<NCs> fldCS;
  TaggedArray<T, Ns> fldS0, piS0, masses2S;
  TaggedArray<ComplexValue<T>, NCs> fldCS0, piCS0, masses2CS;

  // modelparametersbase.h:19-22
  T dx, kIR, dt;
  std::string name;
  T t0 = 0, t = 0;
  T alpha = 1, fStar = 1, omegaStar = 1;
  // scalefactorbase.h:20-21 and the cached averages (scalarbase.h, complexscalarbase.h, u1base.h)
  T aSI = 1, aI = 1, aIM = 1;
  T aDotSI = 0, aDotSIM = 0, aDotI = 0, aDotIM = 0;
  T pi2AvI = 0, pi2AvSI = 0, pi2AvSIM = 0, grad2AvI = 0;
  T CSpi2AvI = 0, CSpi2AvSI = 0, CSpi2AvSIM = 0, CSgrad2AvI = 0;
  T U1pi2AvI = 0, U1pi2AvSI = 0, U1pi2AvSIM = 0, U1Mag2AvI = 0;
  T potAvI = 0, potAvSI = 0, pot0 = 0, pot0SI = 0;
  T gQ = 0;  // effective charge g*Q of complex scalar 0 under U(1) field 0
  ptrdiff_t N;

  AbstractModel(ParameterParser &parser, const LatticeParameters<T> &par,
                device::memory::host_ptr<MemoryToolBox<NDim>> toolBox, T pDt, std::string pName = "")
      : dx(par.getDx()), kIR(par.getKIR()), dt(pDt), name(std::move(pName)), N(toolBox ? toolBox->N : par.getN())
  {
    // complexscalarbase.h:71-73
    if constexpr (NCs > 0 && NU1 > 0 && CsU1Couplings::active) {
      const auto g = parser.get<T, (size_t)CsU1Couplings::nGauge>("gU1s", T(1));
      const auto q = parser.get<T, CsU1Couplings::howManyCouples()>("CSU1_charges", T(1));
      typename CsU1Couplings::template Container<T> cm;
      cm.setEffectiveCharges(std::vector<T>(q.begin(), q.end()), std::vector<T>(g.begin(), g.end()));
      gQ = cm(Tag<0>{}, Tag<0>{});
    }
  }
  AbstractModel(const AbstractModel &) = delete;
  AbstractModel &operator=(const AbstractModel &) = delete;
  ~AbstractModel()
  {
    if (ctx_) clmi_destroy(ctx_);
  }

  // ---- potential ---------------------------------------------------------------------------

  // Sum of potentialTerms(Tag<0..NPotTerms-1>) as a polynomial in the program-unit fields.
  Poly potentialPolynomial() { return sumTerms(std::make_index_sequence<NPotTerms>{}); }

  // Program-unit homogeneous point: phi_n = fldS0(n)/fStar, |phi| = |fldCS0|/fStar with the
  // reference's 1e-14 floor for a vanishing condensate (abstractmodel.h:255-274).
  std::array<double, kNumVars> initialPoint() const
  {
    std::array<double, kNumVars> x{};
    for (size_t j = 0; j < Ns; ++j) x[j] = fldS0(j) / fStar;
    if constexpr (NCs > 0) {
      const double re = fldCS0(0).re, im = fldCS0(0).im;
      x[kNormCSVar] = (re == 0 && im == 0) ? std::sqrt(2.0) * (1e-14 * 0.70710678118654752)
                                           : std::sqrt(re * re + im * im) / fStar;
    }
    return x;
  }

  // abstractmodel.h:220-248: V and the diagonal second derivatives at the homogeneous point.
  void setInitialPotentialAndMassesFromPotential()
  {
    const auto x = initialPoint();
    pot0 = potentialPolynomial().eval(x);
    setMasses(x, std::make_index_sequence<Ns>{});
    if constexpr (NCs > 0) {
      // Potential::deriv2CS (definitions/potential.h:36-43,83-88) per component c:
      //   V''(|phi|) phi_c^2/|phi|^2 + V'(|phi|)/|phi| (1 - phi_c^2/|phi|^2)
      R &self = static_cast<R &>(*this);
      Poly d1 = potentialPolynomial().derivative(kNormCSVar);
      Poly d2 = d1.derivative(kNormCSVar);
      if constexpr (requires(R &r) { r.potDerivNormCS(Tag<0>{}); }) d1 = Poly(self.potDerivNormCS(Tag<0>{}));
      if constexpr (requires(R &r) { r.potDeriv2NormCS(Tag<0>{}); }) d2 = Poly(self.potDeriv2NormCS(Tag<0>{}));
      const double r = x[kNormCSVar];
      double c[2] = {fldCS0(0).re / fStar, fldCS0(0).im / fStar};
      if (c[0] == 0 && c[1] == 0) c[0] = c[1] = 1e-14 * 0.70710678118654752;
      double m2[2];
      for (int k = 0; k < 2; ++k) {
        const double w = c[k] * c[k] / (r * r);
        m2[k] = d2.eval(x) * w + d1.eval(x) / r * (1 - w);
      }
      masses2CS(0) = Complexify<T>(m2[0], m2[1]);
    }
  }
  void setInitialPotentialFromPotential() { pot0 = potentialPolynomial().eval(initialPoint()); }

  // ---- device context ------------------------------------------------------------------------

  // The library configuration of this model for these run parameters (potential as monomials, checked
  // against the header's derivatives).
  clmi_config deviceConfig(const RunParameters<T> &rPar)
  {
    if (rPar.N != N) throw ModelNotSupported("run parameters and model disagree on N");
    clmi_config cfg{};
    cfg.N = (int)N;
    cfg.n_scalars = (int)Ns;
    cfg.n_cscalars = (int)NCs;
    cfg.n_u1 = (int)NU1;
    cfg.precision = rPar.precision;
    cfg.expansion = rPar.expansion ? 1 : 0;
    cfg.device = rPar.device;
    cfg.dx = dx;
    cfg.dt = dt;
    cfg.alpha = alpha;
    cfg.fStar = fStar;
    cfg.omegaStar = omegaStar;
    cfg.MPl = MPl;
    cfg.gQ = gQ;
    fillPotential(cfg);
    return cfg;
  }

  // Create the GPU lattice for these run parameters (idempotent).  All fields start at zero.
  void attachDevice(const RunParameters<T> &rPar)
  {
    if (ctx_) return;
    const clmi_config cfg = deviceConfig(rPar);
    const int rc = clmi_create(&cfg, &ctx_);
    if (rc != CLMI_OK) {
      ctx_ = nullptr;
      throw ClmiError(rc, std::string("clmi_create: ") + clmi_last_error(nullptr));
    }
    expansion_ = rPar.expansion;
    pushBackground();
  }
  bool hasDevice() const { return ctx_ != nullptr; }
  clmi_ctx *context() const
  {
    if (!ctx_) throw ClmiError(CLMI_E_STATE, "the model has no GPU lattice yet (attachDevice)");
    return ctx_;
  }
  void check(int rc, const char *what) const
  {
    if (rc != CLMI_OK) throw ClmiError(rc, std::string(what) + ": " + clmi_last_error(ctx_));
  }

  // Host mirror <-> device background.
  void pullBackground()
  {
    clmi_background b{};
    check(clmi_get_background(context(), &b), "clmi_get_background");
    aI = b.aI, aSI = b.aSI, aIM = b.aIM, aDotI = b.aDotI, aDotSI = b.aDotSI, aDotSIM = b.aDotSIM;
    pi2AvI = b.pi2AvI, pi2AvSI = b.pi2AvSI, pi2AvSIM = b.pi2AvSIM, grad2AvI = b.grad2AvI;
    CSpi2AvI = b.CSpi2AvI, CSpi2AvSI = b.CSpi2AvSI, CSpi2AvSIM = b.CSpi2AvSIM, CSgrad2AvI = b.CSgrad2AvI;
    U1pi2AvI = b.U1pi2AvI, U1pi2AvSI = b.U1pi2AvSI, U1pi2AvSIM = b.U1pi2AvSIM, U1Mag2AvI = b.U1Mag2AvI;
    potAvI = b.potAvI, potAvSI = b.potAvSI;
  }
  void pushBackground()
  {
    clmi_background b{aI,       aSI,       aIM,        aDotI,    aDotSI,    aDotSIM,    pi2AvI,
                      pi2AvSI,  pi2AvSIM,  grad2AvI,   CSpi2AvI, CSpi2AvSI, CSpi2AvSIM, CSgrad2AvI,
                      U1pi2AvI, U1pi2AvSI, U1pi2AvSIM, U1Mag2AvI, potAvI,   potAvSI};
    check(clmi_set_background(context(), &b), "clmi_set_background");
  }

  // Whole-lattice host copies (N^3 doubles, [x][y][z], z fastest) of one real component.
  std::vector<double> getField(int kind, int index, int comp = 0) const
  {
    std::vector<double> h((size_t)N * N * N);
    check(clmi_download_field(context(), kind, index, comp, h.data()), "clmi_download_field");
    return h;
  }
  void setField(int kind, int index, int comp, const std::vector<double> &h)
  {
    if (h.size() != (size_t)N * N * N) throw std::invalid_argument("setField: need N^3 values");
    check(clmi_upload_field(context(), kind, index, comp, h.data()), "clmi_upload_field");
  }
  bool expansionOn() const { return expansion_; }

  // Per potential term k: (index into the library's scalar monomial list, coefficient) of each of its
  // monomials, so that <potentialTerms(Tag<k>)> = sum coefficient * <monomial>.  Filled when the
  // context is created; empty for terms in |phi| of the complex scalar.
  const std::vector<std::vector<std::pair<int, double>>> &termMonomials() const { return termMonomials_; }

private:
  template <size_t... K> Poly sumTerms(std::index_sequence<K...>)
  {
    R &self = static_cast<R &>(*this);
    Poly v;
    ((v += Poly(self.potentialTerms(Tag<(int)K>{}))), ...);
    return v;
  }
  template <size_t... J> void setMasses(const std::array<double, kNumVars> &x, std::index_sequence<J...>)
  {
    R &self = static_cast<R &>(*this);
    ((masses2S((int)J) = (T)secondDerivative<(int)J>(self).eval(x)), ...);
  }
  template <int J> Poly secondDerivative(R &self)
  {
    if constexpr (requires(R &r) { r.potDeriv2(Tag<J>{}); })
      return Poly(self.potDeriv2(Tag<J>{}));
    else
      return potentialPolynomial().derivative(J).derivative(J);
  }
  // potDeriv(Tag<n>) of the header must be dV/dphi_n: the library differentiates the monomials.
  template <size_t... J> void checkDerivatives(const Poly &V, std::index_sequence<J...>)
  {
    R &self = static_cast<R &>(*this);
    (checkOne<(int)J>(self, V), ...);
    if constexpr (NCs > 0) {
      if constexpr (requires(R &r) { r.potDerivNormCS(Tag<0>{}); }) {
        if (!Poly(self.potDerivNormCS(Tag<0>{})).approxEqual(V.derivative(kNormCSVar)))
          throw PotentialDerivativeNotDefined(
              "potDerivNormCS(Tag<0>) is not the derivative of the potential with respect to |phi|");
      }
    }
  }
  template <int J> void checkOne(R &self, const Poly &V)
  {
    if constexpr (requires(R &r) { r.potDeriv(Tag<J>{}); }) {
      if (!Poly(self.potDeriv(Tag<J>{})).approxEqual(V.derivative(J)))
        throw PotentialDerivativeNotDefined("potDeriv(Tag<" + std::to_string(J) +
                                            ">) is not the derivative of the potential terms");
    } else {
      throw PotentialDerivativeNotDefined("potDeriv(Tag<" + std::to_string(J) + ">) is not defined");
    }
  }
  template <size_t... K> void collectTerms(const clmi_config &cfg, std::index_sequence<K...>)
  {
    R &self = static_cast<R &>(*this);
    termMonomials_.clear();
    auto one = [&](const Poly &term) {
      std::vector<std::pair<int, double>> list;
      for (const auto &[m, c] : term.terms())
        for (int t = 0; t < cfg.n_terms; ++t) {
          bool same = m[kNormCSVar] == 0;
          for (int v = 0; v < CLMI_MAX_SCALARS; ++v) same = same && m[v] == cfg.term_exp[t][v];
          if (same) list.push_back({t, c});
        }
      termMonomials_.push_back(list);
    };
    (one(Poly(self.potentialTerms(Tag<(int)K>{}))), ...);
  }
  void fillPotential(clmi_config &cfg)
  {
    const Poly V = potentialPolynomial();
    checkDerivatives(V, std::make_index_sequence<Ns>{});
    for (const auto &[m, c] : V.terms()) {
      bool scalar = false;
      for (int v = 0; v < kNumScalarVars; ++v) scalar = scalar || m[v] != 0;
      const bool cs = m[kNormCSVar] != 0;
      if (scalar && cs) throw ModelNotSupported("potential terms mixing real scalars and |phi| are not supported");
      if (cs || (!scalar && NCs > 0 && Ns == 0)) {
        if (cfg.n_cs_terms >= CLMI_MAX_TERMS) throw ModelNotSupported("too many complex-scalar potential terms");
        cfg.cs_coeff[cfg.n_cs_terms] = c;
        cfg.cs_pow[cfg.n_cs_terms] = m[kNormCSVar];
        ++cfg.n_cs_terms;
      } else {
        if (cfg.n_terms >= CLMI_MAX_TERMS) throw ModelNotSupported("too many scalar potential terms");
        for (int v = 0; v < CLMI_MAX_SCALARS; ++v) cfg.term_exp[cfg.n_terms][v] = m[v];
        cfg.term_coeff[cfg.n_terms] = c;
        ++cfg.n_terms;
      }
    }
    collectTerms(cfg, std::make_index_sequence<NPotTerms>{});
  }

  clmi_ctx *ctx_ = nullptr;
  bool expansion_ = true;
  std::vector<std::vector<std::pair<int, double>>> termMonomials_;
};

}  // namespace TempLat
