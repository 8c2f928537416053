// Evolvers with the reference's public API (evolvers/leapfrog.h:23-288, evolvers/velocityverlet.h:33-278,
// evolvers/positionverlet.h:29-296, evolvers/evolver.h:40-160).  Each call is one C-ABI call into libclmi; the lattice work runs on the
// GPU, asynchronously, on the model's context stream.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "CosmoInterface/abstractmodel.h"

namespace TempLat {

class EvolverTypeNotInEvolver : public std::runtime_error
{
public:
  using std::runtime_error::runtime_error;
};
class NotVVEvolver : public std::runtime_error
{
public:
  using std::runtime_error::runtime_error;
};
class NotPVEvolver : public std::runtime_error
{
public:
  using std::runtime_error::runtime_error;
};

// Composition weights of velocity-Verlet schemes of order 2-10 (evolvers/velocityverletparameters.h:23-47).
// A step runs stages w[0], ..., w[n-1], ..., w[0] (2n-1 stages).  These are the published symmetric
// composition coefficients (Yoshida 1990 for order 4 and 6; Kahan and Li 1997 for 6 (alternative), 8 and 10).
template <typename T> struct VelocityVerletParameters {
  static std::vector<T> getCoeffs(EvolverType ev)
  {
    switch (ev) {
      case VV2: return {T(1.0)};
      case VV4: return {T(1.351207191959657771818), T(-1.702414403875838200264)};
      case VV6:
        return {T(0.78451361047755726382), T(0.23557321335935813368), T(-1.1776799841788710069),
                T(1.3151863206839112189)};
      case VV6_2:
        return {T(0.39216144400731413928), T(0.33259913678935943860), T(-0.70624617255763935981),
                T(0.082213596293550800230), T(0.79854399093482996340)};
      case VV8:
        return {T(0.74167036435061295345), T(-0.40910082580003159400), T(0.19075471029623837995),
                T(-0.57386247111608226666), T(0.29906418130365592384), T(0.33462491824529818378),
                T(0.31529309239676659663), T(-0.79688793935291635402)};
      case VV10:
        return {T(-0.48159895600253002870), T(0.0036303931544595926879), T(0.50180317558723140279),
                T(0.28298402624506254868),  T(0.80702967895372223806),   T(-0.026090580538592205447),
                T(-0.87286590146318071547), T(-0.52373568062510581643),  T(0.44521844299952789252),
                T(0.18612289547097907887),  T(0.23137327866438360633),   T(-0.52191036590418628905),
                T(0.74866113714499296793),  T(0.066736511890604057532),  T(-0.80360324375670830316),
                T(0.91249037635867994571)};
      default: throw NotVVEvolver("not a velocity-Verlet evolver");
    }
  }
  static size_t getOrder(EvolverType ev)
  {
    switch (ev) {
      case VV2: return 2;
      case VV4: return 4;
      case VV6:
      case VV6_2: return 6;
      case VV8: return 8;
      case VV10: return 10;
      default: throw NotVVEvolver("not a velocity-Verlet evolver");
    }
  }
  static bool isVerlet(EvolverType ev)
  {
    return ev == VV2 || ev == VV4 || ev == VV6 || ev == VV8 || ev == VV10 || ev == VV6_2;
  }
};

// Position Verlet uses the same composition weights (evolvers/positionverletparameters.h:18-73).
template <typename T> struct PositionVerletParameters {
  static EvolverType asVV(EvolverType ev)
  {
    switch (ev) {
      case PV2: return VV2;
      case PV4: return VV4;
      case PV6: return VV6;
      case PV6_2: return VV6_2;
      case PV8: return VV8;
      case PV10: return VV10;
      default: throw NotPVEvolver("not a position-Verlet evolver");
    }
  }
  static std::vector<T> getCoeffs(EvolverType ev) { return VelocityVerletParameters<T>::getCoeffs(asVV(ev)); }
  static size_t getOrder(EvolverType ev) { return VelocityVerletParameters<T>::getOrder(asVV(ev)); }
  static bool isVerlet(EvolverType ev)
  {
    return ev == PV2 || ev == PV4 || ev == PV6 || ev == PV8 || ev == PV10 || ev == PV6_2;
  }
};

template <typename T> class LeapFrog
{
public:
  // leapfrog.h:29-38: synced, and aDotSI = aDotI - dt/2 K for a new run.
  template <class Model> LeapFrog(Model &model, RunParameters<T> &rPar) : expansion(rPar.expansion)
  {
    model.attachDevice(rPar);
    model.pushBackground();
    model.check(clmi_lf_init(model.context(), rPar.doWeRestart ? 1 : 0), "clmi_lf_init");
    model.pullBackground();
  }
  // leapfrog.h:42-89, one step of dt.
  template <class Model> void evolve(Model &model, T /*tMinust0*/)
  {
    model.check(clmi_lf_evolve(model.context(), 1), "clmi_lf_evolve");
  }
  // Several steps in one call (same result as repeated evolve; fewer host calls).
  template <class Model> void evolveSteps(Model &model, int nsteps)
  {
    model.check(clmi_lf_evolve(model.context(), nsteps), "clmi_lf_evolve");
  }
  // leapfrog.h:96-116: half-kick the momenta to integer time; refreshes the host background.
  template <class Model> void sync(Model &model, T /*tMinust0*/)
  {
    model.check(clmi_lf_sync(model.context()), "clmi_lf_sync");
    model.pullBackground();
  }
  template <class Model> bool isSynced(const Model &model) const { return clmi_lf_is_synced(model.context()) != 0; }

  bool expansion;
};

template <typename T> class VelocityVerlet
{
public:
  template <class Model>
  VelocityVerlet(Model &model, RunParameters<T> &rPar)
      : ws(VelocityVerletParameters<T>::getCoeffs(rPar.eType)), expansion(rPar.expansion)
  {
    model.attachDevice(rPar);
    model.pushBackground();
  }
  // velocityverlet.h:51-114, one step of dt.
  template <class Model> void evolve(Model &model, T /*tMinust0*/, bool /*evolveGWs*/ = false)
  {
    evolveSteps(model, 1);
  }
  template <class Model> void evolveSteps(Model &model, int nsteps)
  {
    std::vector<double> w(ws.begin(), ws.end());
    model.check(clmi_vv_evolve(model.context(), w.data(), (int)w.size(), nsteps), "clmi_vv_evolve");
  }
  // velocityverlet.h:119-122: fields and momenta already live at integer times.
  template <class Model> void sync(Model &model, T /*tMinust0*/) { model.pullBackground(); }

  std::vector<T> ws;
  bool expansion;
};

template <typename T> class PositionVerlet
{
public:
  template <class Model>
  PositionVerlet(Model &model, RunParameters<T> &rPar)
      : ws(PositionVerletParameters<T>::getCoeffs(rPar.eType)), expansion(rPar.expansion)
  {
    model.attachDevice(rPar);
    model.pushBackground();
  }
  // positionverlet.h:47-111, one step of dt.
  template <class Model> void evolve(Model &model, T /*tMinust0*/, bool /*evolveGWs*/ = false)
  {
    evolveSteps(model, 1);
  }
  template <class Model> void evolveSteps(Model &model, int nsteps)
  {
    std::vector<double> w(ws.begin(), ws.end());
    model.check(clmi_pv_evolve(model.context(), w.data(), (int)w.size(), nsteps), "clmi_pv_evolve");
  }
  // positionverlet.h:115-119: only a fixed background changes aDotI here; refresh the host copy.
  template <class Model> void sync(Model &model, T /*tMinust0*/) { model.pullBackground(); }

  std::vector<T> ws;
  bool expansion;
};

// Dispatcher of evolvers/evolver.h:40-160 for the evolvers this library implements.
template <class Model> class Evolver
{
public:
  using T = typename Model::FloatType;

  template <class... Extra> Evolver(Model &model, RunParameters<T> &rPar, Extra &&...) : type(rPar.eType)
  {
    if (type == LF)
      lf = std::make_unique<LeapFrog<T>>(model, rPar);
    else if (VelocityVerletParameters<T>::isVerlet(type))
      vv = std::make_unique<VelocityVerlet<T>>(model, rPar);
    else if (PositionVerletParameters<T>::isVerlet(type))
      pv = std::make_unique<PositionVerlet<T>>(model, rPar);
    else
      throw EvolverTypeNotInEvolver(
          "The evolver type you specified is not implemented by this library (LF, VV2-VV10 and PV2-PV10 are).");
  }
  void evolve(Model &model, T tMinust0) const
  {
    if (lf)
      lf->evolve(model, tMinust0);
    else if (vv)
      vv->evolve(model, tMinust0);
    else
      pv->evolve(model, tMinust0);
  }
  void sync(Model &model, T tMinust0) const
  {
    if (lf)
      lf->sync(model, tMinust0);
    else if (vv)
      vv->sync(model, tMinust0);
    else
      pv->sync(model, tMinust0);
  }
  void evolveSteps(Model &model, int nsteps) const
  {
    if (lf)
      lf->evolveSteps(model, nsteps);
    else if (vv)
      vv->evolveSteps(model, nsteps);
    else
      pv->evolveSteps(model, nsteps);
  }

  EvolverType type;

private:
  std::unique_ptr<LeapFrog<T>> lf;
  std::unique_ptr<VelocityVerlet<T>> vv;
  std::unique_ptr<PositionVerlet<T>> pv;
};

}  // namespace TempLat
