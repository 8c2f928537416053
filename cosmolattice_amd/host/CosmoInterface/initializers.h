// Initial conditions with the reference's API (initializers/modelinitializer.h:47-81,
// scalefactorinitializer.h:30-89, scalarsingletinitializer.h:41-77, u1initializer.h:77-230).  The Gaussian random fields are
// drawn on the GPU (clmi_grf_scalar) from Philox4x32-10 keyed by FNV-1a-64 of the seed string, so a
// run is reproducible from its parameter file; the stream is not TempLat's.
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "CosmoInterface/abstractmodel.h"

namespace TempLat {

inline uint64_t seedFromString(const std::string &s)
{
  uint64_t h = 1469598103934665603ull;  // FNV-1a 64
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ull;
  }
  return h;
}

template <typename T> class ModelInitializer
{
public:
  template <class Model>
  ModelInitializer(Model &, T pLSide, std::string pSeed) : lSide(pLSide), baseSeed(std::move(pSeed))
  {
  }

  // modelinitializer.h:47-81
  template <class Model, class... Extra> void initialize(Model &model, RunParameters<T> &rPar, Extra &&...)
  {
    if (rPar.a0 != T(1))
      throw RunParametersInconsistent("a0 != 1 is not part of this library's initial conditions");
    model.attachDevice(rPar);
    clmi_ctx *ctx = model.context();
    model.aI = model.aSI = model.aIM = rPar.a0;
    model.aDotI = 0;
    if (rPar.expansion) initializeScaleFactor(model);

    const double fs = model.fStar, fo = model.fStar * model.omegaStar;
    // scalarsingletinitializer.h:41-77: vacuum fluctuations, then the homogeneous values
    const uint64_t key = seedFromString(baseSeed);
    for (size_t f = 0; f < Model::Ns; ++f)
      model.check(clmi_grf_scalar(ctx, (int)f, key, (uint32_t)f, model.masses2S((int)f), model.aDotI, model.kIR,
                                  rPar.kCutoff, model.fldS0((int)f) / fs, model.piS0((int)f) / fo),
                  "clmi_grf_scalar");
    // complex scalar with constrained phases, gauge momenta from the Gauss constraint
    // (u1initializer.h:77-117)
    if constexpr (Model::NCs > 0) {
      const double m2[2] = {model.masses2CS(0).re, model.masses2CS(0).im};
      const double f0[2] = {model.fldCS0(0).re / fs, model.fldCS0(0).im / fs};
      const double p0[2] = {model.piCS0(0).re / fo, model.piCS0(0).im / fo};
      model.check(clmi_grf_u1(ctx, key, (uint32_t)(Model::Ns + 1), m2, model.aDotI, model.kIR, rPar.kCutoff, f0, p0),
                  "clmi_grf_u1");
    }
    model.pushBackground();
    // setAllAverages, then aDotI from the Friedmann constraint (modelinitializer.h:71-80)
    if (rPar.expansion) {
      double hc[3];
      model.check(clmi_hubble_constraint(ctx, hc), "clmi_hubble_constraint");
      model.pullBackground();
      model.aDotI = std::sqrt(hc[2]);
      model.pushBackground();
    } else {
      model.check(clmi_set_all_averages(ctx), "clmi_set_all_averages");
      model.pullBackground();
    }
  }

  // scalefactorinitializer.h:49-64: aDot from the homogeneous energy.
  template <class Model> static void initializeScaleFactor(Model &model)
  {
    const double fo = model.fStar * model.omegaStar;
    double kin = 0;
    for (size_t i = 0; i < Model::Ns; ++i) kin += 0.5 * std::pow(model.piS0((int)i) / fo, 2);
    if constexpr (Model::NCs > 0) {
      const auto &p = model.piCS0(0);
      kin += (p.re * p.re + p.im * p.im) / (fo * fo);
    }
    const double rho = kin + model.pot0;
    model.aDotI = model.fStar / Model::MPl * std::sqrt(rho / 3.0);
  }

  T lSide;
  std::string baseSeed;
};

}  // namespace TempLat
