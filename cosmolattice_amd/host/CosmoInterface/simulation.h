// The time loop of the reference's main program (source/cosmolattice.cpp:60-201) over this library:
// parameters -> model -> initial conditions (or a raw-field load) -> evolver -> measure at cadence
// -> evolve, plus an optional raw dump of the final state.
//
// Raw field files: <dir>/<kind>_<index>_<comp>.f64, N^3 little-endian doubles, [x][y][z] with z
// fastest; <dir>/background.txt: the 20 clmi_background values in declaration order, one per line.
#pragma once

#include <cstdio>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "CosmoInterface/evolvers.h"
#include "CosmoInterface/initializers.h"
#include "CosmoInterface/measurements.h"

namespace TempLat {

template <class Model> std::vector<std::array<int, 3>> modelArrays()
{
  std::vector<std::array<int, 3>> r;
  for (size_t i = 0; i < Model::Ns; ++i) {
    r.push_back({CLMI_FLD_S, (int)i, 0});
    r.push_back({CLMI_PI_S, (int)i, 0});
  }
  if constexpr (Model::NCs > 0)
    for (int c = 0; c < 2; ++c) {
      r.push_back({CLMI_FLD_CS, 0, c});
      r.push_back({CLMI_PI_CS, 0, c});
    }
  if constexpr (Model::NU1 > 0)
    for (int c = 0; c < 3; ++c) {
      r.push_back({CLMI_FLD_U1, 0, c});
      r.push_back({CLMI_PI_U1, 0, c});
    }
  return r;
}
inline std::string rawName(const std::string &dir, const std::array<int, 3> &a)
{
  return dir + "/" + std::to_string(a[0]) + "_" + std::to_string(a[1]) + "_" + std::to_string(a[2]) + ".f64";
}

template <class Model> void dumpState(Model &model, const std::string &dir)
{
  for (const auto &a : modelArrays<Model>()) {
    const auto h = model.getField(a[0], a[1], a[2]);
    std::ofstream f(rawName(dir, a), std::ios::binary);
    f.write(reinterpret_cast<const char *>(h.data()), (std::streamsize)(h.size() * sizeof(double)));
    if (!f) throw std::runtime_error("cannot write " + rawName(dir, a));
  }
  clmi_background b{};
  model.check(clmi_get_background(model.context(), &b), "clmi_get_background");
  std::ofstream f(dir + "/background.txt");
  const double *v = &b.aI;
  f << std::setprecision(17);
  for (size_t i = 0; i < sizeof(b) / sizeof(double); ++i) f << v[i] << "\n";
}

template <class Model> void loadState(Model &model, const std::string &dir)
{
  const size_t V = (size_t)model.N * model.N * model.N;
  for (const auto &a : modelArrays<Model>()) {
    std::vector<double> h(V);
    std::ifstream f(rawName(dir, a), std::ios::binary);
    f.read(reinterpret_cast<char *>(h.data()), (std::streamsize)(V * sizeof(double)));
    if (!f) throw std::runtime_error("cannot read " + rawName(dir, a));
    model.setField(a[0], a[1], a[2], h);
  }
  clmi_background b{};
  std::ifstream f(dir + "/background.txt");
  double *v = &b.aI;
  for (size_t i = 0; i < sizeof(b) / sizeof(double); ++i) f >> v[i];
  if (!f) throw std::runtime_error("cannot read " + dir + "/background.txt");
  model.check(clmi_set_background(model.context(), &b), "clmi_set_background");
  model.pullBackground();
}

template <class ModelType> int runCosmoLattice(ParameterParser &parser)
{
  using FloatType = typename ModelType::FloatType;
  RunParameters<FloatType> runParams(parser);
  const std::string loadDir = parser.get<std::string>("clmi_load_dir", std::string());
  const std::string dumpDir = parser.get<std::string>("clmi_dump_dir", std::string());
  runParams.setDoWeRestart(false);

  auto toolBox = MemoryToolBox<ModelType::NDim>::makeShared(runParams.N, 1);
  ModelType model(parser, runParams, toolBox);
  Measurer<ModelType> measurer(model, runParams, parser);

  double t = runParams.t0;
  model.t0 = runParams.t0;
  if (loadDir.empty()) {
    ModelInitializer<FloatType> initializer(model, runParams.lSide, runParams.baseSeed);
    initializer.initialize(model, runParams);
  } else {
    model.attachDevice(runParams);
    loadState(model, loadDir);
  }

  Evolver<ModelType> evolver(model, runParams);
  for (int i = 0; t < runParams.tMax - runParams.dt / 2; ++i) {
    if (measurer.areWeMeasuring(i)) {
      evolver.sync(model, t - runParams.t0);
      measurer.measure(i, t, model);
    }
    evolver.evolve(model, t - runParams.t0);
    t += model.dt;
  }
  if (!dumpDir.empty()) dumpState(model, dumpDir);
  return 0;
}

}  // namespace TempLat
