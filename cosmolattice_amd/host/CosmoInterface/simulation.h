// The time loop of the reference's main program (source/cosmolattice.cpp:60-201) over this library:
// parameters -> model -> initial conditions (or a restart) -> evolver -> measure at cadence ->
// evolve, with periodic backups and an end-of-run save (simulationmanager.h:69-150).
//
// Saved state (a directory): one raw file per real component, <kind>_<index>_<comp>.f64 with kind in
// fldS, piS, fldCS, piCS, fldU1, piU1 (N^3 little-endian doubles, [x][y][z], z fastest), and
// state.json holding the scale-factor variables the reference saves (aSI, aI, aIM, aDotSI, aDotI,
// aDotSIM, tSave), the lattice size, and the whole device background under "background".
#pragma once

#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "CosmoInterface/evolvers.h"
#include "CosmoInterface/initializers.h"
#include "CosmoInterface/measurements.h"

namespace TempLat {

inline const char *kindName(int kind)
{
  switch (kind) {
    case CLMI_FLD_S: return "fldS";
    case CLMI_PI_S: return "piS";
    case CLMI_FLD_CS: return "fldCS";
    case CLMI_PI_CS: return "piCS";
    case CLMI_FLD_U1: return "fldU1";
    default: return "piU1";
  }
}

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
  return dir + "/" + kindName(a[0]) + "_" + std::to_string(a[1]) + "_" + std::to_string(a[2]) + ".f64";
}

inline const char *const kBgNames[] = {"aI",       "aSI",       "aIM",        "aDotI",     "aDotSI",   "aDotSIM",
                                       "pi2AvI",   "pi2AvSI",   "pi2AvSIM",   "grad2AvI",  "CSpi2AvI", "CSpi2AvSI",
                                       "CSpi2AvSIM", "CSgrad2AvI", "U1pi2AvI", "U1pi2AvSI", "U1pi2AvSIM",
                                       "U1Mag2AvI", "potAvI", "potAvSI"};

// Minimal reader of the flat number-valued JSON written below: "key": value pairs.
inline std::map<std::string, double> readNumbers(const std::string &text)
{
  std::map<std::string, double> out;
  size_t p = 0;
  while ((p = text.find('"', p)) != std::string::npos) {
    const size_t q = text.find('"', p + 1);
    if (q == std::string::npos) break;
    const std::string key = text.substr(p + 1, q - p - 1);
    size_t c = text.find_first_not_of(" \t\n:", q + 1);
    p = q + 1;
    if (c == std::string::npos || text[c] == '"' || text[c] == '{' || text[c] == '[') continue;
    out[key] = std::strtod(text.c_str() + c, nullptr);
  }
  return out;
}

// SimulationManager::saveSim (simulationmanager.h:106-140) in the raw format.
template <class Model> void saveSim(Model &model, const std::string &dir, double t)
{
  std::filesystem::create_directories(dir);
  for (const auto &a : modelArrays<Model>()) {
    const auto h = model.getField(a[0], a[1], a[2]);
    std::ofstream f(rawName(dir, a), std::ios::binary);
    f.write(reinterpret_cast<const char *>(h.data()), (std::streamsize)(h.size() * sizeof(double)));
    if (!f) throw std::runtime_error("cannot write " + rawName(dir, a));
  }
  clmi_background b{};
  model.check(clmi_get_background(model.context(), &b), "clmi_get_background");
  const double *v = &b.aI;
  std::ofstream f(dir + "/state.json");
  f << std::setprecision(17) << "{\n  \"model\": \"" << model.name << "\",\n  \"N\": " << model.N
    << ",\n  \"dx\": " << model.dx << ",\n  \"dt\": " << model.dt << ",\n  \"tSave\": " << t << ",\n";
  for (const char *k : {"aSI", "aI", "aIM", "aDotSI", "aDotI", "aDotSIM"}) {
    for (int i = 0; i < 20; ++i)
      if (std::string(kBgNames[i]) == k) f << "  \"" << k << "\": " << v[i] << ",\n";
  }
  f << "  \"background\": {";
  for (int i = 0; i < 20; ++i) f << (i ? ", " : "") << "\"bg_" << kBgNames[i] << "\": " << v[i];
  f << "}\n}\n";
  if (!f) throw std::runtime_error("cannot write " + dir + "/state.json");
}

template <class Model> std::map<std::string, double> loadFields(Model &model, const std::string &dir)
{
  const size_t V = (size_t)model.N * model.N * model.N;
  std::ifstream js(dir + "/state.json");
  if (!js) throw std::runtime_error("cannot read " + dir + "/state.json");
  std::stringstream ss;
  ss << js.rdbuf();
  auto num = readNumbers(ss.str());
  if (num.count("N") == 0 || (ptrdiff_t)num["N"] != model.N)
    throw std::runtime_error("saved state has a different lattice size");
  for (const auto &a : modelArrays<Model>()) {
    std::vector<double> h(V);
    std::ifstream f(rawName(dir, a), std::ios::binary);
    f.read(reinterpret_cast<char *>(h.data()), (std::streamsize)(V * sizeof(double)));
    if (!f) throw std::runtime_error("cannot read " + rawName(dir, a));
    model.setField(a[0], a[1], a[2], h);
  }
  return num;
}

// SimulationManager::loadSim (simulationmanager.h:69-104): fields, the scale-factor variables and
// tSave; the cached averages are recomputed with setAllAverages, as the reference does.
template <class Model> void loadSim(Model &model, const std::string &dir, double &t)
{
  auto num = loadFields(model, dir);
  model.pullBackground();
  model.aSI = num.at("aSI");
  model.aI = num.at("aI");
  model.aIM = num.at("aIM");
  model.aDotSI = num.at("aDotSI");
  model.aDotI = num.at("aDotI");
  model.aDotSIM = num.at("aDotSIM");
  model.pushBackground();
  model.check(clmi_set_all_averages(model.context()), "clmi_set_all_averages");
  model.pullBackground();
  t = num.at("tSave");
}

// A saved state taken as the initial condition of a new run, with its whole background.
template <class Model> void loadState(Model &model, const std::string &dir)
{
  auto num = loadFields(model, dir);
  clmi_background b{};
  double *v = &b.aI;
  for (int i = 0; i < 20; ++i) v[i] = num.at(std::string("bg_") + kBgNames[i]);
  model.check(clmi_set_background(model.context(), &b), "clmi_set_background");
  model.pullBackground();
}

// SimulationManager::backup (simulationmanager.h:143-149): keep the previous backup as <dir>~.
template <class Model> void backup(Model &model, const std::string &dir, double t)
{
  std::error_code ec;
  if (std::filesystem::exists(dir)) {
    std::filesystem::remove_all(dir + "~", ec);
    std::filesystem::rename(dir, dir + "~", ec);
  }
  saveSim(model, dir, t);
}

template <class ModelType> int runCosmoLattice(ParameterParser &parser)
{
  using FloatType = typename ModelType::FloatType;
  const std::string restartDir = parser.get<std::string>("load_file", std::string());
  RunParameters<FloatType> runParams(parser);
  const std::string loadDir = parser.get<std::string>("clmi_load_dir", std::string());
  const std::string dumpDir = parser.get<std::string>("clmi_dump_dir", std::string());
  const double tBackup = parser.get<double>("tBackupFreq", -1.0);
  const std::string backupDir = parser.get<std::string>("backup_dir", runParams.outFn);
  const std::string saveDir = parser.get<std::string>("save_dir", std::string());
  runParams.setDoWeRestart(!restartDir.empty());

  auto toolBox = MemoryToolBox<ModelType::NDim>::makeShared(runParams.N, 1);
  ModelType model(parser, runParams, toolBox);
  Measurer<ModelType> measurer(model, runParams, parser);

  double t = runParams.t0;
  model.t0 = runParams.t0;
  if (!restartDir.empty()) {
    model.attachDevice(runParams);
    loadSim(model, restartDir, t);
  } else if (!loadDir.empty()) {
    model.attachDevice(runParams);
    loadState(model, loadDir);
  } else {
    ModelInitializer<FloatType> initializer(model, runParams.lSide, runParams.baseSeed);
    initializer.initialize(model, runParams);
  }

  Evolver<ModelType> evolver(model, runParams);
  const int backupEvery = tBackup > 0 ? (int)(tBackup / runParams.dt) : 0;
  for (int i = 0; t < runParams.tMax - runParams.dt / 2; ++i) {
    if (measurer.areWeMeasuring(i)) {
      evolver.sync(model, t - runParams.t0);
      measurer.measure(i, t, model);
    }
    evolver.evolve(model, t - runParams.t0);
    t += model.dt;
    if (backupEvery > 0 && i % backupEvery == 0) backup(model, backupDir + model.name + "_backup", t);
  }
  if (!saveDir.empty()) saveSim(model, saveDir + model.name + "_end", t);
  if (!dumpDir.empty()) saveSim(model, dumpDir, t);
  return 0;
}

}  // namespace TempLat
