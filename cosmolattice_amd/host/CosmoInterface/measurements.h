// Measurements at output cadence with the reference's cadence rule and file format
// (measurements/measurer.h:43-53,181-197, measurements/energiesmeasurer.h:45-166,
// measurements/measurementsIO/std/measurementssaverstd.h:24-90): each line is the values at 15
// significant digits separated by four spaces, with an optional '#' header line.
#pragma once

#include <cmath>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "CosmoInterface/abstractmodel.h"

namespace TempLat {

class AverageFile
{
public:
  AverageFile(const std::string &path, const std::vector<std::string> &headers, bool printHeaders)
      : out_(path, std::ios_base::out)
  {
    if (!out_) throw std::runtime_error("cannot open output file " + path);
    if (printHeaders) {
      out_ << "#";
      for (size_t i = 0; i < headers.size(); ++i) out_ << (i ? "    " : "") << headers[i];
      out_ << "\n";
    }
  }
  void save(const std::vector<double> &values)
  {
    std::ostringstream s;
    s << std::setprecision(15);
    for (size_t i = 0; i < values.size(); ++i) s << (i ? "    " : "") << values[i];
    out_ << s.str() << "\n";
    out_.flush();
  }

private:
  std::ofstream out_;
};

// Energy bookkeeping of definitions/energies.h:37-85 from the device volume averages.
struct EnergyTotals {
  double kinS, gradS, kinCS, gradCS, elec, mag, pot, total;
};
template <class Model> EnergyTotals energyTotals(Model &model, const clmi_averages &av)
{
  const double a = model.aI;
  const double a6 = std::pow(a, -6.0), a2 = std::pow(a, -2.0), a4 = std::pow(a, -4.0);
  const double w2 = (model.omegaStar / model.fStar) * (model.omegaStar / model.fStar);
  EnergyTotals e{};
  e.kinS = 0.5 * av.pi2S * a6;
  e.gradS = 0.5 * av.grad2S * a2;
  e.kinCS = av.pi2CS * a6;
  e.gradCS = av.grad2CS * a2;
  e.elec = 0.5 * w2 * a4 * av.pi2U1;
  e.mag = 0.5 * w2 * a4 * av.B2U1;
  e.pot = av.pot;
  e.total = e.kinS + e.gradS + e.kinCS + e.gradCS + e.elec + e.mag + e.pot;
  return e;
}

template <class Model> class Measurer
{
public:
  using T = typename Model::FloatType;

  Measurer(Model &, const RunParameters<T> &par, ParameterParser &)
      : outputFreq((int)std::lround(par.tOutFreq / par.dt)),
        infreqOutputFreq((int)std::lround(par.tOutInfreq / par.dt)),
        rareOutputFreq((int)std::lround(par.tOutRareFreq / par.dt)), expansion(par.expansion),
        energyCons(par.outFn + "average_energy_conservation.txt",
                   expansion ? std::vector<std::string>{"t", "rel_diff_friedmann", "LHS_friedmann", "RHS_friedmann"}
                             : std::vector<std::string>{"t", "energy_cons"},
                   par.printHeaders),
        scaleFactor(par.outFn + "average_scale_factor.txt", {"t", "a", "aDot", "H"}, par.printHeaders)
  {
    if (outputFreq < 1 || infreqOutputFreq < 1 || rareOutputFreq < 1)
      throw RunParametersInconsistent("output intervals must be at least one time step");
  }

  // measurer.h:181-197
  bool areWeMeasuring(int n) const
  {
    return n % outputFreq == 0 || n % infreqOutputFreq == 0 || n % rareOutputFreq == 0;
  }

  // Frequent output: energy conservation (energiesmeasurer.h:105-122) and the scale factor.
  void measure(int n, T t, Model &model)
  {
    if (n % outputFreq != 0) return;
    if (expansion) {
      double hc[3];
      model.check(clmi_hubble_constraint(model.context(), hc), "clmi_hubble_constraint");
      model.pullBackground();
      energyCons.save({(double)t, hc[0], hc[1], hc[2]});
      scaleFactor.save({(double)t, model.aI, model.aDotI, model.aDotI / model.aI});
    } else {
      clmi_averages av{};
      model.check(clmi_volume_averages(model.context(), &av), "clmi_volume_averages");
      const double etot = energyTotals(model, av).total;
      if (first_) {
        etot0_ = etot;
        first_ = false;
      }
      energyCons.save({(double)t, std::fabs(1.0 - etot / etot0_)});
    }
  }

  int outputFreq, infreqOutputFreq, rareOutputFreq;

private:
  bool expansion;
  AverageFile energyCons, scaleFactor;
  bool first_ = true;
  double etot0_ = 0.0;
};

}  // namespace TempLat
