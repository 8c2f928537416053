// Measurements at output cadence with the reference's cadence rule and file format
// (measurements/measurer.h:43-53,181-197, measurements/energiesmeasurer.h:45-166,
// measurements/measurementsIO/std/measurementssaverstd.h:24-90): each line is the values at 15
// significant digits separated by four spaces, with an optional '#' header line.
#pragma once

#include <cmath>
#include <fstream>
#include <iomanip>
#include <set>
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

// Power spectrum of one real component with the normalisations of PowerSpectrumMeasurer
// (measurements/powerspectrum.h:85-134): returns (k = nominal bin centre x kIR, PS) per bin, and, when
// asked, the mean |k| of the bin's modes (kavg) and the bin's mode count (mult).
// PS_version 3 bins |n| f (PS_type 1) or |n|^3 f (PS_type 2): the |n| weight sits inside the bin sum
// and the prefactor carries neither the bin's |n| nor 1/dk.
template <class Model>
void powerSpectrum(Model &model, int kind, int index, int comp, double dk, int psType, int psVersion,
                   std::vector<double> &k, std::vector<double> &ps, std::vector<double> *kavg = nullptr,
                   std::vector<double> *mult = nullptr)
{
  clmi_ctx *ctx = model.context();
  const int nb = clmi_spectrum_nbins(ctx, dk);
  if (nb < 1) throw std::runtime_error("bad spectrum bin width");
  std::vector<double> s(nb), cnt(nb), mn(nb);
  const int npow = psVersion == 3 ? (psType == 1 ? 1 : 3) : 0;
  model.check(clmi_spectrum_bins_npow(ctx, kind, index, comp, dk, nb, npow, s.data(), cnt.data(), mn.data()),
              "clmi_spectrum_bins_npow");
  const double N = (double)model.N, kIR = model.kIR, dx = model.dx, pi = 3.14159265358979323846;
  k.assign(nb, 0.0);
  ps.assign(nb, 0.0);
  if (kavg) {
    kavg->assign(nb, 0.0);
    for (int b = 0; b < nb; ++b) (*kavg)[b] = mn[b] * kIR;
  }
  if (mult) *mult = cnt;
  for (int b = 0; b < nb; ++b) {
    k[b] = b * dk * kIR;
    if (psVersion == 3) {
      if (psType == 1)
        ps[b] = kIR * dx / std::pow(N, 5) / 2 / pi * s[b];
      else
        ps[b] = std::pow(kIR * dx / N, 3) / 2 / (pi * pi) * (cnt[b] > 0 ? s[b] / cnt[b] : 0.0);
      continue;
    }
    const double nt = psVersion == 1 ? mn[b] : b * dk;
    if (psType == 1)
      ps[b] = kIR * nt * dx / dk / std::pow(N, 5) / 2 / pi * s[b];
    else
      ps[b] = std::pow(kIR * nt * dx / N, 3) / dk / 2 / (pi * pi) * (cnt[b] > 0 ? s[b] / cnt[b] : 0.0);
  }
}

template <class Model> class Measurer
{
public:
  using T = typename Model::FloatType;

  Measurer(Model &, const RunParameters<T> &par, ParameterParser &)
      : outputFreq((int)std::lround(par.tOutFreq / par.dt)),
        infreqOutputFreq((int)std::lround(par.tOutInfreq / par.dt)),
        rareOutputFreq((int)std::lround(par.tOutRareFreq / par.dt)), expansion(par.expansion),
        printHeaders(par.printHeaders), dk(par.deltaKBin), psType(par.powerSpectrumType),
        psVersion(par.powerSpectrumVersion), verbosity(par.spectraVerbosity), dir(par.outFn),
        energyCons(par.outFn + "average_energy_conservation.txt",
                   expansion ? std::vector<std::string>{"t", "rel_diff_friedmann", "LHS_friedmann", "RHS_friedmann"}
                             : std::vector<std::string>{"t", "energy_cons"},
                   par.printHeaders),
        scaleFactor(par.outFn + "average_scale_factor.txt", {"t", "a", "aDot", "H"}, par.printHeaders),
        energies(par.outFn + "average_energies.txt", energyHeaders(), par.printHeaders)
  {
    if (outputFreq < 1 || infreqOutputFreq < 1 || rareOutputFreq < 1)
      throw RunParametersInconsistent("output intervals must be at least one time step");
    const std::vector<std::string> header{"t", "<phi>", "<pi>", "<phi^2>", "<pi^2>", "rms(phi)", "rms(pi)"};
    for (size_t i = 0; i < Model::Ns; ++i)
      means.emplace_back(par.outFn + "average_scalar_" + std::to_string(i) + ".txt", header, par.printHeaders);
    // complexscalarmeasurer.h:35-56: real part, imaginary part and norm of each complex scalar
    for (size_t i = 0; i < Model::NCs; ++i) {
      const std::string n = "cmplx_scalar_" + std::to_string(i);
      csMeans.emplace_back(par.outFn + "average_" + n + "_0.txt", header, par.printHeaders);
      csMeans.emplace_back(par.outFn + "average_" + n + "_1.txt", header, par.printHeaders);
      csMeans.emplace_back(par.outFn + "average_norm_" + n + ".txt", header, par.printHeaders);
    }
    // u1measurer.h:42-62: norms of E and B, and the Gauss law, of each U(1) field
    for (size_t i = 0; i < Model::NU1; ++i) {
      u1Means.emplace_back(par.outFn + "average_norm_U1_" + std::to_string(i) + ".txt",
                           std::vector<std::string>{"t", "<|E|>", "<|B|>", "<|E|^2>", "<|B|^2>", "rms(|E|)", "rms(|B|)"},
                           par.printHeaders);
      gauss.emplace_back(par.outFn + "average_gauss_U1_" + std::to_string(i) + ".txt",
                         std::vector<std::string>{"t", "av|LHS-RHS|_over_av|LHS+RHS|", "av|LHS|", "av|RHS|"},
                         par.printHeaders);
    }
  }

  // measurer.h:181-197
  bool areWeMeasuring(int n) const
  {
    return n % outputFreq == 0 || n % infreqOutputFreq == 0 || n % rareOutputFreq == 0;
  }

  void measure(int n, T t, Model &model)
  {
    if (n % outputFreq == 0) measureStandard(t, model);
    if (n % infreqOutputFreq == 0) measureSpectra(t, model);
  }

  int outputFreq, infreqOutputFreq, rareOutputFreq;

private:
  static std::vector<std::string> energyHeaders()
  {
    std::vector<std::string> h{"t"};
    for (size_t i = 0; i < Model::Ns; ++i) {
      h.push_back("E^kin_scal" + std::to_string(i));
      h.push_back("E^grad_scal" + std::to_string(i));
    }
    if (Model::NCs > 0) {
      h.push_back("E^kin_cmplxscal0");
      h.push_back("E^grad_cmplxscal0");
    }
    if (Model::NU1 > 0) {
      h.push_back("E^kin_U10");
      h.push_back("E^grad_U10");
    }
    for (size_t k = 0; k < Model::NPotTerms; ++k) h.push_back("Vpot_term_" + std::to_string(k));
    h.push_back("E_tot");
    return h;
  }

  // MeansMeasurer::measure (meansmeasurer.h:28-48) from <f>, <f^2>, <p>, <p^2>:
  // t, <f>, <p>, <f^2>, <p^2>, rms(f), rms(p), with a negative variance written as rms 0.
  static std::vector<double> meansRow(double t, double f1, double f2, double p1, double p2)
  {
    const double rf = f2 - f1 * f1, rp = p2 - p1 * p1;
    return {t, f1, p1, f2, p2, rf > 0 ? std::sqrt(rf) : 0.0, rp > 0 ? std::sqrt(rp) : 0.0};
  }

  // Complex scalars (complexscalarmeasurer.h:71-79): the components sqrt2 phi_c and
  // sqrt2 a^(alpha-3) pi_c, then |phi| and a^(alpha-3) |pi|.  U(1) fields (u1measurer.h:70-86):
  // |E| = a^(alpha-1) |pi_A| and |B|, then the Gauss law.
  void measureGaugeSector(T t, Model &model)
  {
    clmi_ctx *ctx = model.context();
    const double a = model.aI, al = model.alpha;
    for (size_t i = 0; i < Model::NCs; ++i) {
      const double sp = std::sqrt(2.0) * std::pow(a, al - 3);
      for (int c = 0; c < 2; ++c) {
        double f[2], p[2];
        model.check(clmi_field_means(ctx, CLMI_FLD_CS, (int)i, c, std::sqrt(2.0), f), "clmi_field_means");
        model.check(clmi_field_means(ctx, CLMI_PI_CS, (int)i, c, sp, p), "clmi_field_means");
        csMeans[3 * i + c].save(meansRow(t, f[0], f[1], p[0], p[1]));
      }
      double f[2], p[2];
      model.check(clmi_norm_averages(ctx, CLMI_FLD_CS, (int)i, f), "clmi_norm_averages");
      model.check(clmi_norm_averages(ctx, CLMI_PI_CS, (int)i, p), "clmi_norm_averages");
      const double s = std::pow(a, al - 3);
      csMeans[3 * i + 2].save(meansRow(t, f[0], f[1], s * p[0], s * s * p[1]));
    }
    for (size_t i = 0; i < Model::NU1; ++i) {
      double e[2], b[2], g[3];
      model.check(clmi_norm_averages(ctx, CLMI_PI_U1, (int)i, e), "clmi_norm_averages");
      model.check(clmi_norm_averages(ctx, CLMI_MAG_U1, (int)i, b), "clmi_norm_averages");
      const double s = std::pow(a, al - 1);
      u1Means[i].save(meansRow(t, s * e[0], s * s * e[1], b[0], b[1]));
      model.check(clmi_gauss_u1(ctx, g), "clmi_gauss_u1");
      gauss[i].save({(double)t, g[0], g[1], g[2]});
    }
  }

  // Frequent output: scalar means (meansmeasurer.h:28-48, scalarsingletmeasurer.h:60-66), energies
  // (energiesmeasurer.h:45-103), energy conservation (energiesmeasurer.h:105-122), scale factor.
  void measureStandard(T t, Model &model)
  {
    clmi_ctx *ctx = model.context();
    model.pullBackground();
    const double a = model.aI, al = model.alpha;
    std::vector<double> row{(double)t};
    double etot = 0;
    for (size_t i = 0; i < Model::Ns; ++i) {
      double sf[3], sp[3];
      model.check(clmi_field_stats(ctx, CLMI_FLD_S, (int)i, 0, sf), "clmi_field_stats");
      model.check(clmi_field_stats(ctx, CLMI_PI_S, (int)i, 0, sp), "clmi_field_stats");
      const double sc = std::pow(a, al - 3);  // momenta as field velocities
      const double p1 = sp[0] * sc, p2 = sp[1] * sc * sc;
      const double rf = sf[1] - sf[0] * sf[0], rp = p2 - p1 * p1;
      means[i].save({(double)t, sf[0], p1, sf[1], p2, rf > 0 ? std::sqrt(rf) : 0.0, rp > 0 ? std::sqrt(rp) : 0.0});
      const double ek = 0.5 * sp[1] * std::pow(a, -6.0), eg = 0.5 * sf[2] * std::pow(a, -2.0);
      row.push_back(ek);
      row.push_back(eg);
      etot += ek + eg;
    }
    clmi_averages av{};
    model.check(clmi_volume_averages(ctx, &av), "clmi_volume_averages");
    const EnergyTotals e = energyTotals(model, av);
    if (Model::NCs > 0) {
      row.push_back(e.kinCS);
      row.push_back(e.gradCS);
      etot += e.kinCS + e.gradCS;
    }
    if (Model::NU1 > 0) {
      row.push_back(e.elec);
      row.push_back(e.mag);
      etot += e.elec + e.mag;
    }
    measureGaugeSector(t, model);
    std::vector<double> mono(CLMI_MAX_TERMS, 0.0);
    if (Model::Ns > 0) model.check(clmi_monomial_averages(ctx, mono.data()), "clmi_monomial_averages");
    const auto &tm = model.termMonomials();
    for (size_t k = 0; k < Model::NPotTerms; ++k) {
      double v = 0;
      if (Model::NCs > 0 && Model::NPotTerms == 1)
        v = av.pot;
      else
        for (const auto &[idx, c] : tm[k]) v += c * mono[idx];
      row.push_back(v);
      etot += v;
    }
    row.push_back(etot);
    energies.save(row);
    if (expansion) {
      double hc[3];
      model.check(clmi_hubble_constraint(ctx, hc), "clmi_hubble_constraint");
      model.pullBackground();
      energyCons.save({(double)t, hc[0], hc[1], hc[2]});
      scaleFactor.save({(double)t, model.aI, model.aDotI, model.aDotI / model.aI});
    } else {
      if (first_) {
        etot0_ = etot;
        first_ = false;
      }
      energyCons.save({(double)t, std::fabs(1.0 - etot / etot0_)});
    }
  }

  // One block of a spectra file (spectrumsaverstd.h:54-96): a row per bin, then an empty line; the
  // header line is written once, before the first block.  The bin columns follow spectraVerbosity
  // (spectrumsaverstd.h:56-82): 0 -> the bin's central k; 1 -> the mean |k| of the bin's modes (the
  // HDF5 writer's momBinAverage, hdf5/spectrumsaverhdf5.h:98-100); 2 -> central k, mean |k| and the
  // bin's mode count (momMultiplicity).  Then one column per spectrum.
  void saveSpectra(const std::string &name, T t, const std::vector<double> &k, const std::vector<double> &kavg,
                   const std::vector<double> &mult, const std::vector<std::vector<double>> &cols,
                   const std::vector<std::string> &names)
  {
    const std::string fn = dir + "spectra_" + name + ".txt";
    const bool started = spectraStarted_.count(fn) > 0;
    std::ofstream out(fn, started ? std::ios_base::app : std::ios_base::out);
    if (!out) throw std::runtime_error("cannot open " + fn);
    out << std::setprecision(15);
    if (!started && printHeaders) {
      if (verbosity == 0) out << "#k";
      else if (verbosity == 1) out << "#k_avg";
      else out << "#k    k_avg    multiplicity";
      for (const auto &n : names) out << "    " << n;
      out << "    t = " << t << "\n";
    }
    for (size_t b = 0; b < k.size(); ++b) {
      if (verbosity == 0) out << k[b];
      else if (verbosity == 1) out << kavg[b];
      else out << k[b] << "    " << kavg[b] << "    " << mult[b];
      for (const auto &c : cols) out << "    " << c[b];
      out << "\n";
    }
    out << "\n";
    spectraStarted_.insert(fn);
  }

  // Infrequent output: spectra of every scalar and its momentum (scalarsingletmeasurer.h:74-83),
  // k, PS(phi), a^(2 alpha - 6) PS(pi) per bin; of the complex scalar's norm
  // (complexscalarmeasurer.h:85-95), summed over its two components; of the U(1) electric and magnetic
  // fields (u1measurer.h:91-105), a^(2 alpha - 2) sum_i PS(pi_A,i) and sum_i PS(B_i).
  void measureSpectra(T t, Model &model)
  {
    const double a = model.aI, al = model.alpha;
    std::vector<double> k, kavg, mult, pf, pp;
    auto sum = [&](int kind, int idx, int ncomp, double scale) {
      std::vector<double> acc, c;
      for (int q = 0; q < ncomp; ++q) {
        powerSpectrum(model, kind, idx, q, dk, psType, psVersion, k, c, &kavg, &mult);
        if (acc.empty()) acc.assign(c.size(), 0.0);
        for (size_t b = 0; b < c.size(); ++b) acc[b] += c[b];
      }
      for (auto &v : acc) v *= scale;
      return acc;
    };
    for (size_t i = 0; i < Model::Ns; ++i) {
      pf = sum(CLMI_FLD_S, (int)i, 1, 1.0);
      pp = sum(CLMI_PI_S, (int)i, 1, std::pow(a, 2 * al - 6));
      saveSpectra("scalar_" + std::to_string(i), t, k, kavg, mult, {pf, pp}, {"PS(phi)", "PS(pi)"});
    }
    for (size_t i = 0; i < Model::NCs; ++i) {
      pf = sum(CLMI_FLD_CS, (int)i, 2, 1.0);
      pp = sum(CLMI_PI_CS, (int)i, 2, std::pow(a, 2 * al - 6));
      saveSpectra("norm_cmplx_scalar_" + std::to_string(i), t, k, kavg, mult, {pf, pp}, {"PS(|phi|)", "PS(|pi|)"});
    }
    for (size_t i = 0; i < Model::NU1; ++i) {
      const auto pe = sum(CLMI_PI_U1, (int)i, 3, std::pow(a, 2 * al - 2));
      const auto pb = sum(CLMI_MAG_U1, (int)i, 3, 1.0);
      saveSpectra("norm_U1_" + std::to_string(i), t, k, kavg, mult, {pe, pb}, {"PS(E)", "PS(B)"});
    }
  }

  bool expansion, printHeaders;
  double dk;
  int psType, psVersion;
  int verbosity;  // spectraVerbosity: which bin columns the spectra files carry
  std::string dir;
  AverageFile energyCons, scaleFactor, energies;
  std::vector<AverageFile> means, csMeans, u1Means, gauss;
  std::set<std::string> spectraStarted_;
  bool first_ = true;
  double etot0_ = 0.0;
};

}  // namespace TempLat
