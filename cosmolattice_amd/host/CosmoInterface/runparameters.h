// Run parameters of the lattice-evolution path (runparameters.h:60-236 of the reference), read
// from a ParameterParser with the reference's keys and consistency rules.
#pragma once

#include <cmath>
#include <string>

#include "CosmoInterface/parameterparser.h"
#include "clmi.h"

namespace TempLat {

// Evolver names of evolvers/evolvertype.h:37.
enum EvolverType { LF, VV2, VV4, VV6, VV8, VV10, VV6_2, RK2, RK3_3, RK3_4, RK4_5, PV2, PV4, PV6, PV8, PV10, PV6_2 };

template <> struct ParamConvert<EvolverType> {
  static EvolverType from(const std::string &k, const std::string &s)
  {
    static const char *names[] = {"LF",  "VV2",   "VV4",   "VV6",   "VV8", "VV10", "VV6_2", "RK2", "RK3_3",
                                  "RK3_4", "RK4_5", "PV2", "PV4", "PV6", "PV8", "PV10", "PV6_2"};
    for (int i = 0; i < (int)(sizeof(names) / sizeof(names[0])); ++i)
      if (s == names[i]) return (EvolverType)i;
    throw ParameterInvalid("parameter '" + k + "': unknown evolver '" + s + "'");
  }
};

class RunParametersInconsistent : public std::runtime_error
{
public:
  using std::runtime_error::runtime_error;
};

template <typename T> class LatticeParameters
{
public:
  LatticeParameters(T dx, T lSide, T kIR, ptrdiff_t N) : dx_(dx), lSide_(lSide), kIR_(kIR), N_(N) {}
  T getDx() const { return dx_; }
  T getLSide() const { return lSide_; }
  T getKIR() const { return kIR_; }
  ptrdiff_t getN() const { return N_; }

private:
  T dx_, lSide_, kIR_;
  ptrdiff_t N_;
};

template <typename T> class RunParameters
{
public:
  explicit RunParameters(ParameterParser &par)
      : N(par.getOverride<ptrdiff_t>("N", Important)),
        kIR(par.getOverride<T>("kIR", -1, Important)),
        lSide(par.getOverride<T>("lSide", -1, Important)),
        dx(par.getOverride<T>("dx", -1, Important)),
        dt(par.get<T>("dt", Important)),
        expansion(par.get<bool>("expansion", true)),
        t0(par.get<T>("t0", 0)),
        tMax(par.get<T>("tMax", 10000 * dt)),
        kCutoff(par.get<T>("kCutOff", -1.0)),
        tOutFreq(par.get<T>("tOutputFreq", 10 * dt)),
        tOutInfreq(par.get<T>("tOutputInfreq", 100 * dt)),
        tOutRareFreq(par.get<T>("tOutputRareFreq", 1000 * dt)),
        outFn(par.get<std::string>("outputfile", "./")),
        baseSeed(par.getSeed("baseSeed")),
        fixedBackground(expansion ? par.get<bool>("fixedBackground", false) : false),
        a0(par.get<T>("a0", 1.0)),
        deltaKBin(par.get<double>("deltaKBin", 1)),
        eType(par.get<EvolverType>("evolver", LF)),
        printHeaders(par.get<bool>("print_headers", false)),
        powerSpectrumType(par.get<int>("PS_type", 1)),
        powerSpectrumVersion(par.get<int>("PS_version", 1)),
        spectraVerbosity(par.get<int>("spectraVerbosity", 0)),
        withGWs(par.get<bool>("withGWs", false)),
        precision(precisionFrom(par.get<std::string>("clmi_precision", "f64"))),
        device(par.get<int>("clmi_device", 0))
  {
    // lattice geometry: any one of kIR, lSide, dx, checked for consistency (runparameters.h:155-186)
    const bool hasKIR = kIR != T(-1), hasLSide = lSide != T(-1), hasDx = dx != T(-1);
    if (!hasKIR && !hasLSide && !hasDx) throw RunParametersInconsistent("You need to specify either kIR, lSide or dx.");
    const T pi = T(3.14159265358979323846);
    if (!hasLSide) lSide = hasKIR ? 2 * pi / kIR : dx * N;
    if (hasKIR && std::fabs(kIR - 2 * pi / lSide) > 1e-10 * std::fabs(kIR))
      throw RunParametersInconsistent("kIR and lSide are not consistent.");
    if (hasDx && std::fabs(dx - lSide / N) > 1e-10 * std::fabs(dx))
      throw RunParametersInconsistent("dx and lSide are not consistent.");
    kIR = 2 * pi / lSide;
    dx = lSide / N;
    if (dt == T(0)) throw RunParametersInconsistent("dt cannot be zero.");
    if (tMax < t0) throw RunParametersInconsistent("tMax must be >= t0.");
    if (tOutFreq < dt) throw RunParametersInconsistent("tOutputFreq must be >= dt.");
    const T kUV = std::sqrt(T(3)) * pi / dx;
    if (kCutoff < 0) kCutoff = 2 * kUV;  // no cutoff
    if (fixedBackground)
      throw RunParametersInconsistent("fixedBackground expansion is not part of this library's lattice path.");
    if (withGWs) throw RunParametersInconsistent("gravitational waves are not part of this library's lattice path.");
    if (powerSpectrumType < 1 || powerSpectrumType > 2)
      throw RunParametersInconsistent("PS_type must be 1 or 2 for the binned spectra of this library.");
    if (powerSpectrumVersion < 1 || powerSpectrumVersion > 3)
      throw RunParametersInconsistent("powerSpectrumVersion " + std::to_string(powerSpectrumVersion) +
                                      " is not a valid powerSpectrumVersion.");
    if (spectraVerbosity < 0 || spectraVerbosity > 2)
      throw RunParametersInconsistent("spectraVerbosity " + std::to_string(spectraVerbosity) +
                                      " is not a valid spectraVerbosity.");
  }

  LatticeParameters<T> getLatParams() const { return LatticeParameters<T>(dx, lSide, kIR, N); }
  void setDoWeRestart(bool r) { doWeRestart = r; }

  ptrdiff_t N;
  T kIR, lSide, dx, dt;
  bool expansion;
  T t0, tMax, kCutoff;
  T tOutFreq, tOutInfreq, tOutRareFreq;
  std::string outFn, baseSeed;
  bool fixedBackground;
  T a0;
  double deltaKBin;
  EvolverType eType;
  bool printHeaders;
  int powerSpectrumType, powerSpectrumVersion;
  int spectraVerbosity;  // columns of the spectra files (runparameters.h:96,233-235): 0, 1 or 2
  bool withGWs;
  int precision;  // CLMI_F64 or CLMI_F32_BULK (key clmi_precision = f64 | f32_bulk)
  int device;     // HIP device ordinal (key clmi_device)
  bool doWeRestart = false;

private:
  static int precisionFrom(const std::string &s)
  {
    if (s == "f64") return CLMI_F64;
    if (s == "f32_bulk") return CLMI_F32_BULK;
    throw ParameterInvalid("clmi_precision must be f64 or f32_bulk, got '" + s + "'");
  }
};

}  // namespace TempLat
