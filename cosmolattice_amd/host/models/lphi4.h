// lphi4: inflaton phi with V = lambda phi^4 / 4 and a daughter field chi coupled through
// g^2 phi^2 chi^2 / 2, q = g^2 / lambda.  Written in the model-header API (same content as the
// reference's models/lphi4.h:24-226): program units alpha = 1, fStar = phi_0,
// omegaStar = sqrt(lambda) phi_0, so V = phi^4/4 + q/2 phi^2 chi^2.
#ifndef CLMI_MODEL_LPHI4_H
#define CLMI_MODEL_LPHI4_H

#include "CosmoInterface/cosmointerface.h"

namespace TempLat
{
  struct ModelPars : public TempLat::DefaultModelPars {
    static constexpr size_t NScalars = 2;
    static constexpr size_t NPotTerms = 2;
    using FloatType = double;
  };

#define MODELNAME lphi4
  template <class R> using Model = MakeModel(R, ModelPars);

  class MODELNAME : public Model<MODELNAME>
  {
  private:
    FloatType lambda, q;

  public:
    MODELNAME(ParameterParser &parser, RunParameters<FloatType> &runPar, auto toolBox)
        : Model<MODELNAME>(parser, runPar.getLatParams(), toolBox, runPar.dt, STRINGIFY(MODELLABEL))
    {
      lambda = parser.get<FloatType>("lambda");
      q = parser.get<FloatType>("q");
      fldS0 = parser.get<FloatType, 2>("initial_amplitudes");                 // GeV
      piS0 = parser.get<FloatType, 2>("initial_momenta", std::array<FloatType, 2>{0, 0});  // GeV^2
      alpha = 1;
      fStar = fldS0[0];
      omegaStar = std::sqrt(lambda) * fStar;
      setInitialPotentialAndMassesFromPotential();
    }

    auto potentialTerms(Tag<0>) const { return FloatType(0.25) * pow<4>(fldS(0_c)); }
    auto potentialTerms(Tag<1>) const { return FloatType(0.5) * q * pow<2>(fldS(0_c) * fldS(1_c)); }

    auto potDeriv(Tag<0>) const { return pow<3>(fldS(0_c)) + q * fldS(0_c) * pow<2>(fldS(1_c)); }
    auto potDeriv(Tag<1>) const { return q * pow<2>(fldS(0_c)) * fldS(1_c); }

    auto potDeriv2(Tag<0>) const { return 3 * pow<2>(fldS(0_c)) + q * pow<2>(fldS(1_c)); }
    auto potDeriv2(Tag<1>) const { return q * pow<2>(fldS(0_c)); }
  };
#undef MODELNAME
} // namespace TempLat

#endif
