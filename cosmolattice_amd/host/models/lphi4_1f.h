// lphi4 with the inflaton alone: V = lambda phi^4 / 4, program units alpha = 1, fStar = phi_0,
// omegaStar = sqrt(lambda) phi_0, so V = phi^4/4.  The single-field variant of the reference's
// models/lphi4.h, written in the same model-header API (NScalars = 1, NPotTerms = 1).
#ifndef CLMI_MODEL_LPHI4_1F_H
#define CLMI_MODEL_LPHI4_1F_H

#include "CosmoInterface/cosmointerface.h"

namespace TempLat
{
  struct ModelParsLphi4Single : public TempLat::DefaultModelPars {
    static constexpr size_t NScalars = 1;
    static constexpr size_t NPotTerms = 1;
    using FloatType = double;
  };

  class lphi4_1f : public MakeModel(lphi4_1f, ModelParsLphi4Single)
  {
  private:
    FloatType lambda;

  public:
    lphi4_1f(ParameterParser &parser, RunParameters<FloatType> &runPar, auto toolBox)
        : AbstractModel(parser, runPar.getLatParams(), toolBox, runPar.dt, "lphi4_1f")
    {
      lambda = parser.get<FloatType>("lambda");
      fldS0 = parser.get<FloatType, 1>("initial_amplitudes");
      piS0 = parser.get<FloatType, 1>("initial_momenta", std::array<FloatType, 1>{0});
      alpha = 1;
      fStar = fldS0[0];
      omegaStar = std::sqrt(lambda) * fStar;
      setInitialPotentialAndMassesFromPotential();
    }

    auto potentialTerms(Tag<0>) const { return FloatType(0.25) * pow<4>(fldS(0_c)); }
    auto potDeriv(Tag<0>) const { return pow<3>(fldS(0_c)); }
    auto potDeriv2(Tag<0>) const { return 3 * pow<2>(fldS(0_c)); }
  };
} // namespace TempLat

#endif
