// lphi4U1: a complex scalar with V = lambda |phi|^4 charged under a U(1) gauge field.  Written in the
// model-header API (same content as the reference's models/lphi4U1.h:20-137): program units
// alpha = 1, fStar = |phi|_0, omegaStar = sqrt(lambda) |phi|_0, so V = |phi|^4.  The gauge coupling
// comes from the parameters gU1s and CSU1_charges.
#ifndef CLMI_MODEL_LPHI4U1_H
#define CLMI_MODEL_LPHI4U1_H

#include "CosmoInterface/cosmointerface.h"

namespace TempLat
{
  struct ModelParsLphi4U1 : public TempLat::DefaultModelPars {
    static constexpr size_t NScalars = 0;
    static constexpr size_t NCScalars = 1;
    static constexpr size_t NU1Flds = 1;
    static constexpr size_t NPotTerms = 1;
    using FloatType = double;
    typedef TempLat::CouplingsManager<NCScalars, NU1Flds, true> CsU1Couplings;
  };

  class lphi4U1 : public MakeModel(lphi4U1, ModelParsLphi4U1)
  {
  private:
    FloatType lambda;

  public:
    lphi4U1(ParameterParser &parser, RunParameters<FloatType> &runPar,
            device::memory::host_ptr<MemoryToolBox<3>> toolBox)
        : AbstractModel(parser, runPar.getLatParams(), toolBox, runPar.dt, "lphi4U1")
    {
      const FloatType norm0 = parser.get<FloatType>("cmplx_field_initial_norm");
      const FloatType normPi0 = parser.get<FloatType>("cmplx_momentum_initial_norm");
      const FloatType r2 = std::sqrt(FloatType(2));
      fldCS0(0_c) = Complexify(norm0 / r2, norm0 / r2);
      piCS0(0_c) = Complexify(normPi0 / r2, normPi0 / r2);
      lambda = parser.get<FloatType>("lambda");
      alpha = 1;
      fStar = norm0;
      omegaStar = std::sqrt(lambda) * norm0;
      setInitialPotentialAndMassesFromPotential();
    }

    auto potentialTerms(Tag<0>) { return pow<4>(norm(fldCS(0_c))); }
    auto potDerivNormCS(Tag<0>) { return 4 * pow<3>(norm(fldCS(0_c))); }
    auto potDeriv2NormCS(Tag<0>) { return 12 * pow<2>(norm(fldCS(0_c))); }
  };
} // namespace TempLat

#endif
