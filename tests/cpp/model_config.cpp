// Prints the library configuration a model header produces (no GPU needed):
//   model_config input=<parameter file> [key=value ...]
// Build with -DCLMI_MODEL_HEADER='"<header>"' -DCLMI_MODEL_CLASS=<class>.
#include <cstdio>
#include <exception>

#include CLMI_MODEL_HEADER

int main(int argc, char **argv)
{
  try {
    using namespace TempLat;
    ParameterParser parser(argc, argv);
    using M = CLMI_MODEL_CLASS;
    RunParameters<M::FloatType> rp(parser);
    auto tb = MemoryToolBox<3>::makeShared(rp.N, 1);
    M model(parser, rp, tb);
    const clmi_config c = model.deviceConfig(rp);
    std::printf("N %d\nn_scalars %d\nn_cscalars %d\nn_u1 %d\n", c.N, c.n_scalars, c.n_cscalars, c.n_u1);
    std::printf("dx %.17g\ndt %.17g\nalpha %.17g\nfStar %.17g\nomegaStar %.17g\nMPl %.17g\ngQ %.17g\n", c.dx, c.dt,
                c.alpha, c.fStar, c.omegaStar, c.MPl, c.gQ);
    for (int t = 0; t < c.n_terms; ++t)
      std::printf("term %.17g %d %d %d %d\n", c.term_coeff[t], c.term_exp[t][0], c.term_exp[t][1], c.term_exp[t][2],
                  c.term_exp[t][3]);
    for (int k = 0; k < c.n_cs_terms; ++k) std::printf("cs_term %.17g %d\n", c.cs_coeff[k], c.cs_pow[k]);
    std::printf("pot0 %.17g\n", model.pot0);
    for (size_t j = 0; j < M::Ns; ++j) std::printf("mass2 %.17g\n", model.masses2S((int)j));
    if constexpr (M::NCs > 0) std::printf("mass2cs %.17g\n", model.masses2CS(0).re);
    for (size_t j = 0; j < M::Ns; ++j) std::printf("fld0 %.17g\npi0 %.17g\n", model.fldS0((int)j), model.piS0((int)j));
    return 0;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
