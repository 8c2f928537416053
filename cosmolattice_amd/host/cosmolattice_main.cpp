// Main program of a model, as the reference's source/cosmolattice.cpp: build with
//   -DCLMI_MODEL_HEADER='"models/lphi4.h"' -DCLMI_MODEL_CLASS=lphi4
// and run as `<exe> input=<parameter file> [key=value ...]`.
#include <cstdio>
#include <exception>

#include CLMI_MODEL_HEADER

int main(int argc, char **argv)
{
  try {
    TempLat::ParameterParser parser(argc, argv);
    return TempLat::runCosmoLattice<TempLat::CLMI_MODEL_CLASS>(parser);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
