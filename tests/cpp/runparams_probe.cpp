// Prints the lattice geometry the CosmoInterface mirror's RunParameters derives from a parameter
// file, or the exception it throws:  runparams_probe input=<file>
#include <cstdio>

#include "CosmoInterface/runparameters.h"

int main(int argc, char **argv)
{
  using namespace TempLat;
  try {
    ParameterParser parser(argc, argv);
    RunParameters<double> rp(parser);
    std::printf("lSide %.17g\nkIR %.17g\ndx %.17g\n", rp.lSide, rp.kIR, rp.dx);
  } catch (const RunParametersInconsistent &) {
    std::printf("throws RunParametersInconsistent\n");
  }
  return 0;
}
