// Prints what the CosmoInterface mirror's CouplingsManager gives for the coupling pattern of the
// reference's couplingsmanager test (2 matter fields x 4 gauge fields):
//   couplings_kat <charges,comma,separated> <couplings,comma,separated>
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>
#include <vector>

#include "CosmoInterface/couplingsmanager.h"

using namespace TempLat;
using CM = CouplingsManager<2, 4, true, false, true, false, true, true, true, false>;

static std::vector<double> parse(const char *s)
{
  std::vector<double> v;
  std::stringstream ss(s);
  std::string item;
  while (std::getline(ss, item, ','))
    if (!item.empty()) v.push_back(std::atof(item.c_str()));
  return v;
}

template <int M, int G> static void pair(const CM::Container<double> &cm)
{
  std::printf("couples %d %d %d\n", M, G, CM::couples(Tag<M>{}, Tag<G>{}) ? 1 : 0);
  std::printf("charge %d %d %.17g\n", M, G, cm(Tag<M>{}, Tag<G>{}));
}

int main(int argc, char **argv)
{
  if (argc != 3) return 2;
  std::printf("how_many_couples %zu\n", CM::howManyCouples());
  CM::Container<double> cm;
  try {
    cm.setEffectiveCharges(parse(argv[1]), parse(argv[2]));
  } catch (const NotEnoughChargesForThisCouplingsManager &) {
    std::printf("throws NotEnoughChargesForThisCouplingsManager\n");
    return 0;
  } catch (const NotEnoughCouplingsForThisCouplingsManager &) {
    std::printf("throws NotEnoughCouplingsForThisCouplingsManager\n");
    return 0;
  }
  pair<0, 0>(cm); pair<0, 1>(cm); pair<0, 2>(cm); pair<0, 3>(cm);
  pair<1, 0>(cm); pair<1, 1>(cm); pair<1, 2>(cm); pair<1, 3>(cm);
  std::printf("coupling 0 %.17g\ncoupling 1 %.17g\ncoupling 2 %.17g\ncoupling 3 %.17g\n", cm.coupling(Tag<0>{}),
              cm.coupling(Tag<1>{}), cm.coupling(Tag<2>{}), cm.coupling(Tag<3>{}));
  return 0;
}
