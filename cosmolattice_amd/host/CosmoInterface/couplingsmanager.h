// Which matter fields couple to which gauge fields, and the effective charges g*Q, with the
// reference's interface (couplingsmanager.h:17-200): CouplingsManager<NMatter, NGauge, pattern...>
// holds the coupling pattern (row-major over (matter, gauge)) at compile time; its Container<T> holds
// the gauge couplings g and the effective charges.  setEffectiveCharges takes one charge Q per coupled
// pair, in pattern order, and one coupling g per gauge field; an uncoupled pair has charge 0.
#pragma once

#include <array>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "CosmoInterface/symbolic.h"

namespace TempLat {

class NotEnoughChargesForThisCouplingsManager : public std::runtime_error
{
public:
  using std::runtime_error::runtime_error;
};
class NotEnoughCouplingsForThisCouplingsManager : public std::runtime_error
{
public:
  using std::runtime_error::runtime_error;
};

template <typename T, int NMatter, int NGauge, bool... Pattern> class CouplingsManagerContainer
{
  static constexpr size_t kPairs = (size_t)NMatter * (size_t)NGauge;
  static_assert(sizeof...(Pattern) == 0 || sizeof...(Pattern) == kPairs,
                "one coupling flag per (matter, gauge) pair, row-major");

public:
  static constexpr int nGauge = NGauge;
  static constexpr bool active = (false || ... || Pattern);

  static constexpr size_t howManyCouples() { return (size_t(0) + ... + (size_t)Pattern); }

  template <int m, int g> static constexpr bool couples(Tag<m>, Tag<g>)
  {
    static_assert(m >= 0 && m < NMatter && g >= 0 && g < NGauge, "field index out of range");
    if constexpr (sizeof...(Pattern) == 0) {
      return false;
    } else {
      constexpr bool pat[] = {Pattern...};
      return pat[m * NGauge + g];
    }
  }

  template <typename U> void setEffectiveCharges(const std::vector<U> &charges, const std::vector<U> &couplings)
  {
    if (charges.size() != howManyCouples() && !couplings.empty())
      throw NotEnoughChargesForThisCouplingsManager(std::to_string(charges.size()) + " charges given, " +
                                                    std::to_string(howManyCouples()) + " coupled pairs expected");
    if (couplings.size() != (size_t)NGauge && !charges.empty())
      throw NotEnoughCouplingsForThisCouplingsManager(std::to_string(couplings.size()) + " couplings given, " +
                                                      std::to_string(NGauge) + " gauge fields present");
    for (size_t g = 0; g < (size_t)NGauge && g < couplings.size(); ++g) gs_[g] = T(couplings[g]);
    if constexpr (sizeof...(Pattern) > 0) {
      constexpr bool pat[] = {Pattern...};
      size_t next = 0;  // the next unused charge
      for (size_t k = 0; k < kPairs; ++k) {
        if (pat[k] && next < charges.size() && !couplings.empty())
          eff_[k] = T(charges[next++]) * T(couplings[k % NGauge]);
        else
          eff_[k] = T(0);
      }
    }
  }

  // g*Q of matter field m under gauge field g (0 for an uncoupled pair)
  template <int m, int g> T operator()(Tag<m>, Tag<g>) const
  {
    static_assert(m >= 0 && m < NMatter && g >= 0 && g < NGauge, "field index out of range");
    return eff_[m * NGauge + g];
  }
  // the coupling g of gauge field g
  template <int g> T coupling(Tag<g>) const
  {
    static_assert(g >= 0 && g < NGauge, "gauge index out of range");
    return gs_[g];
  }

private:
  std::array<T, kPairs> eff_{};
  std::array<T, (size_t)NGauge> gs_{};
};

template <int NMatter, int NGauge, bool... Pattern> struct CouplingsManager {
  template <typename T> using Container = CouplingsManagerContainer<T, NMatter, NGauge, Pattern...>;
  static constexpr int nGauge = NGauge;
  static constexpr bool active = (false || ... || Pattern);
  static constexpr size_t howManyCouples() { return (size_t(0) + ... + (size_t)Pattern); }
  template <int m, int g> static constexpr bool couples(Tag<m> a, Tag<g> b)
  {
    return Container<double>::couples(a, b);
  }
};

}  // namespace TempLat
