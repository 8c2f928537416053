// Instantiations of the full-row scalar pass, FP32-bulk storage, 2 fields.
#include "scalar_rows.hpp"

namespace clmi {

hipError_t launch_scalar_rows_f32_nf2(bool drift, bool kick, bool red, const ScalarPassArgs &a, const RowsPlan &p, hipStream_t s)
{
#define CLMI_SHAPE(NF_, F32_, VZ_, WPR_, TY_)                                                          \
  if (p.VZ == VZ_ && p.WPR == WPR_ && p.TY == TY_)                                                     \
    switch (p.MM) {                                                                                    \
      case 1: return dispatch_rows<NF_, F32_, VZ_, WPR_, TY_, 1>(drift, kick, red, a, p.blocks, p.XC, p.G, s); \
      case 7: return dispatch_rows<NF_, F32_, VZ_, WPR_, TY_, 7>(drift, kick, red, a, p.blocks, p.XC, p.G, s); \
      case 3: return dispatch_rows<NF_, F32_, VZ_, WPR_, TY_, 3>(drift, kick, red, a, p.blocks, p.XC, p.G, s); \
      case 15: return dispatch_rows<NF_, F32_, VZ_, WPR_, TY_, 15>(drift, kick, red, a, p.blocks, p.XC, p.G, s); \
      default: return hipErrorInvalidValue;                                                            \
    }
  CLMI_SHAPE(2, true, 2, 2, 8)
  CLMI_SHAPE(2, true, 4, 1, 8)
  CLMI_SHAPE(2, true, 2, 1, 8)
  CLMI_SHAPE(2, true, 1, 1, 8)
  CLMI_SHAPE(2, true, 3, 1, 8)
  CLMI_SHAPE(2, true, 2, 4, 4)
  CLMI_SHAPE(2, true, 2, 2, 4)
  CLMI_SHAPE(2, true, 4, 1, 4)
#undef CLMI_SHAPE
  return hipErrorInvalidValue;
}

}  // namespace clmi
