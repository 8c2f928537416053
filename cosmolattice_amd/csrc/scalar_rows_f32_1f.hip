// Instantiations of the full-row scalar pass, FP32-bulk storage, 1 field.
#include "scalar_rows.hpp"

namespace clmi {

hipError_t launch_scalar_rows_f32_nf1(bool drift, bool kick, bool red, const ScalarPassArgs &a, const RowsPlan &p, hipStream_t s,
                                      bool kick2)
{
  // The lphi4 potential has a compile-time form (no per-site scalar loops) for the main shapes.
  const int pk = scalar_rows_pk(1, true, a.poly, p);
#define CLMI_MM(NF_, VZ_, WPR_, TY_, PK_)                                                                \
  switch (p.MM) {                                                                                      \
    case 1: return dispatch_rows<NF_, true, VZ_, WPR_, TY_, 1, PK_>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);   \
    case 7: return dispatch_rows<NF_, true, VZ_, WPR_, TY_, 7, PK_>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);   \
    case 3: return dispatch_rows<NF_, true, VZ_, WPR_, TY_, 3, PK_>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2);   \
    case 15: return dispatch_rows<NF_, true, VZ_, WPR_, TY_, 15, PK_>(drift, kick, red, a, p.blocks, p.XC, p.G, s, kick2); \
    default: return hipErrorInvalidValue;                                                              \
  }
#define CLMI_SHAPE(NF_, VZ_, WPR_, TY_)                                                                  \
  if (p.VZ == VZ_ && p.WPR == WPR_ && p.TY == TY_) CLMI_MM(NF_, VZ_, WPR_, TY_, 0)
#define CLMI_SHAPE_PK(NF_, VZ_, WPR_, TY_)                                                               \
  if (p.VZ == VZ_ && p.WPR == WPR_ && p.TY == TY_) {                                                   \
    if (pk == 1) CLMI_MM(NF_, VZ_, WPR_, TY_, 1)                                                        \
    CLMI_MM(NF_, VZ_, WPR_, TY_, 0)                                                                     \
  }
  CLMI_SHAPE_PK(1, 4, 1, 8)
  CLMI_SHAPE_PK(1, 2, 2, 8)
  CLMI_SHAPE_PK(1, 2, 1, 8)
  CLMI_SHAPE_PK(1, 1, 1, 8)
  CLMI_SHAPE_PK(1, 3, 1, 8)
  CLMI_SHAPE_PK(1, 2, 4, 4)
#undef CLMI_SHAPE_PK
#undef CLMI_MM
#undef CLMI_SHAPE
  return hipErrorInvalidValue;
}

}  // namespace clmi
