// Shapes and launch of the full-row scalar pass (kernel in scalar_rows.hpp).
#include "clmi_internal.hpp"
#include "scalar_poly.hpp"

namespace clmi {

hipError_t launch_scalar_rows_f32_nf2(bool drift, bool kick, bool red, const ScalarPassArgs &a, const RowsPlan &p,
                                      hipStream_t s, bool kick2);
hipError_t launch_scalar_rows_f32_nf1(bool drift, bool kick, bool red, const ScalarPassArgs &a, const RowsPlan &p,
                                      hipStream_t s, bool kick2);
hipError_t launch_scalar_rows_f64_nf2(bool drift, bool kick, bool red, const ScalarPassArgs &a, const RowsPlan &p,
                                      hipStream_t s, bool kick2);
hipError_t launch_scalar_rows_f64_nf1(bool drift, bool kick, bool red, const ScalarPassArgs &a, const RowsPlan &p,
                                      hipStream_t s, bool kick2);

// Row-kernel shapes by (N, precision): VZ sites per lane, WPR wavefronts per row, TY rows per
// workgroup.  The shape fixes the summation order; the x-chunk XC and store mode are tuned at run
// time without changing results.
bool scalar_rows_plan(int N, int nf, bool f32, int XC, int mm, int shape, RowsPlan *p)
{
  if (nf != 2 && nf != 1) return false;
  int vz = 0, wpr = 0, ty = 0;
  const bool alt = (shape & 1) != 0, ty4 = (shape & 2) != 0 && nf == 2;
  if (N == 256 && !alt) { vz = 2; wpr = 2; ty = ty4 ? 4 : 8; }
  else if (N == 256) { vz = 4; wpr = 1; ty = ty4 ? 4 : 8; }
  else if (N == 128) { vz = 2; wpr = 1; ty = 8; }
  else if (N == 64) { vz = 1; wpr = 1; ty = 8; }
  else if (N == 192) { vz = 3; wpr = 1; ty = 8; }
  // 1024 threads; 8-row tiles would need 2048.  FP64 at 512^3 runs faster in the tiled pass.
  else if (N == 512 && f32) { vz = 2; wpr = 4; ty = 4; }
  else return false;
  const int G = 8;
  if (XC < G || XC % G != 0 || N % XC != 0) return false;
  if (mm != 1 && mm != 3 && mm != 7 && mm != 15) return false;
  p->VZ = vz;
  p->WPR = wpr;
  p->TY = ty;
  p->XC = XC;
  p->G = G;
  p->MM = mm;
  p->blocks = (N / ty) * (N / XC);
  p->rows = (N / ty) * (N / G);
  return true;
}

int scalar_rows_pk(int nf, bool f32, const ScalarPoly &P, const RowsPlan &p)
{
  if (poly_kind(P, nf) != 1) return 0;
  const int v = p.VZ, w = p.WPR, t = p.TY;
  if (!f32) return (w == 1 && t == 8 && v <= 3) ? 1 : 0;  // FP64: the 512-thread shapes
  if (nf == 1) return 1;  // every 1-field shape
  const bool inst = (v == 2 && w == 2 && t == 8) || (v == 2 && w == 4 && t == 4) || (v == 2 && w == 1 && t == 8) ||
                    (v == 1 && w == 1 && t == 8) || (v == 3 && w == 1 && t == 8);
  return inst ? 1 : 0;
}

hipError_t launch_scalar_rows(int nf, bool f32, bool drift, bool kick, bool red, const ScalarPassArgs &a,
                              const RowsPlan &p, hipStream_t s, bool kick2)
{
  if (nf == 2)
    return f32 ? launch_scalar_rows_f32_nf2(drift, kick, red, a, p, s, kick2)
               : launch_scalar_rows_f64_nf2(drift, kick, red, a, p, s, kick2);
  if (nf == 1)
    return f32 ? launch_scalar_rows_f32_nf1(drift, kick, red, a, p, s, kick2)
               : launch_scalar_rows_f64_nf1(drift, kick, red, a, p, s, kick2);
  return hipErrorInvalidValue;
}

}  // namespace clmi
