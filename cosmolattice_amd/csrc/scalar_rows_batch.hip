// Batched full-row pass: the fused leapfrog pass of several independent lattices of N = 64, 128
// or 192 in one launch.  Workgroups [start, start + blocks) of the grid belong to one lattice (the
// member table lives in device memory, so one captured launch serves any number of lattices) and
// run exactly the single-lattice workgroup code, so every lattice's fields and partial sums are
// bit-identical to its own launch.  Each lattice's workgroups start at a multiple of 8, keeping the
// XCD-aware tile order of the single-lattice pass.
#include "scalar_rows.hpp"

namespace clmi {
namespace {

constexpr int kBatchThreads = 512;  // 64 * WPR * TY for the shapes (VZ, 1, 8)

template <int NF, bool F32, bool D, bool K, bool R>
__global__ __launch_bounds__(kBatchThreads) void k_scalar_rows_batch(const RowsMember *m, int n)
{
  using V = typename Tr<F32>::V;
  __shared__ V lds[2 * NF * 10 * 192];  // the largest tile (VZ = 3); smaller shapes use its prefix
  const int b = blockIdx.x;
  int i = 0;
  while (i + 1 < n && m[i + 1].start <= b) ++i;
  const RowsMember &mm = m[i];
  const int lb = b - mm.start;
  if (lb >= mm.blocks) return;
  switch (mm.vz) {
    case 1:
      rows_body<NF, F32, 1, 1, 8, D, K, R, true>(mm.a, mm.XC, mm.G, lb, mm.blocks,
                                                 *reinterpret_cast<RowsTile<NF, F32, 1, 1, 8> *>(lds));
      break;
    case 2:
      rows_body<NF, F32, 2, 1, 8, D, K, R, true>(mm.a, mm.XC, mm.G, lb, mm.blocks,
                                                 *reinterpret_cast<RowsTile<NF, F32, 2, 1, 8> *>(lds));
      break;
    case 3:
      rows_body<NF, F32, 3, 1, 8, D, K, R, true>(mm.a, mm.XC, mm.G, lb, mm.blocks,
                                                 *reinterpret_cast<RowsTile<NF, F32, 3, 1, 8> *>(lds));
      break;
    default:
      break;
  }
}

template <int NF, bool F32>
hipError_t launch_batch(bool drift, bool kick, bool red, const RowsMember *m, int n, int total, hipStream_t s)
{
  if (drift && kick && red)
    hipLaunchKernelGGL((k_scalar_rows_batch<NF, F32, true, true, true>), dim3(total), dim3(kBatchThreads), 0, s, m, n);
  else if (drift && kick && !red)
    hipLaunchKernelGGL((k_scalar_rows_batch<NF, F32, true, true, false>), dim3(total), dim3(kBatchThreads), 0, s, m,
                       n);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace

hipError_t launch_scalar_rows_batch(int nf, bool f32, bool drift, bool kick, bool red, const RowsMember *d_members,
                                    int n, int total_blocks, hipStream_t s)
{
  if (nf == 1) return f32 ? launch_batch<1, true>(drift, kick, red, d_members, n, total_blocks, s)
                          : launch_batch<1, false>(drift, kick, red, d_members, n, total_blocks, s);
  if (nf == 2) return f32 ? launch_batch<2, true>(drift, kick, red, d_members, n, total_blocks, s)
                          : launch_batch<2, false>(drift, kick, red, d_members, n, total_blocks, s);
  return hipErrorInvalidValue;
}

}  // namespace clmi
