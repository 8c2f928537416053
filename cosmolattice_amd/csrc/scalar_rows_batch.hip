// Batched full-row pass: the fused leapfrog pass of several independent lattices of one size class
// (N = 64 VZ) in one launch.  Workgroups [start, start + blocks) of the grid belong to one lattice
// (the member table lives in device memory, so one captured launch serves any number of lattices)
// and run exactly the single-lattice workgroup code, so every lattice's fields and partial sums are
// bit-identical to its own launch.  Each lattice's workgroups start at a multiple of 8, keeping the
// XCD-aware tile order of the single-lattice pass.  One launch per size class keeps each class at
// its own register and LDS footprint (and occupancy).
#include "scalar_rows.hpp"

namespace clmi {
namespace {

constexpr int kBatchThreads = 512;  // 64 * WPR * TY for the shapes (VZ, 1, 8)

template <int NF, bool F32, int VZ, bool D, bool K, bool R, int PK>
__global__ __launch_bounds__(kBatchThreads) void k_scalar_rows_batch(const RowsMember *__restrict__ m, int n)
{
  __shared__ RowsTile<NF, F32, VZ, 1, 8> lds;
  const int b = blockIdx.x;
  int i = 0;
  while (i + 1 < n && m[i + 1].start <= b) ++i;
  // the table is read-only for the whole launch (restrict): its loads stay scalar and are not
  // repeated after the pass's stores
  const RowsMember &mm = m[i];
  const int lb = b - mm.start;
  if (lb >= mm.blocks) return;
  rows_body<NF, F32, VZ, 1, 8, D, K, R, 15, PK>(mm.a, mm.XC, mm.G, lb, mm.blocks, lds);
}

template <int NF, bool F32, int VZ, int PK>
hipError_t launch_batch(bool drift, bool kick, bool red, const RowsMember *m, int n, int total, hipStream_t s)
{
  if (drift && kick && red)
    hipLaunchKernelGGL((k_scalar_rows_batch<NF, F32, VZ, true, true, true, PK>), dim3(total), dim3(kBatchThreads), 0,
                       s, m, n);
  else if (drift && kick && !red)
    hipLaunchKernelGGL((k_scalar_rows_batch<NF, F32, VZ, true, true, false, PK>), dim3(total), dim3(kBatchThreads), 0,
                       s, m, n);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

template <int NF, bool F32, int PK>
hipError_t launch_batch_vz(int vz, bool drift, bool kick, bool red, const RowsMember *m, int n, int total,
                           hipStream_t s)
{
  switch (vz) {
    case 1: return launch_batch<NF, F32, 1, PK>(drift, kick, red, m, n, total, s);
    case 2: return launch_batch<NF, F32, 2, PK>(drift, kick, red, m, n, total, s);
    case 3: return launch_batch<NF, F32, 3, PK>(drift, kick, red, m, n, total, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t launch_scalar_rows_batch(int nf, bool f32, int pk, int vz, bool drift, bool kick, bool red,
                                    const RowsMember *d_members, int n, int total_blocks, hipStream_t s)
{
  if (nf == 1) {
    if (!f32)
      return pk == 1 ? launch_batch_vz<1, false, 1>(vz, drift, kick, red, d_members, n, total_blocks, s)
                     : launch_batch_vz<1, false, 0>(vz, drift, kick, red, d_members, n, total_blocks, s);
    return pk == 1 ? launch_batch_vz<1, true, 1>(vz, drift, kick, red, d_members, n, total_blocks, s)
                   : launch_batch_vz<1, true, 0>(vz, drift, kick, red, d_members, n, total_blocks, s);
  }
  if (nf == 2) {
    if (!f32)
      return pk == 1 ? launch_batch_vz<2, false, 1>(vz, drift, kick, red, d_members, n, total_blocks, s)
                     : launch_batch_vz<2, false, 0>(vz, drift, kick, red, d_members, n, total_blocks, s);
    return pk == 1 ? launch_batch_vz<2, true, 1>(vz, drift, kick, red, d_members, n, total_blocks, s)
                   : launch_batch_vz<2, true, 0>(vz, drift, kick, red, d_members, n, total_blocks, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace clmi
