// Allocation-free entry point: the statistics of clmi_field_stats on a caller-owned FP64 array,
// for callers that cannot allocate device memory (code under HIP graph capture, serving loops).
//
// The kernels are those of clmi_field_stats (k_means with the gradient term, then the fixed-order
// row sum).  Their partial-sum rows live in `scratch`, a file-scope device array of fixed size:
// kMaxPassBlocks rows of kNumRed doubles (the most k_means ever writes) plus one row for the sums,
// 262 208 bytes.  Each GPU holds its own copy from the library's load until the process ends.  Its
// device address is taken once per device with hipGetSymbolAddress (the array's name is only a host
// stand-in); the kernels receive that address.  Every call on a GPU shares the GPU's copy, so calls
// that overlap on one GPU (different streams, different host threads) race.  Nothing here calls
// hipMalloc, hipFree or an allocating library routine.
#include <mutex>
#include <string>

#include "clmi_internal.hpp"

using namespace clmi;

namespace {

constexpr int kScratchRows = kMaxPassBlocks + 1;
// k_means indexes sites with 32-bit unsigned arithmetic, including its grid stride.
constexpr int kNoAllocMaxN = 1024;

__device__ __align__(256) double scratch[kScratchRows * kNumRed];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};
thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
  g_err = msg;
  return code;
}

// Device address of this device's scratch, looked up on the first call on the device.
int scratch_ptr(int dev, double **out)
{
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(CLMI_E_HIP, std::string("clmi_field_stats_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_ptr[dev] = (double *)p;
  }
  *out = g_ptr[dev];
  return CLMI_OK;
}

// out = {sums[0] / V, sums[1] / V, sums[2] / dx^2 / V}: the host arithmetic of clmi_field_stats.
__global__ void k_stats_finish(const double *sums, double V, double dx, double *out)
{
  if (threadIdx.x == 0) {
    out[0] = sums[0] / V;
    out[1] = sums[1] / V;
    out[2] = sums[2] / (dx * dx) / V;
  }
}

}  // namespace

extern "C" {

const char *clmi_noalloc_last_error(void) { return g_err.c_str(); }

int clmi_field_stats_noalloc(const double *d_field, int N, double dx, double *d_out, void *stream)
{
  if (!d_field || !d_out) return fail(CLMI_E_INVALID, "clmi_field_stats_noalloc: null device pointer");
  if (N < 2 || N > kNoAllocMaxN)
    return fail(CLMI_E_INVALID, "clmi_field_stats_noalloc: N = " + std::to_string(N) + " outside [2, " +
                                    std::to_string(kNoAllocMaxN) + "]");
  if (!(dx > 0)) return fail(CLMI_E_INVALID, "clmi_field_stats_noalloc: dx must be positive");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(CLMI_E_HIP, std::string("clmi_field_stats_noalloc: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(CLMI_E_INVALID, "clmi_field_stats_noalloc: device ordinal too large");
  double *ptr = nullptr;
  const int rc = scratch_ptr(dev, &ptr);
  if (rc != CLMI_OK) return rc;

  const hipStream_t s = (hipStream_t)stream;
  MeansArgs a{};
  a.p = d_field;
  a.lo = nullptr;
  a.V = (uint32_t)N * (uint32_t)N * (uint32_t)N;
  a.f32 = 0;
  a.scale = 1.0;
  a.partials = ptr;
  a.N = (uint32_t)N;
  a.grad = 1;
  const int blocks = scalar_pass_blocks(a.V);
  if (blocks > kMaxPassBlocks) return fail(CLMI_E_INVALID, "clmi_field_stats_noalloc: too many partial rows");
  double *sums = ptr + (size_t)kMaxPassBlocks * kNumRed;
  e = launch_means(a, blocks, s);
  if (e == hipSuccess) e = launch_sum_rows(ptr, blocks, sums, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_stats_finish, dim3(1), dim3(64), 0, s, (const double *)sums, (double)a.V, dx, d_out);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return fail(CLMI_E_HIP, std::string("clmi_field_stats_noalloc: ") + hipGetErrorString(e));
  return CLMI_OK;
}

}  // extern "C"
