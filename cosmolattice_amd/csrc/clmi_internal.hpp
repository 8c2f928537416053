// Internal declarations shared by the HIP translation units of libclmi.so.
//
// Design: every evolved lattice array lives in HBM for the lifetime of a context.  A leapfrog step
// is ONE lattice pass (the deferred drift of the previous step fused with this step's kick and all
// volume reductions), followed by a one-block "background update" kernel that finishes the
// reductions in a fixed order and advances the scale factor on the device.  No host round trip
// happens inside the time loop.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "clmi.h"

namespace clmi {

constexpr int kBlock = 256;        // threads per block of the lattice passes (4 wavefronts)
constexpr int kNumRed = 8;         // reduction slots per block row
constexpr int kMaxPassBlocks = 4096;

// Reduction slots (volume sums; the update kernel divides by N^3).
enum RedSlot : int {
  RED_PI2S = 0,
  RED_GRAD2S = 1,
  RED_POT = 2,
  RED_PI2CS = 3,
  RED_GRAD2CS = 4,
  RED_PI2U1 = 5,
  RED_B2U1 = 6,
  RED_PI2S_KICK2 = 7,  // scalar momenta after the second kick of a double-kick pass
};

// Device-resident homogeneous state plus the per-pass coefficients derived from it.
struct DevState {
  clmi_background bg;
  double K;         // last ScaleFactorKernels::get value
  double aK;        // aI^(1+alpha)   -- Laplacian term of the scalar / complex-scalar kick
  double aV;        // aI^(3+alpha)   -- potential term of the scalar / complex-scalar kick
  double aG;        // aI^(alpha-1)   -- gradient term of the U(1) kick
  double aJ;        // aI^(1+alpha)   -- current term of the U(1) kick
  double fdrift_s;  // aSI^(alpha-3)  -- scalar / complex-scalar drift factor
  double fdrift_u;  // aSI^(alpha-1)  -- U(1) drift factor
  double red[kNumRed];  // last finished volume averages
  double pad[3];
};

// Polynomial potential of the scalar singlets in program units.
struct ScalarPoly {
  int nterms;
  int pad;
  double c[CLMI_MAX_TERMS];
  int e[CLMI_MAX_TERMS][CLMI_MAX_SCALARS];
};

// Complex-scalar potential V = sum_k c_k |phi|^(2 h_k).
struct CsPoly {
  int nterms;
  int pad;
  double c[CLMI_MAX_TERMS];
  int h[CLMI_MAX_TERMS];  // half powers
};

// Model constants needed by the background update.
struct ModelConsts {
  int ns, ncs, nu1, expansion;
  double dt, alpha, fStar, omegaStar, MPl;
  double invV;
};

// Operations of the background update kernel, executed in order by one thread.
enum UpdateOp : int {
  OP_STORE_FIELDS = 1,   // storeFieldsAverages        (leapfrog.h:269-277, velocityverlet.h:258-266)
  OP_LF_MOMENTA = 2,     // storeMomentaAverages + kickScaleFactor + driftScaleFactor (leapfrog.h:67-78)
  OP_LF_SYNC = 3,        // sync's averages + aDotI    (leapfrog.h:105-112)
  OP_LF_INIT = 4,        // aDotSI = aDotI - dt/2 K    (leapfrog.h:35)
  OP_VV_SF_HALF = 5,     // kickScaleFactorHalf        (velocityverlet.h:165-168)
  OP_VV_MOMENTA = 6,     // storeMomentaAverages       (velocityverlet.h:249-256)
  OP_VV_DRIFT_SF = 7,    // driftScaleFactor           (velocityverlet.h:225-243)
  OP_VV_SF_ONE = 8,      // kickScaleFactorOne         (velocityverlet.h:171-174)
  OP_SET_ALL = 9,        // Averages::setAllAverages   (averages.h:86-129)
  OP_KERNEL_ONLY = 10,   // recompute K only
  OP_PV_SF_FIRST = 11,   // driftScaleFactorFirstHalf  (positionverlet.h:196-209)
  OP_PV_MID = 12,        // storeFieldsAveragesHalf + storeMomentaAverages + kickScaleFactor +
                         // driftScaleFactorSecondHalf (positionverlet.h:73-94)
  OP_PV_DRIFT_AI = 13,   // drift coefficients from aI (positionverlet.h:161-192)
  OP_VV_MOMENTA_KICK2 = 14,  // storeMomentaAverages of the second kick of a double-kick pass
};

struct UpdateOps {
  int n;
  int op[6];
  int pad;
  double w[6];
  clmi_background *snap;  // non-null: the background after op 0 is also written here
};

// One real lattice array in either precision.
struct DevArr {
  void *p = nullptr;   // double* (F64) or float* hi (F32 bulk)
  float *lo = nullptr; // F32 bulk fields: low word; nullptr for momenta / F64
};

struct ScalarPassArgs {
  const void *phi_in[CLMI_MAX_SCALARS];
  const float *philo_in[CLMI_MAX_SCALARS];
  const void *pi_in[CLMI_MAX_SCALARS];
  void *phi_out[CLMI_MAX_SCALARS];
  float *philo_out[CLMI_MAX_SCALARS];
  void *pi_out[CLMI_MAX_SCALARS];
  int N;
  uint32_t V;
  double inv_dx2;
  double inv_dx;
  double cdt;    // drift: phi += fdrift_s * (cdt * pi)
  double ckick;  // kick: pi += ckick * (aK * lap - aV * dV)
  double ckick2; // double-kick pass: then pi += ckick2 * (aK * lap - aV * dV) with the same fields
  const DevState *st;
  double *partials;  // this pass's rows, kNumRed doubles each
  ScalarPoly poly;
};

struct GaugePassArgs {
  const void *phi_in[2];
  const float *philo_in[2];
  const void *pi_in[2];
  const void *A_in[3];
  const float *Alo_in[3];
  const void *piA_in[3];
  void *phi_out[2];
  float *philo_out[2];
  void *pi_out[2];
  void *A_out[3];
  float *Alo_out[3];
  void *piA_out[3];
  int N;
  uint32_t V;
  int has_u1;
  int pad;
  double dx, inv_dx, inv_dx2;
  double gQ;
  double jnorm;  // 2/dx * (fStar/omegaStar)^2
  double cdt;
  double ckick;
  const DevState *st;
  double *partials;
  CsPoly poly;
};

// Deterministic block reduction: wavefront shuffles, then a fixed-order sum over the block's
// wavefronts; thread k < K writes slot k of the block's row.
// Sum of v over a block of NTB threads, stored by threads 0..K-1 into row[0..K).
template <int K, int NTB>
__device__ __forceinline__ void block_reduce_store_nt(double (&v)[K], double *row)
{
  __shared__ double sh[NTB / 64][K];
#pragma unroll
  for (int k = 0; k < K; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) sh[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x < K) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NTB / 64; ++w) s += sh[w][threadIdx.x];
    row[threadIdx.x] = s;
  }
}

template <int K>
__device__ __forceinline__ void block_reduce_store(double (&v)[K], double *row)
{
  block_reduce_store_nt<K, kBlock>(v, row);
}

// ---- tiled charged scalar + U(1) pass (gauge_tiled.hip) ----
bool gauge_tiled_shape(int N, int ty, int XC, int *blocks, int *rows);
bool gauge_tiled_plan(int N, int ty, int ncu, int *XC, int *blocks, int *rows);
hipError_t launch_gauge_tiled(bool f32, int ty, bool drift, bool kick, bool red, const GaugePassArgs &a, int blocks,
                              int XC, hipStream_t s);

// ---- ensembles: several lattices in one launch (scalar_rows_batch.hip, background.hip) ----
struct RowsMember {   // one lattice's fused pass inside a batched launch (device-resident table)
  ScalarPassArgs a;
  int XC, G;
  int blocks;         // workgroups of this lattice
  int start;          // first workgroup index of this lattice in the batched grid (multiple of 8)
  int vz;             // row shape: VZ = N / 64 (WPR = 1, TY = 8)
  int pad;
};
// pk: potential form of every member (1: compile-time lphi4 table; see scalar_rows_pk).
hipError_t launch_scalar_rows_batch(int nf, bool f32, int pk, int vz, bool drift, bool kick, bool red,
                                    const RowsMember *d_members, int n, int total_blocks, hipStream_t s);
struct UpdMember {    // one lattice's background update inside a batched launch
  DevState *st;
  const double *partials;
  int rows;
  int pad;
  ModelConsts mc;
};
hipError_t launch_update_batch(const UpdMember *d_members, int n, const UpdateOps &ops, hipStream_t s);

// ---- adjoint (adjoint_kernels.hip) ----
constexpr int kAdjRed = 32;  // partial-sum slots of the adjoint reductions
constexpr int kAdjMaxRows = 1024;  // partial rows (one per workgroup) of an adjoint lattice kernel
constexpr int kAdjFieldSlots = 16;  // first slot of the fields-transpose sums in the fused reverse pass
struct AdjFields {           // forward fields of one leapfrog state, in storage precision
  const void *phi[CLMI_MAX_SCALARS];
  const float *philo[CLMI_MAX_SCALARS];
  const void *pi[CLMI_MAX_SCALARS];
};
// Adjoint of the background chain, resident on the device during the reverse sweep of
// clmi_adjoint_lf: the adjoints of the current background scalars, those of the step being reversed
// (committed once its kick transpose has run), the coefficients the next lattice kernels read, and
// the accumulated derivatives by the potential coefficients.
struct AdjScalars {
  double lam_aI, lam_aDotSI, lam_pi2SI, lam_G, lam_U;
  double n_aI, n_aDotSI, n_pi2SI, n_G, n_U;
  double cF, lamP;    // pi transpose: lam_pi += cF lam_phi + lamP (2/N^3) pi
  double ck, aK, aV;  // kick transpose
  double fG, fU;      // fields transpose of the kicked state (fused with the kick transpose)
  double gc[CLMI_MAX_TERMS];
};
struct AdjArgs {
  AdjFields fwd;
  AdjFields fwd_red;  // momenta for the fields transpose's sum lam_phi . pi (unset, pi[0] null: fwd's)
  double *lam_phi[CLMI_MAX_SCALARS];  // FP64 adjoint fields
  double *lam_pi[CLMI_MAX_SCALARS];
  int nf;
  uint32_t N, V;
  double inv_dx2;
  ScalarPoly poly;
  double *partials;  // kAdjMaxRows rows of kAdjRed
  double *red_out;   // kAdjRed finished sums
  const AdjScalars *sc;  // non-null: the kernels take their coefficients from here, not from arguments
  int no_tiled;      // 1: the kick transpose runs one site per thread even where the tiled form applies
};

// ScaleFactorKernels::get for scalar singlets (evolvers/kernels/scalefactorkernels.h:27-81) and its
// partial derivatives by <pi^2>, <|grad phi|^2>, <V> and a.
struct SFd {
  double K, dpi2, dG, dU, da;
};
// a^e: integer exponents (|e| <= 32, the common case alpha = 1) by binary powering, which is a
// few multiplications instead of a serial log/exp evaluation on a single GPU thread; anything else by
// pow().  Agrees with pow() to a few ulp.
__host__ __device__ inline double rpow(double a, double e)
{
  if (e == rint(e) && fabs(e) <= 32.0) {
    int n = (int)fabs(e);
    double p = 1.0, q = a;
    while (n) {
      if (n & 1) p *= q;
      q *= q;
      n >>= 1;
    }
    return e < 0 ? 1.0 / p : p;
  }
  return pow(a, e);
}

__host__ __device__ inline SFd sf_partials(const ModelConsts &mc, double a, double pi2, double G, double U)
{
  const double al = mc.alpha, fm = mc.fStar / mc.MPl;
  const double A = rpow(a, 2 * al + 1) / 3.0 * fm * fm;
  const double B = (al - 2) * 0.5 * pi2 * rpow(a, -6.0) + al * 0.5 * G * rpow(a, -2.0) + (al + 1) * U;
  SFd d;
  d.K = A * B;
  d.dpi2 = A * (al - 2) * 0.5 * rpow(a, -6.0);
  d.dG = A * al * 0.5 * rpow(a, -2.0);
  d.dU = A * (al + 1);
  d.da = (2 * al + 1) / a * A * B + A * ((al - 2) * 0.5 * pi2 * (-6.0) * rpow(a, -7.0) + al * 0.5 * G * (-2.0) * rpow(a, -3.0));
  return d;
}
// Scalar steps of the reverse sweep, one thread, on the stream after the lattice kernel whose sums
// (red) they consume.  chain_fields: after the fields transpose of state n+1 -- derivatives by the
// potential coefficients, the background adjoint of the drift and scale-factor update of step n, and
// the coefficients of the pi and kick transposes of step n (w: the kick weight, V: N^3; the fields
// sums sit at red[fo], red[fo + 1 + t]).  chain_kick: after the kick transpose of step n -- its a_I
// and coefficient terms, then the step's background adjoints become the current ones.
hipError_t launch_adj_chain_fields(AdjScalars *S, const double *red, int fo, const clmi_background *tape, int n,
                                   double w, double V, const ModelConsts &mc, int nterms, bool ex, hipStream_t s);
// After launch_adj_kick_fields of step n: the fixed-order sum of its partial rows into a.red_out, then
// chain_kick of step n and, for n > 0, chain_fields of step n - 1 (weight w) on the fields sums, on the
// background adjoint S (the a.sc the lattice kernels read), in one launch.
hipError_t launch_adj_finish_chains(const AdjArgs &a, AdjScalars *S, const clmi_background *tape, int n, double w,
                                    double V, const ModelConsts &mc, int nterms, bool ex, hipStream_t s);
hipError_t launch_adj_fields(const AdjArgs &a, bool f32, double lamG, double lamU, double lamP, bool with_pi,
                             hipStream_t s);
// Plan of the tiled kick transpose for a.N, a.nf, a.poly and a.no_tiled: x-chunk and workgroups;
// false where the one-site-per-thread kernel runs instead.
bool adj_kick_tiled_plan(const AdjArgs &a, int *XC, int *blocks);
hipError_t launch_adj_pi(const AdjArgs &a, bool f32, double cF, double lamP, hipStream_t s);
hipError_t launch_adj_kick(const AdjArgs &a, bool f32, double c, double aK, double aV, hipStream_t s);
// Kick transpose of step n fused with the fields transpose of the same state (coefficients from
// a.sc): per-block partial sums of the kick at slots 0 .. 1 + n_terms, those of the fields transpose
// at kAdjFieldSlots (lam_phi . pi) and kAdjFieldSlots + 1 + t (monomials); launch_adj_finish_chains
// sums them.  with_fields = false: the kick transpose alone.
hipError_t launch_adj_kick_fields(const AdjArgs &a, bool f32, bool with_fields, hipStream_t s);

// Launchers (defined in the .hip files).
struct TiledPlan {
  int XC = 0;      // planes per workgroup (x-chunk)
  int G = 0;       // planes per partial-sum group
  bool NT = false; // non-temporal stores
  int blocks = 0;  // workgroups
  int rows = 0;    // partial-sum rows written by a reducing pass
};
bool scalar_tiled_plan(int N, int nf, int XC, bool nt, TiledPlan *p);
struct RowsPlan {
  int VZ = 0, WPR = 0, TY = 0;  // sites per lane, wavefronts per z-row, rows per workgroup
  int XC = 0, G = 0;
  int MM = 1;  // memory mode: bit 0 non-temporal stores, bit 1 NT momentum loads, bit 2 NT field loads
  int blocks = 0, rows = 0;
};
bool scalar_rows_plan(int N, int nf, bool f32, int XC, int mm, int shape, RowsPlan *p);
// Potential form the full-row pass uses with plan p: 1 for the compile-time lphi4 table (a run whose
// exponent table is lphi4's, on the shapes instantiated with it: FP32-bulk main shapes, FP64 512-thread
// shapes), else 0 (runtime table).
int scalar_rows_pk(int nf, bool f32, const ScalarPoly &P, const RowsPlan &p);
// kick2 (with drift and kick): kick a second time, by a.ckick2, from the same drifted fields.
hipError_t launch_scalar_rows(int nf, bool f32, bool drift, bool kick, bool red, const ScalarPassArgs &a,
                              const RowsPlan &p, hipStream_t s, bool kick2 = false);
hipError_t launch_scalar_tiled(int nf, bool f32, bool drift, bool kick, bool red, const ScalarPassArgs &a,
                               const TiledPlan &p, hipStream_t s, bool kick2 = false);
hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t s);
// Copies of up to kCopySetMax device arrays in one launch (sizes in 16-byte units).
constexpr int kCopySetMax = 12;
struct CopySet {
  const void *src[kCopySetMax];
  void *dst[kCopySetMax];
  size_t n16[kCopySetMax];
  int count;
};
hipError_t launch_copy_set(const CopySet &cs, hipStream_t s);
int scalar_pass_blocks(uint32_t V);
hipError_t launch_scalar_pass(int nf, bool f32, bool drift, bool kick, bool red, const ScalarPassArgs &a,
                              int blocks, hipStream_t s);
int gauge_pass_blocks(uint32_t V);
hipError_t launch_gauge_pass(bool f32, bool drift, bool kick, bool red, const GaugePassArgs &a, int blocks,
                             hipStream_t s);
hipError_t launch_update(DevState *st, const double *partials, int rows, const ModelConsts &mc,
                         const UpdateOps &ops, hipStream_t s);

struct GaussArgs {
  void *piA[3];
  void *phi[2];
  float *philo[2];
  void *pi[2];
  int N;
  uint32_t V;
  int f32;
  int has_cs;
  double inv_dx, dx, qnorm_gQ;
  double *partials;
};
hipError_t launch_gauss(const GaussArgs &a, int blocks, hipStream_t s);

struct MeansArgs {
  const void *p;
  const float *lo;
  uint32_t V;
  int f32;
  double scale;
  double *partials;
  uint32_t N;   // with grad: slot 2 = sum over sites of sum_i (f(x+i) - f(x))^2 (unscaled differences)
  int grad;
};
struct MonoArgs {  // slot t = sum over sites of monomial t of the scalar potential
  const void *phi[CLMI_MAX_SCALARS];
  const float *philo[CLMI_MAX_SCALARS];
  int nf, f32;
  uint32_t V;
  ScalarPoly poly;
  double *partials;
};
hipError_t launch_monomials(const MonoArgs &a, int blocks, hipStream_t s);
hipError_t launch_means(const MeansArgs &a, int blocks, hipStream_t s);
// Norm statistics of a 2- or 3-component field, or of the forward-difference curl of a 3-component
// field (curl = 1): slot 0 = sum over sites of |v|, slot 1 = sum of |v|^2.
struct NormArgs {
  const void *p[3];
  const float *lo[3];  // low words of hi/lo arrays (FP32-bulk fields), else null
  int ncomp, curl, f32;
  uint32_t N, V;
  double inv_dx;
  double *partials;
};
hipError_t launch_norms(const NormArgs &a, int blocks, hipStream_t s);
// dst[s] = component comp of the forward-difference curl of (p[0], p[1], p[2]) divided by dx.
hipError_t launch_curl_component(const NormArgs &a, int comp, double *dst, hipStream_t s);
hipError_t launch_sum_rows(const double *partials, int rows, double *out, hipStream_t s);

// Array conversion kernels (upload / download in F32-bulk mode, FFT staging).
hipError_t launch_split_f64(const double *src, float *hi, float *lo, uint32_t V, hipStream_t s);
hipError_t launch_to_f32(const double *src, float *dst, uint32_t V, hipStream_t s);
hipError_t launch_join_f64(const float *hi, const float *lo, double *dst, uint32_t V, hipStream_t s);

// Spectrum binning and Gaussian random fields (fft.hip).
constexpr int kMaxStreamBins = 8192;  // streaming binning: one LDS histogram of <= 64 KiB per wave
struct BinTable {
  int N = 0;
  double dk = 0;
  int nbins = 0;
  // streaming pass (nbins <= kMaxStreamBins)
  uint16_t *d_bin = nullptr;     // bin of each |n|^2, 0 <= |n|^2 <= 3 (N/2)^2
  double *d_rows = nullptr;      // groups x nbins per-workgroup bin sums
  int wpg = 0, groups = 0;       // waves per workgroup, workgroups
  uint32_t chunks_per_wave = 0;  // 64-mode chunks per wave
  // bin-sorted index table (more bins)
  uint32_t *d_modes = nullptr;   // half-space mode indices sorted by bin
  uint32_t *d_start = nullptr;   // nbins + 1 offsets
  std::vector<double> count, mean_n;
};
hipError_t build_bin_table(BinTable &t, int N, double dk, int nbins);
void free_bin_table(BinTable &t);
// out[b] = sum over the modes of bin b of w |f_k|^2 |n|^npow (npow = 0, 1 or 3; w = 2 for the modes the
// R2C half-space stands for twice).
hipError_t launch_bin_sums(const BinTable &t, const void *fk /*double2*/, double *out, hipStream_t s, int npow = 0);

struct GrfArgs {
  int N;
  uint32_t nmodes;
  uint64_t seed;
  uint32_t stream;
  double mass2, aDot, kIR, kCut;
  double sigma_pref;  // (omegaStar/fStar) * (L/dx^2)^(3/2) / sqrt(2)
  void *phik;  // double2
  void *pik;   // double2
};
hipError_t launch_grf_modes(const GrfArgs &a, int64_t n2max, hipStream_t s);
struct GrfCsArgs {
  int N;
  uint32_t nmodes;
  uint64_t seed;
  uint32_t stream;
  int comp;          // complex-scalar component 0 or 1
  double mass2[2];   // effective masses^2 of the two components
  double aDot, kIR;
  double sigma_pref;
  void *phik;
  void *pik;
};
hipError_t launch_grf_cs_modes(const GrfCsArgs &a, int64_t n2max, hipStream_t s);
hipError_t launch_u1_source(const void *f0, const float *f0lo, const void *f1, const float *f1lo, const void *p0,
                            const void *p1, int f32, double coef, double *out, uint32_t V, hipStream_t s);
hipError_t launch_gauss_solve(const void *j0, void *out, int dir, int N, hipStream_t s);
hipError_t launch_grf_finish(const double *src, double scale, double add, void *dst, float *dstlo, int f32,
                             uint32_t V, hipStream_t s);

}  // namespace clmi
