// C ABI of libclmi.so (declared in include/clmi.h): context management, pass orchestration,
// data movement, spectra and initial conditions.
//
// Leapfrog schedule.  The reference's step (evolvers/leapfrog.h:42-89) is
//   kick(aI_n) -> <pi^2> -> scale-factor kick+drift -> drift(aSI) -> <grad^2>, <V>.
// Here the drift of step n is deferred and executed inside the next lattice pass, which then
// computes the drifted fields (at every stencil point), their reductions, the kick of step n+1 on
// them and <pi^2> of the kicked momenta, all in one read and one write of every array.  The
// reductions of that pass are exactly the storeFieldsAverages of step n and the
// storeMomentaAverages of step n+1, so the background update that follows reproduces the
// reference's sequence of scalar operations.  Any call that observes the fields first flushes the
// deferred drift with a drift-only pass.
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "clmi_internal.hpp"
#include "scalar_poly.hpp"

using namespace clmi;

namespace {
thread_local std::string g_create_error;

enum Sector { SEC_S = 0, SEC_CS = 1 };
}  // namespace

struct clmi_ctx {
  clmi_config cfg{};
  int N = 0;
  uint32_t V = 0;
  bool f32 = false;
  int device = 0;
  hipStream_t stream = nullptr;

  // double-buffered arrays [array][buffer]
  DevArr fS[CLMI_MAX_SCALARS][2], pS[CLMI_MAX_SCALARS][2];
  DevArr fCS[2][2], pCS[2][2], fU[3][2], pU[3][2];
  int cur_fS = 0, cur_pS = 0, cur_fG = 0, cur_pG = 0;  // fG: phiCS and A together; pG: piCS and piA

  DevState *dstate = nullptr;
  double *partials = nullptr;
  double *dscratch = nullptr;  // small device scratch for row sums
  int blocks_s = 0, blocks_g = 0;
  int rows_g = 0;  // partial-sum rows a reducing charged-scalar + U(1) pass writes
  clmi_background *snap = nullptr;  // device target of the next storeFieldsAverages snapshot (adjoint tape)
  bool tiled = false;  // scalar passes use the tiled kernel with plan tplan
  TiledPlan tplan;
  bool rowk = false;   // scalar passes use the full-row kernel with plan rplan
  RowsPlan rplan;
  bool gtiled = false; // charged scalar + U(1) passes use the tiled kernel (gTY-row tiles, x-chunk gXC)
  int gTY = 4;
  int gXC = 0;
  size_t partial_rows = 0;

  ScalarPoly spoly{};
  CsPoly cpoly{};
  ModelConsts mc{};

  bool synced = true;
  bool pending_drift = false;

  // FFT resources (lazy)
  double *stage = nullptr;  // N^3 doubles
  void *fk = nullptr;       // N*N*(N/2+1) double2
  void *fk2 = nullptr;
  hipfftHandle plan_r2c = 0, plan_c2r = 0;
  bool have_r2c = false, have_c2r = false;
  BinTable bins;
  double *bin_out = nullptr;
  int bin_out_n = 0;

  // profiling
  bool prof = false;
  int prof_every = 1;     // time every prof_every-th drift + kick pass
  int64_t prof_count = 0; // drift + kick passes since enabling
  std::vector<hipEvent_t> ev;  // pairs
  size_t ev_used = 0;
  double prof_ms = 0.0;
  int64_t prof_launches = 0;

  std::string err;
};

namespace {

int fail(clmi_ctx *c, int code, const std::string &msg)
{
  if (c)
    c->err = msg;
  else
    g_create_error = msg;
  return code;
}

#define HIPCHK(c, expr)                                                                                     \
  do {                                                                                                      \
    hipError_t e_ = (expr);                                                                                 \
    if (e_ != hipSuccess) return fail((c), CLMI_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define FFTCHK(c, expr)                                                                              \
  do {                                                                                               \
    hipfftResult r_ = (expr);                                                                        \
    if (r_ != HIPFFT_SUCCESS) return fail((c), CLMI_E_HIP, std::string(#expr) + " failed: " + std::to_string((int)r_)); \
  } while (0)

#define CHK(expr)            \
  do {                       \
    int rc_ = (expr);        \
    if (rc_ != CLMI_OK) return rc_; \
  } while (0)

size_t bytes_field(bool f32) { return f32 ? 4 : 8; }

// Row-pass shape selector from clmi_config.flags (bit 0: VZ = 4 rows, bit 1: 4-row tiles).
// Memory mode of the full-row pass when neither the configuration nor clmi_autotune chooses one.
constexpr int kDefaultMemMode = 15;

int rows_shape(int flags)
{
  return ((flags & CLMI_FLAG_SHAPE_ALT) ? 1 : 0) | ((flags & CLMI_FLAG_SHAPE_TY4) ? 2 : 0);
}

int alloc_arr(clmi_ctx *c, DevArr &a, bool field)
{
  const size_t n = c->V;
  const size_t b = bytes_field(c->f32);
  HIPCHK(c, hipMalloc(&a.p, n * b));
  HIPCHK(c, hipMemsetAsync(a.p, 0, n * b, c->stream));
  if (c->f32 && field) {
    HIPCHK(c, hipMalloc((void **)&a.lo, n * sizeof(float)));
    HIPCHK(c, hipMemsetAsync(a.lo, 0, n * sizeof(float), c->stream));
  }
  return CLMI_OK;
}

void free_arr(DevArr &a)
{
  if (a.p) (void)hipFree(a.p);
  if (a.lo) (void)hipFree(a.lo);
  a.p = nullptr;
  a.lo = nullptr;
}

// Algorithmic HBM bytes of one fused step pass: every evolved array read once and written once.
double fused_pass_bytes(const clmi_ctx *c)
{
  const double bf = c->f32 ? 8.0 : 8.0;  // field: FP64, or FP32 hi + FP32 lo
  const double bp = c->f32 ? 4.0 : 8.0;  // momentum
  double per_site = c->cfg.n_scalars * 2.0 * (bf + bp);
  if (c->cfg.n_cscalars > 0) per_site += 2 * 2.0 * (bf + bp);
  if (c->cfg.n_u1 > 0) per_site += 3 * 2.0 * (bf + bp);
  return per_site * (double)c->V;
}

int prof_flush(clmi_ctx *c)
{
  if (c->ev_used == 0) return CLMI_OK;
  HIPCHK(c, hipEventSynchronize(c->ev[c->ev_used - 1]));
  for (size_t i = 0; i + 1 < c->ev_used; i += 2) {
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]));
    c->prof_ms += ms;
    c->prof_launches += 1;
  }
  c->ev_used = 0;
  return CLMI_OK;
}

int prof_event(clmi_ctx *c)
{
  if (c->ev_used == c->ev.size()) {
    if (c->ev.size() >= 2048) {
      CHK(prof_flush(c));
    } else {
      hipEvent_t e;
      HIPCHK(c, hipEventCreate(&e));
      c->ev.push_back(e);
    }
  }
  HIPCHK(c, hipEventRecord(c->ev[c->ev_used], c->stream));
  c->ev_used++;
  return CLMI_OK;
}

ScalarPassArgs scalar_args(const clmi_ctx *c, int fi, int fo, int pi, int po, double cdt_s, double ckick_s,
                           int row0)
{
  ScalarPassArgs a{};
  for (int f = 0; f < c->cfg.n_scalars; ++f) {
    a.phi_in[f] = c->fS[f][fi].p;
    a.philo_in[f] = c->fS[f][fi].lo;
    a.pi_in[f] = c->pS[f][pi].p;
    a.phi_out[f] = c->fS[f][fo].p;
    a.philo_out[f] = c->fS[f][fo].lo;
    a.pi_out[f] = c->pS[f][po].p;
  }
  a.N = c->N;
  a.V = c->V;
  a.inv_dx2 = 1.0 / (c->cfg.dx * c->cfg.dx);
  a.inv_dx = 1.0 / c->cfg.dx;
  a.cdt = cdt_s;
  a.ckick = ckick_s;
  a.st = c->dstate;
  a.partials = c->partials + (size_t)row0 * kNumRed;
  a.poly = c->spoly;
  return a;
}

// Arguments of the charged-scalar + U(1) pass from buffers fi/fo (fields) and pi/po (momenta).
GaugePassArgs gauge_args(const clmi_ctx *c, int fi, int fo, int pi, int po, double cdt_s, double ckick_g, int row0)
{
  GaugePassArgs a{};
  for (int k = 0; k < 2; ++k) {
    a.phi_in[k] = c->fCS[k][fi].p;
    a.philo_in[k] = c->fCS[k][fi].lo;
    a.pi_in[k] = c->pCS[k][pi].p;
    a.phi_out[k] = c->fCS[k][fo].p;
    a.philo_out[k] = c->fCS[k][fo].lo;
    a.pi_out[k] = c->pCS[k][po].p;
  }
  for (int d = 0; d < 3; ++d)
    if (c->cfg.n_u1 > 0) {
      a.A_in[d] = c->fU[d][fi].p;
      a.Alo_in[d] = c->fU[d][fi].lo;
      a.piA_in[d] = c->pU[d][pi].p;
      a.A_out[d] = c->fU[d][fo].p;
      a.Alo_out[d] = c->fU[d][fo].lo;
      a.piA_out[d] = c->pU[d][po].p;
    }
  a.N = c->N;
  a.V = c->V;
  a.has_u1 = c->cfg.n_u1 > 0;
  a.dx = c->cfg.dx;
  a.inv_dx = 1.0 / c->cfg.dx;
  a.inv_dx2 = 1.0 / (c->cfg.dx * c->cfg.dx);
  a.gQ = c->cfg.gQ;
  const double fw = c->cfg.fStar / c->cfg.omegaStar;
  a.jnorm = 2.0 / c->cfg.dx * (fw * fw);
  a.cdt = cdt_s;
  a.ckick = ckick_g;
  a.st = c->dstate;
  a.partials = c->partials + (size_t)row0 * kNumRed;
  a.poly = c->cpoly;
  return a;
}

// Launch one lattice pass over every sector; returns the number of partial rows written in *rows.
// ckick2_s != 0 (drift + kick passes of scalar-singlet models on the full-row pass): kick the scalars a
// second time, by ckick2_s, from the same drifted fields (see double_kick_ok).
int run_pass(clmi_ctx *c, bool drift, bool kick, bool red, double cdt_s, double ckick_s, double ckick_g, int *rows,
             double ckick2_s = 0.0)
{
  const bool kick2 = ckick2_s != 0.0;
  if (kick2 && !(drift && kick && (c->rowk || c->tiled) && c->cfg.n_cscalars == 0))
    return fail(c, CLMI_E_INVALID, "double kick outside a scalar-singlet full-row or tiled drift + kick pass");
  const bool timed = c->prof && drift && kick && (c->prof_count++ % c->prof_every == 0);
  if (timed) CHK(prof_event(c));
  int r = 0;
  if (c->cfg.n_scalars > 0) {
    const int fi = c->cur_fS, fo = drift ? 1 - c->cur_fS : c->cur_fS;
    const int pi = c->cur_pS, po = (drift && kick) ? 1 - c->cur_pS : c->cur_pS;
    ScalarPassArgs a = scalar_args(c, fi, fo, pi, po, cdt_s, ckick_s, r);
    a.ckick2 = ckick2_s;
    const int nrows = c->rowk ? c->rplan.rows : c->tiled ? c->tplan.rows : c->blocks_s;
    if (red && (size_t)r + nrows > c->partial_rows)
      return fail(c, CLMI_E_INVALID, "partial-sum buffer too small for the scalar pass");
    if (c->rowk) {
      HIPCHK(c, launch_scalar_rows(c->cfg.n_scalars, c->f32, drift, kick, red, a, c->rplan, c->stream, kick2));
      if (red) r += c->rplan.rows;
    } else if (c->tiled) {
      HIPCHK(c, launch_scalar_tiled(c->cfg.n_scalars, c->f32, drift, kick, red, a, c->tplan, c->stream, kick2));
      if (red) r += c->tplan.rows;
    } else {
      HIPCHK(c, launch_scalar_pass(c->cfg.n_scalars, c->f32, drift, kick, red, a, c->blocks_s, c->stream));
      if (red) r += c->blocks_s;
    }
    c->cur_fS = fo;
    c->cur_pS = po;
  }
  if (c->cfg.n_cscalars > 0) {
    const int fi = c->cur_fG, fo = drift ? 1 - c->cur_fG : c->cur_fG;
    const int pi = c->cur_pG, po = (drift && kick) ? 1 - c->cur_pG : c->cur_pG;
    const GaugePassArgs a = gauge_args(c, fi, fo, pi, po, cdt_s, ckick_g, r);
    if (red && (size_t)r + c->rows_g > c->partial_rows)
      return fail(c, CLMI_E_INVALID, "partial-sum buffer too small for the charged-scalar pass");
    if (c->gtiled)
      HIPCHK(c, launch_gauge_tiled(c->f32, c->gTY, drift, kick, red, a, c->blocks_g, c->gXC, c->stream));
    else
      HIPCHK(c, launch_gauge_pass(c->f32, drift, kick, red, a, c->blocks_g, c->stream));
    if (red) r += c->rows_g;
    c->cur_fG = fo;
    c->cur_pG = po;
  }
  if (timed) CHK(prof_event(c));
  *rows = r;
  return CLMI_OK;
}

UpdateOps make_ops(std::initializer_list<std::pair<int, double>> ops)
{
  UpdateOps u{};
  for (const auto &o : ops) {
    u.op[u.n] = o.first;
    u.w[u.n] = o.second;
    u.n++;
  }
  return u;
}

int update(clmi_ctx *c, int rows, std::initializer_list<std::pair<int, double>> ops)
{
  HIPCHK(c, launch_update(c->dstate, c->partials, rows, c->mc, make_ops(ops), c->stream));
  return CLMI_OK;
}

// A lattice pass followed by the background update `ops` (skipped when do_update is false).  When the
// context holds a snapshot target (c->snap) and the update starts with storeFieldsAverages, the
// background after that operation -- the state a flush at this point would leave -- is written there.
int pass_update(clmi_ctx *c, bool drift, bool kick, bool red, double cdt_s, double ckick_s, double ckick_g,
                bool do_update, std::initializer_list<std::pair<int, double>> ops, double ckick2_s = 0.0)
{
  int rows = 0;
  CHK(run_pass(c, drift, kick, red, cdt_s, ckick_s, ckick_g, &rows, ckick2_s));
  if (do_update) {
    UpdateOps u = make_ops(ops);
    if (c->snap && u.n > 0 && u.op[0] == OP_STORE_FIELDS) {
      u.snap = c->snap;
      c->snap = nullptr;
    }
    HIPCHK(c, launch_update(c->dstate, c->partials, rows, c->mc, u, c->stream));
  }
  return CLMI_OK;
}

// Whether velocity-Verlet stages may share a double-kick pass: scalar singlets only, on the full-row or
// the tiled pass.
bool double_kick_ok(const clmi_ctx *c)
{
  return (c->rowk || c->tiled) && c->cfg.n_scalars > 0 && c->cfg.n_cscalars == 0 && !(c->cfg.flags & CLMI_FLAG_NO_DOUBLE_KICK);
}

// Apply the deferred drift of the last leapfrog step (drift-only pass + storeFieldsAverages).
int flush(clmi_ctx *c)
{
  if (!c->pending_drift) return CLMI_OK;
  const bool ex = c->cfg.expansion != 0;
  CHK(pass_update(c, true, false, ex, c->cfg.dt, 0.0, 0.0, ex, {{OP_STORE_FIELDS, 0.0}}));
  c->pending_drift = false;
  return CLMI_OK;
}

int check_ctx(clmi_ctx *c)
{
  if (!c) return fail(nullptr, CLMI_E_INVALID, "null context");
  HIPCHK(c, hipSetDevice(c->device));
  return CLMI_OK;
}

int select_arr(clmi_ctx *c, int kind, int index, int comp, DevArr **out, bool *is_field)
{
  const clmi_config &g = c->cfg;
  switch (kind) {
    case CLMI_FLD_S:
    case CLMI_PI_S:
      if (index < 0 || index >= g.n_scalars || comp != 0) break;
      *is_field = kind == CLMI_FLD_S;
      *out = kind == CLMI_FLD_S ? &c->fS[index][c->cur_fS] : &c->pS[index][c->cur_pS];
      return CLMI_OK;
    case CLMI_FLD_CS:
    case CLMI_PI_CS:
      if (index != 0 || g.n_cscalars < 1 || comp < 0 || comp > 1) break;
      *is_field = kind == CLMI_FLD_CS;
      *out = kind == CLMI_FLD_CS ? &c->fCS[comp][c->cur_fG] : &c->pCS[comp][c->cur_pG];
      return CLMI_OK;
    case CLMI_FLD_U1:
    case CLMI_PI_U1:
      if (index != 0 || g.n_u1 < 1 || comp < 0 || comp > 2) break;
      *is_field = kind == CLMI_FLD_U1;
      *out = kind == CLMI_FLD_U1 ? &c->fU[comp][c->cur_fG] : &c->pU[comp][c->cur_pG];
      return CLMI_OK;
    default:
      break;
  }
  return fail(c, CLMI_E_INVALID,
              "no such field: kind " + std::to_string(kind) + " index " + std::to_string(index) + " comp " +
                  std::to_string(comp));
}

int ensure_stage(clmi_ctx *c)
{
  if (!c->stage) HIPCHK(c, hipMalloc(&c->stage, sizeof(double) * (size_t)c->V));
  return CLMI_OK;
}

int ensure_fft(clmi_ctx *c, bool r2c, bool c2r, bool two)
{
  CHK(ensure_stage(c));
  const size_t nm = (size_t)c->N * c->N * (c->N / 2 + 1);
  if (!c->fk) HIPCHK(c, hipMalloc(&c->fk, 16 * nm));
  if (two && !c->fk2) HIPCHK(c, hipMalloc(&c->fk2, 16 * nm));
  if (r2c && !c->have_r2c) {
    FFTCHK(c, hipfftPlan3d(&c->plan_r2c, c->N, c->N, c->N, HIPFFT_D2Z));
    FFTCHK(c, hipfftSetStream(c->plan_r2c, c->stream));
    c->have_r2c = true;
  }
  if (c2r && !c->have_c2r) {
    FFTCHK(c, hipfftPlan3d(&c->plan_c2r, c->N, c->N, c->N, HIPFFT_Z2D));
    FFTCHK(c, hipfftSetStream(c->plan_c2r, c->stream));
    c->have_c2r = true;
  }
  return CLMI_OK;
}

int read_state(clmi_ctx *c, DevState *h)
{
  HIPCHK(c, hipMemcpyAsync(h, c->dstate, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CLMI_OK;
}

}  // namespace

extern "C" {

int clmi_create(const clmi_config *cfg, clmi_ctx **out)
{
  if (!cfg || !out) return fail(nullptr, CLMI_E_INVALID, "null argument");
  *out = nullptr;
  const clmi_config &g = *cfg;
  if (g.N < 4 || g.N > 1024 || (g.N & 1)) return fail(nullptr, CLMI_E_INVALID, "N must be even, 4 <= N <= 1024");
  if (g.n_scalars < 0 || g.n_scalars > CLMI_MAX_SCALARS) return fail(nullptr, CLMI_E_INVALID, "n_scalars out of range");
  if (g.n_cscalars < 0 || g.n_cscalars > 1) return fail(nullptr, CLMI_E_INVALID, "n_cscalars must be 0 or 1");
  if (g.n_u1 < 0 || g.n_u1 > 1 || (g.n_u1 == 1 && g.n_cscalars != 1))
    return fail(nullptr, CLMI_E_INVALID, "n_u1 must be 0 or 1 and requires one complex scalar");
  if (g.n_scalars + g.n_cscalars == 0) return fail(nullptr, CLMI_E_INVALID, "model has no fields");
  if (g.precision != CLMI_F64 && g.precision != CLMI_F32_BULK) return fail(nullptr, CLMI_E_INVALID, "bad precision");
  if (!(g.dx > 0) || g.dt == 0.0) return fail(nullptr, CLMI_E_INVALID, "dx must be > 0 and dt != 0");
  if (g.n_terms < 0 || g.n_terms > CLMI_MAX_TERMS || g.n_cs_terms < 0 || g.n_cs_terms > CLMI_MAX_TERMS)
    return fail(nullptr, CLMI_E_INVALID, "too many potential terms");
  for (int t = 0; t < g.n_terms; ++t)
    for (int f = 0; f < CLMI_MAX_SCALARS; ++f)
      if (g.term_exp[t][f] < 0 || g.term_exp[t][f] > 16 || (f >= g.n_scalars && g.term_exp[t][f] != 0))
        return fail(nullptr, CLMI_E_INVALID, "bad exponent in scalar potential term " + std::to_string(t));
  for (int k = 0; k < g.n_cs_terms; ++k)
    if (g.cs_pow[k] < 0 || g.cs_pow[k] > 32 || (g.cs_pow[k] & 1))
      return fail(nullptr, CLMI_E_INVALID, "complex-scalar potential powers must be even, 0..32");
  if (g.expansion && !(g.MPl > 0 && g.fStar > 0)) return fail(nullptr, CLMI_E_INVALID, "expansion needs MPl, fStar > 0");

  clmi_ctx *c = new clmi_ctx();
  c->cfg = g;
  c->N = g.N;
  c->V = (uint32_t)g.N * g.N * g.N;
  c->f32 = g.precision == CLMI_F32_BULK;
  c->device = g.device;
  auto bail = [&](int rc) {
    g_create_error = c->err;
    clmi_destroy(c);
    return rc;
  };
  if (hipSetDevice(g.device) != hipSuccess) {
    fail(c, CLMI_E_HIP, "hipSetDevice failed");
    return bail(CLMI_E_HIP);
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    fail(c, CLMI_E_HIP, "hipStreamCreate failed");
    return bail(CLMI_E_HIP);
  }
  int rc = CLMI_OK;
  for (int f = 0; f < g.n_scalars && rc == CLMI_OK; ++f)
    for (int b = 0; b < 2 && rc == CLMI_OK; ++b) {
      rc = alloc_arr(c, c->fS[f][b], true);
      if (rc == CLMI_OK) rc = alloc_arr(c, c->pS[f][b], false);
    }
  for (int k = 0; k < 2 * g.n_cscalars && rc == CLMI_OK; ++k)
    for (int b = 0; b < 2 && rc == CLMI_OK; ++b) {
      rc = alloc_arr(c, c->fCS[k][b], true);
      if (rc == CLMI_OK) rc = alloc_arr(c, c->pCS[k][b], false);
    }
  for (int d = 0; d < 3 * g.n_u1 && rc == CLMI_OK; ++d)
    for (int b = 0; b < 2 && rc == CLMI_OK; ++b) {
      rc = alloc_arr(c, c->fU[d][b], true);
      if (rc == CLMI_OK) rc = alloc_arr(c, c->pU[d][b], false);
    }
  if (rc != CLMI_OK) return bail(CLMI_E_NOMEM);

  c->blocks_s = scalar_pass_blocks(c->V);
  c->blocks_g = gauge_pass_blocks(c->V);
  c->rows_g = c->blocks_g;
  if (!(g.flags & CLMI_FLAG_NO_TILED)) {
    // Default x-chunk of the full-row pass: the longest of 32, 16, 8 that still gives every CU a
    // workgroup, else the shortest (small lattices).  clmi_autotune refines it; results do not depend on it.
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || ncu < 1)
      ncu = 256;
    RowsPlan rp;
    const int mm_req = (g.flags >> 8) & 15;
    // FP64 at 256^3: the full-row pass needs 1024-thread workgroups, whose 128-VGPR cap makes it spill, and
    // the tiled pass is faster there.  Batched ensembles stop at 192^3, so nothing needs the row plan.  A
    // memory-mode or shape flag still asks for the full-row pass.
    const bool asked = mm_req != 0 || (g.flags & (CLMI_FLAG_SHAPE_ALT | CLMI_FLAG_SHAPE_TY4)) != 0;
    const bool prefer_tiled = g.precision == CLMI_F64 && g.N == 256 && !asked;
    for (int xc : {32, 16, 8})
      if (!prefer_tiled &&
          scalar_rows_plan(g.N, g.n_scalars, g.precision == CLMI_F32_BULK, xc, mm_req ? mm_req : kDefaultMemMode,
                           rows_shape(g.flags), &rp)) {
        c->rowk = true;
        c->rplan = rp;
        if (rp.blocks >= ncu) break;
      }
    if (mm_req && !c->rowk) {
      fail(c, CLMI_E_INVALID, "memory mode " + std::to_string(mm_req) + " is not available for this lattice");
      return bail(CLMI_E_INVALID);
    }
    TiledPlan p;
    if (!c->rowk)
      for (int xc : {16, 8, 4, 32, 64})
        if (scalar_tiled_plan(g.N, g.n_scalars, xc, true, &p)) {
          c->tiled = true;
          c->tplan = p;
          break;
        }
    int gxc = 0, gblocks = 0, grows_t = 0;
    const int gty = (g.flags & CLMI_FLAG_GAUGE_TY8) ? 8 : 4;
    if (g.n_cscalars > 0 && gauge_tiled_plan(g.N, gty, ncu, &gxc, &gblocks, &grows_t)) {
      c->gtiled = true;
      c->gTY = gty;
      c->gXC = gxc;
      c->blocks_g = gblocks;
      c->rows_g = grows_t;
    }
  }
  // Partial-sum rows: a reducing step pass writes the scalar sector's rows, then the charged-scalar + U(1)
  // sector's after them; the measurement kernels write at most kMaxPassBlocks rows each.
  const size_t srows = g.n_scalars == 0 ? 0
                       : c->rowk        ? (size_t)c->rplan.rows
                       : c->tiled       ? (size_t)c->tplan.rows
                                        : (size_t)c->blocks_s;
  const size_t grows = g.n_cscalars > 0 ? (size_t)c->rows_g : 0;
  const size_t rows = std::max(2 * (size_t)kMaxPassBlocks, srows + grows) + 16;
  c->partial_rows = rows;
  if (hipMalloc(&c->partials, rows * kNumRed * sizeof(double)) != hipSuccess ||
      hipMalloc(&c->dscratch, 64 * sizeof(double)) != hipSuccess ||
      hipMalloc(&c->dstate, sizeof(DevState)) != hipSuccess) {
    fail(c, CLMI_E_NOMEM, "device allocation failed");
    return bail(CLMI_E_NOMEM);
  }

  c->spoly.nterms = g.n_terms;
  for (int t = 0; t < g.n_terms; ++t) {
    c->spoly.c[t] = g.term_coeff[t];
    for (int f = 0; f < CLMI_MAX_SCALARS; ++f) c->spoly.e[t][f] = g.term_exp[t][f];
  }
  c->cpoly.nterms = g.n_cs_terms;
  for (int k = 0; k < g.n_cs_terms; ++k) {
    c->cpoly.c[k] = g.cs_coeff[k];
    c->cpoly.h[k] = g.cs_pow[k] / 2;
  }
  c->mc.ns = g.n_scalars;
  c->mc.ncs = g.n_cscalars;
  c->mc.nu1 = g.n_u1;
  c->mc.expansion = g.expansion;
  c->mc.dt = g.dt;
  c->mc.alpha = g.alpha;
  c->mc.fStar = g.fStar;
  c->mc.omegaStar = g.omegaStar;
  c->mc.MPl = g.MPl;
  c->mc.invV = 1.0 / ((double)g.N * g.N * g.N);

  DevState h{};
  h.bg.aI = h.bg.aSI = h.bg.aIM = 1.0;
  if (hipMemcpyAsync(c->dstate, &h, sizeof(DevState), hipMemcpyHostToDevice, c->stream) != hipSuccess) {
    fail(c, CLMI_E_HIP, "state upload failed");
    return bail(CLMI_E_HIP);
  }
  if (update(c, 0, {}) != CLMI_OK || hipStreamSynchronize(c->stream) != hipSuccess) return bail(CLMI_E_HIP);
  *out = c;
  return CLMI_OK;
}

void clmi_destroy(clmi_ctx *c)
{
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int f = 0; f < CLMI_MAX_SCALARS; ++f)
    for (int b = 0; b < 2; ++b) {
      free_arr(c->fS[f][b]);
      free_arr(c->pS[f][b]);
    }
  for (int k = 0; k < 2; ++k)
    for (int b = 0; b < 2; ++b) {
      free_arr(c->fCS[k][b]);
      free_arr(c->pCS[k][b]);
    }
  for (int d = 0; d < 3; ++d)
    for (int b = 0; b < 2; ++b) {
      free_arr(c->fU[d][b]);
      free_arr(c->pU[d][b]);
    }
  if (c->partials) (void)hipFree(c->partials);
  if (c->dscratch) (void)hipFree(c->dscratch);
  if (c->dstate) (void)hipFree(c->dstate);
  if (c->stage) (void)hipFree(c->stage);
  if (c->fk) (void)hipFree(c->fk);
  if (c->fk2) (void)hipFree(c->fk2);
  if (c->bin_out) (void)hipFree(c->bin_out);
  free_bin_table(c->bins);
  if (c->have_r2c) (void)hipfftDestroy(c->plan_r2c);
  if (c->have_c2r) (void)hipfftDestroy(c->plan_c2r);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *clmi_last_error(const clmi_ctx *c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int clmi_synchronize(clmi_ctx *c)
{
  CHK(check_ctx(c));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CLMI_OK;
}

int clmi_upload_field(clmi_ctx *c, int kind, int index, int comp, const double *host)
{
  CHK(check_ctx(c));
  if (!host) return fail(c, CLMI_E_INVALID, "null host buffer");
  CHK(flush(c));
  DevArr *a = nullptr;
  bool is_field = false;
  CHK(select_arr(c, kind, index, comp, &a, &is_field));
  const size_t n = c->V;
  if (!c->f32) {
    HIPCHK(c, hipMemcpyAsync(a->p, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  } else {
    CHK(ensure_stage(c));
    HIPCHK(c, hipMemcpyAsync(c->stage, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (is_field)
      HIPCHK(c, launch_split_f64(c->stage, (float *)a->p, a->lo, c->V, c->stream));
    else
      HIPCHK(c, launch_to_f32(c->stage, (float *)a->p, c->V, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CLMI_OK;
}

int clmi_download_field(clmi_ctx *c, int kind, int index, int comp, double *host)
{
  CHK(check_ctx(c));
  if (!host) return fail(c, CLMI_E_INVALID, "null host buffer");
  CHK(flush(c));
  DevArr *a = nullptr;
  bool is_field = false;
  CHK(select_arr(c, kind, index, comp, &a, &is_field));
  const size_t n = c->V;
  if (!c->f32) {
    HIPCHK(c, hipMemcpyAsync(host, a->p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  } else {
    CHK(ensure_stage(c));
    HIPCHK(c, launch_join_f64((const float *)a->p, a->lo, c->stage, c->V, c->stream));
    HIPCHK(c, hipMemcpyAsync(host, c->stage, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CLMI_OK;
}

int clmi_set_background(clmi_ctx *c, const clmi_background *bg)
{
  CHK(check_ctx(c));
  if (!bg) return fail(c, CLMI_E_INVALID, "null background");
  CHK(flush(c));
  DevState h{};
  CHK(read_state(c, &h));
  h.bg = *bg;
  HIPCHK(c, hipMemcpyAsync(c->dstate, &h, sizeof(DevState), hipMemcpyHostToDevice, c->stream));
  CHK(update(c, 0, {}));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CLMI_OK;
}

int clmi_get_background(clmi_ctx *c, clmi_background *bg)
{
  CHK(check_ctx(c));
  if (!bg) return fail(c, CLMI_E_INVALID, "null background");
  CHK(flush(c));
  DevState h{};
  CHK(read_state(c, &h));
  *bg = h.bg;
  return CLMI_OK;
}

int clmi_lf_init(clmi_ctx *c, int restart)
{
  CHK(check_ctx(c));
  CHK(flush(c));
  c->synced = restart == 0;
  if (c->synced && c->cfg.expansion) CHK(update(c, 0, {{OP_LF_INIT, 0.0}}));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CLMI_OK;
}

int clmi_lf_evolve(clmi_ctx *c, int nsteps)
{
  CHK(check_ctx(c));
  if (nsteps < 0) return fail(c, CLMI_E_INVALID, "nsteps < 0");
  const bool ex = c->cfg.expansion != 0;
  const double dt = c->cfg.dt;
  for (int n = 0; n < nsteps; ++n) {
    const double w = c->synced ? 0.5 : 1.0;
    if (c->pending_drift) {
      CHK(pass_update(c, true, true, ex, dt, w * dt, w * dt, ex, {{OP_STORE_FIELDS, 0.0}, {OP_LF_MOMENTA, 0.0}}));
    } else {
      CHK(pass_update(c, false, true, ex, 0.0, w * dt, w * dt, ex, {{OP_LF_MOMENTA, 0.0}}));
    }
    c->pending_drift = true;
    c->synced = false;
  }
  return CLMI_OK;
}

int clmi_lf_sync(clmi_ctx *c)
{
  CHK(check_ctx(c));
  CHK(flush(c));
  if (!c->synced) {
    const bool ex = c->cfg.expansion != 0;
    CHK(pass_update(c, false, true, ex, 0.0, 0.5 * c->cfg.dt, 0.5 * c->cfg.dt, ex, {{OP_LF_SYNC, 0.0}}));
  }
  c->synced = true;
  return CLMI_OK;
}

int clmi_lf_is_synced(const clmi_ctx *c) { return c && c->synced ? 1 : 0; }

int clmi_vv_evolve(clmi_ctx *c, const double *ws, int nws, int nsteps)
{
  CHK(check_ctx(c));
  if (!ws || nws < 1 || nsteps < 0) return fail(c, CLMI_E_INVALID, "bad composition weights");
  CHK(flush(c));
  const bool ex = c->cfg.expansion != 0;
  const double dt = c->cfg.dt;
  const int stages = (nws - 1) * 2 + 1;
  auto weight = [&](int s) { return ws[s < nws ? s : stages - s - 1]; };
  if (double_kick_ok(c)) {
    // The second half-kick of a stage and the first half-kick of the next read the same drifted fields,
    // and the scale-factor operations between them change aDot but not a_I, so both kicks have the
    // same coefficients: one double-kick pass gives exactly the momenta, and in its two sums exactly
    // the averages, of the two passes they replace (velocityverlet.h:51-114).
    const int K = nsteps * stages;
    auto wk = [&](int k) { return weight(k % stages); };
    if (K > 0) {
      if (ex) CHK(update(c, 0, {{OP_VV_SF_HALF, wk(0)}}));
      CHK(pass_update(c, false, true, ex, 0.0, wk(0) * dt / 2, dt / 2 * wk(0), ex,
                      {{OP_VV_MOMENTA, 0.0}, {OP_VV_DRIFT_SF, wk(0)}}));
    }
    for (int k = 0; k < K; ++k) {
      const double w = wk(k);
      if (k + 1 < K) {
        const double w2 = wk(k + 1);
        CHK(pass_update(c, true, true, ex, dt * w, w * dt / 2, dt / 2 * w, ex,
                        {{OP_STORE_FIELDS, 0.0}, {OP_VV_MOMENTA, 0.0}, {OP_VV_SF_ONE, w}, {OP_VV_SF_HALF, w2},
                         {OP_VV_MOMENTA_KICK2, 0.0}, {OP_VV_DRIFT_SF, w2}},
                        w2 * dt / 2));
      } else {
        CHK(pass_update(c, true, true, ex, dt * w, w * dt / 2, dt / 2 * w, ex,
                        {{OP_STORE_FIELDS, 0.0}, {OP_VV_MOMENTA, 0.0}, {OP_VV_SF_ONE, w}}));
      }
    }
    return CLMI_OK;
  }
  for (int n = 0; n < nsteps; ++n)
    for (int s = 0; s < stages; ++s) {
      const double w = weight(s);
      // the scale-factor half kick that opens a stage; after the first it rides on the previous stage's
      // last update (no lattice work lies between them, so the scalar sequence is the same)
      if (ex && n == 0 && s == 0) CHK(update(c, 0, {{OP_VV_SF_HALF, w}}));
      CHK(pass_update(c, false, true, ex, 0.0, w * dt / 2, dt / 2 * w, ex,
                      {{OP_VV_MOMENTA, 0.0}, {OP_VV_DRIFT_SF, w}}));
      // drift and second half-kick in one pass: the drift factor (a_SI) and the kick factors (a_I) are both
      // set by the update above, and the pass's sums are the drifted fields' and the kicked momenta's
      const bool last = n == nsteps - 1 && s == stages - 1;
      if (ex && !last)
        CHK(pass_update(c, true, true, ex, dt * w, w * dt / 2, dt / 2 * w, ex,
                        {{OP_STORE_FIELDS, 0.0}, {OP_VV_MOMENTA, 0.0}, {OP_VV_SF_ONE, w},
                         {OP_VV_SF_HALF, weight((s + 1) % stages)}}));
      else
        CHK(pass_update(c, true, true, ex, dt * w, w * dt / 2, dt / 2 * w, ex,
                        {{OP_STORE_FIELDS, 0.0}, {OP_VV_MOMENTA, 0.0}, {OP_VV_SF_ONE, w}}));
    }
  return CLMI_OK;
}

int clmi_pv_evolve(clmi_ctx *c, const double *ws, int nws, int nsteps)
{
  CHK(check_ctx(c));
  if (!ws || nws < 1 || nsteps < 0) return fail(c, CLMI_E_INVALID, "bad composition weights");
  CHK(flush(c));
  const bool ex = c->cfg.expansion != 0;
  const double dt = c->cfg.dt;
  const int stages = (nws - 1) * 2 + 1;
  auto weight = [&](int s) { return ws[s < nws ? s : stages - s - 1]; };
  if (!ex) CHK(update(c, 0, {{OP_PV_DRIFT_AI, 0.0}}));
  for (int n = 0; n < nsteps; ++n)
    for (int s = 0; s < stages; ++s) {
      const double w = weight(s);
      // drift (aI of the stage start) + kick (half-step aI) in one pass, then the mid-stage update; the
      // stage's first scale-factor drift rides on the previous stage's last update after the first stage
      if (ex && n == 0 && s == 0) CHK(update(c, 0, {{OP_PV_SF_FIRST, w}}));
      CHK(pass_update(c, true, true, ex, dt / 2 * w, w * dt, w * dt, ex, {{OP_PV_MID, w}}));
      // second drift (new aI) + storeFieldsAveragesInt
      const bool last = n == nsteps - 1 && s == stages - 1;
      if (ex && !last)
        CHK(pass_update(c, true, false, ex, dt / 2 * w, 0.0, 0.0, ex,
                        {{OP_STORE_FIELDS, 0.0}, {OP_PV_SF_FIRST, weight((s + 1) % stages)}}));
      else
        CHK(pass_update(c, true, false, ex, dt / 2 * w, 0.0, 0.0, ex, {{OP_STORE_FIELDS, 0.0}}));
    }
  // static background: restore the aSI-based drift coefficients of the other evolvers (with
  // expansion the last update already did)
  if (!ex) CHK(update(c, 0, {}));
  return CLMI_OK;
}

int clmi_kick(clmi_ctx *c, double ck)
{
  CHK(check_ctx(c));
  CHK(flush(c));
  int rows = 0;
  CHK(run_pass(c, false, true, false, 0.0, ck, ck, &rows));
  return CLMI_OK;
}

int clmi_drift(clmi_ctx *c, double cdt)
{
  CHK(check_ctx(c));
  CHK(flush(c));
  int rows = 0;
  CHK(run_pass(c, true, false, false, cdt, 0.0, 0.0, &rows));
  return CLMI_OK;
}

int clmi_set_all_averages(clmi_ctx *c)
{
  CHK(check_ctx(c));
  CHK(flush(c));
  CHK(pass_update(c, false, false, true, 0.0, 0.0, 0.0, true, {{OP_SET_ALL, 0.0}}));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CLMI_OK;
}

int clmi_volume_averages(clmi_ctx *c, clmi_averages *out)
{
  CHK(check_ctx(c));
  if (!out) return fail(c, CLMI_E_INVALID, "null output");
  CHK(flush(c));
  CHK(pass_update(c, false, false, true, 0.0, 0.0, 0.0, true, {}));
  DevState h{};
  CHK(read_state(c, &h));
  out->pi2S = h.red[RED_PI2S];
  out->grad2S = h.red[RED_GRAD2S];
  out->pot = h.red[RED_POT];
  out->pi2CS = h.red[RED_PI2CS];
  out->grad2CS = h.red[RED_GRAD2CS];
  out->pi2U1 = h.red[RED_PI2U1];
  out->B2U1 = h.red[RED_B2U1];
  out->reserved = 0.0;
  return CLMI_OK;
}

int clmi_scale_factor_kernel(clmi_ctx *c, double *K)
{
  CHK(check_ctx(c));
  if (!K) return fail(c, CLMI_E_INVALID, "null output");
  CHK(flush(c));
  CHK(update(c, 0, {{OP_KERNEL_ONLY, 0.0}}));
  DevState h{};
  CHK(read_state(c, &h));
  *K = h.K;
  return CLMI_OK;
}

int clmi_hubble_constraint(clmi_ctx *c, double out[3])
{
  CHK(clmi_set_all_averages(c));
  DevState h{};
  CHK(read_state(c, &h));
  const clmi_background &b = h.bg;
  const clmi_config &g = c->cfg;
  const double a = b.aI;
  double rho = b.potAvI;
  if (g.n_scalars > 0) rho += 0.5 * b.pi2AvI * std::pow(a, -6) + 0.5 * b.grad2AvI * std::pow(a, -2);
  if (g.n_cscalars > 0) rho += b.CSpi2AvI * std::pow(a, -6) + b.CSgrad2AvI * std::pow(a, -2);
  if (g.n_u1 > 0) {
    const double w2 = (g.omegaStar / g.fStar) * (g.omegaStar / g.fStar);
    rho += 0.5 * w2 * std::pow(a, -4) * (b.U1pi2AvI + b.U1Mag2AvI);
  }
  const double lhs = b.aDotI * b.aDotI;
  const double fm = g.fStar / g.MPl;
  const double rhs = std::pow(a, 2 * g.alpha + 2) / 3.0 * (fm * fm) * rho;
  out[0] = (lhs - rhs) / (lhs + rhs);
  out[1] = lhs;
  out[2] = rhs;
  return CLMI_OK;
}

int clmi_gauss_u1(clmi_ctx *c, double out[3])
{
  CHK(check_ctx(c));
  if (!out) return fail(c, CLMI_E_INVALID, "null output");
  if (c->cfg.n_u1 < 1) return fail(c, CLMI_E_STATE, "model has no U(1) field");
  CHK(flush(c));
  GaussArgs a{};
  for (int d = 0; d < 3; ++d) a.piA[d] = c->pU[d][c->cur_pG].p;
  for (int k = 0; k < 2; ++k) {
    a.phi[k] = c->fCS[k][c->cur_fG].p;
    a.philo[k] = c->fCS[k][c->cur_fG].lo;
    a.pi[k] = c->pCS[k][c->cur_pG].p;
  }
  a.N = c->N;
  a.V = c->V;
  a.f32 = c->f32;
  a.has_cs = c->cfg.n_cscalars > 0;
  a.inv_dx = 1.0 / c->cfg.dx;
  a.dx = c->cfg.dx;
  const double fw = c->cfg.fStar / c->cfg.omegaStar;
  a.qnorm_gQ = 2.0 / c->cfg.dx * (fw * fw) * c->cfg.gQ;
  a.partials = c->partials;
  const int gblocks = gauge_pass_blocks(c->V);
  HIPCHK(c, launch_gauss(a, gblocks, c->stream));
  HIPCHK(c, launch_sum_rows(c->partials, gblocks, c->dscratch, c->stream));
  double h[kNumRed];
  HIPCHK(c, hipMemcpyAsync(h, c->dscratch, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const double V = (double)c->V;
  out[0] = (h[0] / V) / (h[1] / V);
  out[1] = h[3] / V;
  out[2] = h[2] / V;
  return CLMI_OK;
}

int clmi_field_means(clmi_ctx *c, int kind, int index, int comp, double scale, double out[2])
{
  CHK(check_ctx(c));
  if (!out) return fail(c, CLMI_E_INVALID, "null output");
  CHK(flush(c));
  DevArr *arr = nullptr;
  bool is_field = false;
  CHK(select_arr(c, kind, index, comp, &arr, &is_field));
  MeansArgs a{};
  a.p = arr->p;
  a.lo = arr->lo;
  a.V = c->V;
  a.f32 = c->f32;
  a.scale = scale;
  a.partials = c->partials;
  a.N = (uint32_t)c->N;
  a.grad = 0;
  const int blocks = scalar_pass_blocks(c->V);
  HIPCHK(c, launch_means(a, blocks, c->stream));
  HIPCHK(c, launch_sum_rows(c->partials, blocks, c->dscratch, c->stream));
  double h[kNumRed];
  HIPCHK(c, hipMemcpyAsync(h, c->dscratch, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  out[0] = h[0] / (double)c->V;
  out[1] = h[1] / (double)c->V;
  return CLMI_OK;
}

int clmi_field_stats(clmi_ctx *c, int kind, int index, int comp, double out[3])
{
  CHK(check_ctx(c));
  if (!out) return fail(c, CLMI_E_INVALID, "null output");
  CHK(flush(c));
  DevArr *arr = nullptr;
  bool is_field = false;
  CHK(select_arr(c, kind, index, comp, &arr, &is_field));
  MeansArgs a{};
  a.p = arr->p;
  a.lo = arr->lo;
  a.V = c->V;
  a.f32 = c->f32;
  a.scale = 1.0;
  a.partials = c->partials;
  a.N = (uint32_t)c->N;
  a.grad = 1;
  const int blocks = scalar_pass_blocks(c->V);
  HIPCHK(c, launch_means(a, blocks, c->stream));
  HIPCHK(c, launch_sum_rows(c->partials, blocks, c->dscratch, c->stream));
  double h[kNumRed];
  HIPCHK(c, hipMemcpyAsync(h, c->dscratch, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const double dx = c->cfg.dx;
  out[0] = h[0] / (double)c->V;
  out[1] = h[1] / (double)c->V;
  out[2] = h[2] / (dx * dx) / (double)c->V;
  return CLMI_OK;
}

namespace {
// The three U(1) field components as a curl input (FLD_U1, current buffers).
int u1_curl_args(clmi_ctx *c, int index, NormArgs *a)
{
  for (int d = 0; d < 3; ++d) {
    DevArr *arr = nullptr;
    bool is_field = false;
    CHK(select_arr(c, CLMI_FLD_U1, index, d, &arr, &is_field));
    a->p[d] = arr->p;
    a->lo[d] = arr->lo;
  }
  a->ncomp = 3;
  a->curl = 1;
  a->f32 = c->f32;
  a->N = (uint32_t)c->N;
  a->V = c->V;
  a->inv_dx = 1.0 / c->cfg.dx;
  return CLMI_OK;
}
}  // namespace

int clmi_norm_averages(clmi_ctx *c, int kind, int index, double out[2])
{
  CHK(check_ctx(c));
  if (!out) return fail(c, CLMI_E_INVALID, "null output");
  CHK(flush(c));
  NormArgs a{};
  if (kind == CLMI_MAG_U1) {
    CHK(u1_curl_args(c, index, &a));
  } else if (kind == CLMI_FLD_CS || kind == CLMI_PI_CS || kind == CLMI_PI_U1) {
    a.ncomp = kind == CLMI_PI_U1 ? 3 : 2;
    for (int d = 0; d < a.ncomp; ++d) {
      DevArr *arr = nullptr;
      bool is_field = false;
      CHK(select_arr(c, kind, index, d, &arr, &is_field));
      a.p[d] = arr->p;
      a.lo[d] = arr->lo;
    }
    a.f32 = c->f32;
    a.N = (uint32_t)c->N;
    a.V = c->V;
  } else {
    return fail(c, CLMI_E_INVALID, "clmi_norm_averages: kind must be CLMI_FLD_CS, CLMI_PI_CS, CLMI_PI_U1 or CLMI_MAG_U1");
  }
  a.partials = c->partials;
  const int blocks = scalar_pass_blocks(c->V);
  HIPCHK(c, launch_norms(a, blocks, c->stream));
  HIPCHK(c, launch_sum_rows(c->partials, blocks, c->dscratch, c->stream));
  double h[kNumRed];
  HIPCHK(c, hipMemcpyAsync(h, c->dscratch, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  out[0] = h[0] / (double)c->V;
  out[1] = h[1] / (double)c->V;
  return CLMI_OK;
}

int clmi_monomial_averages(clmi_ctx *c, double *out)
{
  CHK(check_ctx(c));
  if (!out) return fail(c, CLMI_E_INVALID, "null output");
  if (c->cfg.n_scalars == 0) return fail(c, CLMI_E_INVALID, "no scalar singlets");
  CHK(flush(c));
  MonoArgs a{};
  for (int f = 0; f < c->cfg.n_scalars; ++f) {
    a.phi[f] = c->fS[f][c->cur_fS].p;
    a.philo[f] = c->fS[f][c->cur_fS].lo;
  }
  a.nf = c->cfg.n_scalars;
  a.f32 = c->f32;
  a.V = c->V;
  a.poly = c->spoly;
  a.partials = c->partials;
  const int blocks = scalar_pass_blocks(c->V);
  HIPCHK(c, launch_monomials(a, blocks, c->stream));
  HIPCHK(c, launch_sum_rows(c->partials, blocks, c->dscratch, c->stream));
  double h[kNumRed];
  HIPCHK(c, hipMemcpyAsync(h, c->dscratch, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int t = 0; t < c->spoly.nterms; ++t) out[t] = h[t] / (double)c->V;
  return CLMI_OK;
}

int clmi_spectrum_nbins(const clmi_ctx *c, double dk)
{
  if (!c || !(dk > 0)) return CLMI_E_INVALID;
  return (int)std::floor(std::sqrt(3.0) / 2.0 * c->N / dk) + 1;
}

int clmi_spectrum_bins(clmi_ctx *c, int kind, int index, int comp, double dk, int nbins, double *sum_f2,
                       double *count, double *mean_n)
{
  return clmi_spectrum_bins_npow(c, kind, index, comp, dk, nbins, 0, sum_f2, count, mean_n);
}

int clmi_spectrum_bins_npow(clmi_ctx *c, int kind, int index, int comp, double dk, int nbins, int npow,
                            double *sum_f2, double *count, double *mean_n)
{
  CHK(check_ctx(c));
  if (!(dk > 0) || nbins != clmi_spectrum_nbins(c, dk) || !sum_f2)
    return fail(c, CLMI_E_INVALID, "nbins must equal clmi_spectrum_nbins(ctx, dk)");
  if (npow != 0 && npow != 1 && npow != 3) return fail(c, CLMI_E_INVALID, "npow must be 0, 1 or 3");
  CHK(flush(c));
  double *fft_in = nullptr;
  if (kind == CLMI_MAG_U1) {
    if (comp < 0 || comp > 2) return fail(c, CLMI_E_INVALID, "CLMI_MAG_U1: comp must be 0, 1 or 2");
    NormArgs a{};
    CHK(u1_curl_args(c, index, &a));
    CHK(ensure_fft(c, true, false, false));
    HIPCHK(c, launch_curl_component(a, comp, c->stage, c->stream));
    fft_in = c->stage;
  } else {
    DevArr *arr = nullptr;
    bool is_field = false;
    CHK(select_arr(c, kind, index, comp, &arr, &is_field));
    CHK(ensure_fft(c, true, false, false));
    if (c->f32) {
      HIPCHK(c, launch_join_f64((const float *)arr->p, arr->lo, c->stage, c->V, c->stream));
      fft_in = c->stage;
    } else {
      fft_in = (double *)arr->p;  // the out-of-place R2C transform reads the array and leaves it unchanged
    }
  }
  FFTCHK(c, hipfftExecD2Z(c->plan_r2c, (hipfftDoubleReal *)fft_in, (hipfftDoubleComplex *)c->fk));
  if (c->bins.nbins != nbins || c->bins.dk != dk || c->bins.N != c->N) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, build_bin_table(c->bins, c->N, dk, nbins));
  }
  if (c->bin_out_n < nbins) {
    if (c->bin_out) HIPCHK(c, hipFree(c->bin_out));
    c->bin_out = nullptr;
    c->bin_out_n = 0;
    HIPCHK(c, hipMalloc(&c->bin_out, sizeof(double) * nbins));
    c->bin_out_n = nbins;
  }
  HIPCHK(c, launch_bin_sums(c->bins, c->fk, c->bin_out, c->stream, npow));
  HIPCHK(c, hipMemcpyAsync(sum_f2, c->bin_out, sizeof(double) * nbins, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int b = 0; b < nbins; ++b) {
    if (count) count[b] = c->bins.count[b];
    if (mean_n) mean_n[b] = c->bins.mean_n[b];
  }
  return CLMI_OK;
}

namespace {
// largest |n|^2 with sqrt(|n|^2) * kIR <= kCutOff (-1: no mode)
int64_t cut_n2max(int N, double kIR, double kCutOff)
{
  const double r = kCutOff / kIR;
  if (!(r >= 0)) return -1;
  if (r > 4.0 * N) return (int64_t)3 * N * N;
  int64_t n2max = (int64_t)std::floor(r * r);
  while (std::sqrt((double)(n2max + 1)) * kIR <= kCutOff) ++n2max;
  while (n2max >= 0 && std::sqrt((double)n2max) * kIR > kCutOff) --n2max;
  return n2max;
}
}  // namespace

int clmi_grf_scalar(clmi_ctx *c, int index, uint64_t seed, uint32_t stream, double mass2, double aDot, double kIR,
                    double kCutOff, double fld0, double pi0)
{
  CHK(check_ctx(c));
  if (index < 0 || index >= c->cfg.n_scalars) return fail(c, CLMI_E_INVALID, "no such scalar field");
  if (!(kIR > 0)) return fail(c, CLMI_E_INVALID, "kIR must be > 0");
  CHK(flush(c));
  CHK(ensure_fft(c, false, true, true));
  const int N = c->N;
  const double dx = c->cfg.dx;
  const double L = N * dx;
  GrfArgs a{};
  a.N = N;
  a.nmodes = (uint32_t)N * N * (N / 2 + 1);
  a.seed = seed;
  a.stream = stream;
  a.mass2 = mass2;
  a.aDot = aDot;
  a.kIR = kIR;
  a.kCut = kCutOff;
  a.sigma_pref = c->cfg.omegaStar / c->cfg.fStar * std::pow(L / (dx * dx), 1.5) / std::sqrt(2.0);
  a.phik = c->fk;
  a.pik = c->fk2;
  HIPCHK(c, launch_grf_modes(a, cut_n2max(N, kIR, kCutOff), c->stream));
  const double scale = 1.0 / ((double)N * N * N);
  DevArr &f = c->fS[index][c->cur_fS];
  DevArr &p = c->pS[index][c->cur_pS];
  FFTCHK(c, hipfftExecZ2D(c->plan_c2r, (hipfftDoubleComplex *)c->fk, (hipfftDoubleReal *)c->stage));
  HIPCHK(c, launch_grf_finish(c->stage, scale, fld0, f.p, f.lo, c->f32, c->V, c->stream));
  FFTCHK(c, hipfftExecZ2D(c->plan_c2r, (hipfftDoubleComplex *)c->fk2, (hipfftDoubleReal *)c->stage));
  HIPCHK(c, launch_grf_finish(c->stage, scale, pi0, p.p, nullptr, c->f32, c->V, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CLMI_OK;
}

int clmi_grf_u1(clmi_ctx *c, uint64_t seed, uint32_t stream, const double mass2[2], double aDot, double kIR,
                double kCutOff, const double fld0[2], const double pi0[2])
{
  CHK(check_ctx(c));
  if (c->cfg.n_cscalars != 1) return fail(c, CLMI_E_INVALID, "the model has no complex scalar");
  if (!mass2 || !fld0 || !pi0) return fail(c, CLMI_E_INVALID, "null argument");
  if (!(kIR > 0)) return fail(c, CLMI_E_INVALID, "kIR must be > 0");
  CHK(flush(c));
  CHK(ensure_fft(c, c->cfg.n_u1 > 0, true, true));
  const int N = c->N;
  const double dx = c->cfg.dx;
  const double L = N * dx;
  const double scale = 1.0 / ((double)N * N * N);
  GrfCsArgs a{};
  a.N = N;
  a.nmodes = (uint32_t)N * N * (N / 2 + 1);
  a.seed = seed;
  a.stream = stream;
  a.mass2[0] = mass2[0];
  a.mass2[1] = mass2[1];
  a.aDot = aDot;
  a.kIR = kIR;
  // the norm of fluctuationsgenerator.h:39-71 (both components share it; their masses differ)
  a.sigma_pref = c->cfg.omegaStar / c->cfg.fStar * std::pow(L / (dx * dx), 1.5) / std::sqrt(2.0);
  a.phik = c->fk;
  a.pik = c->fk2;
  const int64_t n2max = cut_n2max(N, kIR, kCutOff);
  for (int comp = 0; comp < 2; ++comp) {
    a.comp = comp;
    HIPCHK(c, launch_grf_cs_modes(a, n2max, c->stream));
    DevArr &f = c->fCS[comp][c->cur_fG];
    DevArr &p = c->pCS[comp][c->cur_pG];
    FFTCHK(c, hipfftExecZ2D(c->plan_c2r, (hipfftDoubleComplex *)c->fk, (hipfftDoubleReal *)c->stage));
    HIPCHK(c, launch_grf_finish(c->stage, scale, fld0[comp], f.p, f.lo, c->f32, c->V, c->stream));
    FFTCHK(c, hipfftExecZ2D(c->plan_c2r, (hipfftDoubleComplex *)c->fk2, (hipfftDoubleReal *)c->stage));
    HIPCHK(c, launch_grf_finish(c->stage, scale, pi0[comp], p.p, nullptr, c->f32, c->V, c->stream));
  }
  if (c->cfg.n_u1 > 0) {
    // A = 0; pi_A from the Gauss constraint sourced by the complex scalar's charge
    const DevArr &f0 = c->fCS[0][c->cur_fG], &f1 = c->fCS[1][c->cur_fG];
    const DevArr &p0 = c->pCS[0][c->cur_pG], &p1 = c->pCS[1][c->cur_pG];
    const double fw = c->cfg.fStar / c->cfg.omegaStar;
    const double coef = -dx * dx * (2.0 / dx) * fw * fw * c->cfg.gQ;
    HIPCHK(c, launch_u1_source(f0.p, f0.lo, f1.p, f1.lo, p0.p, p1.p, c->f32, coef, c->stage, c->V, c->stream));
    FFTCHK(c, hipfftExecD2Z(c->plan_r2c, (hipfftDoubleReal *)c->stage, (hipfftDoubleComplex *)c->fk));
    for (int d = 0; d < 3; ++d) {
      DevArr &A = c->fU[d][c->cur_fG];
      DevArr &P = c->pU[d][c->cur_pG];
      const size_t eb = c->f32 ? sizeof(float) : sizeof(double);
      HIPCHK(c, hipMemsetAsync(A.p, 0, eb * c->V, c->stream));
      if (A.lo) HIPCHK(c, hipMemsetAsync(A.lo, 0, sizeof(float) * c->V, c->stream));
      HIPCHK(c, launch_gauss_solve(c->fk, c->fk2, d, N, c->stream));
      FFTCHK(c, hipfftExecZ2D(c->plan_c2r, (hipfftDoubleComplex *)c->fk2, (hipfftDoubleReal *)c->stage));
      HIPCHK(c, launch_grf_finish(c->stage, scale, 0.0, P.p, nullptr, c->f32, c->V, c->stream));
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CLMI_OK;
}

}  // extern "C"

namespace {

// ---- adjoint of the leapfrog (scalar singlets) ----------------------------------------------

// Device copy of one leapfrog state (phi, pi of every scalar singlet) in storage precision.
struct AdjState {
  DevArr f[CLMI_MAX_SCALARS], p[CLMI_MAX_SCALARS];
};

// The adjoint's states live in one device allocation, carved into arrays of 256-byte aligned size.
size_t adj_arr_bytes(size_t b) { return (b + 255) / 256 * 256; }
size_t adj_state_bytes(const clmi_ctx *c)
{
  const size_t bf = adj_arr_bytes(bytes_field(c->f32) * c->V);
  const size_t bl = c->f32 ? adj_arr_bytes(sizeof(float) * c->V) : 0;
  return (size_t)c->cfg.n_scalars * (2 * bf + bl);
}
void adj_carve(const clmi_ctx *c, char *&p, AdjState &s)
{
  const size_t bf = adj_arr_bytes(bytes_field(c->f32) * c->V);
  for (int f = 0; f < c->cfg.n_scalars; ++f) {
    s.f[f].p = p;
    p += bf;
    if (c->f32) {
      s.f[f].lo = (float *)p;
      p += adj_arr_bytes(sizeof(float) * c->V);
    }
    s.p[f].p = p;
    p += bf;
  }
}
// One copy launch for a whole state (every array of every scalar singlet).
void copy_add(CopySet &cs, const void *src, void *dst, size_t bytes)
{
  cs.src[cs.count] = src;
  cs.dst[cs.count] = dst;
  cs.n16[cs.count] = bytes / 16;
  cs.count++;
}
int adj_copy_state(clmi_ctx *c, bool save, AdjState &s)
{
  CopySet cs{};
  const size_t bf = bytes_field(c->f32) * c->V;
  for (int f = 0; f < c->cfg.n_scalars; ++f) {
    DevArr &cf = c->fS[f][c->cur_fS], &cp = c->pS[f][c->cur_pS];
    if (save) {
      copy_add(cs, cf.p, s.f[f].p, bf);
      if (c->f32) copy_add(cs, cf.lo, s.f[f].lo, sizeof(float) * c->V);
      copy_add(cs, cp.p, s.p[f].p, bf);
    } else {
      copy_add(cs, s.f[f].p, cf.p, bf);
      if (c->f32) copy_add(cs, s.f[f].lo, cf.lo, sizeof(float) * c->V);
      copy_add(cs, s.p[f].p, cp.p, bf);
    }
  }
  HIPCHK(c, launch_copy_set(cs, c->stream));
  return CLMI_OK;
}
int adj_save(clmi_ctx *c, AdjState &s) { return adj_copy_state(c, true, s); }
int adj_restore(clmi_ctx *c, const AdjState &s) { return adj_copy_state(c, false, const_cast<AdjState &>(s)); }
AdjFields adj_view(const clmi_ctx *c, const AdjState &s)
{
  AdjFields v{};
  for (int f = 0; f < c->cfg.n_scalars; ++f) {
    v.phi[f] = s.f[f].p;
    v.philo[f] = s.f[f].lo;
    v.pi[f] = s.p[f].p;
  }
  return v;
}

// The context's scalar-singlet buffers, restored when the recomputation of a segment ends (it points
// the context's spare buffers at segment slots).
struct AdjBufferGuard {
  clmi_ctx *c;
  DevArr f[CLMI_MAX_SCALARS][2], p[CLMI_MAX_SCALARS][2];
  explicit AdjBufferGuard(clmi_ctx *ctx) : c(ctx)
  {
    for (int i = 0; i < CLMI_MAX_SCALARS; ++i)
      for (int b = 0; b < 2; ++b) {
        f[i][b] = c->fS[i][b];
        p[i][b] = c->pS[i][b];
      }
  }
  ~AdjBufferGuard()
  {
    for (int i = 0; i < CLMI_MAX_SCALARS; ++i)
      for (int b = 0; b < 2; ++b) {
        c->fS[i][b] = f[i][b];
        c->pS[i][b] = p[i][b];
      }
  }
};

AdjFields adj_view_cur(const clmi_ctx *c)
{
  AdjFields v{};
  for (int f = 0; f < c->cfg.n_scalars; ++f) {
    v.phi[f] = c->fS[f][c->cur_fS].p;
    v.philo[f] = c->fS[f][c->cur_fS].lo;
    v.pi[f] = c->pS[f][c->cur_pS].p;
  }
  return v;
}

struct AdjRun {
  clmi_ctx *c = nullptr;
  std::vector<AdjState> checkpoints, segment;
  std::vector<AdjFields> seg_view;  // state after step n0 + i of the segment being reversed
  AdjState final_state;
  void *pool = nullptr;        // every state's arrays
  double *lam = nullptr;       // 2 * n_scalars FP64 arrays
  double *partials = nullptr;  // kAdjMaxRows x kAdjRed
  double *red = nullptr;       // kAdjRed
  AdjScalars *sc = nullptr;    // background adjoint during the reverse sweep
  clmi_background *tape = nullptr;  // device copy of the forward background after every step
  ~AdjRun()
  {
    if (pool) (void)hipFree(pool);
    if (lam) (void)hipFree(lam);
    if (partials) (void)hipFree(partials);
    if (red) (void)hipFree(red);
    if (sc) (void)hipFree(sc);
    if (tape) (void)hipFree(tape);
  }
};

int adj_read_red(clmi_ctx *c, const AdjRun &r, double out[kAdjRed])
{
  HIPCHK(c, hipMemcpyAsync(out, r.red, sizeof(double) * kAdjRed, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CLMI_OK;
}

}  // namespace

extern "C" {

int clmi_adjoint_lf(clmi_ctx *c, int nsteps, int checkpoint_every, int bg_from_fields,
                    const double *const *seed_fld, const double *const *seed_pi, const double seed_bg[2],
                    double *const *grad_fld, double *const *grad_pi, double *grad_coeff, clmi_background *grad_bg)
{
  CHK(check_ctx(c));
  const int nf = c->cfg.n_scalars;
  if (nf < 1 || c->cfg.n_cscalars != 0) return fail(c, CLMI_E_INVALID, "the adjoint covers scalar-singlet models");
  if (nsteps < 1) return fail(c, CLMI_E_INVALID, "nsteps must be >= 1");
  if (!seed_fld || !seed_pi || !seed_bg || !grad_fld || !grad_pi || !grad_coeff)
    return fail(c, CLMI_E_INVALID, "null argument");
  for (int f = 0; f < nf; ++f)
    if (!seed_fld[f] || !seed_pi[f] || !grad_fld[f] || !grad_pi[f]) return fail(c, CLMI_E_INVALID, "null array");
  const int M = checkpoint_every > 0 ? checkpoint_every
                                     : std::max(1, (int)std::lround(std::sqrt((double)nsteps)));
  const bool ex = c->cfg.expansion != 0;
  const double dt = c->cfg.dt, al = c->cfg.alpha;
  const double V = (double)c->V;
  const ModelConsts &mc = c->mc;
  CHK(flush(c));

  AdjRun run;
  run.c = c;
  const int ncp = (nsteps + M - 1) / M;
  run.checkpoints.resize(ncp);
  run.segment.resize(M);
  run.seg_view.resize(M);
  {
    const size_t sb = adj_state_bytes(c);
    HIPCHK(c, hipMalloc(&run.pool, sb * (size_t)(ncp + M + 1)));
    char *p = (char *)run.pool;
    for (auto &s : run.checkpoints) adj_carve(c, p, s);
    for (auto &s : run.segment) adj_carve(c, p, s);
    adj_carve(c, p, run.final_state);
  }
  HIPCHK(c, hipMalloc(&run.lam, sizeof(double) * 2 * nf * c->V));
  HIPCHK(c, hipMalloc(&run.partials, sizeof(double) * kAdjMaxRows * kAdjRed));
  HIPCHK(c, hipMalloc(&run.red, sizeof(double) * kAdjRed));
  HIPCHK(c, hipMalloc(&run.sc, sizeof(AdjScalars)));
  HIPCHK(c, hipMalloc(&run.tape, sizeof(clmi_background) * (size_t)(nsteps + 1)));

  // ---- forward: initial background, lf_init, nsteps steps, sync; record the background tape ----
  std::vector<clmi_background> tape(nsteps + 1);
  clmi_background b_init{}, b_fin{};
  if (bg_from_fields) {
    // ModelInitializer: setAllAverages, then aDotI = sqrt(Friedmann RHS) (modelinitializer.h:71-80)
    double hc[3];
    CHK(clmi_hubble_constraint(c, hc));
    CHK(clmi_get_background(c, &b_init));
    b_init.aDotI = std::sqrt(hc[2]);
    CHK(clmi_set_background(c, &b_init));
  }
  CHK(clmi_get_background(c, &b_init));
  CHK(clmi_lf_init(c, 0));
  // Plain fused leapfrog steps, flushed only where a checkpoint needs the materialised state.  The tape
  // entry of step n is the background after storeFieldsAverages of step n - 1 (what a flush leaves);
  // the update that runs that operation writes it on the device (c->snap), and the tape is read back
  // once.  Without expansion no update runs and the background stays that of the start.
  static_assert(offsetof(DevState, bg) == 0, "DevState must start with the background");
  CHK(flush(c));
  HIPCHK(c, hipMemcpyAsync(run.tape, c->dstate, sizeof(clmi_background), hipMemcpyDeviceToDevice, c->stream));
  for (int n = 0; n < nsteps; ++n) {
    if (n % M == 0) {
      if (c->pending_drift) {
        c->snap = run.tape + n;
        CHK(flush(c));
      }
      CHK(adj_save(c, run.checkpoints[n / M]));
    }
    if (c->pending_drift) c->snap = run.tape + n;
    CHK(clmi_lf_evolve(c, 1));
  }
  c->snap = run.tape + nsteps;
  CHK(flush(c));
  c->snap = nullptr;
  HIPCHK(c, hipMemcpyAsync(tape.data(), run.tape, sizeof(clmi_background) * (size_t)(nsteps + 1),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (!ex) {
    for (int n = 1; n <= nsteps; ++n) tape[n] = tape[0];
    HIPCHK(c, hipMemcpyAsync(run.tape, tape.data(), sizeof(clmi_background) * (size_t)(nsteps + 1),
                             hipMemcpyHostToDevice, c->stream));
  }
  CHK(clmi_lf_sync(c));
  CHK(clmi_get_background(c, &b_fin));
  CHK(adj_save(c, run.final_state));

  // ---- reverse ----
  AdjArgs a{};
  a.nf = nf;
  a.N = (uint32_t)c->N;
  a.V = c->V;
  a.inv_dx2 = 1.0 / (c->cfg.dx * c->cfg.dx);
  a.poly = c->spoly;
  a.partials = run.partials;
  a.red_out = run.red;
  a.no_tiled = (c->cfg.flags & CLMI_FLAG_NO_TILED) ? 1 : 0;
  for (int f = 0; f < nf; ++f) {
    a.lam_phi[f] = run.lam + (size_t)f * c->V;
    a.lam_pi[f] = run.lam + (size_t)(nf + f) * c->V;
    HIPCHK(c, hipMemcpyAsync(a.lam_phi[f], seed_fld[f], sizeof(double) * c->V, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(a.lam_pi[f], seed_pi[f], sizeof(double) * c->V, hipMemcpyHostToDevice, c->stream));
  }
  const int nt = c->spoly.nterms;
  std::vector<double> gc(nt, 0.0);
  double red[kAdjRed];
  // scalar adjoints of the current background state
  double lam_aI = seed_bg[0], lam_aDotSI = 0, lam_pi2SI = 0, lam_G = 0, lam_U = 0;

  // final sync (leapfrog.h:96-116): pi_T = pi + dt/2 kernel(phi_T); pi2AvI = <pi_T^2>;
  // aDotI = aDotSI + dt/2 K
  {
    const clmi_background &bT = tape[nsteps];
    double lam_pi2T = 0.0;
    if (ex) {
      lam_aDotSI = seed_bg[1];
      const double lamK = 0.5 * dt * seed_bg[1];
      const SFd d = sf_partials(mc, bT.aI, b_fin.pi2AvI, bT.grad2AvI, bT.potAvI);
      lam_pi2T = lamK * d.dpi2;
      lam_G = lamK * d.dG;
      lam_U = lamK * d.dU;
      lam_aI += lamK * d.da;
    }
    a.fwd = adj_view(c, run.final_state);
    HIPCHK(c, launch_adj_pi(a, c->f32, 0.0, lam_pi2T, c->stream));
    const double aK = std::pow(bT.aI, 1 + al), aV = std::pow(bT.aI, 3 + al);
    HIPCHK(c, launch_adj_kick(a, c->f32, 0.5 * dt, aK, aV, c->stream));
    CHK(adj_read_red(c, run, red));
    if (ex) lam_aI += 0.5 * dt * (red[0] * (1 + al) * std::pow(bT.aI, al) - red[1] * (3 + al) * std::pow(bT.aI, 2 + al));
    for (int t = 0; t < nt; ++t) gc[t] += -0.5 * dt * aV * red[2 + t];
  }

  // The background adjoint lives on the device for the sweep: one-thread chain kernels between the
  // lattice kernels replace the host arithmetic, so no step waits for the host.
  {
    AdjScalars h{};
    h.lam_aI = lam_aI;
    h.lam_aDotSI = lam_aDotSI;
    h.lam_pi2SI = lam_pi2SI;
    h.lam_G = lam_G;
    h.lam_U = lam_U;
    for (int t = 0; t < nt; ++t) h.gc[t] = gc[t];
    HIPCHK(c, hipMemcpyAsync(run.sc, &h, sizeof(AdjScalars), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  // Per reverse step n: the pi transpose of step n, then one pass with the kick transpose of step n
  // fused with the fields transpose of state n (which the original order runs first thing in step
  // n - 1), each followed by its scalar chain.  The fields transpose of the final state runs alone.
  //
  // A segment is recomputed with plain leapfrog steps from its checkpoint (a materialised state
  // phi_n0, pi_{n0-1/2}), without flushing the deferred drift: after step n the context holds
  // S_n = (phi_n, pi_{n+1/2}), which is what the step's transposes read -- phi_n for the kick and fields
  // transposes, pi_{n+1/2} for the pi transpose; the fields transpose's sum lam_phi . pi_{n-1/2} takes
  // the momentum of S_{n-1} (of the checkpoint for n = n0).  The scalar updates of the fused pass are
  // the reference's sequence, so S_n equals the forward run's state bit for bit.
  AdjArgs ad = a;
  ad.sc = run.sc;
  for (int k = ncp - 1; k >= 0; --k) {
    const int n0 = k * M, n1 = std::min(nsteps, n0 + M);
    // the previous segment left a deferred drift that belongs to states no longer needed: drop it
    // before the background call below, which would otherwise flush it into the restored fields
    c->pending_drift = false;
    CHK(adj_restore(c, run.checkpoints[k]));
    CHK(clmi_set_background(c, &tape[n0]));
    c->synced = (n0 == 0);
    {
      // Each drift + kick step writes its state straight into segment slot n - n0: the context's spare
      // buffers point there for the step.  The first step of a segment is a kick alone, which updates
      // the momenta in place, so that state stays in the context's own buffers, which no later step of
      // the segment writes.
      AdjBufferGuard guard(c);
      for (int n = n0; n < n1; ++n) {
        if (c->pending_drift) {
          const AdjState &q = run.segment[n - n0];
          for (int f = 0; f < nf; ++f) {
            c->fS[f][1 - c->cur_fS] = q.f[f];
            c->pS[f][1 - c->cur_pS] = q.p[f];
          }
        }
        CHK(clmi_lf_evolve(c, 1));
        run.seg_view[n - n0] = adj_view_cur(c);
      }
    }
    if (k == ncp - 1) {
      // storeFieldsAverages of the final phi (unchanged by the sync), then the drift's sum
      // lam_phi . pi_{T-1/2}
      ad.fwd = adj_view(c, run.final_state);
      ad.fwd_red = run.seg_view[n1 - 1 - n0];
      HIPCHK(c, launch_adj_fields(ad, c->f32, 0.0, 0.0, 0.0, false, c->stream));
      HIPCHK(c, launch_adj_chain_fields(run.sc, run.red, 0, run.tape, nsteps - 1, nsteps == 1 ? 0.5 : 1.0, V, mc, nt,
                                        ex, c->stream));
    }
    for (int n = n1 - 1; n >= n0; --n) {
      // drift and <pi^2> transposes of step n (pi_{n+1/2})
      ad.fwd = run.seg_view[n - n0];
      ad.fwd_red = AdjFields{};
      HIPCHK(c, launch_adj_pi(ad, c->f32, 0.0, 0.0, c->stream));
      // kick of step n on phi_n, with the fields transpose of phi_n for step n - 1
      ad.fwd_red = n > n0 ? run.seg_view[n - 1 - n0] : adj_view(c, run.checkpoints[k]);
      HIPCHK(c, launch_adj_kick_fields(ad, c->f32, n > 0, c->stream));
      HIPCHK(c, launch_adj_finish_chains(ad, run.sc, run.tape, n, n - 1 == 0 ? 0.5 : 1.0, V, mc, nt, ex, c->stream));
    }
  }
  c->pending_drift = false;
  {
    AdjScalars h{};
    HIPCHK(c, hipMemcpyAsync(&h, run.sc, sizeof(AdjScalars), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    lam_aI = h.lam_aI;
    lam_aDotSI = h.lam_aDotSI;
    lam_pi2SI = h.lam_pi2SI;
    lam_G = h.lam_G;
    lam_U = h.lam_U;
    for (int t = 0; t < nt; ++t) gc[t] = h.gc[t];
  }

  // lf_init (leapfrog.h:35): aDotSI = aDotI - dt/2 K(initial averages)
  clmi_background gb{};
  if (ex) {
    const double lam_aDotI0 = lam_aDotSI;
    const double lamK0 = -0.5 * dt * lam_aDotSI;
    const SFd d = sf_partials(mc, b_init.aI, b_init.pi2AvI, b_init.grad2AvI, b_init.potAvI);
    double lam_pi2I0 = lamK0 * d.dpi2;
    lam_G += lamK0 * d.dG;
    lam_U += lamK0 * d.dU;
    lam_aI += lamK0 * d.da;
    gb.aI = lam_aI;
    if (bg_from_fields) {
      // aDotI = sqrt(a^(2 alpha + 2)/3 (fStar/MPl)^2 rho), rho = pi2/2 a^-6 + G/2 a^-2 + U, with
      // pi2AvI = pi2AvSI = <pi^2>, grad2AvI = <|grad phi|^2>, potAvI = <V> of the initial fields
      const double a0 = b_init.aI, fm = mc.fStar / mc.MPl;
      const double lamRHS = lam_aDotI0 / (2.0 * b_init.aDotI);
      const double lamRho = lamRHS * std::pow(a0, 2 * al + 2) / 3.0 * fm * fm;
      const double lamP0 = lam_pi2I0 + lam_pi2SI + lamRho * 0.5 * std::pow(a0, -6.0);
      const double lamG0 = lam_G + lamRho * 0.5 * std::pow(a0, -2.0);
      const double lamU0 = lam_U + lamRho;
      a.fwd = adj_view(c, run.checkpoints[0]);
      HIPCHK(c, launch_adj_fields(a, c->f32, lamG0, lamU0, lamP0, true, c->stream));
      CHK(adj_read_red(c, run, red));
      for (int t = 0; t < nt; ++t) gc[t] += lamU0 * red[1 + t] / V;
    } else {
      gb.aDotI = lam_aDotI0;
      gb.pi2AvI = lam_pi2I0;
      gb.pi2AvSI = lam_pi2SI;
      gb.grad2AvI = lam_G;
      gb.potAvI = lam_U;
    }
  }
  for (int f = 0; f < nf; ++f) {
    HIPCHK(c, hipMemcpyAsync(grad_fld[f], a.lam_phi[f], sizeof(double) * c->V, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(grad_pi[f], a.lam_pi[f], sizeof(double) * c->V, hipMemcpyDeviceToHost, c->stream));
  }
  for (int t = 0; t < nt; ++t) grad_coeff[t] = gc[t];
  if (grad_bg) *grad_bg = gb;
  // leave the context at the end of the forward run
  CHK(adj_restore(c, run.final_state));
  CHK(clmi_set_background(c, &b_fin));
  c->synced = true;
  c->pending_drift = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CLMI_OK;
}

}  // extern "C"

// ---- ensembles ----------------------------------------------------------------------------------

struct clmi_ensemble {
  std::vector<clmi_ctx *> m;      // as given
  std::vector<clmi_ctx *> order;  // sorted by size class (row shape VZ = N / 64)
  int device = 0;
  hipStream_t stream = nullptr;
  RowsMember *d_rows[2] = {nullptr, nullptr};  // member tables (order) for the two buffer parities
  UpdMember *d_upd = nullptr;
  size_t group_bytes = 0;
  // A group: consecutive members of one size class (whose evolved arrays fit the byte budget, when
  // one is set).  Each group has its own stream; a captured graph holds two steps of every group as
  // parallel branches, so the size classes share the GPU.
  struct Group {
    int vz, first, count, blocks;
  };
  std::vector<Group> groups;
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<int> captured_blocks;
  int nf = 0;
  int pk = 0;  // potential form of the batched passes (1: every member has the compile-time lphi4 table)
  bool f32 = false, ex = false;
  std::string err;
};

namespace {

int efail(clmi_ensemble *e, int code, const std::string &msg)
{
  if (e)
    e->err = msg;
  else
    g_create_error = msg;
  return code;
}

#define EHIP(e, expr)                                                                                         \
  do {                                                                                                        \
    hipError_t e_ = (expr);                                                                                   \
    if (e_ != hipSuccess) return efail((e), CLMI_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

size_t state_bytes(const clmi_ctx *c) { return fused_pass_bytes(c) / 2; }

// Groups (formed once) and member tables for the current buffer parities (A) and the flipped ones
// (B).  Workgroup starts restart at 0 in every group.
int ens_tables(clmi_ensemble *e)
{
  const size_t n = e->order.size();
  if (e->groups.empty()) {
    size_t bytes = 0;
    for (size_t i = 0; i < n; ++i) {
      const clmi_ctx *c = e->order[i];
      const int vz = c->N / 64;
      const size_t b = state_bytes(c);
      if (e->groups.empty() || e->groups.back().vz != vz || bytes + b > e->group_bytes) {
        clmi_ensemble::Group g{};
        g.vz = vz;
        g.first = (int)i;
        e->groups.push_back(g);
        bytes = 0;
      }
      bytes += b;
      e->groups.back().count++;
    }
  }
  std::vector<RowsMember> ta(n), tb(n);
  std::vector<UpdMember> tu(n);
  for (auto &g : e->groups) {
    int start = 0;
    for (int i = g.first; i < g.first + g.count; ++i) {
      clmi_ctx *c = e->order[i];
      const int fi = c->cur_fS, pi = c->cur_pS;
      for (int par = 0; par < 2; ++par) {
        const int f0 = par ? 1 - fi : fi, p0 = par ? 1 - pi : pi;
        RowsMember r{};
        r.a = scalar_args(c, f0, 1 - f0, p0, 1 - p0, c->cfg.dt, c->cfg.dt, 0);
        r.XC = c->rplan.XC;
        r.G = c->rplan.G;
        r.blocks = c->rplan.blocks;
        r.start = start;
        r.vz = g.vz;
        (par ? tb : ta)[i] = r;
      }
      UpdMember u{};
      u.st = c->dstate;
      u.partials = c->partials;
      u.rows = c->rplan.rows;
      u.mc = c->mc;
      tu[i] = u;
      start += (c->rplan.blocks + 7) / 8 * 8;
    }
    g.blocks = start;
  }
  std::vector<int> blocks;
  for (const auto &g : e->groups) blocks.push_back(g.blocks);
  if (e->exec && blocks != e->captured_blocks) {  // a member's launch plan changed since the capture
    (void)hipGraphExecDestroy(e->exec);
    (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
  }
  while (e->side.size() < e->groups.size()) {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    EHIP(e, hipStreamCreate(&st));
    e->side.push_back(st);
    EHIP(e, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    e->join.push_back(ev);
  }
  if (!e->fork) EHIP(e, hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
  EHIP(e, hipMemcpyAsync(e->d_rows[0], ta.data(), sizeof(RowsMember) * n, hipMemcpyHostToDevice, e->stream));
  EHIP(e, hipMemcpyAsync(e->d_rows[1], tb.data(), sizeof(RowsMember) * n, hipMemcpyHostToDevice, e->stream));
  EHIP(e, hipMemcpyAsync(e->d_upd, tu.data(), sizeof(UpdMember) * n, hipMemcpyHostToDevice, e->stream));
  EHIP(e, hipStreamSynchronize(e->stream));
  return CLMI_OK;
}

// One steady-state leapfrog step of a group: fused drift+kick pass, then the background update.
int ens_step(clmi_ensemble *e, const clmi_ensemble::Group &g, int parity, hipStream_t st)
{
  EHIP(e, launch_scalar_rows_batch(e->nf, e->f32, e->pk, g.vz, true, true, e->ex, e->d_rows[parity] + g.first, g.count,
                                   g.blocks, st));
  if (e->ex)
    EHIP(e, launch_update_batch(e->d_upd + g.first, g.count,
                                make_ops({{OP_STORE_FIELDS, 0.0}, {OP_LF_MOMENTA, 0.0}}), st));
  return CLMI_OK;
}

// Capture two steps of every group, each group on its own stream (fork/join on the main stream).
int ens_capture(clmi_ensemble *e)
{
  EHIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeGlobal));
  int rc = CLMI_OK;
  hipError_t he = hipEventRecord(e->fork, e->stream);
  for (size_t i = 0; i < e->groups.size() && rc == CLMI_OK && he == hipSuccess; ++i) {
    he = hipStreamWaitEvent(e->side[i], e->fork, 0);
    if (he == hipSuccess) rc = ens_step(e, e->groups[i], 0, e->side[i]);
    if (rc == CLMI_OK && he == hipSuccess) rc = ens_step(e, e->groups[i], 1, e->side[i]);
    if (rc == CLMI_OK && he == hipSuccess) he = hipEventRecord(e->join[i], e->side[i]);
    if (rc == CLMI_OK && he == hipSuccess) he = hipStreamWaitEvent(e->stream, e->join[i], 0);
  }
  hipGraph_t gr = nullptr;
  const hipError_t ce = hipStreamEndCapture(e->stream, &gr);
  if (rc != CLMI_OK || he != hipSuccess || ce != hipSuccess) {
    if (gr) (void)hipGraphDestroy(gr);
    if (rc != CLMI_OK) return rc;
    return efail(e, CLMI_E_HIP, std::string("ensemble graph capture: ") +
                                     hipGetErrorString(he != hipSuccess ? he : ce));
  }
  e->graph = gr;
  EHIP(e, hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
  e->captured_blocks.clear();
  for (const auto &g : e->groups) e->captured_blocks.push_back(g.blocks);
  return CLMI_OK;
}

}  // namespace

extern "C" {

int clmi_ensemble_create(clmi_ctx *const *members, int n, size_t group_bytes, clmi_ensemble **out)
{
  if (!out) return efail(nullptr, CLMI_E_INVALID, "null output");
  *out = nullptr;
  if (!members || n < 1) return efail(nullptr, CLMI_E_INVALID, "no members");
  const clmi_ctx *c0 = members[0];
  if (!c0) return efail(nullptr, CLMI_E_INVALID, "null member");
  for (int i = 0; i < n; ++i) {
    const clmi_ctx *c = members[i];
    if (!c) return efail(nullptr, CLMI_E_INVALID, "null member");
    if (c->device != c0->device || c->f32 != c0->f32 || c->cfg.n_scalars != c0->cfg.n_scalars ||
        c->cfg.expansion != c0->cfg.expansion)
      return efail(nullptr, CLMI_E_INVALID, "members must share device, precision, field count and expansion");
    if (c->cfg.n_cscalars != 0 || !c->rowk || c->N > 192 || c->rplan.WPR != 1 || c->rplan.TY != 8 ||
        c->rplan.VZ * 64 != c->N)
      return efail(nullptr, CLMI_E_INVALID,
                   "ensemble members must be scalar-singlet lattices with N = 64, 128 or 192 (row pass)");
    for (int j = 0; j < i; ++j)
      if (members[j] == c) return efail(nullptr, CLMI_E_INVALID, "a context appears twice");
  }
  auto *e = new clmi_ensemble();
  e->m.assign(members, members + n);
  e->order = e->m;
  std::stable_sort(e->order.begin(), e->order.end(), [](const clmi_ctx *a, const clmi_ctx *b) { return a->N < b->N; });
  e->device = c0->device;
  e->nf = c0->cfg.n_scalars;
  e->f32 = c0->f32;
  e->pk = 1;  // the batched shapes are the 512-thread ones, where both precisions use it
  for (clmi_ctx *c : e->m)
    if (poly_kind(c->spoly, e->nf) != 1) e->pk = 0;
  e->ex = c0->cfg.expansion != 0;
  e->group_bytes = group_bytes ? group_bytes : SIZE_MAX;
  int rc = CLMI_OK;
  if (hipSetDevice(e->device) != hipSuccess || hipStreamCreate(&e->stream) != hipSuccess ||
      hipMalloc(&e->d_rows[0], sizeof(RowsMember) * n) != hipSuccess ||
      hipMalloc(&e->d_rows[1], sizeof(RowsMember) * n) != hipSuccess ||
      hipMalloc(&e->d_upd, sizeof(UpdMember) * n) != hipSuccess)
    rc = efail(nullptr, CLMI_E_NOMEM, "ensemble allocation failed");
  if (rc != CLMI_OK) {
    clmi_ensemble_destroy(e);
    return rc;
  }
  *out = e;
  return CLMI_OK;
}

void clmi_ensemble_destroy(clmi_ensemble *e)
{
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  for (hipStream_t st : e->side) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : e->join) (void)hipEventDestroy(ev);
  if (e->fork) (void)hipEventDestroy(e->fork);
  for (auto *p : e->d_rows)
    if (p) (void)hipFree(p);
  if (e->d_upd) (void)hipFree(e->d_upd);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

const char *clmi_ensemble_last_error(const clmi_ensemble *e) { return e ? e->err.c_str() : g_create_error.c_str(); }

int clmi_ensemble_groups(const clmi_ensemble *e) { return e ? (int)e->groups.size() : 0; }

int clmi_ensemble_autotune(clmi_ensemble *e, int reps, char *report, size_t report_len)
{
  if (!e) return efail(nullptr, CLMI_E_INVALID, "null ensemble");
  if (report && report_len) report[0] = '\0';
  if (reps < 1) reps = 1;
  EHIP(e, hipSetDevice(e->device));
  for (clmi_ctx *c : e->m) EHIP(e, hipStreamSynchronize(c->stream));
  CHK(ens_tables(e));
  hipEvent_t t0 = nullptr, t1 = nullptr;
  EHIP(e, hipEventCreate(&t0));
  EHIP(e, hipEventCreate(&t1));
  std::string rep;
  int rc = CLMI_OK;
  // Each group's batched fused pass from the current buffers into the spare ones (parity-0 table),
  // timed per x-chunk; no parity flip and no background update, so the members' state is unchanged.
  for (size_t gi = 0; gi < e->groups.size() && rc == CLMI_OK; ++gi) {
    const clmi_ensemble::Group g0 = e->groups[gi];
    double best = 1e300;
    std::vector<RowsPlan> best_plans;
    std::string line = "# group VZ=" + std::to_string(g0.vz) + " members=" + std::to_string(g0.count);
    for (int xc : {8, 16, 32, 64}) {
      std::vector<RowsPlan> plans;
      bool ok = true;
      for (int i = g0.first; i < g0.first + g0.count && ok; ++i) {
        const clmi_ctx *c = e->order[i];
        RowsPlan rp;
        ok = scalar_rows_plan(c->N, c->cfg.n_scalars, c->f32, xc, c->rplan.MM, rows_shape(c->cfg.flags), &rp);
        plans.push_back(rp);
      }
      if (!ok) continue;
      for (int i = g0.first; i < g0.first + g0.count; ++i) e->order[i]->rplan = plans[i - g0.first];
      if ((rc = ens_tables(e)) != CLMI_OK) break;
      const clmi_ensemble::Group &g = e->groups[gi];
      auto launch = [&]() {
        return launch_scalar_rows_batch(e->nf, e->f32, e->pk, g.vz, true, true, e->ex, e->d_rows[0] + g.first, g.count,
                                        g.blocks, e->stream);
      };
      hipError_t he = launch();
      if (he == hipSuccess) he = hipEventRecord(t0, e->stream);
      for (int r = 0; r < reps && he == hipSuccess; ++r) he = launch();
      if (he == hipSuccess) he = hipEventRecord(t1, e->stream);
      if (he == hipSuccess) he = hipEventSynchronize(t1);
      float ms = 0.0f;
      if (he == hipSuccess) he = hipEventElapsedTime(&ms, t0, t1);
      if (he != hipSuccess) {
        rc = efail(e, CLMI_E_HIP, std::string("ensemble autotune: ") + hipGetErrorString(he));
        break;
      }
      const double us = 1e3 * ms / reps;
      line += " XC=" + std::to_string(xc) + ":" + std::to_string(us);
      if (us < best) {
        best = us;
        best_plans = plans;
      }
    }
    if (rc == CLMI_OK && !best_plans.empty())
      for (int i = g0.first; i < g0.first + g0.count; ++i) e->order[i]->rplan = best_plans[i - g0.first];
    rep += line + " -> XC=" + std::to_string(best_plans.empty() ? 0 : best_plans[0].XC) + "\n";
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (rc != CLMI_OK) return rc;
  CHK(ens_tables(e));
  if (report && report_len) {
    std::strncpy(report, rep.c_str(), report_len - 1);
    report[report_len - 1] = '\0';
  }
  return CLMI_OK;
}

int clmi_ensemble_lf_evolve(clmi_ensemble *e, int nsteps)
{
  if (!e) return efail(nullptr, CLMI_E_INVALID, "null ensemble");
  if (nsteps < 0) return efail(e, CLMI_E_INVALID, "nsteps < 0");
  if (nsteps == 0) return CLMI_OK;
  EHIP(e, hipSetDevice(e->device));
  // members that are synced or fully drifted take their first step on their own
  bool steady = true;
  for (clmi_ctx *c : e->m) steady = steady && c->pending_drift && !c->synced && !c->prof;
  if (!steady) {
    for (clmi_ctx *c : e->m) {
      if (c->prof) return efail(e, CLMI_E_STATE, "disable profiling on ensemble members");
      if (clmi_lf_evolve(c, 1) != CLMI_OK) return efail(e, CLMI_E_HIP, std::string("member step: ") + c->err);
    }
    --nsteps;
  }
  for (clmi_ctx *c : e->m) EHIP(e, hipStreamSynchronize(c->stream));
  CHK(ens_tables(e));
  // two steps (both buffer parities) of every group per graph replay
  if (nsteps >= 2 && !e->exec) CHK(ens_capture(e));
  for (int k = 0; k + 1 < nsteps; k += 2) EHIP(e, hipGraphLaunch(e->exec, e->stream));
  if (nsteps % 2)
    for (const auto &g : e->groups) CHK(ens_step(e, g, 0, e->stream));
  EHIP(e, hipStreamSynchronize(e->stream));
  for (clmi_ctx *c : e->m) {
    if (nsteps % 2) {
      c->cur_fS = 1 - c->cur_fS;
      c->cur_pS = 1 - c->cur_pS;
    }
    c->pending_drift = true;
    c->synced = false;
  }
  return CLMI_OK;
}

int clmi_profile_enable(clmi_ctx *c, int on)
{
  CHK(check_ctx(c));
  CHK(prof_flush(c));
  c->prof = on > 0;
  c->prof_every = on > 0 ? on : 1;
  c->prof_count = 0;
  c->prof_ms = 0.0;
  c->prof_launches = 0;
  return CLMI_OK;
}

namespace {
constexpr double kTuneKeep = 0.98;  // another x-chunk replaces the current one if at least 2% faster

// x-chunk of the tiled charged-scalar + U(1) pass: fused passes from the current buffers into the spare
// ones (no parity flip, no background update, so the context state is unchanged) for every valid XC;
// the partial rows do not depend on XC, so neither do the results.
int autotune_gauge(clmi_ctx *c, int reps, std::string &rep)
{
  const GaugePassArgs a =
      gauge_args(c, c->cur_fG, 1 - c->cur_fG, c->cur_pG, 1 - c->cur_pG, c->cfg.dt, 0.0, 0);
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  double best = 1e300, cur = 1e300;
  int best_xc = c->gXC, best_blocks = c->blocks_g;
  std::string lines;
  hipError_t e = hipSuccess;
  for (int xc = 8; xc <= c->N && e == hipSuccess; xc *= 2) {
    int blocks = 0, rows = 0;
    if (!gauge_tiled_shape(c->N, c->gTY, xc, &blocks, &rows) || rows != c->rows_g) continue;
    auto launch = [&]() { return launch_gauge_tiled(c->f32, c->gTY, true, true, true, a, blocks, xc, c->stream); };
    e = launch();
    if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
    for (int i = 0; i < reps && e == hipSuccess; ++i) e = launch();
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) break;
    const double us = 1e3 * ms / reps;
    lines += "gauge XC=" + std::to_string(xc) + " " + std::to_string(us) + "\n";
    if (xc == c->gXC) cur = us;
    if (us < best) {
      best = us;
      best_xc = xc;
      best_blocks = blocks;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return fail(c, CLMI_E_HIP, std::string("autotune: ") + hipGetErrorString(e));
  // Back-to-back passes over one buffer pair rank the x-chunks to within a few per cent of what they
  // take inside the leapfrog loop, so the current x-chunk stays unless another is clearly faster.
  if (best < kTuneKeep * cur) {
    c->gXC = best_xc;
    c->blocks_g = best_blocks;
  }
  best_xc = c->gXC;
  rep += "# kernel k_gauge_tiled TY=" + std::to_string(c->gTY) + " XC=" + std::to_string(best_xc) + "\n" + lines;
  return CLMI_OK;
}
}  // namespace

int clmi_autotune(clmi_ctx *c, int reps, char *report, size_t report_len)
{
  CHK(check_ctx(c));
  if (report && report_len) report[0] = '\0';
  if (reps < 1) reps = 1;
  if ((!c->tiled && !c->rowk) || c->cfg.n_scalars == 0) {
    if (!c->gtiled) return CLMI_OK;
    CHK(flush(c));
    std::string rep;
    CHK(autotune_gauge(c, reps, rep));
    if (report && report_len) {
      std::strncpy(report, rep.c_str(), report_len - 1);
      report[report_len - 1] = '\0';
    }
    return CLMI_OK;
  }
  CHK(flush(c));
  // fused drift+kick+reduce passes from the current buffers into the spare ones; parities are not
  // flipped and no background update runs, so the context state is unchanged
  const ScalarPassArgs a =
      scalar_args(c, c->cur_fS, 1 - c->cur_fS, c->cur_pS, 1 - c->cur_pS, c->cfg.dt, 0.0, 0);
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  std::string rep;
  double best = 1e300;
  TiledPlan best_plan = c->tplan;
  RowsPlan best_rplan = c->rplan;
  int rc = CLMI_OK;
  // row kernel: memory modes 1 (NT stores), 3 (+ NT momentum loads), 7 (+ NT loads of own field rows),
  // 15 (as 7, but ordinary loads for the tile's edge rows); tiled kernel: NT stores off / on
  const std::vector<int> modes = c->rowk ? std::vector<int>{1, 3, 7, 15} : std::vector<int>{0, 1};
  for (int nt : modes) {
    if (rc != CLMI_OK) break;
    for (int xc : {4, 8, 16, 32, 64}) {
      TiledPlan p;
      RowsPlan rp;
      if (c->rowk ? !scalar_rows_plan(c->N, c->cfg.n_scalars, c->f32, xc, nt, rows_shape(c->cfg.flags), &rp)
                  : !scalar_tiled_plan(c->N, c->cfg.n_scalars, xc, nt != 0, &p))
        continue;
      auto launch = [&]() {
        return c->rowk ? launch_scalar_rows(c->cfg.n_scalars, c->f32, true, true, true, a, rp, c->stream)
                       : launch_scalar_tiled(c->cfg.n_scalars, c->f32, true, true, true, a, p, c->stream);
      };
      hipError_t e = launch();
      if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
      for (int i = 0; i < reps && e == hipSuccess; ++i) e = launch();
      if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) {
        rc = fail(c, CLMI_E_HIP, std::string("autotune: ") + hipGetErrorString(e));
        break;
      }
      const double us = 1e3 * ms / reps;
      rep += "XC=" + std::to_string(xc) + (c->rowk ? " MM=" : " NT=") + std::to_string(nt) + " " +
             std::to_string(us) + "\n";
      if (us < best) {
        best = us;
        best_plan = p;
        best_rplan = rp;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != CLMI_OK) return rc;
  if (c->rowk) {
    c->rplan = best_rplan;
    rep = "# kernel k_scalar_rows VZ=" + std::to_string(best_rplan.VZ) + " WPR=" + std::to_string(best_rplan.WPR) +
          " TY=" + std::to_string(best_rplan.TY) + " XC=" + std::to_string(best_rplan.XC) +
          " MM=" + std::to_string(best_rplan.MM) +
          " PK=" + std::to_string(scalar_rows_pk(c->cfg.n_scalars, c->f32, c->spoly, best_rplan)) + "\n" + rep;
  } else {
    c->tplan = best_plan;
    rep = "# kernel k_scalar_tiled XC=" + std::to_string(best_plan.XC) + " NT=" + std::to_string((int)best_plan.NT) +
          " PK=" + std::to_string(poly_kind(c->spoly, c->cfg.n_scalars)) + "\n" + rep;
  }
  if (c->gtiled) CHK(autotune_gauge(c, reps, rep));
  if (report && report_len) {
    std::strncpy(report, rep.c_str(), report_len - 1);
    report[report_len - 1] = '\0';
  }
  return CLMI_OK;
}

int clmi_plan_report(const clmi_ctx *c, char *report, size_t report_len)
{
  if (!c) return CLMI_E_INVALID;
  if (!report || report_len == 0) return CLMI_E_INVALID;
  std::string rep;
  if (c->cfg.n_scalars > 0) {
    if (c->rowk)
      rep += "k_scalar_rows VZ=" + std::to_string(c->rplan.VZ) + " WPR=" + std::to_string(c->rplan.WPR) +
             " TY=" + std::to_string(c->rplan.TY) + " XC=" + std::to_string(c->rplan.XC) +
             " MM=" + std::to_string(c->rplan.MM) +
             " PK=" + std::to_string(scalar_rows_pk(c->cfg.n_scalars, c->f32, c->spoly, c->rplan)) + "\n";
    else if (c->tiled)
      rep += "k_scalar_tiled XC=" + std::to_string(c->tplan.XC) + " G=" + std::to_string(c->tplan.G) +
             " NT=" + std::to_string((int)c->tplan.NT) + " PK=" + std::to_string(poly_kind(c->spoly, c->cfg.n_scalars)) +
             "\n";
    else
      rep += "k_scalar_pass\n";
  }
  if (c->cfg.n_cscalars > 0) {
    if (c->gtiled)
      rep += "k_gauge_tiled TY=" + std::to_string(c->gTY) + " XC=" + std::to_string(c->gXC) + "\n";
    else
      rep += "k_gauge_pass\n";
  } else if (c->cfg.n_scalars > 0) {
    // the kick transpose of clmi_adjoint_lf
    AdjArgs a{};
    a.nf = c->cfg.n_scalars;
    a.N = (uint32_t)c->N;
    a.V = c->V;
    a.poly = c->spoly;
    a.no_tiled = (c->cfg.flags & CLMI_FLAG_NO_TILED) ? 1 : 0;
    int xc = 0, blocks = 0;
    if (adj_kick_tiled_plan(a, &xc, &blocks))
      rep += "k_adj_kick_tiled XC=" + std::to_string(xc) + " blocks=" + std::to_string(blocks) + "\n";
    else
      rep += "k_adj_kick\n";
  }
  std::strncpy(report, rep.c_str(), report_len - 1);
  report[report_len - 1] = '\0';
  return rep.size() < report_len ? CLMI_OK : CLMI_E_INVALID;
}

int clmi_copy_calibration(clmi_ctx *c, size_t bytes, int iters)
{
  CHK(check_ctx(c));
  if (bytes < 16 || bytes > ((size_t)16 << 30) || iters < 1) return fail(c, CLMI_E_INVALID, "bad copy size");
  bytes &= ~(size_t)15;
  void *s = nullptr, *d = nullptr;
  HIPCHK(c, hipMalloc(&s, bytes));
  if (hipMalloc(&d, bytes) != hipSuccess) {
    (void)hipFree(s);
    return fail(c, CLMI_E_NOMEM, "copy buffer allocation failed");
  }
  hipError_t e = hipMemsetAsync(s, 0, bytes, c->stream);
  for (int i = 0; i < iters && e == hipSuccess; ++i) e = launch_copy(s, d, bytes, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(s);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(c, CLMI_E_HIP, std::string("copy calibration: ") + hipGetErrorString(e));
  return CLMI_OK;
}

int clmi_profile_read(clmi_ctx *c, double *total_ms, int64_t *launches, double *bytes_per_launch)
{
  CHK(check_ctx(c));
  CHK(prof_flush(c));
  if (total_ms) *total_ms = c->prof_ms;
  if (launches) *launches = c->prof_launches;
  if (bytes_per_launch) *bytes_per_launch = fused_pass_bytes(c);
  return CLMI_OK;
}

}  // extern "C"
