/*
 * clmi.h -- C ABI of the MI355X-native CosmoLattice lattice-evolution library (libclmi.so).
 *
 * The reference evaluates its evolution as C++ templates (CosmoInterface over TempLat); there is no
 * C ABI in the reference.  Each entry point below replaces one reference interface, cited as
 * path:line relative to the CosmoLattice source tree.  No C++ or torch types cross this boundary:
 * plain pointers, sizes and POD structs only.  Every function returns CLMI_OK (0) on success or a
 * negative CLMI_E_* code; clmi_last_error() then holds a message.  No exception crosses the ABI.
 *
 * Data format at the boundary: every real lattice array is N*N*N doubles, dense row-major
 * [x][y][z] with z fastest, no ghost layers, periodic.  Direction i = 1, 2, 3 of the reference is
 * x, y, z here.  Host buffers are borrowed for the duration of the call only.
 *
 * Threading: one host thread drives a context.  All work is ordered on the context's own HIP
 * stream.  Calls on the same context are not re-entrant; several contexts may coexist.
 */
#ifndef CLMI_H
#define CLMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLMI_OK 0
#define CLMI_E_INVALID (-1)   /* bad argument or inconsistent configuration */
#define CLMI_E_HIP (-2)       /* a HIP runtime / hipFFT call failed */
#define CLMI_E_STATE (-3)     /* call not valid in the context's current state */
#define CLMI_E_NOMEM (-4)     /* device allocation failed */

#define CLMI_MAX_SCALARS 4
#define CLMI_MAX_TERMS 8

/* Device precision modes.  CLMI_F64: every array and every operation in FP64 (the drop-in mode).
 * CLMI_F32_BULK: fields stored as FP32 hi/lo pairs, momenta FP32, stencil/kick/drift in FP32;
 * volume reductions, scale-factor update, spectra bin sums and constraint checks in FP64. */
#define CLMI_F64 0
#define CLMI_F32_BULK 1

/* clmi_config.flags: use the generic one-site-per-thread lattice pass even where the tiled
 * (x-marching, LDS) pass applies.  Both give bit-identical results; the flag exists for testing. */
#define CLMI_FLAG_NO_TILED 1
/* clmi_config.flags: the alternative workgroup shape of the full-row lattice pass (N = 256).  Fields
 * are bit-identical to the default shape; volume sums differ in summation order only.  Used to
 * compare shapes. */
#define CLMI_FLAG_SHAPE_ALT 2
/* clmi_config.flags: 4-row workgroup tiles of the full-row lattice pass (N = 256, two fields),
 * combinable with CLMI_FLAG_SHAPE_ALT.  Same results as above. */
#define CLMI_FLAG_SHAPE_TY4 4
/* clmi_config.flags: 8-row tiles (512-thread workgroups) in the tiled charged-scalar + U(1) pass
 * instead of 4-row tiles.  Fields bit-identical; volume sums differ in summation order only. */
#define CLMI_FLAG_GAUGE_TY8 16
/* clmi_config.flags: run velocity Verlet as two passes per stage.  By default, for scalar-singlet
 * models on the full-row or tiled lattice pass, the second half-kick of each stage and the first half-kick of
 * the next run in one pass (they read the same fields with the same coefficients); the results are
 * bit-identical either way.  The flag exists for testing. */
#define CLMI_FLAG_NO_DOUBLE_KICK 32
/* clmi_config.flags: memory mode m of the full-row lattice pass (0 = the library's default, 15; 1 =
 * non-temporal stores, 3 = also non-temporal loads of the momenta, 7 = also of the workgroup's own
 * field rows, 15 = as 7 but ordinary loads for the tile's first and last rows).  Results do not
 * depend on it; clmi_autotune chooses one.  This flag, like the shape flags, also selects the
 * full-row pass where another pass is the default (FP64 at 256^3 runs the tiled pass). */
#define CLMI_FLAG_MEMMODE(m) ((m) << 8)

/* Field kinds, numbered as the reference's FieldsNumbering (fieldsnumbering.h:20-40). */
#define CLMI_FLD_S 0
#define CLMI_PI_S 1
#define CLMI_FLD_CS 2
#define CLMI_PI_CS 3
#define CLMI_FLD_U1 6
#define CLMI_PI_U1 7
/* Derived kind (not a FieldsNumbering value) for clmi_norm_averages and clmi_spectrum_bins: the
 * magnetic field of the U(1) field, B_i = (d+_j A_k - d+_k A_j)/dx with (i, j, k) cyclic and forward
 * differences d+_j f = f(x+j) - f(x), so |B|^2 = F_12^2 + F_13^2 + F_23^2, the field strength of
 * definitions/fieldfunctionals.h:140-153 (TempLat's magneticField as measurements/u1measurer.h:72,95-97
 * uses it).  comp selects B_1, B_2, B_3. */
#define CLMI_MAG_U1 64

/* Model and run configuration.  Replaces the model/run state the reference builds in
 * AbstractModel (abstractmodel.h:147-186), RunParameters (runparameters.h:60-236) and the
 * model's potentialTerms/potDeriv (models/lphi4.h:154-206, models/lphi4U1.h:117-129).
 *
 * The scalar-singlet potential is a polynomial in program units,
 *   V(phi) = sum_t term_coeff[t] * prod_f phi_f^term_exp[t][f],
 * and the complex-scalar potential is V(|phi|) = sum_k cs_coeff[k] * |phi|^cs_pow[k] (even powers).
 * The library differentiates both analytically. */
typedef struct clmi_config {
  int N;          /* sites per side */
  int n_scalars;  /* real scalar singlets, 0..CLMI_MAX_SCALARS */
  int n_cscalars; /* complex scalars, 0 or 1 */
  int n_u1;       /* U(1) gauge fields, 0 or 1 (requires n_cscalars == 1) */
  int precision;  /* CLMI_F64 or CLMI_F32_BULK */
  int expansion;  /* 1: self-consistent expansion (runparameters.h:66); 0: static a = aI */
  int device;     /* HIP device ordinal */
  int flags;      /* CLMI_FLAG_* bits */
  double dx, dt;
  double alpha, fStar, omegaStar, MPl; /* program units (modelparametersbase.h:20, scalefactorbase.h:18) */
  double gQ;                           /* effective U(1) charge g*Q (couplingsmanager.h:70-94) */
  int n_terms;
  int n_cs_terms;
  double term_coeff[CLMI_MAX_TERMS];
  int term_exp[CLMI_MAX_TERMS][CLMI_MAX_SCALARS];
  double cs_coeff[CLMI_MAX_TERMS];
  int cs_pow[CLMI_MAX_TERMS];
} clmi_config;

/* Homogeneous state: scale factor and cached volume averages, named as in the reference's
 * ScaleFactorBase (abstractmodel/scalefactorbase.h:20-21), ScalarBase, ComplexScalarBase
 * (abstractmodel/complexscalarbase.h:33-34), U1Base and AbstractModel (abstractmodel.h:142). */
typedef struct clmi_background {
  double aI, aSI, aIM, aDotI, aDotSI, aDotSIM;
  double pi2AvI, pi2AvSI, pi2AvSIM, grad2AvI;
  double CSpi2AvI, CSpi2AvSI, CSpi2AvSIM, CSgrad2AvI;
  double U1pi2AvI, U1pi2AvSI, U1pi2AvSIM, U1Mag2AvI;
  double potAvI, potAvSI;
} clmi_background;

/* Volume averages of the current device fields (no state change).  Replaces
 * Averages::pi2S/grad2S/pi2CS/grad2CS/B2U1/pi2U1 (definitions/averages.h:27-71) and
 * average(Potential::potential) (evolvers/leapfrog.h:276). */
typedef struct clmi_averages {
  double pi2S, grad2S, pot, pi2CS, grad2CS, pi2U1, B2U1;
  double reserved;
} clmi_averages;

typedef struct clmi_ctx clmi_ctx;

/* ---- lifetime ------------------------------------------------------------------------------ */

/* Allocate all device storage for the model (fields zeroed, background a = 1).  Replaces the
 * model constructor's field allocation (abstractmodel.h:147-156, abstractmodel/scalarbase.h:81-88). */
int clmi_create(const clmi_config *cfg, clmi_ctx **out);
void clmi_destroy(clmi_ctx *ctx);
/* Message of the last failed call on ctx, or of the last failed clmi_create when ctx is NULL. */
const char *clmi_last_error(const clmi_ctx *ctx);
/* Block until all work queued on the context's stream has finished. */
int clmi_synchronize(clmi_ctx *ctx);

/* ---- data movement -------------------------------------------------------------------------- */

/* Copy one real component array (N^3 doubles) to / from the device.  kind is a CLMI_FLD_* /
 * CLMI_PI_* value, index the field number, comp the component (0..1 complex, 0..2 gauge vector).
 * Replaces direct Field access in the reference's initializers and measurers.  download flushes a
 * deferred drift first, so the values are at the evolver's current time. */
int clmi_upload_field(clmi_ctx *ctx, int kind, int index, int comp, const double *host);
int clmi_download_field(clmi_ctx *ctx, int kind, int index, int comp, double *host);

/* Background scalars in and out (synchronous).  Replaces direct access to model.aI, model.aDotI,
 * model.pi2AvI, ... (abstractmodel/scalefactorbase.h:20-21). */
int clmi_set_background(clmi_ctx *ctx, const clmi_background *bg);
int clmi_get_background(clmi_ctx *ctx, clmi_background *bg);

/* ---- evolvers ----------------------------------------------------------------------------- */

/* LeapFrog constructor for a new run: synced = true, aDotSI = aDotI - dt/2 * K
 * (evolvers/leapfrog.h:29-38).  restart != 0 starts unsynced and leaves aDotSI alone. */
int clmi_lf_init(clmi_ctx *ctx, int restart);
/* nsteps x LeapFrog::evolve (evolvers/leapfrog.h:42-89).  The drift of the last step is kept
 * deferred and fused into the next step's kick; every observable call flushes it. */
int clmi_lf_evolve(clmi_ctx *ctx, int nsteps);
/* LeapFrog::sync (evolvers/leapfrog.h:96-116): half-kick momenta to integer time. */
int clmi_lf_sync(clmi_ctx *ctx);
/* Whether momenta are currently synced with the fields (leapfrog.h:284). */
int clmi_lf_is_synced(const clmi_ctx *ctx);
/* nsteps x VelocityVerlet::evolve with composition weights ws[0..nws-1]
 * (evolvers/velocityverlet.h:51-114; VV2 = {1.0}, velocityverletparameters.h:26-27). */
int clmi_vv_evolve(clmi_ctx *ctx, const double *ws, int nws, int nsteps);
/* nsteps x PositionVerlet::evolve with composition weights ws[0..nws-1]
 * (evolvers/positionverlet.h:47-111; PV2 = {1.0}, positionverletparameters.h:22-23).  Each stage is
 * one drift+kick pass and one drift pass. */
int clmi_pv_evolve(clmi_ctx *ctx, const double *ws, int nws, int nsteps);

/* ---- primitives (one lattice pass each) ----------------------------------------------------- */

/* All momenta += c * kernel(fields) at the current aI (LeapFrog::kickScalar/kickCS/kickU1Vector,
 * evolvers/leapfrog.h:131-155, with kernels evolvers/kernels/scalarsingletkernels.h:30-40,
 * complexscalarkernels.h:28-35, u1kernels.h:31-52). */
int clmi_kick(clmi_ctx *ctx, double c);
/* All fields += aSI^(alpha-3) * cdt * pi (scalars, complex scalar) and aSI^(alpha-1) * cdt * piU1
 * (LeapFrog::driftScalar/driftCS/driftU1Vector, evolvers/leapfrog.h:185-217). */
int clmi_drift(clmi_ctx *ctx, double cdt);
/* Averages::setAllAverages (definitions/averages.h:86-129), including its overwrite of the
 * half-step caches pi2AvSI, potAvSI, CSpi2AvSI, U1pi2AvSI. */
int clmi_set_all_averages(clmi_ctx *ctx);
/* Volume averages of the current fields; no background change. */
int clmi_volume_averages(clmi_ctx *ctx, clmi_averages *out);
/* ScaleFactorKernels::get from the current background (evolvers/kernels/scalefactorkernels.h:27-81). */
int clmi_scale_factor_kernel(clmi_ctx *ctx, double *K);
/* HubbleConstraint::get (definitions/hubbleconstraint.h:27-36), which calls Energies::rho and so
 * setAllAverages: out = {(LHS-RHS)/(LHS+RHS), LHS, RHS}. */
int clmi_hubble_constraint(clmi_ctx *ctx, double out[3]);
/* GaussLaws::checkU1 (definitions/gausslaws.h:31-53): out = {<|LHS-RHS|>/<sum of squares>, <|LHS|>,
 * <|RHS|>}. */
int clmi_gauss_u1(clmi_ctx *ctx, double out[3]);
/* MeansMeasurer::measure inputs (measurements/meansmeasurer.h:28-48) for one component, with the
 * array multiplied by scale: out = {<f>, <f^2>}. */
int clmi_field_means(clmi_ctx *ctx, int kind, int index, int comp, double scale, double out[2]);

/* Volume statistics of one real component: out = {<f>, <f^2>, <sum_i ((f(x+i) - f(x))/dx)^2>}, the
 * inputs of MeansMeasurer (measurements/meansmeasurer.h:28-48) and of the per-field gradient energy
 * (definitions/fieldfunctionals.h:52-67, energies.h:37-45). */
int clmi_field_stats(clmi_ctx *ctx, int kind, int index, int comp, double out[3]);
/* Norm statistics of a vector-valued field, the inputs MeansMeasurer::measure takes for norm(...)
 * (measurements/meansmeasurer.h:28-48): out = {<|v|>, <|v|^2>}.  kind CLMI_FLD_CS / CLMI_PI_CS:
 * |phi| = sqrt(phi_0^2 + phi_1^2) of complex scalar `index` and the same for its momentum
 * (measurements/complexscalarmeasurer.h:71-79).  kind CLMI_PI_U1: |pi_A| of U(1) field `index`;
 * kind CLMI_MAG_U1: |B| (measurements/u1measurer.h:70-73).  Scale factors (a^(alpha-3),
 * a^(alpha-1)) are the caller's. */
int clmi_norm_averages(clmi_ctx *ctx, int kind, int index, double out[2]);

/* Volume average of every scalar-potential monomial prod_f phi_f^term_exp[t][f] (without its
 * coefficient), t = 0..n_terms-1: the potential terms of EnergiesMeasurer (energiesmeasurer.h:85-90)
 * are sums of these with the model's coefficients. */
int clmi_monomial_averages(clmi_ctx *ctx, double *out);

/* ---- spectra and initial conditions ---------------------------------------------------------- */

/* Number of radial bins used by clmi_spectrum_bins for bin width dk: floor(sqrt(3)/2*N/dk) + 1
 * (runparameters.h:97-98 gives the reference's floor(sqrt(3)/2*N/dk); one more bin holds k = 0). */
int clmi_spectrum_nbins(const clmi_ctx *ctx, double dk);
/* R2C FFT of one component (unnormalised forward transform; kind CLMI_MAG_U1 transforms B_comp) and
 * deterministic |n|-shell binning
 * (measurements/powerspectrum.h:85-134 with TempLat's projectRadially).  Mode n (integer vector,
 * components in (-N/2, N/2]) goes to bin b = min(floor(|n|/dk + 1/2), nbins-1).  Outputs, each of
 * length nbins: sum over the full N^3 mode set of |f_n|^2, number of modes, and the mean |n| of
 * the modes in the bin (0 for an empty bin). */
int clmi_spectrum_bins(clmi_ctx *ctx, int kind, int index, int comp, double dk, int nbins, double *sum_f2,
                       double *count, double *mean_n);
/* clmi_spectrum_bins with every mode's |f_n|^2 weighted by |n|^npow inside the bin sum (npow = 0, 1
 * or 3): the projections PS_version 3 measures, |n| f for PS_type 1 and |n|^3 f for PS_type 2
 * (measurements/powerspectrum.h:111-131).  npow = 0 is clmi_spectrum_bins. */
int clmi_spectrum_bins_npow(clmi_ctx *ctx, int kind, int index, int comp, double dk, int nbins, int npow,
                            double *sum_f2, double *count, double *mean_n);

/* Gaussian random field initial conditions for scalar singlet `index` on the device
 * (initializers/fluctuationsgenerator.h:39-71,103-150, scalarsingletinitializer.h:41-77,283-284):
 * vacuum modes with frequency sqrt(k^2 + max(mass2, 0)), cut at k > kCutOff, zero mode zero,
 * pi_k = i w (fL - fR)/sqrt2 - aDot phi_k, then the homogeneous values fld0 and pi0 (program units)
 * are added.  Random numbers: Philox4x32-10 keyed by (seed, stream) with counter (mode, draw), so
 * the result does not depend on launch shape. */
int clmi_grf_scalar(clmi_ctx *ctx, int index, uint64_t seed, uint32_t stream, double mass2, double aDot,
                    double kIR, double kCutOff, double fld0, double pi0);

/* Complex scalar + U(1) initial conditions (U1Initializer::initializeRandomWithMatterU1,
 * initializers/u1initializer.h:77-117, with addFluctuationsCScalarFromPhases, :157-230): vacuum
 * fluctuations of both complex-scalar components built from Rayleigh amplitudes and phases
 * e00, e01, e10 with e11 = e01 e10 conj(e00), so that every mode carries zero charge; then the
 * homogeneous values fld0[c], pi0[c] (program units) are added.  With a U(1) field, A = 0 and
 * pi_A = conj(k-_i)/|k-|^2 * FFT(-dx^2 rho_Q), k-_i = 1 - exp(-2 pi i n_i/N), which satisfies the
 * Gauss constraint.  mass2[c]: effective mass^2 of component c (Potential::deriv2CS,
 * definitions/potential.h:83-88); the momenta of both components use the frequency of component 0,
 * as the reference does.  Random numbers: Philox4x32-10 keyed by seed, counter (mode, draw, stream). */
int clmi_grf_u1(clmi_ctx *ctx, uint64_t seed, uint32_t stream, const double mass2[2], double aDot, double kIR,
                double kCutOff, const double fld0[2], const double pi0[2]);

/* ---- ensembles --------------------------------------------------------------------------------- */

/* No counterpart in the reference (which runs one realisation per process).  Independent lattices --
 * same device, precision, number of scalar singlets and expansion setting; N = 64, 128 or 192 each;
 * potentials, dt, dx and initial conditions free -- advanced together.  Members of one size form a
 * group (split further when group_bytes > 0 caps a group's evolved arrays); every leapfrog step of a
 * group is one batched lattice launch plus one batched background launch, the groups run as
 * parallel branches of one captured HIP graph (two steps per replay).  Each member's fields and background are
 * bit-identical to clmi_lf_evolve on that member alone.  Members keep their own contexts for initial
 * conditions, observations and sync; the ensemble borrows them and must be destroyed before any
 * member. */
typedef struct clmi_ensemble clmi_ensemble;
int clmi_ensemble_create(clmi_ctx *const *members, int n, size_t group_bytes, clmi_ensemble **out);
/* Number of member groups (formed at the first clmi_ensemble_lf_evolve; 0 before). */
int clmi_ensemble_groups(const clmi_ensemble *ens);
/* Autotuner of the batched launches: for every group, times the batched fused pass at each x-chunk
 * on the members' live data (into their spare buffers, without an update, so no state changes) and
 * keeps the fastest.  Results do not depend on the choice.  A one-line summary per group goes into
 * report (may be NULL).  No counterpart in the reference: its ensembles are separate runs. */
int clmi_ensemble_autotune(clmi_ensemble *ens, int reps, char *report, size_t report_len);
void clmi_ensemble_destroy(clmi_ensemble *ens);
const char *clmi_ensemble_last_error(const clmi_ensemble *ens);
/* nsteps x LeapFrog::evolve on every member (evolvers/leapfrog.h:42-89), as clmi_lf_evolve. */
int clmi_ensemble_lf_evolve(clmi_ensemble *ens, int nsteps);

/* ---- adjoint ----------------------------------------------------------------------------------- */

/* No counterpart in the reference (a reverse-mode derivative of its leapfrog).  Runs, from the state
 * the context holds, the forward map
 *     [bg_from_fields: setAllAverages, aDotI = sqrt(Friedmann RHS)  (modelinitializer.h:71-80)]
 *     LeapFrog constructor, nsteps x evolve, sync          (evolvers/leapfrog.h:29-116)
 * and then its exact reverse-mode derivative.  Given the adjoint seeds of the outputs -- seed_fld[f]
 * = dL/dphi_f(T), seed_pi[f] = dL/dpi_f(T) (synced momenta), seed_bg = {dL/daI, dL/daDotI} at T --
 * it returns dL/dphi_f(0), dL/dpi_f(0), dL/dterm_coeff[t] for every potential monomial, and (when
 * bg_from_fields is 0) the derivatives with respect to the initial background scalars aI, aDotI,
 * pi2AvI, pi2AvSI, grad2AvI, potAvI in grad_bg (the other members zero).  Arrays are N^3 doubles on
 * the host, one per scalar singlet.  The forward states are recomputed from checkpoints every
 * checkpoint_every steps (<= 0: about sqrt(nsteps)); the adjoint fields are FP64 in both precision
 * modes.  On return the context holds the forward run's final (synced) state.  Scalar-singlet models
 * only; expansion on or off. */
int clmi_adjoint_lf(clmi_ctx *ctx, int nsteps, int checkpoint_every, int bg_from_fields,
                    const double *const *seed_fld, const double *const *seed_pi, const double seed_bg[2],
                    double *const *grad_fld, double *const *grad_pi, double *grad_coeff, clmi_background *grad_bg);

/* ---- tuning and timing ------------------------------------------------------------------------ */

/* No counterpart in the reference (tile shapes belong to this library's kernels).  Times every
 * candidate plan of the tiled lattice pass -- x-chunk length and store mode -- with `reps` fused
 * passes on the context's own data, without changing the context's fields or background, keeps
 * the fastest and writes one "XC=<n> NT=<0|1> <microseconds>" line per candidate into report
 * (if non-NULL).  With a charged scalar + U(1) it also times the x-chunk of that sector's tiled
 * pass ("gauge XC=<n> <microseconds>" lines).  All candidates give bit-identical results, including
 * the volume sums.  A no-op returning CLMI_OK where no tiled pass applies. */
int clmi_autotune(clmi_ctx *ctx, int reps, char *report, size_t report_len);

/* No counterpart in the reference.  Names the lattice passes the context runs, one line per sector
 * ("k_scalar_rows VZ=.. WPR=.. TY=.. XC=.. MM=.. PK=..", "k_scalar_tiled XC=.. G=.. NT=.. PK=..",
 * "k_scalar_pass"; "k_gauge_tiled TY=.. XC=..", "k_gauge_pass"; for scalar-singlet models the kick
 * transpose of clmi_adjoint_lf, "k_adj_kick_tiled XC=.. blocks=.." or "k_adj_kick"), into report
 * (NUL-terminated).  Launches nothing.  CLMI_E_INVALID if report is NULL or too short for the whole
 * text. */
int clmi_plan_report(const clmi_ctx *ctx, char *report, size_t report_len);

/* HIP-event timing of the dominant per-step lattice pass (the drift + kick pass, on the context
 * stream): on >= 1 times every on-th such pass (1: every pass; a larger period keeps the events'
 * own cost out of a throughput measurement), on <= 0 switches timing off.  Resets the totals. */
int clmi_profile_enable(clmi_ctx *ctx, int on);
/* Total milliseconds and number of timed launches since enabling; bytes_per_launch is the
 * algorithmic HBM traffic of one such pass (every evolved array read once and written once). */
int clmi_profile_read(clmi_ctx *ctx, double *total_ms, int64_t *launches, double *bytes_per_launch);
/* No counterpart in the reference: runs `iters` launches of a plain device-to-device copy kernel of
 * `bytes` bytes on the context stream, so that hardware byte counters (FETCH_SIZE / WRITE_SIZE)
 * can be scaled against a kernel whose traffic is known.  Allocates and frees its two buffers. */
int clmi_copy_calibration(clmi_ctx *ctx, size_t bytes, int iters);

/* ---- allocation-free variant ------------------------------------------------------------------ */

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): clmi_field_stats on a
 * caller-owned FP64 array d_field of N*N*N doubles (dense [x][y][z], z fastest, periodic, on the
 * current device), with the output d_out = {<f>, <f^2>, <sum_i ((f(x+i) - f(x))/dx)^2>} written to
 * caller-owned device memory (3 doubles), ordered on `stream` (a hipStream_t; NULL = the null stream).
 * The call returns once the work is enqueued.  Same kernels and summation order as clmi_field_stats,
 * so the results are bit-identical to it.  The partial sums go to a fixed 262 208-byte static device
 * buffer that each GPU holds until the process ends; the call allocates nothing.  Calls that overlap
 * on one GPU (other streams or host threads) share that buffer and race.  N must lie in [2, 1024];
 * otherwise, or for a null pointer, it returns CLMI_E_INVALID before launching anything, and
 * clmi_noalloc_last_error() holds the message. */
int clmi_field_stats_noalloc(const double *d_field, int N, double dx, double *d_out, void *stream);
const char *clmi_noalloc_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CLMI_H */
