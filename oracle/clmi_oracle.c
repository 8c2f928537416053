/*
 * clmi_oracle.c -- FP64 CPU restatement of CosmoLattice's lattice-evolution hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is compared with; it
 * is never linked into, called by or shipped with the product library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * The structure is deliberately "reference-structured": every reference expression that the
 * reference evaluates as one lattice sweep (a kick of one field, a drift, one volume average)
 * is one sweep here too, so the same code doubles as the CPU baseline that is timed beside the
 * GPU numbers.  Arithmetic follows the reference's expression order.
 *
 * Layout: every real lattice array is dense row-major [x][y][z], z fastest, no ghost layers,
 * periodic boundaries.  Spatial direction i = 1, 2, 3 is x, y, z.
 *
 * Reference files cited relative to the CosmoLattice source tree.
 *
 * Pinning (tests/test_oracle.py): closed forms and invariants pin the Laplacian, the potentials,
 * kick/drift (exact free-wave recurrence, time reversibility), the scale-factor update and the
 * reductions (O(dt^2) Friedmann and energy constraints), the U(1) sector (gauge invariance, Gauss
 * law, gQ = 0 limit) and the spectrum bins (Parseval, counts, single-mode bin).  The position
 * Verlet (orc_pv_evolve) is pinned by its exact free-wave recurrence (momenta differ from the
 * velocity Verlet's by 1/(1 - w^2 dt^2/4)), time reversibility of PV4/PV6, the dt^2 / dt^4 global
 * error orders of PV2 / PV4 and the O(dt^2) Friedmann constraint with expansion.
 * The Gauss check (orc_gauss_u1, all three outputs: the ratio with its mean-of-sum-of-squares
 * denominator, <|LHS|> and <|RHS|>) is pinned by checkerboard data on which every lattice average
 * is exact, with and without charge, and by uncharged plane waves; a forward difference, a wrong
 * sign of the charge term or any of the plausible wrong denominators fails one of them.
 * The spectrum prefactors and the GRF random stream live in oracle.py (their status is stated
 * there).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAX_SCALARS 4
#define ORC_MAX_TERMS 8

/* Background (homogeneous) state; names follow abstractmodel/scalefactorbase.h:20-21,
 * abstractmodel/scalarbase.h, complexscalarbase.h:33-34, u1base.h and abstractmodel.h:142. */
typedef struct orc_background {
  double aI, aSI, aIM, aDotI, aDotSI, aDotSIM;
  double pi2AvI, pi2AvSI, pi2AvSIM, grad2AvI;
  double CSpi2AvI, CSpi2AvSI, CSpi2AvSIM, CSgrad2AvI;
  double U1pi2AvI, U1pi2AvSI, U1pi2AvSIM, U1Mag2AvI;
  double potAvI, potAvSI;
} orc_background;

typedef struct orc_model {
  int N;          /* sites per side */
  int ns;         /* number of real scalar singlets */
  int ncs;        /* number of complex scalars (0 or 1) */
  int nu1;        /* number of U(1) gauge fields (0 or 1) */
  int expansion;  /* self-consistent expansion on/off (runparameters.h:66) */
  int synced;     /* leapfrog: momenta live at integer time (leapfrog.h:31) */
  double dx, dt;
  double alpha, fStar, omegaStar, MPl;
  /* Scalar-singlet potential V = sum_m c_m prod_f phi_f^{e_mf} (program units). */
  int nterms;
  double coeff[ORC_MAX_TERMS];
  int expo[ORC_MAX_TERMS][ORC_MAX_SCALARS];
  /* Complex-scalar potential V = sum_k d_k |phi|^{p_k}. */
  int ncsterms;
  double cscoeff[ORC_MAX_TERMS];
  int cspow[ORC_MAX_TERMS];
  double gQ; /* effective charge g*Q of the complex scalar under the U(1) (couplingsmanager.h:70-94) */
  double *fldS[ORC_MAX_SCALARS], *piS[ORC_MAX_SCALARS];
  double *fldCS[2], *piCS[2];
  double *fldU1[3], *piU1[3];
  orc_background bg;
} orc_model;

/* ------------------------------------------------------------------------------------------ */
/* Lattice helpers                                                                             */
/* ------------------------------------------------------------------------------------------ */

static inline size_t idx3(int N, int x, int y, int z) { return ((size_t)x * N + y) * N + z; }
static inline int wrap(int a, int N) { return a < 0 ? a + N : (a >= N ? a - N : a); }

/* neighbour of (x,y,z) in direction dir (0,1,2 = x,y,z) with step s (+1 / -1) */
static inline size_t nbr(int N, int x, int y, int z, int dir, int s)
{
  if (dir == 0) return idx3(N, wrap(x + s, N), y, z);
  if (dir == 1) return idx3(N, x, wrap(y + s, N), z);
  return idx3(N, x, y, wrap(z + s, N));
}

/* Volume average with per-row sums combined by Neumaier-compensated summation, so the oracle's
 * own rounding stays far below the 1e-10 parity bar at 256^3. */
typedef double (*site_fn)(const orc_model *m, int x, int y, int z);

static double lattice_average(const orc_model *m, site_fn f)
{
  const int N = m->N;
  double *rows = (double *)malloc(sizeof(double) * (size_t)N * N);
#pragma omp parallel for schedule(static)
  for (int xy = 0; xy < N * N; ++xy) {
    int x = xy / N, y = xy % N;
    double s = 0.0;
    for (int z = 0; z < N; ++z) s += f(m, x, y, z);
    rows[xy] = s;
  }
  double sum = 0.0, c = 0.0;
  for (int r = 0; r < N * N; ++r) {
    double v = rows[r];
    double t = sum + v;
    if (fabs(sum) >= fabs(v)) c += (sum - t) + v;
    else c += (v - t) + sum;
    sum = t;
  }
  free(rows);
  return (sum + c) / ((double)N * N * N);
}

/* ------------------------------------------------------------------------------------------ */
/* Potentials                                                                                  */
/* ------------------------------------------------------------------------------------------ */

static inline double ipow(double v, int e)
{
  double r = 1.0;
  for (int i = 0; i < e; ++i) r *= v;
  return r;
}

/* V(phi) = sum of potential terms (definitions/potential.h:47-50). */
static double scalar_potential(const orc_model *m, const double *phi)
{
  double v = 0.0;
  for (int t = 0; t < m->nterms; ++t) {
    double p = m->coeff[t];
    for (int f = 0; f < m->ns; ++f) p *= ipow(phi[f], m->expo[t][f]);
    v += p;
  }
  return v;
}

/* dV/dphi_n (definitions/potential.h:53-56 -> the model's potDeriv(Tag<n>)). */
static double scalar_potential_deriv(const orc_model *m, const double *phi, int n)
{
  double d = 0.0;
  for (int t = 0; t < m->nterms; ++t) {
    int e = m->expo[t][n];
    if (e == 0) continue;
    double p = m->coeff[t] * e;
    for (int f = 0; f < m->ns; ++f) p *= ipow(phi[f], f == n ? e - 1 : m->expo[t][f]);
    d += p;
  }
  return d;
}

double orc_scalar_potential(const orc_model *m, const double *phi) { return scalar_potential(m, phi); }
double orc_scalar_potential_deriv(const orc_model *m, const double *phi, int n)
{
  return scalar_potential_deriv(m, phi, n);
}

/* Complex scalar: V = sum_k d_k |phi|^{p_k}; derivative w.r.t. component c is
 * (dV/d|phi|) / |phi| * phi_c (definitions/potential.h:30-34, 75-79). */
static double cs_potential(const orc_model *m, double re, double im)
{
  double nrm = sqrt(re * re + im * im);
  double v = 0.0;
  for (int k = 0; k < m->ncsterms; ++k) v += m->cscoeff[k] * ipow(nrm, m->cspow[k]);
  return v;
}

static double cs_norm_deriv(const orc_model *m, double nrm)
{
  double d = 0.0;
  for (int k = 0; k < m->ncsterms; ++k)
    if (m->cspow[k] > 0) d += m->cscoeff[k] * m->cspow[k] * ipow(nrm, m->cspow[k] - 1);
  return d;
}

static double site_potential(const orc_model *m, int x, int y, int z)
{
  size_t i = idx3(m->N, x, y, z);
  double v = 0.0;
  if (m->ns > 0) {
    double phi[ORC_MAX_SCALARS];
    for (int f = 0; f < m->ns; ++f) phi[f] = m->fldS[f][i];
    v += scalar_potential(m, phi);
  }
  if (m->ncs > 0) v += cs_potential(m, m->fldCS[0][i], m->fldCS[1][i]);
  return v;
}

/* ------------------------------------------------------------------------------------------ */
/* Field functionals (definitions/fieldfunctionals.h)                                          */
/* ------------------------------------------------------------------------------------------ */

/* sum_n pi_n^2  (fieldfunctionals.h:196-200, averages.h:32-35) */
static double site_pi2S(const orc_model *m, int x, int y, int z)
{
  size_t i = idx3(m->N, x, y, z);
  double s = 0.0;
  for (int f = 0; f < m->ns; ++f) s += m->piS[f][i] * m->piS[f][i];
  return s;
}

/* sum_n sum_i ((phi_n(x+i) - phi_n(x)) / dx)^2 -- forward differences (fieldfunctionals.h:185-189,
 * averages.h:27-30) */
static double site_grad2S(const orc_model *m, int x, int y, int z)
{
  const int N = m->N;
  size_t i = idx3(N, x, y, z);
  double s = 0.0;
  for (int f = 0; f < m->ns; ++f) {
    const double *p = m->fldS[f];
    for (int d = 0; d < 3; ++d) {
      double g = (p[nbr(N, x, y, z, d, +1)] - p[i]) / m->dx;
      s += g * g;
    }
  }
  return s;
}

/* U(1) link for the complex scalar in direction d at site i: exp(-i dx gQ A_d(x))
 * (definitions/gaugederivatives.h:92-102). */
static inline void u1_link(const orc_model *m, size_t i, int d, double *c, double *s)
{
  double th = -m->dx * m->gQ * m->fldU1[d][i];
  *c = cos(th);
  *s = sin(th);
}

/* |pi_CS|^2 (fieldfunctionals.h:225-229) */
static double site_pi2CS(const orc_model *m, int x, int y, int z)
{
  size_t i = idx3(m->N, x, y, z);
  return m->piCS[0][i] * m->piCS[0][i] + m->piCS[1][i] * m->piCS[1][i];
}

/* sum_j |(U_j(x) phi(x+j) - phi(x)) / dx|^2 (fieldfunctionals.h:213-217,
 * gaugederivatives.h:62-65) */
static double site_grad2CS(const orc_model *m, int x, int y, int z)
{
  const int N = m->N;
  size_t i = idx3(N, x, y, z);
  double s = 0.0;
  for (int d = 0; d < 3; ++d) {
    size_t j = nbr(N, x, y, z, d, +1);
    double c = 1.0, sn = 0.0;
    if (m->nu1 > 0) u1_link(m, i, d, &c, &sn);
    double ur = c * m->fldCS[0][j] - sn * m->fldCS[1][j];
    double ui = c * m->fldCS[1][j] + sn * m->fldCS[0][j];
    double gr = (ur - m->fldCS[0][i]) / m->dx;
    double gi = (ui - m->fldCS[1][i]) / m->dx;
    s += gr * gr + gi * gi;
  }
  return s;
}

/* sum_i pi_A,i^2 (fieldfunctionals.h:282-286) */
static double site_pi2U1(const orc_model *m, int x, int y, int z)
{
  size_t i = idx3(m->N, x, y, z);
  return m->piU1[0][i] * m->piU1[0][i] + m->piU1[1][i] * m->piU1[1][i] + m->piU1[2][i] * m->piU1[2][i];
}

/* Field strength F_ij = (A_j(x+i) - A_j(x) - A_i(x+j) + A_i(x)) / dx, forward differences. */
static inline double field_strength(const orc_model *m, int x, int y, int z, int i, int j)
{
  const int N = m->N;
  size_t s = idx3(N, x, y, z);
  double dAj = m->fldU1[j][nbr(N, x, y, z, i, +1)] - m->fldU1[j][s];
  double dAi = m->fldU1[i][nbr(N, x, y, z, j, +1)] - m->fldU1[i][s];
  return (dAj - dAi) / m->dx;
}

/* F_21^2 + F_31^2 + F_32^2 (fieldfunctionals.h:266-274) */
static double site_B2U1(const orc_model *m, int x, int y, int z)
{
  double s = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < i; ++j) {
      double f = field_strength(m, x, y, z, i, j);
      s += f * f;
    }
  return s;
}

double orc_avg_pi2S(const orc_model *m) { return lattice_average(m, site_pi2S); }
double orc_avg_grad2S(const orc_model *m) { return lattice_average(m, site_grad2S); }
double orc_avg_potential(const orc_model *m) { return lattice_average(m, site_potential); }
double orc_avg_pi2CS(const orc_model *m) { return lattice_average(m, site_pi2CS); }
double orc_avg_grad2CS(const orc_model *m) { return lattice_average(m, site_grad2CS); }
double orc_avg_pi2U1(const orc_model *m) { return lattice_average(m, site_pi2U1); }
double orc_avg_B2U1(const orc_model *m) { return lattice_average(m, site_B2U1); }

/* Averages::setAllAverages (definitions/averages.h:86-129).  Note that it overwrites pi2AvSI
 * with the integer-time value; the leapfrog's next storeMomentaAverages picks that up. */
void orc_set_all_averages(orc_model *m)
{
  orc_background *b = &m->bg;
  b->potAvI = orc_avg_potential(m);
  b->potAvSI = b->potAvI;
  if (m->ns > 0) {
    b->pi2AvI = orc_avg_pi2S(m);
    b->pi2AvSI = b->pi2AvI;
    b->grad2AvI = orc_avg_grad2S(m);
  }
  if (m->nu1 > 0) {
    b->U1pi2AvI = orc_avg_pi2U1(m);
    b->U1pi2AvSI = b->U1pi2AvI;
    b->U1Mag2AvI = orc_avg_B2U1(m);
  }
  if (m->ncs > 0) {
    b->CSpi2AvI = orc_avg_pi2CS(m);
    b->CSpi2AvSI = b->CSpi2AvI;
    b->CSgrad2AvI = orc_avg_grad2CS(m);
  }
}

/* ------------------------------------------------------------------------------------------ */
/* Energies and scale factor                                                                    */
/* ------------------------------------------------------------------------------------------ */

/* ScaleFactorKernels::get for a minimally coupled model (evolvers/kernels/scalefactorkernels.h:
 * 27-81) with energies from definitions/energies.h:37-85. */
double orc_scale_factor_kernel(const orc_model *m)
{
  const orc_background *b = &m->bg;
  const double a = m->bg.aI;
  double Eks = 0, Egs = 0, Ekcs = 0, Egcs = 0, Eel = 0, Emag = 0;
  if (m->ns > 0) {
    Eks = 0.5 * b->pi2AvI * pow(a, -6);
    Egs = 0.5 * b->grad2AvI * pow(a, -2);
  }
  if (m->ncs > 0) {
    Ekcs = b->CSpi2AvI * pow(a, -6);
    Egcs = b->CSgrad2AvI * pow(a, -2);
  }
  if (m->nu1 > 0) {
    double w2 = (m->omegaStar / m->fStar) * (m->omegaStar / m->fStar);
    Eel = 0.5 * w2 * pow(a, -4) * b->U1pi2AvI;
    Emag = 0.5 * w2 * pow(a, -4) * b->U1Mag2AvI;
  }
  double fm = m->fStar / m->MPl;
  return pow(a, 2 * m->alpha + 1) / 3.0 * (fm * fm) *
         ((m->alpha - 2) * (Eks + Ekcs) + m->alpha * (Egs + Egcs) + (m->alpha - 1) * (Eel + Emag) +
          (m->alpha + 1) * b->potAvI);
}

/* Energies::rho after setAllAverages (definitions/energies.h:217-239). */
double orc_rho(orc_model *m)
{
  orc_set_all_averages(m);
  const orc_background *b = &m->bg;
  const double a = b->aI;
  double rho = b->potAvI;
  if (m->ns > 0) rho += 0.5 * b->pi2AvI * pow(a, -6) + 0.5 * b->grad2AvI * pow(a, -2);
  if (m->ncs > 0) rho += b->CSpi2AvI * pow(a, -6) + b->CSgrad2AvI * pow(a, -2);
  if (m->nu1 > 0) {
    double w2 = (m->omegaStar / m->fStar) * (m->omegaStar / m->fStar);
    rho += 0.5 * w2 * pow(a, -4) * (b->U1pi2AvI + b->U1Mag2AvI);
  }
  return rho;
}

/* HubbleConstraint::get (definitions/hubbleconstraint.h:27-36): out = {(L-R)/(L+R), L, R}. */
void orc_hubble_constraint(orc_model *m, double *out)
{
  double rho = orc_rho(m);
  double lhs = m->bg.aDotI * m->bg.aDotI;
  double fm = m->fStar / m->MPl;
  double rhs = pow(m->bg.aI, 2 * m->alpha + 2) / 3.0 * (fm * fm) * rho;
  out[0] = (lhs - rhs) / (lhs + rhs);
  out[1] = lhs;
  out[2] = rhs;
}

/* ------------------------------------------------------------------------------------------ */
/* Kicks and drifts -- one lattice sweep per reference expression                              */
/* ------------------------------------------------------------------------------------------ */

/* piS(n) += c * (aI^(1+alpha) * LatLapl(phi_n) - aI^(3+alpha) * dV/dphi_n)
 * (evolvers/kernels/scalarsingletkernels.h:30-40; LatLapl written as in
 * definitions/gaugederivatives.h:36-40). */
static void kick_scalar_field(orc_model *m, int n, double c)
{
  const int N = m->N;
  const double aK = pow(m->bg.aI, 1 + m->alpha);
  const double aV = pow(m->bg.aI, 3 + m->alpha);
  const double idx2 = 1.0 / (m->dx * m->dx);
  const double *phi = m->fldS[n];
  double *pi = m->piS[n];
#pragma omp parallel for schedule(static)
  for (int xy = 0; xy < N * N; ++xy) {
    int x = xy / N, y = xy % N;
    for (int z = 0; z < N; ++z) {
      size_t i = idx3(N, x, y, z);
      double lap = 0.0;
      for (int d = 0; d < 3; ++d) lap += phi[nbr(N, x, y, z, d, +1)] + phi[nbr(N, x, y, z, d, -1)] - 2 * phi[i];
      lap *= idx2;
      double f[ORC_MAX_SCALARS];
      for (int g = 0; g < m->ns; ++g) f[g] = m->fldS[g][i];
      pi[i] += c * (aK * lap - aV * scalar_potential_deriv(m, f, n));
    }
  }
}

/* piCS += c * (aI^(1+alpha) * covLapl(phi) - aI^(3+alpha)/2 * derivCS)
 * (evolvers/kernels/complexscalarkernels.h:28-35, definitions/gaugederivatives.h:48-53). */
static void kick_cs(orc_model *m, double c)
{
  const int N = m->N;
  const double aK = pow(m->bg.aI, 1 + m->alpha);
  const double aV = pow(m->bg.aI, 3 + m->alpha);
  const double idx2 = 1.0 / (m->dx * m->dx);
  const double *pr = m->fldCS[0], *pim = m->fldCS[1];
  double *out_r = (double *)malloc(sizeof(double) * (size_t)N * N * N);
  double *out_i = (double *)malloc(sizeof(double) * (size_t)N * N * N);
#pragma omp parallel for schedule(static)
  for (int xy = 0; xy < N * N; ++xy) {
    int x = xy / N, y = xy % N;
    for (int z = 0; z < N; ++z) {
      size_t i = idx3(N, x, y, z);
      double plus_r = 0.0, plus_i = 0.0, minus_r = 0.0, minus_i = 0.0;
      for (int d = 0; d < 3; ++d) {
        size_t jp = nbr(N, x, y, z, d, +1), jm = nbr(N, x, y, z, d, -1);
        double c1 = 1.0, s1 = 0.0, c2 = 1.0, s2 = 0.0;
        if (m->nu1 > 0) {
          u1_link(m, i, d, &c1, &s1);  /* U_d(x) */
          u1_link(m, jm, d, &c2, &s2); /* U_d(x-d) */
        }
        /* U_d(x) phi(x+d) */
        plus_r += c1 * pr[jp] - s1 * pim[jp];
        plus_i += c1 * pim[jp] + s1 * pr[jp];
        /* conj(U_d(x-d)) phi(x-d) */
        minus_r += c2 * pr[jm] + s2 * pim[jm];
        minus_i += c2 * pim[jm] - s2 * pr[jm];
      }
      double lap_r = (plus_r + minus_r - 6 * pr[i]) * idx2;
      double lap_i = (plus_i + minus_i - 6 * pim[i]) * idx2;
      double nrm = sqrt(pr[i] * pr[i] + pim[i] * pim[i]);
      double nd = cs_norm_deriv(m, nrm);
      double dr = nd / nrm * pr[i];
      double di = nd / nrm * pim[i];
      out_r[i] = aK * lap_r - aV / 2 * dr;
      out_i[i] = aK * lap_i - aV / 2 * di;
    }
  }
  size_t V = (size_t)N * N * N;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < V; ++i) {
    m->piCS[0][i] += c * out_r[i];
    m->piCS[1][i] += c * out_i[i];
  }
  free(out_r);
  free(out_i);
}

/* piU1(i) += c * (aI^(alpha-1) * (LatLapl A_i - sum_j backDiff(forwDiff(A_j, i), j)) - aI^(1+alpha) J_i)
 * with J_i = 2/dx (fStar/omegaStar)^2 gQ Im[conj(U_i(x) phi(x+i)) phi(x)]
 * (evolvers/kernels/u1kernels.h:31-52, definitions/mattercurrents.h:52-76). */
static void kick_u1(orc_model *m, double c)
{
  const int N = m->N;
  const double aG = pow(m->bg.aI, -1 + m->alpha);
  const double aJ = pow(m->bg.aI, 1 + m->alpha);
  const double idx2 = 1.0 / (m->dx * m->dx);
  const double fw = m->fStar / m->omegaStar;
  const double jnorm = 2.0 / m->dx * fw * fw;
  size_t V = (size_t)N * N * N;
  double *out[3];
  for (int d = 0; d < 3; ++d) out[d] = (double *)malloc(sizeof(double) * V);
#pragma omp parallel for schedule(static)
  for (int xy = 0; xy < N * N; ++xy) {
    int x = xy / N, y = xy % N;
    for (int z = 0; z < N; ++z) {
      size_t s = idx3(N, x, y, z);
      for (int i = 0; i < 3; ++i) {
        const double *Ai = m->fldU1[i];
        double lap = 0.0;
        for (int d = 0; d < 3; ++d) lap += Ai[nbr(N, x, y, z, d, +1)] + Ai[nbr(N, x, y, z, d, -1)] - 2 * Ai[s];
        lap *= idx2;
        /* sum_j backDiff_j(forwDiff_i A_j) = sum_j [A_j(x+i) - A_j(x) - A_j(x+i-j) + A_j(x-j)] / dx^2 */
        double gd = 0.0;
        for (int j = 0; j < 3; ++j) {
          const double *Aj = m->fldU1[j];
          int xp = x, yp = y, zp = z;
          if (i == 0) xp = wrap(x + 1, N);
          if (i == 1) yp = wrap(y + 1, N);
          if (i == 2) zp = wrap(z + 1, N);
          size_t s_pi = idx3(N, xp, yp, zp);
          size_t s_pimj = nbr(N, xp, yp, zp, j, -1);
          size_t s_mj = nbr(N, x, y, z, j, -1);
          gd += (Aj[s_pi] - Aj[s]) - (Aj[s_pimj] - Aj[s_mj]);
        }
        gd *= idx2;
        double J = 0.0;
        if (m->ncs > 0) {
          size_t sp = nbr(N, x, y, z, i, +1);
          double cu, su;
          u1_link(m, s, i, &cu, &su);
          double ur = cu * m->fldCS[0][sp] - su * m->fldCS[1][sp];
          double ui = cu * m->fldCS[1][sp] + su * m->fldCS[0][sp];
          /* Im[conj(u) * phi(x)] = ur * phi_i - ui * phi_r */
          J = jnorm * m->gQ * (ur * m->fldCS[1][s] - ui * m->fldCS[0][s]);
        }
        out[i][s] = aG * (lap - gd) - aJ * J;
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (size_t s = 0; s < V; ++s)
    for (int i = 0; i < 3; ++i) m->piU1[i][s] += c * out[i][s];
  for (int d = 0; d < 3; ++d) free(out[d]);
}

/* Kick every sector by weight c = w*dt (leapfrog) or w*dt/2 (velocity Verlet). */
void orc_kick(orc_model *m, double c)
{
  for (int n = 0; n < m->ns; ++n) kick_scalar_field(m, n, c);
  if (m->ncs > 0) kick_cs(m, c);
  if (m->nu1 > 0) kick_u1(m, c);
}

/* fldS += a^(alpha-3) * (cdt * piS); fldCS likewise; fldU1 += a^(alpha-1) * cdt * piU1.  The leapfrog
 * and velocity Verlet drift with a = aSI (evolvers/leapfrog.h:185-217, velocityverlet.h:181-209), the
 * position Verlet with a = aI (evolvers/positionverlet.h:161-192). */
static void drift_at(orc_model *m, double cdt, double a)
{
  const size_t V = (size_t)m->N * m->N * m->N;
  const double fs = pow(a, m->alpha - 3);
  const double fu = pow(a, m->alpha - 1);
  for (int n = 0; n < m->ns; ++n) {
    double *phi = m->fldS[n];
    const double *pi = m->piS[n];
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < V; ++i) phi[i] += fs * (cdt * pi[i]);
  }
  if (m->ncs > 0)
    for (int c = 0; c < 2; ++c) {
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < V; ++i) m->fldCS[c][i] += fs * cdt * m->piCS[c][i];
    }
  if (m->nu1 > 0)
    for (int d = 0; d < 3; ++d) {
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < V; ++i) m->fldU1[d][i] += fu * cdt * m->piU1[d][i];
    }
}

void orc_drift(orc_model *m, double cdt) { drift_at(m, cdt, m->bg.aSI); }

/* ------------------------------------------------------------------------------------------ */
/* Leapfrog (evolvers/leapfrog.h)                                                              */
/* ------------------------------------------------------------------------------------------ */

/* LeapFrog constructor (leapfrog.h:29-38): momenta start synced, aDotSI set at t = -dt/2. */
void orc_lf_init(orc_model *m)
{
  m->synced = 1;
  if (m->expansion) m->bg.aDotSI = m->bg.aDotI - m->dt / 2.0 * orc_scale_factor_kernel(m);
}

static void store_momenta_averages_lf(orc_model *m)
{
  orc_background *b = &m->bg;
  if (m->ns > 0) {
    b->pi2AvSIM = b->pi2AvSI;
    b->pi2AvSI = orc_avg_pi2S(m);
    b->pi2AvI = 0.5 * (b->pi2AvSIM + b->pi2AvSI);
  }
  if (m->ncs > 0) {
    b->CSpi2AvSIM = b->CSpi2AvSI;
    b->CSpi2AvSI = orc_avg_pi2CS(m);
    b->CSpi2AvI = 0.5 * (b->CSpi2AvSIM + b->CSpi2AvSI);
  }
  if (m->nu1 > 0) {
    b->U1pi2AvSIM = b->U1pi2AvSI;
    b->U1pi2AvSI = orc_avg_pi2U1(m);
    b->U1pi2AvI = 0.5 * (b->U1pi2AvSIM + b->U1pi2AvSI);
  }
}

static void store_fields_averages(orc_model *m)
{
  orc_background *b = &m->bg;
  if (m->ns > 0) b->grad2AvI = orc_avg_grad2S(m);
  if (m->ncs > 0) b->CSgrad2AvI = orc_avg_grad2CS(m);
  if (m->nu1 > 0) b->U1Mag2AvI = orc_avg_B2U1(m);
  b->potAvI = orc_avg_potential(m);
}

/* LeapFrog::evolve (leapfrog.h:42-89) for a self-consistent or static background. */
void orc_lf_evolve(orc_model *m)
{
  orc_background *b = &m->bg;
  double w = m->synced ? 0.5 : 1.0;
  orc_kick(m, w * m->dt);
  if (m->expansion) {
    store_momenta_averages_lf(m);
    /* kickScaleFactor (leapfrog.h:124-129) */
    b->aDotSIM = b->aDotSI;
    b->aDotSI += m->dt * orc_scale_factor_kernel(m);
    b->aDotI = (b->aDotSIM + b->aDotSI) / 2.0;
    /* driftScaleFactor (leapfrog.h:168-184) */
    b->aIM = b->aI;
    b->aI += m->dt * b->aDotSI;
    b->aSI = (b->aIM + b->aI) / 2.0;
  }
  orc_drift(m, m->dt);
  if (m->expansion) store_fields_averages(m);
  m->synced = 0;
}

/* LeapFrog::sync (leapfrog.h:96-116). */
void orc_lf_sync(orc_model *m)
{
  if (!m->synced) {
    orc_kick(m, 0.5 * m->dt);
    if (m->expansion) {
      orc_background *b = &m->bg;
      if (m->ns > 0) b->pi2AvI = orc_avg_pi2S(m);
      if (m->ncs > 0) b->CSpi2AvI = orc_avg_pi2CS(m);
      if (m->nu1 > 0) b->U1pi2AvI = orc_avg_pi2U1(m);
      b->aDotI = b->aDotSI + m->dt / 2.0 * orc_scale_factor_kernel(m);
    }
  }
  m->synced = 1;
}

/* ------------------------------------------------------------------------------------------ */
/* Velocity Verlet (evolvers/velocityverlet.h:51-114)                                          */
/* ------------------------------------------------------------------------------------------ */

static void store_momenta_averages_vv(orc_model *m)
{
  orc_background *b = &m->bg;
  if (m->ns > 0) b->pi2AvI = orc_avg_pi2S(m);
  if (m->ncs > 0) b->CSpi2AvI = orc_avg_pi2CS(m);
  if (m->nu1 > 0) b->U1pi2AvI = orc_avg_pi2U1(m);
}

void orc_vv_evolve(orc_model *m, const double *ws, int nws)
{
  orc_background *b = &m->bg;
  int stages = (nws - 1) * 2 + 1;
  for (int s = 0; s < stages; ++s) {
    double w = ws[s < nws ? s : stages - s - 1];
    if (m->expansion) b->aDotSI = b->aDotI + (m->dt / 2) * orc_scale_factor_kernel(m) * w;
    orc_kick(m, w * m->dt / 2);
    if (m->expansion) {
      store_momenta_averages_vv(m);
      b->aIM = b->aI;
      b->aI += m->dt * b->aDotSI * w;
      b->aSI = (b->aIM + b->aI) / 2.0;
    }
    orc_drift(m, m->dt * w);
    if (m->expansion) store_fields_averages(m);
    orc_kick(m, w * m->dt / 2);
    if (m->expansion) {
      store_momenta_averages_vv(m);
      b->aDotI = b->aDotSI + (m->dt / 2) * orc_scale_factor_kernel(m) * w;
    }
  }
}

/* ------------------------------------------------------------------------------------------ */
/* Position Verlet (evolvers/positionverlet.h:47-111)                                          */
/* ------------------------------------------------------------------------------------------ */

/* ScaleFactorKernels::getSI (evolvers/kernels/scalefactorkernels.h:84-124): the kernel of get() with
 * the half-step averages, the current aI in every energy normalisation (definitions/energies.h:37-85,
 * 104-145). */
static double scale_factor_kernel_si(const orc_model *m, double grad2SI, double CSgrad2SI, double U1Mag2SI)
{
  const orc_background *b = &m->bg;
  const double a = b->aI;
  double Eks = 0, Egs = 0, Ekcs = 0, Egcs = 0, Eel = 0, Emag = 0;
  if (m->ns > 0) {
    Eks = 0.5 * b->pi2AvSI * pow(a, -6);
    Egs = 0.5 * grad2SI * pow(a, -2);
  }
  if (m->ncs > 0) {
    Ekcs = b->CSpi2AvSI * pow(a, -6);
    Egcs = CSgrad2SI * pow(a, -2);
  }
  if (m->nu1 > 0) {
    double w2 = (m->omegaStar / m->fStar) * (m->omegaStar / m->fStar);
    Eel = 0.5 * w2 * pow(a, -4) * b->U1pi2AvSI;
    Emag = 0.5 * w2 * pow(a, -4) * U1Mag2SI;
  }
  double fm = m->fStar / m->MPl;
  return pow(a, 2 * m->alpha + 1) / 3.0 * (fm * fm) *
         ((m->alpha - 2) * (Eks + Ekcs) + m->alpha * (Egs + Egcs) + (m->alpha - 1) * (Eel + Emag) +
          (m->alpha + 1) * b->potAvSI);
}

/* One step of PositionVerlet::evolve: for each stage weight w, drift by w*dt/2 (with aI), the
 * half-step scale-factor drift, half-step field averages, kick by w*dt (with the half-step aI),
 * storeMomentaAverages (positionverlet.h:238-256: previous integer-time value and the new one give
 * the half-step value), kickScaleFactor with getSI, second half-step scale-factor drift
 * (aSI = (aIM + aI)/2), drift by w*dt/2 with the new aI, integer-time field averages. */
void orc_pv_evolve(orc_model *m, const double *ws, int nws)
{
  orc_background *b = &m->bg;
  const int stages = (nws - 1) * 2 + 1;
  const double dt = m->dt;
  for (int s = 0; s < stages; ++s) {
    const double w = ws[s < nws ? s : stages - s - 1];
    drift_at(m, dt / 2.0 * w, b->aI);
    double grad2SI = 0, CSgrad2SI = 0, U1Mag2SI = 0;
    if (m->expansion) {
      b->aIM = b->aI; /* driftScaleFactorFirstHalf (positionverlet.h:196-209) */
      b->aI += dt / 2.0 * b->aDotI * w;
      b->aSI = b->aI;
      /* storeFieldsAveragesHalf (positionverlet.h:258-266) */
      if (m->ns > 0) grad2SI = orc_avg_grad2S(m);
      if (m->ncs > 0) CSgrad2SI = orc_avg_grad2CS(m);
      if (m->nu1 > 0) U1Mag2SI = orc_avg_B2U1(m);
      b->potAvSI = orc_avg_potential(m);
    }
    orc_kick(m, w * dt);
    if (m->expansion) {
      /* storeMomentaAverages (positionverlet.h:238-256) */
      const double pIM = b->pi2AvI, cIM = b->CSpi2AvI, uIM = b->U1pi2AvI;
      if (m->ns > 0) b->pi2AvI = orc_avg_pi2S(m);
      if (m->ncs > 0) b->CSpi2AvI = orc_avg_pi2CS(m);
      if (m->nu1 > 0) b->U1pi2AvI = orc_avg_pi2U1(m);
      b->pi2AvSI = 0.5 * (b->pi2AvI + pIM);
      b->CSpi2AvSI = 0.5 * (b->CSpi2AvI + cIM);
      b->U1pi2AvSI = 0.5 * (b->U1pi2AvI + uIM);
      /* kickScaleFactor (positionverlet.h:151-154) */
      b->aDotI += dt * scale_factor_kernel_si(m, grad2SI, CSgrad2SI, U1Mag2SI) * w;
      /* driftScaleFactorSecondHalf (positionverlet.h:212-223) */
      b->aI += dt / 2.0 * b->aDotI * w;
      b->aSI = (b->aIM + b->aI) / 2.0;
    }
    drift_at(m, dt / 2.0 * w, b->aI);
    if (m->expansion) store_fields_averages(m); /* storeFieldsAveragesInt (positionverlet.h:268-276) */
  }
}

/* ------------------------------------------------------------------------------------------ */
/* Gauss law (definitions/gausslaws.h:31-53, mattercurrents.h:27-48)                           */
/* ------------------------------------------------------------------------------------------ */

/* out = {<|LHS-RHS|> / <LHS1^2+LHS2^2+LHS3^2+RHS^2>, <|LHS|>, <|RHS|>} */
void orc_gauss_u1(const orc_model *m, double *out)
{
  const int N = m->N;
  const double fw = m->fStar / m->omegaStar;
  const double qnorm = 2.0 / m->dx * fw * fw;
  double sd = 0, ss = 0, sr = 0, sl = 0;
#pragma omp parallel for reduction(+ : sd, ss, sr, sl) schedule(static)
  for (int xy = 0; xy < N * N; ++xy) {
    int x = xy / N, y = xy % N;
    for (int z = 0; z < N; ++z) {
      size_t s = idx3(N, x, y, z);
      double l[3], lhs = 0.0;
      for (int i = 0; i < 3; ++i) {
        l[i] = (m->piU1[i][s] - m->piU1[i][nbr(N, x, y, z, i, -1)]) / m->dx;
      }
      lhs = (m->piU1[0][s] - m->piU1[0][nbr(N, x, y, z, 0, -1)] + m->piU1[1][s] -
             m->piU1[1][nbr(N, x, y, z, 1, -1)] + m->piU1[2][s] - m->piU1[2][nbr(N, x, y, z, 2, -1)]) /
            m->dx;
      double rho = 0.0;
      if (m->ncs > 0) /* Im[conj(piCS) * fldCS] */
        rho = qnorm * m->gQ * (m->piCS[0][s] * m->fldCS[1][s] - m->piCS[1][s] * m->fldCS[0][s]);
      double rhs = -m->dx * rho;
      double diff = lhs - rhs;
      double sum = l[0] * l[0] + l[1] * l[1] + l[2] * l[2] + rhs * rhs;
      sd += fabs(diff);
      ss += fabs(sum);
      sr += fabs(rhs);
      sl += fabs(lhs);
    }
  }
  double V = (double)N * N * N;
  out[0] = (sd / V) / (ss / V);
  out[1] = sl / V;
  out[2] = sr / V;
}

/* ------------------------------------------------------------------------------------------ */
/* Small helpers exported for tests                                                             */
/* ------------------------------------------------------------------------------------------ */

/* Plain lattice Laplacian of one array (definitions/gaugederivatives.h:36-40). */
void orc_laplacian(int N, double dx, const double *f, double *out)
{
  const double idx2 = 1.0 / (dx * dx);
#pragma omp parallel for schedule(static)
  for (int xy = 0; xy < N * N; ++xy) {
    int x = xy / N, y = xy % N;
    for (int z = 0; z < N; ++z) {
      size_t i = idx3(N, x, y, z);
      double lap = 0.0;
      for (int d = 0; d < 3; ++d) lap += f[nbr(N, x, y, z, d, +1)] + f[nbr(N, x, y, z, d, -1)] - 2 * f[i];
      out[i] = lap * idx2;
    }
  }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

size_t orc_sizeof_model(void) { return sizeof(orc_model); }
