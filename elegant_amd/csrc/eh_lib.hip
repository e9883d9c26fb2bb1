/* eh_lib.hip -- HIP kernels and C ABI of the gfx950 element-tracking library.
 *
 * Residency: particles stay on the GPU between calls as structure-of-arrays
 * (x, x', y, y', s, delta as separate double arrays, plus the particle's initial
 * row and its loss record).  Survivors live in one of two "alive" buffers;
 * lost particles are appended to a graveyard.  Compaction moves survivors with
 * a wavefront ballot + mbcnt prefix and one atomic per wavefront, so the device
 * order is not the reference's order.  The reference's row order (swap-to-top
 * compaction in every element, src/multipole.c:1032-1047,
 * src/limit_amplitudes.c:262-283) is reconstructed exactly on the host when it
 * is needed -- for downloads and for the RFCA TMEAN fiducial, whose Kahan sum
 * runs in row order -- by replaying the recorded loss events.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "elegant_hip.h"
#include "eh_csbend_terms.h"
#include "eh_device.h"

/* ------------------------------------------------------------------------- */
/* errors */
static char g_err[512];
extern "C" const char *eh_last_error(void) { return g_err; }
static long fail(long code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}
#define HIPCHK(call)                                                                                                  \
  do {                                                                                                                \
    hipError_t e_ = (call);                                                                                           \
    if (e_ != hipSuccess)                                                                                             \
      return fail(EH_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));                                        \
  } while (0)

static const double EH_PI = 3.14159265358979323846;

/* ------------------------------------------------------------------------- */
/* integrator tables, src/multipole.c:1087-1141 and src/csbend.c:1478-1500 */
#define EH_BETA 1.25992104989487316477
static const double driftFrac2[2] = {0.5, 0.5};
static const double kickFrac2[2] = {1.0, 0.0};
static const double driftFrac4[4] = {0.5 / (2 - EH_BETA), (1 - EH_BETA) / (2 - EH_BETA) / 2,
                                     (1 - EH_BETA) / (2 - EH_BETA) / 2, 0.5 / (2 - EH_BETA)};
static const double kickFrac4[4] = {1. / (2 - EH_BETA), -EH_BETA / (2 - EH_BETA), 1 / (2 - EH_BETA), 0};
static const double driftFrac6[8] = {0.39225680523878,   0.5100434119184585, -0.47105338540975655,
                                     0.0687531682525181, 0.0687531682525181, -0.47105338540975655,
                                     0.5100434119184585, 0.39225680523878};
static const double kickFrac6[8] = {0.784513610477560,  0.235573213359357, -1.17767998417887, 1.3151863206839063,
                                    -1.17767998417887, 0.235573213359357, 0.784513610477560,  0};

static int integ_tables(long order, const double **drift, const double **kick, int *nsub) {
  switch (order) {
  case 2:
    *drift = driftFrac2;
    *kick = kickFrac2;
    *nsub = 2;
    return 1;
  case 4:
    *drift = driftFrac4;
    *kick = kickFrac4;
    *nsub = 4;
    return 1;
  case 6:
    *drift = driftFrac6;
    *kick = kickFrac6;
    *nsub = 8;
    return 1;
  }
  return 0;
}

/* cos/sin of a rotation angle with the exact special angles of
 * rotateBeamCoordinatesForMisalignment, src/tilt_matrices.c:156-175 */
static void rot_cs(double angle, double *c, double *s) {
  if (std::fabs(std::fabs(angle) - EH_PI) < 1e-12) {
    *c = -1;
    *s = 0;
  } else if (std::fabs(angle - EH_PI / 2) < 1e-12) {
    *c = 0;
    *s = 1;
  } else if (std::fabs(angle + EH_PI / 2) < 1e-12) {
    *c = 0;
    *s = -1;
  } else {
    *c = eh_cos(angle);
    *s = eh_sin(angle);
  }
}

static double hsqr(double x) { return x * x; }
static double hpow3(double x) { return x * x * x; }

/* rad_coef of multipole_tracking and multipole_tracking2 (src/multipole.c:501,
 * :756), the bend's with sqr(1 + fse)/sqr(rho0) (src/csbend.c:957); the same
 * expression, constant for constant, as the oracle */
static double rad_coef_multipole(double Po) {
  return hsqr(EH_E_MKS) * hpow3(Po) / (6 * EH_PI * EH_EPSILON_O * hsqr(EH_C_MKS) * EH_ME_MKS);
}

/* rotateBeamCoordinatesForMisalignment leaves the beam alone for angle 0 and
 * for |angle| within 1e-12 of 2*pi (src/malign_mat.c) */
static int tilt_on(double angle) { return angle != 0 && !(std::fabs(std::fabs(angle) - 2 * EH_PI) < 1e-12); }

/* ------------------------------------------------------------------------- */
/* CSBEND field table: computeCSBENDFieldCoefficients (src/csbend.c:175-652),
 * evaluated term by term from the generated Maxwell-recursion table */
extern "C" long eh_csbend_field_table(const double *b, const double *c, double h1, int nonlinear,
                                      int expansionOrder, double *Fx, double *Fy, int *hasNormal, int *hasSkew) {
  long i;
  double h[20];
  if (expansionOrder == 0) {
    for (i = 8; i >= 0; i--)
      if (b[i] || c[i])
        break;
    if ((expansionOrder = (int)(i + 2)) < 4)
      expansionOrder = 4;
  }
  long eo1 = expansionOrder + 1;
  if (eo1 > 11)
    return fail(EH_ERR_UNSUPPORTED, "expansion order >10 for CSBEND");
  int hn = 0, hs = 0;
  for (i = 0; i < 9; i++) {
    if (b[i])
      hn = 1;
    if (c[i])
      hs = 1;
  }
  if (hasNormal)
    *hasNormal = hn;
  if (hasSkew)
    *hasSkew = hs;
  std::memset(Fx, 0, sizeof(double) * 121);
  std::memset(Fy, 0, sizeof(double) * 121);
  h[0] = 1;
  for (i = 1; i < 20; i++)
    h[i] = h[i - 1] * h1;
  Fx[0] = c[0];
  Fy[0] = 1 - b[0];
  int prev_w = -1, prev_i = -1, prev_j = -1;
  for (long t = 0; t < EH_CSBEND_NTERMS; t++) {
    int w = eh_csbend_terms[t].which, ii = eh_csbend_terms[t].i, jj = eh_csbend_terms[t].j;
    bool first = !(w == prev_w && ii == prev_i && jj == prev_j);
    prev_w = w;
    prev_i = ii;
    prev_j = jj;
    if (!nonlinear && ii + jj > 1)
      continue;
    const double *v = eh_csbend_terms[t].kind ? c + 1 : b + 1;
    double term = ((eh_csbend_terms[t].num * h[eh_csbend_terms[t].hpow]) * v[eh_csbend_terms[t].m]) /
                  eh_csbend_terms[t].den;
    double *dst = (w ? Fy : Fx) + ii * 11 + jj;
    if (first)
      *dst = term;
    else
      *dst += term;
  }
  return eo1;
}

/* ------------------------------------------------------------------------- */
/* descriptor construction.  ap[pi] is the MAXAMP passed to the element routine
 * and post[pi] the limits of the sweep that follows it, for pass 0 (pi = 0) and
 * later passes (pi = 1). */
struct eh_ctx {
  double Po, z_start, z_end;
  const EH_MAXAMP *ap[2];
  double post_x[2], post_y[2];
  int post_on[2];
  /* an RFCA in a phase-reference group set by another cavity: the group's
   * phase replaces the cavity's own phase_fiducial (src/simple_rfca.c:419-421) */
  int rf_linked;
  double rf_ref_phase;
  /* CHANGE_T: the time offset found at this visit (src/simple_rfca.c:488-498) */
  double rf_toff;
  /* LINEARIZE / LOCK_PHASE: the beam's mean arrival time at this visit */
  double rf_tave;
};

static long build_multipole(eh_desc *d, const EH_ELEMENT_LIST *e, const eh_ctx &c) {
  double KnL[3] = {0, 0, 0};
  long order[3] = {0, 0, 0};
  int skew[3] = {0, 0, 0};
  double dx = 0, dy = 0, dz = 0, tilt = 0, xkick = 0, ykick = 0, drift = 0;
  long nSlices = 0, n_kicks = 0, integ_order = 0;
  int expand = 0, synch_rad = 0;
  switch (e->type) {
  case EH_T_KQUAD: {
    const EH_KQUAD *q = (const EH_KQUAD *)e->p_elem;
    synch_rad = q->synch_rad;
    nSlices = q->nSlices;
    n_kicks = q->n_kicks;
    expand = q->expandHamiltonian;
    order[0] = 1;
    /* KnL = n! B / BORE^n e / (m c P) L (1 + FSE), src/multipole.c:739-742 */
    KnL[0] = q->bore ? q->B / q->bore * (EH_E_MKS / (EH_ME_MKS * EH_C_MKS * c.Po)) * q->length * (1 + q->fse)
                     : q->k1 * q->length * (1 + q->fse);
    drift = q->length;
    tilt = q->tilt;
    dx = q->dx;
    dy = q->dy;
    dz = q->dz;
    xkick = q->xkick * q->xKickCalibration;
    ykick = q->ykick * q->yKickCalibration;
    integ_order = q->integration_order;
    break;
  }
  case EH_T_KSEXT: {
    const EH_KSEXT *q = (const EH_KSEXT *)e->p_elem;
    synch_rad = q->synch_rad;
    nSlices = q->nSlices;
    n_kicks = q->n_kicks;
    expand = q->expandHamiltonian;
    order[0] = 2;
    KnL[0] = (q->bore ? 2 * q->B / (q->bore * q->bore) * (EH_E_MKS / (EH_ME_MKS * EH_C_MKS * c.Po)) * q->length *
                                  (1 + q->fse)
                           : q->k2 * q->length * (1 + q->fse));
    drift = q->length;
    tilt = q->tilt;
    dx = q->dx;
    dy = q->dy;
    dz = q->dz;
    xkick = q->xkick * q->xKickCalibration;
    ykick = q->ykick * q->yKickCalibration;
    integ_order = q->integration_order;
    if (q->k1) {
      KnL[1] = q->k1 * q->length;
      order[1] = 1;
    }
    if (q->j1) {
      KnL[2] = q->j1 * q->length;
      order[2] = 1;
      skew[2] = 1;
    }
    break;
  }
  case EH_T_KOCT: {
    const EH_KOCT *q = (const EH_KOCT *)e->p_elem;
    synch_rad = q->synch_rad;
    nSlices = q->nSlices;
    n_kicks = q->n_kicks;
    expand = q->expandHamiltonian;
    order[0] = 3;
    KnL[0] = (q->bore ? 6 * q->B / (q->bore * q->bore * q->bore) * (EH_E_MKS / (EH_ME_MKS * EH_C_MKS * c.Po)) *
                                  q->length * (1 + q->fse)
                           : q->k3 * q->length * (1 + q->fse));
    drift = q->length;
    tilt = q->tilt;
    dx = q->dx;
    dy = q->dy;
    dz = q->dz;
    integ_order = q->integration_order;
    break;
  }
  case EH_T_KQUSE: {
    /* src/multipole.c:923-949 */
    const EH_KQUSE *q = (const EH_KQUSE *)e->p_elem;
    synch_rad = q->synch_rad;
    nSlices = q->nSlices;
    n_kicks = q->n_kicks;
    expand = q->expandHamiltonian;
    order[0] = 1;
    KnL[0] = q->k1 * q->length * (1 + q->fse1);
    drift = q->length;
    tilt = q->tilt;
    dx = q->dx;
    dy = q->dy;
    dz = q->dz;
    integ_order = q->integration_order;
    KnL[1] = q->k2 * q->length * (1 + q->fse2);
    order[1] = 2;
    break;
  }
  default:
    return fail(EH_ERR_UNSUPPORTED, "multipole_tracking2: element type not supported");
  }
  const double *driftFrac, *kickFrac;
  int nsub;
  if (!integ_tables(integ_order, &driftFrac, &kickFrac, &nsub))
    return fail(EH_ERR_UNSUPPORTED, "multipole integration_order must be 2, 4, or 6");
  if (n_kicks <= 0) {
    if (nSlices <= 0)
      return fail(EH_ERR_UNSUPPORTED, "N_KICKS<=0 and N_SLICES<=0 in multipole tracking");
  } else {
    if (integ_order > 2) {
      if ((nSlices = (long)std::ceil(n_kicks / (1.0 * integ_order))) < 1)
        nSlices = 1;
    } else
      nSlices = n_kicks;
  }
  d->nslices = (int32_t)nSlices;
  d->nsub = nsub;
  d->expand = expand;
  double mdrift = drift / nSlices;
  d->mdrift = mdrift;
  double xk = xkick / nSlices, yk = ykick / nSlices;
  d->has_xkick = xk != 0;
  d->has_ykick = yk != 0;
  double KnLp[3];
  for (int i = 0; i < 3; i++) {
    KnLp[i] = KnL[i] / nSlices;
    d->mp_order[i] = KnL[i] ? (int32_t)order[i] : 0;
    d->mp_skew[i] = skew[i];
  }
  d->adj_order0 = (int32_t)order[0];
  const double kfac = drift * (1 + (e->type == EH_T_KQUAD   ? ((const EH_KQUAD *)e->p_elem)->fse
                                    : e->type == EH_T_KSEXT ? ((const EH_KSEXT *)e->p_elem)->fse
                                    : e->type == EH_T_KQUSE ? ((const EH_KQUSE *)e->p_elem)->fse1
                                                            : ((const EH_KOCT *)e->p_elem)->fse)) /
                      nSlices;
  /* classical radiation: none for elements shorter than 1e-6 m (src/multipole.c:761-763) */
  d->rad_coef = synch_rad && drift >= 1e-6 ? rad_coef_multipole(c.Po) : 0.0;
  d->rad = d->rad_coef != 0;
  d->rad_xk = xk;
  d->rad_yk = yk;
  for (int s = 0; s < nsub; s++) {
    d->kfrac[s] = kickFrac[s];
    d->adj_kfac[s] = kfac * kickFrac[s];
    d->dsh[s] = mdrift * driftFrac[s];
    for (int i = 0; i < 3; i++)
      d->kick[s][i] = KnLp[i] * kickFrac[s];
    d->xk[s] = -xk * kickFrac[s];
    d->yk[s] = -yk * kickFrac[s];
  }
  d->mdx = dx;
  d->mdy = dy;
  d->mdz = dz;
  d->has_offset = (dx || dy || dz);
  d->has_tilt = tilt_on(tilt);
  rot_cs(tilt + 0.0, &d->ncos, &d->nsin);
  rot_cs(-(tilt + 0.0), &d->xcos, &d->xsin);
  double rt = -tilt;
  d->ap_rtilt = rt != 0;
  d->ap_cs0 = eh_cos(rt);
  d->ap_cs1 = eh_sin(rt);
  d->length = drift;
  return EH_OK;
}

/* the bend's b[] and c[] coefficients before the negative-angle flip
 * (src/csbend.c:800-857: K1..K8 times rho0 or B1..B8, the FSE adjustments of
 * the quadrupole and sextupole terms, the steering kicks) */
static void csbend_coefficients(const EH_CSBEND *cb, double b[9], double cc[9]) {
  double rho0 = cb->length / cb->angle;
  b[0] = 0;
  if (cb->use_bn) {
    b[1] = cb->b1;
    b[2] = cb->b2;
    b[3] = cb->b3;
    b[4] = cb->b4;
    b[5] = cb->b5;
    b[6] = cb->b6;
    b[7] = cb->b7;
    b[8] = cb->b8;
  } else {
    b[1] = cb->k1 * rho0;
    b[2] = cb->k2 * rho0;
    b[3] = cb->k3 * rho0;
    b[4] = cb->k4 * rho0;
    b[5] = cb->k5 * rho0;
    b[6] = cb->k6 * rho0;
    b[7] = cb->k7 * rho0;
    b[8] = cb->k8 * rho0;
  }
  for (int j = 0; j < 9; j++)
    cc[j] = 0;
  b[1] *= (1 + cb->fse + cb->fseQuadrupole) / (1 + cb->fse + cb->fseDipole);
  cc[1] *= (1 + cb->fse + cb->fseQuadrupole) / (1 + cb->fse + cb->fseDipole);
  b[2] *= (1 + cb->fse) / (1 + cb->fse + cb->fseDipole);
  cc[2] *= (1 + cb->fse) / (1 + cb->fse + cb->fseDipole);
  b[0] = cb->xKick / cb->angle;
  cc[0] = cb->yKick / cb->angle;
}

static long build_csbend(eh_desc *d, const EH_CSBEND *cb, const eh_ctx &c, std::vector<eh_fterm> &pool) {
  if (cb->angle == 0) {
    d->type = EH_T_EDRIFT;
    d->length = cb->length;
    return EH_OK;
  }
  const double *driftFrac, *kickFrac;
  int nsub;
  if (!integ_tables(cb->integration_order, &driftFrac, &kickFrac, &nsub))
    return fail(EH_ERR_UNSUPPORTED, "CSBEND integration_order is invalid");
  if (cb->malignMethod != 0 || cb->etilt != 0 || cb->expandHamiltonian)
    return fail(EH_ERR_UNSUPPORTED, "CSBEND option not supported");
  if (cb->edge1_effects < 0 || cb->edge1_effects > 4 || cb->edge2_effects < 0 || cb->edge2_effects > 4)
    return fail(EH_ERR_UNSUPPORTED, "CSBEND edge method not supported");
  if (cb->edge_order > 1 && (cb->edge1_effects == 2 || cb->edge2_effects == 2) && cb->hgap == 0)
    return fail(EH_ERR_ARG, "CSBEND has EDGE_ORDER>1 and EDGE[12]_EFFECTS==2, but HGAP=0. This gives undefined "
                            "results.");
  if (cb->nSlices < 1)
    return fail(EH_ERR_UNSUPPORTED, "invalid number of steps (integrate_csbend_ordn)");
  double b[9], cc[9];
  double rho0 = cb->length / cb->angle;
  csbend_coefficients(cb, b, cc);
  double angle, e1, e2, tilt, etilt;
  if (cb->angle < 0) {
    angle = -cb->angle;
    e1 = -cb->e1;
    e2 = -cb->e2;
    etilt = cb->etilt * cb->etiltSign;
    tilt = cb->tilt + EH_PI;
    rho0 = cb->length / angle;
    for (int j = 1; j < 9; j += 2) {
      b[j] *= -1;
      cc[j] *= -1;
    }
  } else {
    angle = cb->angle;
    e1 = cb->e1;
    e2 = cb->e2;
    etilt = cb->etilt * cb->etiltSign;
    tilt = cb->tilt;
    rho0 = cb->length / angle;
  }
  (void)angle;
  double rt = -tilt;
  d->ap_rtilt = rt != 0;
  d->ap_cs0 = eh_cos(rt);
  d->ap_cs1 = eh_sin(rt);
  if (std::fabs(rho0) > 1e6)
    return fail(EH_ERR_UNSUPPORTED, "CSBEND with radius > 1e6 not supported");
  double fse = cb->fse + cb->fseDipole;
  double h = 1 / rho0;
  if (std::fabs(fse + 1) < 1e-12)
    fse = -1 + 1e-12;
  double rho_actual = 1 / ((1 + fse) * h);
  /* classical radiation, src/csbend.c:956-959 (the oracle's expression) */
  d->rad_coef = cb->synch_rad ? hsqr(EH_E_MKS) * hpow3(c.Po) * hsqr(1 + fse) /
                                    (6 * EH_PI * EH_EPSILON_O * hsqr(EH_C_MKS) * EH_ME_MKS * hsqr(rho0))
                              : 0.0;
  d->rad = d->rad_coef != 0;
  d->csb_irho0 = 1. / rho0;
  double sgn = rho0 < 0 ? -1.0 : (rho0 > 0 ? 1.0 : 0.0);
  double Kg1 = 2 * cb->hgap * cb->fint * sgn;
  double psi1 = Kg1 / std::fabs(rho_actual) / eh_cos(e1) * (1 + eh_sqr(eh_sin(e1)));
  double Kg2 = 2 * cb->hgap * cb->fint * sgn;
  double psi2 = Kg2 / std::fabs(rho_actual) / eh_cos(e2) * (1 + eh_sqr(eh_sin(e2)));
  if (cb->length < 0) {
    psi1 *= -1;
    psi2 *= -1;
  }
  double Fx[121], Fy[121];
  int hn, hs;
  long eo1 = eh_csbend_field_table(b, cc, h, cb->nonlinear, cb->expansionOrder, Fx, Fy, &hn, &hs);
  if (eo1 < 0)
    return eo1;
  d->csb_field = hn || hs;
  /* non-zero terms in the summation order of computeCSBENDFields (i outer, j
   * inner; without skew terms Fx only takes odd j and Fy even j) */
  std::vector<eh_fterm> fx, fy;
  for (int i = 0; i < eo1; i++)
    for (int j = 0; j < eo1 - i; j++) {
      double fxv = Fx[i * 11 + j], fyv = Fy[i * 11 + j];
      if (hs) {
        if (fxv != 0)
          fx.push_back({i, j, fxv});
        if (fyv != 0)
          fy.push_back({i, j, fyv});
      } else if (j & 1) {
        if (fxv != 0)
          fx.push_back({i, j, fxv});
      } else {
        if (fyv != 0)
          fy.push_back({i, j, fyv});
      }
    }
  d->csb_nfx = (int32_t)fx.size();
  d->csb_nfy = (int32_t)fy.size();
  d->csb_maxj = 0;
  for (const eh_fterm &t : fx)
    d->csb_maxj = std::max(d->csb_maxj, t.j);
  for (const eh_fterm &t : fy)
    d->csb_maxj = std::max(d->csb_maxj, t.j);
  /* dense blocks for the unrolled evaluation of the common orders */
  d->csb_dense = 0;
  d->csb_live = 0;
  if (!hs && (eo1 == 5 || eo1 == 6)) {
    d->csb_dense = (int32_t)eo1;
    for (int i = 0, t = 0; i < eo1; i++)
      for (int j = 0; j < eo1 - i; j++, t++)
        if (((j & 1) ? Fx[i * 11 + j] : Fy[i * 11 + j]) != 0)
          d->csb_live |= 1u << t;
    for (int c = 0; c < 2; c++) {
      (c ? d->csb_dy0 : d->csb_dx0) = (int32_t)pool.size();
      for (int i = 0; i < eo1; i++)
        for (int j = 0; j < eo1; j++)
          pool.push_back({i, j, (i + j < eo1) ? (c ? Fy[i * 11 + j] : Fx[i * 11 + j]) : 0.0});
    }
  }
  d->csb_fx0 = (int32_t)pool.size();
  pool.insert(pool.end(), fx.begin(), fx.end());
  d->csb_fy0 = (int32_t)pool.size();
  pool.insert(pool.end(), fy.begin(), fy.end());
  double ttilt = tilt + etilt, cos_tt, sin_tt;
  if (ttilt == 0) {
    cos_tt = 1;
    sin_tt = 0;
  } else if (std::fabs(std::fabs(ttilt) - EH_PI) < 1e-12) {
    cos_tt = -1;
    sin_tt = 0;
  } else if (std::fabs(ttilt - EH_PI / 2) < 1e-12) {
    cos_tt = 0;
    sin_tt = 1;
  } else if (std::fabs(ttilt + EH_PI / 2) < 1e-12) {
    cos_tt = 0;
    sin_tt = -1;
  } else {
    cos_tt = eh_cos(ttilt);
    sin_tt = eh_sin(ttilt);
  }
  d->cos_tt = cos_tt;
  d->sin_tt = sin_tt;
  d->dxi = -cb->dx;
  d->dzi = cb->dz;
  d->dyi = -cb->dy;
  d->dxf = cb->dx * eh_cos(cb->angle) + cb->dz * eh_sin(cb->angle);
  d->dzf = cb->dx * eh_sin(cb->angle) - cb->dz * eh_cos(cb->angle);
  d->dyf = cb->dy;
  d->rho0 = rho0;
  d->rho_actual = rho_actual;
  /* reciprocal pairs for the kick's divisions (see eh_div_recip): valid for
   * divisors well inside the normal range */
  d->csb_recip = 0;
  if (std::fabs(rho0) >= 0x1p-500 && std::fabs(rho0) <= 0x1p500 && std::fabs(rho_actual) >= 0x1p-500 &&
      std::fabs(rho_actual) <= 0x1p500) {
    d->csb_recip = 1;
    d->rho0_rh = 1.0 / rho0;
    d->rho0_rl = std::fma(-d->rho0_rh, rho0, 1.0) / rho0;
    d->rhoa_rh = 1.0 / rho_actual;
    d->rhoa_rl = std::fma(-d->rhoa_rh, rho_actual, 1.0) / rho_actual;
  }
  d->csb_nslices = cb->nSlices;
  d->nsub = nsub;
  double s = cb->length / cb->nSlices;
  for (int j = 0; j < nsub; j++) {
    d->csb_dsh[j] = s * driftFrac[j];
    d->csb_dshr[j] = d->csb_dsh[j] / rho0;
    d->adj_sin_dshr[j] = eh_sin(d->csb_dshr[j]);
    d->csb_ds[j] = s * kickFrac[j];
  }
  d->csb_dzl = cb->nSlices * s;
  d->csb_edge1 = (cb->edgeFlags & EH_BEND_EDGE1_EFFECTS) && cb->edge1_effects >= 1;
  d->csb_edge2 = (cb->edgeFlags & EH_BEND_EDGE2_EFFECTS) && cb->edge2_effects >= 1;
  d->csb_esimple1 = d->csb_edge1 && cb->edge1_effects == 3;
  d->csb_esimple2 = d->csb_edge2 && cb->edge2_effects == 3;
  /* EDGE_EFFECTS 2, dipoleFringeKHwang (src/csbend.c:1139-1159, :1319-1338):
   * the edge angle's functions, the full gap, FINT (both edges) and H */
  d->csb_hwang1 = d->csb_edge1 && cb->edge1_effects == 2;
  d->csb_hwang2 = d->csb_edge2 && cb->edge2_effects == 2;
  d->csb_hw_order2 = cb->edge_order > 1;
  /* EDGE_EFFECTS 4, dipoleFringeKHwangRLindberg (:1160-1181, :1340-1360) */
  d->csb_hl1 = d->csb_edge1 && cb->edge1_effects == 4;
  d->csb_hl2 = d->csb_edge2 && cb->edge2_effects == 4;
  for (int k = 0; k < 2; k++) {
    const double e = k ? e2 : e1;
    double *hw = k ? d->csb_hw2 : d->csb_hw1;
    hw[0] = 1. / eh_cos(e);
    hw[1] = eh_tan(e);
    hw[2] = eh_sin(e);
    hw[3] = eh_cos(e);
    hw[4] = 2 * cb->hgap;
    hw[5] = cb->fint;
    hw[6] = k ? cb->h2 : cb->h1;
  }
  /* edge kick limits, src/csbend.c:932-937 (they act in methods 1 with
   * EDGE_ORDER 1, and 3) */
  d->csb_klim1 = cb->edge1_kick_limit;
  d->csb_klim2 = cb->edge2_kick_limit;
  if (cb->kick_limit_scaling) {
    d->csb_klim1 *= rho0 / rho_actual;
    d->csb_klim2 *= rho0 / rho_actual;
  }
  d->full_only = d->full_only || d->csb_esimple1 || d->csb_esimple2 || d->csb_hwang1 || d->csb_hwang2 ||
                 d->csb_hl1 || d->csb_hl2 ||
                 (d->csb_edge1 && d->csb_klim1 > 0 && (d->csb_esimple1 || cb->edge_order <= 1)) ||
                 (d->csb_edge2 && d->csb_klim2 > 0 && (d->csb_esimple2 || cb->edge_order <= 1));
  d->csb_e1r = (cb->edgeFlags & EH_BEND_EDGE1_EFFECTS) && e1 != 0;
  d->csb_e2r = (cb->edgeFlags & EH_BEND_EDGE2_EFFECTS) && e2 != 0;
  d->e1 = e1;
  d->e2 = e2;
  d->tan_e1 = eh_tan(e1);
  d->tan_e2 = eh_tan(e2);
  /* EDGE_ORDER >= 2: apply_edge_effects (src/csbend.c:1135-1138, :1316-1318) */
  d->csb_eorder2 = cb->edge_order >= 2;
  d->csb_n = -b[1] / h;
  d->csb_sec1 = 1. / eh_cos(e1);
  d->csb_sec2 = 1. / eh_cos(e2);
  d->csb_he1 = cb->h1;
  d->csb_he2 = cb->h2;
  d->psi1 = psi1;
  d->psi2 = psi2;
  d->length = cb->length;
  return EH_OK;
}

/* RFCA constants, src/simple_rfca.c:378-470, once the fiducial phase is known */
static long build_rfca(eh_desc *d, const EH_RFCA *rf, const eh_ctx &c) {
  if (rf->change_t && rf->Q)
    return fail(EH_ERR_ARG, "incompatible RF cavity parameters: change_t!=0 && Q!=0");
  d->rf_flen = rf->length;
  if (rf->volt == 0) {
    d->rf_zero_volt = 1;
    return EH_OK;
  }
  double length = rf->length;
  double omega = 2 * EH_PI * rf->freq;
  double volt = rf->volt / (1e6 * EH_ME_MEV * 1.0);
  long nKicks = length ? rf->nKicks : 1;
  int matrix = 0;
  if (nKicks <= 0) {
    matrix = 1;
    nKicks = 1;
  }
  length /= nKicks;
  volt /= nKicks;
  d->rf_matrix = matrix;
  d->rf_nkicks = (int32_t)nKicks;
  d->rf_len = length;
  d->rf_volt = volt;
  d->rf_omega = omega;
  d->rf_dtLight = length / EH_C_MKS;
  double phase = c.rf_linked ? c.rf_ref_phase : rf->phase_fiducial;
  d->rf_t0 = omega ? -rf->phase_fiducial / omega : 0;
  phase += rf->phase * EH_PI / 180.0;
  d->rf_phase = phase;
  d->rf_same = (omega == 0);
  d->rf_dgamma_same = d->rf_same ? volt * eh_sin(phase) : 0;
  d->rf_standing = rf->standingWave;
  d->rf_end1 = rf->end1Focus;
  d->rf_end2 = rf->end2Focus;
  d->rf_dx = rf->dx;
  d->rf_dy = rf->dy;
  d->rf_toff = rf->change_t ? c.rf_toff : 0.0;
  d->rf_tau = omega ? rf->Q / omega : 0;
  /* LINEARIZE / LOCK_PHASE, src/simple_rfca.c:500-556, with the beam's mean
   * arrival time found at this visit */
  if (rf->linearize || rf->lockPhase) {
    if (nKicks != 1)
      return fail(EH_ERR_ARG, "Must use n_kicks=1 for linearized rf cavity");
    const double toff = d->rf_toff, tAve = c.rf_tave;
    double dgammaAve;
    if (rf->lockPhase) {
      phase = EH_PI / 180 * rf->phase;
      d->rf_phase = phase;
      dgammaAve = volt * eh_sin(phase);
    } else
      dgammaAve = volt * eh_sin(omega * (tAve - toff) + phase);
    d->rf_lin = rf->linearize;
    d->rf_tlock = rf->lockPhase ? tAve : 0.0;
    d->rf_tave = tAve;
    d->rf_dgave = dgammaAve;
    d->rf_cosk = eh_cos(omega * (tAve - toff) + phase);
    d->rf_cosm = eh_cos(omega * tAve + phase);
  }
  return EH_OK;
}

/* MULT, src/multipole.c:439-658: the slice strength, the half drift, the
 * misalignment faces and the expansion coefficients of its order
 * (expansion_coefficients, src/multipole.h:76-145); a zero strength is the
 * exact drift the reference takes instead */
static long build_mult(eh_desc *d, const EH_MULT *m, double Po) {
  if (m->nSlices <= 0)
    return fail(EH_ERR_UNSUPPORTED, "N_SLICES (or N_KICKS) <=0  in multipole_tracking()");
  if (m->order < 0)
    return fail(EH_ERR_UNSUPPORTED, "order < 0 in multipole_tracking()");
  if (m->order > EH_MULT_MAXORDER)
    return fail(EH_ERR_UNSUPPORTED, "MULT order above %d not supported", EH_MULT_MAXORDER);
  if (m->bore)
    return fail(EH_ERR_UNSUPPORTED, "MULT with BORE/BTIPL not supported");
  const long nSlices = m->nSlices;
  const double drift = m->length / nSlices / 2;
  const double KnL = m->KnL * m->factor / nSlices;
  d->length = m->length;
  if (KnL == 0) {
    /* exactDrift over the length; with no length the routine returns at once */
    if (m->length != 0)
      d->type = EH_T_EDRIFT;
    return EH_OK;
  }
  d->nslices = (int32_t)nSlices;
  d->expand = m->expandHamiltonian;
  d->mdrift = drift;
  d->rad_coef = m->synch_rad ? rad_coef_multipole(Po) : 0.0;
  d->rad = d->rad_coef != 0;
  d->kick[0][0] = KnL;
  d->mp_order[0] = m->order;
  const int n = m->order;
  for (int i = 0; i <= n; i++) {
    double fi = 1, fn = 1;
    for (long k = i; k > 1; k--)
      fi *= k;
    for (long k = n - i; k > 1; k--)
      fn *= k;
    d->mult_coef[i] = n == 0   ? 1.0
                      : n == 1 ? 1.0
                      : n == 2 ? (i == 0 ? 1. / 2 : i == 1 ? 1. : -1. / 2)
                               : (((i / 2) & 1) ? -1.0 : 1.0) / (fi * fn);
  }
  d->mdx = m->dx;
  d->mdy = m->dy;
  d->mdz = m->dz;
  d->has_offset = (m->dx || m->dy || m->dz);
  d->has_tilt = tilt_on(m->tilt);
  rot_cs(m->tilt, &d->ncos, &d->nsin);
  rot_cs(-m->tilt, &d->xcos, &d->xsin);
  return EH_OK;
}

/* ROTATE, rotation_matrix (src/tilt_matrices.c:81-117): its cos and sin entries */
static long build_rotate(eh_desc *d, const EH_ROTATE *r) {
  const double tilt = r->tilt;
  d->on_pass = r->on_pass < 0 ? -1 : r->on_pass;
  if (std::fabs(tilt - EH_PI / 2) < 1e-12) {
    d->nsin = 1;
    d->ncos = 0;
  } else if (std::fabs(tilt - EH_PI) < 1e-12) {
    d->nsin = 0;
    d->ncos = -1;
  } else {
    d->nsin = eh_sin(tilt);
    d->ncos = eh_cos(tilt);
  }
  return EH_OK;
}

/* MALIGN, offset_beam (src/do_tracking.c:3311-3372), all particles */
static long build_malign(eh_desc *d, const EH_MALIGN *m) {
  if (m->startPID != -1 || m->endPID != -1)
    return fail(EH_ERR_UNSUPPORTED, "MALIGN particle-ID ranges not supported");
  d->on_pass = m->on_pass == -1 ? -1 : m->on_pass < -1 ? -2 : m->on_pass;
  d->mdx = m->dx;
  d->mdy = m->dy;
  d->mdz = m->dz;
  d->mal_dxp = m->dxp;
  d->mal_dyp = m->dyp;
  d->mal_dt = m->dt;
  d->mal_dp = m->dp;
  d->mal_de = m->de;
  d->mal_energy = (m->dt || m->dp || m->de);
  return EH_OK;
}

/* the reference momentum an ENERGY element sets (src/do_tracking.c:1861-1881):
 * 1 with *P_new, 0 if it sets none; EH_MATCH_BEAMLINE / EH_MATCH_PARTICLES for
 * the matching modes (do_match_energy with change_beam 0 / 1; no fiducial beam
 * flags, so MATCH_BEAMLINE always matches the reference momentum to the beam) */
enum { EH_MATCH_BEAMLINE = 2, EH_MATCH_PARTICLES = 3 };
static long energy_target(const EH_ENERGY *en, double *P_new) {
  if (en->match_beamline && en->match_particles)
    return fail(EH_ERR_ARG, "can't match_beamline AND match_particles for ENERGY element");
  if (en->match_beamline)
    return EH_MATCH_BEAMLINE;
  if (en->match_particles)
    return EH_MATCH_PARTICLES;
  if (en->central_energy) {
    *P_new = std::sqrt(hsqr(en->central_energy + 1) - 1);
    return 1;
  }
  if (en->central_momentum) {
    *P_new = en->central_momentum;
    return 1;
  }
  return 0;
}

/* the exponents of an elliptical MAXAMP, checked as elimit_amplitudes does
 * (src/limit_amplitudes.c:521-536): even and >= 2, YEXPONENT 0 meaning EXPONENT */
static long ell_exponents(const EH_MAXAMP *m, int *xe, int *ye) {
  if (m->exponent < 2 || m->exponent % 2)
    return fail(EH_ERR_ARG, "MAXAMP: EXPONENT=%d is not valid; give an even integer >= 2", (int)m->exponent);
  *xe = *ye = m->exponent;
  if (m->y_exponent) {
    if (m->y_exponent < 2 || m->y_exponent % 2)
      return fail(EH_ERR_ARG, "MAXAMP: YEXPONENT=%d is not valid; give an even integer >= 2", (int)m->y_exponent);
    *ye = m->y_exponent;
  }
  return EH_OK;
}

static double elem_length(const EH_ELEMENT_LIST *e) {
  if (e->type == EH_T_MARK || e->type == EH_T_MAXAMP || e->type == EH_T_ROTATE || e->type == EH_T_MALIGN ||
      e->type == EH_T_ENERGY)
    return 0;
  return *(const double *)e->p_elem;
}

static long build_desc(eh_desc *d, const EH_ELEMENT_LIST *e, const eh_ctx &c, std::vector<eh_fterm> &pool) {
  std::memset(d, 0, sizeof *d);
  d->type = e->type;
  d->counted = !(e->type == EH_T_MARK || e->type == EH_T_MAXAMP);
  d->adj_pk = d->adj_pdx = d->adj_pdy = d->adj_pfse = d->adj_anchor = -1;
  d->Po = c.Po;
  d->zpos[0] = c.z_start;
  d->zpos[1] = c.z_end;
  for (int pi = 0; pi < 2; pi++) {
    const EH_MAXAMP *m = c.ap[pi];
    d->ap_xmax[pi] = m ? m->x_max : 0;
    d->ap_ymax[pi] = m ? m->y_max : 0;
    d->post_ap[pi] = c.post_on[pi];
    d->post_xmax[pi] = c.post_x[pi];
    d->post_ymax[pi] = c.post_y[pi];
    if (m && m->elliptical && (m->x_max || m->y_max)) {
      int xe, ye;
      long r = ell_exponents(m, &xe, &ye);
      if (r < 0)
        return r;
      /* with a non-positive limit both checks are the rectangular ones
       * (src/multipole.c:1775, src/limit_amplitudes.c:538-543) */
      const int both = m->x_max > 0 && m->y_max > 0;
      d->ap_ell[pi] = both;
      d->post_ell[pi] = both && c.post_on[pi];
      d->ell_xe[pi] = xe;
      d->ell_ye[pi] = ye;
      d->ell_a2[pi] = eh_ipow(m->x_max, xe);
      d->ell_b2[pi] = eh_ipow(m->y_max, ye);
      d->ell_any = d->ell_any || both;
    }
  }
  d->full_only = d->ell_any;
  d->extrap = e->type == EH_T_DRIF;
  switch (e->type) {
  case EH_T_MARK:
  case EH_T_MAXAMP:
    return EH_OK;
  case EH_T_DRIF: {
    const EH_DRIFT *dr = (const EH_DRIFT *)e->p_elem;
    long order = dr->order ? dr->order : 2;
    d->length = dr->length;
    d->T11 = order > 1 ? dr->length / 2 : 0;
    return EH_OK;
  }
  case EH_T_EDRIFT:
    d->length = ((const EH_EDRIFT *)e->p_elem)->length;
    return EH_OK;
  case EH_T_KQUAD:
  case EH_T_KSEXT:
  case EH_T_KOCT:
  case EH_T_KQUSE:
    return build_multipole(d, e, c);
  case EH_T_CSBEND:
    return build_csbend(d, (const EH_CSBEND *)e->p_elem, c, pool);
  case EH_T_RFCA:
    d->length = ((const EH_RFCA *)e->p_elem)->length;
    return build_rfca(d, (const EH_RFCA *)e->p_elem, c);
  case EH_T_MULT:
    return build_mult(d, (const EH_MULT *)e->p_elem, c.Po);
  case EH_T_ROTATE:
    return build_rotate(d, (const EH_ROTATE *)e->p_elem);
  case EH_T_MALIGN:
    return build_malign(d, (const EH_MALIGN *)e->p_elem);
  case EH_T_ENERGY: {
    double Pn = c.Po;
    long r = energy_target((const EH_ENERGY *)e->p_elem, &Pn);
    if (r < 0)
      return r;
    d->en_pold = c.Po;
    d->en_pnew = Pn;
    d->en_apply = r == 1 && c.Po != Pn; /* the matching modes run between launches */
    return EH_OK;
  }
  }
  return fail(EH_ERR_UNSUPPORTED, "element type %d not supported", e->type);
}

/* ------------------------------------------------------------------------- */
/* loss positions settled on the host.  multipole_tracking2 records a lost
 * particle at z_start + dzLoss, then its exit offset adds a path-length term
 * to every row (src/multipole.c:1043, :1069-1070); a particle lost before the
 * integrator's first step keeps the dzLoss the routine's previous particle left
 * behind (dzLoss is reset only after the entry checks, :1160-1232).  The kernels
 * store z_start + the particle's own dzLoss without the offset term and flag the
 * entry losses; the download replays each routine call's particle sequence and
 * settles both, so lost rows match the reference bit for bit. */
struct eh_kmp {
  int32_t kick;       /* KQUAD/KSEXT/KOCT/KQUSE */
  int32_t has_offset; /* exit offset (dx, dy or dz) */
  double z_start;
  double zfull; /* z_start + the dzLoss of a particle that ran the whole element */
  double mdz;
};

static eh_kmp kmp_of(const eh_desc &d) {
  eh_kmp k = {};
  k.kick = d.type == EH_T_KQUAD || d.type == EH_T_KSEXT || d.type == EH_T_KOCT || d.type == EH_T_KQUSE;
  if (!k.kick)
    return k;
  /* the survivor's dzLoss: every drift sub-step of every slice, in order */
  double dz = 0;
  for (int ik = 0; ik < d.nslices; ik++)
    for (int st = 0; st < d.nsub; st++)
      if (d.mdrift != 0)
        dz += d.dsh[st];
  k.has_offset = d.has_offset;
  k.z_start = d.zpos[0];
  k.zfull = d.zpos[0] + dz;
  k.mdz = d.mdz;
  return k;
}

typedef std::shared_ptr<const std::vector<eh_kmp>> eh_kmp_tab;
typedef std::shared_ptr<const std::vector<int32_t>> eh_kmp_idx;

static eh_kmp_tab kmp_table(const std::vector<eh_desc> &hd) {
  auto t = std::make_shared<std::vector<eh_kmp>>(hd.size());
  for (size_t i = 0; i < hd.size(); i++)
    (*t)[i] = kmp_of(hd[i]);
  return t;
}

/* the events [base, base + span) of one launch: (event - base) = ((pass - p0) *
 * nelem + element) * 3 + stage; a plain routine call is one event (span 1) */
struct eh_evrec {
  int32_t base, span, p0, nelem;
  eh_kmp_tab tab; /* per element, or per descriptor-pool entry when idx is set */
  eh_kmp_idx idx; /* batches: variant * nelem + element -> tab entry */
};

/* the kick-multipole record of an element-stage loss event, or null */
static const eh_kmp *kmp_at(const std::vector<eh_evrec> &log, int32_t ev, long variant) {
  auto it = std::upper_bound(log.begin(), log.end(), ev,
                             [](int32_t e, const eh_evrec &r) { return e < r.base; });
  if (it == log.begin())
    return nullptr;
  --it;
  const long rel = (long)ev - it->base;
  if (rel < 0 || rel >= it->span || rel % 3 != 0)
    return nullptr;
  const long el = (rel / 3) % it->nelem;
  long k = el;
  if (it->idx) {
    const size_t j = (size_t)variant * it->nelem + el;
    if (variant < 0 || j >= it->idx->size())
      return nullptr;
    k = (*it->idx)[j];
  }
  if (k < 0 || (size_t)k >= it->tab->size() || !(*it->tab)[k].kick)
    return nullptr;
  return &(*it->tab)[k];
}

/* order: lost identities in the sequence the reference's routines processed
 * them; pred[id]: -2 the first particle of its routine call, -1 processed after
 * a survivor, else the lost identity processed just before.  G holds the
 * graveyard's six columns (6 x nl), gk[id] the slot of identity id, gev and
 * gpass its event and pass word (whose flag is cleared here). */
static void settle_losses(const std::vector<eh_evrec> &log, long variant, const std::vector<int32_t> &order,
                          const std::vector<int32_t> &pred, const std::vector<int32_t> &gk,
                          const std::vector<int32_t> &gev, std::vector<double> &G, long nl,
                          std::vector<int32_t> &gpass) {
  if (!log.empty()) {
    std::vector<double> base(pred.size(), 0.0);
    for (int32_t id : order) {
      const long k = gk[id];
      if (k < 0)
        continue;
      const eh_kmp *m = kmp_at(log, gev[k], variant);
      if (!m)
        continue;
      double b;
      if (gpass[k] & EH_LPASS_INHERIT) {
        const int32_t q = pred[id];
        b = q == -2 ? m->z_start + 0.0 : q == -1 ? m->zfull : base[q];
      } else {
        b = G[4 * (size_t)nl + k];
      }
      base[id] = b;
      double sv = b;
      if (m->has_offset) {
        const double xp = G[1 * (size_t)nl + k], yp = G[3 * (size_t)nl + k];
        sv += (-m->mdz) * std::sqrt(1 + xp * xp + yp * yp);
      }
      G[4 * (size_t)nl + k] = sv;
    }
  }
  for (long k = 0; k < nl; k++)
    gpass[k] &= ~EH_LPASS_INHERIT;
}

/* ------------------------------------------------------------------------- */
/* device storage */
/* one position of a fused line: its z pair (z_start, z_end) and the byte
 * offset of its descriptor from the start of the position table */
struct eh_pos {
  double z[2];
  int64_t off;
  int64_t pad_;
};

struct eh_soa {
  double *x, *xp, *y, *yp, *s, *dp;
  int32_t *row0;   /* row index at upload: the particle's identity */
  int32_t *lpass;  /* pass of loss (lost particles) */
  int32_t *levent; /* loss event number, -1 while alive */
};

struct eh_beam {
  long n_total = 0, n_alive = 0, n_lost = 0;
  eh_soa buf[2] = {}; /* alive particles, double-buffered for compaction */
  int cur = 0;
  eh_soa grave = {};  /* lost particles, append-only */
  void *mem = nullptr;
  unsigned long long *d_cnt = nullptr; /* [0] alive out, [1] graveyard append, [2] lost in place, [3] effort */
  eh_desc *d_desc = nullptr;
  long desc_cap = 0;
  /* fused tracking: the position table and the distinct descriptors it points
   * into (upload_fused_descs) */
  char *d_fused = nullptr;
  long fused_cap = 0; /* bytes */
  eh_fterm *d_terms = nullptr; /* CSBEND field-term pool */
  long terms_cap = 0;
  EH_VMATRIX *d_mat = nullptr; /* transport map of eh_track_particles */
  std::vector<double> cols; /* columns 6..9 of each uploaded row, by row0 */
  long pass = 0;
  int32_t event = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double kernel_ms = 0;
  long launches = 0;
  double last_ms = 0; /* the most recent timed launch */
  double main_ms = 0, main_units = 0;
  long main_launches = 0;
  std::vector<eh_evrec> evlog; /* launches whose losses need settling */
  bool noalloc = false;         /* descriptors, terms, counters in the static scratch */
};

static int g_device = -1;

/* The static scratch of eh_do_tracking_noalloc: per GPU, about 9 MiB holding
 * the launch counters (first 256 B), the fused line's position table and
 * distinct descriptors for up to EH_SCRATCH_NDESC positions, one descriptor for
 * the per-element sweeps and the field terms of EH_SCRATCH_NBENDS CSBENDs at
 * their largest.  In host code `scratch` names
 * a host-side stand-in, not device memory: its device address on each GPU is
 * taken once with hipGetSymbolAddress (eh_init) and kept in g_scratch. */
#define EH_SCRATCH_NDESC 4096l
#define EH_SCRATCH_SINGLE_BYTES (((long)sizeof(eh_desc) + 255l) / 256l * 256l)
#define EH_SCRATCH_NBENDS 421l
#define EH_SCRATCH_FUSED_OFF 256l
#define EH_SCRATCH_FUSED_BYTES (EH_SCRATCH_NDESC * ((long)sizeof(eh_pos) + (long)sizeof(eh_desc)))
#define EH_SCRATCH_DESC_OFF (EH_SCRATCH_FUSED_OFF + EH_SCRATCH_FUSED_BYTES)
#define EH_SCRATCH_TERMS_OFF (EH_SCRATCH_DESC_OFF + EH_SCRATCH_SINGLE_BYTES)
#define EH_SCRATCH_NTERMS (EH_SCRATCH_NBENDS * EH_MAX_TERMS_PER_BEND)
#define EH_SCRATCH_BYTES (EH_SCRATCH_TERMS_OFF + EH_SCRATCH_NTERMS * (long)sizeof(eh_fterm))
/* field terms of one CSBEND at most: both sparse lists and both dense blocks */
#define EH_MAX_TERMS_PER_BEND (2l * EH_MAXFTERMS + 2l * 36l)
static __device__ __align__(256) unsigned char scratch[EH_SCRATCH_BYTES];
#define EH_MAX_DEVICES 64
static void *g_scratch[EH_MAX_DEVICES];

extern "C" int eh_init(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return (int)fail(EH_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= n)
    return (int)fail(EH_ERR_ARG, "device %d out of range (%d devices)", device, n);
  if (hipSetDevice(device) != hipSuccess)
    return (int)fail(EH_ERR_HIP, "hipSetDevice(%d) failed", device);
  if (device < EH_MAX_DEVICES && !g_scratch[device]) {
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
      return (int)fail(EH_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d", device);
    g_scratch[device] = ptr;
  }
  g_device = device;
  return EH_OK;
}

static size_t soa_bytes(long n) { return (size_t)n * (6 * sizeof(double) + 3 * sizeof(int32_t)); }

static void soa_carve(eh_soa *s, char *base, long n) {
  double *d = (double *)base;
  s->x = d;
  s->xp = d + n;
  s->y = d + 2 * n;
  s->yp = d + 3 * n;
  s->s = d + 4 * n;
  s->dp = d + 5 * n;
  int32_t *i = (int32_t *)(d + 6 * n);
  s->row0 = i;
  s->lpass = i + n;
  s->levent = i + 2 * n;
}

extern "C" eh_beam *eh_beam_create(long n_capacity) {
  if (g_device < 0 && eh_init(0) != EH_OK)
    return nullptr;
  if (n_capacity < 1)
    n_capacity = 1;
  eh_beam *b = new eh_beam();
  size_t per = soa_bytes(n_capacity);
  per = (per + 255) & ~(size_t)255;
  if (hipMalloc(&b->mem, 3 * per) != hipSuccess) {
    fail(EH_ERR_NOMEM, "hipMalloc of %zu bytes failed", 3 * per);
    delete b;
    return nullptr;
  }
  soa_carve(&b->buf[0], (char *)b->mem, n_capacity);
  soa_carve(&b->buf[1], (char *)b->mem + per, n_capacity);
  soa_carve(&b->grave, (char *)b->mem + 2 * per, n_capacity);
  if (hipMalloc(&b->d_cnt, 8 * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&b->d_mat, sizeof(EH_VMATRIX)) != hipSuccess ||
      hipStreamCreate(&b->stream) != hipSuccess || hipEventCreate(&b->ev0) != hipSuccess ||
      hipEventCreate(&b->ev1) != hipSuccess) {
    fail(EH_ERR_HIP, "beam setup failed");
    (void)hipFree(b->mem);
    delete b;
    return nullptr;
  }
  b->n_total = n_capacity;
  return b;
}

extern "C" void eh_beam_free(eh_beam *b) {
  if (!b)
    return;
  if (b->stream)
    (void)hipStreamSynchronize(b->stream);
  (void)hipFree(b->mem);
  (void)hipFree(b->d_cnt);
  if (b->d_mat)
    (void)hipFree(b->d_mat);
  if (b->d_desc)
    (void)hipFree(b->d_desc);
  if (b->d_fused)
    (void)hipFree(b->d_fused);
  if (b->d_terms)
    (void)hipFree(b->d_terms);
  if (b->ev0)
    (void)hipEventDestroy(b->ev0);
  if (b->ev1)
    (void)hipEventDestroy(b->ev1);
  if (b->stream)
    (void)hipStreamDestroy(b->stream);
  delete b;
}

extern "C" long eh_beam_n_alive(const eh_beam *b) { return b ? b->n_alive : EH_ERR_ARG; }
extern "C" long eh_beam_n_total(const eh_beam *b) { return b ? b->n_alive + b->n_lost : EH_ERR_ARG; }
extern "C" void eh_beam_set_pass(eh_beam *b, long pass) {
  if (b)
    b->pass = pass;
}

extern "C" long eh_beam_upload(eh_beam *b, double **coord, long n_part) {
  if (!b || !coord || n_part < 0 || n_part > b->n_total)
    return fail(EH_ERR_ARG, "eh_beam_upload: bad arguments (n_part=%ld, capacity=%ld)", n_part,
                b ? b->n_total : 0L);
  std::vector<double> h(6 * (size_t)n_part);
  std::vector<int32_t> r(n_part), m1(n_part, -1), z(n_part, 0);
  b->cols.assign(4 * (size_t)n_part, 0.0);
  for (long i = 0; i < n_part; i++) {
    for (int k = 0; k < 6; k++)
      h[k * (size_t)n_part + i] = coord[i][k];
    for (int k = 0; k < 4; k++)
      b->cols[4 * (size_t)i + k] = coord[i][6 + k];
    r[i] = (int32_t)i;
  }
  eh_soa &a = b->buf[0];
  b->cur = 0;
  double *dst[6] = {a.x, a.xp, a.y, a.yp, a.s, a.dp};
  for (int k = 0; k < 6; k++)
    HIPCHK(hipMemcpy(dst[k], h.data() + k * (size_t)n_part, sizeof(double) * n_part, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(a.row0, r.data(), sizeof(int32_t) * n_part, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(a.levent, m1.data(), sizeof(int32_t) * n_part, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(a.lpass, z.data(), sizeof(int32_t) * n_part, hipMemcpyHostToDevice));
  b->n_alive = n_part;
  b->n_lost = 0;
  b->event = 0;
  b->evlog.clear();
  b->pass = 0;
  b->kernel_ms = 0;
  b->launches = 0;
  return n_part;
}

/* ------------------------------------------------------------------------- */
/* kernels */
#define EH_BLOCK 256
/* Occupancy target for the tracking kernels: six wavefronts per SIMD caps
 * them at 80 VGPRs (512 / 6, in granules of 8).  The compiler's own choice
 * (~90 VGPRs, five wavefronts) leaves the FP64 pipe idle while wavefronts wait
 * on their loads and dependent FP64 chains; the few registers it then spills
 * cost less than the sixth wavefront gains (measured on the MBA ring). */
#define EH_TRACK_WAVES __attribute__((amdgpu_waves_per_eu(6)))

__device__ __forceinline__ eh_p6 load_p(const eh_soa a, long i) {
  eh_p6 p;
  p.x = a.x[i];
  p.xp = a.xp[i];
  p.y = a.y[i];
  p.yp = a.yp[i];
  p.s = a.s[i];
  p.dp = a.dp[i];
  return p;
}

__device__ __forceinline__ void store_p(const eh_soa a, long i, const eh_p6 &p) {
  a.x[i] = p.x;
  a.xp[i] = p.xp;
  a.y[i] = p.y;
  a.yp[i] = p.yp;
  a.s[i] = p.s;
  a.dp[i] = p.dp;
}

/* wavefront-aggregated atomic add: one atomic per wavefront */
__device__ __forceinline__ unsigned long long wave_add(unsigned long long *ctr, unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off, 64);
  unsigned long long base = 0;
  if (__lane_id() == 0 && v)
    base = atomicAdd(ctr, v);
  return base;
}

/* exclusive prefix of a predicate within the wavefront, and the wavefront total */
__device__ __forceinline__ unsigned wave_prefix(bool pred, unsigned *total) {
  unsigned long long m = __ballot(pred);
  *total = (unsigned)__popcll(m);
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

/* one element routine (op 0), one limit_amplitudes sweep (op 1), one
 * removeInvalidParticles sweep (op 2) or one elimit_amplitudes sweep (op 3, its
 * exponents and a2, b2 in the descriptor) over the alive particles, in place;
 * losses are marked with (pass, event) */
__global__ void __launch_bounds__(EH_BLOCK) EH_TRACK_WAVES k_plain(eh_soa a, long n, const eh_desc *__restrict__ desc,
                                                     const eh_fterm *__restrict__ terms, int op,
                                                     double xmax, double ymax, double z, double Po, int extrap,
                                                     int32_t pass, int32_t event, unsigned long long *cnt) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  bool lost = false, inherit = false;
  if (i < n) {
    eh_p6 p = load_p(a, i);
    if (op == 0) {
      const eh_desc &d = *desc;
      switch (d.type) {
      case EH_T_DRIF:
        eh_drif(p, d.length, d.T11);
        break;
      case EH_T_EDRIFT:
        eh_edrift(p, d.length);
        break;
      case EH_T_KQUAD:
      case EH_T_KSEXT:
      case EH_T_KOCT:
      case EH_T_KQUSE: {
        const int m = eh_multipole(p, d, 0, d.zpos);
        lost = m != 0;
        inherit = m == 2;
        break;
      }
      case EH_T_CSBEND:
        lost = eh_csbend(p, d, 0, terms, d.zpos);
        break;
      case EH_T_RFCA:
        lost = eh_rfca(p, d, d.zpos);
        break;
      case EH_T_MULT:
        if (d.nslices)
          lost = eh_mult(p, d, d.zpos);
        break;
      case EH_T_ROTATE:
        eh_rotate(p, d.ncos, d.nsin);
        break;
      case EH_T_MALIGN:
        eh_offset(p, d);
        break;
      case EH_T_ENERGY:
        if (d.en_apply)
          p.dp = ((1 + p.dp) * d.en_pold - d.en_pnew) / d.en_pnew;
        break;
      default:
        break;
      }
    } else if (op == 1) {
      lost = eh_limit_one(p, xmax, ymax, z, Po, extrap);
    } else if (op == 3) {
      const eh_desc &d = *desc;
      lost = eh_elimit_one(p, d.ell_a2[0], d.ell_b2[0], d.ell_xe[0], d.ell_ye[0], z, Po, extrap);
    } else {
      /* removeInvalidParticles, src/limit_amplitudes.c:289-330 */
      if (eh_isnan(p.x) || eh_isnan(p.xp) || eh_isnan(p.y) || eh_isnan(p.yp) || eh_isnan(p.s) || eh_isnan(p.dp) ||
          p.dp <= -1) {
        p.s = z;
        p.dp = Po * (1 + p.dp);
        lost = true;
      }
    }
    store_p(a, i, p);
    if (lost) {
      a.lpass[i] = pass | (inherit ? EH_LPASS_INHERIT : 0);
      a.levent[i] = event;
    }
  }
  wave_add(&cnt[2], lost ? 1ull : 0ull);
}

/* do_match_energy's particle step (src/do_tracking.c:3474-3484, :3503-3512):
 * mode 0 re-expresses delta against the new reference momentum (dp = c1 + dp*c2),
 * mode 1 moves every momentum by c1 keeping the particle's time of flight;
 * cnt[2] counts non-finite time coordinates */
__global__ void __launch_bounds__(EH_BLOCK) k_match(eh_soa a, long n, int mode, double c1, double c2, double Pc,
                                                    unsigned long long *cnt) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const double dp = a.dp[i];
  if (mode == 0) {
    a.dp[i] = c1 + dp * c2;
  } else {
    double P = (1 + dp) * Pc;
    const double t = a.s[i] / (P / eh_sqrt(P * P + 1));
    P += c1;
    a.dp[i] = (P - Pc) / Pc;
    const double s = t * (P / eh_sqrt(P * P + 1));
    a.s[i] = s;
    if (isnan(s) || isinf(s))
      atomicAdd(cnt + 2, 1ull);
  }
}

/* track_particles with one transport map over the alive particles, in place */
__global__ void __launch_bounds__(EH_BLOCK) k_matrix(eh_soa a, long n, const EH_VMATRIX *__restrict__ M) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  eh_p6 p = load_p(a, i);
  eh_vmatrix_apply(p, *M);
  store_p(a, i, p);
}

/* move survivors of a to the front of b and lost particles to the graveyard */
__global__ void __launch_bounds__(EH_BLOCK) k_compact(eh_soa a, long n, eh_soa b, eh_soa g,
                                                       unsigned long long *cnt) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  bool in = i < n;
  bool alive = in && a.levent[i] < 0;
  bool lost = in && !alive;
  unsigned na, nl;
  unsigned pa = wave_prefix(alive, &na);
  unsigned pl = wave_prefix(lost, &nl);
  unsigned long long ba = 0, bl = 0;
  if (__lane_id() == 0) {
    if (na)
      ba = atomicAdd(&cnt[0], (unsigned long long)na);
    if (nl)
      bl = atomicAdd(&cnt[1], (unsigned long long)nl);
  }
  ba = __shfl(ba, 0, 64);
  bl = __shfl(bl, 0, 64);
  if (alive) {
    long j = (long)(ba + pa);
    store_p(b, j, load_p(a, i));
    b.row0[j] = a.row0[i];
    b.lpass[j] = 0;
    b.levent[j] = -1;
  } else if (lost) {
    long j = (long)(bl + pl);
    store_p(g, j, load_p(a, i));
    g.row0[j] = a.row0[i];
    g.lpass[j] = a.lpass[i];
    g.levent[j] = a.levent[i];
  }
}

/* one particle through the element sequence from (p0, e0) up to (p1, e1)
 * exclusive, in registers.  The loop is wavefront-uniform (the element index is
 * the same in every lane), so descriptors are read with scalar loads; the
 * wavefront leaves early once all its lanes are lost.  `units` (wavefront-
 * uniform) counts the wavefront's particle-element units; a lost lane keeps
 * where and how in one word (eh_loss_code, decoded by eh_loss_words): a single
 * loop-carried register. */
template <bool RAD>
__device__ __forceinline__ void track_span(eh_p6 &p, bool &alive, const eh_pos *__restrict__ pos,
                                           const eh_fterm *__restrict__ terms, int nelem, int p0, int e0, int p1,
                                           int e1, int32_t &lcode, unsigned long long &units) {
  int pp = p0, ee = e0;
  while (pp < p1 || (pp == p1 && ee < e1)) {
    const unsigned long long live = __ballot(alive);
    if (live == 0)
      break;
    const eh_pos *q = pos + ee;
    const eh_desc *d = (const eh_desc *)((const char *)pos + q->off);
    units += (unsigned long long)__popcll(live) * (unsigned)d->counted;
    if (alive) {
      int st = eh_apply_element_t<RAD>(p, *d, pp, terms, q->z);
      if (st >= 0) {
        alive = false;
        lcode = eh_loss_code(pp - p0, nelem, ee, st);
      }
    }
    if (++ee == nelem) {
      ee = 0;
      pp++;
    }
  }
}

/* write a wavefront's particles after tracking: survivors to out at
 * out_base + atomicAdd(out_count), lost ones to the graveyard */
__device__ __forceinline__ void wave_store(const eh_p6 &p, bool in, bool alive, int32_t row0, int32_t lpass,
                                           int32_t levent, const eh_soa out, long out_base,
                                           unsigned long long *out_count, const eh_soa g,
                                           unsigned long long *grave_count) {
  bool lost = in && !alive;
  unsigned na, nl;
  unsigned pa = wave_prefix(alive, &na);
  unsigned pl = wave_prefix(lost, &nl);
  unsigned long long ba = 0, bl = 0;
  if (__lane_id() == 0) {
    if (na)
      ba = atomicAdd(out_count, (unsigned long long)na);
    if (nl)
      bl = atomicAdd(grave_count, (unsigned long long)nl);
  }
  ba = __shfl(ba, 0, 64);
  bl = __shfl(bl, 0, 64);
  if (alive) {
    long j = out_base + (long)(ba + pa);
    store_p(out, j, p);
    out.row0[j] = row0;
    out.lpass[j] = 0;
    out.levent[j] = -1;
  } else if (lost) {
    long j = (long)(bl + pl);
    store_p(g, j, p);
    g.row0[j] = row0;
    g.lpass[j] = lpass;
    g.levent[j] = levent;
  }
}

/* fused tracking of one lattice: each thread takes one particle through the span,
 * then survivors are compacted into b and lost particles appended to the graveyard.
 * Position ee runs the descriptor at byte offset pos[ee].off from pos, with its
 * z pair pos[ee].z. */
template <bool RAD>
__global__ void __launch_bounds__(EH_BLOCK) EH_TRACK_WAVES k_fused(eh_soa a, long n, eh_soa b, eh_soa g,
                                                     const eh_pos *__restrict__ pos,
                                                     const eh_fterm *__restrict__ terms, int nelem, int p0, int e0,
                                                     int p1, int e1, int32_t event_base, unsigned long long *cnt) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  bool in = i < n;
  bool alive = in;
  eh_p6 p = {0, 0, 0, 0, 0, 0};
  if (in)
    p = load_p(a, i);
  int32_t lcode = 0;
  unsigned long long units = 0;
  track_span<RAD>(p, alive, pos, terms, nelem, p0, e0, p1, e1, lcode, units);
  if (__lane_id() == 0 && units)
    atomicAdd(&cnt[3], units);
  /* identity and loss words are read and formed only now, so that none of them
   * is carried through the element loop */
  int32_t row0 = 0, lpass = 0, levent = -1;
  if (in)
    row0 = a.row0[i];
  if (in && !alive)
    eh_loss_words(lcode, nelem, p0, event_base, lpass, levent);
  wave_store(p, in, alive, row0, lpass, levent, b, 0, &cnt[0], g, &cnt[1]);
}

/* ------------------------------------------------------------------------- */
/* batched lattice variants.  Variant v owns the particle region
 * [region_off[v], region_off[v] + capacity_v) (64-aligned), its alive particles at
 * the start of it.  Work items are wavefronts of one variant each, so every
 * wavefront reads a single variant's descriptors (uniform, scalar loads).  The
 * kernel is persistent: a fixed grid pulls work items from a device counter until
 * the list is exhausted, so uneven per-variant counts need no host involvement and
 * a whole run can be captured in one graph. */
struct eh_batch_dev {
  const long *region_off;   /* [nvar] */
  const int32_t *alive_in;  /* [nvar] */
  unsigned long long *alive_out; /* [nvar] */
  const int32_t *wave_start; /* [nvar + 1] */
  int32_t *queue;
  const int32_t *desc_idx;  /* [nvar * nelem] index into the descriptor pool */
  int nvar;
};

/* the batched kernel's work list: variant v's alive particles form ceil(alive/64)
 * one-wavefront items numbered from wave_start[v]; wave_start[nvar] is the total.
 * Host and device code, so that the work list can be checked on the CPU. */
EH_HD int32_t eh_batch_wave_starts(const int32_t *alive, int nvar, int32_t *wave_start) {
  int32_t acc = 0;
  for (int v = 0; v < nvar; v++) {
    wave_start[v] = acc;
    acc += (alive[v] + 63) / 64;
  }
  wave_start[nvar] = acc;
  return acc;
}

/* the variant of work item w (0 <= w < wave_start[nvar]): the v with
 * wave_start[v] <= w < wave_start[v + 1]; variants without particles own no item */
EH_HD int eh_batch_item_variant(const int32_t *wave_start, int nvar, int32_t w) {
  int lo = 0, hi = nvar; /* wave_start[lo] <= w < wave_start[hi] */
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (wave_start[mid] <= w)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

/* work list for the next span from the previous span's survivor counts (or, on
 * the first span of a call, from alive_in as uploaded); clears the output
 * counters and the queue (a kernel, not a memset, so it can live in a graph) */
__global__ void k_batch_prep(const unsigned long long *alive_prev, int32_t *alive_in, int32_t *wave_start,
                             unsigned long long *alive_out, int32_t *queue, int nvar, int first) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    for (int v = 0; v < nvar; v++) {
      alive_in[v] = first ? alive_in[v] : (int32_t)alive_prev[v];
      alive_out[v] = 0;
    }
    eh_batch_wave_starts(alive_in, nvar, wave_start);
    *queue = 0;
  }
}

template <bool RAD>
__global__ void __launch_bounds__(EH_BLOCK) EH_TRACK_WAVES k_batch(eh_soa a, eh_soa b, eh_soa g, eh_batch_dev bd,
                                                     const eh_desc *__restrict__ pool,
                                                     const eh_fterm *__restrict__ terms, int nelem, int p0, int e0,
                                                     int p1, int e1, int32_t event_base, unsigned long long *cnt) {
  const int lane = __lane_id();
  const int32_t total = bd.wave_start[bd.nvar];
  unsigned long long units = 0;
  for (;;) {
    int32_t w = 0;
    if (lane == 0)
      w = atomicAdd(bd.queue, 1);
    /* the work item is wavefront-uniform: keep it (and all that follows from
     * it: variant, descriptor indices) in scalar registers so that descriptor
     * fields are read with scalar loads */
    w = __builtin_amdgcn_readfirstlane(__shfl(w, 0, 64));
    if (w >= total)
      break;
    const int v = eh_batch_item_variant(bd.wave_start, bd.nvar, w);
    const long local = (long)(w - bd.wave_start[v]) * 64 + lane;
    const bool in = local < bd.alive_in[v];
    const long i = bd.region_off[v] + local;
    bool alive = in;
    eh_p6 p = {0, 0, 0, 0, 0, 0};
    if (in)
      p = load_p(a, i);
    int32_t lcode = 0; /* as in track_span */
    /* descriptors of variant v: pool entries listed by desc_idx (uniform) */
    const int32_t *idx = bd.desc_idx + (long)v * nelem;
    int pp = p0, ee = e0;
    while (pp < p1 || (pp == p1 && ee < e1)) {
      const unsigned long long live = __ballot(alive);
      if (live == 0)
        break;
      const eh_desc *d = pool + __builtin_amdgcn_readfirstlane(idx[ee]);
      units += (unsigned long long)__popcll(live) * (unsigned)d->counted;
      if (alive) {
        int st = eh_apply_element_t<RAD>(p, *d, pp, terms, d->zpos);
        if (st >= 0) {
          alive = false;
          lcode = eh_loss_code(pp - p0, nelem, ee, st);
        }
      }
      if (++ee == nelem) {
        ee = 0;
        pp++;
      }
    }
    int32_t row0 = 0, lpass = 0, levent = -1;
    if (in)
      row0 = a.row0[i];
    if (in && !alive)
      eh_loss_words(lcode, nelem, p0, event_base, lpass, levent);
    wave_store(p, in, alive, row0, lpass, levent, b, bd.region_off[v], &bd.alive_out[v], g, &cnt[1]);
  }
  if (__lane_id() == 0 && units)
    atomicAdd(&cnt[3], units);
}

/* ------------------------------------------------------------------------- */
/* host helpers */
static long ensure_desc(eh_beam *b, long n) {
  if (b->desc_cap >= n)
    return EH_OK;
  if (b->noalloc)
    return fail(EH_ERR_ARG, "%ld element descriptors exceed the static scratch (%ld)", n, b->desc_cap);
  if (b->d_desc)
    HIPCHK(hipFree(b->d_desc));
  b->d_desc = nullptr;
  HIPCHK(hipMalloc(&b->d_desc, sizeof(eh_desc) * n));
  b->desc_cap = n;
  return EH_OK;
}

/* the line's descriptors for the fused kernel, in one buffer: a table of
 * positions (eh_pos: z pair, byte offset of the descriptor) followed by the
 * distinct descriptors.  Positions whose descriptors are equal apart from their
 * z share one, so a ring of identical cells reads a few dozen descriptors
 * instead of one per position. */
static long upload_fused_descs(eh_beam *b, const std::vector<eh_desc> &hd) {
  const long ne = (long)hd.size();
  std::vector<eh_desc> uniq;
  std::vector<eh_pos> pos(ne);
  std::unordered_map<std::string, int64_t> seen;
  const int64_t base = (int64_t)sizeof(eh_pos) * ne;
  for (long ie = 0; ie < ne; ie++) {
    eh_desc k;
    std::memcpy(&k, &hd[ie], sizeof k);
    pos[ie].z[0] = k.zpos[0];
    pos[ie].z[1] = k.zpos[1];
    pos[ie].pad_ = 0;
    k.zpos[0] = k.zpos[1] = 0;
    std::string key(reinterpret_cast<const char *>(&k), sizeof k);
    auto it = seen.find(key);
    if (it == seen.end()) {
      it = seen.emplace(std::move(key), base + (int64_t)sizeof(eh_desc) * (int64_t)uniq.size()).first;
      uniq.push_back(k);
    }
    pos[ie].off = it->second;
  }
  const long bytes = (long)base + (long)(sizeof(eh_desc) * uniq.size());
  if (b->fused_cap < bytes) {
    if (b->noalloc)
      return fail(EH_ERR_ARG, "the line's %ld positions exceed the static scratch", ne);
    if (b->d_fused)
      HIPCHK(hipFree(b->d_fused));
    b->d_fused = nullptr;
    b->fused_cap = 0;
    HIPCHK(hipMalloc(&b->d_fused, bytes));
    b->fused_cap = bytes;
  }
  if (ne) {
    HIPCHK(hipMemcpy(b->d_fused, pos.data(), (size_t)base, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->d_fused + base, uniq.data(), sizeof(eh_desc) * uniq.size(), hipMemcpyHostToDevice));
  }
  return EH_OK;
}

static long upload_terms(eh_beam *b, const std::vector<eh_fterm> &pool) {
  if (pool.empty())
    return EH_OK;
  if (b->terms_cap < (long)pool.size()) {
    if (b->noalloc)
      return fail(EH_ERR_ARG, "%zu field terms exceed the static scratch (%ld)", pool.size(), b->terms_cap);
    if (b->d_terms)
      HIPCHK(hipFree(b->d_terms));
    b->d_terms = nullptr;
    HIPCHK(hipMalloc(&b->d_terms, sizeof(eh_fterm) * pool.size()));
    b->terms_cap = (long)pool.size();
  }
  HIPCHK(hipMemcpy(b->d_terms, pool.data(), sizeof(eh_fterm) * pool.size(), hipMemcpyHostToDevice));
  return EH_OK;
}

static unsigned grid_for(long n) { return (unsigned)((n + EH_BLOCK - 1) / EH_BLOCK); }

static long timed_begin(eh_beam *b) {
  HIPCHK(hipEventRecord(b->ev0, b->stream));
  return EH_OK;
}

static long timed_end(eh_beam *b) {
  HIPCHK(hipEventRecord(b->ev1, b->stream));
  HIPCHK(hipEventSynchronize(b->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->kernel_ms += ms;
  b->last_ms = ms;
  b->launches++;
  return EH_OK;
}

static int32_t next_event(eh_beam *b) { return b->event++; }

/* compact after an in-place kernel that lost `lost` particles */
static long compact(eh_beam *b, long lost) {
  if (lost <= 0)
    return b->n_alive;
  unsigned long long init[2] = {0, (unsigned long long)b->n_lost};
  HIPCHK(hipMemcpyAsync(b->d_cnt, init, sizeof init, hipMemcpyHostToDevice, b->stream));
  eh_soa &src = b->buf[b->cur], &dst = b->buf[1 - b->cur];
  long n = b->n_alive;
  hipLaunchKernelGGL(k_compact, dim3(grid_for(n)), dim3(EH_BLOCK), 0, b->stream, src, n, dst, b->grave, b->d_cnt);
  HIPCHK(hipGetLastError());
  unsigned long long out[2] = {0, 0};
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipMemcpy(out, b->d_cnt, sizeof out, hipMemcpyDeviceToHost));
  b->cur = 1 - b->cur;
  b->n_alive = (long)out[0];
  b->n_lost = (long)out[1];
  return b->n_alive;
}

/* run one plain op over the alive particles; returns survivors */
static long run_plain(eh_beam *b, const eh_desc *hd, const std::vector<eh_fterm> *pool, int op, double xmax,
                      double ymax, double z, double Po, int extrap) {
  long n = b->n_alive;
  int32_t ev = next_event(b);
  if (n == 0)
    return 0;
  if (op == 0 || op == 3) {
    long r = ensure_desc(b, 1);
    if (r < 0)
      return r;
    HIPCHK(hipMemcpyAsync(b->d_desc, hd, sizeof(eh_desc), hipMemcpyHostToDevice, b->stream));
    if (pool) {
      r = upload_terms(b, *pool);
      if (r < 0)
        return r;
    }
  }
  unsigned long long zero = 0;
  HIPCHK(hipMemcpyAsync(b->d_cnt + 2, &zero, sizeof zero, hipMemcpyHostToDevice, b->stream));
  long r = timed_begin(b);
  if (r < 0)
    return r;
  hipLaunchKernelGGL(k_plain, dim3(grid_for(n)), dim3(EH_BLOCK), 0, b->stream, b->buf[b->cur], n, b->d_desc, b->d_terms, op,
                     xmax, ymax, z, Po, extrap, (int32_t)b->pass, ev, b->d_cnt);
  HIPCHK(hipGetLastError());
  r = timed_end(b);
  if (r < 0)
    return r;
  if (op == 0 && hd && hd->counted) {
    b->main_ms += b->last_ms;
    b->main_launches++;
    b->main_units += (double)n;
  }
  unsigned long long lost = 0;
  HIPCHK(hipMemcpy(&lost, b->d_cnt + 2, sizeof lost, hipMemcpyDeviceToHost));
  if (op == 0 && lost && hd) {
    const eh_kmp k = kmp_of(*hd);
    if (k.kick) {
      eh_evrec rec = {ev, 1, 0, 1, std::make_shared<const std::vector<eh_kmp>>(1, k), nullptr};
      b->evlog.push_back(rec);
    }
  }
  long before_lost = b->n_lost;
  long na = compact(b, (long)lost);
  if (na < 0)
    return na;
  if (lost && (b->n_lost - before_lost != (long)lost || na != n - (long)lost))
    return fail(EH_ERR_STATE, "compaction count mismatch: %ld in, %llu marked lost, %ld kept, %ld buried", n, lost, na,
                b->n_lost - before_lost);
  return na;
}

static long check_np(eh_beam *b, long np) {
  if (!b)
    return fail(EH_ERR_ARG, "null beam");
  if (np != b->n_alive)
    return fail(EH_ERR_ARG, "n_part=%ld but %ld particles are resident", np, b->n_alive);
  return EH_OK;
}

/* ------------------------------------------------------------------------- */
/* per-element entry points */
extern "C" long eh_multipole_tracking2(eh_beam *b, long n_part, const EH_ELEMENT_LIST *elem, double p_error,
                                       double Po, double z_start, const EH_MAXAMP *maxamp, long iSlice) {
  (void)p_error;
  long r = check_np(b, n_part);
  if (r < 0)
    return r;
  if (iSlice >= 0)
    return fail(EH_ERR_UNSUPPORTED, "slice-by-slice integration (iSlice>=0) not supported");
  eh_desc d;
  std::vector<eh_fterm> pool;
  eh_ctx c = {};
  c.Po = Po;
  c.z_start = z_start;
  c.z_end = z_start + elem_length(elem);
  c.ap[0] = c.ap[1] = maxamp;
  r = build_desc(&d, elem, c, pool);
  if (r < 0)
    return r;
  return run_plain(b, &d, &pool, 0, 0, 0, 0, Po, 0);
}

extern "C" long eh_track_through_csbend(eh_beam *b, long n_part, const EH_CSBEND *csbend, double p_error, double Po,
                                        double z_start, const EH_MAXAMP *maxamp, long iSlice) {
  (void)p_error;
  long r = check_np(b, n_part);
  if (r < 0)
    return r;
  if (iSlice >= 0)
    return fail(EH_ERR_UNSUPPORTED, "slice-by-slice integration (iSlice>=0) not supported");
  EH_ELEMENT_LIST e = {EH_T_CSBEND, 0, "CSBEND", (void *)csbend};
  eh_desc d;
  std::vector<eh_fterm> pool;
  eh_ctx c = {};
  c.Po = Po;
  c.z_start = z_start;
  c.z_end = z_start + csbend->length;
  c.ap[0] = c.ap[1] = maxamp;
  r = build_desc(&d, &e, c, pool);
  if (r < 0)
    return r;
  return run_plain(b, &d, &pool, 0, 0, 0, 0, Po, 0);
}

extern "C" long eh_multipole_tracking(eh_beam *b, long n_part, const EH_MULT *mult, double p_error, double Po,
                                      double z_start) {
  (void)p_error;
  long r = check_np(b, n_part);
  if (r < 0)
    return r;
  if (!mult)
    return fail(EH_ERR_ARG, "null MULT pointer (multipole_tracking)");
  eh_desc d;
  std::memset(&d, 0, sizeof d);
  d.type = EH_T_MULT;
  d.Po = Po;
  d.zpos[0] = z_start;
  d.zpos[1] = z_start + mult->length;
  r = build_mult(&d, mult, Po);
  if (r < 0)
    return r;
  if (d.type == EH_T_MULT && d.nslices == 0)
    return n_part;
  return run_plain(b, &d, nullptr, 0, 0, 0, 0, Po, 0);
}

extern "C" long eh_offset_beam(eh_beam *b, long n_part, const EH_MALIGN *offset, double P_central) {
  long r = check_np(b, n_part);
  if (r < 0)
    return r;
  if (!offset)
    return fail(EH_ERR_ARG, "null MALIGN pointer (offset_beam)");
  eh_desc d;
  std::memset(&d, 0, sizeof d);
  d.type = EH_T_MALIGN;
  d.Po = P_central;
  r = build_malign(&d, offset);
  if (r < 0)
    return r;
  return run_plain(b, &d, nullptr, 0, 0, 0, 0, P_central, 0);
}

extern "C" long eh_set_central_momentum(eh_beam *b, long np, double P_new, double *P_central) {
  long r = check_np(b, np);
  if (r < 0)
    return r;
  if (!P_central)
    return fail(EH_ERR_ARG, "null P_central (set_central_momentum)");
  if (np && *P_central != P_new) {
    eh_desc d;
    std::memset(&d, 0, sizeof d);
    d.type = EH_T_ENERGY;
    d.en_apply = 1;
    d.en_pold = *P_central;
    d.en_pnew = P_new;
    r = run_plain(b, &d, nullptr, 0, 0, 0, 0, P_new, 0);
    if (r < 0)
      return r;
  }
  if (!np || *P_central != P_new)
    *P_central = P_new;
  return np;
}

extern "C" long eh_rotate_beam(eh_beam *b, long n_part, const EH_ROTATE *rotate) {
  long r = check_np(b, n_part);
  if (r < 0)
    return r;
  if (!rotate)
    return fail(EH_ERR_ARG, "null ROTATE pointer");
  eh_desc d;
  std::memset(&d, 0, sizeof d);
  d.type = EH_T_ROTATE;
  r = build_rotate(&d, rotate);
  if (r < 0)
    return r;
  return run_plain(b, &d, nullptr, 0, 0, 0, 0, 0, 0);
}

extern "C" long eh_exactDrift(eh_beam *b, long np, double length) {
  long r = check_np(b, np);
  if (r < 0)
    return r;
  eh_desc d;
  std::memset(&d, 0, sizeof d);
  d.type = EH_T_EDRIFT;
  d.length = length;
  return run_plain(b, &d, nullptr, 0, 0, 0, 0, 0, 0);
}

extern "C" long eh_drift_beam(eh_beam *b, long np, double length, long order) {
  long r = check_np(b, np);
  if (r < 0)
    return r;
  if (!length)
    return np; /* drift_beam does nothing for zero length, src/do_tracking.c:3759 */
  eh_desc d;
  std::memset(&d, 0, sizeof d);
  d.type = EH_T_DRIF;
  d.length = length;
  if (order <= 0)
    order = 2;
  d.T11 = order > 1 ? length / 2 : 0;
  return run_plain(b, &d, nullptr, 0, 0, 0, 0, 0, 0);
}

extern "C" long eh_track_particles(eh_beam *b, long n_part, const EH_VMATRIX *M) {
  long r = check_np(b, n_part);
  if (r < 0)
    return r;
  if (!M)
    return fail(EH_ERR_ARG, "eh_track_particles: null matrix");
  if (M->order < 1 || M->order > 2)
    return fail(EH_ERR_UNSUPPORTED, "eh_track_particles: order %d not supported (1 and 2 are)", M->order);
  long n = b->n_alive;
  if (n == 0)
    return 0;
  HIPCHK(hipMemcpyAsync(b->d_mat, M, sizeof *M, hipMemcpyHostToDevice, b->stream));
  r = timed_begin(b);
  if (r < 0)
    return r;
  hipLaunchKernelGGL(k_matrix, dim3(grid_for(n)), dim3(EH_BLOCK), 0, b->stream, b->buf[b->cur], n, b->d_mat);
  HIPCHK(hipGetLastError());
  r = timed_end(b);
  if (r < 0)
    return r;
  return n;
}

extern "C" long eh_limit_amplitudes(eh_beam *b, double xmax, double ymax, long np, double z, double Po,
                                    long extrapolate_z, long openCode) {
  long r = check_np(b, np);
  if (r < 0)
    return r;
  if (openCode)
    return fail(EH_ERR_UNSUPPORTED, "open-sided apertures not supported");
  if (xmax < 0 && ymax < 0)
    return np;
  return run_plain(b, nullptr, nullptr, 1, xmax, ymax, z, Po, (int)extrapolate_z);
}

extern "C" long eh_elimit_amplitudes(eh_beam *b, double xmax, double ymax, long np, double z, double Po,
                                     long extrapolate_z, long openCode, long exponent, long yexponent) {
  long r = check_np(b, np);
  if (r < 0)
    return r;
  if (openCode)
    return fail(EH_ERR_UNSUPPORTED, "open-sided apertures not supported");
  EH_MAXAMP m = {};
  m.x_max = xmax;
  m.y_max = ymax;
  m.elliptical = 1;
  m.exponent = (int32_t)exponent;
  m.y_exponent = (int32_t)yexponent;
  if (exponent != m.exponent || yexponent != m.y_exponent)
    return fail(EH_ERR_ARG, "MAXAMP exponent out of range");
  int xe, ye;
  r = ell_exponents(&m, &xe, &ye);
  if (r < 0)
    return r;
  if (xmax <= 0 || ymax <= 0) {
    /* src/limit_amplitudes.c:538-543 */
    if (xmax > 0 || ymax > 0)
      return eh_limit_amplitudes(b, xmax, ymax, np, z, Po, extrapolate_z, 0);
    return np;
  }
  eh_desc d;
  std::memset(&d, 0, sizeof d);
  d.type = EH_T_MAXAMP;
  d.ell_a2[0] = eh_ipow(xmax, xe);
  d.ell_b2[0] = eh_ipow(ymax, ye);
  d.ell_xe[0] = xe;
  d.ell_ye[0] = ye;
  return run_plain(b, &d, nullptr, 3, xmax, ymax, z, Po, (int)extrapolate_z);
}

/* the beam's fiducial time at a cavity (findFiducialTime, src/simple_rfca.c:62-262,
 * serial elegant), over the alive particles in the reference's row order: TMEAN
 * as a Kahan-compensated mean (KahanPlus, src/kahan.c:27-35); PMAXIMUM from the
 * first particle whose momentum exceeds every earlier one, starting from row 0
 * (an error if row 0 has the largest); FIRST always fails, as the reference's
 * loop, which never stops early, does (src/simple_rfca.c:101-111) */
static long reference_order(eh_beam *b, std::vector<int32_t> &slot_at_row, std::vector<int32_t> &alive_row0,
                            std::vector<int32_t> &dev_index_alive, std::vector<int32_t> *pred = nullptr,
                            std::vector<int32_t> *order = nullptr);
static long fiducial_from_beam(eh_beam *b, int mode, double sOffset, double p0, double *tFid) {
  if (mode == EH_FID_FIRST)
    return fail(EH_ERR_ARG, "No available particle for the FID_MODE_FIRST mode in findFiducialTime");
  std::vector<int32_t> slot_at_row, alive_row0, idx;
  long r = reference_order(b, slot_at_row, alive_row0, idx);
  if (r < 0)
    return r;
  long n = b->n_alive;
  std::vector<double> s(n), dp(n);
  eh_soa &a = b->buf[b->cur];
  HIPCHK(hipMemcpy(s.data(), a.s, sizeof(double) * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(dp.data(), a.dp, sizeof(double) * n, hipMemcpyDeviceToHost));
  /* idx[row0] = device index of the alive particle with that identity */
  std::vector<long> dev(n);
  for (long row = 0; row < n; row++) {
    dev[row] = idx[slot_at_row[row]];
    if (dev[row] < 0)
      return fail(EH_ERR_STATE, "replay produced a lost particle among survivors");
  }
  auto t_of = [&](long k) {
    double p = p0;
    p *= 1 + dp[k];
    double beta = p / std::sqrt(p * p + 1);
    return (s[k] + sOffset) / (EH_C_MKS * beta);
  };
  if (mode == EH_FID_PMAX) {
    long ibest = -1;
    double best = n > 0 ? dp[dev[0]] : 0;
    for (long row = 0; row < n; row++)
      if (best < dp[dev[row]]) {
        best = dp[dev[row]];
        ibest = row;
      }
    if (ibest < 0)
      return fail(EH_ERR_ARG, "No available particle for RF cavity fiducialization");
    *tFid = t_of(dev[ibest]);
    return EH_OK;
  }
  double tsum = 0, error = 0.0;
  long nsum = 0;
  for (long row = 0; row < n; row++) {
    double Y = t_of(dev[row]) + error;
    double sum = tsum + Y;
    error = Y - (sum - tsum);
    tsum = sum;
    nsum++;
  }
  *tFid = nsum > 0 ? tsum / nsum : 0.0;
  return EH_OK;
}

/* do_match_energy (src/do_tracking.c:3376-3549, serial elegant with USE_KAHAN):
 * the mean of P_central*(1+delta) over the alive particles as a Kahan sum in the
 * reference's row order; change_beam 0 moves the reference momentum to it when
 * they differ by more than 1e-14 relative, change_beam 1 moves every particle's
 * momentum so that the mean is the reference momentum.  A non-finite time
 * coordinate is an error (the reference aborts). */
static long match_energy(eh_beam *b, double *Pc, long change_beam) {
  const long n = b->n_alive;
  if (n == 0)
    return 0;
  std::vector<int32_t> slot_at_row, alive_row0, idx;
  long r = reference_order(b, slot_at_row, alive_row0, idx);
  if (r < 0)
    return r;
  std::vector<double> dp(n);
  eh_soa &a = b->buf[b->cur];
  HIPCHK(hipMemcpy(dp.data(), a.dp, sizeof(double) * n, hipMemcpyDeviceToHost));
  const double P0 = *Pc;
  double P_average = 0, error = 0.0;
  for (long row = 0; row < n; row++) {
    const int32_t k = idx[slot_at_row[row]];
    if (k < 0)
      return fail(EH_ERR_STATE, "replay produced a lost particle among survivors");
    double Y = P0 * (1 + dp[k]) + error;
    double sum = P_average + Y;
    error = Y - (sum - P_average);
    P_average = sum;
  }
  P_average /= n;
  int mode;
  double c1, c2 = 0;
  if (!change_beam) {
    if (!(std::fabs(P_average - P0) / P0 > 1e-14))
      return n;
    mode = 0;
    c1 = (P0 - P_average) / P_average;
    c2 = P0 / P_average;
  } else {
    mode = 1;
    c1 = P0 - P_average;
  }
  unsigned long long zero = 0;
  HIPCHK(hipMemcpyAsync(b->d_cnt + 2, &zero, sizeof zero, hipMemcpyHostToDevice, b->stream));
  r = timed_begin(b);
  if (r < 0)
    return r;
  hipLaunchKernelGGL(k_match, dim3(grid_for(n)), dim3(EH_BLOCK), 0, b->stream, a, n, mode, c1, c2, P0, b->d_cnt);
  HIPCHK(hipGetLastError());
  r = timed_end(b);
  if (r < 0)
    return r;
  if (mode == 1) {
    unsigned long long bad = 0;
    HIPCHK(hipMemcpy(&bad, b->d_cnt + 2, sizeof bad, hipMemcpyDeviceToHost));
    if (bad)
      return fail(EH_ERR_STATE, "bad time coordinate for %llu particles in do_match_energy", bad);
  } else
    *Pc = P_average;
  return n;
}

extern "C" long eh_do_match_energy(eh_beam *b, long np, double *P_central, long change_beam) {
  long r = check_np(b, np);
  if (r < 0)
    return r;
  if (!P_central)
    return fail(EH_ERR_ARG, "null P_central (do_match_energy)");
  return match_energy(b, P_central, change_beam);
}

/* CHANGE_T's time offset at a cavity visit (src/simple_rfca.c:488-498): the
 * whole RF periods in the arrival time of the particle in row 0 of the
 * reference's arrangement, when that time exceeds 0.9 periods */
static long rf_time_offset(eh_beam *b, const EH_RFCA *rf, double Pc, double *toff) {
  *toff = 0;
  const double omega = 2 * EH_PI * rf->freq;
  if (!(omega && rf->change_t && rf->volt != 0 && b->n_alive > 0))
    return EH_OK;
  std::vector<int32_t> slot_at_row, alive_row0, idx;
  long r = reference_order(b, slot_at_row, alive_row0, idx);
  if (r < 0)
    return r;
  const int32_t k = idx[slot_at_row[0]];
  if (k < 0)
    return fail(EH_ERR_STATE, "replay produced a lost particle in row 0");
  double s, dp;
  eh_soa &a = b->buf[b->cur];
  HIPCHK(hipMemcpy(&s, a.s + k, sizeof s, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&dp, a.dp + k, sizeof dp, hipMemcpyDeviceToHost));
  const double To = 2 * EH_PI / omega;
  double P = Pc * (1 + dp);
  double beta_i = P / std::sqrt(P * P + 1);
  double t = s / (EH_C_MKS * beta_i);
  if (t > (0.9 * To))
    *toff = ((long)(t / To + 0.5)) * To;
  return EH_OK;
}

/* LINEARIZE / LOCK_PHASE's mean arrival time at a cavity visit
 * (src/simple_rfca.c:503-521): a Kahan sum in the reference's row order of
 * (s + L/2) / (c beta) - timeOffset over the alive particles */
static long rf_mean_time(eh_beam *b, const EH_RFCA *rf, double Pc, double toff, double *tAve) {
  *tAve = 0;
  if (!(rf->linearize || rf->lockPhase) || rf->volt == 0 || b->n_alive == 0)
    return EH_OK;
  std::vector<int32_t> slot_at_row, alive_row0, idx;
  long r = reference_order(b, slot_at_row, alive_row0, idx);
  if (r < 0)
    return r;
  const long n = b->n_alive;
  std::vector<double> s(n), dp(n);
  eh_soa &a = b->buf[b->cur];
  HIPCHK(hipMemcpy(s.data(), a.s, sizeof(double) * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(dp.data(), a.dp, sizeof(double) * n, hipMemcpyDeviceToHost));
  const double length = rf->length; /* n_kicks is 1 */
  double sum = 0, error = 0.0;
  for (long row = 0; row < n; row++) {
    const int32_t k = idx[slot_at_row[row]];
    if (k < 0)
      return fail(EH_ERR_STATE, "replay produced a lost particle among survivors");
    double P = Pc * (1 + dp[k]);
    double beta_i = P / std::sqrt(P * P + 1);
    double t = (s[k] + length / 2) / (EH_C_MKS * beta_i) - toff;
    double Y = t + error;
    double s1 = sum + Y;
    error = Y - (s1 - sum);
    sum = s1;
  }
  *tAve = sum / n;
  return EH_OK;
}

/* phase references, src/phase_reference.c:28-143: numbered groups of
 * time-dependent elements sharing one reference phase, set by the first cavity
 * of the group tracked with particles and kept until eh_delete_phase_references.
 * A cavity with PHASE_REFERENCE 0 is its own group (the reference gives it an
 * unused number; the phase it then tracks with is always its own fiducial). */
struct eh_phase_ref {
  int64_t ref;
  int set;
  double phase;
};
static std::vector<eh_phase_ref> g_phase_refs;

extern "C" long eh_get_phase_reference(int64_t ref, double *phase) {
  if (ref == 0 || !phase)
    return 0;
  for (const eh_phase_ref &g : g_phase_refs)
    if (g.ref == ref) {
      if (!g.set)
        return EH_REF_PHASE_NOT_SET;
      *phase = g.phase;
      return EH_REF_PHASE_RETURNED;
    }
  return EH_REF_PHASE_NONEXISTENT;
}

extern "C" long eh_set_phase_reference(int64_t ref, double phase) {
  if (ref == 0)
    return 0;
  if (ref > INT64_MAX / 2 || ref < 0)
    return fail(EH_ERR_ARG, "please use a small integer for the phase_reference number");
  for (eh_phase_ref &g : g_phase_refs)
    if (g.ref == ref) {
      g.phase = phase;
      g.set = 1;
      return 1;
    }
  g_phase_refs.push_back({ref, 1, phase});
  return 1;
}

extern "C" long eh_delete_phase_references(void) {
  for (eh_phase_ref &g : g_phase_refs)
    g.set = 0;
  return EH_OK;
}

/* an RF cavity that may need the beam for its phase on its first pass: it
 * takes a fiducial there unless it has one or its group's phase is set by
 * then (src/simple_rfca.c:414-463; a cavity without voltage takes none) */
static bool rf_pending(const EH_ELEMENT_LIST *e) {
  if (e->type != EH_T_RFCA)
    return false;
  const EH_RFCA *rf = (const EH_RFCA *)e->p_elem;
  double ph;
  if (rf->phase_reference && eh_get_phase_reference(rf->phase_reference, &ph) == EH_REF_PHASE_RETURNED)
    return false;
  return !rf->fiducial_seen && rf->volt != 0;
}

/* fiducial time of a cavity's first pass: from the beam (fid_of, for TMEAN,
 * PMAXIMUM and FIRST), LIGHT from its position, or T_REFERENCE
 * (src/simple_rfca.c:414-463) */
template <class BeamFid>
static long rf_take_fiducial(EH_RFCA *rf, double zEnd, BeamFid fid_of) {
  if (rf->fiducial_mode < EH_FID_TMEAN || rf->fiducial_mode > EH_FID_PMAX)
    return fail(EH_ERR_ARG, "invalid fiducial mode for RFCA element");
  double length = rf->length;
  long nk = length ? rf->nKicks : 1;
  if (nk <= 0)
    nk = 1;
  length /= nk;
  double t0;
  if (rf->tReference != -1)
    t0 = rf->tReference;
  else if (rf->fiducial_mode == EH_FID_LIGHT)
    t0 = ((zEnd - rf->length) + length / 2) / EH_C_MKS;
  else {
    long r = fid_of(rf->fiducial_mode, length / 2, &t0);
    if (r < 0)
      return r;
  }
  rf->phase_fiducial = -(2 * EH_PI * rf->freq) * t0;
  rf->fiducial_seen = 1;
  if (rf->freq != 0)
    rf->t_fiducial = -rf->phase_fiducial / (2 * EH_PI * rf->freq);
  return EH_OK;
}

/* the phase a cavity with voltage tracks with at a call with particles
 * (src/simple_rfca.c:419-466): its group's if set, else its own fiducial (taken
 * from the beam b if it has none), which then sets the group.  c gets the
 * group's phase for build_desc. */
static long rf_resolve(eh_beam *b, EH_RFCA *rf, double Pc, double zEnd, eh_ctx &c) {
  double ph = 0;
  c.rf_linked = 0;
  if (rf->phase_reference && eh_get_phase_reference(rf->phase_reference, &ph) == EH_REF_PHASE_RETURNED) {
    c.rf_linked = 1;
    c.rf_ref_phase = ph;
    return EH_OK;
  }
  if (!rf->fiducial_seen) {
    long r = rf_take_fiducial(rf, zEnd, [&](int mode, double sOffset, double *t0) {
      return fiducial_from_beam(b, mode, sOffset, Pc, t0);
    });
    if (r < 0)
      return r;
  }
  if (rf->phase_reference) {
    long r = eh_set_phase_reference(rf->phase_reference, rf->phase_fiducial);
    if (r < 0)
      return r;
  }
  return EH_OK;
}

extern "C" long eh_simple_rf_cavity(eh_beam *b, long np, EH_RFCA *rfca, double *P_central, double zEnd) {
  long r = check_np(b, np);
  if (r < 0)
    return r;
  if (np <= 0)
    return np;
  eh_ctx c = {};
  if (rfca->volt != 0) {
    r = rf_resolve(b, rfca, *P_central, zEnd, c);
    if (r < 0)
      return r;
    if (rfca->freq != 0)
      rfca->t_fiducial = -rfca->phase_fiducial / (2 * EH_PI * rfca->freq);
    r = rf_time_offset(b, rfca, *P_central, &c.rf_toff);
    if (r < 0)
      return r;
    r = rf_mean_time(b, rfca, *P_central, c.rf_toff, &c.rf_tave);
    if (r < 0)
      return r;
  }
  EH_ELEMENT_LIST e = {EH_T_RFCA, 0, "RFCA", (void *)rfca};
  eh_desc d;
  std::vector<eh_fterm> pool;
  c.Po = *P_central;
  c.z_start = zEnd - rfca->length;
  c.z_end = zEnd;
  r = build_desc(&d, &e, c, pool);
  if (r < 0)
    return r;
  long nLeft = run_plain(b, &d, nullptr, 0, 0, 0, 0, *P_central, 0);
  if (nLeft > 0 && rfca->change_p0 && rfca->volt != 0) {
    r = match_energy(b, P_central, 0);
    if (r < 0)
      return r;
  }
  return nLeft;
}

extern "C" long eh_removeInvalidParticles(eh_beam *b, long np, double z, double Po) {
  long r = check_np(b, np);
  if (r < 0)
    return r;
  return run_plain(b, nullptr, nullptr, 2, 0, 0, z, Po, 0);
}

/* ------------------------------------------------------------------------- */
/* reference row order: replay of swap-to-top compaction
 * (src/multipole.c:1013-1049, src/csbend.c:1088-1267, src/limit_amplitudes.c:226-285,
 * :294-327).  Within one compaction event the reference scans rows upward and
 * swaps every lost row with the current top row, re-examining the row it just
 * filled; the chain below reproduces that sequence touching only O(lost) rows. */
static void replay_order(long N, std::vector<std::pair<int32_t, int32_t>> &lost, std::vector<int32_t> &slot_at,
                         std::vector<int32_t> *pred = nullptr, std::vector<int32_t> *order = nullptr) {
  if (pred) {
    pred->assign(N, -1);
    order->clear();
  }
  slot_at.resize(N);
  std::vector<int32_t> pos(N);
  for (long r = 0; r < N; r++)
    slot_at[r] = pos[r] = (int32_t)r;
  std::vector<char> lostnow(N, 0);
  std::stable_sort(lost.begin(), lost.end(),
                   [](const std::pair<int32_t, int32_t> &a, const std::pair<int32_t, int32_t> &b) {
                     return a.first < b.first;
                   });
  long nTo = N;
  std::vector<int32_t> rows;
  size_t g0 = 0;
  while (g0 < lost.size()) {
    size_t g1 = g0;
    while (g1 < lost.size() && lost[g1].first == lost[g0].first)
      g1++;
    rows.clear();
    for (size_t k = g0; k < g1; k++) {
      int32_t id = lost[k].second;
      lostnow[id] = 1;
      rows.push_back(pos[id]);
    }
    std::sort(rows.begin(), rows.end());
    long top = nTo - 1;
    for (size_t k = 0; k < rows.size(); k++) {
      long i = rows[k];
      if (i > top)
        break;
      int32_t prev = i == 0 ? -2 : -1; /* rows below i ended with survivors */
      for (;;) {
        if (pred) {
          const int32_t cur = slot_at[i];
          (*pred)[cur] = prev;
          order->push_back(cur);
          prev = cur;
        }
        if (i == top) {
          top--;
          break;
        }
        int32_t a = slot_at[i], c = slot_at[top];
        slot_at[i] = c;
        slot_at[top] = a;
        pos[c] = (int32_t)i;
        pos[a] = (int32_t)top;
        top--;
        if (!lostnow[slot_at[i]])
          break;
      }
      if (i > top)
        break;
    }
    for (size_t k = g0; k < g1; k++)
      lostnow[lost[k].second] = 0;
    nTo -= (long)(g1 - g0);
    g0 = g1;
  }
}

/* slot_at_row[r] = identity (upload row) of the particle in row r of the
 * reference's arrangement; idx[identity] = index in the alive buffer or -1 */
static long reference_order(eh_beam *b, std::vector<int32_t> &slot_at_row, std::vector<int32_t> &alive_row0,
                            std::vector<int32_t> &idx, std::vector<int32_t> *pred, std::vector<int32_t> *order) {
  long N = b->n_alive + b->n_lost;
  alive_row0.resize(b->n_alive);
  std::vector<int32_t> grow0(b->n_lost), gev(b->n_lost);
  HIPCHK(hipStreamSynchronize(b->stream));
  if (b->n_alive)
    HIPCHK(hipMemcpy(alive_row0.data(), b->buf[b->cur].row0, sizeof(int32_t) * b->n_alive, hipMemcpyDeviceToHost));
  if (b->n_lost) {
    HIPCHK(hipMemcpy(grow0.data(), b->grave.row0, sizeof(int32_t) * b->n_lost, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(gev.data(), b->grave.levent, sizeof(int32_t) * b->n_lost, hipMemcpyDeviceToHost));
  }
  idx.assign(N, -1);
  for (long k = 0; k < b->n_alive; k++) {
    if (alive_row0[k] < 0 || alive_row0[k] >= N)
      return fail(EH_ERR_STATE, "corrupt particle identity");
    idx[alive_row0[k]] = (int32_t)k;
  }
  std::vector<std::pair<int32_t, int32_t>> lost(b->n_lost);
  for (long k = 0; k < b->n_lost; k++) {
    if (grow0[k] < 0 || grow0[k] >= N)
      return fail(EH_ERR_STATE, "corrupt particle identity");
    lost[k] = std::make_pair(gev[k], grow0[k]);
  }
  replay_order(N, lost, slot_at_row, pred, order);
  return EH_OK;
}

extern "C" long eh_beam_download(eh_beam *b, double **coord, long n_part) {
  if (!b || !coord)
    return fail(EH_ERR_ARG, "eh_beam_download: null argument");
  long N = b->n_alive + b->n_lost;
  if (n_part != N)
    return fail(EH_ERR_ARG, "eh_beam_download: n_part=%ld but the beam holds %ld particles", n_part, N);
  std::vector<int32_t> slot_at_row, alive_row0, idx, pred, order;
  long r = reference_order(b, slot_at_row, alive_row0, idx, &pred, &order);
  if (r < 0)
    return r;
  long na = b->n_alive, nl = b->n_lost;
  std::vector<double> A(6 * (size_t)na), G(6 * (size_t)nl);
  std::vector<int32_t> grow0(nl), gpass(nl);
  const eh_soa &a = b->buf[b->cur], &g = b->grave;
  const double *as[6] = {a.x, a.xp, a.y, a.yp, a.s, a.dp}, *gs[6] = {g.x, g.xp, g.y, g.yp, g.s, g.dp};
  for (int k = 0; k < 6; k++) {
    if (na)
      HIPCHK(hipMemcpy(A.data() + k * (size_t)na, as[k], sizeof(double) * na, hipMemcpyDeviceToHost));
    if (nl)
      HIPCHK(hipMemcpy(G.data() + k * (size_t)nl, gs[k], sizeof(double) * nl, hipMemcpyDeviceToHost));
  }
  if (nl) {
    HIPCHK(hipMemcpy(grow0.data(), g.row0, sizeof(int32_t) * nl, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(gpass.data(), g.lpass, sizeof(int32_t) * nl, hipMemcpyDeviceToHost));
  }
  std::vector<int32_t> gidx(N, -1), gev(nl);
  for (long k = 0; k < nl; k++)
    gidx[grow0[k]] = (int32_t)k;
  if (nl)
    HIPCHK(hipMemcpy(gev.data(), g.levent, sizeof(int32_t) * nl, hipMemcpyDeviceToHost));
  settle_losses(b->evlog, -1, order, pred, gidx, gev, G, nl, gpass);
  for (long row = 0; row < N; row++) {
    int32_t id = slot_at_row[row];
    double *c = coord[row];
    for (int k = 0; k < 4; k++)
      c[6 + k] = b->cols[4 * (size_t)id + k];
    if (idx[id] >= 0) {
      long k = idx[id];
      for (int q = 0; q < 6; q++)
        c[q] = A[q * (size_t)na + k];
    } else {
      long k = gidx[id];
      if (k < 0)
        return fail(EH_ERR_STATE, "particle %d missing on download", id);
      for (int q = 0; q < 6; q++)
        c[q] = G[q * (size_t)nl + k];
      c[EH_LOSS_PASS_INDEX] = gpass[k];
    }
  }
  return na;
}

extern "C" long eh_beam_row_order(eh_beam *b, int32_t *upload_row, long n_part) {
  if (!b || !upload_row)
    return fail(EH_ERR_ARG, "eh_beam_row_order: null argument");
  long N = b->n_alive + b->n_lost;
  if (n_part != N)
    return fail(EH_ERR_ARG, "eh_beam_row_order: n_part=%ld but the beam holds %ld particles", n_part, N);
  std::vector<int32_t> slot_at_row, alive_row0, idx;
  long r = reference_order(b, slot_at_row, alive_row0, idx);
  if (r < 0)
    return r;
  for (long row = 0; row < N; row++)
    upload_row[row] = slot_at_row[row];
  return b->n_alive;
}

/* ------------------------------------------------------------------------- */
/* do_tracking */
static long track_plain(eh_beam *b, EH_ELEMENT_LIST *bl, long ne, long n_passes, double *Pc, double *effort) {
  long nToTrack = b->n_alive, nLeft;
  const EH_MAXAMP *maxamp = nullptr;
  double x_max = 0, y_max = 0;
  long elliptical = 0, maxampExponent = 0, maxampYExponent = 0;
  b->pass = 0;
  nLeft = eh_limit_amplitudes(b, EH_DBL_MAX, EH_DBL_MAX, nToTrack, 0.0, *Pc, 0, 0);
  if (nLeft < 0)
    return nLeft;
  nToTrack = nLeft;
  for (long pass = 0; pass < n_passes; pass++) {
    b->pass = pass;
    double z = 0, z_travel = 0, last_z = 0;
    for (long ie = 0; ie < ne; ie++) {
      EH_ELEMENT_LIST *e = bl + ie;
      double len = elem_length(e);
      last_z = z;
      z += len;
      z_travel += len;
      nLeft = nToTrack;
      if (e->type != EH_T_MARK && e->type != EH_T_MAXAMP)
        *effort += nToTrack;
      switch (e->type) {
      case EH_T_MARK:
        break;
      case EH_T_MAXAMP:
        maxamp = (const EH_MAXAMP *)e->p_elem;
        x_max = maxamp->x_max;
        y_max = maxamp->y_max;
        elliptical = maxamp->elliptical;
        maxampExponent = maxamp->exponent;
        maxampYExponent = maxamp->y_exponent;
        break;
      case EH_T_DRIF: {
        const EH_DRIFT *d = (const EH_DRIFT *)e->p_elem;
        eh_desc dd;
        std::memset(&dd, 0, sizeof dd);
        dd.type = EH_T_DRIF;
        dd.length = d->length;
        long order = d->order ? d->order : 2;
        dd.T11 = order > 1 ? d->length / 2 : 0;
        nLeft = run_plain(b, &dd, nullptr, 0, 0, 0, 0, 0, 0);
        break;
      }
      case EH_T_EDRIFT:
        nLeft = eh_exactDrift(b, nToTrack, ((const EH_EDRIFT *)e->p_elem)->length);
        break;
      case EH_T_KQUAD:
      case EH_T_KSEXT:
      case EH_T_KOCT:
      case EH_T_KQUSE:
        nLeft = eh_multipole_tracking2(b, nToTrack, e, 0.0, *Pc, last_z, maxamp, -1);
        break;
      case EH_T_CSBEND:
        nLeft = eh_track_through_csbend(b, nToTrack, (const EH_CSBEND *)e->p_elem, 0.0, *Pc, last_z, maxamp, -1);
        break;
      case EH_T_RFCA:
        nLeft = eh_simple_rf_cavity(b, nToTrack, (EH_RFCA *)e->p_elem, Pc, z_travel);
        break;
      case EH_T_MULT:
        nLeft = eh_multipole_tracking(b, nToTrack, (const EH_MULT *)e->p_elem, 0.0, *Pc, last_z);
        break;
      case EH_T_ROTATE: {
        const EH_ROTATE *r = (const EH_ROTATE *)e->p_elem;
        if (r->on_pass < 0 || r->on_pass == pass)
          nLeft = eh_rotate_beam(b, nToTrack, r);
        break;
      }
      case EH_T_MALIGN: {
        const EH_MALIGN *m = (const EH_MALIGN *)e->p_elem;
        if (m->on_pass == -1 || m->on_pass == pass)
          nLeft = eh_offset_beam(b, nToTrack, m, *Pc);
        break;
      }
      case EH_T_ENERGY: {
        double Pn = *Pc;
        long t = energy_target((const EH_ENERGY *)e->p_elem, &Pn);
        if (t < 0)
          return t;
        if (t == EH_MATCH_BEAMLINE || t == EH_MATCH_PARTICLES)
          nLeft = match_energy(b, Pc, t == EH_MATCH_PARTICLES);
        else if (t > 0)
          nLeft = eh_set_central_momentum(b, nToTrack, Pn, Pc);
        break;
      }
      default:
        return fail(EH_ERR_UNSUPPORTED, "element type %d not supported", e->type);
      }
      if (nLeft < 0)
        return nLeft;
      if (e->type != EH_T_MARK && (x_max || y_max)) {
        /* src/do_tracking.c:2508-2517 */
        if (!elliptical)
          nLeft = eh_limit_amplitudes(b, x_max, y_max, nLeft, z, *Pc, e->type == EH_T_DRIF, 0);
        else
          nLeft = eh_elimit_amplitudes(b, x_max, y_max, nLeft, z, *Pc, e->type == EH_T_DRIF, 0, maxampExponent,
                                       maxampYExponent);
        if (nLeft < 0)
          return nLeft;
      }
      nLeft = eh_limit_amplitudes(b, EH_DBL_MAX, EH_DBL_MAX, nLeft, z, *Pc, 0, 0);
      if (nLeft < 0)
        return nLeft;
      nToTrack = nLeft;
    }
  }
  return nToTrack;
}

static long launch_fused(eh_beam *b, long ne, int p0, int e0, int p1, int e1, double *effort,
                         const eh_kmp_tab &tab, bool rad) {
  long n = b->n_alive;
  long span = ((long)(p1 - p0) * ne + ne) * 3 + 3;
  if ((long)b->event + span >= 2147483647L)
    return fail(EH_ERR_STATE, "loss event counter overflow");
  if (((long)(p1 - p0) * ne + ne) * 8 >= 2147483647L)
    return fail(EH_ERR_ARG, "span of %ld passes x %ld elements too long for one launch", (long)(p1 - p0 + 1), ne);
  int32_t base = b->event;
  b->event += (int32_t)span;
  if (n > 0 && tab) {
    eh_evrec rec = {base, (int32_t)span, p0, (int32_t)ne, tab, nullptr};
    b->evlog.push_back(rec);
  }
  if (n == 0 || (p0 == p1 && e0 >= e1))
    return b->n_alive;
  unsigned long long init[4] = {0, (unsigned long long)b->n_lost, 0, 0};
  HIPCHK(hipMemcpyAsync(b->d_cnt, init, sizeof init, hipMemcpyHostToDevice, b->stream));
  long r = timed_begin(b);
  if (r < 0)
    return r;
  if (rad)
    hipLaunchKernelGGL(k_fused<true>, dim3(grid_for(n)), dim3(EH_BLOCK), 0, b->stream, b->buf[b->cur], n,
                       b->buf[1 - b->cur], b->grave, (const eh_pos *)b->d_fused, b->d_terms, (int)ne, p0, e0, p1,
                       e1, base, b->d_cnt);
  else
    hipLaunchKernelGGL(k_fused<false>, dim3(grid_for(n)), dim3(EH_BLOCK), 0, b->stream, b->buf[b->cur], n,
                       b->buf[1 - b->cur], b->grave, (const eh_pos *)b->d_fused, b->d_terms, (int)ne, p0, e0, p1,
                       e1, base, b->d_cnt);
  HIPCHK(hipGetLastError());
  r = timed_end(b);
  if (r < 0)
    return r;
  unsigned long long out[4];
  HIPCHK(hipMemcpy(out, b->d_cnt, sizeof out, hipMemcpyDeviceToHost));
  if ((long)out[0] + (long)out[1] != n + b->n_lost)
    return fail(EH_ERR_STATE, "fused compaction count mismatch");
  b->cur = 1 - b->cur;
  b->n_alive = (long)out[0];
  b->n_lost = (long)out[1];
  *effort += (double)out[3];
  b->main_ms += b->last_ms;
  b->main_launches++;
  b->main_units += (double)out[3];
  return b->n_alive;
}

static long track_fused(eh_beam *b, EH_ELEMENT_LIST *bl, long ne, long n_passes, double *Pc, int tpl,
                        double *effort) {
  if (tpl < 1)
    tpl = 1;
  long r;
  std::vector<eh_desc> hd(ne);
  std::vector<eh_fterm> pool;
  std::vector<double> zs(ne), ze(ne);
  /* positions (wrap_around resets z every pass) and apertures for pass 0 / later */
  const EH_MAXAMP *last_amp = nullptr;
  {
    double z = 0;
    for (long ie = 0; ie < ne; ie++) {
      zs[ie] = z;
      z += elem_length(bl + ie);
      ze[ie] = z;
      if (bl[ie].type == EH_T_MAXAMP)
        last_amp = (const EH_MAXAMP *)bl[ie].p_elem;
    }
  }
  std::vector<const EH_MAXAMP *> amp0(ne), amp1(ne);
  {
    const EH_MAXAMP *cur0 = nullptr, *cur1 = last_amp;
    for (long ie = 0; ie < ne; ie++) {
      if (bl[ie].type == EH_T_MAXAMP)
        cur0 = cur1 = (const EH_MAXAMP *)bl[ie].p_elem;
      amp0[ie] = cur0;
      amp1[ie] = cur1;
    }
  }
  /* the reference momentum in force at each element on pass 0 (Pe0) and on
   * later passes (Pe1): ENERGY elements change it for the elements after them */
  std::vector<double> Pe0(ne), Pe1(ne);
  double P_end = *Pc;
  bool two_phase = false;
  for (int phase = 0; phase < 2; phase++) {
    double P = phase == 0 ? *Pc : P_end;
    std::vector<double> &Pe = phase == 0 ? Pe0 : Pe1;
    for (long ie = 0; ie < ne; ie++) {
      Pe[ie] = P;
      if (bl[ie].type == EH_T_ENERGY) {
        double Pn = P;
        r = energy_target((const EH_ENERGY *)bl[ie].p_elem, &Pn);
        if (r < 0)
          return r;
        if (r == 1)
          P = Pn;
      }
    }
    if (phase == 0)
      P_end = P;
  }
  for (long ie = 0; ie < ne; ie++)
    two_phase = two_phase || Pe0[ie] != Pe1[ie];
  const std::vector<double> *Pe = &Pe0;
  /* elements that need the beam at every pass: an RF cavity with CHANGE_T (its
   * time offset, judged on the particle in row 0 as the beam reaches it) and the
   * momentum matching after an RF cavity with CHANGE_P0 or an ENERGY element in
   * a matching mode.  Lines holding them are tracked in launches that end there. */
  std::vector<char> stop_before(ne, 0), stop_after(ne, 0);
  std::vector<long> match_mode(ne, 0);
  std::vector<double> toff(ne, 0.0), tave(ne, 0.0);
  bool dynamic = false;
  for (long ie = 0; ie < ne; ie++) {
    if (bl[ie].type == EH_T_RFCA) {
      const EH_RFCA *rf = (const EH_RFCA *)bl[ie].p_elem;
      stop_before[ie] = rf->volt != 0 && ((rf->change_t && rf->freq != 0) || rf->linearize || rf->lockPhase);
      stop_after[ie] = rf->volt != 0 && rf->change_p0;
    } else if (bl[ie].type == EH_T_ENERGY) {
      double Pn = 0;
      r = energy_target((const EH_ENERGY *)bl[ie].p_elem, &Pn);
      stop_after[ie] = r == EH_MATCH_BEAMLINE || r == EH_MATCH_PARTICLES;
      match_mode[ie] = r == EH_MATCH_PARTICLES;
    }
    dynamic = dynamic || stop_before[ie] || stop_after[ie];
  }
  std::vector<double> Pdyn;
  if (dynamic) {
    Pdyn = Pe0;
    Pe = &Pdyn;
  }
  /* RF cavities whose phase needs the beam at their first pass (pend), and
   * those tracking with another cavity's group phase (link, lph) */
  std::vector<char> pend(ne, 0), link(ne, 0);
  std::vector<double> lph(ne, 0.0);
  auto build = [&](long ie) -> long {
    eh_ctx c = {};
    c.rf_linked = link[ie];
    c.rf_ref_phase = lph[ie];
    c.rf_toff = toff[ie];
    c.rf_tave = tave[ie];
    c.Po = (*Pe)[ie];
    c.z_start = zs[ie];
    c.z_end = ze[ie];
    c.ap[0] = amp0[ie];
    c.ap[1] = amp1[ie];
    for (int pi = 0; pi < 2; pi++) {
      const EH_MAXAMP *m = pi ? amp1[ie] : amp0[ie];
      c.post_x[pi] = m ? m->x_max : 0;
      c.post_y[pi] = m ? m->y_max : 0;
      c.post_on[pi] = (bl[ie].type != EH_T_MARK) && m && (m->x_max || m->y_max);
    }
    long r2 = build_desc(&hd[ie], bl + ie, c, pool);
    /* an element followed by a matching: the reference sweeps its MAXAMP after
     * the matching, so the sweep runs on the host side after it */
    if (r2 >= 0 && stop_after[ie])
      hd[ie].post_ap[0] = hd[ie].post_ap[1] = 0;
    return r2;
  };
  /* initial NaN sweep, src/do_tracking.c:478-483 */
  b->pass = 0;
  r = eh_limit_amplitudes(b, EH_DBL_MAX, EH_DBL_MAX, b->n_alive, 0.0, *Pc, 0, 0);
  if (r < 0)
    return r;
  /* the cavities' phases in pass-0 order (src/simple_rfca.c:419-466): a group
   * phase already set; a cavity with a fiducial sets its group; one without
   * waits for the beam, and so do the later cavities of its group.  Groups are
   * set only when the pass is tracked with particles, as in the reference. */
  const bool tracks = n_passes > 0 && b->n_alive > 0;
  {
    std::vector<int64_t> claimed;
    for (long ie = 0; ie < ne; ie++) {
      if (bl[ie].type != EH_T_RFCA)
        continue;
      EH_RFCA *rf = (EH_RFCA *)bl[ie].p_elem;
      const int64_t ref = rf->phase_reference;
      double ph;
      if (rf->volt == 0)
        continue;
      if (ref && std::find(claimed.begin(), claimed.end(), ref) != claimed.end()) {
        pend[ie] = 1;
      } else if (ref && eh_get_phase_reference(ref, &ph) == EH_REF_PHASE_RETURNED) {
        link[ie] = 1;
        lph[ie] = ph;
      } else if (rf->fiducial_seen) {
        if (ref && tracks) {
          r = eh_set_phase_reference(ref, rf->phase_fiducial);
          if (r < 0)
            return r;
        }
      } else {
        pend[ie] = 1;
        if (ref)
          claimed.push_back(ref);
      }
    }
  }
  for (long ie = 0; ie < ne; ie++) {
    if (pend[ie])
      continue; /* built once its phase is known */
    r = build(ie);
    if (r < 0)
      return r;
  }
  r = upload_fused_descs(b, hd);
  if (r < 0)
    return r;
  r = upload_terms(b, pool);
  if (r < 0)
    return r;
  /* kick-multipole records for settling loss positions (RF cavities built later
   * are not kick multipoles) */
  eh_kmp_tab tab = kmp_table(hd);
  bool any_kick = false;
  for (const eh_kmp &k : *tab)
    any_kick = any_kick || k.kick;
  if (!any_kick)
    tab = nullptr;
  /* the full kernel only for lines that radiate or hold MULT, ROTATE, MALIGN
   * or ENERGY (RF cavities still pending are built later; they are lean) */
  bool rad = false;
  for (long ie = 0; ie < ne; ie++)
    rad = rad || !eh_lean_capable(hd[ie]) || bl[ie].type >= EH_T_MULT;
  /* a pending cavity's phase from the beam that reaches it on pass 0; then
   * this cavity, the other positions of the same cavity, and the later
   * cavities of its group have their phase */
  auto resolve_pending = [&](long ie) -> long {
    EH_RFCA *rf = (EH_RFCA *)bl[ie].p_elem;
    eh_ctx rc = {};
    long r2 = rf_resolve(b, rf, (*Pe)[ie], ze[ie], rc);
    if (r2 < 0)
      return r2;
    for (long je = ie; je < ne; je++) {
      if (!pend[je])
        continue;
      const EH_RFCA *rj = (const EH_RFCA *)bl[je].p_elem;
      double ph;
      if (bl[je].p_elem == bl[ie].p_elem) {
        link[je] = (char)rc.rf_linked;
        lph[je] = rc.rf_ref_phase;
      } else if (rj->phase_reference &&
                 eh_get_phase_reference(rj->phase_reference, &ph) == EH_REF_PHASE_RETURNED) {
        link[je] = 1;
        lph[je] = ph;
      } else
        continue;
      pend[je] = 0;
      r2 = build(je);
      if (r2 < 0)
        return r2;
    }
    return upload_fused_descs(b, hd);
  };
  int p = 0, e = 0;
  if (dynamic) {
    /* the reference momentum P at the position (p, e); the momentum each
     * element sees at its next visit (Pdyn) follows from it through the
     * ENERGY targets, and every descriptor is rebuilt when that changes */
    double P = *Pc;
    auto advance = [&](long from, long to, double q) {
      for (long ie = from; ie < to; ie++)
        if (bl[ie].type == EH_T_ENERGY) {
          double Pn = q;
          if (energy_target((const EH_ENERGY *)bl[ie].p_elem, &Pn) == 1)
            q = Pn;
        }
      return q;
    };
    auto repropagate = [&](long from, double q) -> long {
      bool changed = false;
      for (long k = 0; k < ne; k++) {
        const long ie = (from + k) % ne;
        changed = changed || Pdyn[ie] != q;
        Pdyn[ie] = q;
        q = advance(ie, ie + 1, q);
      }
      if (!changed)
        return EH_OK;
      pool.clear();
      for (long ie = 0; ie < ne; ie++) {
        if (pend[ie])
          continue;
        long r2 = build(ie);
        if (r2 < 0)
          return r2;
      }
      long r2 = upload_fused_descs(b, hd);
      if (r2 < 0)
        return r2;
      return upload_terms(b, pool);
    };
    while (p < n_passes) {
      if (e == ne) {
        p++;
        e = 0;
        r = repropagate(0, P);
        if (r < 0)
          return r;
        continue;
      }
      if (b->n_alive == 0) {
        /* matching does nothing without particles; ENERGY targets still apply */
        P = advance(e, ne, P);
        if (p + 1 < n_passes)
          P = advance(0, ne, P);
        break;
      }
      if (Pdyn[e] != P)
        return fail(EH_ERR_STATE, "reference momentum bookkeeping out of step at element %d", e);
      if (p == 0 && pend[e]) {
        r = resolve_pending(e);
        if (r < 0)
          return r;
      }
      if (stop_before[e]) {
        r = rf_time_offset(b, (const EH_RFCA *)bl[e].p_elem, P, &toff[e]);
        if (r < 0)
          return r;
        const EH_RFCA *rf = (const EH_RFCA *)bl[e].p_elem;
        r = rf_mean_time(b, rf, P, rf->change_t ? toff[e] : 0.0, &tave[e]);
        if (r < 0)
          return r;
        r = build(e);
        if (r < 0)
          return r;
        r = upload_fused_descs(b, hd);
        if (r < 0)
          return r;
      }
      int e1 = e;
      while (e1 < ne) {
        if (e1 > e && ((p == 0 && pend[e1]) || stop_before[e1]))
          break;
        e1++;
        if (stop_after[e1 - 1])
          break;
      }
      r = launch_fused(b, ne, p, e, p, e1, effort, tab, rad);
      if (r < 0)
        return r;
      P = advance(e, e1, P);
      if (stop_after[e1 - 1]) {
        const long je = e1 - 1;
        r = match_energy(b, &P, match_mode[je]);
        if (r < 0)
          return r;
        /* then the element's MAXAMP sweep (src/do_tracking.c:2506-2517; its NaN
         * sweep ran in the launch and loses nothing here: the cavity removed
         * every invalid particle and a matching keeps finite values finite) */
        const EH_MAXAMP *m = p == 0 ? amp0[je] : amp1[je];
        b->pass = p;
        if (m && (m->x_max || m->y_max)) {
          r = m->elliptical ? eh_elimit_amplitudes(b, m->x_max, m->y_max, b->n_alive, ze[je], P, 0, 0, m->exponent,
                                                   m->y_exponent)
                            : eh_limit_amplitudes(b, m->x_max, m->y_max, b->n_alive, ze[je], P, 0, 0);
          if (r < 0)
            return r;
          if (m->elliptical) { /* the per-element sweep used descriptor slot 0 */
            r = upload_fused_descs(b, hd);
            if (r < 0)
              return r;
          }
        }
        r = repropagate(e1 % ne, P);
        if (r < 0)
          return r;
      }
      e = e1;
    }
    if (n_passes > 0)
      *Pc = P;
    b->pass = n_passes > 0 ? n_passes - 1 : 0;
    return b->n_alive;
  }
  if (n_passes > 0) {
    /* pass 0: stop before each RFCA whose phase waits for the beam */
    for (long ie = 0; ie < ne; ie++) {
      if (!pend[ie])
        continue;
      r = launch_fused(b, ne, p, e, 0, (int)ie, effort, tab, rad);
      if (r < 0)
        return r;
      e = (int)ie;
      if (b->n_alive == 0)
        break; /* no particle reaches it: the reference takes no fiducial */
      r = resolve_pending(ie);
      if (r < 0)
        return r;
    }
  }
  while (p < n_passes) {
    int p1 = p + tpl;
    if (p1 > n_passes)
      p1 = (int)n_passes;
    if (two_phase && p == 0)
      p1 = 1; /* pass 0 alone: later passes see other reference momenta */
    int pe = p1, ee = 0;
    if (e != 0) { /* finish the partial pass first */
      pe = p;
      ee = (int)ne;
      p1 = p + 1;
    }
    r = launch_fused(b, ne, p, e, pe, ee, effort, tab, rad);
    if (r < 0)
      return r;
    if (two_phase && p == 0 && p1 < n_passes) {
      /* the descriptors of the later passes */
      Pe = &Pe1;
      pool.clear();
      for (long ie = 0; ie < ne; ie++) {
        r = build(ie);
        if (r < 0)
          return r;
      }
      r = upload_fused_descs(b, hd);
      if (r < 0)
        return r;
      r = upload_terms(b, pool);
      if (r < 0)
        return r;
    }
    p = p1;
    e = 0;
  }
  if (n_passes > 0)
    *Pc = P_end;
  b->pass = n_passes > 0 ? n_passes - 1 : 0;
  return b->n_alive;
}

extern "C" long eh_do_tracking(eh_beam *b, EH_ELEMENT_LIST *beamline, long n_elements, long n_passes,
                               double *P_central, int mode, int turns_per_launch, eh_track_stats *stats) {
  if (!b || !beamline || !P_central || n_elements < 0 || n_passes < 0)
    return fail(EH_ERR_ARG, "eh_do_tracking: bad arguments");
  double effort = 0;
  double ms0 = b->kernel_ms, mms0 = b->main_ms, mu0 = b->main_units;
  long l0 = b->launches, ml0 = b->main_launches;
  long r;
  if (mode == EH_MODE_PLAIN)
    r = track_plain(b, beamline, n_elements, n_passes, P_central, &effort);
  else if (mode == EH_MODE_FUSED)
    r = track_fused(b, beamline, n_elements, n_passes, P_central, turns_per_launch, &effort);
  else
    return fail(EH_ERR_ARG, "unknown mode %d", mode);
  if (stats) {
    stats->particle_element_turns = effort;
    stats->kernel_ms = b->kernel_ms - ms0;
    stats->launches = b->launches - l0;
    stats->n_alive = b->n_alive;
    stats->main_ms = b->main_ms - mms0;
    stats->main_launches = b->main_launches - ml0;
    stats->main_units = b->main_units - mu0;
  }
  return r;
}

/* eh_do_tracking in fused mode with every device buffer beyond the beam (the
 * element descriptors, the CSBEND field terms, the launch counters) in the
 * static scratch of the current GPU: no allocation.  Lines larger than the
 * scratch are refused before anything runs.  Calls that overlap on one GPU
 * share the scratch and race. */
extern "C" long eh_do_tracking_noalloc(eh_beam *b, EH_ELEMENT_LIST *beamline, long n_elements, long n_passes,
                                       double *P_central, int turns_per_launch, eh_track_stats *stats) {
  if (!b || !beamline || !P_central || n_elements < 0 || n_passes < 0)
    return fail(EH_ERR_ARG, "eh_do_tracking_noalloc: bad arguments");
  if (g_device < 0 || g_device >= EH_MAX_DEVICES || !g_scratch[g_device])
    return fail(EH_ERR_STATE, "eh_do_tracking_noalloc: no scratch on this device (call eh_init first)");
  if (n_elements > EH_SCRATCH_NDESC)
    return fail(EH_ERR_ARG, "eh_do_tracking_noalloc: %ld elements, the static scratch holds %ld", n_elements,
                EH_SCRATCH_NDESC);
  long bends = 0;
  for (long e = 0; e < n_elements; e++)
    bends += beamline[e].type == EH_T_CSBEND;
  if (bends * EH_MAX_TERMS_PER_BEND > EH_SCRATCH_NTERMS)
    return fail(EH_ERR_ARG, "eh_do_tracking_noalloc: %ld CSBENDs may need more field terms than the static "
                            "scratch holds (%ld)", bends, EH_SCRATCH_NTERMS);
  char *base = (char *)g_scratch[g_device];
  unsigned long long *cnt0 = b->d_cnt;
  eh_desc *desc0 = b->d_desc;
  eh_fterm *terms0 = b->d_terms;
  long dcap0 = b->desc_cap, tcap0 = b->terms_cap;
  b->d_cnt = (unsigned long long *)base;
  char *fused0 = b->d_fused;
  long fcap0 = b->fused_cap;
  b->d_desc = (eh_desc *)(base + EH_SCRATCH_DESC_OFF);
  b->desc_cap = 1;
  b->d_fused = base + EH_SCRATCH_FUSED_OFF;
  b->fused_cap = EH_SCRATCH_FUSED_BYTES;
  b->d_terms = (eh_fterm *)(base + EH_SCRATCH_TERMS_OFF);
  b->terms_cap = EH_SCRATCH_NTERMS;
  b->noalloc = true;
  long r = eh_do_tracking(b, beamline, n_elements, n_passes, P_central, EH_MODE_FUSED, turns_per_launch, stats);
  b->d_cnt = cnt0;
  b->d_desc = desc0;
  b->d_fused = fused0;
  b->fused_cap = fcap0;
  b->d_terms = terms0;
  b->desc_cap = dcap0;
  b->terms_cap = tcap0;
  b->noalloc = false;
  return r;
}

/* contiguous AoS variants: rows of EH_NPROP doubles, as czarray_2d lays them out */
extern "C" long eh_beam_upload_flat(eh_beam *b, const double *flat, long n_part) {
  if (!flat || n_part < 0)
    return fail(EH_ERR_ARG, "eh_beam_upload_flat: bad arguments");
  std::vector<double *> rows(n_part > 0 ? n_part : 1);
  for (long i = 0; i < n_part; i++)
    rows[i] = const_cast<double *>(flat) + i * EH_NPROP;
  return eh_beam_upload(b, rows.data(), n_part);
}

extern "C" long eh_beam_download_flat(eh_beam *b, double *flat, long n_part) {
  if (!flat || n_part < 0)
    return fail(EH_ERR_ARG, "eh_beam_download_flat: bad arguments");
  std::vector<double *> rows(n_part > 0 ? n_part : 1);
  for (long i = 0; i < n_part; i++)
    rows[i] = flat + i * EH_NPROP;
  return eh_beam_download(b, rows.data(), n_part);
}

/* ------------------------------------------------------------------------- */
/* diagnostics */

/* reference row order for n particles given the loss event and identity of each
 * lost particle (host routine; exposed for testing the replay) */
extern "C" long eh_replay_order(long n, const int32_t *lost_event, const int32_t *lost_id, long n_lost,
                                int32_t *slot_at_row) {
  if (n < 0 || n_lost < 0 || n_lost > n || (n_lost && (!lost_event || !lost_id)) || !slot_at_row)
    return fail(EH_ERR_ARG, "eh_replay_order: bad arguments");
  std::vector<std::pair<int32_t, int32_t>> lost(n_lost);
  for (long k = 0; k < n_lost; k++) {
    if (lost_id[k] < 0 || lost_id[k] >= n)
      return fail(EH_ERR_ARG, "eh_replay_order: identity out of range");
    lost[k] = std::make_pair(lost_event[k], lost_id[k]);
  }
  std::vector<int32_t> s;
  replay_order(n, lost, s);
  std::memcpy(slot_at_row, s.data(), sizeof(int32_t) * n);
  return EH_OK;
}

extern "C" long eh_replay_sequence(long n, const int32_t *lost_event, const int32_t *lost_id, long n_lost,
                                   int32_t *pred, int32_t *order) {
  if (n < 0 || n_lost < 0 || n_lost > n || (n_lost && (!lost_event || !lost_id)) || !pred || !order)
    return fail(EH_ERR_ARG, "eh_replay_sequence: bad arguments");
  std::vector<std::pair<int32_t, int32_t>> lost(n_lost);
  for (long k = 0; k < n_lost; k++) {
    if (lost_id[k] < 0 || lost_id[k] >= n)
      return fail(EH_ERR_ARG, "eh_replay_sequence: identity out of range");
    lost[k] = std::make_pair(lost_event[k], lost_id[k]);
  }
  std::vector<int32_t> s, pr, od;
  replay_order(n, lost, s, &pr, &od);
  if ((long)od.size() != n_lost)
    return fail(EH_ERR_STATE, "eh_replay_sequence: %zu of %ld lost particles sequenced", od.size(), n_lost);
  std::memcpy(pred, pr.data(), sizeof(int32_t) * n);
  std::memcpy(order, od.data(), sizeof(int32_t) * n_lost);
  return EH_OK;
}

__global__ void k_selftest(const double *x, const double *y, long n, double *out) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  double a = x[i], b = y[i];
  out[i] = eh_sin(a);
  out[n + i] = eh_cos(a);
  out[2 * n + i] = eh_asin(b);
  out[3 * n + i] = eh_log(eh_fabs(a));
  out[4 * n + i] = eh_sqrt(eh_fabs(a));
  out[5 * n + i] = a / b;
  out[6 * n + i] = eh_tan(a);
}

/* evaluates sin(x), cos(x), asin(y), log|x|, sqrt|x|, x/y, tan(x) on the GPU into
 * out[7*n] (for checking that device arithmetic matches the host bit for bit) */
extern "C" long eh_selftest_math(const double *x, const double *y, long n, double *out) {
  if (g_device < 0 && eh_init(0) != EH_OK)
    return EH_ERR_HIP;
  if (n <= 0)
    return 0;
  double *d = nullptr;
  HIPCHK(hipMalloc(&d, sizeof(double) * 9 * n));
  hipError_t e = hipMemcpy(d, x, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(d + n, y, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_selftest, dim3(grid_for(n)), dim3(EH_BLOCK), 0, 0, d, d + n, n, d + 2 * n);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpy(out, d + 2 * n, sizeof(double) * 7 * n, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess)
    return fail(EH_ERR_HIP, "eh_selftest_math: %s", hipGetErrorString(e));
  return n;
}

__global__ void __launch_bounds__(EH_BLOCK) k_copy(const double *__restrict__ a, double *__restrict__ b, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    b[i] = a[i];
}

/* copies n doubles device-to-device with a plain kernel, reps times: a launch of
 * known HBM traffic (8n bytes read, 8n written) for calibrating memory counters */
extern "C" long eh_diag_copy(long n, int reps) {
  if (g_device < 0 && eh_init(0) != EH_OK)
    return EH_ERR_HIP;
  if (n <= 0 || reps < 1)
    return fail(EH_ERR_ARG, "eh_diag_copy: n and reps must be positive");
  double *d = nullptr;
  HIPCHK(hipMalloc(&d, sizeof(double) * 2 * n));
  hipError_t e = hipMemset(d, 0, sizeof(double) * 2 * n);
  for (int k = 0; k < reps && e == hipSuccess; k++) {
    hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(EH_BLOCK), 0, 0, d, d + n, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipDeviceSynchronize();
  (void)hipFree(d);
  if (e != hipSuccess)
    return fail(EH_ERR_HIP, "eh_diag_copy: %s", hipGetErrorString(e));
  return n;
}

/* ------------------------------------------------------------------------- */
/* batched lattice variants: host side */
#include <string>
#include <unordered_map>

struct eh_batch {
  long nvar = 0;
  std::vector<long> n_v, id_off, region_off;
  long n_total = 0, cap = 0, n_lost = 0;
  eh_soa buf[2] = {}, grave = {};
  int cur = 0;
  void *mem = nullptr;
  long *d_region_off = nullptr;
  int32_t *d_alive_in = nullptr, *d_wave_start = nullptr, *d_queue = nullptr, *d_desc_idx = nullptr;
  unsigned long long *d_alive_out = nullptr, *d_cnt = nullptr;
  eh_desc *d_pool = nullptr;
  eh_fterm *d_terms = nullptr;
  std::vector<double> cols;
  std::vector<long> alive_v;
  int32_t event = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double kernel_ms = 0;
  long launches = 0;
  int n_cu = 256;
  /* descriptor tables of the last call, reused while the beamlines are unchanged */
  uint64_t tables_hash = 0;
  long tables_ne = 0;
  bool rad = false;    /* some descriptor of the pool radiates */
  eh_kmp_tab kmp_pool; /* kick-multipole records of the descriptor pool */
  eh_kmp_idx kmp_idx;  /* variant * ne + element -> pool entry */
  std::vector<eh_evrec> evlog;
};

static size_t elem_struct_size(int32_t type) {
  switch (type) {
  case EH_T_DRIF:
    return sizeof(EH_DRIFT);
  case EH_T_EDRIFT:
    return sizeof(EH_EDRIFT);
  case EH_T_KQUAD:
    return sizeof(EH_KQUAD);
  case EH_T_KSEXT:
    return sizeof(EH_KSEXT);
  case EH_T_KOCT:
    return sizeof(EH_KOCT);
  case EH_T_KQUSE:
    return sizeof(EH_KQUSE);
  case EH_T_CSBEND:
    return sizeof(EH_CSBEND);
  case EH_T_RFCA:
    return sizeof(EH_RFCA);
  case EH_T_MAXAMP:
    return sizeof(EH_MAXAMP);
  case EH_T_MULT:
    return sizeof(EH_MULT);
  case EH_T_ROTATE:
    return sizeof(EH_ROTATE);
  case EH_T_MALIGN:
    return sizeof(EH_MALIGN);
  case EH_T_ENERGY:
    return sizeof(EH_ENERGY);
  default:
    return 0;
  }
}

/* FNV-1a over every element struct of every variant, and P_central */
static uint64_t beamlines_hash(EH_ELEMENT_LIST *const *bl, long nvar, long ne, double Pc) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](const void *p, size_t n) {
    const unsigned char *c = (const unsigned char *)p;
    for (size_t i = 0; i < n; i++) {
      h ^= c[i];
      h *= 1099511628211ull;
    }
  };
  mix(&Pc, sizeof Pc);
  for (long v = 0; v < nvar; v++)
    for (long e = 0; e < ne; e++) {
      mix(&bl[v][e].type, sizeof(int32_t));
      if (bl[v][e].p_elem)
        mix(bl[v][e].p_elem, elem_struct_size(bl[v][e].type));
    }
  return h;
}

extern "C" eh_batch *eh_batch_create(long nvar, const long *n_particles) {
  if (nvar < 1 || !n_particles) {
    fail(EH_ERR_ARG, "eh_batch_create: bad arguments");
    return nullptr;
  }
  if (g_device < 0 && eh_init(0) != EH_OK)
    return nullptr;
  eh_batch *b = new eh_batch();
  b->nvar = nvar;
  b->n_v.assign(n_particles, n_particles + nvar);
  long idn = 0, off = 0;
  for (long v = 0; v < nvar; v++) {
    if (n_particles[v] < 0) {
      fail(EH_ERR_ARG, "eh_batch_create: negative particle count");
      delete b;
      return nullptr;
    }
    b->id_off.push_back(idn);
    b->region_off.push_back(off);
    idn += n_particles[v];
    off += (n_particles[v] + 63) / 64 * 64;
  }
  b->n_total = idn;
  b->cap = off > 0 ? off : 64;
  if (b->n_total > 2147483647L) {
    fail(EH_ERR_ARG, "eh_batch_create: more than 2^31 particles");
    delete b;
    return nullptr;
  }
  size_t per = (soa_bytes(b->cap) + 255) & ~(size_t)255;
  size_t small = sizeof(long) * nvar + sizeof(int32_t) * (3 * nvar + 8) + sizeof(unsigned long long) * (nvar + 8);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, g_device) == hipSuccess)
    b->n_cu = prop.multiProcessorCount;
  if (hipMalloc(&b->mem, 3 * per) != hipSuccess || hipMalloc(&b->d_region_off, sizeof(long) * nvar) != hipSuccess ||
      hipMalloc(&b->d_alive_in, sizeof(int32_t) * nvar) != hipSuccess ||
      hipMalloc(&b->d_wave_start, sizeof(int32_t) * (nvar + 1)) != hipSuccess ||
      hipMalloc(&b->d_queue, sizeof(int32_t) * 4) != hipSuccess ||
      hipMalloc(&b->d_alive_out, sizeof(unsigned long long) * nvar) != hipSuccess ||
      hipMalloc(&b->d_cnt, sizeof(unsigned long long) * 8) != hipSuccess ||
      hipStreamCreate(&b->stream) != hipSuccess || hipEventCreate(&b->ev0) != hipSuccess ||
      hipEventCreate(&b->ev1) != hipSuccess) {
    (void)small;
    fail(EH_ERR_NOMEM, "eh_batch_create: device allocation failed");
    delete b;
    return nullptr;
  }
  soa_carve(&b->buf[0], (char *)b->mem, b->cap);
  soa_carve(&b->buf[1], (char *)b->mem + per, b->cap);
  soa_carve(&b->grave, (char *)b->mem + 2 * per, b->cap);
  (void)hipMemcpy(b->d_region_off, b->region_off.data(), sizeof(long) * nvar, hipMemcpyHostToDevice);
  b->cols.assign(4 * (size_t)b->n_total, 0.0);
  b->alive_v.assign(nvar, 0);
  return b;
}

extern "C" void eh_batch_free(eh_batch *b) {
  if (!b)
    return;
  if (b->stream)
    (void)hipStreamSynchronize(b->stream);
  void *ptrs[] = {b->mem, b->d_region_off, b->d_alive_in, b->d_wave_start, b->d_queue, b->d_alive_out, b->d_cnt,
                  b->d_pool, b->d_terms, b->d_desc_idx};
  for (void *q : ptrs)
    if (q)
      (void)hipFree(q);
  if (b->ev0)
    (void)hipEventDestroy(b->ev0);
  if (b->ev1)
    (void)hipEventDestroy(b->ev1);
  if (b->stream)
    (void)hipStreamDestroy(b->stream);
  delete b;
}

extern "C" long eh_batch_upload(eh_batch *b, long v, const double *flat, long n) {
  if (!b || v < 0 || v >= b->nvar || !flat || n != b->n_v[v])
    return fail(EH_ERR_ARG, "eh_batch_upload: bad arguments");
  if (b->cur != 0 || b->n_lost)
    return fail(EH_ERR_STATE, "eh_batch_upload: upload every variant before tracking");
  std::vector<double> h(6 * (size_t)(n > 0 ? n : 1));
  std::vector<int32_t> r(n > 0 ? n : 1), m1(n > 0 ? n : 1, -1), z(n > 0 ? n : 1, 0);
  for (long i = 0; i < n; i++) {
    for (int k = 0; k < 6; k++)
      h[k * (size_t)n + i] = flat[i * EH_NPROP + k];
    for (int k = 0; k < 4; k++)
      b->cols[4 * (size_t)(b->id_off[v] + i) + k] = flat[i * EH_NPROP + 6 + k];
    r[i] = (int32_t)(b->id_off[v] + i);
  }
  if (n) {
    eh_soa &a = b->buf[0];
    long o = b->region_off[v];
    double *dst[6] = {a.x, a.xp, a.y, a.yp, a.s, a.dp};
    for (int k = 0; k < 6; k++)
      HIPCHK(hipMemcpy(dst[k] + o, h.data() + k * (size_t)n, sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(a.row0 + o, r.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(a.levent + o, m1.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(a.lpass + o, z.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  }
  b->alive_v[v] = n;
  return n;
}

extern "C" long eh_batch_n_alive(const eh_batch *b, long v) {
  if (!b || v < 0 || v >= b->nvar)
    return EH_ERR_ARG;
  return b->alive_v[v];
}

/* reference row order of variant v (local identities 0..n_v-1); idx[local id] =
 * index in the alive region of v or -1; the graveyard slots of v go to gidx */
static long batch_order(eh_batch *b, long v, std::vector<int32_t> &slot_at_row, std::vector<int32_t> &idx,
                        std::vector<int32_t> &gidx, std::vector<int32_t> *pred = nullptr,
                        std::vector<int32_t> *order = nullptr) {
  long nv = b->n_v[v], na = b->alive_v[v], off = b->id_off[v];
  std::vector<int32_t> arow(na > 0 ? na : 1), grow0(b->n_lost > 0 ? b->n_lost : 1),
      gev(b->n_lost > 0 ? b->n_lost : 1);
  HIPCHK(hipStreamSynchronize(b->stream));
  if (na)
    HIPCHK(hipMemcpy(arow.data(), b->buf[b->cur].row0 + b->region_off[v], sizeof(int32_t) * na,
                     hipMemcpyDeviceToHost));
  if (b->n_lost) {
    HIPCHK(hipMemcpy(grow0.data(), b->grave.row0, sizeof(int32_t) * b->n_lost, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(gev.data(), b->grave.levent, sizeof(int32_t) * b->n_lost, hipMemcpyDeviceToHost));
  }
  idx.assign(nv, -1);
  gidx.assign(nv, -1);
  for (long k = 0; k < na; k++) {
    long id = arow[k] - off;
    if (id < 0 || id >= nv)
      return fail(EH_ERR_STATE, "corrupt particle identity in variant %ld", v);
    idx[id] = (int32_t)k;
  }
  std::vector<std::pair<int32_t, int32_t>> lost;
  for (long k = 0; k < b->n_lost; k++) {
    long id = grow0[k] - off;
    if (id >= 0 && id < nv) {
      lost.push_back(std::make_pair(gev[k], (int32_t)id));
      gidx[id] = (int32_t)k;
    }
  }
  if ((long)lost.size() + na != nv)
    return fail(EH_ERR_STATE, "variant %ld: %ld alive + %zu lost != %ld", v, na, lost.size(), nv);
  replay_order(nv, lost, slot_at_row, pred, order);
  return EH_OK;
}

extern "C" long eh_batch_download(eh_batch *b, long v, double *flat, long n) {
  if (!b || v < 0 || v >= b->nvar || !flat || n != b->n_v[v])
    return fail(EH_ERR_ARG, "eh_batch_download: bad arguments");
  std::vector<int32_t> slot, idx, gidx, pred, order;
  long r = batch_order(b, v, slot, idx, gidx, &pred, &order);
  if (r < 0)
    return r;
  long na = b->alive_v[v], nl = b->n_lost;
  std::vector<double> A(6 * (size_t)(na > 0 ? na : 1)), G(6 * (size_t)(nl > 0 ? nl : 1));
  std::vector<int32_t> gpass(nl > 0 ? nl : 1);
  const eh_soa &a = b->buf[b->cur], &g = b->grave;
  const double *as[6] = {a.x, a.xp, a.y, a.yp, a.s, a.dp}, *gs[6] = {g.x, g.xp, g.y, g.yp, g.s, g.dp};
  for (int k = 0; k < 6; k++) {
    if (na)
      HIPCHK(hipMemcpy(A.data() + k * (size_t)na, as[k] + b->region_off[v], sizeof(double) * na,
                       hipMemcpyDeviceToHost));
    if (nl)
      HIPCHK(hipMemcpy(G.data() + k * (size_t)nl, gs[k], sizeof(double) * nl, hipMemcpyDeviceToHost));
  }
  std::vector<int32_t> gev(nl > 0 ? nl : 1);
  if (nl) {
    HIPCHK(hipMemcpy(gpass.data(), g.lpass, sizeof(int32_t) * nl, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(gev.data(), g.levent, sizeof(int32_t) * nl, hipMemcpyDeviceToHost));
  }
  settle_losses(b->evlog, v, order, pred, gidx, gev, G, nl, gpass);
  for (long row = 0; row < n; row++) {
    int32_t id = slot[row];
    double *c = flat + row * EH_NPROP;
    for (int k = 0; k < 4; k++)
      c[6 + k] = b->cols[4 * (size_t)(b->id_off[v] + id) + k];
    if (idx[id] >= 0) {
      for (int q = 0; q < 6; q++)
        c[q] = A[q * (size_t)na + idx[id]];
    } else {
      long k = gidx[id];
      for (int q = 0; q < 6; q++)
        c[q] = G[q * (size_t)nl + k];
      c[EH_LOSS_PASS_INDEX] = gpass[k];
    }
  }
  return na;
}

/* Kahan TMEAN fiducial of variant v in its reference row order */
static long batch_fiducial(eh_batch *b, long v, double sOffset, double p0, double *tFid) {
  std::vector<int32_t> slot, idx, gidx;
  long r = batch_order(b, v, slot, idx, gidx);
  if (r < 0)
    return r;
  long na = b->alive_v[v];
  std::vector<double> s(na > 0 ? na : 1), dp(na > 0 ? na : 1);
  if (na) {
    HIPCHK(hipMemcpy(s.data(), b->buf[b->cur].s + b->region_off[v], sizeof(double) * na, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(dp.data(), b->buf[b->cur].dp + b->region_off[v], sizeof(double) * na, hipMemcpyDeviceToHost));
  }
  double tsum = 0, error = 0.0;
  long nsum = 0;
  for (long row = 0; row < na; row++) {
    long k = idx[slot[row]];
    if (k < 0)
      return fail(EH_ERR_STATE, "replay placed a lost particle among survivors");
    double p = p0;
    p *= 1 + dp[k];
    double beta = p / std::sqrt(p * p + 1);
    double term = (s[k] + sOffset) / (EH_C_MKS * beta);
    double Y = term + error;
    double sum = tsum + Y;
    error = Y - (sum - tsum);
    tsum = sum;
    nsum++;
  }
  *tFid = nsum > 0 ? tsum / nsum : 0.0;
  return EH_OK;
}

/* one span (p0,e0)->(p1,e1): prep + persistent batch kernel, on the batch stream */
static long batch_enqueue_span(eh_batch *b, long ne, int p0, int e0, int p1, int e1, int first, int block,
                               int waves_per_cu) {
  long span = ((long)(p1 - p0) * ne + ne) * 3 + 3;
  if ((long)b->event + span >= 2147483647L)
    return fail(EH_ERR_STATE, "loss event counter overflow");
  if (((long)(p1 - p0) * ne + ne) * 8 >= 2147483647L)
    return fail(EH_ERR_ARG, "span of %ld passes x %ld elements too long for one launch", (long)(p1 - p0 + 1), ne);
  int32_t base = b->event;
  b->event += (int32_t)span;
  if (b->kmp_pool) {
    eh_evrec rec = {base, (int32_t)span, p0, (int32_t)ne, b->kmp_pool, b->kmp_idx};
    b->evlog.push_back(rec);
  }
  hipLaunchKernelGGL(k_batch_prep, dim3(1), dim3(64), 0, b->stream, b->d_alive_out, b->d_alive_in, b->d_wave_start,
                     b->d_alive_out, b->d_queue, (int)b->nvar, first);
  eh_batch_dev bd;
  bd.region_off = b->d_region_off;
  bd.alive_in = b->d_alive_in;
  bd.alive_out = b->d_alive_out;
  bd.wave_start = b->d_wave_start;
  bd.queue = b->d_queue;
  bd.desc_idx = b->d_desc_idx;
  bd.nvar = (int)b->nvar;
  long threads = (long)b->n_cu * waves_per_cu * 64;
  unsigned grid = (unsigned)((threads + block - 1) / block);
  if (b->rad)
    hipLaunchKernelGGL(k_batch<true>, dim3(grid), dim3(block), 0, b->stream, b->buf[b->cur], b->buf[1 - b->cur],
                       b->grave, bd, b->d_pool, b->d_terms, (int)ne, p0, e0, p1, e1, base, b->d_cnt);
  else
    hipLaunchKernelGGL(k_batch<false>, dim3(grid), dim3(block), 0, b->stream, b->buf[b->cur], b->buf[1 - b->cur],
                       b->grave, bd, b->d_pool, b->d_terms, (int)ne, p0, e0, p1, e1, base, b->d_cnt);
  b->cur = 1 - b->cur;
  b->launches++;
  return EH_OK;
}

static long batch_sync_counts(eh_batch *b) {
  std::vector<unsigned long long> out(b->nvar);
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(out.data(), b->d_alive_out, sizeof(unsigned long long) * b->nvar, hipMemcpyDeviceToHost));
  unsigned long long cnt[4];
  HIPCHK(hipMemcpy(cnt, b->d_cnt, sizeof cnt, hipMemcpyDeviceToHost));
  long alive = 0;
  for (long v = 0; v < b->nvar; v++) {
    b->alive_v[v] = (long)out[v];
    alive += b->alive_v[v];
  }
  b->n_lost = (long)cnt[1];
  if (alive + b->n_lost != b->n_total)
    return fail(EH_ERR_STATE, "batch count mismatch: %ld alive + %ld lost != %ld", alive, b->n_lost, b->n_total);
  return alive;
}

/* the context of every element of one beamline: its s range from the line start,
 * the MAXAMP in force on pass 0 and on later passes (a MAXAMP stays in force
 * across the end of a pass), and the limits of the sweep that follows it */
static void line_contexts(const EH_ELEMENT_LIST *bl, long ne, double Pc, std::vector<eh_ctx> &ctx) {
  ctx.assign(ne, eh_ctx{});
  const EH_MAXAMP *last_amp = nullptr;
  double z = 0;
  for (long e = 0; e < ne; e++) {
    ctx[e].Po = Pc;
    ctx[e].z_start = z;
    z += elem_length(bl + e);
    ctx[e].z_end = z;
    if (bl[e].type == EH_T_MAXAMP)
      last_amp = (const EH_MAXAMP *)bl[e].p_elem;
  }
  const EH_MAXAMP *cur0 = nullptr, *cur1 = last_amp;
  for (long e = 0; e < ne; e++) {
    if (bl[e].type == EH_T_MAXAMP)
      cur0 = cur1 = (const EH_MAXAMP *)bl[e].p_elem;
    eh_ctx &c = ctx[e];
    c.ap[0] = cur0;
    c.ap[1] = cur1;
    for (int pi = 0; pi < 2; pi++) {
      const EH_MAXAMP *m = pi ? cur1 : cur0;
      c.post_x[pi] = m ? m->x_max : 0;
      c.post_y[pi] = m ? m->y_max : 0;
      c.post_on[pi] = (bl[e].type != EH_T_MARK) && m && (m->x_max || m->y_max);
    }
  }
}

/* append one descriptor (and its field terms) to the pool; returns its index */
static int32_t pool_push(eh_desc d, const std::vector<eh_fterm> &ft, std::vector<eh_desc> &pool,
                         std::vector<eh_fterm> &terms) {
  if (!ft.empty()) {
    /* this element's term offsets are relative to its own list */
    int32_t shift = (int32_t)terms.size();
    d.csb_fx0 += shift;
    d.csb_fy0 += shift;
    if (d.csb_dense) {
      d.csb_dx0 += shift;
      d.csb_dy0 += shift;
    }
    terms.insert(terms.end(), ft.begin(), ft.end());
  }
  pool.push_back(d);
  return (int32_t)pool.size() - 1;
}

/* the descriptor pool of all variants, de-duplicated by content (descriptor
 * bytes and field terms), with idx[v*ne + e] the descriptor of element e of
 * variant v.  RF cavities whose fiducial is still pending get a MARK placeholder
 * (never shared); batch_rf_update replaces it once the fiducial is known.  Host
 * code only (no HIP calls), so that tests can run it on the CPU. */
static long batch_tables(EH_ELEMENT_LIST *const *beamlines, long nvar, long ne, double Pc,
                         std::vector<eh_desc> &pool, std::vector<eh_fterm> &terms, std::vector<int32_t> &idx) {
  std::unordered_map<std::string, int32_t> seen;
  pool.clear();
  terms.clear();
  idx.assign((size_t)nvar * ne, 0);
  std::vector<eh_ctx> ctx;
  for (long v = 0; v < nvar; v++) {
    const EH_ELEMENT_LIST *bl = beamlines[v];
    line_contexts(bl, ne, Pc, ctx);
    for (long e = 0; e < ne; e++) {
      eh_desc d;
      std::vector<eh_fterm> ft;
      long r;
      if (rf_pending(bl + e)) {
        std::memset(&d, 0, sizeof d);
        d.type = EH_T_MARK;
        d.adj_pk = d.adj_pdx = d.adj_pdy = d.adj_pfse = d.adj_anchor = -1;
        r = EH_OK;
      } else
        r = build_desc(&d, bl + e, ctx[e], ft);
      if (r < 0)
        return r;
      std::string key((const char *)&d, sizeof d);
      key.append((const char *)ft.data(), sizeof(eh_fterm) * ft.size());
      auto it = seen.find(key);
      if (it != seen.end() && bl[e].type != EH_T_RFCA) {
        idx[(size_t)v * ne + e] = it->second;
      } else {
        int32_t id = pool_push(d, ft, pool, terms);
        seen.emplace(key, id);
        idx[(size_t)v * ne + e] = id;
      }
    }
  }
  return EH_OK;
}

/* once the RF cavity at position er has its fiducial in every variant: a fresh
 * descriptor for every position of each variant's line that holds the same
 * cavity (the same element struct, as track_fused does).  Host code only. */
static long batch_rf_update(EH_ELEMENT_LIST *const *beamlines, long nvar, long ne, long er, double Pc,
                            std::vector<eh_desc> &pool, std::vector<eh_fterm> &terms, std::vector<int32_t> &idx) {
  std::vector<eh_ctx> ctx;
  for (long v = 0; v < nvar; v++) {
    const EH_ELEMENT_LIST *bl = beamlines[v];
    line_contexts(bl, ne, Pc, ctx);
    for (long je = 0; je < ne; je++) {
      if (bl[je].p_elem != bl[er].p_elem || bl[je].type != EH_T_RFCA)
        continue;
      eh_desc d;
      std::vector<eh_fterm> ft;
      long r = build_desc(&d, bl + je, ctx[je], ft);
      if (r < 0)
        return r;
      idx[(size_t)v * ne + je] = pool_push(d, ft, pool, terms);
    }
  }
  return EH_OK;
}

/* the device tables of the last call can serve this one only if the beamlines
 * are byte-for-byte unchanged and no RF cavity still waits for its fiducial
 * (whose placeholder the pass-0 stop has to replace in the host copy) */
static bool batch_tables_reusable(bool have_tables, long tables_ne, uint64_t tables_hash,
                                  EH_ELEMENT_LIST *const *beamlines, long nvar, long ne, uint64_t hash) {
  if (!have_tables || tables_ne != ne || tables_hash != hash)
    return false;
  for (long v = 0; v < nvar; v++)
    for (long e = 0; e < ne; e++)
      if (rf_pending(beamlines[v] + e))
        return false;
  return true;
}

extern "C" long eh_batch_track(eh_batch *b, EH_ELEMENT_LIST *const *beamlines, long ne, long n_passes,
                               double *P_central, const eh_batch_tune *tune, eh_track_stats *stats) {
  if (!b || !beamlines || ne < 1 || n_passes < 0 || !P_central)
    return fail(EH_ERR_ARG, "eh_batch_track: bad arguments");
  eh_batch_tune t = {256, (int32_t)ne, 8, 1};
  if (tune)
    t = *tune;
  if (t.block != 64 && t.block != 128 && t.block != 256)
    return fail(EH_ERR_ARG, "eh_batch_track: block must be 64, 128 or 256");
  if (t.span_elements < 1)
    t.span_elements = (int32_t)ne;
  if (t.waves_per_cu < 1)
    t.waves_per_cu = 8;
  for (long v = 1; v < b->nvar; v++)
    for (long e = 0; e < ne; e++)
      if (beamlines[v][e].type != beamlines[0][e].type)
        return fail(EH_ERR_ARG, "variant %ld element %ld: types differ from variant 0", v, e);
  for (long e = 0; e < ne; e++)
    if (beamlines[0][e].type == EH_T_ENERGY)
      return fail(EH_ERR_UNSUPPORTED, "element %ld: ENERGY is not supported in batched tracking", e);
  /* every variant is a run of its own: no phase-reference groups across them,
   * and the batch takes TMEAN and LIGHT fiducials */
  for (long v = 0; v < b->nvar; v++)
    for (long e = 0; e < ne; e++)
      if (beamlines[v][e].type == EH_T_RFCA) {
        const EH_RFCA *rf = (const EH_RFCA *)beamlines[v][e].p_elem;
        if (rf->phase_reference)
          return fail(EH_ERR_UNSUPPORTED, "element %ld: PHASE_REFERENCE is not supported in batched tracking", e);
        if (rf->change_t || rf->change_p0 || rf->linearize || rf->lockPhase)
          return fail(EH_ERR_UNSUPPORTED,
                      "element %ld: CHANGE_T, CHANGE_P0, LINEARIZE and LOCK_PHASE are not supported in batched "
                      "tracking", e);
        if (rf->fiducial_mode != EH_FID_TMEAN && rf->fiducial_mode != EH_FID_LIGHT && !rf->fiducial_seen)
          return fail(EH_ERR_UNSUPPORTED, "element %ld: batched tracking takes TMEAN and LIGHT fiducials only", e);
      }
  double Pc = *P_central;
  /* descriptors, deduplicated across variants */
  std::vector<eh_desc> pool;
  std::vector<eh_fterm> terms;
  std::vector<int32_t> idx((size_t)b->nvar * ne);
  std::vector<long> rf_elems;
  for (long e = 0; e < ne; e++)
    if (beamlines[0][e].type == EH_T_RFCA)
      rf_elems.push_back(e);
  const uint64_t hash = beamlines_hash(beamlines, b->nvar, ne, Pc);
  const bool reuse =
      batch_tables_reusable(b->d_pool != nullptr, b->tables_ne, b->tables_hash, beamlines, b->nvar, ne, hash);
  if (!reuse) {
    long r = batch_tables(beamlines, b->nvar, ne, Pc, pool, terms, idx);
    if (r < 0)
      return r;
    eh_kmp_tab tab = kmp_table(pool);
    bool any_kick = false;
    for (const eh_kmp &k : *tab)
      any_kick = any_kick || k.kick;
    b->kmp_pool = any_kick ? tab : nullptr;
    b->rad = false;
    for (const eh_desc &d : pool)
      b->rad = b->rad || !eh_lean_capable(d);
    b->kmp_idx = any_kick ? std::make_shared<const std::vector<int32_t>>(idx) : nullptr;
  }
  auto upload_tables = [&]() -> long {
    b->tables_hash = 0; /* valid again only once this call completes */
    if (b->d_pool)
      HIPCHK(hipFree(b->d_pool));
    if (b->d_terms)
      HIPCHK(hipFree(b->d_terms));
    if (b->d_desc_idx)
      HIPCHK(hipFree(b->d_desc_idx));
    b->d_pool = nullptr;
    b->d_terms = nullptr;
    b->d_desc_idx = nullptr;
    HIPCHK(hipMalloc(&b->d_pool, sizeof(eh_desc) * pool.size()));
    HIPCHK(hipMemcpy(b->d_pool, pool.data(), sizeof(eh_desc) * pool.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&b->d_terms, sizeof(eh_fterm) * (terms.size() + 1)));
    if (!terms.empty())
      HIPCHK(hipMemcpy(b->d_terms, terms.data(), sizeof(eh_fterm) * terms.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&b->d_desc_idx, sizeof(int32_t) * idx.size()));
    HIPCHK(hipMemcpy(b->d_desc_idx, idx.data(), sizeof(int32_t) * idx.size(), hipMemcpyHostToDevice));
    return EH_OK;
  };
  long r = reuse ? EH_OK : upload_tables();
  if (r < 0)
    return r;
  std::vector<int32_t> alive32(b->nvar);
  for (long v = 0; v < b->nvar; v++)
    alive32[v] = (int32_t)b->alive_v[v];
  HIPCHK(hipMemcpy(b->d_alive_in, alive32.data(), sizeof(int32_t) * b->nvar, hipMemcpyHostToDevice));
  unsigned long long cnt0[4] = {0, (unsigned long long)b->n_lost, 0, 0};
  HIPCHK(hipMemcpy(b->d_cnt, cnt0, sizeof cnt0, hipMemcpyHostToDevice));
  double ms0 = b->kernel_ms;
  long l0 = b->launches;
  int first = 1;
  /* initial NaN sweep of do_tracking is folded into nothing here: the batch is
   * uploaded from finite generated beams; a NaN row is still removed by the first
   * element's NaN sweep.  Pass 0 stops before each RFCA whose fiducial is unknown. */
  int p = 0, e = 0;
  double span_ms = 0; /* stream time of the tracking launches, host work excluded */
  auto elapsed = [&]() -> long {
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    HIPCHK(hipEventSynchronize(b->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    span_ms += ms;
    return EH_OK;
  };
  if (n_passes > 0) {
    for (long er : rf_elems) {
      bool pending = false;
      for (long v = 0; v < b->nvar; v++)
        pending = pending || rf_pending(beamlines[v] + er);
      if (!pending)
        continue;
      if (reuse)
        return fail(EH_ERR_STATE, "eh_batch_track: reused tables with a pending RF fiducial");
      if (e < er) {
        HIPCHK(hipEventRecord(b->ev0, b->stream));
        r = batch_enqueue_span(b, ne, p, e, 0, (int)er, first, t.block, t.waves_per_cu);
        if (r < 0)
          return r;
        first = 0;
        r = elapsed();
        if (r < 0)
          return r;
        r = batch_sync_counts(b);
        if (r < 0)
          return r;
      }
      for (long v = 0; v < b->nvar; v++) {
        if (!rf_pending(beamlines[v] + er))
          continue;
        double zEnd = 0;
        for (long k = 0; k <= er; k++)
          zEnd += elem_length(beamlines[v] + k);
        r = rf_take_fiducial((EH_RFCA *)beamlines[v][er].p_elem, zEnd, [&](int, double sOffset, double *t0) {
          return batch_fiducial(b, v, sOffset, Pc, t0);
        });
        if (r < 0)
          return r;
      }
      r = batch_rf_update(beamlines, b->nvar, ne, er, Pc, pool, terms, idx);
      if (r < 0)
        return r;
      r = upload_tables();
      if (r < 0)
        return r;
      e = (int)er;
    }
  }
  /* the rest as spans of t.span_elements elements, optionally one graph */
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool capture = t.use_graph && n_passes > 0;
  if (capture)
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  else
    HIPCHK(hipEventRecord(b->ev0, b->stream));
  while (p < n_passes) {
    long pos = (long)p * ne + e + t.span_elements;
    long endpos = (long)n_passes * ne;
    if (pos > endpos)
      pos = endpos;
    int p1 = (int)(pos / ne), e1 = (int)(pos % ne);
    r = batch_enqueue_span(b, ne, p, e, p1, e1, first, t.block, t.waves_per_cu);
    if (r < 0) {
      if (capture) {
        (void)hipStreamEndCapture(b->stream, &graph);
        if (graph)
          (void)hipGraphDestroy(graph);
      }
      return r;
    }
    first = 0;
    p = p1;
    e = e1;
  }
  if (capture) {
    HIPCHK(hipStreamEndCapture(b->stream, &graph));
    HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    HIPCHK(hipGraphLaunch(exec, b->stream));
  }
  r = elapsed();
  if (r < 0)
    return r;
  b->kernel_ms += span_ms;
  if (exec)
    HIPCHK(hipGraphExecDestroy(exec));
  if (graph)
    HIPCHK(hipGraphDestroy(graph));
  if (first) {
    /* nothing was tracked */
  } else {
    r = batch_sync_counts(b);
    if (r < 0)
      return r;
  }
  unsigned long long cnt[4];
  HIPCHK(hipMemcpy(cnt, b->d_cnt, sizeof cnt, hipMemcpyDeviceToHost));
  long alive = 0;
  for (long v = 0; v < b->nvar; v++)
    alive += b->alive_v[v];
  /* the tables now on the device describe the beamlines as they stand (RF fiducials set) */
  b->tables_hash = beamlines_hash(beamlines, b->nvar, ne, Pc);
  b->tables_ne = ne;
  if (stats) {
    stats->particle_element_turns = (double)cnt[3];
    stats->kernel_ms = b->kernel_ms - ms0;
    stats->launches = b->launches - l0;
    stats->n_alive = alive;
    stats->main_ms = stats->kernel_ms; /* spans run back to back in one graph */
    stats->main_launches = stats->launches;
    stats->main_units = stats->particle_element_turns;
  }
  return alive;
}

/* ------------------------------------------------------------------------- */
/* reverse-mode adjoint of multi-turn tracking                                 */
/*                                                                           */
/* Forward: particles stay in fixed rows (no compaction); every C turns the   */
/* turn-start state is kept (checkpoint).  Backward, segment by segment from  */
/* the last: the survivors are re-tracked from the segment's checkpoint while */
/* the entry states of anchor elements (turn start, RF cavities) are stored,  */
/* then one kernel walks the segment's turns and elements in reverse,         */
/* recovering each element's entry state from its exit state with the        */
/* inverse maps (re-anchored at every anchor) and pulling the adjoint back.   */
/* With C = 1 the forward pass stores the anchors itself and nothing is       */
/* re-tracked.  Parameter gradients are summed per wavefront in LDS, then per */
/* block, then over blocks in a fixed order: the result is deterministic.     */
#include "eh_adjoint.h"

#define EH_ADJ_BLOCK 64

/* forward over turns [t0, t1) for particles i < n in fixed rows; anchors (if
 * given) receive the entry state of element 0 (slot 0) and of each anchor
 * element (slot d.adj_anchor) for every turn, at index
 * ((t - t0) * n_anchor + slot) * 6 * a_ld + c * a_ld + i */
__global__ void __launch_bounds__(EH_BLOCK) EH_TRACK_WAVES k_adj_fwd(double *z, long ld, int32_t *alive, long n,
                                                       const eh_desc *__restrict__ descs,
                                                       const eh_fterm *__restrict__ terms, int ne, int t0, int t1,
                                                       double *anchors, int n_anchor, long a_ld,
                                                       unsigned long long *units, const int32_t *cvar) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  bool in = i < n;
  /* lattice variants: rows come in 64-row chunks of one variant each (cvar),
   * whose descriptors are the variant's block of ne; the chunk is the wavefront */
  if (cvar) {
    const long chunk = i >> 6;
    const int v = __builtin_amdgcn_readfirstlane(chunk < (n + 63) / 64 ? cvar[chunk] : 0);
    descs += (long)v * ne;
  }
  bool live = in && alive[i];
  eh_p6 p = {0, 0, 0, 0, 0, 0};
  if (live)
    p = {z[i], z[ld + i], z[2 * ld + i], z[3 * ld + i], z[4 * ld + i], z[5 * ld + i]};
  unsigned long long u = 0;
  for (int t = t0; t < t1; t++) {
    if (__ballot(live) == 0)
      break;
    for (int e = 0; e < ne; e++) {
      const eh_desc *d = descs + e;
      if (live) {
        int slot = e == 0 ? 0 : d->adj_anchor;
        if (anchors && slot >= 0) {
          double *a = anchors + ((long)(t - t0) * n_anchor + slot) * 6 * a_ld + i;
          a[0] = p.x;
          a[a_ld] = p.xp;
          a[2 * a_ld] = p.y;
          a[3 * a_ld] = p.yp;
          a[4 * a_ld] = p.s;
          a[5 * a_ld] = p.dp;
        }
        u += d->counted;
        if (eh_apply_element_t<false>(p, *d, t, terms, d->zpos) >= 0)
          live = false;
      }
    }
  }
  if (in) {
    if (alive[i] && !live)
      alive[i] = 0;
    z[i] = p.x;
    z[ld + i] = p.xp;
    z[2 * ld + i] = p.y;
    z[3 * ld + i] = p.yp;
    z[4 * ld + i] = p.s;
    z[5 * ld + i] = p.dp;
  }
  wave_add(units, u);
}

/* out[c*m + i] = in[c*ld + rows[i]] */
__global__ void k_adj_gather(const double *in, long ld, const int32_t *rows, long m, double *out) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m)
    for (int c = 0; c < 6; c++)
      out[c * m + i] = in[c * ld + rows[i]];
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1)
    v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ eh_p6 load_anchor(const double *anchors, long idx, long a_ld, long i) {
  const double *a = anchors + idx * 6 * a_ld + i;
  return {a[0], a[a_ld], a[2 * a_ld], a[3 * a_ld], a[4 * a_ld], a[5 * a_ld]};
}

/* reverse sweep over turns [t0, t1) for m survivors, one wavefront per block.
 * zb/lam hold the state and adjoint at the end of turn t1-1 on entry and at
 * the start of turn t0 on exit.  Survivor i's anchors sit at column arow[i]
 * of turn (t - ta0). */
/* three wavefronts per SIMD: at most 168 VGPRs (the RF cavity's dual-number
 * Jacobian and the bend's parameter gradients sit at that limit) */
__global__ void __launch_bounds__(EH_ADJ_BLOCK) __attribute__((amdgpu_waves_per_eu(3)))
k_adj_bwd(double *zb, double *lam, long m, const int32_t *arow,
                                                           const double *__restrict__ anchors, int n_anchor,
                                                           long a_ld, const eh_desc *__restrict__ descs,
                                                           const eh_fterm *__restrict__ terms, int ne, int t0,
                                                           int t1, int ta0, int n_params, double *gpart,
                                                           double *errpart, const int32_t *cvar) {
  extern __shared__ double g_lds[];
  const int lane = threadIdx.x;
  /* lattice variants: the block (one wavefront) is one 64-entry chunk of one variant */
  if (cvar)
    descs += (long)cvar[blockIdx.x] * ne;
  for (int q = lane; q < n_params; q += EH_ADJ_BLOCK)
    g_lds[q] = 0;
  __syncthreads();
  long i = (long)blockIdx.x * EH_ADJ_BLOCK + lane;
  const bool act = i < m;
  const long ii = act ? i : 0; /* idle lanes follow survivor 0 with a zero adjoint */
  eh_p6 z = {zb[ii], zb[m + ii], zb[2 * m + ii], zb[3 * m + ii], zb[4 * m + ii], zb[5 * m + ii]};
  eh_p6 l = {0, 0, 0, 0, 0, 0};
  if (act)
    l = {lam[i], lam[m + i], lam[2 * m + i], lam[3 * m + i], lam[4 * m + i], lam[5 * m + i]};
  const long col = arow[ii];
  double err = 0, err_s = 0;
  for (int t = t1 - 1; t >= t0; t--) {
    const long tb = (long)(t - ta0) * n_anchor;
    for (int e = ne - 1; e >= 0; e--) {
      const eh_desc &d = descs[e];
      switch (d.type) {
      case EH_T_DRIF:
        eh_adj_drif(z, l, d.length, d.T11);
        break;
      case EH_T_EDRIFT:
        eh_adj_edrift(z, l, d.length);
        break;
      case EH_T_KQUAD:
      case EH_T_KSEXT:
      case EH_T_KOCT:
      case EH_T_KQUSE: {
        eh_pgrad g = {0, 0, 0, 0};
        eh_adj_multipole(z, l, d, g);
        if (d.adj_pk >= 0) {
          double v = wave_sum(g.k);
          if (lane == 0)
            g_lds[d.adj_pk] += v;
        }
        if (d.adj_pdx >= 0) {
          double v = wave_sum(g.dx);
          if (lane == 0)
            g_lds[d.adj_pdx] += v;
        }
        if (d.adj_pdy >= 0) {
          double v = wave_sum(g.dy);
          if (lane == 0)
            g_lds[d.adj_pdy] += v;
        }
        break;
      }
      case EH_T_CSBEND: {
        eh_pgrad g = {0, 0, 0, 0};
        eh_adj_csbend(z, l, d, terms, g);
        if (d.adj_pk >= 0) {
          double v = wave_sum(g.k);
          if (lane == 0)
            g_lds[d.adj_pk] += v;
        }
        if (d.adj_pfse >= 0) {
          double v = wave_sum(g.fse);
          if (lane == 0)
            g_lds[d.adj_pfse] += v;
        }
        break;
      }
      case EH_T_RFCA: {
        eh_p6 zin = load_anchor(anchors, tb + (e == 0 ? 0 : d.adj_anchor), a_ld, col);
        eh_adj_rfca(z, l, zin, d);
        break;
      }
      default:
        break;
      }
      if (e == 0) {
        eh_p6 za = load_anchor(anchors, tb, a_ld, col);
        err = fmax(err, fmax(fmax(eh_fabs(z.x - za.x), eh_fabs(z.xp - za.xp)),
                             fmax(fmax(eh_fabs(z.y - za.y), eh_fabs(z.yp - za.yp)), eh_fabs(z.dp - za.dp))));
        err_s = fmax(err_s, eh_fabs(z.s - za.s));
        z = za;
      }
    }
  }
  if (act) {
    zb[i] = z.x;
    zb[m + i] = z.xp;
    zb[2 * m + i] = z.y;
    zb[3 * m + i] = z.yp;
    zb[4 * m + i] = z.s;
    zb[5 * m + i] = z.dp;
    lam[i] = l.x;
    lam[m + i] = l.xp;
    lam[2 * m + i] = l.y;
    lam[3 * m + i] = l.yp;
    lam[4 * m + i] = l.s;
    lam[5 * m + i] = l.dp;
  }
  err = wave_max(err);
  err_s = wave_max(err_s);
  __syncthreads();
  double *gp = gpart + (long)blockIdx.x * n_params;
  for (int q = lane; q < n_params; q += EH_ADJ_BLOCK)
    gp[q] += g_lds[q];
  if (lane == 0) {
    errpart[2 * blockIdx.x] = fmax(errpart[2 * blockIdx.x], err);
    errpart[2 * blockIdx.x + 1] = fmax(errpart[2 * blockIdx.x + 1], err_s);
  }
}

/* out[p] = sum over blocks b (in order) of part[b * np + p] */
__global__ void k_adj_reduce(const double *part, long nb, int np, double *out) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < np) {
    double acc = 0;
    for (long b = 0; b < nb; b++)
      acc += part[b * np + q];
    out[q] = acc;
  }
}

__global__ void k_fill(double *p, long n, double v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = v;
}

struct eh_adjoint {
  long n = 0, ne = 0, n_params = 0, n_turns = 0, C = 1;
  int n_anchor = 1;
  bool store_all = false;
  std::vector<eh_desc> hd;
  std::vector<eh_fterm> pool;
  double *d_z = nullptr;       /* 6 x n, fixed rows */
  int32_t *d_alive = nullptr;  /* n */
  double *d_ckpt = nullptr;    /* (n_ckpt) x 6 x n turn-start states */
  double *d_anchor = nullptr;  /* anchors: all turns x n (store_all) or one segment x survivors */
  size_t anchor_bytes = 0;
  eh_desc *d_desc = nullptr;
  eh_fterm *d_terms = nullptr;
  unsigned long long *d_cnt = nullptr;
  std::vector<double> z_final; /* 6 x n */
  std::vector<int32_t> alive;
  bool have_forward = false;
  double fwd_ms = 0;
  double units = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  /* lattice variants (eh_adjoint_forward_batch): variant v's particles occupy
   * rows [var_first[v], var_first[v] + var_count[v]) of a layout padded to 64
   * rows per variant; cvar[chunk] is the variant of each 64-row chunk, and the
   * caller's k-th particle sits at row row_of[k].  Empty for one lattice. */
  long nvar = 1, n_caller = 0;
  std::vector<long> var_first, var_count, row_of;
  std::vector<int32_t> cvar;
  int32_t *d_cvar = nullptr;
};

static void adj_release(eh_adjoint *a) {
  void *ptrs[] = {a->d_z, a->d_alive, a->d_ckpt, a->d_anchor, a->d_desc, a->d_terms, a->d_cnt, a->d_cvar};
  for (void *p : ptrs)
    if (p)
      (void)hipFree(p);
  a->d_z = nullptr;
  a->d_alive = nullptr;
  a->d_ckpt = nullptr;
  a->d_anchor = nullptr;
  a->d_desc = nullptr;
  a->d_terms = nullptr;
  a->d_cnt = nullptr;
  a->d_cvar = nullptr;
  a->anchor_bytes = 0;
  a->have_forward = false;
}

extern "C" eh_adjoint *eh_adjoint_create(void) {
  if (g_device < 0 && eh_init(0) != EH_OK)
    return nullptr;
  eh_adjoint *a = new eh_adjoint();
  if (hipStreamCreate(&a->stream) != hipSuccess || hipEventCreate(&a->ev0) != hipSuccess ||
      hipEventCreate(&a->ev1) != hipSuccess) {
    fail(EH_ERR_HIP, "eh_adjoint_create: stream setup failed");
    delete a;
    return nullptr;
  }
  return a;
}

extern "C" void eh_adjoint_free(eh_adjoint *a) {
  if (!a)
    return;
  adj_release(a);
  if (a->ev0)
    (void)hipEventDestroy(a->ev0);
  if (a->ev1)
    (void)hipEventDestroy(a->ev1);
  if (a->stream)
    (void)hipStreamDestroy(a->stream);
  delete a;
}

static long adj_timed(eh_adjoint *a, double *acc) {
  HIPCHK(hipEventRecord(a->ev1, a->stream));
  HIPCHK(hipEventSynchronize(a->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, a->ev0, a->ev1));
  *acc += ms;
  return EH_OK;
}

/* descriptors of a beamline as one do_tracking call sees them */
/* gradient slots per element: strength (K of the kick multipoles, K1 of a
 * bend), DX, DY (kick multipoles), FSE (bends) */
#define EH_ADJ_KINDS 4

/* the field of a bend per unit K1 and per unit FSE, for the adjoint's
 * parameter gradients: the field table is linear in the b[] coefficients, so
 * each is the table of the coefficients' derivatives (src/csbend.c:800-857:
 * b1 = K1 rho0 (1 + FSE + FSE_QUADRUPOLE) / (1 + FSE + FSE_DIPOLE), b2 = K2 rho0
 * (1 + FSE) / (1 + FSE + FSE_DIPOLE), odd terms negated for a negative angle),
 * at the bend's own expansion order, without the constant dipole term.  Appends
 * the term lists to pool and points the descriptor at them. */
static long adj_bend_params(eh_desc &d, const EH_CSBEND *cb, std::vector<eh_fterm> &pool) {
  d.adj_dk_nfx = d.adj_dk_nfy = d.adj_df_nfx = d.adj_df_nfy = 0;
  d.adj_dk_fx0 = d.adj_dk_fy0 = d.adj_df_fx0 = d.adj_df_fy0 = 0;
  if (d.adj_pk < 0 && d.adj_pfse < 0)
    return EH_OK;
  if (d.type != EH_T_CSBEND)
    return fail(EH_ERR_UNSUPPORTED, "a CSBEND without angle has no K1 or FSE gradient");
  if (d.adj_pk >= 0 && cb->use_bn)
    return fail(EH_ERR_UNSUPPORTED, "CSBEND with USE_BN: K1 is not a parameter");
  const double rho0s = cb->length / cb->angle, flip = cb->angle < 0 ? -1.0 : 1.0;
  const double F = cb->fse, fd = cb->fseDipole, fq = cb->fseQuadrupole, den = 1 + F + fd;
  double b[9], cc[9], bk[9], bf[9], zero[9];
  csbend_coefficients(cb, b, cc);
  for (int j = 0; j < 9; j++)
    bk[j] = bf[j] = zero[j] = 0;
  if (cb->angle < 0)
    for (int j = 1; j < 9; j += 2) {
      b[j] *= -1;
      cc[j] *= -1;
    }
  bk[1] = rho0s * ((1 + F + fq) / den) * flip;
  const double s1 = cb->use_bn ? cb->b1 : cb->k1 * rho0s, s2 = cb->use_bn ? cb->b2 : cb->k2 * rho0s;
  bf[1] = s1 * ((fd - fq) / (den * den)) * flip;
  bf[2] = s2 * (fd / (den * den));
  const double h = 1 / (cb->length / std::fabs(cb->angle)); /* as build_csbend */
  double Fx[121], Fy[121];
  long eo1 = eh_csbend_field_table(b, cc, h, cb->nonlinear, cb->expansionOrder, Fx, Fy, nullptr, nullptr);
  if (eo1 < 0)
    return eo1;
  for (int which = 0; which < 2; which++) {
    if ((which ? d.adj_pfse : d.adj_pk) < 0)
      continue;
    long r = eh_csbend_field_table(which ? bf : bk, zero, h, cb->nonlinear, (int)eo1 - 1, Fx, Fy, nullptr, nullptr);
    if (r < 0)
      return r;
    Fy[0] -= 1; /* the dipole's own field, 1 - b[0] with b[0] = 0 */
    for (int c = 0; c < 2; c++) {
      (which ? (c ? d.adj_df_fy0 : d.adj_df_fx0) : (c ? d.adj_dk_fy0 : d.adj_dk_fx0)) = (int32_t)pool.size();
      int32_t n = 0;
      for (int i = 0; i < eo1; i++)
        for (int j = 0; j < eo1 - i; j++) {
          const double v = c ? Fy[i * 11 + j] : Fx[i * 11 + j];
          if (v != 0) {
            pool.push_back({i, j, v});
            n++;
          }
        }
      (which ? (c ? d.adj_df_nfy : d.adj_df_nfx) : (c ? d.adj_dk_nfy : d.adj_dk_nfx)) = n;
    }
  }
  d.adj_fse1 = 1 / den;
  return EH_OK;
}

static long build_line(EH_ELEMENT_LIST *bl, long ne, double Pc, std::vector<eh_desc> &hd,
                       std::vector<eh_fterm> &pool) {
  hd.assign(ne, eh_desc());
  pool.clear();
  std::vector<double> zs(ne), ze(ne);
  const EH_MAXAMP *last_amp = nullptr;
  double z = 0;
  for (long ie = 0; ie < ne; ie++) {
    zs[ie] = z;
    z += elem_length(bl + ie);
    ze[ie] = z;
    if (bl[ie].type == EH_T_MAXAMP)
      last_amp = (const EH_MAXAMP *)bl[ie].p_elem;
  }
  const EH_MAXAMP *cur0 = nullptr, *cur1 = last_amp;
  for (long ie = 0; ie < ne; ie++) {
    if (bl[ie].type == EH_T_MAXAMP)
      cur0 = cur1 = (const EH_MAXAMP *)bl[ie].p_elem;
    eh_ctx c = {};
    c.Po = Pc;
    c.z_start = zs[ie];
    c.z_end = ze[ie];
    c.ap[0] = cur0;
    c.ap[1] = cur1;
    for (int pi = 0; pi < 2; pi++) {
      const EH_MAXAMP *m = pi ? cur1 : cur0;
      c.post_x[pi] = m ? m->x_max : 0;
      c.post_y[pi] = m ? m->y_max : 0;
      c.post_on[pi] = (bl[ie].type != EH_T_MARK) && m && (m->x_max || m->y_max);
      if (m && m->elliptical && m->x_max > 0 && m->y_max > 0)
        return fail(EH_ERR_UNSUPPORTED, "the adjoint does not support elliptical MAXAMP");
    }
    long r = build_desc(&hd[ie], bl + ie, c, pool);
    if (r < 0)
      return r;
  }
  return EH_OK;
}

/* descriptors of nvar variants of one beamline structure, variant-major, with
 * the adjoint's fields set (parameter slots, RF anchor slots); every variant's
 * field-term offsets are shifted into one shared pool */
static long adj_prepare(eh_adjoint *a, EH_ELEMENT_LIST *const *beamlines, long nvar, long ne,
                        const int32_t *param_slots, long n_params, double Pc, const long *counts) {
  a->hd.clear();
  a->pool.clear();
  int n_anchor = 1;
  for (long v = 0; v < nvar; v++) {
    std::vector<eh_desc> hd;
    std::vector<eh_fterm> pool;
    long r = build_line(beamlines[v], ne, Pc, hd, pool);
    if (r < 0)
      return r;
    int na = 1;
    for (long e = 0; e < ne; e++) {
      eh_desc &d = hd[e];
      const int t = beamlines[v][e].type;
      if (t != beamlines[0][e].type)
        return fail(EH_ERR_ARG, "variant %ld element %ld: types differ from variant 0", v, e);
      if ((t == EH_T_KQUAD && ((const EH_KQUAD *)beamlines[v][e].p_elem)->bore) ||
          (t == EH_T_KSEXT && ((const EH_KSEXT *)beamlines[v][e].p_elem)->bore) ||
          (t == EH_T_KOCT && ((const EH_KOCT *)beamlines[v][e].p_elem)->bore))
        return fail(EH_ERR_UNSUPPORTED, "element %ld: the adjoint does not support strengths set by BORE and B", e);
      if (t == EH_T_RFCA) {
        const EH_RFCA *rf = (const EH_RFCA *)beamlines[v][e].p_elem;
        if (rf->phase_reference)
          return fail(EH_ERR_UNSUPPORTED, "element %ld: the adjoint does not support PHASE_REFERENCE", e);
        if (rf->change_t || rf->change_p0 || rf->Q != 0 || rf->linearize || rf->lockPhase)
          return fail(EH_ERR_UNSUPPORTED,
                      "element %ld: the adjoint does not support CHANGE_T, CHANGE_P0, Q, LINEARIZE or LOCK_PHASE", e);
        /* a variant without particles never reads its descriptors */
        if (!rf->fiducial_seen && rf->volt != 0 && !(counts && counts[v] == 0))
          return fail(EH_ERR_ARG, "element %ld: RFCA fiducial unknown; track one pass first to fix it", e);
        if (!d.rf_zero_volt && !d.rf_matrix)
          return fail(EH_ERR_UNSUPPORTED, "element %ld: adjoint supports the RFCA matrix method only", e);
        if (e > 0)
          d.adj_anchor = na++;
      }
      /* k_adj_fwd runs the lean map */
      if (t == EH_T_MULT || t == EH_T_ROTATE || t == EH_T_MALIGN || t == EH_T_ENERGY)
        return fail(EH_ERR_UNSUPPORTED, "element %ld: the adjoint does not support MULT, ROTATE, MALIGN or ENERGY",
                    e);
      if (d.rad)
        return fail(EH_ERR_UNSUPPORTED, "element %ld: the adjoint does not support SYNCH_RAD", e);
      if (d.full_only)
        return fail(EH_ERR_UNSUPPORTED,
                    "element %ld: the adjoint does not support elliptical MAXAMP, CSBEND edge methods 2 to 4 or "
                    "edge kick limits", e);
      if (!eh_lean_capable(d))
        return fail(EH_ERR_STATE, "element %ld: not trackable by the adjoint's map", e);
      if ((t == EH_T_KQUAD || t == EH_T_KSEXT || t == EH_T_KOCT || t == EH_T_KQUSE) && d.expand)
        return fail(EH_ERR_UNSUPPORTED, "element %ld: adjoint needs EXPAND_HAMILTONIAN=0", e);
      const bool kick = t == EH_T_KQUAD || t == EH_T_KSEXT || t == EH_T_KOCT || t == EH_T_KQUSE;
      for (int k = 0; k < EH_ADJ_KINDS; k++) {
        int32_t slot = param_slots ? param_slots[EH_ADJ_KINDS * e + k] : -1;
        if (slot < -1 || slot >= n_params)
          return fail(EH_ERR_ARG, "element %ld: parameter slot %d out of range", e, slot);
        if (slot >= 0 && !(kick && k <= 2) && !(t == EH_T_CSBEND && (k == 0 || k == 3)))
          return fail(EH_ERR_UNSUPPORTED,
                      "element %ld: parameters are K, DX, DY of KQUAD/KSEXT/KOCT/KQUSE and K1, FSE of CSBEND", e);
        (k == 0 ? d.adj_pk : k == 1 ? d.adj_pdx : k == 2 ? d.adj_pdy : d.adj_pfse) = slot;
      }
      if (t == EH_T_CSBEND) {
        long r2 = adj_bend_params(d, (const EH_CSBEND *)beamlines[v][e].p_elem, pool);
        if (r2 < 0)
          return r2;
      }
      if (t == EH_T_CSBEND && !pool.empty()) {
        const int32_t shift = (int32_t)a->pool.size();
        d.csb_fx0 += shift;
        d.csb_fy0 += shift;
        if (d.csb_dense) {
          d.csb_dx0 += shift;
          d.csb_dy0 += shift;
        }
        d.adj_dk_fx0 += shift;
        d.adj_dk_fy0 += shift;
        d.adj_df_fx0 += shift;
        d.adj_df_fy0 += shift;
      }
    }
    n_anchor = na;
    a->hd.insert(a->hd.end(), hd.begin(), hd.end());
    a->pool.insert(a->pool.end(), pool.begin(), pool.end());
  }
  a->n_anchor = n_anchor;
  return EH_OK;
}

/* the forward pass over n rows already laid out (6 x n, row-major by
 * coordinate in h); alive0 marks rows to track; a->cvar (if any) the variants */
static long adj_forward_rows(eh_adjoint *a, const std::vector<double> &h, const std::vector<int32_t> &alive0, long n,
                             long ne, long n_params, long n_turns, long checkpoint_every) {
  a->n = n;
  a->ne = ne;
  a->n_params = n_params;
  a->n_turns = n_turns;
  a->C = checkpoint_every;
  a->store_all = checkpoint_every == 1;
  const int n_anchor = a->n_anchor;
  const long n_ckpt = (n_turns + a->C - 1) / a->C;
  HIPCHK(hipMalloc(&a->d_z, sizeof(double) * 6 * n));
  HIPCHK(hipMalloc(&a->d_alive, sizeof(int32_t) * n));
  HIPCHK(hipMalloc(&a->d_cnt, sizeof(unsigned long long) * 4));
  HIPCHK(hipMalloc(&a->d_desc, sizeof(eh_desc) * a->hd.size()));
  HIPCHK(hipMalloc(&a->d_terms, sizeof(eh_fterm) * (a->pool.size() + 1)));
  if (!a->cvar.empty()) {
    HIPCHK(hipMalloc(&a->d_cvar, sizeof(int32_t) * a->cvar.size()));
    HIPCHK(hipMemcpy(a->d_cvar, a->cvar.data(), sizeof(int32_t) * a->cvar.size(), hipMemcpyHostToDevice));
  }
  if (a->store_all) {
    a->anchor_bytes = sizeof(double) * 6 * (size_t)n * n_anchor * n_turns;
    HIPCHK(hipMalloc(&a->d_anchor, a->anchor_bytes));
  } else {
    HIPCHK(hipMalloc(&a->d_ckpt, sizeof(double) * 6 * (size_t)n * n_ckpt));
  }
  HIPCHK(hipMemcpy(a->d_desc, a->hd.data(), sizeof(eh_desc) * a->hd.size(), hipMemcpyHostToDevice));
  if (!a->pool.empty())
    HIPCHK(hipMemcpy(a->d_terms, a->pool.data(), sizeof(eh_fterm) * a->pool.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(a->d_z, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(a->d_alive, alive0.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(a->d_cnt, 0, sizeof(unsigned long long) * 4));
  a->fwd_ms = 0;
  HIPCHK(hipEventRecord(a->ev0, a->stream));
  for (long k = 0; k < n_ckpt; k++) {
    int t0 = (int)(k * a->C), t1 = (int)std::min(n_turns, (k + 1) * a->C);
    double *anc = nullptr;
    if (a->store_all)
      anc = a->d_anchor + (size_t)t0 * n_anchor * 6 * n;
    else
      HIPCHK(hipMemcpyAsync(a->d_ckpt + (size_t)k * 6 * n, a->d_z, sizeof(double) * 6 * n, hipMemcpyDeviceToDevice,
                            a->stream));
    hipLaunchKernelGGL(k_adj_fwd, dim3(grid_for(n)), dim3(EH_BLOCK), 0, a->stream, a->d_z, n, a->d_alive, n,
                       a->d_desc, a->d_terms, (int)ne, t0, t1, anc, n_anchor, n, a->d_cnt, a->d_cvar);
    HIPCHK(hipGetLastError());
  }
  long r = adj_timed(a, &a->fwd_ms);
  if (r < 0)
    return r;
  unsigned long long cnt[4];
  HIPCHK(hipMemcpy(cnt, a->d_cnt, sizeof cnt, hipMemcpyDeviceToHost));
  a->units = (double)cnt[0];
  a->z_final.resize(6 * (size_t)n);
  a->alive.resize(n);
  HIPCHK(hipMemcpy(a->z_final.data(), a->d_z, sizeof(double) * 6 * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(a->alive.data(), a->d_alive, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  a->have_forward = true;
  return EH_OK;
}

/* rows of 10 doubles -> 6 x n by coordinate at the given rows; the initial NaN
 * sweep of do_tracking (src/do_tracking.c:478-483) clears alive */
static void adj_scatter(const double *coord, long k, long row, long n, std::vector<double> &h,
                        std::vector<int32_t> &al) {
  al[row] = 1;
  for (int c = 0; c < 6; c++) {
    double x = coord[10 * k + c];
    h[c * (size_t)n + row] = x;
    if (std::isnan(x) || std::isinf(x))
      al[row] = 0;
  }
}

extern "C" long eh_adjoint_forward(eh_adjoint *a, const double *coord, long n, EH_ELEMENT_LIST *beamline, long ne,
                                   const int32_t *param_slots, long n_params, long n_turns, double P_central,
                                   long checkpoint_every, double *final_coord, int32_t *alive_out) {
  if (!a || !coord || n < 1 || !beamline || ne < 1 || n_turns < 1 || n_params < 0 || checkpoint_every < 1)
    return fail(EH_ERR_ARG, "eh_adjoint_forward: bad arguments");
  if (n > 2147483647L)
    return fail(EH_ERR_ARG, "eh_adjoint_forward: too many particles");
  if (n_params > 8000)
    return fail(EH_ERR_ARG, "eh_adjoint_forward: at most 8000 parameters (one LDS slot each)");
  adj_release(a);
  a->nvar = 1;
  a->n_caller = n;
  a->var_first.clear();
  a->var_count.clear();
  a->row_of.clear();
  a->cvar.clear();
  EH_ELEMENT_LIST *const lines[1] = {beamline};
  long r = adj_prepare(a, lines, 1, ne, param_slots, n_params, P_central, nullptr);
  if (r < 0)
    return r;
  std::vector<double> h(6 * (size_t)n);
  std::vector<int32_t> al(n, 0);
  for (long i = 0; i < n; i++)
    adj_scatter(coord, i, i, n, h, al);
  r = adj_forward_rows(a, h, al, n, ne, n_params, n_turns, checkpoint_every);
  if (r < 0)
    return r;
  long n_alive = 0;
  for (long i = 0; i < n; i++) {
    n_alive += a->alive[i] != 0;
    if (final_coord)
      for (int c = 0; c < 6; c++)
        final_coord[6 * i + c] = a->z_final[c * (size_t)n + i];
    if (alive_out)
      alive_out[i] = a->alive[i];
  }
  return n_alive;
}

extern "C" long eh_adjoint_forward_batch(eh_adjoint *a, const double *coord, const long *n_per_variant, long nvar,
                                         EH_ELEMENT_LIST *const *beamlines, long ne, const int32_t *param_slots,
                                         long n_params, long n_turns, double P_central, long checkpoint_every,
                                         double *final_coord, int32_t *alive_out) {
  if (!a || !coord || !n_per_variant || nvar < 1 || !beamlines || ne < 1 || n_turns < 1 || n_params < 0 ||
      checkpoint_every < 1)
    return fail(EH_ERR_ARG, "eh_adjoint_forward_batch: bad arguments");
  if (n_params > 8000)
    return fail(EH_ERR_ARG, "eh_adjoint_forward_batch: at most 8000 parameters (one LDS slot each)");
  adj_release(a);
  a->nvar = nvar;
  a->var_first.assign(nvar, 0);
  a->var_count.assign(nvar, 0);
  a->cvar.clear();
  long rows = 0, total = 0;
  for (long v = 0; v < nvar; v++) {
    if (n_per_variant[v] < 0)
      return fail(EH_ERR_ARG, "eh_adjoint_forward_batch: negative particle count");
    a->var_first[v] = rows;
    a->var_count[v] = n_per_variant[v];
    long chunks = (n_per_variant[v] + 63) / 64;
    for (long c = 0; c < chunks; c++)
      a->cvar.push_back((int32_t)v);
    rows += chunks * 64;
    total += n_per_variant[v];
  }
  if (total < 1 || rows > 2147483647L)
    return fail(EH_ERR_ARG, "eh_adjoint_forward_batch: particle count out of range");
  a->n_caller = total;
  long r = adj_prepare(a, beamlines, nvar, ne, param_slots, n_params, P_central, n_per_variant);
  if (r < 0)
    return r;
  a->row_of.assign(total, 0);
  std::vector<double> h(6 * (size_t)rows, 0.0);
  std::vector<int32_t> al(rows, 0);
  for (long v = 0, k = 0; v < nvar; v++)
    for (long j = 0; j < n_per_variant[v]; j++, k++) {
      a->row_of[k] = a->var_first[v] + j;
      adj_scatter(coord, k, a->row_of[k], rows, h, al);
    }
  r = adj_forward_rows(a, h, al, rows, ne, n_params, n_turns, checkpoint_every);
  if (r < 0)
    return r;
  long n_alive = 0;
  for (long k = 0; k < total; k++) {
    const long i = a->row_of[k];
    n_alive += a->alive[i] != 0;
    if (final_coord)
      for (int c = 0; c < 6; c++)
        final_coord[6 * k + c] = a->z_final[c * (size_t)rows + i];
    if (alive_out)
      alive_out[k] = a->alive[i];
  }
  return n_alive;
}

extern "C" long eh_adjoint_backward(eh_adjoint *a, const double *grad_final, double *grad_params,
                                    double *grad_initial, eh_adjoint_stats *stats) {
  if (!a || !grad_final || (a->n_params > 0 && !grad_params))
    return fail(EH_ERR_ARG, "eh_adjoint_backward: bad arguments");
  if (!a->have_forward)
    return fail(EH_ERR_STATE, "eh_adjoint_backward: no forward pass recorded");
  const long n = a->n, ne = a->ne;
  /* survivors in row order; with lattice variants, each variant's survivors
   * padded to whole 64-entry chunks with copies of its first survivor that
   * carry a zero adjoint (they contribute exactly nothing), so that every
   * wavefront of the reverse sweep reads one variant's descriptors */
  std::vector<int32_t> surv, cvar_b;
  std::vector<char> real;
  long n_real = 0;
  if (a->cvar.empty()) {
    for (long i = 0; i < n; i++)
      if (a->alive[i])
        surv.push_back((int32_t)i);
    real.assign(surv.size(), 1);
    n_real = (long)surv.size();
  } else {
    for (long v = 0; v < a->nvar; v++) {
      const size_t start = surv.size();
      for (long j = 0; j < a->var_count[v]; j++) {
        const long i = a->var_first[v] + j;
        if (a->alive[i]) {
          surv.push_back((int32_t)i);
          real.push_back(1);
        }
      }
      const size_t got = surv.size() - start;
      if (!got)
        continue;
      n_real += (long)got;
      while ((surv.size() - start) % 64) {
        surv.push_back(surv[start]);
        real.push_back(0);
      }
      for (size_t c = 0; c < (surv.size() - start) / 64; c++)
        cvar_b.push_back((int32_t)v);
    }
  }
  const long m = (long)surv.size();
  const int np = (int)a->n_params;
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->n_survivors = n_real;
    stats->forward_ms = a->fwd_ms;
    stats->forward_units = a->units;
  }
  if (grad_initial)
    std::fill(grad_initial, grad_initial + 6 * a->n_caller, 0.0);
  /* the caller's index of each row (identity for one lattice) */
  std::vector<long> caller_of;
  if (!a->row_of.empty()) {
    caller_of.assign(n, -1);
    for (long k = 0; k < a->n_caller; k++)
      caller_of[a->row_of[k]] = k;
  }
  auto caller = [&](long row) { return caller_of.empty() ? row : caller_of[row]; };
  if (m == 0) {
    if (np)
      std::fill(grad_params, grad_params + np, 0.0);
    return 0;
  }
  const long nb = (m + EH_ADJ_BLOCK - 1) / EH_ADJ_BLOCK;
  const int npl = np > 0 ? np : 1;
  double *d_zb = nullptr, *d_lam = nullptr, *d_gpart = nullptr, *d_err = nullptr, *d_grad = nullptr,
         *d_seg = nullptr;
  int32_t *d_surv = nullptr, *d_ident = nullptr, *d_cvar_b = nullptr;
  auto cleanup = [&]() {
    void *ptrs[] = {d_zb, d_lam, d_gpart, d_err, d_grad, d_seg, d_surv, d_ident, d_cvar_b};
    for (void *p : ptrs)
      if (p)
        (void)hipFree(p);
  };
  long r = EH_OK;
  do {
    std::vector<double> h(6 * (size_t)m);
    if (hipMalloc(&d_zb, sizeof(double) * 6 * m) != hipSuccess || hipMalloc(&d_lam, sizeof(double) * 6 * m) ||
        hipMalloc(&d_gpart, sizeof(double) * nb * npl) || hipMalloc(&d_err, sizeof(double) * 2 * nb) ||
        hipMalloc(&d_grad, sizeof(double) * npl) || hipMalloc(&d_surv, sizeof(int32_t) * m) ||
        hipMalloc(&d_ident, sizeof(int32_t) * m)) {
      r = fail(EH_ERR_NOMEM, "eh_adjoint_backward: device allocation failed");
      break;
    }
    for (long j = 0; j < m; j++)
      for (int c = 0; c < 6; c++)
        h[c * (size_t)m + j] = a->z_final[c * (size_t)n + surv[j]];
    if (hipMemcpy(d_zb, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
      r = fail(EH_ERR_HIP, "eh_adjoint_backward: upload failed");
      break;
    }
    for (long j = 0; j < m; j++)
      for (int c = 0; c < 6; c++)
        h[c * (size_t)m + j] = real[j] ? grad_final[6 * (size_t)caller(surv[j]) + c] : 0.0;
    if (!cvar_b.empty()) {
      if (hipMalloc(&d_cvar_b, sizeof(int32_t) * cvar_b.size()) != hipSuccess ||
          hipMemcpy(d_cvar_b, cvar_b.data(), sizeof(int32_t) * cvar_b.size(), hipMemcpyHostToDevice) != hipSuccess) {
        r = fail(EH_ERR_HIP, "eh_adjoint_backward: variant table upload failed");
        break;
      }
    }
    std::vector<int32_t> ident(m);
    for (long j = 0; j < m; j++)
      ident[j] = (int32_t)j;
    if (hipMemcpy(d_lam, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_surv, surv.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_ident, ident.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice) != hipSuccess) {
      r = fail(EH_ERR_HIP, "eh_adjoint_backward: upload failed");
      break;
    }
    if (!a->store_all) {
      size_t seg = sizeof(double) * 6 * (size_t)m * a->n_anchor * a->C;
      if (hipMalloc(&d_seg, seg) != hipSuccess) {
        r = fail(EH_ERR_NOMEM, "eh_adjoint_backward: %zu bytes for the segment anchors", seg);
        break;
      }
      a->anchor_bytes = seg;
    }
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((nb * npl + 255) / 256)), dim3(256), 0, a->stream, d_gpart, nb * npl,
                       0.0);
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((2 * nb + 255) / 256)), dim3(256), 0, a->stream, d_err, 2 * nb, 0.0);
    const long n_ckpt = (a->n_turns + a->C - 1) / a->C;
    double bwd_ms = 0, rec_ms = 0;
    const size_t lds = sizeof(double) * npl;
    for (long k = n_ckpt - 1; k >= 0 && r == EH_OK; k--) {
      int t0 = (int)(k * a->C), t1 = (int)std::min(a->n_turns, (k + 1) * a->C);
      const double *anc;
      const int32_t *arow;
      long a_ld;
      int ta0;
      if (a->store_all) {
        anc = a->d_anchor;
        arow = d_surv;
        a_ld = n;
        ta0 = 0;
      } else {
        /* re-track the survivors over this segment, storing their anchors */
        HIPCHK(hipEventRecord(a->ev0, a->stream));
        double *zs = d_seg; /* gather into the first slot, then track in place there */
        double *ztmp = nullptr;
        if (hipMalloc(&ztmp, sizeof(double) * 6 * m) != hipSuccess) {
          r = fail(EH_ERR_NOMEM, "eh_adjoint_backward: re-tracking buffer");
          break;
        }
        int32_t *al1 = nullptr;
        if (hipMalloc(&al1, sizeof(int32_t) * m) != hipSuccess) {
          (void)hipFree(ztmp);
          r = fail(EH_ERR_NOMEM, "eh_adjoint_backward: re-tracking buffer");
          break;
        }
        (void)zs;
        std::vector<int32_t> ones(m, 1);
        hipError_t e1 = hipMemcpyAsync(al1, ones.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, a->stream);
        if (e1 == hipSuccess) {
          hipLaunchKernelGGL(k_adj_gather, dim3(grid_for(m)), dim3(EH_BLOCK), 0, a->stream,
                             a->d_ckpt + (size_t)k * 6 * n, n, d_surv, m, ztmp);
          hipLaunchKernelGGL(k_adj_fwd, dim3(grid_for(m)), dim3(EH_BLOCK), 0, a->stream, ztmp, m, al1, m,
                             a->d_desc, a->d_terms, (int)ne, t0, t1, d_seg, a->n_anchor, m, a->d_cnt + 1,
                             d_cvar_b);
          e1 = hipGetLastError();
        }
        if (e1 == hipSuccess)
          e1 = hipStreamSynchronize(a->stream);
        (void)hipFree(ztmp);
        (void)hipFree(al1);
        if (e1 != hipSuccess) {
          r = fail(EH_ERR_HIP, "eh_adjoint_backward: re-tracking failed: %s", hipGetErrorString(e1));
          break;
        }
        r = adj_timed(a, &rec_ms);
        if (r < 0)
          break;
        anc = d_seg;
        arow = d_ident;
        a_ld = m;
        ta0 = t0;
      }
      HIPCHK(hipEventRecord(a->ev0, a->stream));
      hipLaunchKernelGGL(k_adj_bwd, dim3((unsigned)nb), dim3(EH_ADJ_BLOCK), lds, a->stream, d_zb, d_lam, m, arow, anc,
                         a->n_anchor, a_ld, a->d_desc, a->d_terms, (int)ne, t0, t1, ta0, np, d_gpart, d_err,
                         d_cvar_b);
      hipError_t e2 = hipGetLastError();
      if (e2 != hipSuccess) {
        r = fail(EH_ERR_HIP, "k_adj_bwd launch: %s", hipGetErrorString(e2));
        break;
      }
      r = adj_timed(a, &bwd_ms);
    }
    if (r < 0)
      break;
    if (np > 0) {
      hipLaunchKernelGGL(k_adj_reduce, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, a->stream, d_gpart, nb, np,
                         d_grad);
      if (hipMemcpyAsync(grad_params, d_grad, sizeof(double) * np, hipMemcpyDeviceToHost, a->stream) != hipSuccess ||
          hipStreamSynchronize(a->stream) != hipSuccess) {
        r = fail(EH_ERR_HIP, "eh_adjoint_backward: gradient download failed");
        break;
      }
    }
    std::vector<double> errs(2 * nb);
    if (hipMemcpy(errs.data(), d_err, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(h.data(), d_lam, sizeof(double) * 6 * m, hipMemcpyDeviceToHost) != hipSuccess) {
      r = fail(EH_ERR_HIP, "eh_adjoint_backward: download failed");
      break;
    }
    if (grad_initial)
      for (long j = 0; j < m; j++)
        if (real[j])
          for (int c = 0; c < 6; c++)
            grad_initial[6 * (size_t)caller(surv[j]) + c] = h[c * (size_t)m + j];
    if (stats) {
      stats->backward_ms = bwd_ms;
      stats->retrack_ms = rec_ms;
      for (long b = 0; b < nb; b++) {
        stats->max_recon_err = std::max(stats->max_recon_err, errs[2 * b]);
        stats->max_recon_err_s = std::max(stats->max_recon_err_s, errs[2 * b + 1]);
      }
      stats->anchor_bytes = (double)a->anchor_bytes;
      stats->checkpoint_bytes = a->store_all ? 0.0 : (double)(sizeof(double) * 6 * (size_t)n * n_ckpt);
    }
  } while (0);
  cleanup();
  return r < 0 ? r : n_real;
}
