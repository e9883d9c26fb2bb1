/* elegant_oracle.c -- CPU restatement of elegant's per-particle element tracking.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is
 * compared against; nothing in the product links, loads or calls it.  It keeps
 * the reference's data layout (AoS rows of EH_NPROP doubles addressed through
 * double **), its per-element passes (element routine, MAXAMP sweep, NaN sweep),
 * its swap-to-top loss compaction and its operation order, so that with the same
 * inputs it produces the same doubles as elegant compiled with gcc -O3 and no
 * FMA contraction (Makefile.rules:218).
 *
 * Each function names the reference lines it restates.  Transcendentals come
 * from eh_math.h by default (bit-identical with the GPU build); compiled with
 * -DOR_LIBM they come from the C library instead, which is what an unmodified
 * elegant build would call -- used to measure the drift between the two.
 *
 * Parameter subset: see include/elegant_hip.h.  Unsupported values set an error
 * and make the routine return a negative count.
 *
 * Parity unpinned: the round trips the reference's radiation blocks leave when
 * no radiation is on -- q / (1 + delta) * (1 + delta) after every kick of a
 * kick multipole or bend and beta * s / beta at a bend's exit -- are the
 * identity in exact arithmetic, so no closed form or invariant can check them;
 * they are restated from the reference's text (cited where they are applied)
 * and change results only in the last bit.
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "elegant_hip.h"

#if defined(OR_DUAL)
/* every transcendental computed both ways: the C library's value is used (the
 * trajectory is elegant's as built) and the calls where the shared math
 * library would have returned another double are counted (test builds) */
#include "eh_math.h"
static long or_dual_diff;
long or_dual_count(void) { return or_dual_diff; }
void or_dual_reset(void) { or_dual_diff = 0; }
static double or_dual(double mine, double libc) {
  if (mine != libc && !(mine != mine && libc != libc))
    or_dual_diff++;
  return libc;
}
static double or_dual_sin(double x) { return or_dual(eh_sin(x), sin(x)); }
static double or_dual_cos(double x) { return or_dual(eh_cos(x), cos(x)); }
static double or_dual_tan(double x) { return or_dual(eh_tan(x), tan(x)); }
static double or_dual_asin(double x) { return or_dual(eh_asin(x), asin(x)); }
static double or_dual_log(double x) { return or_dual(eh_log(x), log(x)); }
static double or_dual_exp(double x) { return or_dual(eh_exp(x), exp(x)); }
#define OR_SIN or_dual_sin
#define OR_COS or_dual_cos
#define OR_TAN or_dual_tan
#define OR_ASIN or_dual_asin
#define OR_LOG or_dual_log
#define OR_EXP or_dual_exp
#elif defined(OR_LIBM)
#define OR_SIN sin
#define OR_COS cos
#define OR_TAN tan
#define OR_ASIN asin
#define OR_LOG log
#define OR_EXP exp
#else
#include "eh_math.h"
#define OR_SIN eh_sin
#define OR_COS eh_cos
#define OR_TAN eh_tan
#define OR_ASIN eh_asin
#define OR_LOG eh_log
#define OR_EXP eh_exp
#endif

#include "eh_csbend_terms.h"

/* OR_PLANT (tests only: checks build the oracle with a planted mistake and show
 * that they reject it): 1 R12 with the momentum ratio inverted, 2 R22 growing
 * with the energy, 3 exit path-length rescaling inverted, 4 the kick method's
 * slope ratio inverted, 5 the FINT term of Hwang's first-order fringe map with
 * cos^2 for cos^3.  0 (the default) is the reference's code. */
#ifndef OR_PLANT
#define OR_PLANT 0
#endif

/* the counting build (or_count.h) tallies the generic transport-matrix loops
 * by their non-trivial terms: these mark those loops, and do nothing here */
#ifndef OR_COUNT_MATRIX_BEGIN
#define OR_COUNT_MATRIX_BEGIN()
#define OR_COUNT_MATRIX_END()
#endif

#define PI 3.14159265358979323846
#define PIx2 (2 * PI)
#define PIo2 (PI / 2)
#define c_mks EH_C_MKS
#define particleMassMV EH_ME_MEV
#define particleRelSign 1.0
#define COORD_LIMIT 10.0 /* src/track.h:1251, src/track_data.c:52 */
#define SLOPE_LIMIT 1.0  /* src/track.h:1250, src/track_data.c:51 */

static double sqr(double x) { return x * x; }
/* SDDS's pow3 is not vendored: (x*x)*x here and in the library (unpinned) */
static double pow3(double x) { return x * x * x; }

static char or_err[256];
const char *or_last_error(void) { return or_err; }
static long or_fail(const char *msg) {
  snprintf(or_err, sizeof or_err, "%s", msg);
  return EH_ERR_UNSUPPORTED;
}

/* swapParticles, src/elegant.c:2677-2684 */
static void swapParticles(double *p1, double *p2) {
  double buffer[EH_NPROP];
  if (p1 == p2)
    return;
  memcpy(buffer, p1, sizeof buffer);
  memcpy(p1, p2, sizeof buffer);
  memcpy(p2, buffer, sizeof buffer);
}

static double SIGN(double x) { return x < 0 ? -1.0 : (x > 0 ? 1.0 : 0.0); }

/* ------------------------------------------------------------------------- */
/* limit_amplitudes, src/limit_amplitudes.c:197-287 (openCode 0) */
long or_limit_amplitudes(double **coord, double xmax, double ymax, long np, double z, double Po, long extrapolate_z) {
  long ip, itop, is_out;
  double *part;
  double dz, dzx, dzy;

  if (xmax < 0 && ymax < 0)
    return np;
  itop = np - 1;
  for (ip = 0; ip < np; ip++) {
    part = coord[ip];
    is_out = 0;
    if (xmax > 0 && fabs(part[0]) > xmax)
      is_out += 1;
    if (ymax > 0 && fabs(part[2]) > ymax)
      is_out += 2;
    if (isinf(part[0]) || isinf(part[2]) || isnan(part[0]) || isnan(part[2]))
      is_out += 4;
    dz = 0;
    if (is_out && !(is_out & 4) && extrapolate_z) {
      dzx = dzy = -DBL_MAX;
      if (is_out & 1 && part[1] != 0)
        dzx = (part[0] - SIGN(part[1]) * xmax) / part[1];
      if (is_out & 2 && part[3] != 0)
        dzy = (part[2] - SIGN(part[3]) * ymax) / part[3];
      if (dzx > dzy)
        dz = -dzx;
      else
        dz = -dzy;
      if (dz == -DBL_MAX)
        dz = 0;
    }
    if (is_out) {
      if (ip != itop)
        swapParticles(coord[ip], coord[itop]);
      coord[itop][0] += dz * coord[itop][1];
      coord[itop][2] += dz * coord[itop][3];
      coord[itop][4] = z + dz;
      coord[itop][5] = Po * (1 + coord[itop][5]);
      --itop;
      --ip;
      np--;
    }
  }
  return np;
}

/* ipow of the SDDS toolkit's mdbmth library (ipow.c), which elegant links and
 * calls for the elliptical apertures; the library is not part of the reference
 * sources.  Restated from its published algorithm for p >= 0: exponents up to 8
 * by fixed products of x^2 (x^4 = x^2 x^2, x^6 = x^2 x^2 x^2, x^7 = x^3 x^3 x,
 * x^8 = (x^2 x^2)(x^2 x^2)), larger ones by halving, x^p = h*h or h*h*x with
 * h = x^(p/2).  x == 0 gives 1 for p == 0, else 0. */
double or_ipow(double x, long p) {
  double hp;
  long n;
  if (x == 0)
    return p == 0 ? 1.0 : 0.0;
  switch (p) {
  case 0:
    return 1.0;
  case 1:
    return x;
  case 2:
    return x * x;
  case 3:
    hp = x * x;
    return hp * x;
  case 4:
    hp = x * x;
    return hp * hp;
  case 5:
    hp = x * x;
    return hp * hp * x;
  case 6:
    hp = x * x;
    return hp * hp * hp;
  case 7:
    hp = x * x * x;
    return hp * hp * x;
  case 8:
    hp = x * x;
    hp = hp * hp;
    return hp * hp;
  default:
    n = p / 2;
    hp = or_ipow(x, n);
    return (p - 2 * n) ? hp * hp * x : hp * hp;
  }
}

/* the exponent checks of elimit_amplitudes, src/limit_amplitudes.c:521-536:
 * 0 with the x and y exponents, or the error code */
static long or_ell_exponents(long exponent, long yexponent, long *xe, long *ye) {
  if ((*xe = exponent) < 2 || *xe % 2)
    return or_fail("MAXAMP EXPONENT must be an even integer >= 2");
  *ye = *xe;
  if (yexponent && ((*ye = yexponent) < 2 || *ye % 2))
    return or_fail("MAXAMP YEXPONENT must be an even integer >= 2");
  return 0;
}

/* elimit_amplitudes, src/limit_amplitudes.c:498-617 (openCode 0, no global loss
 * coordinates): non-finite particles, then particles with
 * x^xe/a2 + y^ye/b2 > 1, are moved to the top; a lost particle's dp becomes
 * sqrt((Po(1+dp))^2 + 1) and its s the loss position, extrapolated back along
 * a straight line for drifts with a true ellipse */
long or_elimit_amplitudes(double **coord, double xmax, double ymax, long np, double z, double Po, long extrapolate_z,
                          long exponent, long yexponent) {
  long ip, itop, xe, ye, r;
  double *part;
  double a2, b2, c1, c2, c0, dz, det;

  if ((r = or_ell_exponents(exponent, yexponent, &xe, &ye)) < 0)
    return r;
  if (xmax <= 0 || ymax <= 0) {
    if (xmax > 0 || ymax > 0)
      return or_limit_amplitudes(coord, xmax, ymax, np, z, Po, extrapolate_z);
    return np;
  }
  a2 = or_ipow(xmax, xe);
  b2 = or_ipow(ymax, ye);
  itop = np - 1;
  for (ip = 0; ip < np; ip++) {
    part = coord[ip];
    if (isinf(part[0]) || isinf(part[2]) || isnan(part[0]) || isnan(part[2])) {
      swapParticles(coord[ip], coord[itop]);
      coord[itop][4] = z;
      coord[itop][5] = sqrt(sqr(Po * (1 + coord[itop][5])) + 1);
      --itop;
      --ip;
      np--;
      continue;
    }
    if (!((or_ipow(part[0], xe) / a2 + or_ipow(part[2], ye) / b2) > 1))
      continue;
    dz = 0;
    if (extrapolate_z && xe == 2 && ye == 2) {
      c0 = sqr(part[0]) / a2 + sqr(part[2]) / b2 - 1;
      c1 = 2 * (part[0] * part[1] / a2 + part[2] * part[3] / b2);
      c2 = sqr(part[1]) / a2 + sqr(part[3]) / b2;
      det = sqr(c1) - 4 * c0 * c2;
      if (z > 0 && c2 && det >= 0) {
        if ((dz = (-c1 + sqrt(det)) / (2 * c2)) > 0)
          dz = (-c1 - sqrt(det)) / (2 * c2);
        if ((z + dz) < 0)
          dz = -z;
        part[0] += dz * part[1];
        part[2] += dz * part[3];
      }
    }
    swapParticles(coord[ip], coord[itop]);
    coord[itop][4] = z + dz;
    coord[itop][5] = sqrt(sqr(Po * (1 + coord[itop][5])) + 1);
    --itop;
    --ip;
    np--;
  }
  return np;
}

/* removeInvalidParticles, src/limit_amplitudes.c:289-330 */
long or_removeInvalidParticles(double **coord, long np, double z, double Po) {
  long ip, itop, is_out, ic;
  double *part;
  itop = np - 1;
  for (ip = 0; ip < np; ip++) {
    part = coord[ip];
    is_out = 0;
    for (ic = 0; ic < 6; ic++)
      if (isnan(part[ic])) {
        is_out = 1;
        break;
      }
    if (part[5] <= -1)
      is_out = 1;
    if (is_out) {
      swapParticles(coord[ip], coord[itop]);
      coord[itop][4] = z;
      coord[itop][5] = Po * (1 + coord[itop][5]);
      --itop;
      --ip;
      np--;
    }
  }
  return np;
}

/* exactDrift, src/csbend.c:3792-3816 */
void or_exactDrift(double **part, long np, double length) {
  long i;
  double *coord;
  for (i = 0; i < np; i++) {
    coord = part[i];
    coord[0] += coord[1] * length;
    coord[2] += coord[3] * length;
    coord[4] += length * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
  }
}

/* ------------------------------------------------------------------------- */
/* second-order transport map: drift_matrix (src/compute_matrices.c:391-410)
 * applied by the generic order-2 loop of track_particles (src/matrix.c:334-381,
 * TURBO_FASTMATTRACK=3 branch :373-385) */
typedef struct {
  double C[6], R[6][6], T[6][6][6];
} or_vmatrix2;

static void or_drift_matrix(or_vmatrix2 *M, double length, long order) {
  memset(M, 0, sizeof *M);
  M->C[4] = length;
  M->R[0][0] = M->R[1][1] = M->R[2][2] = M->R[3][3] = M->R[4][4] = M->R[5][5] = 1;
  M->R[0][1] = M->R[2][3] = length;
  if (order > 1)
    M->T[4][1][1] = M->T[4][3][3] = length / 2;
}

static void or_track_particles2(double **part, const or_vmatrix2 *M, long n_part) {
  double temp[6], sum, sum1;
  long i, j, k, i_part;
  OR_COUNT_MATRIX_BEGIN();
  for (i_part = 0; i_part < n_part; i_part++) {
    double *fin = part[i_part], *ini = part[i_part];
    for (i = 5; i >= 0; i--) {
      sum = M->C[i];
      for (j = 5; j >= 0; j--) {
        sum1 = M->R[i][j];
        for (k = j; k >= 0; k--)
          sum1 += M->T[i][j][k] * ini[k];
        sum += sum1 * ini[j];
      }
      temp[i] = sum;
    }
    for (i = 5; i >= 0; i--)
      fin[i] = temp[i];
  }
  OR_COUNT_MATRIX_END();
}

/* track_particles for a packed first- or second-order map (EH_VMATRIX), rows
 * from the top down as the reference loops (src/matrix.c:334-381 order 2,
 * TURBO_FASTMATTRACK=3; :424-460 order 1) */
long or_track_particles_flat(double *flat, long n_part, const EH_VMATRIX *M) {
  double temp[6], sum, sum1;
  long i, j, k, i_part;
  if (M->order < 1 || M->order > 2)
    return or_fail("track_particles: only orders 1 and 2 are supported");
  OR_COUNT_MATRIX_BEGIN();
  for (i_part = n_part - 1; i_part >= 0; i_part--) {
    double *ini = flat + i_part * EH_NPROP;
    for (i = 5; i >= 0; i--) {
      sum = M->C[i];
      if (M->order == 2) {
        for (j = 5; j >= 0; j--) {
          sum1 = M->R[i][j];
          for (k = j; k >= 0; k--)
            sum1 += M->T[i][j * (j + 1) / 2 + k] * ini[k];
          sum += sum1 * ini[j];
        }
      } else {
        for (j = 5; j >= 0; j--)
          sum += M->R[i][j] * ini[j];
      }
      temp[i] = sum;
    }
    for (i = 5; i >= 0; i--)
      ini[i] = temp[i];
  }
  OR_COUNT_MATRIX_END();
  return n_part;
}

/* drift_beam, src/do_tracking.c:3754-3767; also the DRIF element path
 * (matrix tracking, src/do_tracking.c:1239-1311) with order 0 -> default_order 2 */
void or_drift_beam(double **part, long np, double length, long order) {
  or_vmatrix2 M;
  if (order <= 0)
    order = 2; /* run_setup default_order, src/elegant.nl:43 */
  or_drift_matrix(&M, length, order > 2 ? 2 : order);
  or_track_particles2(part, &M, np);
}

/* ------------------------------------------------------------------------- */
/* misalignment method 0: src/malign_mat.c:257-267, :565-588; tilt_matrices.c:156-187 */
static void offsetBeamCoordinatesForMisalignment(double **coord, long np, double dx, double dy, double dz) {
  long ip;
  double *part;
  for (ip = np - 1; ip >= 0; ip--) {
    part = coord[ip];
    part[4] += dz * sqrt(1 + sqr(part[1]) + sqr(part[3]));
    part[0] = part[0] - dx + dz * part[1];
    part[2] = part[2] - dy + dz * part[3];
  }
}

static void rotateBeamCoordinatesForMisalignment(double **part, long np, double angle) {
  double x, xp, y, yp, *coord, sin_a, cos_a;
  long i;
  if (!angle || fabs(fabs(angle) - PIx2) < 1e-12)
    return;
  if (fabs(fabs(angle) - PI) < 1e-12) {
    cos_a = -1;
    sin_a = 0;
  } else if (fabs(angle - PIo2) < 1e-12) {
    cos_a = 0;
    sin_a = 1;
  } else if (fabs(angle + PIo2) < 1e-12) {
    cos_a = 0;
    sin_a = -1;
  } else {
    cos_a = OR_COS(angle);
    sin_a = OR_SIN(angle);
  }
  for (i = 0; i < np; i++) {
    coord = part[i];
    x = coord[0];
    xp = coord[1];
    y = coord[2];
    yp = coord[3];
    coord[0] = x * cos_a + y * sin_a;
    coord[2] = -x * sin_a + y * cos_a;
    coord[1] = xp * cos_a + yp * sin_a;
    coord[3] = -xp * sin_a + yp * cos_a;
  }
}

/* ------------------------------------------------------------------------- */
/* aperture data for in-element checks: setupMultApertureData / checkMultAperture,
 * src/multipole.c:1690-1798, MAXAMP without open sides, contours or files */
typedef struct {
  double xMax, yMax, reverseTilt, cs0, cs1;
  long elliptical, xExponent, yExponent;
} or_apdata;

static void or_setup_ap(or_apdata *ap, double reverseTilt, const EH_MAXAMP *maxamp) {
  ap->reverseTilt = reverseTilt;
  ap->cs0 = OR_COS(reverseTilt);
  ap->cs1 = OR_SIN(reverseTilt);
  ap->xMax = ap->yMax = 0;
  ap->elliptical = ap->xExponent = ap->yExponent = 0;
  if (maxamp) {
    ap->xMax = maxamp->x_max;
    ap->yMax = maxamp->y_max;
    ap->elliptical = maxamp->elliptical;
    ap->xExponent = ap->yExponent = maxamp->exponent;
    if (maxamp->y_exponent)
      ap->yExponent = maxamp->y_exponent;
  }
}

static long or_check_ap(double xInput, double yInput, const or_apdata *ap) {
  double x = xInput - 0.0, y = yInput - 0.0; /* xCen = yCen = 0 */
  if (ap->reverseTilt) {
    double x0 = x, y0 = y;
    x = x0 * ap->cs0 + y0 * ap->cs1;
    y = -x0 * ap->cs1 + y0 * ap->cs0;
  }
  if (ap->elliptical == 0 || ap->xMax <= 0 || ap->yMax <= 0) {
    if ((ap->xMax > 0 && fabs(x) > ap->xMax) || (ap->yMax > 0 && fabs(y) > ap->yMax))
      return 0;
  } else {
    double xa = x / ap->xMax, yb = y / ap->yMax;
    if ((or_ipow(xa, ap->xExponent) + or_ipow(yb, ap->yExponent)) >= 1)
      return 0;
  }
  return 1;
}

/* ------------------------------------------------------------------------- */
/* multipole kicks: expansion_coefficients (src/multipole.h:76-145) and
 * apply_canonical_multipole_kicks_ret/_noret (src/multipole.h:391-445) */
static double dblfactorial(long n) {
  double prod = 1;
  while (n > 1)
    prod *= n--;
  return prod;
}

static void or_coefs(long n, double *coef) {
  long i;
  if (n == 0) {
    coef[0] = 1;
  } else if (n == 1) {
    coef[0] = 1;
    coef[1] = 1;
  } else if (n == 2) {
    coef[0] = 1. / 2;
    coef[1] = 1;
    coef[2] = -1. / 2;
  } else {
    for (i = 0; i <= n; i++)
      coef[i] = (((i / 2) & 1) ? -1.0 : 1.0) / (dblfactorial(i) * dblfactorial(n - i));
  }
}

static void or_kick_ret(double *qx, double *qy, double *dqx, double *dqy, const double *xpow, const double *ypow,
                        long order, double KnL, long skew);

static void or_kick(double *qx, double *qy, const double *xpow, const double *ypow, long order, double KnL, long skew) {
  or_kick_ret(qx, qy, NULL, NULL, xpow, ypow, order, KnL, skew);
}

/* the kick with the increments accumulated into dqx, dqy when given
 * (apply_canonical_multipole_kicks_ret, src/multipole.h:391-420) */
static void or_kick_ret(double *qx, double *qy, double *dqx, double *dqy, const double *xpow, const double *ypow,
                        long order, double KnL, long skew) {
  double coef[16], sum_Fx = 0, sum_Fy = 0, tmp;
  long i;
  or_coefs(order, coef);
  for (i = 0; i <= order; i++) {
    double f = coef[i] * xpow[order - i] * ypow[i];
    if (i & 1)
      sum_Fx += f;
    else
      sum_Fy += f;
  }
  if (skew) {
    tmp = sum_Fx;
    sum_Fx = sum_Fy;
    sum_Fy = tmp;
    sum_Fx = -sum_Fx;
  }
  *qx -= KnL * sum_Fy;
  *qy += KnL * sum_Fx;
  if (dqx) {
    *dqx -= KnL * sum_Fy;
    *dqy += KnL * sum_Fx;
  }
}

static void fillPowerArray(double x, double *xpow, long order) {
  long i;
  xpow[0] = 1;
  for (i = 1; i <= order; i++)
    xpow[i] = xpow[i - 1] * x;
}

/* convertSlopesToMomenta / convertMomentaToSlopes, src/multipole.h:448-486 */
static int slopes_to_momenta(double *qx, double *qy, double xp, double yp, double delta, int expand) {
  if (expand) {
    *qx = (1 + delta) * xp;
    *qy = (1 + delta) * yp;
  } else {
    double denom = sqrt(1 + sqr(xp) + sqr(yp));
    *qx = (1 + delta) * xp / denom;
    *qy = (1 + delta) * yp / denom;
  }
  return 1;
}

static int momenta_to_slopes(double *xp, double *yp, double qx, double qy, double delta, int expand) {
  if (expand) {
    *xp = qx / (1 + delta);
    *yp = qy / (1 + delta);
  } else {
    double denom = sqr(1 + delta) - sqr(qx) - sqr(qy);
    if (denom <= 0)
      return 0;
    denom = sqrt(denom);
    *xp = qx / denom;
    *yp = qy / denom;
  }
  return 1;
}

#define BETA 1.25992104989487316477
static const double driftFrac2[2] = {0.5, 0.5};
static const double kickFrac2[2] = {1.0, 0.0};
static const double driftFrac4[4] = {0.5 / (2 - BETA), (1 - BETA) / (2 - BETA) / 2, (1 - BETA) / (2 - BETA) / 2,
                                     0.5 / (2 - BETA)};
static const double kickFrac4[4] = {1. / (2 - BETA), -BETA / (2 - BETA), 1 / (2 - BETA), 0};
static const double driftFrac6[8] = {0.39225680523878,      0.5100434119184585, -0.47105338540975655,
                                     0.0687531682525181,    0.0687531682525181, -0.47105338540975655,
                                     0.5100434119184585,    0.39225680523878};
static const double kickFrac6[8] = {0.784513610477560,  0.235573213359357,  -1.17767998417887, 1.3151863206839063,
                                    -1.17767998417887, 0.235573213359357, 0.784513610477560,  0};

static int or_integrator_tables(long order, const double **drift, const double **kick, long *nsub) {
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

/* integrate_kick_multipole_ordn, src/multipole.c:1099-1456, for i_part = -1, no
 * radiation, no error/edge/steering multipoles, non-radial */
static int or_integrate_kick_multipole(double *coord, double dx, double dy, double xkick, double ykick, double Po,
                                       double rad_coef,
                                       const long *order, const double *KnL, const short *skew, long n_parts,
                                       double drift, long integration_order, const or_apdata *apData, double *dzLoss,
                                       int expand) {
  double qx, qy, s, dp, p, beta0, beta1, dqx = 0, dqy = 0;
  double x, y, xp, yp, xpow[16], ypow[16], KnLp[3];
  long i_kick, step, iOrder, maxOrder, nSubsteps, i;
  double dsh;
  const double *driftFrac, *kickFrac;

  or_integrator_tables(integration_order, &driftFrac, &kickFrac, &nSubsteps);
  drift = drift / n_parts;
  xkick = xkick / n_parts;
  ykick = ykick / n_parts;

  x = coord[0];
  xp = coord[1];
  y = coord[2];
  yp = coord[3];
  s = 0;
  dp = coord[5];
  p = Po * (1 + dp);
  beta0 = p / sqrt(sqr(p) + 1);

  if (isnan(x) || isnan(xp) || isnan(y) || isnan(yp))
    return 0;
  if (fabs(x) > COORD_LIMIT || fabs(y) > COORD_LIMIT || fabs(xp) > SLOPE_LIMIT || fabs(yp) > SLOPE_LIMIT)
    return 0;

  slopes_to_momenta(&qx, &qy, xp, yp, dp, expand);
  maxOrder = order[0] > (order[1] > order[2] ? order[1] : order[2]) ? order[0] : (order[1] > order[2] ? order[1] : order[2]);
  for (i = 0; i < 3; i++)
    KnLp[i] = KnL[i] / n_parts;

  if (!momenta_to_slopes(&xp, &yp, qx, qy, dp, expand))
    return 0;

  *dzLoss = 0;
  for (i_kick = 0; i_kick < n_parts; i_kick++) {
    if (!or_check_ap(x + dx, y + dy, apData)) {
      coord[0] = x;
      coord[2] = y;
      return 0;
    }
    for (step = 0; step < nSubsteps; step++) {
      if (drift) {
        dsh = drift * driftFrac[step];
        x += xp * dsh;
        y += yp * dsh;
        if (expand)
          s += dsh * (1 + (sqr(xp) + sqr(yp)) / 2);
        else
          s += dsh * sqrt(1 + sqr(xp) + sqr(yp));
        *dzLoss += dsh;
      }
      if (!kickFrac[step])
        break;
      fillPowerArray(x, xpow, maxOrder);
      fillPowerArray(y, ypow, maxOrder);
      dqx = dqy = 0;
      for (iOrder = 0; iOrder < 3; iOrder++)
        if (KnL[iOrder])
          or_kick_ret(&qx, &qy, &dqx, &dqy, xpow, ypow, order[iOrder], KnLp[iOrder] * kickFrac[step], skew[iOrder]);
      if (xkick)
        or_kick(&qx, &qy, xpow, ypow, 0, -xkick * kickFrac[step], 0);
      if (ykick)
        or_kick(&qx, &qy, xpow, ypow, 0, -ykick * kickFrac[step], 1);
      if (!momenta_to_slopes(&xp, &yp, qx, qy, dp, expand))
        return 0;
      /* the radiation block runs whenever the element has length
       * (src/multipole.c:1360-1390): isr_coef is non-zero even without ISR
       * (negative, :757-760); without radiation it leaves qx, qy divided by
       * and multiplied by (1 + dp), and converts them to slopes again */
      if (drift) {
        double deltaFactor, F2, dsFactor;
        qx /= (1 + dp);
        qy /= (1 + dp);
        deltaFactor = sqr(1 + dp);
        /* the kick increments carry kickFrac[step]; remove it */
        dqx /= kickFrac[step];
        dqy /= kickFrac[step];
        F2 = sqr(dqx / drift - xkick / drift) + sqr(dqy / drift + ykick / drift);
        dqx = 0;
        dqy = 0;
        dsFactor = sqrt(1 + sqr(xp) + sqr(yp));
        dsFactor *= drift * kickFrac[step];
        if (rad_coef)
          dp -= rad_coef * deltaFactor * F2 * dsFactor;
        qx *= (1 + dp);
        qy *= (1 + dp);
        if (!momenta_to_slopes(&xp, &yp, qx, qy, dp, expand))
          return 0;
      }
    }
  }
  if (!or_check_ap(x + dx, y + dy, apData)) {
    coord[0] = x;
    coord[2] = y;
    return 0;
  }
  if (!momenta_to_slopes(&xp, &yp, qx, qy, dp, expand))
    return 0;
  coord[0] = x;
  coord[1] = xp;
  coord[2] = y;
  coord[3] = yp;
  if (rad_coef) {
    p = Po * (1 + dp);
    beta1 = p / sqrt(sqr(p) + 1);
    coord[4] = beta1 * (coord[4] / beta0 + 2 * s / (beta0 + beta1));
  } else
    coord[4] += s;
  coord[5] = dp;
  if (isnan(x) || isnan(xp) || isnan(y) || isnan(yp))
    return 0;
  if (fabs(x) > COORD_LIMIT || fabs(y) > COORD_LIMIT || fabs(xp) > SLOPE_LIMIT || fabs(yp) > SLOPE_LIMIT)
    return 0;
  return 1;
}

/* multipole_tracking2, src/multipole.c:660-1084 (KQUAD/KSEXT/KOCT, iSlice = -1,
 * malignMethod 0, no bore/B, no radiation, no multipole files, no fringe) */
long or_multipole_tracking2(double **particle, long n_part, const EH_ELEMENT_LIST *elem, double Po, double z_start,
                            const EH_MAXAMP *maxamp) {
  double KnL[3] = {0, 0, 0};
  long order[3] = {0, 0, 0};
  short skew[3] = {0, 0, 0};
  double dx = 0, dy = 0, dz = 0, tilt = 0, xkick = 0, ykick = 0, drift = 0, dzLoss = 0, rad_coef = 0;
  long nSlices = 0, n_kicks = 0, integ_order = 0, i_part, i_top, synch_rad = 0;
  int expand = 0;
  or_apdata ap;

  switch (elem->type) {
  case EH_T_KQUAD: {
    const EH_KQUAD *q = (const EH_KQUAD *)elem->p_elem;
    nSlices = q->nSlices;
    n_kicks = q->n_kicks;
    expand = q->expandHamiltonian;
    order[0] = 1;
    KnL[0] = q->bore ? q->B / q->bore * (EH_E_MKS / (EH_ME_MKS * c_mks * Po)) * q->length * (1 + q->fse)
                     : q->k1 * q->length * (1 + q->fse);
    drift = q->length;
    tilt = q->tilt;
    dx = q->dx;
    dy = q->dy;
    dz = q->dz;
    xkick = q->xkick * q->xKickCalibration;
    ykick = q->ykick * q->yKickCalibration;
    integ_order = q->integration_order;
    synch_rad = q->synch_rad;
    break;
  }
  case EH_T_KSEXT: {
    const EH_KSEXT *q = (const EH_KSEXT *)elem->p_elem;
    n_kicks = q->n_kicks;
    nSlices = q->nSlices;
    expand = q->expandHamiltonian;
    order[0] = 2;
    KnL[0] = q->bore ? 2 * q->B / sqr(q->bore) * (EH_E_MKS / (EH_ME_MKS * c_mks * Po)) * q->length * (1 + q->fse)
                     : q->k2 * q->length * (1 + q->fse);
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
    synch_rad = q->synch_rad;
    break;
  }
  case EH_T_KOCT: {
    const EH_KOCT *q = (const EH_KOCT *)elem->p_elem;
    n_kicks = q->n_kicks;
    nSlices = q->nSlices;
    expand = q->expandHamiltonian;
    order[0] = 3;
    KnL[0] = q->bore ? 6 * q->B / (q->bore * q->bore * q->bore) * (EH_E_MKS / (EH_ME_MKS * c_mks * Po)) * q->length * (1 + q->fse)
                     : q->k3 * q->length * (1 + q->fse);
    drift = q->length;
    tilt = q->tilt;
    dx = q->dx;
    dy = q->dy;
    dz = q->dz;
    integ_order = q->integration_order;
    synch_rad = q->synch_rad;
    break;
  }
  case EH_T_KQUSE: {
    /* src/multipole.c:923-949: quadrupole with the sextupole as second multipole */
    const EH_KQUSE *q = (const EH_KQUSE *)elem->p_elem;
    n_kicks = q->n_kicks;
    nSlices = q->nSlices;
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
    synch_rad = q->synch_rad;
    break;
  }
  default:
    return or_fail("multipole_tracking2: element type not supported");
  }
  if (integ_order != 2 && integ_order != 4 && integ_order != 6)
    return or_fail("multipole integration_order must be 2, 4, or 6");
  /* classical radiation, src/multipole.c:755-763 (and :827-836, :888-897,
   * :938-947): none for elements shorter than 1e-6 m */
  if (synch_rad && drift >= 1e-6)
    rad_coef = sqr(EH_E_MKS) * pow3(Po) / (6 * PI * EH_EPSILON_O * sqr(c_mks) * EH_ME_MKS);

  i_top = n_part - 1;
  if (n_kicks <= 0) {
    if (nSlices <= 0)
      return or_fail("N_KICKS<=0 and N_SLICES<=0 in multipole tracking");
  } else {
    if (integ_order > 2) {
      if ((nSlices = (long)ceil(n_kicks / (1.0 * integ_order))) < 1)
        nSlices = 1;
      n_kicks = nSlices * integ_order;
    } else
      nSlices = n_kicks;
  }

  or_setup_ap(&ap, -tilt, maxamp);

  /* offsetParticlesForMisalignment(0, ..., face 1), src/malign_mat.c:580-585 */
  if (dx || dy || dz)
    offsetBeamCoordinatesForMisalignment(particle, n_part, dx, dy, dz);
  if (tilt)
    rotateBeamCoordinatesForMisalignment(particle, n_part, tilt + 0.0);

  for (i_part = 0; i_part <= i_top; i_part++) {
    if (!or_integrate_kick_multipole(particle[i_part], dx, dy, xkick, ykick, Po, rad_coef, order, KnL, skew, nSlices,
                                     drift,
                                     integ_order, &ap, &dzLoss, expand)) {
      swapParticles(particle[i_part], particle[i_top]);
      particle[i_top][4] = z_start + dzLoss;
      particle[i_top][5] = Po * (1 + particle[i_top][5]);
      i_top--;
      i_part--;
      continue;
    }
  }

  /* face 2, src/multipole.c:1069-1070, src/malign_mat.c:586-590: all n_part rows */
  if (dx || dy || dz || tilt) {
    if (tilt)
      rotateBeamCoordinatesForMisalignment(particle, n_part, -(tilt + 0.0));
    if (dx || dy || dz)
      offsetBeamCoordinatesForMisalignment(particle, n_part, -dx, -dy, -dz);
  }
  return i_top + 1;
}

/* ------------------------------------------------------------------------- */
/* MULT: multipole_tracking, src/multipole.c:439-658, without BORE/BTIPL and
 * without radiation: N_SLICES thin kicks KnL*FACTOR/N_SLICES, each between
 * drifts of L/N_SLICES (half drifts at the ends); the NaN and coordinate/slope
 * limit losses record coord[4] = z_start; the entry offset/rotation act on the
 * rows passed in and the exit ones on all of them, lost rows included. */
long or_multipole_tracking(double **particle, long n_part, const EH_MULT *multipole, double Po, double z_start) {
  long order = multipole->order, nSlices = multipole->nSlices, i_part, i_kick, i_top, is_lost;
  double KnL, drift, qx, qy, x, xp, y, yp, s, dp, xpow[16], ypow[16], rad_coef, p, beta0, beta1;
  const long expand = multipole->expandHamiltonian;
  if (nSlices <= 0)
    return or_fail("N_SLICES (or N_KICKS) <=0  in multipole_tracking()");
  if (order < 0)
    return or_fail("order < 0 in multipole_tracking()");
  if (order > 8)
    return or_fail("MULT order above 8 not supported");
  if (multipole->bore)
    return or_fail("MULT with BORE/BTIPL not supported");
  drift = multipole->length / nSlices / 2;
  KnL = multipole->KnL * multipole->factor / nSlices;
  if (KnL == 0) {
    if (multipole->length == 0)
      return n_part;
    or_exactDrift(particle, n_part, multipole->length);
    return n_part;
  }
  if (multipole->synch_rad)
    rad_coef = sqr(EH_E_MKS) * pow3(Po) / (6 * PI * EH_EPSILON_O * sqr(c_mks) * EH_ME_MKS);
  else
    rad_coef = 0;
  if (multipole->dx || multipole->dy || multipole->dz)
    offsetBeamCoordinatesForMisalignment(particle, n_part, multipole->dx, multipole->dy, multipole->dz);
  if (multipole->tilt)
    rotateBeamCoordinatesForMisalignment(particle, n_part, multipole->tilt);
  i_top = n_part - 1;
  for (i_part = 0; i_part <= i_top; i_part++) {
    double *coord = particle[i_part];
    x = coord[0];
    xp = coord[1];
    y = coord[2];
    yp = coord[3];
    s = 0;
    dp = coord[5];
    p = Po * (1 + dp);
    beta0 = p / sqrt(sqr(p) + 1);
    if (isnan(x) || isnan(xp) || isnan(y) || isnan(yp) || fabs(x) > 10.0 || fabs(y) > 10.0 || fabs(xp) > 1.0 ||
        fabs(yp) > 1.0) {
      swapParticles(particle[i_part], particle[i_top]);
      particle[i_top][4] = z_start;
      particle[i_top][5] = Po * (1 + particle[i_top][5]);
      i_top--;
      i_part--;
      continue;
    }
    slopes_to_momenta(&qx, &qy, xp, yp, dp, expand);
    is_lost = 0;
    for (i_kick = 0; i_kick < nSlices; i_kick++) {
      if (drift) {
        x += xp * drift * (i_kick ? 2 : 1);
        y += yp * drift * (i_kick ? 2 : 1);
        if (expand)
          s += (i_kick ? 2 : 1) * drift * (1 + (sqr(xp) + sqr(yp)) / 2);
        else
          s += (i_kick ? 2 : 1) * drift * sqrt(1 + sqr(xp) + sqr(yp));
      }
      fillPowerArray(x, xpow, order);
      fillPowerArray(y, ypow, order);
      or_kick(&qx, &qy, xpow, ypow, order, KnL, 0);
      if (!momenta_to_slopes(&xp, &yp, qx, qy, dp, expand)) {
        is_lost = 1;
        break;
      }
      if (rad_coef && drift) {
        /* src/multipole.c:599-611, with the kick's field sums */
        double coef[16], sum_Fx = 0, sum_Fy = 0;
        long i;
        or_coefs(order, coef);
        for (i = 0; i <= order; i++) {
          if (i & 1)
            sum_Fx += coef[i] * xpow[order - i] * ypow[i];
          else
            sum_Fy += coef[i] * xpow[order - i] * ypow[i];
        }
        qx /= (1 + dp);
        qy /= (1 + dp);
        dp -= rad_coef * sqr(KnL * (1 + dp)) * (sqr(sum_Fy) + sqr(sum_Fx)) * sqrt(1 + sqr(xp) + sqr(yp)) /
              (2 * drift);
        qx *= (1 + dp);
        qy *= (1 + dp);
        if (!momenta_to_slopes(&xp, &yp, qx, qy, dp, expand)) {
          is_lost = 1;
          break;
        }
      }
    }
    if (drift && !is_lost) {
      x += xp * drift;
      y += yp * drift;
      if (expand)
        s += drift * (1 + (sqr(xp) + sqr(yp)) / 2);
      else
        s += drift * sqrt(1 + sqr(xp) + sqr(yp));
    }
    coord[0] = x;
    coord[1] = xp;
    coord[2] = y;
    coord[3] = yp;
    coord[5] = dp;
    if (rad_coef) {
      p = Po * (1 + dp);
      beta1 = p / sqrt(sqr(p) + 1);
      coord[4] = beta1 * (coord[4] / beta0 + 2 * s / (beta0 + beta1));
    } else
      coord[4] += s;
    if (isnan(x) || isnan(xp) || isnan(y) || isnan(yp) || fabs(x) > 10.0 || fabs(y) > 10.0 || fabs(xp) > 1.0 ||
        fabs(yp) > 1.0 || is_lost) {
      swapParticles(particle[i_part], particle[i_top]);
      particle[i_top][4] = z_start;
      particle[i_top][5] = Po * (1 + particle[i_top][5]);
      i_top--;
      i_part--;
      continue;
    }
  }
  if (multipole->tilt)
    rotateBeamCoordinatesForMisalignment(particle, n_part, -multipole->tilt);
  if (multipole->dx || multipole->dy || multipole->dz)
    offsetBeamCoordinatesForMisalignment(particle, n_part, -multipole->dx, -multipole->dy, -multipole->dz);
  return i_top + 1;
}

/* ROTATE: rotation_matrix(tilt) (src/tilt_matrices.c:81-117) applied by the
 * order-1 loop of track_particles (src/matrix.c:424-460): every output is
 * C[i] plus R[i][j]*ini[j] for j = 5 down to 0, zero entries included */
void or_rotate(double **coord, long np, const EH_ROTATE *rot) {
  double R[6][6], sin_tilt, cos_tilt, temp[6];
  long i, j, ip;
  const double tilt = rot->tilt;
  if (fabs(tilt - PI / 2) < 1e-12) {
    sin_tilt = 1;
    cos_tilt = 0;
  } else if (fabs(tilt - PI) < 1e-12) {
    sin_tilt = 0;
    cos_tilt = -1;
  } else {
    sin_tilt = OR_SIN(tilt);
    cos_tilt = OR_COS(tilt);
  }
  memset(R, 0, sizeof R);
  R[0][0] = cos_tilt;
  R[0][2] = sin_tilt;
  R[2][0] = -sin_tilt;
  R[2][2] = cos_tilt;
  R[1][1] = cos_tilt;
  R[1][3] = sin_tilt;
  R[3][1] = -sin_tilt;
  R[3][3] = cos_tilt;
  R[4][4] = R[5][5] = 1;
  OR_COUNT_MATRIX_BEGIN();
  for (ip = np - 1; ip >= 0; ip--) {
    double *ini = coord[ip];
    for (i = 5; i >= 0; i--) {
      double sum = 0;
      for (j = 5; j >= 0; j--)
        sum += R[i][j] * ini[j];
      temp[i] = sum;
    }
    for (i = 5; i >= 0; i--)
      ini[i] = temp[i];
  }
  OR_COUNT_MATRIX_END();
}

/* set_central_momentum, src/do_tracking.c:3563-3596 (no MPI) */
void or_set_central_momentum(double **coord, long np, double P_new, double *P_central) {
  long ip;
  if (!np) {
    *P_central = P_new;
    return;
  }
  if (*P_central != P_new) {
    for (ip = 0; ip < np; ip++)
      coord[ip][5] = ((1 + coord[ip][5]) * (*P_central) - P_new) / P_new;
    *P_central = P_new;
  }
}

long or_do_match_energy(double **coord, long np, double *P_central, long change_beam);

/* ENERGY, src/do_tracking.c:1861-1881 (no fiducial beam flags: MATCH_BEAMLINE
 * always matches the reference momentum to the beam) */
long or_energy(double **coord, long np, const EH_ENERGY *energy, double *P_central) {
  if (energy->match_beamline) {
    long r = or_do_match_energy(coord, np, P_central, 0);
    if (r < 0)
      return r;
    if (energy->match_particles)
      return or_fail("can't match_beamline AND match_particles for ENERGY element");
  } else if (energy->match_particles) {
    long r = or_do_match_energy(coord, np, P_central, 1);
    if (r < 0)
      return r;
  } else if (energy->central_energy)
    or_set_central_momentum(coord, np, sqrt(sqr(energy->central_energy + 1) - 1), P_central);
  else if (energy->central_momentum)
    or_set_central_momentum(coord, np, energy->central_momentum, P_central);
  return np;
}

/* MALIGN: offset_beam, src/do_tracking.c:3311-3366, every particle */
long or_offset_beam(double **coord, long nToTrack, const EH_MALIGN *offset, double P_central) {
  long i_part;
  double *part, pc, beta, gamma, t, ds;
  if (offset->startPID != -1 || offset->endPID != -1)
    return or_fail("MALIGN particle ranges not supported");
  for (i_part = nToTrack - 1; i_part >= 0; i_part--) {
    part = coord[i_part];
    if (offset->dz)
      ds = offset->dz * sqrt(1 + sqr(part[1]) + sqr(part[3]));
    else
      ds = 0;
    part[0] += offset->dx + offset->dz * part[1];
    part[1] += offset->dxp;
    part[2] += offset->dy + offset->dz * part[3];
    part[3] += offset->dyp;
    part[4] += ds;
    if (offset->dt || offset->dp || offset->de) {
      pc = P_central * (1 + part[5]);
      beta = pc / (gamma = sqrt(1 + pc * pc));
      t = part[4] / (beta * c_mks) + offset->dt;
      if (offset->dp) {
        part[5] += offset->dp;
        pc = P_central * (1 + part[5]);
        beta = pc / sqrt(1 + pc * pc);
      }
      if (offset->de) {
        gamma += offset->de * gamma;
        pc = sqrt(gamma * gamma - 1);
        beta = pc / gamma;
        part[5] = (pc - P_central) / P_central;
      }
      part[4] = t * beta * c_mks;
    }
  }
  return nToTrack;
}

/* ------------------------------------------------------------------------- */
/* CSBEND field expansion (computeCSBENDFieldCoefficients, src/csbend.c:175-652):
 * coefficients evaluated term by term from the generated table, whose entries
 * equal the reference's formulas */
typedef struct {
  double Fx[11][11], Fy[11][11];
  long expansionOrder1;
  int hasSkew, hasNormal;
} or_csbend_fields;

static long or_field_coefficients(or_csbend_fields *F, const double *b, const double *c, double h1, long nonlinear,
                                  long expansionOrder) {
  long i, t;
  double h[20];
  if (expansionOrder == 0) {
    for (i = 8; i >= 0; i--)
      if (b[i] || c[i])
        break;
    if ((expansionOrder = i + 2) < 4)
      expansionOrder = 4;
  }
  F->expansionOrder1 = expansionOrder + 1;
  if (F->expansionOrder1 > 11)
    return or_fail("expansion order >10 for CSBEND");
  F->hasSkew = F->hasNormal = 0;
  for (i = 0; i < 9; i++) {
    if (b[i])
      F->hasNormal = 1;
    if (c[i])
      F->hasSkew = 1;
  }
  memset(F->Fx, 0, sizeof F->Fx);
  memset(F->Fy, 0, sizeof F->Fy);
  h[0] = 1;
  for (i = 1; i < 20; i++)
    h[i] = h[i - 1] * h1;
  F->Fx[0][0] = c[0];
  F->Fy[0][0] = 1 - b[0];
  for (t = 0; t < EH_CSBEND_NTERMS; t++) {
    int ii = eh_csbend_terms[t].i, jj = eh_csbend_terms[t].j;
    /* the linear block (entries of degree 1) is set regardless of NONLINEAR */
    if (!nonlinear && ii + jj > 1)
      continue;
    if (ii + jj >= 11)
      continue;
    const double *v = eh_csbend_terms[t].kind ? c + 1 : b + 1;
    double term = ((eh_csbend_terms[t].num * h[eh_csbend_terms[t].hpow]) * v[eh_csbend_terms[t].m]) /
                  eh_csbend_terms[t].den;
    double *dst = eh_csbend_terms[t].which ? &F->Fy[ii][jj] : &F->Fx[ii][jj];
    /* first term of an entry initialises it (the reference writes a single expression) */
    if (t == 0 || eh_csbend_terms[t - 1].which != eh_csbend_terms[t].which || eh_csbend_terms[t - 1].i != ii ||
        eh_csbend_terms[t - 1].j != jj)
      *dst = term;
    else
      *dst += term;
  }
  return F->expansionOrder1;
}

/* computeCSBENDFields, src/csbend.c:99-172 */
static void or_csbend_fields_eval(const or_csbend_fields *F, double *Fx, double *Fy, double x, double y) {
  double yp[11], sumFx = 0, sumFy = 0, xt = 1;
  long i, j;
  if (!F->hasSkew && !F->hasNormal) {
    *Fx = 0;
    *Fy = 1;
    return;
  }
  yp[0] = 1;
  for (i = 1; i < F->expansionOrder1; i++)
    yp[i] = yp[i - 1] * y;
  if (F->hasSkew) {
    for (i = 0; i < F->expansionOrder1; i++) {
      for (j = 0; j < F->expansionOrder1 - i; j++) {
        sumFx += F->Fx[i][j] * xt * yp[j];
        sumFy += F->Fy[i][j] * xt * yp[j];
      }
      xt *= x;
    }
  } else {
    for (i = 0; i < F->expansionOrder1; i++) {
      for (j = 0; j < F->expansionOrder1 - i; j++) {
        if (j & 1)
          sumFx += F->Fx[i][j] * xt * yp[j];
        else
          sumFy += F->Fy[i][j] * xt * yp[j];
      }
      xt *= x;
    }
  }
  *Fx = sumFx;
  *Fy = sumFy;
}

long or_csbend_field_table(const double *b, const double *c, double h, int nonlinear, int expansionOrder, double *Fx,
                           double *Fy) {
  or_csbend_fields F;
  long r = or_field_coefficients(&F, b, c, h, nonlinear, expansionOrder);
  if (r < 0)
    return r;
  memcpy(Fx, F.Fx, sizeof F.Fx);
  memcpy(Fy, F.Fy, sizeof F.Fy);
  return r;
}

/* integrate_csbend_ordn, src/csbend.c:1459-1720 (no radiation, no reference
 * trajectory correction, iSlice = -1) */
static long or_integrate_csbend(double *Qf, const double *Qi, double s, long n, double rho0, double rho_actual,
                                double *dz_lost, const or_apdata *apData, long integration_order,
                                const or_csbend_fields *F, double rad_coef) {
  long i, j, nSubsteps;
  double factor, f, phi, ds, dsh, dist, Fx, Fy, x, y, sine, cosi, tang, sin_phi, cos_phi;
  const double *driftFrac, *kickFrac;
  or_integrator_tables(integration_order, &driftFrac, &kickFrac, &nSubsteps);
#define X Qf[0]
#define QX Qf[1]
#define Y Qf[2]
#define QY Qf[3]
#define DPoP Qf[5]
  memcpy(Qf, Qi, sizeof(*Qi) * 6);
  dist = 0;
  s /= n;
  *dz_lost = 0;
  for (i = 0; i < n; i++) {
    if (!or_check_ap(X, Y, apData))
      return 0;
    for (j = 0; j < nSubsteps; j++) {
      dsh = s * driftFrac[j];
      f = sqr(1 + DPoP) - sqr(QY);
      if (f <= 0)
        return 0;
      f = sqrt(f);
      if (fabs(QX / f) > 1)
        return 0;
      sin_phi = QX / f;
      phi = OR_ASIN(sin_phi);
      sine = OR_SIN(dsh / rho0 + phi);
      cosi = OR_COS(dsh / rho0 + phi);
      if (cosi == 0)
        return 0;
      tang = sine / cosi;
      cos_phi = OR_COS(phi);
      QX = f * sine;
      factor = (rho0 + X) * cos_phi / f * (tang - sin_phi / cos_phi);
      Y += QY * factor;
      dist += factor * (1 + DPoP);
      *dz_lost += dsh;
      f = cos_phi / cosi;
      X = rho0 * (f - 1) + f * X;
      if (kickFrac[j] == 0)
        break;
      ds = s * kickFrac[j];
      x = X;
      y = Y;
      or_csbend_fields_eval(F, &Fx, &Fy, x, y);
      QX += -ds * (1 + X / rho0) * Fy / rho_actual;
      QY += ds * (1 + X / rho0) * Fx / rho_actual;
      /* the radiation block runs for every bend (src/csbend.c:1627-1660):
       * isrConstant is non-zero even without ISR (negative, :961-964); without
       * radiation it leaves QX, QY divided by and multiplied by (1 + DPoP) */
      {
        double fr = (1 + X * (1. / rho0)) / sqrt(sqr(1 + DPoP) - sqr(QX) - sqr(QY));
        double xpr = QX * fr, ypr = QY * fr;
        double dsFactor = sqrt(sqr(1 + X * (1. / rho0)) + sqr(xpr) + sqr(ypr));
        double F2 = sqr(Fx) + sqr(Fy);
        double deltaFactor = sqr(1 + DPoP);
        QX /= (1 + DPoP);
        QY /= (1 + DPoP);
        if (rad_coef)
          DPoP -= rad_coef * deltaFactor * F2 * ds * dsFactor;
        QX *= (1 + DPoP);
        QY *= (1 + DPoP);
      }
    }
  }
  *dz_lost = n * s;
  if (!or_check_ap(X, Y, apData))
    return 0;
  Qf[4] += dist;
  return 1;
#undef X
#undef QX
#undef Y
#undef QY
#undef DPoP
}

/* convertToDipoleCanonicalCoordinates / convertFrom..., src/csbend.c:1436-1454 */
static void toDipoleCanonical(double *Qi) {
  double f = (1 + Qi[5]) / sqrt(1 + sqr(Qi[1]) + sqr(Qi[3]));
  Qi[1] *= f;
  Qi[3] *= f;
}

static void fromDipoleCanonical(double *Qi) {
  double f = 1 / sqrt(sqr(1 + Qi[5]) - sqr(Qi[1]) - sqr(Qi[3]));
  Qi[1] *= f;
  Qi[3] *= f;
}

/* dipoleFringeKHwang, src/csbend.c:4336-4410 (EDGE_EFFECTS 2): K. Hwang's
 * dipole fringe map on the canonical coordinates Q, at the entrance
 * (in_fringe -1) or the exit (+1), with the terms of higher order in the
 * transverse coordinates for EDGE_ORDER > 1.  The SDDS toolkit's ipow1/2/3 (not
 * vendored) are x, x*x and (x*x)*x here. */
#define OR_P2(v) ((v) * (v))
#define OR_P3(v) ((v) * (v) * (v))
static void or_fringe_hwang(double *Q, double rho, double in_fringe, long edge_order, double edge, double gap,
                            double fint, double Rhe) {
  const double k0 = sqr(PI) / 6., k2 = fint, k3 = 1.0 * 1. / 6., k4 = -1.0 * sqr(PI) / 3., k5 = 0.0, k6 = -1.0;
  const double x0 = Q[0], px0 = Q[1], y0 = Q[2], py0 = Q[3], dp0 = Q[5];
  const double sec_e = 1. / OR_COS(edge), tan_e = OR_TAN(edge), sin_e = OR_SIN(edge), cos_e = OR_COS(edge);
  double dx = 0, dpx = 0, dy = 0, dpy = 0;
  if (edge_order > 1) {
    if (in_fringe == -1.) {
      dx = in_fringe * OR_P2(sec_e) * OR_P2(gap) * k0 / rho / (1 + dp0) +
           in_fringe * OR_P2(x0) * OR_P2(tan_e) / 2 / rho / (1 + dp0) -
           in_fringe * OR_P2(y0) * OR_P2(sec_e) / 2 / rho / (1 + dp0);
      dy = -in_fringe * x0 * y0 * OR_P2(tan_e) / rho / (1 + dp0);
      dpx = -1. * OR_P3(sec_e) * sin_e * OR_P2(gap) * k0 / rho / rho / (1 + dp0) + tan_e * x0 / rho +
            OR_P2(y0) / 2 * (2 * OR_P3(tan_e)) / OR_P2(rho) / (1 + dp0) +
            OR_P2(y0) / 2 * (tan_e) / OR_P2(rho) / (1 + dp0) -
            in_fringe * (x0 * px0 - y0 * py0) * OR_P2(tan_e) / rho / (1 + dp0) +
            k4 * OR_P2(sin_e) * OR_P2(gap) / 2 / OR_P3(cos_e) / rho * Rhe -
            k5 * x0 * sin_e * gap / OR_P3(cos_e) / rho * Rhe +
            k6 * (y0 * y0 - x0 * x0) / 2 / OR_P3(cos_e) / rho * Rhe;
      dpy = -1. * tan_e * y0 / rho + k2 * y0 * (1 + OR_P2(sin_e)) * gap / (1 + dp0) / OR_P2(rho) / OR_P3(cos_e) +
            in_fringe * (x0 * py0 + y0 * px0) * OR_P2(tan_e) / rho / (1 + dp0) +
            in_fringe * y0 * px0 / rho / (1 + dp0) +
            k3 * OR_P3(y0) * (2. / 3. / cos_e - 4. / 3. / OR_P3(cos_e)) / (1 + dp0) / rho / rho / gap +
            k6 * x0 * y0 / OR_P3(cos_e) / rho * Rhe;
    }
    if (in_fringe == 1.) {
      dx = in_fringe * OR_P2(sec_e) * OR_P2(gap) * k0 / rho / (1 + dp0) +
           in_fringe * OR_P2(x0) * OR_P2(tan_e) / 2 / rho / (1 + dp0) -
           in_fringe * OR_P2(y0) * OR_P2(sec_e) / 2 / rho / (1 + dp0);
      dy = -in_fringe * x0 * y0 * OR_P2(tan_e) / rho / (1 + dp0);
      dpx = tan_e * x0 / rho - OR_P2(y0) / 2 * (1 * OR_P3(tan_e)) / OR_P2(rho) / (1 + dp0) -
            OR_P2(x0) / 2 * (1 * OR_P3(tan_e)) / OR_P2(rho) / (1 + dp0) -
            in_fringe * (x0 * px0 - y0 * py0) * OR_P2(tan_e) / rho / (1 + dp0) +
            k4 * OR_P2(sin_e) * OR_P2(gap) / 2 / OR_P3(cos_e) / rho * Rhe -
            k5 * x0 * sin_e * gap / OR_P3(cos_e) / rho * Rhe +
            k6 * (y0 * y0 - x0 * x0) / 2 / OR_P3(cos_e) / rho * Rhe;
      dpy = -1. * tan_e * y0 / rho + k2 * y0 * (1 + OR_P2(sin_e)) * gap / (1 + dp0) / OR_P2(rho) / OR_P3(cos_e) +
            in_fringe * (x0 * py0 + y0 * px0) * OR_P2(tan_e) / rho / (1 + dp0) +
            in_fringe * y0 * px0 / rho / (1 + dp0) + x0 * y0 * OR_P2(sec_e) * tan_e / OR_P2(rho) / (1 + dp0) +
            k3 * OR_P3(y0) * (2. / 3. / cos_e - 4. / 3. / OR_P3(cos_e)) / (1 + dp0) / rho / rho / gap -
            k5 * y0 * sin_e * gap / OR_P3(cos_e) / rho * Rhe + k6 * x0 * y0 / OR_P3(cos_e) / rho * Rhe;
    }
  } else {
    if (in_fringe == -1.) {
      dx = in_fringe * OR_P2(sec_e) * OR_P2(gap) * k0 / rho / (1 + dp0);
      dy = 0;
      dpx = -1. * OR_P3(sec_e) * sin_e * OR_P2(gap) * k0 / rho / rho / (1 + dp0) + tan_e * x0 / rho +
            k4 * OR_P2(sin_e) * OR_P2(gap) / 2 / OR_P3(cos_e) / rho * Rhe -
            k5 * x0 * sin_e * gap / OR_P3(cos_e) / rho * Rhe;
      dpy = -1. * tan_e * y0 / rho +
            k2 * y0 * (1 + OR_P2(sin_e)) * gap / (1 + dp0) / OR_P2(rho) / (OR_PLANT == 5 ? OR_P2(cos_e) : OR_P3(cos_e));
    }
    if (in_fringe == 1.) {
      dx = in_fringe * OR_P2(sec_e) * OR_P2(gap) * k0 / rho / (1 + dp0);
      dy = 0;
      dpx = tan_e * x0 / rho + k4 * OR_P2(sin_e) * OR_P2(gap) / 2 / OR_P3(cos_e) / rho * Rhe -
            k5 * x0 * sin_e * gap / OR_P3(cos_e) / rho * Rhe;
      dpy = -1. * tan_e * y0 / rho +
            k2 * y0 * (1 + OR_P2(sin_e)) * gap / (1 + dp0) / OR_P2(rho) / (OR_PLANT == 5 ? OR_P2(cos_e) : OR_P3(cos_e)) -
            k5 * y0 * sin_e * gap / OR_P3(cos_e) / rho * Rhe;
    }
  }
  Q[0] = x0 + dx;
  Q[1] = px0 + dpx;
  Q[2] = y0 + dy;
  Q[3] = py0 + dpy;
}

/* dipoleFringeKHwangRLindberg, src/csbend.c:4413-4520 (EDGE_EFFECTS 4): Hwang's
 * fringe map made symplectic by R. Lindberg, a sequence of exact maps on the
 * canonical coordinates Q (one of them exponential in x), at the entrance
 * (in_fringe -1) or the exit (+1) */
static void or_fringe_hl(double *Q, double rho, double in_fringe, double edge, double gap, double fint, double Rhe) {
  const double k0 = sqr(PI) / 6., k2 = fint, k3 = 1.0 * 1. / 6., k4 = -1.0 * sqr(PI) / 3., k5 = 0.0, k6 = -1.0;
  const double rho2 = sqr(rho);
  const double x0 = Q[0], px0 = Q[1], y0 = Q[2], py0 = Q[3], dp0 = Q[5];
  const double sec_e = 1. / OR_COS(edge), tan_e = OR_TAN(edge), sin_e = OR_SIN(edge), cos_e = OR_COS(edge);
  const double sec2 = OR_P2(sec_e), cos3 = OR_P3(cos_e), tan2 = OR_P2(tan_e), tan3 = OR_P3(tan_e);
  double x1, px1, y1, py1, x2, px2, y2, py2, x3, px3, y3, py3, x4, px4, y4, py4, x5, px5, y5, py5, t1, t2;
  if (in_fringe == -1) {
    x1 = x0;
    px1 = px0 + tan_e / rho * x0 + tan_e / (2 * rho2 * (1 + dp0)) * sqr(y0) - tan3 / (rho2 * (1 + dp0)) * sqr(x0) -
          gap * k5 * sin_e * Rhe / (rho * cos3) * x0 + k6 * sec2 * Rhe / (2 * rho) * (sqr(y0) - sqr(x0));
    y1 = y0;
    py1 = py0 - tan_e / rho * y0 + tan_e / (rho2 * (1 + dp0)) * x0 * y0 +
          gap * k2 * (1 + sqr(sin_e)) / (rho2 * (1 + dp0) * cos3) * y0 +
          2 * k3 * (sqr(cos_e) - 2) / (3 * gap * rho2 * cos3) * OR_P3(y0) + gap * k5 * sin_e * Rhe / (rho * cos3) * y0 +
          k6 * OR_P3(sec_e) * Rhe / rho * x0 * y0;
    t1 = (1 + tan2 / (2 * rho * (1 + dp0)) * x1);
    x2 = x1 / t1;
    px2 = px1 * sqr(t1);
    y2 = y1;
    py2 = py1;
    t1 = sec2 / (rho * (1 + dp0));
    x3 = x2 + t1 / 2 * sqr(y2);
    px3 = px2;
    y3 = y2;
    py3 = py2 - t1 * y2 * px2;
    t1 = tan2 / (rho * (1 + dp0));
    t2 = OR_EXP(t1 * x3);
    x4 = x3;
    px4 = px3 - t1 * y3 * py3;
    y4 = y3 * t2;
    py4 = py3 / t2;
    t1 = sqr(gap) * k0 * sec2 / (rho * (1 + dp0));
    x5 = x4 - t1;
    px5 = px4 - t1 * tan_e / rho + sqr(gap) * k4 * sqr(sin_e) * Rhe / (2 * rho * cos3);
    y5 = y4;
    py5 = py4;
  } else {
    x1 = x0;
    px1 = px0 + tan_e / rho * x0 + tan3 / (2 * rho2 * (1 + dp0)) * sqr(y0) + tan3 / (2 * rho2 * (1 + dp0)) * sqr(x0) -
          gap * k5 * sin_e * Rhe / (rho * cos3) * x0 + k6 * sec2 * Rhe / (2 * rho) * (sqr(y0) - sqr(x0));
    y1 = y0;
    py1 = py0 - tan_e / rho * y0 + tan3 / (rho2 * (1 + dp0)) * x0 * y0 +
          gap * k2 * (1 + sqr(sin_e)) / (rho2 * (1 + dp0) * cos3) * y0 +
          2 * k3 * (sqr(cos_e) - 2) / (3 * gap * rho2 * cos3) * OR_P3(y0) + gap * k5 * sin_e * Rhe / (rho * cos3) * y0 +
          k6 * OR_P3(sec_e) * Rhe / rho * x0 * y0;
    t1 = (1 - tan2 / (2 * rho * (1 + dp0)) * x1);
    x2 = x1 / t1;
    px2 = px1 * sqr(t1);
    y2 = y1;
    py2 = py1;
    t1 = sec2 / (rho * (1 + dp0));
    x3 = x2 - t1 / 2 * sqr(y2);
    px3 = px2;
    y3 = y2;
    py3 = py2 + t1 * y2 * px2;
    t1 = -tan2 / (rho * (1 + dp0));
    t2 = OR_EXP(t1 * x3);
    x4 = x3;
    px4 = px3 - t1 * y3 * py3;
    y4 = y3 * t2;
    py4 = py3 / t2;
    x5 = x4 + sqr(gap) * k0 * sec2 / (rho * (1 + dp0));
    px5 = px4 + sqr(gap) * k4 * sqr(sin_e) * Rhe / (2 * rho * cos3);
    y5 = y4;
    py5 = py4;
  }
  Q[0] = x5;
  Q[1] = px5;
  Q[2] = y5;
  Q[3] = py5;
}

/* EDGE_EFFECTS 2 or 4 at one edge, src/csbend.c:1139-1181 and :1319-1360 */
static void or_edge_hwang(double *x, double *xp, double *y, double *yp, double dp, double rho_actual,
                          double in_fringe, long edge_order, double edge, double hgap, double fint, double h,
                          int lindberg) {
  double Q[6];
  Q[0] = *x;
  Q[1] = *xp;
  Q[2] = *y;
  Q[3] = *yp;
  Q[4] = 0;
  Q[5] = dp;
  toDipoleCanonical(Q);
  if (lindberg)
    or_fringe_hl(Q, rho_actual, in_fringe, edge, 2 * hgap, fint, h);
  else
    or_fringe_hwang(Q, rho_actual, in_fringe, edge_order, edge, 2 * hgap, fint, h);
  fromDipoleCanonical(Q);
  *x = Q[0];
  *xp = Q[1];
  *y = Q[2];
  *yp = Q[3];
}

/* applySimpleDipoleEdgeKick, src/csbend.c:5664-5690 (EDGE_EFFECTS 3, the
 * Hamiltonian not expanded) */
static void or_simple_edge_kick(double *xp, double *yp, double x, double y, double delta, double rho, double ea,
                                double psi, double kickLimit) {
  double Qi[6], dqx, dqy;
  Qi[0] = x;
  Qi[1] = *xp;
  Qi[2] = y;
  Qi[3] = *yp;
  Qi[4] = 0;
  Qi[5] = delta;
  toDipoleCanonical(Qi);
  dqx = OR_TAN(ea) / rho * x;
  if (kickLimit > 0 && fabs(dqx) > kickLimit)
    dqx = SIGN(dqx) * kickLimit;
  dqy = -OR_TAN(ea - psi / (1 + delta)) / rho * y;
  Qi[1] += dqx;
  Qi[3] += dqy;
  fromDipoleCanonical(Qi);
  *xp = Qi[1];
  *yp = Qi[3];
}

/* track_through_csbend, src/csbend.c:654-1434 (malignMethod 0, ETILT 0,
 * EDGE*_EFFECTS 0 to 4, no reference correction) */
/* apply_edge_effects, src/csbend.c:4291-4332: K. L. Brown's second-order edge
 * map (EDGE_ORDER >= 2); which_edge -1 at the entrance, 1 at the exit */
static void or_apply_edge_effects(double *x, double *xp, double *y, double *yp, double rho, double n, double beta,
                                  double he, double psi, long which_edge) {
  double h, tan_beta, tan2_beta, sec_beta, sec2_beta, h2;
  double R21, R43;
  double T111, T133, T211, T441, T331, T221, T233, T243, T431, T432;
  double x0, xp0, y0, yp0;
  h = 1 / rho;
  R21 = h * (tan_beta = OR_TAN(beta));
  R43 = -h * OR_TAN(beta - psi);
  h2 = sqr(h);
  T111 = which_edge * h / 2 * (tan2_beta = sqr(tan_beta));
  sec_beta = 1. / OR_COS(beta);
  sec2_beta = sqr(sec_beta);
  T133 = -which_edge * h / 2 * sec2_beta;
  T211 = which_edge == -1 ? -n * h2 * tan_beta : -h2 * (n + tan2_beta / 2) * tan_beta;
  T441 = -(T331 = T221 = -which_edge * h * tan2_beta);
  T233 = which_edge == -1 ? h2 * (n + .5 + tan2_beta) * tan_beta : h2 * (n - tan2_beta / 2) * tan_beta;
  T243 = which_edge * h * tan2_beta;
  T431 = h2 * (2 * n + (which_edge == 1 ? sec2_beta : 0)) * tan_beta;
  T432 = which_edge * h * sec2_beta;
  if (he != 0) {
    double term = h / 2 * he * sec2_beta * sec_beta;
    T211 += term;
    T233 -= term;
    T431 -= 2 * term;
  }
  x0 = *x;
  xp0 = *xp;
  y0 = *y;
  yp0 = *yp;
  *x = x0 + T111 * sqr(x0) + T133 * sqr(y0);
  *xp = xp0 + R21 * x0 + T211 * sqr(x0) + T221 * x0 * xp0 + T233 * sqr(y0) + T243 * y0 * yp0;
  *y = y0 + T331 * x0 * y0;
  *yp = yp0 + R43 * y0 + T441 * yp0 * x0 + T431 * x0 * y0 + T432 * xp0 * y0;
}

long or_track_through_csbend(double **part, long n_part, const EH_CSBEND *csbend, double Po, double z_start,
                             const EH_MAXAMP *maxamp) {
  double b[9], c[9], h, rho0, rho_actual, fse, angle, e1, e2, tilt, etilt, ttilt, cos_ttilt, sin_ttilt;
  double psi1, psi2, Kg1, Kg2, dxi, dyi, dzi, dxf, dyf, dzf, dz_lost, rho, delta_xp, e1_kick_limit, e2_kick_limit;
  double x, xp, y, yp, s, dp, dp0, Qi[6], Qf[6], *coord, rad_coef, n;
  long i_part, i_top, j, particle_lost;
  or_csbend_fields F;
  or_apdata ap;

  if (csbend->angle == 0) {
    or_exactDrift(part, n_part, csbend->length);
    return n_part;
  }
  if (csbend->integration_order != 2 && csbend->integration_order != 4 && csbend->integration_order != 6)
    return or_fail("CSBEND integration_order is invalid");
  if (csbend->malignMethod != 0 || csbend->etilt != 0 || csbend->expandHamiltonian)
    return or_fail("CSBEND option not supported");
  if ((csbend->edge1_effects < 0 || csbend->edge1_effects > 4) ||
      (csbend->edge2_effects < 0 || csbend->edge2_effects > 4))
    return or_fail("CSBEND edge method not supported");
  if (csbend->edge_order > 1 && (csbend->edge1_effects == 2 || csbend->edge2_effects == 2) && csbend->hgap == 0)
    return or_fail("CSBEND has EDGE_ORDER>1 and EDGE[12]_EFFECTS==2, but HGAP=0");

  rho0 = csbend->length / csbend->angle;
  b[0] = 0;
  if (csbend->use_bn) {
    b[1] = csbend->b1;
    b[2] = csbend->b2;
    b[3] = csbend->b3;
    b[4] = csbend->b4;
    b[5] = csbend->b5;
    b[6] = csbend->b6;
    b[7] = csbend->b7;
    b[8] = csbend->b8;
  } else {
    b[1] = csbend->k1 * rho0;
    b[2] = csbend->k2 * rho0;
    b[3] = csbend->k3 * rho0;
    b[4] = csbend->k4 * rho0;
    b[5] = csbend->k5 * rho0;
    b[6] = csbend->k6 * rho0;
    b[7] = csbend->k7 * rho0;
    b[8] = csbend->k8 * rho0;
  }
  for (j = 0; j < 9; j++)
    c[j] = 0;
  b[1] *= (1 + csbend->fse + csbend->fseQuadrupole) / (1 + csbend->fse + csbend->fseDipole);
  c[1] *= (1 + csbend->fse + csbend->fseQuadrupole) / (1 + csbend->fse + csbend->fseDipole);
  b[2] *= (1 + csbend->fse) / (1 + csbend->fse + csbend->fseDipole);
  c[2] *= (1 + csbend->fse) / (1 + csbend->fse + csbend->fseDipole);
  b[0] = csbend->xKick / csbend->angle;
  c[0] = csbend->yKick / csbend->angle;

  if (csbend->angle < 0) {
    angle = -csbend->angle;
    e1 = -csbend->e1;
    e2 = -csbend->e2;
    etilt = csbend->etilt * csbend->etiltSign;
    tilt = csbend->tilt + PI;
    rho0 = csbend->length / angle;
    for (j = 1; j < 9; j += 2) {
      b[j] *= -1;
      c[j] *= -1;
    }
  } else {
    angle = csbend->angle;
    e1 = csbend->e1;
    e2 = csbend->e2;
    etilt = csbend->etilt * csbend->etiltSign;
    tilt = csbend->tilt;
    rho0 = csbend->length / angle;
  }
  or_setup_ap(&ap, -tilt, maxamp);
  if (fabs(rho0) > 1e6)
    return or_fail("CSBEND with radius > 1e6 not supported");

  fse = csbend->fse + csbend->fseDipole;
  h = 1 / rho0;
  n = -b[1] / h;
  if (fabs(fse + 1) < 1e-12)
    fse = -1 + 1e-12;
  rho_actual = 1 / ((1 + fse) * h);
  /* classical radiation, src/csbend.c:956-959 */
  if (csbend->synch_rad)
    rad_coef = sqr(EH_E_MKS) * pow3(Po) * sqr(1 + fse) / (6 * PI * EH_EPSILON_O * sqr(c_mks) * EH_ME_MKS * sqr(rho0));
  else
    rad_coef = 0;
  /* edge kick limits, src/csbend.c:932-937 */
  e1_kick_limit = csbend->edge1_kick_limit;
  e2_kick_limit = csbend->edge2_kick_limit;
  if (csbend->kick_limit_scaling) {
    e1_kick_limit *= rho0 / rho_actual;
    e2_kick_limit *= rho0 / rho_actual;
  }
  Kg1 = 2 * csbend->hgap * csbend->fint * SIGN(rho0);
  psi1 = Kg1 / fabs(rho_actual) / OR_COS(e1) * (1 + sqr(OR_SIN(e1)));
  Kg2 = 2 * csbend->hgap * csbend->fint * SIGN(rho0);
  psi2 = Kg2 / fabs(rho_actual) / OR_COS(e2) * (1 + sqr(OR_SIN(e2)));
  if (csbend->length < 0) {
    psi1 *= -1;
    psi2 *= -1;
  }
  if (or_field_coefficients(&F, b, c, h, csbend->nonlinear, csbend->expansionOrder) < 0)
    return EH_ERR_UNSUPPORTED;

  ttilt = tilt + etilt;
  if (ttilt == 0) {
    cos_ttilt = 1;
    sin_ttilt = 0;
  } else if (fabs(fabs(ttilt) - PI) < 1e-12) {
    cos_ttilt = -1;
    sin_ttilt = 0;
  } else if (fabs(ttilt - PIo2) < 1e-12) {
    cos_ttilt = 0;
    sin_ttilt = 1;
  } else if (fabs(ttilt + PIo2) < 1e-12) {
    cos_ttilt = 0;
    sin_ttilt = -1;
  } else {
    cos_ttilt = OR_COS(ttilt);
    sin_ttilt = OR_SIN(ttilt);
  }
  /* computeEtiltCentroidOffset with etilt = 0 gives zero offsets */
  dxi = -csbend->dx;
  dzi = csbend->dz;
  dyi = -csbend->dy;
  dxf = csbend->dx * OR_COS(csbend->angle) + csbend->dz * OR_SIN(csbend->angle);
  dzf = csbend->dx * OR_SIN(csbend->angle) - csbend->dz * OR_COS(csbend->angle);
  dyf = csbend->dy;

  i_top = n_part - 1;
  for (i_part = 0; i_part <= i_top; i_part++) {
    coord = part[i_part];
    coord[4] += dzi * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
    coord[0] = coord[0] + dxi + dzi * coord[1];
    coord[2] = coord[2] + dyi + dzi * coord[3];
    x = coord[0] * cos_ttilt + coord[2] * sin_ttilt;
    y = -coord[0] * sin_ttilt + coord[2] * cos_ttilt;
    xp = coord[1] * cos_ttilt + coord[3] * sin_ttilt;
    yp = -coord[1] * sin_ttilt + coord[3] * cos_ttilt;
    s = coord[4];
    dp = dp0 = coord[5];

    /* src/csbend.c:1125-1200 */
    if ((csbend->edgeFlags & EH_BEND_EDGE1_EFFECTS) && csbend->edge1_effects == 1) {
      rho = (1 + dp) * rho_actual;
      if (csbend->edge_order <= 1) {
        delta_xp = OR_TAN(e1) / rho * x;
        if (e1_kick_limit > 0 && fabs(delta_xp) > e1_kick_limit)
          delta_xp = SIGN(delta_xp) * e1_kick_limit;
        xp += delta_xp;
        yp -= OR_TAN(e1 - psi1 / (1 + dp)) / rho * y;
      } else
        or_apply_edge_effects(&x, &xp, &y, &yp, rho, n, e1, csbend->h1, psi1 * (1 + dp), -1);
    } else if ((csbend->edgeFlags & EH_BEND_EDGE1_EFFECTS) &&
               (csbend->edge1_effects == 2 || csbend->edge1_effects == 4))
      or_edge_hwang(&x, &xp, &y, &yp, dp, rho_actual, -1., csbend->edge_order, e1, csbend->hgap, csbend->fint,
                    csbend->h1, csbend->edge1_effects == 4);
    else if ((csbend->edgeFlags & EH_BEND_EDGE1_EFFECTS) && csbend->edge1_effects == 3)
      or_simple_edge_kick(&xp, &yp, x, y, dp, rho_actual, e1, psi1, e1_kick_limit);
    Qi[0] = x;
    Qi[1] = xp;
    Qi[2] = y;
    Qi[3] = yp;
    Qi[4] = 0;
    Qi[5] = dp;
    if ((csbend->edgeFlags & EH_BEND_EDGE1_EFFECTS) && e1 != 0 && rad_coef) {
      /* pre-adjust delta for the sector integration, src/csbend.c:1214-1222 */
      double Fx, Fy, dp_prime;
      or_csbend_fields_eval(&F, &Fx, &Fy, x, y);
      dp_prime = -rad_coef * (sqr(Fx) + sqr(Fy)) * sqr(1 + dp) * sqrt(sqr(1 + x / rho0) + sqr(xp) + sqr(yp));
      Qi[5] -= dp_prime * x * OR_TAN(e1);
    }
    toDipoleCanonical(Qi);
    particle_lost = !or_integrate_csbend(Qf, Qi, csbend->length, csbend->nSlices, rho0, rho_actual, &dz_lost, &ap,
                                         csbend->integration_order, &F, rad_coef);
    fromDipoleCanonical(Qf);
    if (particle_lost) {
      double X0 = Qf[0], XP0 = Qf[1], Y0 = Qf[2], YP0 = Qf[3];
      memcpy(part[i_part], Qf, sizeof(double) * 6);
      /* convertFromCSBendCoords, src/csbend.c:4527-4550 */
      coord[0] = X0 * cos_ttilt - Y0 * sin_ttilt;
      coord[2] = X0 * sin_ttilt + Y0 * cos_ttilt;
      coord[1] = XP0 * cos_ttilt - YP0 * sin_ttilt;
      coord[3] = XP0 * sin_ttilt + YP0 * cos_ttilt;
      swapParticles(part[i_part], part[i_top]);
      part[i_top][4] = z_start + dz_lost;
      part[i_top][5] = Po * (1 + part[i_top][5]);
      i_top--;
      i_part--;
      continue;
    }
    if ((csbend->edgeFlags & EH_BEND_EDGE2_EFFECTS) && e2 != 0 && rad_coef) {
      /* post-adjust delta, src/csbend.c:1270-1282 */
      double Fx, Fy, dp_prime;
      or_csbend_fields_eval(&F, &Fx, &Fy, Qf[0], Qf[2]);
      dp_prime = -rad_coef * (sqr(Fx) + sqr(Fy)) * sqr(1 + Qf[5]) *
                 sqrt(sqr(1 + Qf[0] / rho0) + sqr(Qf[1]) + sqr(Qf[3]));
      Qf[5] -= dp_prime * Qf[0] * OR_TAN(e2);
    }
    {
      /* the time-of-flight rescaling of src/csbend.c:1285-1296 runs for every
       * bend (rad_coef || isrConstant, isrConstant non-zero): without radiation
       * beta1 = beta0, leaving the round trip beta0 * s / beta0 */
      double p0 = Po * (1 + dp0), beta0 = p0 / sqrt(sqr(p0) + 1);
      double p1 = Po * (1 + Qf[5]), beta1 = p1 / sqrt(sqr(p1) + 1);
      s = beta1 * s / beta0 + Qf[4];
    }
    x = Qf[0];
    xp = Qf[1];
    y = Qf[2];
    yp = Qf[3];
    dp = Qf[5];
    /* src/csbend.c:1306-1340 */
    if ((csbend->edgeFlags & EH_BEND_EDGE2_EFFECTS) && csbend->edge2_effects == 1) {
      rho = (1 + dp) * rho_actual;
      if (csbend->edge_order <= 1) {
        delta_xp = OR_TAN(e2) / rho * x;
        if (e2_kick_limit > 0 && fabs(delta_xp) > e2_kick_limit)
          delta_xp = SIGN(delta_xp) * e2_kick_limit;
        xp += delta_xp;
        yp -= OR_TAN(e2 - psi2 / (1 + dp)) / rho * y;
      } else
        or_apply_edge_effects(&x, &xp, &y, &yp, rho, n, e2, csbend->h2, psi2 * (1 + dp), 1);
    } else if ((csbend->edgeFlags & EH_BEND_EDGE2_EFFECTS) &&
               (csbend->edge2_effects == 2 || csbend->edge2_effects == 4))
      or_edge_hwang(&x, &xp, &y, &yp, dp, rho_actual, 1., csbend->edge_order, e2, csbend->hgap, csbend->fint,
                    csbend->h2, csbend->edge2_effects == 4);
    else if ((csbend->edgeFlags & EH_BEND_EDGE2_EFFECTS) && csbend->edge2_effects == 3)
      or_simple_edge_kick(&xp, &yp, x, y, dp, rho_actual, e2, psi2, e2_kick_limit);
    coord[0] = x * cos_ttilt - y * sin_ttilt + 0.0;
    coord[2] = x * sin_ttilt + y * cos_ttilt + 0.0;
    coord[1] = xp * cos_ttilt - yp * sin_ttilt + 0.0;
    coord[3] = xp * sin_ttilt + yp * cos_ttilt + 0.0;
    coord[4] = s + 0.0;
    coord[5] = dp;
    coord[0] += dxf + dzf * coord[1];
    coord[2] += dyf + dzf * coord[3];
    coord[4] += dzf * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
  }
  return i_top + 1;
}

/* ------------------------------------------------------------------------- */
/* RFCA */
static double KahanPlus(double oldSum, double b, double *error) { /* src/kahan.c:27-35 */
  double sum, Y;
  Y = b + *error;
  sum = oldSum + Y;
  *error = Y - (sum - oldSum);
  return sum;
}

/* do_match_energy, src/do_tracking.c:3376-3549 (serial, USE_KAHAN): the mean
 * momentum of the np particles in row order as a Kahan sum; change_beam 0 moves
 * the reference momentum to it (when they differ by more than 1e-14 relative)
 * and re-expresses every delta, change_beam 1 moves every particle's momentum by
 * the same amount so that the mean is the reference momentum, keeping its time
 * of flight.  A non-finite time coordinate is an error (the reference aborts). */
long or_do_match_energy(double **coord, long np, double *P_central, long change_beam) {
  long ip;
  double P_average, dP_centroid, P, t, dp, dr, error = 0.0;
  if (!np)
    return 0;
  P_average = 0;
  for (ip = 0; ip < np; ip++)
    P_average = KahanPlus(P_average, (*P_central) * (1 + coord[ip][5]), &error);
  P_average /= np;
  if (!change_beam) {
    if (fabs(P_average - (*P_central)) / (*P_central) > 1e-14) {
      dp = (*P_central - P_average) / P_average;
      dr = (*P_central) / P_average;
      for (ip = 0; ip < np; ip++)
        coord[ip][5] = dp + coord[ip][5] * dr;
      *P_central = P_average;
    }
  } else {
    dP_centroid = *P_central - P_average;
    for (ip = 0; ip < np; ip++) {
      P = (1 + coord[ip][5]) * (*P_central);
      t = coord[ip][4] / (P / sqrt(P * P + 1));
      P += dP_centroid;
      coord[ip][5] = (P - *P_central) / (*P_central);
      coord[ip][4] = t * (P / sqrt(P * P + 1));
      if (isnan(coord[ip][4]) || isinf(coord[ip][4]))
        return or_fail("bad time coordinate in do_match_energy");
    }
  }
  return np;
}

static double beta_from_delta(double p, double delta) { /* src/do_tracking.c:95-98 */
  p *= 1 + delta;
  return (p / sqrt(p * p + 1));
}

/* findFiducialTime, src/simple_rfca.c:62-262 (serial elegant, the whole beam):
 * LIGHT from the position; FIRST, whose loop never stops early and so always
 * ends in the reference's error; PMAXIMUM from the first particle whose
 * momentum exceeds every earlier one, starting from particle 0 (an error when
 * particle 0 has the largest); TMEAN as a Kahan sum.  0, or the error code. */
long or_findFiducialTime(double **part, long np, double s0, double sOffset, double p0, int mode, double *tFid) {
  *tFid = 0.0;
  if (mode == EH_FID_LIGHT) {
    *tFid = (s0 + sOffset) / c_mks;
  } else if (mode == EH_FID_FIRST) {
    long i;
    for (i = 0; i < np; i++)
      *tFid = (part[i][4] + sOffset) / (c_mks * beta_from_delta(p0, part[i][5]));
    if (i == np)
      return or_fail("No available particle for the FID_MODE_FIRST mode in findFiducialTime");
  } else if (mode == EH_FID_PMAX) {
    long i, ibest = -1;
    double best = part[0][5];
    for (i = 0; i < np; i++)
      if (best < part[i][5]) {
        best = part[i][5];
        ibest = i;
      }
    if (ibest >= 0 && ibest < np)
      *tFid = (part[ibest][4] + sOffset) / (c_mks * beta_from_delta(p0, part[ibest][5]));
    else
      return or_fail("No available particle for RF cavity fiducialization");
  } else {
    double tsum = 0, error = 0.0;
    long ip, nsum = 0;
    for (ip = 0; ip < np; ip++) {
      tsum = KahanPlus(tsum, (part[ip][4] + sOffset) / (c_mks * beta_from_delta(p0, part[ip][5])), &error);
      nsum++;
    }
    if (nsum > 0)
      *tFid = tsum / nsum;
  }
  return 0;
}

/* phase references, src/phase_reference.c:28-143: numbered groups of
 * time-dependent elements that share one reference phase, set by the first
 * element of the group tracked (with particles) and kept until deleted.  The
 * reference gives an RFCA without a number an unused one (unused_phase_reference,
 * :129-141), a group only that cavity can set, so it always tracks with its own
 * fiducial; the checker keeps no entry for it and writes no number back. */
#define OR_REF_PHASE_RETURNED 1
#define OR_REF_PHASE_NOT_SET 2
#define OR_REF_PHASE_NONEXISTENT 3
#define OR_MAX_PHASE_REFS 4096
static struct {
  long ref_number, set;
  double phase;
} or_refs[OR_MAX_PHASE_REFS];
static long or_n_refs;

long or_get_phase_reference(double *phase, long ref) {
  long i;
  if (ref == 0)
    return 0;
  for (i = 0; i < or_n_refs; i++)
    if (or_refs[i].ref_number == ref) {
      if (!or_refs[i].set)
        return OR_REF_PHASE_NOT_SET;
      *phase = or_refs[i].phase;
      return OR_REF_PHASE_RETURNED;
    }
  return OR_REF_PHASE_NONEXISTENT;
}

long or_set_phase_reference(long ref, double phase) {
  long i;
  if (ref == 0)
    return 0;
  for (i = 0; i < or_n_refs; i++)
    if (or_refs[i].ref_number == ref)
      break;
  if (i == or_n_refs) {
    if (ref > LONG_MAX / 2)
      return or_fail("please use a small integer for the phase_reference number");
    if (or_n_refs == OR_MAX_PHASE_REFS)
      return or_fail("too many phase references");
    or_n_refs++;
  }
  or_refs[i].ref_number = ref;
  or_refs[i].phase = phase;
  or_refs[i].set = 1;
  return 1;
}

void or_delete_phase_references(void) {
  long i;
  for (i = 0; i < or_n_refs; i++)
    or_refs[i].set = 0;
}


/* add_to_particle_energy, src/simple_rfca.c:987-1006 */
static void add_to_particle_energy(double *coord, double timeOfFlight, double Po, double dgamma) {
  double gamma, gamma1, PRatio, P, P1, Pz1, Pz;
  P = Po * (1 + coord[5]);
  gamma1 = (gamma = sqrt(P * P + 1)) + dgamma;
  if (gamma1 <= 1)
    gamma1 = 1 + 1e-7;
  P1 = sqrt(gamma1 * gamma1 - 1);
  coord[5] = (P1 - Po) / Po;
  coord[4] = timeOfFlight * c_mks * P1 / gamma1;
  Pz = P / sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
  Pz1 = sqrt(Pz * Pz + gamma1 * gamma1 - gamma * gamma);
  PRatio = OR_PLANT == 4 ? Pz1 / Pz : Pz / Pz1;
  coord[1] *= PRatio;
  coord[3] *= PRatio;
}

/* simple_rf_cavity -> trackRfCavityWithWakes, src/simple_rfca.c:264-985 (no wakes,
 * no LSC, no backtrack, body focus model none; LINEARIZE and LOCK_PHASE through
 * the beam's mean arrival time; Q gives
 * the fill factor sqrt(1 - exp(-dt/tau)), tau = Q/omega, dt the arrival after
 * the fiducial time, :399-401, :594-599, :872);
 * CHANGE_T takes whole RF periods off the time coordinates, judged on the
 * particle in row 0 (:488-498, :586-588, :864-865); CHANGE_P0 matches the
 * reference momentum to the surviving beam (:977-982) */
long or_simple_rf_cavity(double **part, long np, EH_RFCA *rfca, double *P_central, double zEnd) {
  long ip, nKicks, ik;
  double length, omega, volt, tau, dtLight, phase, t0, To, timeOffset, dgamma = 0, dgammaMax = 0;
  double P, gamma, gamma1, beta_i, t, dt, x, xp, dc4, *coord;
  int matrixMethod = 0, same_dgamma = 0;

  double tAve = 0, dgammaAve = 0, cos_phase;
  if (np <= 0)
    return np;
  length = rfca->length;
  if (rfca->volt == 0) {
    if (rfca->length) {
      for (ip = 0; ip < np; ip++) {
        coord = part[ip];
        coord[0] += coord[1] * length;
        coord[2] += coord[3] * length;
        coord[4] += length * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
      }
    }
    return np;
  }
  omega = PIx2 * rfca->freq;
  volt = rfca->volt / (1e6 * particleMassMV * particleRelSign);
  tau = omega ? rfca->Q / omega : 0;
  if (rfca->change_t && rfca->Q)
    return or_fail("incompatible RF cavity parameters: change_t!=0 && Q!=0");
  nKicks = length ? rfca->nKicks : 1;
  if (nKicks <= 0) {
    matrixMethod = 1;
    nKicks = 1;
  }
  length /= nKicks;
  volt /= nKicks;
  dtLight = length / c_mks;

  /* src/simple_rfca.c:412-466 */
  phase = 0;
  if (!rfca->phase_reference || or_get_phase_reference(&phase, rfca->phase_reference) != OR_REF_PHASE_RETURNED) {
    if (!rfca->fiducial_seen) {
      if (rfca->fiducial_mode < EH_FID_TMEAN || rfca->fiducial_mode > EH_FID_PMAX)
        return or_fail("invalid fiducial mode for RFCA element");
      if (rfca->tReference != -1)
        t0 = rfca->tReference;
      else {
        long r = or_findFiducialTime(part, np, zEnd - rfca->length, length / 2, *P_central, rfca->fiducial_mode,
                                     &t0);
        if (r < 0)
          return r;
      }
      rfca->phase_fiducial = -omega * t0;
      rfca->fiducial_seen = 1;
    }
    long r = or_set_phase_reference(rfca->phase_reference, phase = rfca->phase_fiducial);
    if (r < 0)
      return r;
  }
  if (omega) {
    t0 = -rfca->phase_fiducial / omega;
    rfca->t_fiducial = t0;
    To = PIx2 / omega;
  } else
    t0 = To = 0;
  phase += rfca->phase * PI / 180.0;

  if (omega == 0 && tau == 0) {
    dgamma = volt * OR_SIN(phase);
    dgammaMax = volt;
    same_dgamma = 1;
  }
  timeOffset = 0;
  if (omega && rfca->change_t && np != 0) {
    coord = part[0];
    P = *P_central * (1 + coord[5]);
    beta_i = P / (gamma = sqrt(sqr(P) + 1));
    t = coord[4] / (c_mks * beta_i);
    if (omega != 0 && t > (0.9 * To) && rfca->change_t)
      timeOffset = ((long)(t / To + 0.5)) * To;
  }
  /* LINEARIZE, LOCK_PHASE: the mean arrival time (a Kahan sum in row order),
   * src/simple_rfca.c:500-556 */
  if (rfca->linearize || rfca->lockPhase) {
    double error = 0.0;
    if (nKicks != 1)
      return or_fail("Must use n_kicks=1 for linearized rf cavity");
    for (ip = 0; ip < np; ip++) {
      coord = part[ip];
      P = *P_central * (1 + coord[5]);
      beta_i = P / (gamma = sqrt(sqr(P) + 1));
      t = (coord[4] + length / 2) / (c_mks * beta_i) - timeOffset;
      tAve = KahanPlus(tAve, t, &error);
    }
    tAve /= np;
    if (rfca->lockPhase) {
      phase = PI / 180 * rfca->phase;
      dgammaAve = volt * OR_SIN(phase);
    } else
      dgammaAve = volt * OR_SIN(omega * (tAve - timeOffset) + phase);
  }

  for (ip = 0; ip < np; ip++) {
    coord = part[ip];
    coord[0] -= rfca->dx;
    coord[2] -= rfca->dy;
  }

  if (!matrixMethod) {
    double *inverseF = calloc(sizeof(*inverseF), np > 0 ? np : 1);
    for (ik = 0; ik < nKicks; ik++) {
      for (ip = 0; ip < np; ip++) {
        coord = part[ip];
        if (coord[5] == -1)
          continue;
        if (length)
          dc4 = length / 2 * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
        else
          dc4 = 0;
        P = *P_central * (1 + coord[5]);
        beta_i = P / (gamma = sqrt(sqr(P) + 1));
        t = (coord[4] + dc4) / (c_mks * beta_i) - timeOffset;
        if (ik == 0 && timeOffset && rfca->change_t)
          coord[4] = t * c_mks * beta_i - dc4;
        if ((dt = t - t0) < 0)
          dt = 0;
        if (!same_dgamma) {
          if (!rfca->linearize) {
            if (rfca->standingWave)
              dgamma = volt * OR_SIN(omega * (t - (rfca->lockPhase ? tAve : 0)) + phase) *
                       (tau ? sqrt(1 - OR_EXP(-dt / tau)) : 1);
            else {
              double phi = omega * (t - (rfca->lockPhase ? tAve : 0) - ik * dtLight) + phase;
              double dfactor = tau ? sqrt(1 - OR_EXP(-dt / tau)) : 1;
              dgamma = volt * OR_SIN(phi) * dfactor;
            }
          } else
            dgamma = dgammaAve + volt * omega * (t - tAve) * OR_COS(omega * (tAve - timeOffset) + phase);
        }
        if (length) {
          if (rfca->end1Focus && ik == 0) {
            inverseF[ip] = dgamma / (2 * gamma * length);
            coord[1] -= coord[0] * inverseF[ip];
            coord[3] -= coord[2] * inverseF[ip];
          }
          coord[0] += coord[1] * length / 2;
          coord[2] += coord[3] * length / 2;
          coord[4] += length / 2 * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
        }
        add_to_particle_energy(coord, t, *P_central, dgamma);
        if ((gamma1 = gamma + dgamma) <= 1)
          coord[5] = -1;
        else
          inverseF[ip] = -dgamma / (2 * gamma1 * length);
      }
      if (length) {
        for (ip = 0; ip < np; ip++) {
          coord = part[ip];
          coord[0] += coord[1] * length / 2;
          coord[2] += coord[3] * length / 2;
          coord[4] += length / 2 * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
          if (rfca->end2Focus && (ik == nKicks - 1)) {
            coord[1] -= coord[0] * inverseF[ip];
            coord[3] -= coord[2] * inverseF[ip];
          }
        }
      }
    }
    free(inverseF);
  } else {
    double sin_phase = 0.0, inverseF;
    double R11 = 1, R21 = 0, R22, R12, dP, ds1;
    for (ik = 0; ik < nKicks; ik++) {
      for (ip = 0; ip < np; ip++) {
        coord = part[ip];
        P = *P_central * (1 + coord[5]);
        beta_i = P / (gamma = sqrt(sqr(P) + 1));
        ds1 = length / 2 * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
        t = (coord[4] + ds1) / (c_mks * beta_i) - timeOffset;
        if (timeOffset && rfca->change_t)
          coord[4] = t * c_mks * beta_i - ds1;
        if ((dt = t - t0) < 0)
          dt = 0;
        if (!same_dgamma) {
          if (!rfca->linearize) {
            sin_phase = OR_SIN(omega * (t - (rfca->lockPhase ? tAve : 0) - ik * dtLight) + phase);
            dgamma = (dgammaMax = volt * (tau ? sqrt(1 - OR_EXP(-dt / tau)) : 1)) * sin_phase;
          } else {
            cos_phase = OR_COS(omega * tAve + phase);
            sin_phase = omega * (t - tAve) * cos_phase;
            dgamma = (dgammaMax = volt * (tau ? sqrt(1 - OR_EXP(-dt / tau)) : 1)) * sin_phase + dgammaAve;
          }
        }
        if (rfca->end1Focus && length && ik == 0) {
          inverseF = dgamma / (2 * gamma * length);
          coord[1] -= coord[0] * inverseF;
          coord[3] -= coord[2] * inverseF;
        }
        dP = sqrt(sqr(gamma + dgamma) - 1) - P;
        R22 = OR_PLANT == 2 ? 1 + dP / P : 1 / (1 + dP / P);
        if (fabs(dP / P) > 1e-14)
          R12 = OR_PLANT == 1 ? length * (dP / P * OR_LOG(1 + dP / P)) : length * (P / dP * OR_LOG(1 + dP / P));
        else
          R12 = length;
        coord[4] += ds1;
        x = coord[0];
        xp = coord[1];
        coord[0] = x * R11 + xp * R12;
        coord[1] = x * R21 + xp * R22;
        x = coord[2];
        xp = coord[3];
        coord[2] = x * R11 + xp * R12;
        coord[3] = x * R21 + xp * R22;
        coord[4] += length / 2 * sqrt(1 + sqr(coord[1]) + sqr(coord[3]));
        coord[5] = (P + dP - (*P_central)) / (*P_central);
        if ((gamma += dgamma) <= 1)
          coord[5] = -1;
        if (rfca->end2Focus && length && ik == (nKicks - 1)) {
          inverseF = -dgamma / (2 * gamma * length);
          coord[1] -= coord[0] * inverseF;
          coord[3] -= coord[2] * inverseF;
        }
        coord[4] = OR_PLANT == 3 ? beta_i * coord[4] / ((P + dP) / gamma) : (P + dP) / gamma * coord[4] / beta_i;
      }
    }
    (void)dgammaMax;
  }
  for (ip = 0; ip < np; ip++) {
    coord = part[ip];
    coord[0] += rfca->dx;
    coord[2] += rfca->dy;
  }
  np = or_removeInvalidParticles(part, np, zEnd, *P_central);
  if (rfca->change_p0) {
    long r = or_do_match_energy(part, np, P_central, 0);
    if (r < 0)
      return r;
  }
  return np;
}

/* ------------------------------------------------------------------------- */
/* do_tracking's pass and element loops for this element set,
 * src/do_tracking.c:144 (initial NaN sweep :478-483, pass loop :498, z bookkeeping
 * :1101-1108, dispatch :1352-2495, MAXAMP sweep :2506-2517, NaN sweep and loss
 * recording :2675-2684, nToTrack = nLeft :2750).  wrap_around = 1. */
static double or_elem_length(const EH_ELEMENT_LIST *e) {
  switch (e->type) {
  case EH_T_MARK:
  case EH_T_MAXAMP:
  case EH_T_ROTATE:
  case EH_T_MALIGN:
  case EH_T_ENERGY:
    return 0;
  default:
    return *(const double *)e->p_elem; /* length is the first member, as in the reference */
  }
}

static void recordLostParticles(double **coord, long nLeft, long nToTrack, long pass) {
  long i;
  for (i = nLeft; i < nToTrack; i++)
    coord[i][EH_LOSS_PASS_INDEX] = pass;
}

long or_do_tracking(double **coord, long np, const EH_ELEMENT_LIST *beamline, long n_elements, long n_passes,
                    double *P_central, double *effort) {
  long nToTrack = np, nLeft, i_pass, ie;
  const EH_MAXAMP *maxamp = NULL;
  double x_max = 0, y_max = 0, z, last_z, z_travel;
  long elliptical = 0, maxampExponent = 0, maxampYExponent = 0;

  if (effort)
    *effort = 0;
  nLeft = or_limit_amplitudes(coord, DBL_MAX, DBL_MAX, nToTrack, 0.0, *P_central, 0);
  if (nLeft != nToTrack)
    recordLostParticles(coord, nLeft, nToTrack, 0);
  nToTrack = nLeft;

  for (i_pass = 0; i_pass < n_passes; i_pass++) {
    z = z_travel = last_z = 0;
    for (ie = 0; ie < n_elements; ie++) {
      const EH_ELEMENT_LIST *eptr = beamline + ie;
      double len = or_elem_length(eptr);
      last_z = z;
      z += len;
      z_travel += len;
      nLeft = nToTrack;
      if (effort && eptr->type != EH_T_MARK && eptr->type != EH_T_MAXAMP)
        *effort += nToTrack;
      switch (eptr->type) {
      case EH_T_MARK:
        break;
      case EH_T_MAXAMP:
        maxamp = (const EH_MAXAMP *)eptr->p_elem;
        x_max = maxamp->x_max;
        y_max = maxamp->y_max;
        elliptical = maxamp->elliptical;
        maxampExponent = maxamp->exponent;
        maxampYExponent = maxamp->y_exponent;
        break;
      case EH_T_DRIF: {
        const EH_DRIFT *d = (const EH_DRIFT *)eptr->p_elem;
        or_drift_beam(coord, nToTrack, d->length, d->order);
        break;
      }
      case EH_T_EDRIFT:
        or_exactDrift(coord, nToTrack, ((const EH_EDRIFT *)eptr->p_elem)->length);
        break;
      case EH_T_KQUAD:
      case EH_T_KSEXT:
      case EH_T_KOCT:
      case EH_T_KQUSE:
        nLeft = or_multipole_tracking2(coord, nToTrack, eptr, *P_central, last_z, maxamp);
        break;
      case EH_T_CSBEND:
        nLeft = or_track_through_csbend(coord, nToTrack, (const EH_CSBEND *)eptr->p_elem, *P_central, last_z, maxamp);
        break;
      case EH_T_MULT:
        nLeft = or_multipole_tracking(coord, nToTrack, (const EH_MULT *)eptr->p_elem, *P_central, last_z);
        break;
      case EH_T_ROTATE: {
        const EH_ROTATE *r = (const EH_ROTATE *)eptr->p_elem;
        if (r->on_pass < 0 || r->on_pass == i_pass)
          or_rotate(coord, nToTrack, r);
        break;
      }
      case EH_T_MALIGN: {
        const EH_MALIGN *m = (const EH_MALIGN *)eptr->p_elem;
        if (m->on_pass == -1 || m->on_pass == i_pass)
          nLeft = or_offset_beam(coord, nToTrack, m, *P_central);
        break;
      }
      case EH_T_ENERGY:
        nLeft = or_energy(coord, nToTrack, (const EH_ENERGY *)eptr->p_elem, P_central);
        break;
      case EH_T_RFCA:
        nLeft = or_simple_rf_cavity(coord, nToTrack, (EH_RFCA *)eptr->p_elem, P_central, z_travel);
        break;
      default:
        return or_fail("element type not supported");
      }
      if (nLeft < 0)
        return nLeft;
      if (eptr->type != EH_T_MARK) {
        if (x_max || y_max) {
          if (!elliptical)
            nLeft = or_limit_amplitudes(coord, x_max, y_max, nLeft, z, *P_central, eptr->type == EH_T_DRIF);
          else
            nLeft = or_elimit_amplitudes(coord, x_max, y_max, nLeft, z, *P_central, eptr->type == EH_T_DRIF,
                                         maxampExponent, maxampYExponent);
          if (nLeft < 0)
            return nLeft;
        }
      }
      nLeft = or_limit_amplitudes(coord, DBL_MAX, DBL_MAX, nLeft, z, *P_central, 0);
      if (nLeft != nToTrack)
        recordLostParticles(coord, nLeft, nToTrack, i_pass);
      nToTrack = nLeft;
    }
  }
  return nToTrack;
}

/* contiguous AoS wrapper for ctypes: rows of EH_NPROP doubles */
static double **or_rows(double *flat, long n) {
  double **rows = malloc(sizeof(double *) * (n > 0 ? n : 1));
  long i;
  for (i = 0; i < n; i++)
    rows[i] = flat + i * EH_NPROP;
  return rows;
}

long or_do_tracking_flat(double *flat, long np, const EH_ELEMENT_LIST *beamline, long n_elements, long n_passes,
                         double *P_central, double *effort) {
  double **rows = or_rows(flat, np);
  long r = or_do_tracking(rows, np, beamline, n_elements, n_passes, P_central, effort);
  /* rows were permuted through pointer swaps of row contents, so flat is already in row order */
  free(rows);
  return r;
}

/* single-element helpers for tests */
long or_element_flat(double *flat, long np, const EH_ELEMENT_LIST *eptr, double *P_central, double z_start,
                     double z_end, const EH_MAXAMP *maxamp) {
  double **rows = or_rows(flat, np);
  long r = np;
  switch (eptr->type) {
  case EH_T_DRIF:
    or_drift_beam(rows, np, ((const EH_DRIFT *)eptr->p_elem)->length, ((const EH_DRIFT *)eptr->p_elem)->order);
    break;
  case EH_T_EDRIFT:
    or_exactDrift(rows, np, ((const EH_EDRIFT *)eptr->p_elem)->length);
    break;
  case EH_T_KQUAD:
  case EH_T_KSEXT:
  case EH_T_KOCT:
  case EH_T_KQUSE:
    r = or_multipole_tracking2(rows, np, eptr, *P_central, z_start, maxamp);
    break;
  case EH_T_CSBEND:
    r = or_track_through_csbend(rows, np, (const EH_CSBEND *)eptr->p_elem, *P_central, z_start, maxamp);
    break;
  case EH_T_MULT:
    r = or_multipole_tracking(rows, np, (const EH_MULT *)eptr->p_elem, *P_central, z_start);
    break;
  case EH_T_ROTATE:
    or_rotate(rows, np, (const EH_ROTATE *)eptr->p_elem);
    break;
  case EH_T_MALIGN:
    r = or_offset_beam(rows, np, (const EH_MALIGN *)eptr->p_elem, *P_central);
    break;
  case EH_T_ENERGY:
    r = or_energy(rows, np, (const EH_ENERGY *)eptr->p_elem, P_central);
    break;
  case EH_T_RFCA:
    r = or_simple_rf_cavity(rows, np, (EH_RFCA *)eptr->p_elem, P_central, z_end);
    break;
  default:
    r = or_fail("element type not supported");
  }
  free(rows);
  return r;
}

long or_do_match_energy_flat(double *flat, long np, double *P_central, long change_beam) {
  double **rows = or_rows(flat, np);
  long r = or_do_match_energy(rows, np, P_central, change_beam);
  free(rows);
  return r;
}

long or_limit_amplitudes_flat(double *flat, double xmax, double ymax, long np, double z, double Po, long extrapolate_z) {
  double **rows = or_rows(flat, np);
  long r = or_limit_amplitudes(rows, xmax, ymax, np, z, Po, extrapolate_z);
  free(rows);
  return r;
}

long or_elimit_amplitudes_flat(double *flat, double xmax, double ymax, long np, double z, double Po,
                               long extrapolate_z, long exponent, long yexponent) {
  double **rows = or_rows(flat, np);
  long r = or_elimit_amplitudes(rows, xmax, ymax, np, z, Po, extrapolate_z, exponent, yexponent);
  free(rows);
  return r;
}

/* math entry points exposed for tests */
double or_sin(double x) { return OR_SIN(x); }
double or_cos(double x) { return OR_COS(x); }
double or_asin(double x) { return OR_ASIN(x); }
double or_exp(double x) { return OR_EXP(x); }
double or_log(double x) { return OR_LOG(x); }
