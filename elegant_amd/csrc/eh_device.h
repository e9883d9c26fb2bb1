/* eh_device.h -- per-particle element maps for gfx950, and the element
 * descriptor the host prepares for them.
 *
 * Every map keeps the operation order of the elegant routine it implements
 * (cited per function) so that, compiled with -ffp-contract=off, a particle's
 * six coordinates come out bit-identical to the CPU routine.  The maps work on
 * one particle held in registers; loss bookkeeping writes exactly the values the
 * CPU routine leaves in a lost row (before its swap to the top of the array).
 *
 * Per-element constants that the reference recomputes per particle from
 * element parameters (e.g. KnL/n_parts * kickFrac[step], drift*driftFrac[step])
 * are computed once on the host with the same expression, which gives the same
 * double.
 */
#ifndef EH_DEVICE_H
#define EH_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "eh_math.h"

/* the per-particle maps compile for the GPU and, for test harnesses, the host */
#define EH_HD __host__ __device__ __forceinline__
#define EH_PI_D 3.14159265358979323846

#define EH_MAXSUB 8     /* integrator sub-steps (order 6) */
#define EH_MAXFTERMS 66 /* field terms per component (expansion order 10) */
#define EH_MULT_MAXORDER 8 /* highest MULT order */

#define EH_DBL_MAX 1.7976931348623157e308

/* loss stages inside one element, in the order the reference compacts them */
#define EH_STAGE_ELEMENT 0 /* swap-to-top inside the element routine */
#define EH_STAGE_MAXAMP 1  /* limit_amplitudes with MAXAMP limits after the element */
#define EH_STAGE_NAN 2     /* limit_amplitudes(DBL_MAX) NaN sweep */
/* flag on EH_STAGE_ELEMENT: a kick multipole lost the particle before its first
 * step; the reference records z_start plus the dzLoss its previous particle left
 * behind (integrate_kick_multipole_ordn resets dzLoss only after the entry
 * checks, src/multipole.c:1160-1232), which the host resolves in row order */
#define EH_STAGE_INHERIT 4
/* loss pass word: the pass, with this bit set for a loss that inherits its
 * position (EH_STAGE_INHERIT); cleared before the word reaches the caller */
#define EH_LPASS_INHERIT 0x40000000
/* returned by the lean map (RAD = false) for an element type it does not carry
 * (MULT, ROTATE, MALIGN, ENERGY) or a descriptor with an elliptical aperture;
 * the host never launches a lean kernel on a line that holds one
 * (eh_lean_capable) */
#define EH_STAGE_REFUSED 8

typedef struct {
  int32_t i, j;
  double coef;
} eh_fterm;

/* element descriptor: everything the maps need, precomputed on the host */
typedef struct {
  int32_t type; /* eh_type */
  int32_t counted; /* counts as a particle-element unit (not MARK/MAXAMP) */
  /* in-element aperture (MAXAMP passed to the routine) for pass 0 / later passes */
  double ap_xmax[2], ap_ymax[2];
  /* post-element MAXAMP sweep limits for pass 0 / later passes; extrapolate for DRIF */
  double post_xmax[2], post_ymax[2];
  int32_t post_ap[2];
  int32_t extrap;
  /* elliptical MAXAMP for pass 0 / later passes: the in-element check of
   * checkMultAperture (ap_ell: ELLIPTICAL with both limits positive) and the
   * elimit_amplitudes sweep (post_ell), with exponents xe, ye and
   * a2 = ipow(x_max, xe), b2 = ipow(y_max, ye); ell_any if any is set (such
   * descriptors run only through the full instantiation of the maps) */
  int32_t ap_ell[2], post_ell[2];
  int32_t ell_xe[2], ell_ye[2];
  double ell_a2[2], ell_b2[2];
  int32_t ell_any;
  /* the descriptor needs the full instantiation of the maps: an elliptical
   * aperture, a bend edge method 3 or an edge kick limit */
  int32_t full_only;
  /* last_z and z of do_tracking for this element (z_start, z_end).  The maps
   * take them through a pointer (zp) and read them only for lost particles:
   * the fused kernel passes a per-position table instead, so that positions
   * with equal constants share one descriptor. */
  double zpos[2];
  double Po;

  /* ---- drifts ---- */
  double length;
  double T11; /* DRIF second-order coefficient length/2 (0 for order 1) */

  /* ---- kick multipoles ---- */
  int32_t nslices, nsub, expand;
  int32_t mp_order[3], mp_skew[3];
  int32_t maxOrder;
  int32_t has_xkick, has_ykick;
  int32_t has_offset, has_tilt, ap_rtilt;
  double mdrift; /* drift/n_parts */
  double dsh[EH_MAXSUB];
  double kick[EH_MAXSUB][3];
  double xk[EH_MAXSUB], yk[EH_MAXSUB];
  double mdx, mdy, mdz; /* element offsets */
  double ncos, nsin;   /* entry rotation (angle tilt) */
  double xcos, xsin;   /* exit rotation (angle -tilt) */
  double ap_cs0, ap_cs1; /* aperture reverse-tilt rotation */

  /* ---- CSBEND ---- */
  int32_t csb_nslices;
  int32_t csb_field; /* hasSkew || hasNormal */
  int32_t csb_nfx, csb_nfy;
  int32_t csb_edge1, csb_edge2;  /* the edge applies (EDGE*_EFFECTS 1 or 3) */
  int32_t csb_esimple1, csb_esimple2; /* EDGE*_EFFECTS 3: applySimpleDipoleEdgeKick */
  /* EDGE*_EFFECTS 2 (dipoleFringeKHwang): per edge 1/cos, tan, sin and cos of
   * the edge angle, the full gap 2*HGAP, FINT and the pole-face curvature H;
   * csb_hw_order2 for EDGE_ORDER > 1 */
  int32_t csb_hwang1, csb_hwang2, csb_hw_order2;
  int32_t csb_hl1, csb_hl2; /* EDGE*_EFFECTS 4: dipoleFringeKHwangRLindberg, same constants */
  double csb_hw1[7], csb_hw2[7];
  double csb_klim1, csb_klim2;   /* edge x' kick limits (<= 0: none) */
  double rho0, rho_actual;
  int32_t csb_recip;        /* 1: divide by rho0 and rho_actual through the reciprocals below */
  double rho0_rh, rho0_rl, rhoa_rh, rhoa_rl; /* RN(1/rho) and RN(1/rho - RN(1/rho)) */
  double csb_dsh[EH_MAXSUB], csb_dshr[EH_MAXSUB], csb_ds[EH_MAXSUB];
  double csb_dzl; /* n * s */
  double cos_tt, sin_tt;
  double dxi, dyi, dzi, dxf, dyf, dzf;
  double tan_e1, tan_e2, e1, e2, psi1, psi2;
  int32_t csb_eorder2;      /* EDGE_ORDER >= 2: apply_edge_effects at the edges */
  double csb_n, csb_sec1, csb_sec2, csb_he1, csb_he2; /* its n, 1/cos(e), pole-face curvatures */
  int32_t csb_fx0, csb_fy0; /* offsets of this element's field terms in the term pool */
  int32_t csb_maxj;         /* largest power of y in the field terms */
  int32_t csb_dense;        /* EO1 of the dense blocks at csb_dx0/csb_dy0 (5 or 6), else 0 */
  int32_t csb_dx0, csb_dy0;
  uint32_t csb_live;        /* dense blocks: bit t set if the t-th term (i-outer, j-inner) is non-zero */

  /* ---- RFCA ---- */
  int32_t rf_matrix, rf_nkicks, rf_same, rf_standing, rf_end1, rf_end2, rf_zero_volt;
  double rf_len;  /* length / nKicks */
  double rf_flen; /* full element length (volt == 0 drift) */
  double rf_volt, rf_omega, rf_phase, rf_dtLight, rf_t0, rf_dx, rf_dy, rf_dgamma_same;
  double rf_toff; /* CHANGE_T: whole RF periods taken off the arrival times at this visit (0: none) */
  double rf_tau;  /* Q / omega, the cavity's fill time (0: none) */
  /* LOCK_PHASE / LINEARIZE at this visit: the beam's mean arrival time tave,
   * tlock = tave with LOCK_PHASE (else 0), the mean energy gain and the cosines
   * of the linearized kick (kick and matrix methods) */
  int32_t rf_lin, rf_pad_;
  double rf_tlock, rf_tave, rf_dgave, rf_cosk, rf_cosm;

  /* ---- MULT, ROTATE, MALIGN ---- */
  int32_t on_pass;                      /* ROTATE/MALIGN: -1 every pass, -2 none, else the one pass acted on */
  double mult_coef[EH_MULT_MAXORDER + 1]; /* MULT: expansion coefficients of its order */
  double mal_dxp, mal_dyp, mal_dt, mal_dp, mal_de; /* MALIGN (offsets in mdx, mdy, mdz) */
  int32_t mal_energy;                   /* MALIGN: dt, dp or de non-zero */
  int32_t en_apply;                     /* ENERGY: the reference momentum changes here */
  double en_pold, en_pnew;              /* ENERGY: P_central before and after */

  /* ---- classical radiation (SYNCH_RAD) ---- */
  int32_t rad;                 /* rad_coef != 0: the radiating instantiation of the map */
  double rad_coef;
  double kfrac[EH_MAXSUB];     /* multipoles: the integrator's kick fractions */
  double rad_xk, rad_yk;       /* multipoles: xkick, ykick / n_parts */
  double csb_irho0;            /* bends: 1./rho0 as the radiation block forms it */
  int32_t csb_e1r, csb_e2r;    /* bends: the entry/exit delta adjustments apply */

  /* ---- adjoint (reverse-mode) extras ---- */
  int32_t adj_order0;              /* order of the main multipole term, also when its strength is 0 */
  int32_t adj_pk, adj_pdx, adj_pdy; /* gradient slots of strength, DX, DY (-1: not a parameter) */
  int32_t adj_anchor;              /* slot of the stored entry state (RF cavity), -1 if none */
  double adj_kfac[EH_MAXSUB];      /* d kick[step][0] / d strength */
  double adj_sin_dshr[EH_MAXSUB];  /* sin of the CSBEND sub-step angles */
  /* bends as parameters: adj_pk is the slot of K1, adj_pfse that of FSE; the
   * field per unit K1 and per unit FSE as term lists (i, j, coefficient) in
   * the term pool; adj_fse1 = 1 / (1 + FSE + FSE_DIPOLE), the relative change
   * of 1/rho_actual and of the edge angles psi per unit FSE */
  int32_t adj_pfse;
  int32_t adj_dk_fx0, adj_dk_nfx, adj_dk_fy0, adj_dk_nfy;
  int32_t adj_df_fx0, adj_df_nfx, adj_df_fy0, adj_df_nfy;
  double adj_fse1;
} eh_desc;

typedef struct {
  double x, xp, y, yp, s, dp;
} eh_p6;

EH_FN double eh_sqr(double x) { return x * x; }

EH_HD bool eh_isnan(double x) { return x != x; }
EH_HD bool eh_isinf(double x) { return eh_fabs(x) == __builtin_inf(); }

/* ------------------------------------------------------------------------- */
/* DRIF through drift_matrix + generic order-2 loop (src/compute_matrices.c:391-410,
 * src/matrix.c:370-385).  For finite inputs the generic loop reduces exactly to
 * the expressions below (terms with zero matrix entries add signed zeros, which
 * only normalise -0 to +0 where written as 0.0 + ...); any non-finite input makes
 * every output NaN through 0*inf or 0*NaN. */
EH_HD void eh_drif(eh_p6 &p, double L, double T) {
  bool finite = !(eh_isnan(p.x) || eh_isnan(p.xp) || eh_isnan(p.y) || eh_isnan(p.yp) || eh_isnan(p.s) ||
                  eh_isnan(p.dp) || eh_isinf(p.x) || eh_isinf(p.xp) || eh_isinf(p.y) || eh_isinf(p.yp) ||
                  eh_isinf(p.s) || eh_isinf(p.dp));
  if (!finite) {
    double n = eh_nan();
    p.x = p.xp = p.y = p.yp = p.s = p.dp = n;
    return;
  }
  double x = (0.0 + L * p.xp) + p.x;
  double xp = 0.0 + p.xp;
  double y = (0.0 + L * p.yp) + p.y;
  double yp = 0.0 + p.yp;
  double s = ((L + p.s) + (0.0 + T * p.yp) * p.yp) + (0.0 + T * p.xp) * p.xp;
  double dp = 0.0 + p.dp;
  p.x = x;
  p.xp = xp;
  p.y = y;
  p.yp = yp;
  p.s = s;
  p.dp = dp;
}

/* exactDrift, src/csbend.c:3811-3815 */
EH_HD void eh_edrift(eh_p6 &p, double L) {
  p.x += p.xp * L;
  p.y += p.yp * L;
  p.s += L * eh_sqrt(1 + eh_sqr(p.xp) + eh_sqr(p.yp));
}

/* ------------------------------------------------------------------------- */
/* limit_amplitudes for one particle, src/limit_amplitudes.c:226-284 (openCode 0).
 * Returns true if the particle is lost and applies the loss assignments. */
/* track_particles' loops (src/matrix.c:373-385 for order 2, the
 * TURBO_FASTMATTRACK=3 branch; :438-445 for order 1): zero entries included, in
 * the reference's summation order */
EH_HD void eh_vmatrix_apply(eh_p6 &p, const EH_VMATRIX &M) {
  const double ini[6] = {p.x, p.xp, p.y, p.yp, p.s, p.dp};
  double fin[6];
  if (M.order >= 2) {
    for (int i = 5; i >= 0; i--) {
      double sum = M.C[i];
      for (int j = 5; j >= 0; j--) {
        double sum1 = M.R[i][j];
        const double *Tij = M.T[i] + j * (j + 1) / 2;
        for (int k = j; k >= 0; k--)
          sum1 += Tij[k] * ini[k];
        sum += sum1 * ini[j];
      }
      fin[i] = sum;
    }
  } else {
    for (int i = 5; i >= 0; i--) {
      double sum = M.C[i];
      for (int j = 5; j >= 0; j--)
        sum += M.R[i][j] * ini[j];
      fin[i] = sum;
    }
  }
  p.x = fin[0];
  p.xp = fin[1];
  p.y = fin[2];
  p.yp = fin[3];
  p.s = fin[4];
  p.dp = fin[5];
}

EH_HD bool eh_limit_one(eh_p6 &p, double xmax, double ymax, double z, double Po, int extrap) {
  int is_out = 0;
  if (xmax > 0 && eh_fabs(p.x) > xmax)
    is_out += 1;
  if (ymax > 0 && eh_fabs(p.y) > ymax)
    is_out += 2;
  if (eh_isinf(p.x) || eh_isinf(p.y) || eh_isnan(p.x) || eh_isnan(p.y))
    is_out += 4;
  if (!is_out)
    return false;
  double dz = 0;
  if (!(is_out & 4) && extrap) {
    double dzx = -EH_DBL_MAX, dzy = -EH_DBL_MAX;
    if ((is_out & 1) && p.xp != 0)
      dzx = (p.x - (p.xp < 0 ? -1.0 : 1.0) * xmax) / p.xp;
    if ((is_out & 2) && p.yp != 0)
      dzy = (p.y - (p.yp < 0 ? -1.0 : 1.0) * ymax) / p.yp;
    if (dzx > dzy)
      dz = -dzx;
    else
      dz = -dzy;
    if (dz == -EH_DBL_MAX)
      dz = 0;
  }
  p.x += dz * p.xp;
  p.y += dz * p.yp;
  p.s = z + dz;
  p.dp = Po * (1 + p.dp);
  return true;
}

/* x^p for p >= 0 as the SDDS toolkit's ipow (mdbmth/ipow.c, which elegant
 * links for its elliptical apertures): fixed products of x^2 up to p = 8 (x^4 =
 * x^2 x^2, x^6 = x^2 x^2 x^2, x^7 = x^3 x^3 x, x^8 = (x^2 x^2)(x^2 x^2)), larger
 * p by halving, x^p = h h or h h x with h = x^(p/2), unrolled here from the
 * innermost level outwards; 0^p is 0 (p > 0) or 1.  p is wavefront-uniform. */
EH_HD double eh_ipow(double x, int p) {
  if (x == 0)
    return p == 0 ? 1.0 : 0.0;
  uint32_t bits = 0; /* parities of the halving levels, the innermost in bit 0 */
  int depth = 0;
  while (p > 8) {
    bits = (bits << 1) | (uint32_t)(p & 1);
    p >>= 1;
    depth++;
  }
  double h, r;
  switch (p) {
  case 0:
    r = 1.0;
    break;
  case 1:
    r = x;
    break;
  case 2:
    r = x * x;
    break;
  case 3:
    h = x * x;
    r = h * x;
    break;
  case 4:
    h = x * x;
    r = h * h;
    break;
  case 5:
    h = x * x;
    r = h * h * x;
    break;
  case 6:
    h = x * x;
    r = h * h * h;
    break;
  case 7:
    h = x * x * x;
    r = h * h * x;
    break;
  default:
    h = x * x;
    h = h * h;
    r = h * h;
    break;
  }
  for (int i = 0; i < depth; i++) {
    r = r * r;
    if ((bits >> i) & 1)
      r = r * x;
  }
  return r;
}

/* elimit_amplitudes for one particle, src/limit_amplitudes.c:548-616 (no open
 * side): a non-finite x or y, or x^xe/a2 + y^ye/b2 > 1, loses the particle;
 * with a true ellipse after a drift (extrap) the loss point is found by
 * solving the straight line back to the boundary.  A lost particle's dp
 * becomes sqrt((Po (1 + dp))^2 + 1) and its s the loss position. */
EH_HD bool eh_elimit_one(eh_p6 &p, double a2, double b2, int xe, int ye, double z, double Po, int extrap) {
  if (eh_isinf(p.x) || eh_isinf(p.y) || eh_isnan(p.x) || eh_isnan(p.y)) {
    p.s = z;
    p.dp = eh_sqrt(eh_sqr(Po * (1 + p.dp)) + 1);
    return true;
  }
  if (!((eh_ipow(p.x, xe) / a2 + eh_ipow(p.y, ye) / b2) > 1))
    return false;
  double dz = 0;
  if (extrap && xe == 2 && ye == 2) {
    const double c0 = eh_sqr(p.x) / a2 + eh_sqr(p.y) / b2 - 1;
    const double c1 = 2 * (p.x * p.xp / a2 + p.y * p.yp / b2);
    const double c2 = eh_sqr(p.xp) / a2 + eh_sqr(p.yp) / b2;
    const double det = eh_sqr(c1) - 4 * c0 * c2;
    if (z > 0 && c2 != 0 && det >= 0) {
      if ((dz = (-c1 + eh_sqrt(det)) / (2 * c2)) > 0)
        dz = (-c1 - eh_sqrt(det)) / (2 * c2);
      if ((z + dz) < 0)
        dz = -z;
      p.x += dz * p.xp;
      p.y += dz * p.yp;
    }
  }
  p.s = z + dz;
  p.dp = eh_sqrt(eh_sqr(Po * (1 + p.dp)) + 1);
  return true;
}

/* ------------------------------------------------------------------------- */
/* in-element aperture, checkMultAperture (src/multipole.c:1751-1798): the
 * rectangular limits, or with ELL and ell set the (super-)ellipse
 * (x/xmax)^xe + (y/ymax)^ye >= 1 (ell is set only with xmax, ymax > 0).  ELL is
 * false in the lean instantiation of the maps, which never runs a descriptor
 * with an elliptical aperture. */
template <bool ELL = false>
EH_HD bool eh_ap_ok(double xIn, double yIn, double xmax, double ymax, int rtilt, double cs0, double cs1, int ell = 0,
                    int xe = 0, int ye = 0) {
  double x = xIn - 0.0, y = yIn - 0.0;
  if (rtilt) {
    double x0 = x, y0 = y;
    x = x0 * cs0 + y0 * cs1;
    y = -x0 * cs1 + y0 * cs0;
  }
  if (ELL && ell) {
    const double xa = x / xmax, yb = y / ymax;
    return !((eh_ipow(xa, xe) + eh_ipow(yb, ye)) >= 1);
  }
  if ((xmax > 0 && eh_fabs(x) > xmax) || (ymax > 0 && eh_fabs(y) > ymax))
    return false;
  return true;
}

/* powers x^0..x^3 and y^0..y^3 of fillPowerArray (src/multipole.h:41-46), kept
 * in registers */
struct eh_pow4 {
  double x0, x1, x2, x3, y0, y1, y2, y3;
};

template <int N> EH_HD double eh_px(const eh_pow4 &w) { return N == 0 ? w.x0 : N == 1 ? w.x1 : N == 2 ? w.x2 : w.x3; }
template <int N> EH_HD double eh_py(const eh_pow4 &w) { return N == 0 ? w.y0 : N == 1 ? w.y1 : N == 2 ? w.y2 : w.y3; }

/* coefficient (-1)^floor(i/2) / (i! (n-i)!) of expansion_coefficients
 * (src/multipole.h:76-145) */
template <int ORDER, int I> EH_HD double eh_mcoef() {
  if (ORDER <= 1)
    return 1.0;
  if (ORDER == 2)
    return I == 0 ? 1. / 2 : (I == 1 ? 1.0 : -1. / 2);
  return (((I / 2) & 1) ? -1.0 : 1.0) /
         ((I == 0 ? 1.0 : (I == 1 ? 1.0 : (I == 2 ? 2.0 : 6.0))) *
          ((ORDER - I) == 0 ? 1.0 : ((ORDER - I) == 1 ? 1.0 : ((ORDER - I) == 2 ? 2.0 : 6.0))));
}

template <int ORDER, int I> EH_HD void eh_mterm(const eh_pow4 &w, double &sum_Fx, double &sum_Fy) {
  double f = eh_mcoef<ORDER, I>() * eh_px<ORDER - I>(w) * eh_py<I>(w);
  if (I & 1)
    sum_Fx += f;
  else
    sum_Fy += f;
}

/* apply_canonical_multipole_kicks_ret/_noret, src/multipole.h:391-445: the terms
 * summed in index order as (coef*x^(n-i))*y^i */
template <int ORDER>
EH_HD void eh_mkick(double &qx, double &qy, const eh_pow4 &w, double K, int skew) {
  double sum_Fx = 0, sum_Fy = 0;
  eh_mterm<ORDER, 0>(w, sum_Fx, sum_Fy);
  if (ORDER >= 1)
    eh_mterm<ORDER, (ORDER >= 1 ? 1 : 0)>(w, sum_Fx, sum_Fy);
  if (ORDER >= 2)
    eh_mterm<ORDER, (ORDER >= 2 ? 2 : 0)>(w, sum_Fx, sum_Fy);
  if (ORDER >= 3)
    eh_mterm<ORDER, (ORDER >= 3 ? 3 : 0)>(w, sum_Fx, sum_Fy);
  if (skew) {
    double t = sum_Fx;
    sum_Fx = sum_Fy;
    sum_Fy = t;
    sum_Fx = -sum_Fx;
  }
  qx -= K * sum_Fy;
  qy += K * sum_Fx;
}

/* the kick with its increments accumulated into dqx, dqy
 * (apply_canonical_multipole_kicks_ret, src/multipole.h:391-420) */
template <int ORDER>
EH_HD void eh_mkick_ret(double &qx, double &qy, double &dqx, double &dqy, const eh_pow4 &w, double K, int skew) {
  double sum_Fx = 0, sum_Fy = 0;
  eh_mterm<ORDER, 0>(w, sum_Fx, sum_Fy);
  if (ORDER >= 1)
    eh_mterm<ORDER, (ORDER >= 1 ? 1 : 0)>(w, sum_Fx, sum_Fy);
  if (ORDER >= 2)
    eh_mterm<ORDER, (ORDER >= 2 ? 2 : 0)>(w, sum_Fx, sum_Fy);
  if (ORDER >= 3)
    eh_mterm<ORDER, (ORDER >= 3 ? 3 : 0)>(w, sum_Fx, sum_Fy);
  if (skew) {
    double t = sum_Fx;
    sum_Fx = sum_Fy;
    sum_Fy = t;
    sum_Fx = -sum_Fx;
  }
  qx -= K * sum_Fy;
  qy += K * sum_Fx;
  dqx -= K * sum_Fy;
  dqy += K * sum_Fx;
}

EH_HD void eh_mkick_ret_n(int order, double &qx, double &qy, double &dqx, double &dqy, const eh_pow4 &w, double K,
                          int skew) {
  switch (order) {
  case 0:
    eh_mkick_ret<0>(qx, qy, dqx, dqy, w, K, skew);
    break;
  case 1:
    eh_mkick_ret<1>(qx, qy, dqx, dqy, w, K, skew);
    break;
  case 2:
    eh_mkick_ret<2>(qx, qy, dqx, dqy, w, K, skew);
    break;
  default:
    eh_mkick_ret<3>(qx, qy, dqx, dqy, w, K, skew);
    break;
  }
}

EH_HD void eh_mkick_n(int order, double &qx, double &qy, const eh_pow4 &w, double K, int skew) {
  switch (order) {
  case 0:
    eh_mkick<0>(qx, qy, w, K, skew);
    break;
  case 1:
    eh_mkick<1>(qx, qy, w, K, skew);
    break;
  case 2:
    eh_mkick<2>(qx, qy, w, K, skew);
    break;
  default:
    eh_mkick<3>(qx, qy, w, K, skew);
    break;
  }
}

/* convertMomentaToSlopes, src/multipole.h:466-486 */
EH_HD bool eh_m2s(double &xp, double &yp, double qx, double qy, double dp, int expand) {
  if (expand) {
    xp = qx / (1 + dp);
    yp = qy / (1 + dp);
    return true;
  }
  double denom = eh_sqr(1 + dp) - eh_sqr(qx) - eh_sqr(qy);
  if (denom <= 0)
    return false;
  denom = eh_sqrt(denom);
  xp = qx / denom;
  yp = qy / denom;
  return true;
}

/* the reciprocal pair of b for eh_round_trip, formed per particle: rh = RN(1/b)
 * by a correctly rounded division and rl = RN(e*rh) with e = 1 - b*rh exact (one
 * fma).  rh + rl is then within ~2^-104 of 1/b relative, enough for the first
 * quotient of Markstein's method to be within one ulp.  False (use a division)
 * unless b lies in [2^-500, 2^500]. */
EH_HD bool eh_recip_pair(double b, double &rh, double &rl) {
  const double ab = eh_fabs(b);
  if (!(ab >= 0x1p-500 && ab <= 0x1p500))
    return false;
  rh = 1.0 / b;
  rl = eh_fma(-b, rh, 1.0) * rh;
  return true;
}

/* RN(RN(a / b) * b): what the reference's radiation block leaves on a canonical
 * momentum when no radiation is on (it divides by 1 + delta and multiplies back,
 * src/multipole.c:1362-1385, src/csbend.c:1645-1658).  The quotient is correctly
 * rounded through the pair (Markstein, as eh_div_recip); a zero numerator comes
 * back unchanged, as a/b*b does for finite non-zero b; other numerators outside
 * [2^-960, 2^960], or no pair, take the division. */
EH_HD double eh_round_trip(double a, double b, double rh, double rl, bool ok) {
  const double aa = eh_fabs(a);
  if (ok && aa >= 0x1p-960 && aa <= 0x1p960) {
    const double q0 = eh_fma(a, rh, a * rl);
    const double r = eh_fma(-q0, b, a);
    return eh_fma(r, rh, q0) * b;
  }
  if (a == 0)
    return a;
#if defined(__HIP_DEVICE_COMPILE__)
  __asm__ volatile("");
#endif
  return a / b * b;
}

/* entry face of a misaligned multipole: offsetBeamCoordinatesForMisalignment
 * (dx, dy, dz) then rotateBeamCoordinatesForMisalignment (tilt) */
EH_HD void eh_mface1(eh_p6 &p, const eh_desc &d) {
  if (d.has_offset) {
    p.s += d.mdz * eh_sqrt(1 + eh_sqr(p.xp) + eh_sqr(p.yp));
    p.x = p.x - d.mdx + d.mdz * p.xp;
    p.y = p.y - d.mdy + d.mdz * p.yp;
  }
  if (d.has_tilt) {
    double x = p.x, xp = p.xp, y = p.y, yp = p.yp;
    p.x = x * d.ncos + y * d.nsin;
    p.y = -x * d.nsin + y * d.ncos;
    p.xp = xp * d.ncos + yp * d.nsin;
    p.yp = -xp * d.nsin + yp * d.ncos;
  }
}

/* exit face: rotation by -tilt, then the offset by (-dx, -dy, -dz); every row,
 * lost ones included.  With s_too false the path-length term of the offset is
 * left out (the host adds it to the loss position it settles) */
EH_HD void eh_mface2(eh_p6 &p, const eh_desc &d, bool s_too) {
  if (d.has_offset || d.has_tilt) {
    if (d.has_tilt) {
      double x0 = p.x, xp0 = p.xp, y0 = p.y, yp0 = p.yp;
      p.x = x0 * d.xcos + y0 * d.xsin;
      p.y = -x0 * d.xsin + y0 * d.xcos;
      p.xp = xp0 * d.xcos + yp0 * d.xsin;
      p.yp = -xp0 * d.xsin + yp0 * d.xcos;
    }
    if (d.has_offset) {
      if (s_too)
        p.s += (-d.mdz) * eh_sqrt(1 + eh_sqr(p.xp) + eh_sqr(p.yp));
      p.x = p.x - (-d.mdx) + (-d.mdz) * p.xp;
      p.y = p.y - (-d.mdy) + (-d.mdz) * p.yp;
    }
  }
}

/* KQUAD/KSEXT/KOCT/KQUSE: multipole_tracking2 (src/multipole.c:660-1084) for one
 * particle, including the misalignment passes around the particle loop (method 0,
 * src/malign_mat.c:257-267, :580-590, tilt_matrices.c:156-187) and the loss
 * assignments of :1032-1047.  integrate_kick_multipole_ordn, src/multipole.c:1099-1456.
 * Returns 0 if the particle survives, 1 if lost, 2 if lost before the first
 * step (its loss position is z_start plus the previous particle's dzLoss).  A
 * lost particle keeps s = z_start + dzLoss without the exit offset's path-length
 * term.  RAD: the element radiates (d.rad); FULL: the options only the full
 * instantiation of the maps carries (elliptical apertures; for bends also edge
 * method 3 and edge kick limits) are compiled in; the lean one has neither. */
template <bool RAD, bool FULL = RAD>
EH_HD int eh_multipole_t(eh_p6 &p, const eh_desc &d, int pi, const double *__restrict__ zp) {
  eh_mface1(p, d);
  const double apx = d.ap_xmax[pi], apy = d.ap_ymax[pi];
  const int ell = FULL ? d.ap_ell[pi] : 0;
  double x = p.x, xp = p.xp, y = p.y, yp = p.yp, s = 0, dp = p.dp, qx, qy;
  double beta0 = 0;
  if constexpr (RAD) {
    const double P = d.Po * (1 + dp);
    beta0 = P / eh_sqrt(eh_sqr(P) + 1);
  }
  double dzLoss = 0;
  bool lost = false, entry = false;
  int how = 0; /* 0: coord untouched, 1: x,y written, 2: all written */
  if (eh_isnan(x) || eh_isnan(xp) || eh_isnan(y) || eh_isnan(yp) || eh_fabs(x) > 10.0 || eh_fabs(y) > 10.0 ||
      eh_fabs(xp) > 1.0 || eh_fabs(yp) > 1.0) {
    lost = entry = true;
  } else {
    /* convertSlopesToMomenta, src/multipole.h:448-464 */
    if (d.expand) {
      qx = (1 + dp) * xp;
      qy = (1 + dp) * yp;
    } else {
      double denom = eh_sqrt(1 + eh_sqr(xp) + eh_sqr(yp));
      qx = (1 + dp) * xp / denom;
      qy = (1 + dp) * yp / denom;
    }
    /* 1 + delta and its reciprocal pair for the radiation block's round trip */
    const double b1 = 1 + dp;
    double rh = 0, rl = 0;
    const bool rok = d.mdrift != 0 && eh_recip_pair(b1, rh, rl);
    if (!eh_m2s(xp, yp, qx, qy, dp, d.expand)) {
      lost = entry = true;
    } else {
      eh_pow4 w;
      for (int ik = 0; ik < d.nslices && !lost; ik++) {
        if (!eh_ap_ok<FULL>(x + d.mdx, y + d.mdy, apx, apy, d.ap_rtilt, d.ap_cs0, d.ap_cs1, ell, d.ell_xe[pi],
                           d.ell_ye[pi])) {
          lost = true;
          how = 1;
          break;
        }
        for (int step = 0; step < d.nsub; step++) {
          if (d.mdrift != 0) {
            double dsh = d.dsh[step];
            x += xp * dsh;
            y += yp * dsh;
            if (d.expand)
              s += dsh * (1 + (eh_sqr(xp) + eh_sqr(yp)) / 2);
            else
              s += dsh * eh_sqrt(1 + eh_sqr(xp) + eh_sqr(yp));
            dzLoss += dsh;
          }
          if (step == d.nsub - 1)
            break;
          w.x0 = 1;
          w.y0 = 1;
          w.x1 = w.x0 * x;
          w.y1 = w.y0 * y;
          w.x2 = w.x1 * x;
          w.y2 = w.y1 * y;
          w.x3 = w.x2 * x;
          w.y3 = w.y2 * y;
          double dqx = 0, dqy = 0;
          for (int io = 0; io < 3; io++)
            if (d.mp_order[io]) {
              if constexpr (RAD)
                eh_mkick_ret_n(d.mp_order[io], qx, qy, dqx, dqy, w, d.kick[step][io], d.mp_skew[io]);
              else
                eh_mkick_n(d.mp_order[io], qx, qy, w, d.kick[step][io], d.mp_skew[io]);
            }
          if (d.has_xkick)
            eh_mkick<0>(qx, qy, w, d.xk[step], 0);
          if (d.has_ykick)
            eh_mkick<0>(qx, qy, w, d.yk[step], 1);
          if constexpr (RAD) {
            /* convertMomentaToSlopes, then the radiation block with rad_coef
             * (src/multipole.c:1356-1390): the kick increments without
             * kickFrac give the field, the energy loss follows */
            if (!eh_m2s(xp, yp, qx, qy, dp, d.expand)) {
              lost = true;
              break;
            }
            const double drift = d.mdrift, kf = d.kfrac[step];
            qx /= (1 + dp);
            qy /= (1 + dp);
            const double deltaFactor = eh_sqr(1 + dp);
            dqx /= kf;
            dqy /= kf;
            const double F2 = eh_sqr(dqx / drift - d.rad_xk / drift) + eh_sqr(dqy / drift + d.rad_yk / drift);
            double dsFactor = eh_sqrt(1 + eh_sqr(xp) + eh_sqr(yp));
            dsFactor *= drift * kf;
            dp -= d.rad_coef * deltaFactor * F2 * dsFactor;
            qx *= (1 + dp);
            qy *= (1 + dp);
          } else if (d.mdrift != 0) {
            /* convertMomentaToSlopes, whose slopes the radiation block's second
             * conversion replaces: only its failure test remains; then the
             * block's round trip of the momenta (src/multipole.c:1356-1390; it
             * runs for any element with length: isr_coef is non-zero even
             * without ISR, :757-760) */
            if (!d.expand && eh_sqr(1 + dp) - eh_sqr(qx) - eh_sqr(qy) <= 0) {
              lost = true;
              break;
            }
            qx = eh_round_trip(qx, b1, rh, rl, rok);
            qy = eh_round_trip(qy, b1, rh, rl, rok);
          }
          if (!eh_m2s(xp, yp, qx, qy, dp, d.expand)) {
            lost = true;
            break;
          }
        }
      }
      if (!lost) {
        if (!eh_ap_ok<FULL>(x + d.mdx, y + d.mdy, apx, apy, d.ap_rtilt, d.ap_cs0, d.ap_cs1, ell, d.ell_xe[pi],
                           d.ell_ye[pi])) {
          lost = true;
          how = 1;
        } else if (!eh_m2s(xp, yp, qx, qy, dp, d.expand)) {
          lost = true;
        } else {
          how = 2;
          if (eh_isnan(x) || eh_isnan(xp) || eh_isnan(y) || eh_isnan(yp) || eh_fabs(x) > 10.0 ||
              eh_fabs(y) > 10.0 || eh_fabs(xp) > 1.0 || eh_fabs(yp) > 1.0)
            lost = true;
        }
      }
    }
  }
  if (how == 1) {
    p.x = x;
    p.y = y;
  } else if (how == 2) {
    p.x = x;
    p.xp = xp;
    p.y = y;
    p.yp = yp;
    if constexpr (RAD) {
      /* src/multipole.c:1436-1440 */
      const double P = d.Po * (1 + dp);
      const double beta1 = P / eh_sqrt(eh_sqr(P) + 1);
      p.s = beta1 * (p.s / beta0 + 2 * s / (beta0 + beta1));
    } else {
      p.s += s;
    }
    p.dp = dp;
  }
  if (lost) {
    p.s = zp[0] + dzLoss;
    p.dp = d.Po * (1 + p.dp);
  }
  eh_mface2(p, d, !lost);
  return lost ? (entry ? 2 : 1) : 0;
}

EH_HD int eh_multipole(eh_p6 &p, const eh_desc &d, int pi, const double *__restrict__ zp) {
  return d.rad ? eh_multipole_t<true, true>(p, d, pi, zp) : eh_multipole_t<false, true>(p, d, pi, zp);
}

/* ------------------------------------------------------------------------- */
/* the field sums of one MULT kick of order N: fillPowerArray for x and y, then
 * the terms (coef[i]*x^(N-i))*y^i summed in index order, odd i into Fx, even i
 * into Fy (src/multipole.c:578-586); the kick is qx -= K*Fy, qy += K*Fx */
template <int N>
EH_HD void eh_mult_sums(double x, double y, const double *coef, double &sum_Fx, double &sum_Fy) {
  double xp[N + 1], yp[N + 1];
  xp[0] = 1;
  yp[0] = 1;
#pragma unroll
  for (int i = 1; i <= N; i++) {
    xp[i] = xp[i - 1] * x;
    yp[i] = yp[i - 1] * y;
  }
  sum_Fx = 0;
  sum_Fy = 0;
#pragma unroll
  for (int i = 0; i <= N; i++) {
    if (i & 1)
      sum_Fx += coef[i] * xp[N - i] * yp[i];
    else
      sum_Fy += coef[i] * xp[N - i] * yp[i];
  }
}

EH_HD void eh_mult_sums_n(int order, double x, double y, const double *coef, double &Fx, double &Fy) {
  switch (order) {
  case 0:
    eh_mult_sums<0>(x, y, coef, Fx, Fy);
    break;
  case 1:
    eh_mult_sums<1>(x, y, coef, Fx, Fy);
    break;
  case 2:
    eh_mult_sums<2>(x, y, coef, Fx, Fy);
    break;
  case 3:
    eh_mult_sums<3>(x, y, coef, Fx, Fy);
    break;
  case 4:
    eh_mult_sums<4>(x, y, coef, Fx, Fy);
    break;
  case 5:
    eh_mult_sums<5>(x, y, coef, Fx, Fy);
    break;
  case 6:
    eh_mult_sums<6>(x, y, coef, Fx, Fy);
    break;
  case 7:
    eh_mult_sums<7>(x, y, coef, Fx, Fy);
    break;
  default:
    eh_mult_sums<8>(x, y, coef, Fx, Fy);
    break;
  }
}

/* the coordinate and slope limits of multipole_tracking (coordLimit 10 m,
 * slopeLimit 1), NaN included */
EH_HD bool eh_mult_out(double x, double xp, double y, double yp) {
  return eh_isnan(x) || eh_isnan(xp) || eh_isnan(y) || eh_isnan(yp) || eh_fabs(x) > 10.0 || eh_fabs(y) > 10.0 ||
         eh_fabs(xp) > 1.0 || eh_fabs(yp) > 1.0;
}

/* MULT: multipole_tracking (src/multipole.c:439-658) for one particle, without
 * BORE/BTIPL, with classical radiation.  n_slices kicks of strength kick[0][0]
 * (KnL*FACTOR/N_SLICES) between drifts of mdrift (L/N_SLICES/2, doubled between
 * kicks).  A particle outside the limits on entry keeps its coordinates; one
 * lost in the kicks keeps the values reached; both get s = z_start and
 * delta = Po*(1+delta).  Zero strength is an exact drift, set up by the host.
 * Returns true if lost. */
template <bool RAD> EH_HD bool eh_mult_t(eh_p6 &p, const eh_desc &d, const double *__restrict__ zp) {
  eh_mface1(p, d);
  const double drift = d.mdrift;
  double x = p.x, xp = p.xp, y = p.y, yp = p.yp, s = 0, dp = p.dp, qx, qy;
  const double P0 = d.Po * (1 + dp);
  const double beta0 = RAD && d.rad ? P0 / eh_sqrt(eh_sqr(P0) + 1) : 0.0;
  bool lost;
  if (eh_mult_out(x, xp, y, yp)) {
    lost = true;
  } else {
    if (d.expand) {
      qx = (1 + dp) * xp;
      qy = (1 + dp) * yp;
    } else {
      double denom = eh_sqrt(1 + eh_sqr(xp) + eh_sqr(yp));
      qx = (1 + dp) * xp / denom;
      qy = (1 + dp) * yp / denom;
    }
    bool is_lost = false;
    for (int ik = 0; ik < d.nslices; ik++) {
      if (drift) {
        x += xp * drift * (ik ? 2 : 1);
        y += yp * drift * (ik ? 2 : 1);
        if (d.expand)
          s += (ik ? 2 : 1) * drift * (1 + (eh_sqr(xp) + eh_sqr(yp)) / 2);
        else
          s += (ik ? 2 : 1) * drift * eh_sqrt(1 + eh_sqr(xp) + eh_sqr(yp));
      }
      double sum_Fx, sum_Fy;
      const double K = d.kick[0][0];
      eh_mult_sums_n(d.mp_order[0], x, y, d.mult_coef, sum_Fx, sum_Fy);
      qx -= K * sum_Fy;
      qy += K * sum_Fx;
      if (!eh_m2s(xp, yp, qx, qy, dp, d.expand)) {
        is_lost = true;
        break;
      }
      if (RAD && d.rad && drift) {
        /* classical radiation, src/multipole.c:599-611 */
        qx /= (1 + dp);
        qy /= (1 + dp);
        dp -= d.rad_coef * eh_sqr(K * (1 + dp)) * (eh_sqr(sum_Fy) + eh_sqr(sum_Fx)) *
              eh_sqrt(1 + eh_sqr(xp) + eh_sqr(yp)) / (2 * drift);
        qx *= (1 + dp);
        qy *= (1 + dp);
        if (!eh_m2s(xp, yp, qx, qy, dp, d.expand)) {
          is_lost = true;
          break;
        }
      }
    }
    if (drift && !is_lost) {
      x += xp * drift;
      y += yp * drift;
      if (d.expand)
        s += drift * (1 + (eh_sqr(xp) + eh_sqr(yp)) / 2);
      else
        s += drift * eh_sqrt(1 + eh_sqr(xp) + eh_sqr(yp));
    }
    p.x = x;
    p.xp = xp;
    p.y = y;
    p.yp = yp;
    p.dp = dp;
    if (RAD && d.rad) {
      /* src/multipole.c:623-628 */
      const double P1 = d.Po * (1 + dp);
      const double beta1 = P1 / eh_sqrt(eh_sqr(P1) + 1);
      p.s = beta1 * (p.s / beta0 + 2 * s / (beta0 + beta1));
    } else {
      p.s += s;
    }
    lost = is_lost || eh_mult_out(x, xp, y, yp);
  }
  if (lost) {
    p.s = zp[0];
    p.dp = d.Po * (1 + p.dp);
  }
  eh_mface2(p, d, true);
  return lost;
}

EH_HD bool eh_mult(eh_p6 &p, const eh_desc &d, const double *__restrict__ zp) { return eh_mult_t<true>(p, d, zp); }

/* one row of the order-1 loop of track_particles (src/matrix.c:424-460):
 * C[i] (zero here) plus R[i][j]*ini[j] for j = 5 down to 0, zero entries
 * included, so that a non-finite coordinate makes every output NaN as there */
EH_HD double eh_mrow(double r0, double r1, double r2, double r3, double r4, double r5, const eh_p6 &p) {
  double sum = 0.0;
  sum += r5 * p.dp;
  sum += r4 * p.s;
  sum += r3 * p.yp;
  sum += r2 * p.y;
  sum += r1 * p.xp;
  sum += r0 * p.x;
  return sum;
}

/* ROTATE: rotation_matrix(tilt) (src/tilt_matrices.c:81-117) applied through
 * track_particles; c and s are its cos and sin entries */
EH_HD void eh_rotate(eh_p6 &p, double c, double s) {
  const double z = 0.0, ms = -s;
  eh_p6 q;
  q.dp = eh_mrow(z, z, z, z, z, 1.0, p);
  q.s = eh_mrow(z, z, z, z, 1.0, z, p);
  q.yp = eh_mrow(z, ms, z, c, z, z, p);
  q.y = eh_mrow(ms, z, c, z, z, z, p);
  q.xp = eh_mrow(z, c, z, s, z, z, p);
  q.x = eh_mrow(c, z, s, z, z, z, p);
  p = q;
}

/* MALIGN: offset_beam (src/do_tracking.c:3311-3372) for one particle */
EH_HD void eh_offset(eh_p6 &p, const eh_desc &d) {
  const double c_mks = 2.99792458e8;
  const double Pc = d.Po;
  double ds;
  if (d.mdz)
    ds = d.mdz * eh_sqrt(1 + eh_sqr(p.xp) + eh_sqr(p.yp));
  else
    ds = 0;
  p.x += d.mdx + d.mdz * p.xp;
  p.xp += d.mal_dxp;
  p.y += d.mdy + d.mdz * p.yp;
  p.yp += d.mal_dyp;
  p.s += ds;
  if (d.mal_energy) {
    double pc = Pc * (1 + p.dp);
    double gamma = eh_sqrt(1 + pc * pc);
    double beta = pc / gamma;
    double t = p.s / (beta * c_mks) + d.mal_dt;
    if (d.mal_dp) {
      p.dp += d.mal_dp;
      pc = Pc * (1 + p.dp);
      beta = pc / eh_sqrt(1 + pc * pc);
    }
    if (d.mal_de) {
      gamma += d.mal_de * gamma;
      pc = eh_sqrt(gamma * gamma - 1);
      beta = pc / gamma;
      p.dp = (pc - Pc) / Pc;
    }
    p.s = t * beta * c_mks;
  }
}

/* ------------------------------------------------------------------------- */
/* y^0 .. y^10 as computeCSBENDFields builds them (yp[j] = yp[j-1]*y), in
 * registers; only the powers up to maxj are formed */
struct eh_ypow {
  double p[11];
};

EH_HD void eh_ypow_fill(eh_ypow &w, double y, int maxj) {
  w.p[0] = 1;
#pragma unroll
  for (int j = 1; j <= 10; j++)
    if (j <= maxj)
      w.p[j] = w.p[j - 1] * y;
}

/* the power for a wave-uniform exponent: a scalar branch, no indexed register */
EH_HD double eh_ypick(const eh_ypow &w, int j) {
  switch (j) {
  case 0:
    return w.p[0];
  case 1:
    return w.p[1];
  case 2:
    return w.p[2];
  case 3:
    return w.p[3];
  case 4:
    return w.p[4];
  case 5:
    return w.p[5];
  case 6:
    return w.p[6];
  case 7:
    return w.p[7];
  case 8:
    return w.p[8];
  case 9:
    return w.p[9];
  default:
    return w.p[10];
  }
}

/* computeCSBENDFields for expansion order EO1-1 without skew terms, as the
 * reference evaluates it (src/csbend.c:141-166): every (i, j) with i + j < EO1,
 * Fx taking odd j and Fy even j, in i-outer j-inner order, each term
 * (coef * x^i) * y^j.  Fully unrolled: no indexed registers, no data-dependent
 * branches.  cx/cy hold the dense EO1 x EO1 coefficient blocks (row i, column
 * j).  A term whose coefficient is zero adds a signed zero to a sum that starts
 * at +0, which never changes a finite sum, so it is skipped: bit t of `live`
 * says whether term t is non-zero.  The test is the same for every lane (a
 * scalar branch); the empty volatile asm keeps the compiler from turning it
 * into per-lane selects that would evaluate every term anyway. */
template <int EO1>
EH_HD void eh_csb_fields_dense(const eh_fterm *__restrict__ cx, const eh_fterm *__restrict__ cy, uint32_t live,
                               double x, double y, double &Fx, double &Fy) {
  double yp[EO1];
  yp[0] = 1;
#pragma unroll
  for (int j = 1; j < EO1; j++)
    yp[j] = yp[j - 1] * y;
  double sumFx = 0, sumFy = 0, xt = 1;
#pragma unroll
  for (int i = 0; i < EO1; i++) {
    /* the row's coefficients are loaded together, ahead of the branches */
    double c[EO1];
#pragma unroll
    for (int j = 0; j < EO1 - i; j++)
      c[j] = (j & 1) ? cx[i * EO1 + j].coef : cy[i * EO1 + j].coef;
#pragma unroll
    for (int j = 0; j < EO1 - i; j++) {
      const int t = i * EO1 - i * (i - 1) / 2 + j;
      if (!((live >> t) & 1u))
        continue;
#if defined(__HIP_DEVICE_COMPILE__)
      __asm__ volatile("");
#endif
      if (j & 1)
        sumFx += c[j] * xt * yp[j];
      else
        sumFy += c[j] * xt * yp[j];
    }
    xt *= x;
  }
  Fx = sumFx;
  Fy = sumFy;
}

/* computeCSBENDFields, src/csbend.c:99-172.  The reference adds
 * coef * x^i * y^j over all (i, j) in i-outer, j-inner order; terms with a zero
 * coefficient are skipped here, which leaves every finite sum unchanged (adding
 * a signed zero to a sum that starts at +0 never changes it). */
EH_HD void eh_csb_fields(const eh_desc &d, const eh_fterm *__restrict__ terms, double x, double y, double &Fx,
                         double &Fy) {
  if (!d.csb_field) {
    Fx = 0;
    Fy = 1;
    return;
  }
  if (d.csb_dense == 5) {
    eh_csb_fields_dense<5>(terms + d.csb_dx0, terms + d.csb_dy0, d.csb_live, x, y, Fx, Fy);
    return;
  }
  if (d.csb_dense == 6) {
    eh_csb_fields_dense<6>(terms + d.csb_dx0, terms + d.csb_dy0, d.csb_live, x, y, Fx, Fy);
    return;
  }
  eh_ypow w;
  eh_ypow_fill(w, y, d.csb_maxj);
  double sumFx = 0, sumFy = 0;
  double xt = 1;
  int ic = 0;
  const eh_fterm *fx = terms + d.csb_fx0, *fy = terms + d.csb_fy0;
  for (int t = 0; t < d.csb_nfx; t++) {
    int ti = fx[t].i;
    while (ic < ti) {
      xt *= x;
      ic++;
    }
    sumFx += fx[t].coef * xt * eh_ypick(w, fx[t].j);
  }
  xt = 1;
  ic = 0;
  for (int t = 0; t < d.csb_nfy; t++) {
    int ti = fy[t].i;
    while (ic < ti) {
      xt *= x;
      ic++;
    }
    sumFy += fy[t].coef * xt * eh_ypick(w, fy[t].j);
  }
  Fx = sumFx;
  Fy = sumFy;
}

/* a / b, correctly rounded, for a divisor known ahead through rh = RN(1/b) and
 * rl = RN(1/b - rh).  q0 = RN(a*rh + RN(a*rl)) is within one ulp of a/b (the
 * pair carries 1/b to ~2^-105), the remainder r = a - b*q0 is then exact (one
 * fma), and RN(q0 + r*rh) is the correctly rounded quotient: Markstein's
 * theorem (P. Markstein, IBM J. Res. Develop. 34 (1990) 111-119) for a
 * reciprocal within half an ulp and a quotient within one ulp, without
 * underflow or overflow.  Numerators outside [2^-960, 2^960] in magnitude
 * (zeros, whose sign the formula would not keep, non-finite values, extremes)
 * take the ordinary division; the empty volatile asm keeps that division a
 * branch taken only by such lanes rather than work every lane does. */
EH_HD double eh_div_recip(double a, double b, double rh, double rl) {
  if (eh_fabs(a) >= 0x1p-960 && eh_fabs(a) <= 0x1p960) {
    double q0 = eh_fma(a, rh, a * rl);
    double r = eh_fma(-q0, b, a);
    return eh_fma(r, rh, q0);
  }
#if defined(__HIP_DEVICE_COMPILE__)
  __asm__ volatile("");
#endif
  return a / b;
}

/* apply_edge_effects, src/csbend.c:4291-4332: K. L. Brown's second-order edge
 * map; which = -1 at the entrance, 1 at the exit; tan_beta and sec_beta are
 * tan(beta) and 1./cos(beta) as the routine forms them */
EH_HD void eh_edge2(double &x, double &xp, double &y, double &yp, double rho, double n, double beta,
                    double tan_beta, double sec_beta, double he, double psi, int which) {
  const double h = 1 / rho;
  const double R21 = h * tan_beta;
  const double R43 = -h * eh_tan(beta - psi);
  const double h2 = eh_sqr(h);
  const double tan2_beta = eh_sqr(tan_beta);
  const double w = (double)which, mw = (double)(-which);
  const double T111 = w * h / 2 * tan2_beta;
  const double sec2_beta = eh_sqr(sec_beta);
  const double T133 = mw * h / 2 * sec2_beta;
  double T211 = which == -1 ? -n * h2 * tan_beta : -h2 * (n + tan2_beta / 2) * tan_beta;
  const double T221 = mw * h * tan2_beta, T331 = T221, T441 = -T221;
  double T233 = which == -1 ? h2 * (n + .5 + tan2_beta) * tan_beta : h2 * (n - tan2_beta / 2) * tan_beta;
  const double T243 = w * h * tan2_beta;
  double T431 = h2 * (2 * n + (which == 1 ? sec2_beta : 0)) * tan_beta;
  const double T432 = w * h * sec2_beta;
  if (he != 0) {
    const double term = h / 2 * he * sec2_beta * sec_beta;
    T211 += term;
    T233 -= term;
    T431 -= 2 * term;
  }
  const double x0 = x, xp0 = xp, y0 = y, yp0 = yp;
  x = x0 + T111 * eh_sqr(x0) + T133 * eh_sqr(y0);
  xp = xp0 + R21 * x0 + T211 * eh_sqr(x0) + T221 * x0 * xp0 + T233 * eh_sqr(y0) + T243 * y0 * yp0;
  y = y0 + T331 * x0 * y0;
  yp = yp0 + R43 * y0 + T441 * yp0 * x0 + T431 * x0 * y0 + T432 * xp0 * y0;
}

/* EDGE_EFFECTS 3, applySimpleDipoleEdgeKick (src/csbend.c:5664-5690): the
 * first-order edge kick applied to the canonical momenta (dq_x = tan(e)/rho x,
 * limited to klim if positive; dq_y = -tan(e - psi/(1+dp))/rho y), with rho the
 * bend's rho_actual, not scaled by 1 + dp */
EH_HD void eh_edge_simple(double &xp, double &yp, double x, double y, double dp, double rho, double tan_e, double e,
                          double psi, double klim) {
  double f = (1 + dp) / eh_sqrt(1 + eh_sqr(xp) + eh_sqr(yp));
  double qx = xp * f, qy = yp * f;
  double dqx = tan_e / rho * x;
  if (klim > 0 && eh_fabs(dqx) > klim)
    dqx = (dqx < 0 ? -1.0 : 1.0) * klim;
  double dqy = -eh_tan(e - psi / (1 + dp)) / rho * y;
  qx += dqx;
  qy += dqy;
  f = 1 / eh_sqrt(eh_sqr(1 + dp) - eh_sqr(qx) - eh_sqr(qy));
  xp = qx * f;
  yp = qy * f;
}

/* dipoleFringeKHwang (src/csbend.c:4336-4410, EDGE_EFFECTS 2) on the canonical
 * momenta: K. Hwang's dipole fringe map at the entrance (in_fringe -1) or the
 * exit (+1), with c = {1/cos e, tan e, sin e, cos e, gap, FINT, H} and the
 * higher-order terms when higher; the reference's operation order, SDDS's
 * ipow2/ipow3 as x*x and (x*x)*x */
EH_HD void eh_fringe_hwang(double &x, double &px, double &y, double &py, double dp0, double rho,
                           const double *__restrict__ c, double in_fringe, bool higher) {
  const double k0 = (EH_PI_D * EH_PI_D) / 6., k2 = c[5], k3 = 1.0 * 1. / 6., k4 = -1.0 * (EH_PI_D * EH_PI_D) / 3.,
               k5 = 0.0, k6 = -1.0;
  const double sec_e = c[0], tan_e = c[1], sin_e = c[2], cos_e = c[3], gap = c[4], Rhe = c[6];
  const double x0 = x, px0 = px, y0 = y, py0 = py;
  const double sec2 = sec_e * sec_e, tan2 = tan_e * tan_e, gap2 = gap * gap, cos3 = cos_e * cos_e * cos_e;
  const double sin2 = sin_e * sin_e, rho2 = rho * rho;
  double dx, dpx, dy, dpy;
  if (higher) {
    dx = in_fringe * sec2 * gap2 * k0 / rho / (1 + dp0) + in_fringe * (x0 * x0) * tan2 / 2 / rho / (1 + dp0) -
         in_fringe * (y0 * y0) * sec2 / 2 / rho / (1 + dp0);
    dy = -in_fringe * x0 * y0 * tan2 / rho / (1 + dp0);
    if (in_fringe == -1.) {
      dpx = -1. * (sec_e * sec_e * sec_e) * sin_e * gap2 * k0 / rho / rho / (1 + dp0) + tan_e * x0 / rho +
            (y0 * y0) / 2 * (2 * (tan_e * tan_e * tan_e)) / rho2 / (1 + dp0) + (y0 * y0) / 2 * (tan_e) / rho2 / (1 + dp0) -
            in_fringe * (x0 * px0 - y0 * py0) * tan2 / rho / (1 + dp0) + k4 * sin2 * gap2 / 2 / cos3 / rho * Rhe -
            k5 * x0 * sin_e * gap / cos3 / rho * Rhe + k6 * (y0 * y0 - x0 * x0) / 2 / cos3 / rho * Rhe;
      dpy = -1. * tan_e * y0 / rho + k2 * y0 * (1 + sin2) * gap / (1 + dp0) / rho2 / cos3 +
            in_fringe * (x0 * py0 + y0 * px0) * tan2 / rho / (1 + dp0) + in_fringe * y0 * px0 / rho / (1 + dp0) +
            k3 * (y0 * y0 * y0) * (2. / 3. / cos_e - 4. / 3. / cos3) / (1 + dp0) / rho / rho / gap +
            k6 * x0 * y0 / cos3 / rho * Rhe;
    } else {
      dpx = tan_e * x0 / rho - (y0 * y0) / 2 * (1 * (tan_e * tan_e * tan_e)) / rho2 / (1 + dp0) -
            (x0 * x0) / 2 * (1 * (tan_e * tan_e * tan_e)) / rho2 / (1 + dp0) -
            in_fringe * (x0 * px0 - y0 * py0) * tan2 / rho / (1 + dp0) + k4 * sin2 * gap2 / 2 / cos3 / rho * Rhe -
            k5 * x0 * sin_e * gap / cos3 / rho * Rhe + k6 * (y0 * y0 - x0 * x0) / 2 / cos3 / rho * Rhe;
      dpy = -1. * tan_e * y0 / rho + k2 * y0 * (1 + sin2) * gap / (1 + dp0) / rho2 / cos3 +
            in_fringe * (x0 * py0 + y0 * px0) * tan2 / rho / (1 + dp0) + in_fringe * y0 * px0 / rho / (1 + dp0) +
            x0 * y0 * sec2 * tan_e / rho2 / (1 + dp0) +
            k3 * (y0 * y0 * y0) * (2. / 3. / cos_e - 4. / 3. / cos3) / (1 + dp0) / rho / rho / gap -
            k5 * y0 * sin_e * gap / cos3 / rho * Rhe + k6 * x0 * y0 / cos3 / rho * Rhe;
    }
  } else {
    dx = in_fringe * sec2 * gap2 * k0 / rho / (1 + dp0);
    dy = 0;
    if (in_fringe == -1.) {
      dpx = -1. * (sec_e * sec_e * sec_e) * sin_e * gap2 * k0 / rho / rho / (1 + dp0) + tan_e * x0 / rho +
            k4 * sin2 * gap2 / 2 / cos3 / rho * Rhe - k5 * x0 * sin_e * gap / cos3 / rho * Rhe;
      dpy = -1. * tan_e * y0 / rho + k2 * y0 * (1 + sin2) * gap / (1 + dp0) / rho2 / cos3;
    } else {
      dpx = tan_e * x0 / rho + k4 * sin2 * gap2 / 2 / cos3 / rho * Rhe - k5 * x0 * sin_e * gap / cos3 / rho * Rhe;
      dpy = -1. * tan_e * y0 / rho + k2 * y0 * (1 + sin2) * gap / (1 + dp0) / rho2 / cos3 -
            k5 * y0 * sin_e * gap / cos3 / rho * Rhe;
    }
  }
  x = x0 + dx;
  px = px0 + dpx;
  y = y0 + dy;
  py = py0 + dpy;
}

/* dipoleFringeKHwangRLindberg (src/csbend.c:4413-4520, EDGE_EFFECTS 4): Hwang's
 * map made symplectic by R. Lindberg, a sequence of exact maps on the canonical
 * coordinates (one exponential in x), with the constants of eh_fringe_hwang */
EH_HD void eh_fringe_hl(double &x, double &px, double &y, double &py, double dp0, double rho,
                        const double *__restrict__ c, double in_fringe) {
  const double k0 = (EH_PI_D * EH_PI_D) / 6., k2 = c[5], k3 = 1.0 * 1. / 6., k4 = -1.0 * (EH_PI_D * EH_PI_D) / 3.,
               k5 = 0.0, k6 = -1.0;
  const double sec_e = c[0], tan_e = c[1], sin_e = c[2], cos_e = c[3], gap = c[4], Rhe = c[6];
  const double rho2 = rho * rho, sec2 = sec_e * sec_e, cos3 = cos_e * cos_e * cos_e, tan2 = tan_e * tan_e;
  const double tan3 = tan_e * tan_e * tan_e;
  const double x0 = x, px0 = px, y0 = y, py0 = py;
  double px1, py1, t1, t2;
  const bool in = in_fringe == -1;
  if (in)
    px1 = px0 + tan_e / rho * x0 + tan_e / (2 * rho2 * (1 + dp0)) * (y0 * y0) - tan3 / (rho2 * (1 + dp0)) * (x0 * x0) -
          gap * k5 * sin_e * Rhe / (rho * cos3) * x0 + k6 * sec2 * Rhe / (2 * rho) * ((y0 * y0) - (x0 * x0));
  else
    px1 = px0 + tan_e / rho * x0 + tan3 / (2 * rho2 * (1 + dp0)) * (y0 * y0) +
          tan3 / (2 * rho2 * (1 + dp0)) * (x0 * x0) - gap * k5 * sin_e * Rhe / (rho * cos3) * x0 +
          k6 * sec2 * Rhe / (2 * rho) * ((y0 * y0) - (x0 * x0));
  py1 = py0 - tan_e / rho * y0 + (in ? tan_e : tan3) / (rho2 * (1 + dp0)) * x0 * y0 +
        gap * k2 * (1 + (sin_e * sin_e)) / (rho2 * (1 + dp0) * cos3) * y0 +
        2 * k3 * ((cos_e * cos_e) - 2) / (3 * gap * rho2 * cos3) * (y0 * y0 * y0) +
        gap * k5 * sin_e * Rhe / (rho * cos3) * y0 + k6 * (sec_e * sec_e * sec_e) * Rhe / rho * x0 * y0;
  t1 = in ? (1 + tan2 / (2 * rho * (1 + dp0)) * x0) : (1 - tan2 / (2 * rho * (1 + dp0)) * x0);
  const double x2 = x0 / t1, px2 = px1 * (t1 * t1);
  t1 = sec2 / (rho * (1 + dp0));
  const double x3 = in ? x2 + t1 / 2 * (y0 * y0) : x2 - t1 / 2 * (y0 * y0);
  const double py3 = in ? py1 - t1 * y0 * px2 : py1 + t1 * y0 * px2;
  t1 = in ? tan2 / (rho * (1 + dp0)) : -tan2 / (rho * (1 + dp0));
  t2 = eh_exp(t1 * x3);
  const double px4 = px2 - t1 * y0 * py3, y4 = y0 * t2, py4 = py3 / t2;
  if (in) {
    t1 = (gap * gap) * k0 * sec2 / (rho * (1 + dp0));
    x = x3 - t1;
    px = px4 - t1 * tan_e / rho + (gap * gap) * k4 * (sin_e * sin_e) * Rhe / (2 * rho * cos3);
  } else {
    x = x3 + (gap * gap) * k0 * sec2 / (rho * (1 + dp0));
    px = px4 + (gap * gap) * k4 * (sin_e * sin_e) * Rhe / (2 * rho * cos3);
  }
  y = y4;
  py = py4;
}

/* EDGE_EFFECTS 2 or 4 at one edge (src/csbend.c:1139-1181, :1319-1360): the
 * slopes to canonical momenta, Hwang's map (or Lindberg's form), and back */
EH_HD void eh_edge_hwang(double &x, double &xp, double &y, double &yp, double dp, double rho_actual,
                         const double *__restrict__ c, double in_fringe, bool higher, bool lindberg) {
  double f = (1 + dp) / eh_sqrt(1 + eh_sqr(xp) + eh_sqr(yp));
  double qx = xp * f, qy = yp * f;
  if (lindberg)
    eh_fringe_hl(x, qx, y, qy, dp, rho_actual, c, in_fringe);
  else
    eh_fringe_hwang(x, qx, y, qy, dp, rho_actual, c, in_fringe, higher);
  f = 1 / eh_sqrt(eh_sqr(1 + dp) - eh_sqr(qx) - eh_sqr(qy));
  xp = qx * f;
  yp = qy * f;
}

/* CSBEND: track_through_csbend (src/csbend.c:1088-1403) for one particle, with
 * integrate_csbend_ordn (src/csbend.c:1459-1720) and the canonical conversions
 * (:1436-1454).  Returns true if lost.  RAD and FULL as for eh_multipole_t. */
template <bool RAD, bool FULL = RAD>
EH_HD bool eh_csbend_t(eh_p6 &p, const eh_desc &d, int pi, const eh_fterm *__restrict__ terms,
                       const double *__restrict__ zp) {
  const double apx = d.ap_xmax[pi], apy = d.ap_ymax[pi];
  const int ell = FULL ? d.ap_ell[pi] : 0;
  const double rho0 = d.rho0, rho_actual = d.rho_actual;
  double x, xp, y, yp, s, dp;
  p.s += d.dzi * eh_sqrt(1 + eh_sqr(p.xp) + eh_sqr(p.yp));
  p.x = p.x + d.dxi + d.dzi * p.xp;
  p.y = p.y + d.dyi + d.dzi * p.yp;
  x = p.x * d.cos_tt + p.y * d.sin_tt;
  y = -p.x * d.sin_tt + p.y * d.cos_tt;
  xp = p.xp * d.cos_tt + p.yp * d.sin_tt;
  yp = -p.xp * d.sin_tt + p.yp * d.cos_tt;
  s = p.s;
  dp = p.dp;
  const double dp0 = dp;
  if (d.csb_edge1) {
    double rho = (1 + dp) * rho_actual;
    if (FULL && (d.csb_hwang1 || d.csb_hl1)) {
      eh_edge_hwang(x, xp, y, yp, dp, rho_actual, d.csb_hw1, -1., d.csb_hw_order2, d.csb_hl1);
    } else if (FULL && d.csb_esimple1) {
      eh_edge_simple(xp, yp, x, y, dp, rho_actual, d.tan_e1, d.e1, d.psi1, d.csb_klim1);
    } else if (!d.csb_eorder2) {
      double delta_xp = d.tan_e1 / rho * x;
      if (FULL && d.csb_klim1 > 0 && eh_fabs(delta_xp) > d.csb_klim1)
        delta_xp = (delta_xp < 0 ? -1.0 : 1.0) * d.csb_klim1;
      xp += delta_xp;
      yp -= eh_tan(d.e1 - d.psi1 / (1 + dp)) / rho * y;
    } else {
      eh_edge2(x, xp, y, yp, rho, d.csb_n, d.e1, d.tan_e1, d.csb_sec1, d.csb_he1, d.psi1 * (1 + dp), -1);
    }
  }
  double dpi = dp;
  if constexpr (RAD) {
    if (d.csb_e1r) {
      /* pre-adjust delta for the sector integration, src/csbend.c:1214-1222 */
      double Fx, Fy;
      eh_csb_fields(d, terms, x, y, Fx, Fy);
      const double dp_prime = -d.rad_coef * (eh_sqr(Fx) + eh_sqr(Fy)) * eh_sqr(1 + dp) *
                              eh_sqrt(eh_sqr(1 + x / rho0) + eh_sqr(xp) + eh_sqr(yp));
      dpi -= dp_prime * x * d.tan_e1;
    }
  }
  /* convertToDipoleCanonicalCoordinates */
  double X = x, QX, Y = y, QY, DP = dpi, S4 = 0;
  {
    double f = (1 + DP) / eh_sqrt(1 + eh_sqr(xp) + eh_sqr(yp));
    QX = xp * f;
    QY = yp * f;
  }
  /* integrate_csbend_ordn; 1 + DPoP and its reciprocal pair for the radiation
   * block's round trip after each kick */
  const double b1 = 1 + DP;
  double rh = 0, rl = 0;
  const bool rok = !RAD && eh_recip_pair(b1, rh, rl);
  double dist = 0, dz_lost = 0;
  bool lost = false;
  for (int i = 0; i < d.csb_nslices && !lost; i++) {
    if (!eh_ap_ok<FULL>(X, Y, apx, apy, d.ap_rtilt, d.ap_cs0, d.ap_cs1, ell, d.ell_xe[pi], d.ell_ye[pi])) {
      lost = true;
      break;
    }
    for (int j = 0; j < d.nsub; j++) {
      double f = eh_sqr(1 + DP) - eh_sqr(QY);
      if (f <= 0) {
        lost = true;
        break;
      }
      f = eh_sqrt(f);
      if (eh_fabs(QX / f) > 1) {
        lost = true;
        break;
      }
      double sin_phi = QX / f;
      double phi = eh_asin(sin_phi);
      double sine, cosi;
      eh_sincos(d.csb_dshr[j] + phi, &sine, &cosi);
      if (cosi == 0) {
        lost = true;
        break;
      }
      double tang = sine / cosi;
      double cos_phi = eh_cos(phi);
      QX = f * sine;
      double factor = (rho0 + X) * cos_phi / f * (tang - sin_phi / cos_phi);
      Y += QY * factor;
      dist += factor * (1 + DP);
      dz_lost += d.csb_dsh[j];
      f = cos_phi / cosi;
      X = rho0 * (f - 1) + f * X;
      if (j == d.nsub - 1)
        break;
      double ds = d.csb_ds[j];
      double Fx, Fy;
      eh_csb_fields(d, terms, X, Y, Fx, Fy);
      if (d.csb_recip) {
        const double xr = eh_div_recip(X, rho0, d.rho0_rh, d.rho0_rl);
        QX += eh_div_recip(-ds * (1 + xr) * Fy, rho_actual, d.rhoa_rh, d.rhoa_rl);
        QY += eh_div_recip(ds * (1 + xr) * Fx, rho_actual, d.rhoa_rh, d.rhoa_rl);
      } else {
        QX += -ds * (1 + X / rho0) * Fy / rho_actual;
        QY += ds * (1 + X / rho0) * Fx / rho_actual;
      }
      /* the radiation block runs for every bend (src/csbend.c:1627-1660;
       * isrConstant is non-zero even without ISR, :961-964) */
      if constexpr (RAD) {
        /* with rad_coef: the energy loss of the sub-step (src/csbend.c:1627-1651) */
        const double h1 = 1 + X * d.csb_irho0;
        const double fr = h1 / eh_sqrt(eh_sqr(1 + DP) - eh_sqr(QX) - eh_sqr(QY));
        const double xpr = QX * fr, ypr = QY * fr;
        const double dsFactor = eh_sqrt(eh_sqr(1 + X * d.csb_irho0) + eh_sqr(xpr) + eh_sqr(ypr));
        const double F2 = eh_sqr(Fx) + eh_sqr(Fy);
        const double deltaFactor = eh_sqr(1 + DP);
        QX /= (1 + DP);
        QY /= (1 + DP);
        DP -= d.rad_coef * deltaFactor * F2 * ds * dsFactor;
        QX *= (1 + DP);
        QY *= (1 + DP);
      } else {
        QX = eh_round_trip(QX, b1, rh, rl, rok);
        QY = eh_round_trip(QY, b1, rh, rl, rok);
      }
    }
  }
  if (!lost) {
    dz_lost = d.csb_dzl;
    if (!eh_ap_ok<FULL>(X, Y, apx, apy, d.ap_rtilt, d.ap_cs0, d.ap_cs1, ell, d.ell_xe[pi], d.ell_ye[pi]))
      lost = true;
    else
      S4 += dist;
  }
  /* convertFromDipoleCanonicalCoordinates */
  {
    double f = 1 / eh_sqrt(eh_sqr(1 + DP) - eh_sqr(QX) - eh_sqr(QY));
    QX *= f;
    QY *= f;
  }
  if (lost) {
    /* memcpy(part, Qf) then convertFromCSBendCoords (src/csbend.c:4527-4550) */
    p.x = X * d.cos_tt - Y * d.sin_tt;
    p.y = X * d.sin_tt + Y * d.cos_tt;
    p.xp = QX * d.cos_tt - QY * d.sin_tt;
    p.yp = QX * d.sin_tt + QY * d.cos_tt;
    p.s = zp[0] + dz_lost;
    p.dp = d.Po * (1 + DP);
    return true;
  }
  if constexpr (RAD) {
    if (d.csb_e2r) {
      /* post-adjust delta, src/csbend.c:1270-1282 */
      double Fx, Fy;
      eh_csb_fields(d, terms, X, Y, Fx, Fy);
      const double dp_prime = -d.rad_coef * (eh_sqr(Fx) + eh_sqr(Fy)) * eh_sqr(1 + DP) *
                              eh_sqrt(eh_sqr(1 + X / rho0) + eh_sqr(QX) + eh_sqr(QY));
      DP -= dp_prime * X * d.tan_e2;
    }
    /* time-of-flight rescaling, src/csbend.c:1285-1296 */
    const double P0 = d.Po * (1 + dp0), P1 = d.Po * (1 + DP);
    const double beta0 = P0 / eh_sqrt(eh_sqr(P0) + 1), beta1 = P1 / eh_sqrt(eh_sqr(P1) + 1);
    s = beta1 * s / beta0 + S4;
  } else {
    /* time-of-flight rescaling (src/csbend.c:1285-1296, taken for every bend):
     * without radiation delta is unchanged, so beta1 = beta0 */
    const double P0 = d.Po * (1 + dp0);
    const double beta = P0 / eh_sqrt(eh_sqr(P0) + 1);
    s = beta * s / beta + S4;
  }
  x = X;
  xp = QX;
  y = Y;
  yp = QY;
  dp = DP;
  if (d.csb_edge2) {
    double rho = (1 + dp) * rho_actual;
    if (FULL && (d.csb_hwang2 || d.csb_hl2)) {
      eh_edge_hwang(x, xp, y, yp, dp, rho_actual, d.csb_hw2, 1., d.csb_hw_order2, d.csb_hl2);
    } else if (FULL && d.csb_esimple2) {
      eh_edge_simple(xp, yp, x, y, dp, rho_actual, d.tan_e2, d.e2, d.psi2, d.csb_klim2);
    } else if (!d.csb_eorder2) {
      double delta_xp = d.tan_e2 / rho * x;
      if (FULL && d.csb_klim2 > 0 && eh_fabs(delta_xp) > d.csb_klim2)
        delta_xp = (delta_xp < 0 ? -1.0 : 1.0) * d.csb_klim2;
      xp += delta_xp;
      yp -= eh_tan(d.e2 - d.psi2 / (1 + dp)) / rho * y;
    } else {
      eh_edge2(x, xp, y, yp, rho, d.csb_n, d.e2, d.tan_e2, d.csb_sec2, d.csb_he2, d.psi2 * (1 + dp), 1);
    }
  }
  p.x = x * d.cos_tt - y * d.sin_tt + 0.0;
  p.y = x * d.sin_tt + y * d.cos_tt + 0.0;
  p.xp = xp * d.cos_tt - yp * d.sin_tt + 0.0;
  p.yp = xp * d.sin_tt + yp * d.cos_tt + 0.0;
  p.s = s + 0.0;
  p.dp = dp;
  p.x += d.dxf + d.dzf * p.xp;
  p.y += d.dyf + d.dzf * p.yp;
  p.s += d.dzf * eh_sqrt(1 + eh_sqr(p.xp) + eh_sqr(p.yp));
  return false;
}

EH_HD bool eh_csbend(eh_p6 &p, const eh_desc &d, int pi, const eh_fterm *__restrict__ terms,
                     const double *__restrict__ zp) {
  return d.rad ? eh_csbend_t<true, true>(p, d, pi, terms, zp) : eh_csbend_t<false, true>(p, d, pi, terms, zp);
}

/* ------------------------------------------------------------------------- */
/* RFCA for one particle: trackRfCavityWithWakes (src/simple_rfca.c:295-985,
 * matrix method :824-967, kick method :567-660, add_to_particle_energy
 * :987-1006) followed by removeInvalidParticles (src/limit_amplitudes.c:289-330).
 * Returns true if lost. */
EH_HD bool eh_rfca(eh_p6 &p, const eh_desc &d, const double *__restrict__ zp) {
  const double c_mks = 2.99792458e8;
  const double Pc = d.Po;
  if (d.rf_zero_volt) {
    double L = d.rf_flen;
    if (L) {
      p.x += p.xp * L;
      p.y += p.yp * L;
      p.s += L * eh_sqrt(1 + eh_sqr(p.xp) + eh_sqr(p.yp));
    }
    return false;
  }
  const double length = d.rf_len, volt = d.rf_volt, omega = d.rf_omega, phase = d.rf_phase;
  p.x -= d.rf_dx;
  p.y -= d.rf_dy;
  if (!d.rf_matrix) {
    double inverseF = 0;
    for (int ik = 0; ik < d.rf_nkicks; ik++) {
      if (!(p.dp == -1)) {
        double dc4 = length ? length / 2 * eh_sqrt(1 + eh_sqr(p.xp) + eh_sqr(p.yp)) : 0;
        double P = Pc * (1 + p.dp);
        double gamma = eh_sqrt(eh_sqr(P) + 1);
        double beta_i = P / gamma;
        double t = (p.s + dc4) / (c_mks * beta_i) - d.rf_toff;
        if (ik == 0 && d.rf_toff != 0)
          p.s = t * c_mks * beta_i - dc4;
        double dt = t - d.rf_t0;
        if (dt < 0)
          dt = 0;
        const double tau = d.rf_tau;
        double dgamma = d.rf_dgamma_same;
        if (!d.rf_same) {
          if (!d.rf_lin) {
            if (d.rf_standing)
              dgamma = volt * eh_sin(omega * (t - d.rf_tlock) + phase) * (tau ? eh_sqrt(1 - eh_exp(-dt / tau)) : 1);
            else {
              double phi = omega * (t - d.rf_tlock - ik * d.rf_dtLight) + phase;
              double dfactor = tau ? eh_sqrt(1 - eh_exp(-dt / tau)) : 1;
              dgamma = volt * eh_sin(phi) * dfactor;
            }
          } else
            dgamma = d.rf_dgave + volt * omega * (t - d.rf_tave) * d.rf_cosk;
        }
        if (length) {
          if (d.rf_end1 && ik == 0) {
            inverseF = dgamma / (2 * gamma * length);
            p.xp -= p.x * inverseF;
            p.yp -= p.y * inverseF;
          }
          p.x += p.xp * length / 2;
          p.y += p.yp * length / 2;
          p.s += length / 2 * eh_sqrt(1 + eh_sqr(p.xp) + eh_sqr(p.yp));
        }
        {
          /* add_to_particle_energy */
          double P0 = Pc * (1 + p.dp);
          double gam = eh_sqrt(P0 * P0 + 1);
          double gamma1 = gam + dgamma;
          if (gamma1 <= 1)
            gamma1 = 1 + 1e-7;
          double P1 = eh_sqrt(gamma1 * gamma1 - 1);
          p.dp = (P1 - Pc) / Pc;
          p.s = t * c_mks * P1 / gamma1;
          double Pz = P0 / eh_sqrt(1 + eh_sqr(p.xp) + eh_sqr(p.yp));
          double Pz1 = eh_sqrt(Pz * Pz + gamma1 * gamma1 - gam * gam);
          double PRatio = Pz / Pz1;
          p.xp *= PRatio;
          p.yp *= PRatio;
        }
        double g1 = gamma + dgamma;
        if (g1 <= 1)
          p.dp = -1;
        else
          inverseF = -dgamma / (2 * g1 * length);
      }
      if (length) {
        p.x += p.xp * length / 2;
        p.y += p.yp * length / 2;
        p.s += length / 2 * eh_sqrt(1 + eh_sqr(p.xp) + eh_sqr(p.yp));
        if (d.rf_end2 && (ik == d.rf_nkicks - 1)) {
          p.xp -= p.x * inverseF;
          p.yp -= p.y * inverseF;
        }
      }
    }
  } else {
    const double R11 = 1, R21 = 0;
    double P = Pc * (1 + p.dp);
    double gamma = eh_sqrt(eh_sqr(P) + 1);
    double beta_i = P / gamma;
    double ds1 = length / 2 * eh_sqrt(1 + eh_sqr(p.xp) + eh_sqr(p.yp));
    double t = (p.s + ds1) / (c_mks * beta_i) - d.rf_toff;
    if (d.rf_toff != 0)
      p.s = t * c_mks * beta_i - ds1;
    double dt = t - d.rf_t0;
    if (dt < 0)
      dt = 0;
    const double tau = d.rf_tau;
    double dgamma = d.rf_dgamma_same;
    if (!d.rf_same) {
      if (!d.rf_lin) {
        double sin_phase = eh_sin(omega * (t - d.rf_tlock - 0 * d.rf_dtLight) + phase);
        dgamma = (volt * (tau ? eh_sqrt(1 - eh_exp(-dt / tau)) : 1)) * sin_phase;
      } else {
        double sin_phase = omega * (t - d.rf_tave) * d.rf_cosm;
        dgamma = (volt * (tau ? eh_sqrt(1 - eh_exp(-dt / tau)) : 1)) * sin_phase + d.rf_dgave;
      }
    }
    if (d.rf_end1 && length) {
      double inverseF = dgamma / (2 * gamma * length);
      p.xp -= p.x * inverseF;
      p.yp -= p.y * inverseF;
    }
    double dP = eh_sqrt(eh_sqr(gamma + dgamma) - 1) - P;
    double R22 = 1 / (1 + dP / P);
    double R12;
    if (eh_fabs(dP / P) > 1e-14)
      R12 = length * (P / dP * eh_log(1 + dP / P));
    else
      R12 = length;
    p.s += ds1;
    double x = p.x, xp = p.xp;
    p.x = x * R11 + xp * R12;
    p.xp = x * R21 + xp * R22;
    x = p.y;
    xp = p.yp;
    p.y = x * R11 + xp * R12;
    p.yp = x * R21 + xp * R22;
    p.s += length / 2 * eh_sqrt(1 + eh_sqr(p.xp) + eh_sqr(p.yp));
    p.dp = (P + dP - Pc) / Pc;
    if ((gamma += dgamma) <= 1)
      p.dp = -1;
    if (d.rf_end2 && length) {
      double inverseF = -dgamma / (2 * gamma * length);
      p.xp -= p.x * inverseF;
      p.yp -= p.y * inverseF;
    }
    p.s = (P + dP) / gamma * p.s / beta_i;
  }
  p.x += d.rf_dx;
  p.y += d.rf_dy;
  /* removeInvalidParticles */
  if (eh_isnan(p.x) || eh_isnan(p.xp) || eh_isnan(p.y) || eh_isnan(p.yp) || eh_isnan(p.s) || eh_isnan(p.dp) ||
      p.dp <= -1) {
    p.s = zp[1];
    p.dp = Pc * (1 + p.dp);
    return true;
  }
  return false;
}

/* one element followed by do_tracking's post-element sweeps
 * (src/do_tracking.c:2506-2524, :2675-2684).  Returns -1 if the particle
 * survives, else the loss stage (EH_STAGE_ELEMENT possibly with
 * EH_STAGE_INHERIT). */
template <bool RAD>
EH_HD int eh_apply_element_t(eh_p6 &p, const eh_desc &d, int pass, const eh_fterm *__restrict__ terms,
                             const double *__restrict__ zp) {
  const int pi = pass > 0 ? 1 : 0;
  bool lost = false;
  int flag = 0;
  if constexpr (!RAD) {
    if (d.full_only)
      return EH_STAGE_REFUSED;
  }
  switch (d.type) {
  case 1: /* DRIF */
    eh_drif(p, d.length, d.T11);
    break;
  case 2: /* EDRIFT */
    eh_edrift(p, d.length);
    break;
  case 3:
  case 4:
  case 5:
  case 9: { /* KQUSE */
    const int m = RAD ? eh_multipole(p, d, pi, zp) : eh_multipole_t<false>(p, d, pi, zp);
    lost = m != 0;
    flag = m == 2 ? EH_STAGE_INHERIT : 0;
    break;
  }
  case 6:
    lost = RAD ? eh_csbend(p, d, pi, terms, zp) : eh_csbend_t<false>(p, d, pi, terms, zp);
    break;
  case 7:
    lost = eh_rfca(p, d, zp);
    break;
  /* the full instantiation (RAD) also carries the element types below; the
   * lean one refuses them (the host runs it only on lines without them) */
  case 10: /* MULT (no slices: zero strength and length, nothing to do) */
    if constexpr (!RAD)
      return EH_STAGE_REFUSED;
    if (d.nslices)
      lost = eh_mult_t<RAD>(p, d, zp);
    break;
  case 11: /* ROTATE, src/do_tracking.c:1836-1852 */
    if constexpr (!RAD)
      return EH_STAGE_REFUSED;
    if (d.on_pass == -1 || d.on_pass == pass)
      eh_rotate(p, d.ncos, d.nsin);
    break;
  case 12: /* MALIGN, src/do_tracking.c:1853-1857 */
    if constexpr (!RAD)
      return EH_STAGE_REFUSED;
    if (d.on_pass == -1 || d.on_pass == pass)
      eh_offset(p, d);
    break;
  case 13: /* ENERGY: set_central_momentum, src/do_tracking.c:3591-3595 */
    if constexpr (!RAD)
      return EH_STAGE_REFUSED;
    if (d.en_apply)
      p.dp = ((1 + p.dp) * d.en_pold - d.en_pnew) / d.en_pnew;
    break;
  default: /* MARK, MAXAMP */
    break;
  }
  if (lost)
    return EH_STAGE_ELEMENT | flag;
  if (d.type == 0)
    return -1; /* MARK: NO_APERTURE, and nothing can be NaN-lost here that survived */
  if (d.post_ap[pi]) {
    bool out;
    if (RAD && d.post_ell[pi])
      out = eh_elimit_one(p, d.ell_a2[pi], d.ell_b2[pi], d.ell_xe[pi], d.ell_ye[pi], zp[1], d.Po, d.extrap);
    else
      out = eh_limit_one(p, d.post_xmax[pi], d.post_ymax[pi], zp[1], d.Po, d.extrap);
    if (out)
      return EH_STAGE_MAXAMP;
  }
  if (eh_limit_one(p, EH_DBL_MAX, EH_DBL_MAX, zp[1], d.Po, 0))
    return EH_STAGE_NAN;
  return -1;
}

/* where and how a particle was lost inside a fused span, in one word: element
 * ee of pass p0 + dpass (nelem elements per pass) at stage st (EH_STAGE_*; the
 * element, sweep and NaN stages and the inherit flag are kept).  The host keeps
 * (dpass + 1) * nelem * 8 below 2^31. */
EH_HD int32_t eh_loss_code(int dpass, int nelem, int ee, int st) {
  return ((dpass * nelem + ee) << 3) | (st & (3 | EH_STAGE_INHERIT));
}

/* the loss pass and loss event of a loss code: events are
 * event_base + (dpass * nelem + ee) * 3 + stage, in the order the reference
 * compacts; the pass carries EH_LPASS_INHERIT for an inherited loss position */
EH_HD void eh_loss_words(int32_t lcode, int nelem, int p0, int32_t event_base, int32_t &lpass, int32_t &levent) {
  const int32_t k = lcode >> 3;
  lpass = (p0 + k / nelem) | ((lcode & EH_STAGE_INHERIT) ? EH_LPASS_INHERIT : 0);
  levent = event_base + k * 3 + (lcode & 3);
}

/* whether a descriptor can run through the lean map (RAD = false): kernels for
 * lines without radiating elements and without MULT, ROTATE, MALIGN or ENERGY
 * carry none of that code (and none of its registers) */
EH_HD bool eh_lean_capable(const eh_desc &d) { return !d.rad && d.type < 10 && !d.full_only; }


EH_HD int eh_apply_element(eh_p6 &p, const eh_desc &d, int pass, const eh_fterm *__restrict__ terms) {
  return eh_apply_element_t<true>(p, d, pass, terms, d.zpos);
}

#endif
