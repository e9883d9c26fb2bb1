/* Reverse-mode adjoint of the element maps.
 *
 * Each routine takes the state z at an element's EXIT and the adjoint l = dL/dz
 * there, and returns the state at the element's ENTRY and the adjoint at the
 * entry, adding the element's parameter derivatives to g[].  The entry state is
 * not stored: every sub-step of these maps is a shear or a flow (drift, kick,
 * rotation, offset, exact drift in the curved frame), so its input is recovered
 * from its output by the inverse sub-step, and the adjoint of the sub-step is
 * evaluated at that input.  Rounding makes the recovered state differ from the
 * forward one by a few ulp per sub-step; the caller re-anchors to stored states
 * at turn starts and RF cavities, which bounds the drift to one turn.
 *
 * The RF cavity (matrix method) is not inverted: its entry state is stored by
 * the forward pass, and its 6x6 Jacobian is evaluated there with dual numbers.
 *
 * Conventions (src/multipole.h:448-486, src/csbend.c:1436-1454): slopes
 * xp = dx/ds, momenta qx = (1+dp) xp / sqrt(1+xp^2+yp^2), and back
 * xp = qx / sqrt((1+dp)^2 - qx^2 - qy^2).  Derivatives are those of the exact
 * real functions the forward code evaluates.
 */
#pragma once

#include "eh_device.h"

/* parameter gradients of one element: d/d(strength), d/dDX, d/dDY, d/dFSE
 * (the strength of a bend is its K1) */
struct eh_pgrad {
  double k, dx, dy, fse;
};

/* ---- slopes <-> momenta --------------------------------------------------- */

/* adjoint through xp = m2s(q): given lxp, lyp at the slopes, add to lq, ldp */
EH_HD void eh_adj_m2s(double qx, double qy, double dp, double lxp, double lyp, double &lqx,
                                           double &lqy, double &ldp) {
  double den = eh_sqrt(eh_sqr(1 + dp) - eh_sqr(qx) - eh_sqr(qy));
  double den3 = den * den * den;
  double u = (lxp * qx + lyp * qy) / den3;
  lqx += lxp / den + qx * u;
  lqy += lyp / den + qy * u;
  ldp += -(1 + dp) * u;
}

/* adjoint through q = s2m(xp): given lqx, lqy, add to lxp, lyp, ldp */
EH_HD void eh_adj_s2m(double xp, double yp, double dp, double lqx, double lqy, double &lxp,
                                           double &lyp, double &ldp) {
  double w = eh_sqrt(1 + eh_sqr(xp) + eh_sqr(yp));
  double w3 = w * w * w;
  double v = (lqx * xp + lqy * yp) / w3;
  lxp += (1 + dp) * (lqx / w - xp * v);
  lyp += (1 + dp) * (lqy / w - yp * v);
  ldp += (lqx * xp + lqy * yp) / w;
}

EH_HD void eh_s2m(double xp, double yp, double dp, double &qx, double &qy) {
  double w = eh_sqrt(1 + eh_sqr(xp) + eh_sqr(yp));
  qx = (1 + dp) * xp / w;
  qy = (1 + dp) * yp / w;
}

EH_HD void eh_m2s_raw(double qx, double qy, double dp, double &xp, double &yp) {
  double den = eh_sqrt(eh_sqr(1 + dp) - eh_sqr(qx) - eh_sqr(qy));
  xp = qx / den;
  yp = qy / den;
}

/* ---- frame changes ---------------------------------------------------------- */

/* forward: (x,y) <- (x c + y s, -x s + y c), same for slopes.  Undo and pull the
 * adjoint back. */
EH_HD void eh_adj_rot(eh_p6 &z, eh_p6 &l, double c, double s) {
  double xo = z.x, yo = z.y, xpo = z.xp, ypo = z.yp;
  z.x = xo * c - yo * s;
  z.y = xo * s + yo * c;
  z.xp = xpo * c - ypo * s;
  z.yp = xpo * s + ypo * c;
  double lx = l.x, ly = l.y, lxp = l.xp, lyp = l.yp;
  l.x = c * lx - s * ly;
  l.y = s * lx + c * ly;
  l.xp = c * lxp - s * lyp;
  l.yp = s * lxp + c * lyp;
}

/* forward: s += dz w; x += ox + dz xp; y += oy + dz yp with w = sqrt(1+xp^2+yp^2)
 * (slopes unchanged).  Undo and pull the adjoint back. */
EH_HD void eh_adj_offset(eh_p6 &z, eh_p6 &l, double ox, double oy, double dz) {
  double w = eh_sqrt(1 + eh_sqr(z.xp) + eh_sqr(z.yp));
  z.x = z.x - ox - dz * z.xp;
  z.y = z.y - oy - dz * z.yp;
  z.s = z.s - dz * w;
  l.xp += dz * l.x + dz * l.s * z.xp / w;
  l.yp += dz * l.y + dz * l.s * z.yp / w;
}

/* ---- drifts ------------------------------------------------------------------ */

/* DRIF order-2 map: x += L xp, y += L yp, s += L + T (xp^2 + yp^2) */
EH_HD void eh_adj_drif(eh_p6 &z, eh_p6 &l, double L, double T) {
  z.x -= L * z.xp;
  z.y -= L * z.yp;
  z.s -= L + T * z.yp * z.yp + T * z.xp * z.xp;
  l.xp += L * l.x + 2 * T * z.xp * l.s;
  l.yp += L * l.y + 2 * T * z.yp * l.s;
}

/* exact drift: x += L xp, y += L yp, s += L sqrt(1+xp^2+yp^2) */
EH_HD void eh_adj_edrift(eh_p6 &z, eh_p6 &l, double L) {
  double w = eh_sqrt(1 + eh_sqr(z.xp) + eh_sqr(z.yp));
  z.x -= L * z.xp;
  z.y -= L * z.yp;
  z.s -= L * w;
  l.xp += L * l.x + l.s * L * z.xp / w;
  l.yp += L * l.y + l.s * L * z.yp / w;
}

/* ---- kick multipoles ---------------------------------------------------------- */

/* (x+iy)^n/n! = re + i im and its derivative (x+iy)^(n-1)/(n-1)! = a + i b */
EH_HD void eh_cpow_fact(int n, double x, double y, double &re, double &im, double &a,
                                             double &b) {
  double pr = 1, pi = 0; /* (x+iy)^k / k! */
  a = 0;
  b = 0;
  for (int k = 1; k <= n; k++) {
    a = pr;
    b = pi;
    double nr = (pr * x - pi * y) / k;
    double ni = (pr * y + pi * x) / k;
    pr = nr;
    pi = ni;
  }
  re = pr;
  im = pi;
}

/* field sums of one multipole term: sum_Fy, sum_Fx (multipole.h:391-420) and
 * their partial derivatives */
EH_HD void eh_mfield(int order, int skew, double x, double y, double &Fy, double &Fx,
                                          double &Fy_x, double &Fy_y, double &Fx_x, double &Fx_y) {
  double re, im, a, b;
  eh_cpow_fact(order, x, y, re, im, a, b);
  if (order == 0) {
    a = 0;
    b = 0;
  }
  if (!skew) {
    Fy = re;
    Fx = im;
    Fy_x = a;
    Fy_y = -b;
    Fx_x = b;
    Fx_y = a;
  } else {
    Fy = im;
    Fx = -re;
    Fy_x = b;
    Fy_y = a;
    Fx_x = -a;
    Fx_y = b;
  }
}

/* one kick of the integrator: qx -= k Fy, qy += k Fx over the element's terms and
 * steering.  Undo it and pull the adjoint back (x, y are unchanged by it). */
EH_HD void eh_adj_mkick(const eh_desc &d, int step, double x, double y, double &qx,
                                             double &qy, eh_p6 &l, double &lqx, double &lqy, eh_pgrad &g,
                                             int want_k) {
  double dqx = 0, dqy = 0, lx = 0, ly = 0;
  for (int io = 0; io < 3; io++) {
    if (!d.mp_order[io])
      continue;
    double Fy, Fx, Fy_x, Fy_y, Fx_x, Fx_y;
    eh_mfield(d.mp_order[io], d.mp_skew[io], x, y, Fy, Fx, Fy_x, Fy_y, Fx_x, Fx_y);
    double k = d.kick[step][io];
    dqx -= k * Fy;
    dqy += k * Fx;
    lx += k * (-lqx * Fy_x + lqy * Fx_x);
    ly += k * (-lqx * Fy_y + lqy * Fx_y);
  }
  if (d.has_xkick)
    dqx -= d.xk[step];
  if (d.has_ykick)
    dqy -= d.yk[step];
  if (want_k) {
    /* d kick[step][0] / d K = adj_kfac[step], also when K is zero */
    double Fy, Fx, a1, a2, a3, a4;
    eh_mfield(d.adj_order0, 0, x, y, Fy, Fx, a1, a2, a3, a4);
    g.k += d.adj_kfac[step] * (-lqx * Fy + lqy * Fx);
  }
  qx -= dqx;
  qy -= dqy;
  l.x += lx;
  l.y += ly;
}

/* KQUAD/KSEXT/KOCT survivor path of eh_multipole (expandHamiltonian 0) */
EH_HD void eh_adj_multipole(eh_p6 &z, eh_p6 &l, const eh_desc &d, eh_pgrad &g) {
  /* face 2 (applied last): tilt back, then offset back.  DX and DY enter as
   * x_out = x + DX at the exit and x1 = x - DX at the entry; their derivatives
   * are taken also when the offsets are zero and the steps are identities. */
  g.dx += l.x;
  g.dy += l.y;
  if (d.has_offset)
    eh_adj_offset(z, l, d.mdx, d.mdy, -d.mdz);
  if (d.has_tilt)
    eh_adj_rot(z, l, d.xcos, d.xsin);
  /* body output slopes come from the momenta: q = s2m(xp) inverts xp = m2s(q) */
  const double dp = z.dp;
  double qx, qy;
  eh_s2m(z.xp, z.yp, dp, qx, qy);
  double lqx = 0, lqy = 0;
  eh_adj_m2s(qx, qy, dp, l.xp, l.yp, lqx, lqy, l.dp);
  const int want_k = d.adj_pk >= 0;
  for (int ik = d.nslices - 1; ik >= 0; ik--) {
    for (int step = d.nsub - 1; step >= 0; step--) {
      if (step < d.nsub - 1)
        eh_adj_mkick(d, step, z.x, z.y, qx, qy, l, lqx, lqy, g, want_k);
      if (d.mdrift != 0) {
        double h = d.dsh[step];
        double xp, yp;
        eh_m2s_raw(qx, qy, dp, xp, yp);
        double w = eh_sqrt(1 + eh_sqr(xp) + eh_sqr(yp));
        double lxpe = l.x * h + l.s * h * xp / w;
        double lype = l.y * h + l.s * h * yp / w;
        eh_adj_m2s(qx, qy, dp, lxpe, lype, lqx, lqy, l.dp);
        z.x -= xp * h;
        z.y -= yp * h;
        z.s -= h * w;
      }
    }
  }
  /* entry: q = s2m(xp_in); the slopes the first drift used are m2s(q) */
  eh_m2s_raw(qx, qy, dp, z.xp, z.yp);
  l.xp = 0;
  l.yp = 0;
  eh_adj_s2m(z.xp, z.yp, dp, lqx, lqy, l.xp, l.yp, l.dp);
  /* face 1: offset, then tilt */
  if (d.has_tilt)
    eh_adj_rot(z, l, d.ncos, d.nsin);
  g.dx -= l.x;
  g.dy -= l.y;
  if (d.has_offset)
    eh_adj_offset(z, l, -d.mdx, -d.mdy, d.mdz);
}

/* ---- CSBEND ------------------------------------------------------------------- */

/* field expansion (computeCSBENDFields) with partial derivatives */
/* the dense-block form for expansion orders 4 and 5 (the blocks and live-term
 * mask the forward evaluation uses): powers once, then each live term; the
 * adjoint needs accuracy, not the reference's summation order */
template <int EO1>
EH_HD void eh_csb_fields_d_dense(const eh_fterm *__restrict__ cx, const eh_fterm *__restrict__ cy, uint32_t live,
                                 double x, double y, double &Fx, double &Fy, double &Fx_x, double &Fx_y,
                                 double &Fy_x, double &Fy_y) {
  double xp[EO1], yp[EO1];
  xp[0] = yp[0] = 1;
#pragma unroll
  for (int k = 1; k < EO1; k++) {
    xp[k] = xp[k - 1] * x;
    yp[k] = yp[k - 1] * y;
  }
  double F[2] = {0, 0}, F_x[2] = {0, 0}, F_y[2] = {0, 0};
#pragma unroll
  for (int i = 0; i < EO1; i++) {
#pragma unroll
    for (int j = 0; j < EO1 - i; j++) {
      const int t = i * EO1 - i * (i - 1) / 2 + j;
      if (!((live >> t) & 1u))
        continue;
#if defined(__HIP_DEVICE_COMPILE__)
      __asm__ volatile("");
#endif
      const int c = (j & 1) ? 0 : 1;
      const double k = ((j & 1) ? cx : cy)[i * EO1 + j].coef;
      F[c] += k * xp[i] * yp[j];
      if (i > 0)
        F_x[c] += (k * i) * xp[i - 1] * yp[j];
      if (j > 0)
        F_y[c] += (k * j) * xp[i] * yp[j - 1];
    }
  }
  Fx = F[0];
  Fx_x = F_x[0];
  Fx_y = F_y[0];
  Fy = F[1];
  Fy_x = F_x[1];
  Fy_y = F_y[1];
}

EH_HD void eh_csb_fields_d(const eh_desc &d, const eh_fterm *__restrict__ terms, double x,
                                                double y, double &Fx, double &Fy, double &Fx_x, double &Fx_y,
                                                double &Fy_x, double &Fy_y) {
  Fx = Fy = Fx_x = Fx_y = Fy_x = Fy_y = 0;
  if (!d.csb_field) {
    Fy = 1;
    return;
  }
  if (d.csb_dense == 5) {
    eh_csb_fields_d_dense<5>(terms + d.csb_dx0, terms + d.csb_dy0, d.csb_live, x, y, Fx, Fy, Fx_x, Fx_y, Fy_x, Fy_y);
    return;
  }
  if (d.csb_dense == 6) {
    eh_csb_fields_d_dense<6>(terms + d.csb_dx0, terms + d.csb_dy0, d.csb_live, x, y, Fx, Fy, Fx_x, Fx_y, Fy_x, Fy_y);
    return;
  }
  for (int c = 0; c < 2; c++) {
    const eh_fterm *t = terms + (c ? d.csb_fy0 : d.csb_fx0);
    int n = c ? d.csb_nfy : d.csb_nfx;
    double F = 0, F_x = 0, F_y = 0;
    for (int k = 0; k < n; k++) {
      int i = t[k].i, j = t[k].j;
      double xi = 1, xim = 0, yj = 1, yjm = 0;
      for (int m = 0; m < i; m++) {
        xim = xi;
        xi *= x;
      }
      for (int m = 0; m < j; m++) {
        yjm = yj;
        yj *= y;
      }
      F += t[k].coef * xi * yj;
      F_x += t[k].coef * i * xim * yj;
      F_y += t[k].coef * j * xi * yjm;
    }
    if (c) {
      Fy = F;
      Fy_x = F_x;
      Fy_y = F_y;
    } else {
      Fx = F;
      Fx_x = F_x;
      Fx_y = F_y;
    }
  }
}

/* edge kick xp += a(dp) x, yp -= b(dp) y with a = tan(e)/rho, b = tan(e - psi/(1+dp))/rho,
 * rho = (1+dp) rho_actual (src/csbend.c:1126-1134, :1355-1363); undo and pull back.
 * With fse1 = 1/(1 + FSE + FSE_DIPOLE), 1/rho_actual and psi both change by
 * the factor fse1 per unit FSE; their effect is added to gfse. */
EH_HD void eh_adj_edge(double &x, double &xp, double &y, double &yp, double dp, double &lx,
                                            double lxp, double &ly, double lyp, double &ldp, double tan_e, double e,
                                            double psi, double rho_act, double fse1 = 0, double *gfse = nullptr) {
  double op = 1 + dp;
  double rho = op * rho_act;
  double a = tan_e / rho;
  double u = e - psi / op;
  double tu = eh_tan(u);
  double b = tu / rho;
  double a_dp = -a / op;
  double b_dp = (1 + tu * tu) * (psi / (op * op)) / rho - b / op;
  xp -= a * x;
  yp += b * y;
  lx += a * lxp;
  ly += -b * lyp;
  ldp += lxp * x * a_dp - lyp * y * b_dp;
  if (gfse) {
    const double a_f = a * fse1;
    const double b_f = (1 + tu * tu) * (-psi * fse1 / op) / rho + b * fse1;
    *gfse += lxp * x * a_f - lyp * y * b_f;
  }
}

/* a field component from a term list: sum of coef x^i y^j */
EH_HD double eh_terms_eval(const eh_fterm *__restrict__ t, int n, double x, double y) {
  double F = 0;
  for (int k = 0; k < n; k++) {
    double v = t[k].coef;
    for (int m = 0; m < t[k].i; m++)
      v *= x;
    for (int m = 0; m < t[k].j; m++)
      v *= y;
    F += v;
  }
  return F;
}

EH_HD void eh_adj_csbend(eh_p6 &z, eh_p6 &l, const eh_desc &d, const eh_fterm *__restrict__ terms,
                          eh_pgrad &g) {
  const double rho0 = d.rho0, rho_a = d.rho_actual;
  const bool want_k = d.adj_pk >= 0, want_f = d.adj_pfse >= 0;
  double *gf = want_f ? &g.fse : nullptr;
  /* exit offset, then rotation back into the bend frame */
  eh_adj_offset(z, l, d.dxf, d.dyf, d.dzf);
  {
    /* forward: x = X c - Y s, y = X s + Y c: the rotation with sine -sin_tt */
    eh_adj_rot(z, l, d.cos_tt, -d.sin_tt);
  }
  double X = z.x, Y = z.y, xp = z.xp, yp = z.yp, DP = z.dp;
  double lX = l.x, lY = l.y, lxp = l.xp, lyp = l.yp, lDP = l.dp;
  const double ls = l.s;
  if (d.csb_edge2)
    eh_adj_edge(X, xp, Y, yp, DP, lX, lxp, lY, lyp, lDP, d.tan_e2, d.e2, d.psi2, rho_a, d.adj_fse1, gf);
  /* slopes from the canonical momenta at the exit */
  double QX, QY;
  eh_s2m(xp, yp, DP, QX, QY);
  double lQX = 0, lQY = 0;
  eh_adj_m2s(QX, QY, DP, lxp, lyp, lQX, lQY, lDP);
  double s = z.s;
  /* The reverse sweep needs accuracy, not the forward map's roundings (the
   * recovered states are re-anchored every turn), so divisions become
   * multiplications by reciprocals formed once per sub-step. */
  const double irho0 = 1 / rho0, irho_a = 1 / rho_a;
  for (int i = d.csb_nslices - 1; i >= 0; i--) {
    for (int j = d.nsub - 1; j >= 0; j--) {
      if (j < d.nsub - 1) {
        /* kick: QX -= ds h1 Fy / rho_a, QY += ds h1 Fx / rho_a with h1 = 1 + X/rho0 */
        double ds = d.csb_ds[j];
        double Fx, Fy, Fx_x, Fx_y, Fy_x, Fy_y;
        eh_csb_fields_d(d, terms, X, Y, Fx, Fy, Fx_x, Fx_y, Fy_x, Fy_y);
        double h1 = 1 + X * irho0;
        double c = ds * irho_a;
        lX += c * (-lQX * (Fy * irho0 + h1 * Fy_x) + lQY * (Fx * irho0 + h1 * Fx_x));
        lY += c * (-lQX * h1 * Fy_y + lQY * h1 * Fx_y);
        if (want_k) {
          /* the field per unit K1 */
          const double Fxk = eh_terms_eval(terms + d.adj_dk_fx0, d.adj_dk_nfx, X, Y);
          const double Fyk = eh_terms_eval(terms + d.adj_dk_fy0, d.adj_dk_nfy, X, Y);
          g.k += c * h1 * (-lQX * Fyk + lQY * Fxk);
        }
        if (want_f) {
          /* the field per unit FSE, and 1/rho_actual growing by the factor adj_fse1 */
          const double Fxf = eh_terms_eval(terms + d.adj_df_fx0, d.adj_df_nfx, X, Y);
          const double Fyf = eh_terms_eval(terms + d.adj_df_fy0, d.adj_df_nfy, X, Y);
          g.fse += c * h1 * (-lQX * (Fyf + Fy * d.adj_fse1) + lQY * (Fxf + Fx * d.adj_fse1));
        }
        QX += c * h1 * Fy;
        QY -= c * h1 * Fx;
      }
      /* exact drift through angle th: undo (the same flow through -th), then pull back */
      const double th = d.csb_dshr[j];
      const double sth = d.adj_sin_dshr[j];
      double f = eh_sqrt(eh_sqr(1 + DP) - eh_sqr(QY));
      const double rf = 1 / f;
      double sps = QX * rf;
      double psi = eh_asin(sps);
      double cps = eh_cos(psi);
      double sph, cph;
      eh_sincos(psi - th, &sph, &cph);
      const double rcps = 1 / cps, rcph = 1 / cph;
      double R = (rho0 + X) * cps * rcph; /* rho0 + X at the sub-step input */
      double F = R * sth * rf * rcps;     /* Y and path increment per unit QY, (1+DP) */
      double QXin = f * sph;
      X = R - rho0;
      Y = Y - QY * F;
      s = s - (1 + DP) * F;
      /* adjoint at the input (X, QXin, QY, DP) */
      double lF = QY * lY + (1 + DP) * ls;
      double lR = lX * (cph * rcps) + lF * sth * rf * rcps;
      double tps = sps * rcps, tph = sph * rcph;
      double lphi = lQX * f * cps + lX * R * sth * rcps * rcps + lF * F * tps;
      double lf = lQX * sps - lF * F * rf - lphi * tph * rf;
      double nlQX = lphi * rf * rcph;
      lQY = lQY + lY * F - lf * QY * rf;
      lDP = lDP + ls * F + lf * (1 + DP) * rf;
      lQX = nlQX;
      lX = lR;
      QX = QXin;
    }
  }
  /* entry: canonical momenta from slopes */
  eh_m2s_raw(QX, QY, DP, xp, yp);
  lxp = 0;
  lyp = 0;
  eh_adj_s2m(xp, yp, DP, lQX, lQY, lxp, lyp, lDP);
  if (d.csb_edge1)
    eh_adj_edge(X, xp, Y, yp, DP, lX, lxp, lY, lyp, lDP, d.tan_e1, d.e1, d.psi1, rho_a, d.adj_fse1, gf);
  z.x = X;
  z.y = Y;
  z.xp = xp;
  z.yp = yp;
  z.s = s;
  z.dp = DP;
  l.x = lX;
  l.y = lY;
  l.xp = lxp;
  l.yp = lyp;
  l.dp = lDP;
  l.s = ls;
  /* forward: x_ = x c + y s, y_ = -x s + y c */
  eh_adj_rot(z, l, d.cos_tt, d.sin_tt);
  eh_adj_offset(z, l, d.dxi, d.dyi, d.dzi);
}

/* ---- RF cavity (matrix method), by forward-mode dual numbers ------------------- */

struct eh_dual {
  double v, d;
};
EH_HD eh_dual operator+(eh_dual a, eh_dual b) { return {a.v + b.v, a.d + b.d}; }
EH_HD eh_dual operator-(eh_dual a, eh_dual b) { return {a.v - b.v, a.d - b.d}; }
EH_HD eh_dual operator*(eh_dual a, eh_dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
EH_HD eh_dual operator/(eh_dual a, eh_dual b) {
  double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
EH_HD eh_dual operator+(eh_dual a, double b) { return {a.v + b, a.d}; }
EH_HD eh_dual operator+(double a, eh_dual b) { return {a + b.v, b.d}; }
EH_HD eh_dual operator-(eh_dual a, double b) { return {a.v - b, a.d}; }
EH_HD eh_dual operator-(double a, eh_dual b) { return {a - b.v, -b.d}; }
EH_HD eh_dual operator*(eh_dual a, double b) { return {a.v * b, a.d * b}; }
EH_HD eh_dual operator*(double a, eh_dual b) { return {a * b.v, a * b.d}; }
EH_HD eh_dual operator/(eh_dual a, double b) { return {a.v / b, a.d / b}; }
EH_HD eh_dual operator/(double a, eh_dual b) {
  double q = a / b.v;
  return {q, -q * b.d / b.v};
}
EH_HD eh_dual dsqrt(eh_dual a) {
  double r = eh_sqrt(a.v);
  return {r, a.d / (2 * r)};
}
EH_HD eh_dual dsin(eh_dual a) {
  double s, c;
  eh_sincos(a.v, &s, &c);
  return {s, c * a.d};
}
EH_HD eh_dual dlog(eh_dual a) { return {eh_log(a.v), a.d / a.v}; }

/* the matrix-method body of eh_rfca for a survivor, on dual numbers
 * (src/simple_rfca.c:824-967) */
EH_HD void eh_rfca_dual(eh_dual v[6], const eh_desc &d) {
  const double c_mks = 2.99792458e8;
  const double Pc = d.Po;
  eh_dual &x = v[0], &xp = v[1], &y = v[2], &yp = v[3], &s = v[4], &dp = v[5];
  if (d.rf_zero_volt) {
    double L = d.rf_flen;
    x = x + xp * L;
    y = y + yp * L;
    s = s + L * dsqrt(1 + xp * xp + yp * yp);
    return;
  }
  const double length = d.rf_len;
  x = x - d.rf_dx;
  y = y - d.rf_dy;
  eh_dual P = Pc * (1 + dp);
  eh_dual gamma = dsqrt(P * P + 1);
  eh_dual beta_i = P / gamma;
  eh_dual ds1 = length / 2 * dsqrt(1 + xp * xp + yp * yp);
  eh_dual t = (s + ds1) / (c_mks * beta_i);
  eh_dual dgamma = {d.rf_dgamma_same, 0};
  if (!d.rf_same)
    dgamma = d.rf_volt * dsin(d.rf_omega * t + d.rf_phase);
  if (d.rf_end1 && length) {
    eh_dual iF = dgamma / (2 * gamma * length);
    xp = xp - x * iF;
    yp = yp - y * iF;
  }
  eh_dual g1 = gamma + dgamma;
  eh_dual dP = dsqrt(g1 * g1 - 1) - P;
  eh_dual R22 = 1 / (1 + dP / P);
  eh_dual R12;
  if (eh_fabs(dP.v / P.v) > 1e-14)
    R12 = length * (P / dP * dlog(1 + dP / P));
  else
    R12 = {length, 0};
  s = s + ds1;
  eh_dual x0 = x, xp0 = xp;
  x = x0 + xp0 * R12;
  xp = xp0 * R22;
  x0 = y;
  xp0 = yp;
  y = x0 + xp0 * R12;
  yp = xp0 * R22;
  s = s + length / 2 * dsqrt(1 + xp * xp + yp * yp);
  dp = (P + dP - Pc) / Pc;
  if (d.rf_end2 && length) {
    eh_dual iF = (-1.0 * dgamma) / (2 * g1 * length);
    xp = xp - x * iF;
    yp = yp - y * iF;
  }
  s = (P + dP) / g1 * s / beta_i;
  x = x + d.rf_dx;
  y = y + d.rf_dy;
}

/* l <- J^T l at the stored entry state zin (z becomes zin) */
EH_HD void eh_adj_rfca(eh_p6 &z, eh_p6 &l, const eh_p6 &zin, const eh_desc &d) {
  const double in[6] = {zin.x, zin.xp, zin.y, zin.yp, zin.s, zin.dp};
  const double lo[6] = {l.x, l.xp, l.y, l.yp, l.s, l.dp};
  double li[6] = {0, 0, 0, 0, 0, 0};
  /* one column of the Jacobian at a time, not unrolled: the cavity is one
   * element per turn, and six inlined copies of the dual-number map would set
   * the register budget of the whole reverse-sweep kernel */
#pragma unroll 1
  for (int j = 0; j < 6; j++) {
    eh_dual v[6];
#pragma unroll
    for (int k = 0; k < 6; k++)
      v[k] = {in[k], k == j ? 1.0 : 0.0};
    eh_rfca_dual(v, d);
    double acc = 0;
#pragma unroll
    for (int k = 0; k < 6; k++)
      acc += v[k].d * lo[k];
#pragma unroll
    for (int k = 0; k < 6; k++)
      li[k] = k == j ? acc : li[k];
  }
  l = {li[0], li[1], li[2], li[3], li[4], li[5]};
  z = zin;
}
