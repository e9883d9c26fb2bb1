/* Test harness (not part of the library): runs the library's per-particle maps
 * and their reverse-mode adjoints on the host CPU, with the same turn/element
 * loop, anchors and re-anchoring as the GPU kernels k_adj_fwd / k_adj_bwd, so
 * that the adjoint arithmetic can be checked against PyTorch autograd without a
 * GPU.  It includes the library source to reach its descriptor builder. */
#include "eh_lib.hip"

extern "C" long adjh_gradient(const double *coord, long n, EH_ELEMENT_LIST *bl, long ne, const int32_t *slots,
                              long np, long turns, double Pc, const double *grad_final, double *final_out,
                              int32_t *alive_out, double *grad_params, double *grad_z0, double *max_err) {
  std::vector<eh_desc> hd;
  std::vector<eh_fterm> pool;
  long r = build_line(bl, ne, Pc, hd, pool);
  if (r < 0)
    return r;
  int n_anchor = 1;
  for (long e = 0; e < ne; e++) {
    if (bl[e].type == EH_T_RFCA && e > 0)
      hd[e].adj_anchor = n_anchor++;
    hd[e].adj_pk = slots[EH_ADJ_KINDS * e];
    hd[e].adj_pdx = slots[EH_ADJ_KINDS * e + 1];
    hd[e].adj_pdy = slots[EH_ADJ_KINDS * e + 2];
    hd[e].adj_pfse = slots[EH_ADJ_KINDS * e + 3];
    if (bl[e].type == EH_T_CSBEND) {
      r = adj_bend_params(hd[e], (const EH_CSBEND *)bl[e].p_elem, pool);
      if (r < 0)
        return r;
    }
  }
  std::fill(grad_params, grad_params + np, 0.0);
  *max_err = 0;
  std::vector<eh_p6> anchors((size_t)turns * n_anchor);
  for (long i = 0; i < n; i++) {
    eh_p6 p = {coord[10 * i], coord[10 * i + 1], coord[10 * i + 2], coord[10 * i + 3], coord[10 * i + 4],
               coord[10 * i + 5]};
    bool live = true;
    for (long t = 0; t < turns && live; t++)
      for (long e = 0; e < ne && live; e++) {
        int slot = e == 0 ? 0 : hd[e].adj_anchor;
        if (slot >= 0)
          anchors[t * n_anchor + slot] = p;
        if (eh_apply_element(p, hd[e], t, pool.data()) >= 0)
          live = false;
      }
    alive_out[i] = live;
    final_out[6 * i] = p.x;
    final_out[6 * i + 1] = p.xp;
    final_out[6 * i + 2] = p.y;
    final_out[6 * i + 3] = p.yp;
    final_out[6 * i + 4] = p.s;
    final_out[6 * i + 5] = p.dp;
    for (int c = 0; c < 6; c++)
      grad_z0[6 * i + c] = 0;
    if (!live)
      continue;
    eh_p6 z = p;
    const double *g = grad_final + 6 * i;
    eh_p6 l = {g[0], g[1], g[2], g[3], g[4], g[5]};
    for (long t = turns - 1; t >= 0; t--)
      for (long e = ne - 1; e >= 0; e--) {
        const eh_desc &d = hd[e];
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
          eh_pgrad pg = {0, 0, 0, 0};
          eh_adj_multipole(z, l, d, pg);
          if (d.adj_pk >= 0)
            grad_params[d.adj_pk] += pg.k;
          if (d.adj_pdx >= 0)
            grad_params[d.adj_pdx] += pg.dx;
          if (d.adj_pdy >= 0)
            grad_params[d.adj_pdy] += pg.dy;
          break;
        }
        case EH_T_CSBEND: {
          eh_pgrad pg = {0, 0, 0, 0};
          eh_adj_csbend(z, l, d, pool.data(), pg);
          if (d.adj_pk >= 0)
            grad_params[d.adj_pk] += pg.k;
          if (d.adj_pfse >= 0)
            grad_params[d.adj_pfse] += pg.fse;
          break;
        }
        case EH_T_RFCA:
          eh_adj_rfca(z, l, anchors[t * n_anchor + (e == 0 ? 0 : d.adj_anchor)], d);
          break;
        default:
          break;
        }
        if (e == 0) {
          const eh_p6 &za = anchors[t * n_anchor];
          double err = std::max(std::max(std::fabs(z.x - za.x), std::fabs(z.xp - za.xp)),
                                std::max(std::fabs(z.y - za.y), std::fabs(z.yp - za.yp)));
          *max_err = std::max(*max_err, err);
          z = za;
        }
      }
    grad_z0[6 * i] = l.x;
    grad_z0[6 * i + 1] = l.xp;
    grad_z0[6 * i + 2] = l.y;
    grad_z0[6 * i + 3] = l.yp;
    grad_z0[6 * i + 4] = l.s;
    grad_z0[6 * i + 5] = l.dp;
  }
  return n;
}

/* the de-duplicated descriptor tables of the batched path describe every
 * variant's element exactly as the single-lattice path does: same fields (term
 * offsets aside) and the same field values at sample points.  Returns the
 * number of mismatching (variant, element) pairs. */
extern "C" long adjh_batch_tables_check(EH_ELEMENT_LIST *const *beamlines, long nvar, long ne, double Pc) {
  std::vector<eh_desc> pool;
  std::vector<eh_fterm> terms;
  std::vector<int32_t> idx;
  long r = batch_tables(beamlines, nvar, ne, Pc, pool, terms, idx);
  if (r < 0)
    return r;
  long bad = 0;
  const double pts[4][2] = {{0, 0}, {1e-3, -2e-3}, {-4e-3, 1e-3}, {2.5e-3, 3e-3}};
  for (long v = 0; v < nvar; v++) {
    std::vector<eh_desc> hd;
    std::vector<eh_fterm> own;
    r = build_line(beamlines[v], ne, Pc, hd, own);
    if (r < 0)
      return r;
    for (long e = 0; e < ne; e++) {
      eh_desc a = pool[idx[(size_t)v * ne + e]], b = hd[e];
      if (beamlines[v][e].type == EH_T_RFCA)
        continue; /* placeholder until the fiducial is known */
      bool ok = true;
      if (a.type == EH_T_CSBEND && a.csb_field) {
        for (int k = 0; k < 4; k++) {
          double fx1, fy1, fx2, fy2;
          eh_csb_fields(a, terms.data(), pts[k][0], pts[k][1], fx1, fy1);
          eh_csb_fields(b, own.data(), pts[k][0], pts[k][1], fx2, fy2);
          if (!(fx1 == fx2 && fy1 == fy2))
            ok = false;
        }
      }
      a.csb_fx0 = b.csb_fx0 = a.csb_fy0 = b.csb_fy0 = a.csb_dx0 = b.csb_dx0 = a.csb_dy0 = b.csb_dy0 = 0;
      if (std::memcmp(&a, &b, sizeof a) != 0)
        ok = false;
      bad += !ok;
    }
  }
  return bad;
}

/* eh_div_recip against IEEE division on n random (a, b) pairs: half with random
 * operands over wide exponent ranges (significands near 1 and near 2 included),
 * half with quotients placed next to a rounding midpoint, the hardest case for
 * a division.  Returns the number of results that differ in any bit. */
extern "C" long adjh_div_recip_check(long n, uint64_t seed) {
  uint64_t st = seed | 1;
  auto rnd = [&]() {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return st;
  };
  auto mk = [](uint64_t u, int e, uint64_t m) {
    return eh_from_bits(((uint64_t)(1023 + e) << 52) | (m & 0xfffffffffffffULL)) * ((u >> 63) ? -1.0 : 1.0);
  };
  long bad = 0;
  for (long k = 0; k < n; k++) {
    uint64_t ub = rnd(), ua = rnd();
    uint64_t mb = ub;
    if (((ub >> 56) & 3) == 1)
      mb &= 0xff;
    else if (((ub >> 56) & 3) == 2)
      mb = ~(mb & 0xff);
    double b = mk(ub, (int)((ub >> 40) % 81) - 40, mb);
    double a;
    if (k & 1) {
      double q = mk(0, (int)((ua >> 40) % 41) - 20, ua);
      double mid = q + std::ldexp((ua >> 63) ? 1.0 : -1.0, std::ilogb(q) - 53);
      a = b * mid;
      if ((ua & 7) == 0)
        a = std::nextafter(a, ((ua >> 3) & 1) ? INFINITY : -INFINITY);
    } else {
      a = mk(ua, (int)((ua >> 40) % 1801) - 900, ua);
    }
    if ((ua & 0xff) == 3)
      a = (ua & 0x100) ? 0.0 : -0.0;
    double rh = 1.0 / b, rl = std::fma(-rh, b, 1.0) / b;
    double q1 = eh_div_recip(a, b, rh, rl), q2 = a / b;
    if (eh_bits(q1) != eh_bits(q2))
      bad++;
  }
  return bad;
}
