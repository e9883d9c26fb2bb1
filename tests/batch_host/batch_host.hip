/* Test harness (not part of the library): the batched path's host-side tables
 * (batch_tables, batch_rf_update, batch_tables_reusable) and its work list
 * (eh_batch_wave_starts, eh_batch_item_variant) run on the host CPU, and a span
 * of k_batch is emulated with the same work list, descriptor tables and
 * per-particle maps, taking the work items in a shuffled order (the order in
 * which wavefronts happen to pull them from the device queue).  It includes the
 * library source to reach those functions. */
#include "eh_lib.hip"

#include <random>

extern "C" long bh_worklist(const int32_t *alive, int nvar, int32_t *wave_start, int32_t *item_variant) {
  int32_t total = eh_batch_wave_starts(alive, nvar, wave_start);
  for (int32_t w = 0; w < total; w++)
    item_variant[w] = eh_batch_item_variant(wave_start, nvar, w);
  return total;
}

/* Planted mistake (CPU only): when set, bh_track leaves the dense field-block
 * offsets of every pool descriptor where build_csbend put them for the
 * element's own term list, i.e. pool_push shifts csb_fx0/csb_fy0 but not
 * csb_dx0/csb_dy0.  That was the state of the batched tables while the dense
 * CSBEND blocks were being introduced (their shift came with them later); the
 * test shows what it does to a tracked batch. */
static int g_plant_unshifted_dense = 0;
extern "C" void bh_plant_unshifted_dense(int on) { g_plant_unshifted_dense = on; }

static void plant_unshift(std::vector<eh_desc> &pool) {
  if (!g_plant_unshifted_dense)
    return;
  for (eh_desc &d : pool)
    if (d.csb_dense && d.csb_dx0 != 0) {
      /* build_csbend lays the dense blocks out first: dx0 = 0, dy0 = eo1^2 */
      d.csb_dy0 -= d.csb_dx0;
      d.csb_dx0 = 0;
    }
}

/* the lean map's answer for a descriptor of the given type (with an elliptical
 * aperture if ell), and whether the host would send it to the lean map */
extern "C" long bh_lean_stage(int type, int ell, int *capable) {
  eh_desc d;
  std::memset(&d, 0, sizeof d);
  d.type = type;
  d.adj_pk = d.adj_pdx = d.adj_pdy = d.adj_pfse = d.adj_anchor = -1;
  if (ell) {
    d.ap_ell[0] = d.ap_ell[1] = d.post_ell[0] = d.post_ell[1] = d.ell_any = d.full_only = 1;
    d.ap_xmax[0] = d.ap_xmax[1] = d.ap_ymax[0] = d.ap_ymax[1] = 1.0;
    d.ell_xe[0] = d.ell_xe[1] = d.ell_ye[0] = d.ell_ye[1] = 2;
  }
  *capable = eh_lean_capable(d);
  eh_p6 p = {1e-3, 0, 0, 0, 0, 0};
  return eh_apply_element_t<false>(p, d, 0, nullptr, d.zpos);
}

/* the loss code of track_span and k_batch decoded back (host check of the
 * encoding) */
extern "C" void bh_loss_words(int dpass, int nelem, int ee, int st, int p0, int32_t event_base, int32_t *lpass,
                              int32_t *levent) {
  eh_loss_words(eh_loss_code(dpass, nelem, ee, st), nelem, p0, event_base, *lpass, *levent);
}

/* a pool descriptor equals a descriptor built alone for the same position:
 * the same bytes except that its field-term offsets are shifted by where its
 * terms start in the shared term list, and the same terms there */
static bool same_desc(const eh_desc &got, const std::vector<eh_fterm> &terms, const eh_desc &ref,
                      const std::vector<eh_fterm> &ref_terms) {
  const int32_t shift = ref_terms.empty() ? 0 : got.csb_fx0 - ref.csb_fx0;
  eh_desc x = got;
  if (!ref_terms.empty()) {
    x.csb_fx0 -= shift;
    x.csb_fy0 -= shift;
    if (x.csb_dense) {
      x.csb_dx0 -= shift;
      x.csb_dy0 -= shift;
    }
  }
  if (std::memcmp(&x, &ref, sizeof x) != 0)
    return false;
  if (ref_terms.empty())
    return true;
  if (shift < 0 || (size_t)shift + ref_terms.size() > terms.size())
    return false;
  return std::memcmp(terms.data() + shift, ref_terms.data(), sizeof(eh_fterm) * ref_terms.size()) == 0;
}

/* Runs the table sequence of eh_batch_track on the host: batch_tables, then for
 * each RF cavity position with a pending fiducial in some variant, a fiducial
 * set from T_REFERENCE-like data (phase_fiducial = -2 pi f * t_set[v]) and
 * batch_rf_update.  Then every (variant, position) of the result is compared
 * with a descriptor built afresh for that position.  Returns the number of
 * mismatching entries, or a negative code.  reusable_before/after report
 * batch_tables_reusable with the tables' own hash before and after the fiducials
 * were set. */
extern "C" long bh_tables(EH_ELEMENT_LIST *const *beamlines, long nvar, long ne, double Pc, const double *t_set,
                          int32_t *reusable_before, int32_t *reusable_after, int32_t *n_pool) {
  std::vector<eh_desc> pool;
  std::vector<eh_fterm> terms;
  std::vector<int32_t> idx;
  long r = batch_tables(beamlines, nvar, ne, Pc, pool, terms, idx);
  if (r < 0)
    return r;
  uint64_t h0 = beamlines_hash(beamlines, nvar, ne, Pc);
  *reusable_before = batch_tables_reusable(true, ne, h0, beamlines, nvar, ne, h0);
  for (long er = 0; er < ne; er++) {
    bool pending = false;
    for (long v = 0; v < nvar; v++)
      pending = pending || rf_pending(beamlines[v] + er);
    if (!pending)
      continue;
    for (long v = 0; v < nvar; v++) {
      if (!rf_pending(beamlines[v] + er))
        continue;
      EH_RFCA *rf = (EH_RFCA *)beamlines[v][er].p_elem;
      r = rf_take_fiducial(rf, 0.0, [&](int, double, double *t0) {
        *t0 = t_set[v];
        return (long)EH_OK;
      });
      if (r < 0)
        return r;
    }
    r = batch_rf_update(beamlines, nvar, ne, er, Pc, pool, terms, idx);
    if (r < 0)
      return r;
  }
  uint64_t h1 = beamlines_hash(beamlines, nvar, ne, Pc);
  *reusable_after = batch_tables_reusable(true, ne, h1, beamlines, nvar, ne, h1);
  *n_pool = (int32_t)pool.size();
  long bad = 0;
  std::vector<eh_ctx> ctx;
  for (long v = 0; v < nvar; v++) {
    line_contexts(beamlines[v], ne, Pc, ctx);
    for (long e = 0; e < ne; e++) {
      eh_desc ref;
      std::vector<eh_fterm> ft;
      r = build_desc(&ref, beamlines[v] + e, ctx[e], ft);
      if (r < 0)
        return r;
      int32_t k = idx[(size_t)v * ne + e];
      if (k < 0 || k >= (int32_t)pool.size()) {
        bad++;
        continue;
      }
      if (!same_desc(pool[k], terms, ref, ft))
        bad++;
    }
  }
  return bad;
}

/* host emulation of one k_batch span over SoA buffers: items in a shuffled
 * order, each wavefront's lanes tracked through (p0, e0) -> (p1, e1) with the
 * pool descriptors, survivors written at region_off[v] + a per-variant counter
 * in lane order, lost particles appended to the graveyard */
struct bh_soa {
  std::vector<double> c[6];
  std::vector<int32_t> row0, lpass, levent;
  void resize(long n) {
    for (auto &v : c)
      v.assign(n, 0.0);
    row0.assign(n, 0);
    lpass.assign(n, 0);
    levent.assign(n, -1);
  }
};

static void bh_span(const bh_soa &a, bh_soa &out, bh_soa &grave, long &n_grave, const std::vector<long> &region_off,
                    std::vector<int32_t> &alive, const std::vector<eh_desc> &pool, const std::vector<eh_fterm> &terms,
                    const std::vector<int32_t> &idx, int nelem, int p0, int e0, int p1, int e1, int32_t event_base,
                    std::mt19937 &rng) {
  const int nvar = (int)alive.size();
  std::vector<int32_t> ws(nvar + 1);
  const int32_t total = eh_batch_wave_starts(alive.data(), nvar, ws.data());
  std::vector<int32_t> items(total);
  for (int32_t w = 0; w < total; w++)
    items[w] = w;
  std::shuffle(items.begin(), items.end(), rng);
  std::vector<int32_t> alive_out(nvar, 0);
  for (int32_t w : items) {
    const int v = eh_batch_item_variant(ws.data(), nvar, w);
    for (int lane = 0; lane < 64; lane++) {
      const long local = (long)(w - ws[v]) * 64 + lane;
      if (local >= alive[v])
        continue;
      const long i = region_off[v] + local;
      eh_p6 p = {a.c[0][i], a.c[1][i], a.c[2][i], a.c[3][i], a.c[4][i], a.c[5][i]};
      bool live = true;
      int32_t lpass = 0, levent = -1, lcode = 0;
      int pp = p0, ee = e0;
      while (live && (pp < p1 || (pp == p1 && ee < e1))) {
        const eh_desc &d = pool[idx[(size_t)v * nelem + ee]];
        int st = eh_apply_element_t<true>(p, d, pp, terms.data(), d.zpos);
        if (st >= 0) {
          live = false;
          lcode = eh_loss_code(pp - p0, nelem, ee, st);
          eh_loss_words(lcode, nelem, p0, event_base, lpass, levent);
        }
        if (++ee == nelem) {
          ee = 0;
          pp++;
        }
      }
      bh_soa &dst = live ? out : grave;
      const long j = live ? region_off[v] + alive_out[v]++ : n_grave++;
      const double pv[6] = {p.x, p.xp, p.y, p.yp, p.s, p.dp};
      for (int q = 0; q < 6; q++)
        dst.c[q][j] = pv[q];
      dst.row0[j] = a.row0[i];
      dst.lpass[j] = live ? 0 : lpass;
      dst.levent[j] = live ? -1 : levent;
    }
  }
  alive = alive_out;
}

/* Tracks nvar variants n_passes passes through the host emulation with the
 * span boundaries of eh_batch_track (pass 0 stops before each RF cavity with a
 * pending fiducial; the rest in spans of span_elements elements).  RF cavities
 * must use the LIGHT fiducial or T_REFERENCE (a TMEAN fiducial needs the
 * reference row order, which the device path replays on download).  Writes per
 * particle (in input order) its final 6 coordinates and its loss pass or -1
 * (alive); a loss pass with the bit 0x40000000 set marks a kick-multipole entry
 * loss whose s the device path settles on download.  Returns the survivors. */
extern "C" long bh_track(const double *flat, const long *counts, long nvar, EH_ELEMENT_LIST *const *beamlines,
                         long ne, long n_passes, double Pc, long span_elements, unsigned seed, double *out6,
                         int32_t *loss_pass) {
  std::vector<long> region_off(nvar), id_off(nvar);
  long off = 0, idn = 0;
  for (long v = 0; v < nvar; v++) {
    region_off[v] = off;
    id_off[v] = idn;
    off += (counts[v] + 63) / 64 * 64;
    idn += counts[v];
  }
  const long cap = off > 0 ? off : 64;
  bh_soa buf[2], grave;
  buf[0].resize(cap);
  buf[1].resize(cap);
  grave.resize(cap);
  std::vector<int32_t> alive(nvar);
  for (long v = 0; v < nvar; v++) {
    alive[v] = (int32_t)counts[v];
    for (long i = 0; i < counts[v]; i++) {
      for (int q = 0; q < 6; q++)
        buf[0].c[q][region_off[v] + i] = flat[(id_off[v] + i) * EH_NPROP + q];
      buf[0].row0[region_off[v] + i] = (int32_t)(id_off[v] + i);
    }
  }
  std::vector<eh_desc> pool;
  std::vector<eh_fterm> terms;
  std::vector<int32_t> idx;
  long r = batch_tables(beamlines, nvar, ne, Pc, pool, terms, idx);
  if (r < 0)
    return r;
  plant_unshift(pool);
  std::mt19937 rng(seed);
  int cur = 0, p = 0, e = 0;
  long n_grave = 0;
  int32_t event = 0;
  auto span = [&](int p1, int e1) {
    int32_t base = event;
    event += (int32_t)(((long)(p1 - p) * ne + ne) * 3 + 3);
    bh_span(buf[cur], buf[1 - cur], grave, n_grave, region_off, alive, pool, terms, idx, (int)ne, p, e, p1, e1, base,
            rng);
    cur = 1 - cur;
  };
  if (n_passes > 0) {
    for (long er = 0; er < ne; er++) {
      bool pending = false;
      for (long v = 0; v < nvar; v++)
        pending = pending || rf_pending(beamlines[v] + er);
      if (!pending)
        continue;
      if (e < er)
        span(0, (int)er);
      e = (int)er;
      for (long v = 0; v < nvar; v++) {
        if (!rf_pending(beamlines[v] + er))
          continue;
        double zEnd = 0;
        for (long k = 0; k <= er; k++)
          zEnd += elem_length(beamlines[v] + k);
        r = rf_take_fiducial((EH_RFCA *)beamlines[v][er].p_elem, zEnd, [&](int, double, double *) {
          return fail(EH_ERR_UNSUPPORTED, "bh_track: TMEAN fiducial needs the device path");
        });
        if (r < 0)
          return r;
      }
      r = batch_rf_update(beamlines, nvar, ne, er, Pc, pool, terms, idx);
      if (r < 0)
        return r;
      plant_unshift(pool);
    }
  }
  while (p < n_passes) {
    long pos = (long)p * ne + e + span_elements;
    long endpos = (long)n_passes * ne;
    if (pos > endpos)
      pos = endpos;
    int p1 = (int)(pos / ne), e1 = (int)(pos % ne);
    span(p1, e1);
    p = p1;
    e = e1;
  }
  long total = 0;
  for (long v = 0; v < nvar; v++) {
    total += alive[v];
    for (long k = 0; k < alive[v]; k++) {
      long j = region_off[v] + k, id = buf[cur].row0[j];
      for (int q = 0; q < 6; q++)
        out6[id * 6 + q] = buf[cur].c[q][j];
      loss_pass[id] = -1;
    }
  }
  for (long k = 0; k < n_grave; k++) {
    long id = grave.row0[k];
    for (int q = 0; q < 6; q++)
      out6[id * 6 + q] = grave.c[q][k];
    loss_pass[id] = grave.lpass[k]; /* with EH_LPASS_INHERIT: s is settled on download */
  }
  if (total + n_grave != idn)
    return fail(EH_ERR_STATE, "bh_track: %ld alive + %ld lost != %ld", total, n_grave, idn);
  return total;
}
