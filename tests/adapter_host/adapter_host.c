/* Test harness (not part of the library): the elegant adapter
 * (elegant_amd/adapter/gpu_adapter.c) compiled against elegant's own struct
 * definitions (a header the test extracts from src/track.h into a temporary
 * directory), with elegant's CPU routines and the library's entry points
 * replaced by recording stand-ins: a host array plays the device beam and
 * eh_beam_row_order returns a scripted permutation.  Each ah_* scenario returns
 * 0 on success or the number of the check that failed. */
#include "gpu_adapter.c"

/* ---- recording stand-ins for the library ---- */
struct eh_beam {
  long cap, n, alive;
  double rows[64][EH_NPROP];
};
static struct {
  long uploads, downloads, mp2, csb, rf, drift, tp, lim, elim, rip, off, scm, match, pass_set;
  long last_pass;
  EH_MAXAMP last_amp; /* the aperture of the last element call */
  long linked;        /* cavity calls that found their group's phase set */
  int32_t last_rf_mode;
  int perm[64]; /* download row r <- uploaded row perm[r] */
  long lose;    /* particles each tracking call loses */
} M;

int eh_init(int device) { return device == 0 ? EH_OK : EH_ERR_ARG; }
const char *eh_last_error(void) { return "stand-in"; }
eh_beam *eh_beam_create(long n) {
  eh_beam *b = (eh_beam *)calloc(1, sizeof(eh_beam));
  b->cap = n;
  return b;
}
void eh_beam_free(eh_beam *b) { free(b); }
long eh_beam_upload(eh_beam *b, double **coord, long n) {
  long k;
  if (n > b->cap || n > 64)
    return EH_ERR_ARG;
  for (k = 0; k < n; k++)
    memcpy(b->rows[k], coord[k], sizeof(double) * EH_NPROP);
  b->n = b->alive = n;
  M.uploads++;
  return n;
}
long eh_beam_download(eh_beam *b, double **coord, long n) {
  long r;
  if (n != b->n)
    return EH_ERR_ARG;
  for (r = 0; r < n; r++)
    memcpy(coord[r], b->rows[M.perm[r]], sizeof(double) * EH_NPROP);
  M.downloads++;
  return b->alive;
}
long eh_beam_row_order(eh_beam *b, int32_t *rows, long n) {
  long r;
  if (n != b->n)
    return EH_ERR_ARG;
  for (r = 0; r < n; r++)
    rows[r] = M.perm[r];
  return n;
}
long eh_beam_n_alive(const eh_beam *b) { return b->alive; }
void eh_beam_set_pass(eh_beam *b, long pass) {
  (void)b;
  M.pass_set++;
  M.last_pass = pass;
}
static long track_stand_in(eh_beam *b, long n) {
  long k;
  if (n != b->alive)
    return EH_ERR_ARG;
  for (k = 0; k < b->n; k++)
    b->rows[k][0] += 1; /* "tracking" */
  b->alive -= M.lose;
  return b->alive;
}
long eh_multipole_tracking2(eh_beam *b, long n, const EH_ELEMENT_LIST *e, double pe, double Po, double z,
                            const EH_MAXAMP *a, long iSlice) {
  (void)e, (void)pe, (void)Po, (void)z, (void)iSlice;
  M.mp2++;
  if (a)
    M.last_amp = *a;
  return track_stand_in(b, n);
}
long eh_track_through_csbend(eh_beam *b, long n, const EH_CSBEND *c, double pe, double Po, double z,
                             const EH_MAXAMP *a, long iSlice) {
  (void)c, (void)pe, (void)Po, (void)z, (void)a, (void)iSlice;
  M.csb++;
  return track_stand_in(b, n);
}
/* the library's phase-reference table: one group */
static struct {
  int64_t ref;
  int set;
  double phase;
} LT;
long eh_get_phase_reference(int64_t ref, double *phase) {
  if (!ref)
    return 0;
  if (ref != LT.ref)
    return EH_REF_PHASE_NONEXISTENT;
  if (!LT.set)
    return EH_REF_PHASE_NOT_SET;
  *phase = LT.phase;
  return EH_REF_PHASE_RETURNED;
}
long eh_set_phase_reference(int64_t ref, double phase) {
  LT.ref = ref;
  LT.set = 1;
  LT.phase = phase;
  return 1;
}
long eh_delete_phase_references(void) {
  LT.set = 0;
  return EH_OK;
}
long eh_simple_rf_cavity(eh_beam *b, long n, EH_RFCA *rf, double *P, double zEnd) {
  double ph;
  (void)P, (void)zEnd;
  if (rf->fiducial_mode == EH_FID_FIRST)
    return EH_ERR_ARG; /* as the library: FIRST never finds its particle */
  M.rf++;
  M.last_rf_mode = rf->fiducial_mode;
  if (rf->phase_reference && eh_get_phase_reference(rf->phase_reference, &ph) == EH_REF_PHASE_RETURNED) {
    M.linked++;
    return track_stand_in(b, n);
  }
  rf->fiducial_seen = 1;
  rf->phase_fiducial = 1.25;
  rf->t_fiducial = 2.5;
  if (rf->phase_reference)
    eh_set_phase_reference(rf->phase_reference, rf->phase_fiducial);
  return track_stand_in(b, n);
}
long eh_exactDrift(eh_beam *b, long n, double L) {
  (void)L;
  M.drift++;
  return track_stand_in(b, n);
}
long eh_track_particles(eh_beam *b, long n, const EH_VMATRIX *m) {
  (void)m;
  M.tp++;
  return track_stand_in(b, n);
}
long eh_limit_amplitudes(eh_beam *b, double x, double y, long n, double z, double Po, long ex, long oc) {
  (void)x, (void)y, (void)z, (void)Po, (void)ex;
  if (oc)
    return EH_ERR_UNSUPPORTED;
  M.lim++;
  return track_stand_in(b, n);
}
long eh_elimit_amplitudes(eh_beam *b, double x, double y, long n, double z, double Po, long ex, long oc, long xe,
                          long ye) {
  (void)x, (void)y, (void)z, (void)Po, (void)ex, (void)ye;
  if (oc)
    return EH_ERR_UNSUPPORTED;
  if (xe < 2 || xe % 2)
    return EH_ERR_ARG;
  M.elim++;
  return track_stand_in(b, n);
}
long eh_removeInvalidParticles(eh_beam *b, long n, double z, double Po) {
  (void)z, (void)Po;
  M.rip++;
  return track_stand_in(b, n);
}
long eh_offset_beam(eh_beam *b, long n, const EH_MALIGN *m, double P) {
  (void)P;
  if (m->on_pass != -1)
    return EH_ERR_ARG;
  M.off++;
  return track_stand_in(b, n);
}
long eh_set_central_momentum(eh_beam *b, long n, double Pn, double *P) {
  *P = Pn;
  M.scm++;
  return track_stand_in(b, n);
}
long eh_do_match_energy(eh_beam *b, long n, double *P, long change_beam) {
  if (n != b->alive)
    return EH_ERR_ARG;
  if (!change_beam)
    *P += 1.0;
  M.match++;
  return n;
}

/* ---- recording stand-ins for elegant's routines ---- */
static struct {
  long bomb, mp2, csb, rf, tp, lim, elim, rip, off, fid, match, cent, center, sums, imp;
  int on_gpu_during; /* getElementOnGpu() seen inside the last CPU routine */
  double **coord_seen;
} E;
ELEMENT_DESCRIPTION entity_description[N_TYPES];
int totalPropertiesPerParticle = EH_NPROP, globalLossCoordOffset = -1;
void bombElegant(const char *error, const char *usage) {
  (void)error, (void)usage;
  E.bomb++;
}
static long cpu_stand_in(double **coord, long n) {
  E.on_gpu_during = (int)getElementOnGpu();
  E.coord_seen = coord;
  return n;
}
long multipole_tracking2(double **particle, long n_part, ELEMENT_LIST *elem, double p_error, double Po,
                         double **accepted, double z_start, MAXAMP *maxamp, APCONTOUR *apcontour,
                         APERTURE_DATA *apFileData, double *sigmaDelta2, long iSlice) {
  E.mp2++;
  return cpu_stand_in(particle, n_part);
}
long track_through_csbend(double **part, long n_part, CSBEND *csbend, double p_error, double Po, double **accepted,
                          double z_start, double *sigmaDelta2, char *rootname, MAXAMP *maxamp, APCONTOUR *apContour,
                          APERTURE_DATA *apFileData, long iSlice, ELEMENT_LIST *eptr) {
  E.csb++;
  return cpu_stand_in(part, n_part);
}
long trackRfCavityWithWakes(double **part, long np, RFCA *rfca, double **accepted, double *P_central, double zEnd,
                            long iPass, RUN *run, CHARGE *charge, WAKE *wake, TRWAKE *trwake, LSCKICK *LSCKick,
                            long wakesAtEnd) {
  E.rf++;
  return cpu_stand_in(part, np);
}
void exactDrift(double **part, long np, double length) { cpu_stand_in(part, np); }
void track_particles(double **final, VMATRIX *Mx, double **initial, long n_part) {
  E.tp++;
  cpu_stand_in(final, n_part);
}
long limit_amplitudes(double **coord, double xmax, double ymax, long np, double **accepted, double z, double P_central,
                      long extrapolate_z, long openCode, ELEMENT_LIST *eptr) {
  E.lim++;
  return cpu_stand_in(coord, np);
}
long elimit_amplitudes(double **coord, double xmax, double ymax, long np, double **accepted, double z,
                       double P_central, long extrapolate_z, long openCode, long exponent, long yexponent,
                       ELEMENT_LIST *eptr) {
  E.elim++;
  return cpu_stand_in(coord, np);
}
long removeInvalidParticles(double **coord, long np, double **accepted, double z, double Po) {
  E.rip++;
  return cpu_stand_in(coord, np);
}
void offset_beam(double **coord, long n_to_track, MALIGN *offset, double P_central) {
  E.off++;
  cpu_stand_in(coord, n_to_track);
}
void set_central_momentum(double **coord, long np, double P_new, double *P_central) { cpu_stand_in(coord, np); }
double findFiducialTime(double **part, long np, double s0, double sOffset, double p0, unsigned long mode) {
  E.fid++;
  cpu_stand_in(part, np);
  return 7.0;
}
void do_match_energy(double **coord, long np, double *P_central, long change_beam) {
  E.match++;
  cpu_stand_in(coord, np);
}
void compute_centroids(double *centroid, double **coordinates, long n_part) {
  E.cent++;
  cpu_stand_in(coordinates, n_part);
}
void center_beam(double **part, CENTER *center, long np, long iPass, double P0) {
  E.center++;
  cpu_stand_in(part, np);
}
void accumulate_beam_sums(BEAM_SUMS *sums, double **coords, long n_part, double p_central, double mp_charge,
                          double *timeValue, double tMin, double tMax, long startPID, long endPID,
                          unsigned long flags) {
  E.sums++;
  cpu_stand_in(coords, n_part);
}
long imposeApertureData(double **coord, long np, double **accepted, double z, double Po, APERTURE_DATA *apData,
                        ELEMENT_LIST *eptr) {
  E.imp++;
  return cpu_stand_in(coord, np);
}
unsigned long parseFiducialMode(char *mode) {
  if (!mode)
    return FID_MODE_TMEAN;
  if (!strcmp(mode, "light"))
    return FID_MODE_LIGHT;
  if (!strcmp(mode, "tmean"))
    return FID_MODE_TMEAN;
  if (!strcmp(mode, "first"))
    return FID_MODE_FIRST;
  if (!strcmp(mode, "pmaximum"))
    return FID_MODE_PMAX;
  return 0;
}
/* elegant's phase-reference table: one group */
static struct {
  long ref, set;
  double phase;
} ET;
long get_phase_reference(double *phase, long ref) {
  if (!ref)
    return 0;
  if (ref != ET.ref)
    return REF_PHASE_NONEXISTENT;
  if (!ET.set)
    return REF_PHASE_NOT_SET;
  *phase = ET.phase;
  return REF_PHASE_RETURNED;
}
long set_phase_reference(long ref, double phase) {
  ET.ref = ref;
  ET.set = 1;
  ET.phase = phase;
  return 1;
}

/* ---- fixtures ---- */
static double rows_h[16][EH_NPROP], acc_h[16][EH_NPROP];
static double *coord_h[16], *accp_h[16];

static void reset(long n) {
  long k, q;
  memset(&M, 0, sizeof M);
  memset(&E, 0, sizeof E);
  memset(&LT, 0, sizeof LT);
  memset(&ET, 0, sizeof ET);
  for (k = 0; k < 64; k++)
    M.perm[k] = (int)k;
  for (k = 0; k < 16; k++) {
    for (q = 0; q < EH_NPROP; q++) {
      rows_h[k][q] = 100 * k + q;
      acc_h[k][q] = -(100 * k + q);
    }
    coord_h[k] = rows_h[k];
    accp_h[k] = acc_h[k];
  }
  gpuBaseInit(coord_h, n, accp_h, NULL, 1);
}

static KQUAD kq;
static ELEMENT_LIST el_kq, el_other, el_kq2;

static void default_kquad(KQUAD *q) {
  memset(q, 0, sizeof *q);
  q->length = 0.3;
  q->k1 = 1.2;
  q->xKickCalibration = q->yKickCalibration = 1;
  q->nSlices = 1;
  q->integration_order = 4;
  q->isr1Particle = 1;
  q->edge1Linear = q->edge2Linear = 1;
  q->edge1NonlinearFactor = q->edge2NonlinearFactor = 1;
  q->systematicMultipoleFactor = q->randomMultipoleFactor = q->steeringMultipoleFactor = 1;
  q->minMultipoleOrder[0] = q->minMultipoleOrder[1] = q->maxMultipoleOrder[0] = q->maxMultipoleOrder[1] = -1;
}

static void setup_elements(void) {
  default_kquad(&kq);
  memset(&el_kq, 0, sizeof el_kq);
  el_kq.type = T_KQUAD;
  el_kq.p_elem = (char *)&kq;
  el_kq2 = el_kq;
  memset(&el_other, 0, sizeof el_other);
  el_other.type = T_WATCH; /* no GPU routine */
}

/* a KQUAD on the GPU, then an element on the CPU: one upload, one download in
 * the scripted row order, accepted permuted alike, the pass counted */
int ah_scenario_residency(void) {
  long n, k;
  reset(8);
  setup_elements();
  setElementGpuData(&el_kq);
  if (!getElementOnGpu())
    return 1;
  M.lose = 2;
  for (k = 0; k < 8; k++)
    M.perm[k] = (int)(7 - k); /* download reverses the rows */
  n = gpu_multipole_tracking2(8, &el_kq, 0.0, 5000.0, accp_h, 0.0, NULL, NULL, NULL, NULL, -1);
  if (n != 6 || M.mp2 != 1 || M.uploads != 1 || M.downloads != 0)
    return 2;
  setElementGpuData(&el_other);
  if (getElementOnGpu() || M.downloads != 1)
    return 3;
  /* row r holds uploaded row 7 - r, tracked (x + 1); accepted moved alike */
  for (k = 0; k < 8; k++) {
    if (coord_h[k][0] != 100 * (7 - k) + 1 || coord_h[k][1] != 100 * (7 - k) + 1)
      return 4;
    if (accp_h[k][0] != -(100 * (7 - k)) || accp_h[k][6] != -(100 * (7 - k) + 6))
      return 5;
  }
  /* back on the GPU: the host rows (changed by the CPU element) are uploaded */
  setElementGpuData(&el_kq2);
  M.lose = 0;
  for (k = 0; k < 8; k++)
    M.perm[k] = (int)k;
  n = gpu_multipole_tracking2(6, &el_kq2, 0.0, 5000.0, accp_h, 0.0, NULL, NULL, NULL, NULL, -1);
  if (n != 6 || M.uploads != 2)
    return 6;
  /* second pass: the first element comes round again */
  setElementGpuData(&el_kq);
  n = gpu_multipole_tracking2(6, &el_kq, 0.0, 5000.0, accp_h, 0.0, NULL, NULL, NULL, NULL, -1);
  if (M.last_pass != 1 || M.uploads != 2)
    return 7;
  gpuBaseDealloc();
  if (M.downloads != 2 || E.bomb)
    return 8;
  return 0;
}

/* a hook called with an argument the library does not carry runs elegant's
 * CPU routine on the host copy, with getElementOnGpu() 0 during the call, and
 * the device copy is refreshed before the next GPU routine */
int ah_scenario_host_fallback(void) {
  double sd2 = 0;
  long n;
  reset(8);
  setup_elements();
  setElementGpuData(&el_kq);
  n = gpu_multipole_tracking2(8, &el_kq, 0.0, 5000.0, accp_h, 0.0, NULL, NULL, NULL, NULL, -1);
  if (n != 8 || M.uploads != 1)
    return 1;
  n = gpu_multipole_tracking2(8, &el_kq, 0.0, 5000.0, accp_h, 0.0, NULL, NULL, NULL, &sd2, -1);
  if (E.mp2 != 1 || E.on_gpu_during != 0 || E.coord_seen != coord_h || M.downloads != 1 || !getElementOnGpu())
    return 2;
  n = gpu_multipole_tracking2(8, &el_kq, 0.0, 5000.0, accp_h, 0.0, NULL, NULL, NULL, NULL, 3);
  if (E.mp2 != 2)
    return 3;
  n = gpu_multipole_tracking2(8, &el_kq, 0.0, 5000.0, accp_h, 0.0, NULL, NULL, NULL, NULL, -1);
  if (M.mp2 != 2 || M.uploads != 2)
    return 4;
  {
    MAXAMP amp;
    memset(&amp, 0, sizeof amp);
    amp.x_max = amp.y_max = 0.02;
    amp.elliptical = 1;
    amp.exponent = 4;
    amp.yExponent = 2;
    n = gpu_multipole_tracking2(8, &el_kq, 0.0, 5000.0, accp_h, 0.0, &amp, NULL, NULL, NULL, -1);
    if (M.mp2 != 3 || E.mp2 != 2 || M.last_amp.elliptical != 1 || M.last_amp.exponent != 4 ||
        M.last_amp.y_exponent != 2)
      return 5;
    /* an open side stays on the host */
    amp.openSide = (char *)"+x";
    n = gpu_multipole_tracking2(8, &el_kq, 0.0, 5000.0, accp_h, 0.0, &amp, NULL, NULL, NULL, -1);
    if (E.mp2 != 3)
      return 51;
    amp.openSide = NULL;
    /* the elliptical sweep runs on the device; open sides and exponents the
     * library rejects reach elegant's routine */
    n = gpu_elimit_amplitudes(0.02, 0.02, 8, accp_h, 0.0, 5000.0, 0, 0, 2, 2, NULL);
    if (M.elim != 1 || E.elim != 0)
      return 6;
    n = gpu_elimit_amplitudes(0.02, 0.02, 8, accp_h, 0.0, 5000.0, 0, 1, 2, 2, NULL);
    if (E.elim != 1)
      return 61;
    n = gpu_elimit_amplitudes(0.02, 0.02, 8, accp_h, 0.0, 5000.0, 0, 0, 3, 0, NULL);
    if (E.elim != 2 || M.elim != 1)
      return 62;
    amp.elliptical = 0;
    n = gpu_multipole_tracking2(8, &el_kq, 0.0, 5000.0, accp_h, 0.0, &amp, NULL, NULL, NULL, -1);
    if (M.mp2 != 4)
      return 7;
  }
  /* a third-order matrix stays on the host; order 2 runs on the device */
  {
    VMATRIX m3;
    memset(&m3, 0, sizeof m3);
    m3.order = 3;
    gpu_track_particles(&m3, 8);
    if (E.tp != 1 || M.tp != 0)
      return 8;
  }
  n = gpu_limit_amplitudes(0.02, 0.02, 8, accp_h, 0.0, 5000.0, 0, 1);
  if (E.lim != 1)
    return 9;
  n = gpu_limit_amplitudes(0.02, 0.02, 8, accp_h, 0.0, 5000.0, 0, 0);
  if (M.lim != 1)
    return 10;
  if (gpu_findFiducialTime(8, 0, 0, 5000.0, FID_MODE_TMEAN) != 7.0 || E.fid != 1)
    return 11;
  /* momentum matching runs on the device copy */
  {
    double P = 5000.0;
    gpu_do_match_energy(8, &P, 0);
    if (M.match != 1 || E.match != 0 || P != 5001.0)
      return 13;
  }
  gpuBaseDealloc();
  return E.bomb ? 12 : 0;
}

/* the RF cavity's fiducial state is written back to elegant's element */
int ah_scenario_rfca(void) {
  RFCA rf;
  ELEMENT_LIST e;
  double P = 5000.0;
  long n;
  reset(4);
  memset(&rf, 0, sizeof rf);
  rf.length = 0.3;
  rf.volt = 2e6;
  rf.phase = 170;
  rf.freq = 6e8;
  rf.tReference = -1;
  memset(&e, 0, sizeof e);
  e.type = T_RFCA;
  e.p_elem = (char *)&rf;
  setElementGpuData(&e);
  if (!getElementOnGpu())
    return 1;
  n = gpu_trackRfCavityWithWakes(4, &rf, accp_h, &P, 1.0, 0, NULL, NULL, NULL, NULL, NULL, 0);
  if (n != 4 || M.rf != 1 || rf.fiducial_seen != 1 || rf.phase_fiducial != 1.25 || rf.t_fiducial != 2.5)
    return 2;
  /* PHASE_REFERENCE 3: the first cavity of the group sets elegant's table,
   * a second one tracks with that phase and keeps its own fiducial unset */
  {
    RFCA rf2 = rf;
    rf.phase_reference = rf2.phase_reference = 3;
    rf.fiducial_seen = 0;
    rf2.fiducial_seen = 0;
    rf2.phase_fiducial = 0;
    if (!element_on_gpu(&e))
      return 3;
    n = gpu_trackRfCavityWithWakes(4, &rf, accp_h, &P, 1.0, 0, NULL, NULL, NULL, NULL, NULL, 0);
    if (n != 4 || M.rf != 2 || E.rf != 0 || ET.ref != 3 || !ET.set || ET.phase != 1.25)
      return 4;
    n = gpu_trackRfCavityWithWakes(4, &rf2, accp_h, &P, 1.0, 0, NULL, NULL, NULL, NULL, NULL, 0);
    if (M.rf != 3 || M.linked != 1 || rf2.fiducial_seen != 0)
      return 41;
    /* elegant deleted its phase references (a new step): the library's copy
     * is not used */
    ET.set = 0;
    rf2.phase_fiducial = 0;
    n = gpu_trackRfCavityWithWakes(4, &rf2, accp_h, &P, 1.0, 0, NULL, NULL, NULL, NULL, NULL, 0);
    if (M.linked != 1 || rf2.fiducial_seen != 1 || !ET.set)
      return 42;
  }
  rf.phase_reference = 0;
  /* FIRST: the library fails as elegant's own routine does, which then reports it */
  rf.fiducial = "first";
  n = gpu_trackRfCavityWithWakes(4, &rf, accp_h, &P, 1.0, 0, NULL, NULL, NULL, NULL, NULL, 0);
  if (E.rf != 1)
    return 5;
  rf.fiducial = "pmaximum";
  n = gpu_trackRfCavityWithWakes(4, &rf, accp_h, &P, 1.0, 0, NULL, NULL, NULL, NULL, NULL, 0);
  if (M.last_rf_mode != EH_FID_PMAX || E.rf != 1)
    return 51;
  rf.fiducial = "light";
  if (!element_on_gpu(&e))
    return 6;
  gpuBaseDealloc();
  return E.bomb ? 7 : 0;
}

/* parameters the library does not carry keep the element on the CPU; the
 * CSBEND edge arrays are read through e1Index/e2Index */
int ah_scenario_fill(void) {
  KQUAD q;
  EH_KQUAD dq;
  CSBEND cb;
  EH_CSBEND db;
  MALIGN ma;
  EH_MALIGN dm;
  default_kquad(&q);
  if (!fill_kquad(&q, &dq) || dq.k1 != 1.2 || dq.integration_order != 4)
    return 1;
  q.pitch = 1e-3;
  if (fill_kquad(&q, &dq))
    return 2;
  default_kquad(&q);
  q.systematic_multipoles = "sys.sdds";
  if (fill_kquad(&q, &dq))
    return 3;
  default_kquad(&q);
  q.edge1_effects = 1;
  if (fill_kquad(&q, &dq))
    return 4;
  default_kquad(&q);
  q.lEffective = 0.28;
  if (fill_kquad(&q, &dq))
    return 5;
  default_kquad(&q);
  q.bore = 0.01; /* the pole-tip field sets the strength in the library */
  q.B = 0.4;
  if (!fill_kquad(&q, &dq) || dq.bore != 0.01 || dq.B != 0.4)
    return 6;
  memset(&cb, 0, sizeof cb);
  cb.length = 0.8;
  cb.angle = 0.03;
  cb.e[0] = 0.01;
  cb.e[1] = 0.02;
  cb.h[0] = 0.5;
  cb.h[1] = 0.25;
  cb.edge_effects[0] = 1;
  cb.edge_effects[1] = 0;
  cb.fintBoth = 0.5;
  cb.fint[0] = cb.fint[1] = -1;
  cb.nSlices = 8;
  cb.integration_order = 4;
  cb.nonlinear = 1;
  cb.edge_order = 1;
  cb.etiltSign = 1;
  cb.e1Index = 0;
  cb.e2Index = 1;
  if (!fill_csbend(&cb, &db) || db.e1 != 0.01 || db.h2 != 0.25 || db.edge1_effects != 1 || db.fint != 0.5)
    return 7;
  cb.e1Index = 1;
  cb.e2Index = 0; /* reversed element */
  if (!fill_csbend(&cb, &db) || db.e1 != 0.02 || db.e2 != 0.01 || db.h1 != 0.25 || db.edge1_effects != 0 ||
      db.edge2_effects != 1)
    return 8;
  cb.fint[0] = 0.4; /* FINT1 differs from FINT2 (= FINT) */
  if (fill_csbend(&cb, &db))
    return 9;
  cb.fint[0] = -1;
  cb.edge_effects[0] = 2; /* Hwang's fringe map is carried */
  if (!fill_csbend(&cb, &db) || db.edge2_effects != 2)
    return 10;
  cb.edge_effects[0] = 5; /* Lindberg's curved-dipole fringe is not */
  if (fill_csbend(&cb, &db))
    return 10;
  cb.edge_effects[0] = 1;
  cb.isr = 1;
  if (fill_csbend(&cb, &db))
    return 11;
  cb.isr = 0;
  cb.f2 = 1e-4;
  if (fill_csbend(&cb, &db))
    return 12;
  memset(&ma, 0, sizeof ma);
  ma.dx = 1e-4;
  ma.on_pass = 3;
  ma.startPID = ma.endPID = -1;
  if (!fill_malign(&ma, &dm) || dm.on_pass != -1 || dm.dx != 1e-4)
    return 13;
  ma.startPID = 5;
  if (fill_malign(&ma, &dm))
    return 14;
  return 0;
}

/* second-order maps are packed T[i][j(j+1)/2 + k] for k <= j */
int ah_scenario_vmatrix(void) {
  double C[6], R[6][6], T[6][6][6], *Rp[6], **Tp[6], *Tq[6][6];
  VMATRIX m;
  EH_VMATRIX d;
  int i, j, k;
  for (i = 0; i < 6; i++) {
    C[i] = i;
    Rp[i] = R[i];
    Tp[i] = Tq[i];
    for (j = 0; j < 6; j++) {
      R[i][j] = 10 * i + j;
      Tq[i][j] = T[i][j];
      for (k = 0; k < 6; k++)
        T[i][j][k] = 100 * i + 10 * j + k;
    }
  }
  memset(&m, 0, sizeof m);
  m.C = C;
  m.R = Rp;
  m.T = Tp;
  m.order = 2;
  if (!fill_vmatrix(&m, &d) || d.order != 2)
    return 1;
  for (i = 0; i < 6; i++) {
    if (d.C[i] != i)
      return 2;
    for (j = 0; j < 6; j++) {
      if (d.R[i][j] != 10 * i + j)
        return 3;
      for (k = 0; k <= j; k++)
        if (d.T[i][j * (j + 1) / 2 + k] != 100 * i + 10 * j + k)
          return 4;
    }
  }
  m.order = 1;
  if (!fill_vmatrix(&m, &d) || d.T[5][20] != 0)
    return 5;
  return 0;
}
