/* gpu_adapter.c -- elegant's HAVE_GPU interface (gpu_base.h) on the gfx950
 * element-tracking library (include/elegant_hip.h).
 *
 * Built inside elegant's source tree, next to its objects:
 *
 *     cc -DHAVE_GPU -I<elegant>/src -I<this dir> -I<repo>/include -c gpu_adapter.c
 *     link elegant with gpu_adapter.o and libelegant_hip.so
 *
 * with elegant's own sources compiled -DHAVE_GPU and this directory on their
 * include path (they include gpu_base.h and the per-routine headers beside it).
 *
 * How it works.  do_tracking calls gpuBaseInit with the host particle array,
 * then setElementGpuData(eptr) before every element (src/do_tracking.c:925).
 * setElementGpuData decides whether the element runs on the GPU: its type is
 * one the library tracks and every parameter that changes the tracked
 * coordinates is one the library honours (the fill_* functions below copy each
 * such field or refuse the element when it has a value the library does not
 * carry).  For a GPU element do_tracking calls the element routine, whose
 * HAVE_GPU branch calls the gpu_<routine> defined here with the CPU routine's
 * arguments minus the particle array.  For any other element getElementOnGpu()
 * is 0 and elegant runs its CPU routine on getGpuBase()->coord.
 *
 * Two copies of the particles exist: the host array and the device beam.  Each
 * is marked current or stale; a GPU routine first makes the device current
 * (uploading the n_part rows do_tracking is tracking), a CPU element or a host
 * read first makes the host current (downloading in the reference's exact row
 * order: survivors first, lost rows above them in the order the CPU routines'
 * swap-to-top loops leave them, with the loss pass in column 7).  The caller's
 * `accepted` rows move with the particle rows (src/multipole.c:1045-1046,
 * src/limit_amplitudes.c:270-271): the adapter keeps the rows as of the last
 * upload and permutes them by eh_beam_row_order on every download.
 *
 * A gpu_<routine> called with an argument the library does not carry (an
 * aperture contour or file, sigmaDelta2, a slice index, an open-sided MAXAMP,
 * wakes, a third-order matrix, ...) runs the CPU routine itself on the host copy
 * with elementOnGpu cleared for the duration of the call.
 *
 * Passes: the loss pass of a particle lost on the GPU is the number of times
 * the element sequence has restarted since gpuBaseInit (an element seen again
 * starts a new pass), i.e. do_tracking's i_pass when its pass loop starts at 0.
 *
 * Limits: the RF fiducial is taken over all particles (elegant's default; a
 * fiducialization bunch set by setFiducializationBunch is not seen here); with
 * losses_include_global_coordinates (rows of 13 doubles, global loss
 * coordinates, src/limit_amplitudes.c:269-276) every element stays on the CPU;
 * one beam at a time; not reentrant (nor is elegant's tracking).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "track.h"

#include "elegant_hip.h"
#include "gpu_base.h"

/* defined in src/simple_rfca.c:295-306 and declared only there */
long trackRfCavityWithWakes(double **part, long np, RFCA *rfca, double **accepted, double *P_central, double zEnd,
                            long iPass, RUN *run, CHARGE *charge, WAKE *wake, TRWAKE *trwake, LSCKICK *LSCKick,
                            long wakesAtEnd);

/* ------------------------------------------------------------------------- */
/* state */

static struct {
  GPUBASE pub;
  int lib_ready;
  eh_beam *beam;
  long beam_cap;
  double **accepted; /* the caller's array (rows of EH_NPROP), or NULL */
  double *acc0;      /* accepted rows as of the last upload */
  long n_dev;        /* rows the device beam holds (alive + lost since upload) */
  int dev_ok;        /* device copy current */
  int host_ok;       /* host rows [0, n_dev) current */
  long pass;
  void **seen; /* elements visited in the current pass */
  long n_seen, cap_seen;
} G;

static void ad_fatal(const char *what) {
  char buf[640];
  snprintf(buf, sizeof buf, "gfx950 tracking library: %s (%s)", what, eh_last_error());
  bombElegant(buf, NULL);
}

static void ad_check(long r, const char *what) {
  if (r < 0)
    ad_fatal(what);
}

/* make the host rows [0, n_dev) current: download in the reference row order
 * and move the accepted rows along */
static void ad_to_host(void) {
  long k, q;
  int32_t *rows;
  if (G.host_ok)
    return;
  if (!G.dev_ok || !G.beam)
    ad_fatal("no current copy of the particles");
  ad_check(eh_beam_download(G.beam, G.pub.coord, G.n_dev), "download");
  if (G.accepted && G.acc0 && G.n_dev) {
    rows = (int32_t *)malloc(sizeof(int32_t) * G.n_dev);
    if (!rows)
      ad_fatal("out of memory");
    ad_check(eh_beam_row_order(G.beam, rows, G.n_dev), "row order");
    for (k = 0; k < G.n_dev; k++)
      for (q = 0; q < EH_NPROP; q++)
        G.accepted[k][q] = G.acc0[(long)rows[k] * EH_NPROP + q];
    free(rows);
  }
  G.host_ok = 1;
}

/* make the device hold exactly the n_part particles do_tracking tracks */
static void ad_to_device(long n_part) {
  long k;
  if (G.dev_ok) {
    if (eh_beam_n_alive(G.beam) != n_part)
      ad_fatal("particle count on the device differs from the tracked count");
    eh_beam_set_pass(G.beam, G.pass);
    return;
  }
  if (!G.host_ok)
    ad_fatal("no current copy of the particles");
  if (n_part > G.beam_cap) {
    if (G.beam)
      eh_beam_free(G.beam);
    G.beam = eh_beam_create(n_part);
    if (!G.beam)
      ad_fatal("device allocation");
    G.beam_cap = n_part;
  }
  ad_check(eh_beam_upload(G.beam, G.pub.coord, n_part), "upload");
  if (G.accepted) {
    free(G.acc0);
    G.acc0 = (double *)malloc(sizeof(double) * EH_NPROP * (n_part > 0 ? n_part : 1));
    if (!G.acc0)
      ad_fatal("out of memory");
    for (k = 0; k < n_part; k++)
      memcpy(G.acc0 + k * EH_NPROP, G.accepted[k], sizeof(double) * EH_NPROP);
  }
  G.n_dev = n_part;
  G.dev_ok = 1;
  eh_beam_set_pass(G.beam, G.pass);
}

/* after a GPU routine changed the particles */
static void ad_device_changed(void) { G.host_ok = 0; }

/* run a CPU routine on the host copy from inside a GPU hook */
static double **ad_host_begin(void) {
  ad_to_host();
  G.pub.elementOnGpu = 0;
  return G.pub.coord;
}

static void ad_host_end(int changed) {
  G.pub.elementOnGpu = 1;
  if (changed)
    G.dev_ok = 0;
}

/* ------------------------------------------------------------------------- */
/* element parameters.  Each fill_* copies every field of the element struct
 * (src/track.h) that changes the tracked coordinates into the library's struct
 * or returns 0 (the element stays on the CPU) when the field has a value the
 * library does not honour.  The *_untracked lists name the fields that do not
 * change tracked coordinates (correction flags, matrix-only options, caches
 * derived from the fields above, file contents read only with a file name). */

static int no_file(const char *name) { return name == NULL || name[0] == 0; }

/* the error-multipole and steering-multipole files of KQUAD/KSEXT/KOCT
 * (src/multipole.c:766-791) are not carried */
static const char *const kquad_untracked[] __attribute__((unused)) = {
    "xSteering", "ySteering", "synchRadInOrdinaryMatrix", "trackingBasedMatrix", "sqrtOrder",
    "systematicMultipoleFactor", "randomMultipoleFactor", "steeringMultipoleFactor", "minMultipoleOrder",
    "maxMultipoleOrder", "multipolesInitialized", "totalMultipolesComputed", "systematicMultipoleData",
    "edgeMultipoleData", "randomMultipoleData", "totalMultipoleData", "steeringMultipoleData", "isr1Particle",
    "fringeIntP", "fringeIntM", "edge1Linear", "edge2Linear", "edge1NonlinearFactor", "edge2NonlinearFactor", NULL};

/* KQUAD, src/track.h:2204-2228, as multipole_tracking2 reads it
 * (src/multipole.c:726-791): the fringe integrals
 * and edge factors only with EDGE1/2_EFFECTS, the multipole factors and order
 * limits only with multipole files */
static int fill_kquad(const KQUAD *q, EH_KQUAD *d) {
  if (q->pitch != 0 || q->yaw != 0 || q->isr || q->edge1_effects || q->edge2_effects ||
      q->lEffective > 0 || q->radial || q->malignMethod != 0)
    return 0;
  if (!no_file(q->systematic_multipoles) || !no_file(q->edge_multipoles) || !no_file(q->random_multipoles) ||
      !no_file(q->steering_multipoles))
    return 0;
  memset(d, 0, sizeof *d);
  d->length = q->length;
  d->k1 = q->k1;
  d->tilt = q->tilt;
  d->dx = q->dx;
  d->dy = q->dy;
  d->dz = q->dz;
  d->fse = q->fse;
  d->bore = q->bore;
  d->B = q->B;
  d->xkick = q->xkick;
  d->ykick = q->ykick;
  d->xKickCalibration = q->xKickCalibration;
  d->yKickCalibration = q->yKickCalibration;
  d->n_kicks = (int32_t)q->n_kicks;
  d->nSlices = (int32_t)q->nSlices;
  d->integration_order = q->integration_order;
  d->expandHamiltonian = q->expandHamiltonian;
  d->synch_rad = q->synch_rad;
  return 1;
}

static const char *const ksext_untracked[] __attribute__((unused)) = {
    "xSteering", "ySteering", "sqrtOrder", "systematicMultipoleFactor", "randomMultipoleFactor",
    "steeringMultipoleFactor", "minMultipoleOrder", "maxMultipoleOrder", "multipolesInitialized",
    "totalMultipolesComputed", "systematicMultipoleData", "edgeMultipoleData", "randomMultipoleData",
    "totalMultipoleData", "steeringMultipoleData", "isr1Particle", NULL};

/* KSEXT, src/track.h:2149-2166, src/multipole.c:796-862 */
static int fill_ksext(const KSEXT *q, EH_KSEXT *d) {
  if (q->pitch != 0 || q->yaw != 0 || q->isr || q->malignMethod != 0)
    return 0;
  if (!no_file(q->systematic_multipoles) || !no_file(q->edge_multipoles) || !no_file(q->random_multipoles) ||
      !no_file(q->steering_multipoles))
    return 0;
  memset(d, 0, sizeof *d);
  d->length = q->length;
  d->k2 = q->k2;
  d->k1 = q->k1;
  d->j1 = q->j1;
  d->tilt = q->tilt;
  d->dx = q->dx;
  d->dy = q->dy;
  d->dz = q->dz;
  d->fse = q->fse;
  d->bore = q->bore;
  d->B = q->B;
  d->xkick = q->xkick;
  d->ykick = q->ykick;
  d->xKickCalibration = q->xKickCalibration;
  d->yKickCalibration = q->yKickCalibration;
  d->n_kicks = (int32_t)q->n_kicks;
  d->nSlices = (int32_t)q->nSlices;
  d->integration_order = q->integration_order;
  d->expandHamiltonian = q->expandHamiltonian;
  d->synch_rad = q->synch_rad;
  return 1;
}

static const char *const koct_untracked[] __attribute__((unused)) = {"sqrtOrder", "multipolesInitialized", "totalMultipolesComputed",
                                              "systematicMultipoleData", "randomMultipoleData",
                                              "totalMultipoleData", "isr1Particle", NULL};

/* KOCT, src/track.h:2171-2182, src/multipole.c:867-922 */
static int fill_koct(const KOCT *q, EH_KOCT *d) {
  if (q->pitch != 0 || q->yaw != 0 || q->isr || q->malignMethod != 0)
    return 0;
  if (!no_file(q->systematic_multipoles) || !no_file(q->random_multipoles))
    return 0;
  memset(d, 0, sizeof *d);
  d->length = q->length;
  d->k3 = q->k3;
  d->tilt = q->tilt;
  d->dx = q->dx;
  d->dy = q->dy;
  d->dz = q->dz;
  d->fse = q->fse;
  d->bore = q->bore;
  d->B = q->B;
  d->n_kicks = (int32_t)q->n_kicks;
  d->nSlices = (int32_t)q->nSlices;
  d->integration_order = q->integration_order;
  d->expandHamiltonian = q->expandHamiltonian;
  d->synch_rad = q->synch_rad;
  return 1;
}

static const char *const kquse_untracked[] __attribute__((unused)) = {"isr1Particle", NULL};

/* KQUSE, src/track.h:3498-3503, src/multipole.c:923-949; MATRIX_TRACKING
 * (matrixTracking != 0) tracks through the element's matrix instead */
static int fill_kquse(const KQUSE *q, EH_KQUSE *d) {
  if (q->isr || q->matrixTracking)
    return 0;
  memset(d, 0, sizeof *d);
  d->length = q->length;
  d->k1 = q->k1;
  d->k2 = q->k2;
  d->tilt = q->tilt;
  d->dx = q->dx;
  d->dy = q->dy;
  d->dz = q->dz;
  d->fse1 = q->fse1;
  d->fse2 = q->fse2;
  d->n_kicks = (int32_t)q->n_kicks;
  d->nSlices = (int32_t)q->nSlices;
  d->integration_order = q->integration_order;
  d->expandHamiltonian = q->expandHamiltonian;
  d->synch_rad = q->synch_rad;
  return 1;
}

static const char *const csbend_untracked[] __attribute__((unused)) = {
    "synchRadInOrdinaryMatrix", "sqrtOrder", "photonLowEnergyCutoff", "xSteering", "ySteering", "b", "c",
    "fseCorrectionValue", "fseCorrectionPathError", "refTrajectoryChangeSet", "refLength", "refAngle",
    "refTrajectoryChange", "refSlices", "photonFileActive", "SDDSphotons", "expansionOrder_ref", "nonlinear_ref",
    "expansionOrder1_ref", "hasSkew_ref", "hasNormal_ref", "h_ref", "b_ref", "c_ref", "Fx_xy_ref", "Fy_xy_ref",
    "isr1Particle", "fringeInt", NULL};

/* CSBEND, src/track.h:2462-2502, as track_through_csbend reads it
 * (src/csbend.c:654-1000).  e[], h[], edge_effects[], fint[] are read through
 * e1Index/e2Index (a reversed element swaps them, :859-876); Lindberg's fringe
 * integrals only act with edge method 5, which the library does not carry
 * (and refuses); the b/c arrays and the
 * *_ref fields are caches of the strengths above; the reference-trajectory
 * fields act only with REFERENCE_CORRECTION and the FSE correction value only
 * with FSE_CORRECTION, both refused */
static int fill_csbend(const CSBEND *cb, EH_CSBEND *d) {
  double f1, f2;
  int i1 = cb->e1Index, i2 = cb->e2Index;
  if (i1 < 0 || i1 > 1 || i2 < 0 || i2 > 1)
    return 0;
  if (cb->epitch != 0 || cb->eyaw != 0 || cb->isr || cb->distributionBasedRadiation || cb->includeOpeningAngle ||
      !no_file(cb->photonOutputFile) || cb->referenceCorrection || cb->trackingMatrix || cb->fseCorrection ||
      cb->edgeFlip)
    return 0;
  if (cb->edge_effects[i1] < 0 || cb->edge_effects[i2] < 0 || cb->edge_effects[i1] > 4 || cb->edge_effects[i2] > 4)
    return 0;
  if (cb->xReference != 0 || cb->f1 || cb->f2 || cb->f3 || cb->f4 || cb->f5 || cb->f6 || cb->f7 || cb->f8 || cb->g1 ||
      cb->g2 || cb->g3 || cb->g4 || cb->g5 || cb->g6 || cb->g7 || cb->g8)
    return 0;
  /* one FINT for both edges in the library: FINT1/FINT2 (if set) must agree */
  f1 = cb->fint[i1] >= 0 ? cb->fint[i1] : cb->fintBoth;
  f2 = cb->fint[i2] >= 0 ? cb->fint[i2] : cb->fintBoth;
  if (f1 != f2)
    return 0;
  memset(d, 0, sizeof *d);
  d->length = cb->length;
  d->angle = cb->angle;
  d->k1 = cb->k1;
  d->k2 = cb->k2;
  d->k3 = cb->k3;
  d->k4 = cb->k4;
  d->k5 = cb->k5;
  d->k6 = cb->k6;
  d->k7 = cb->k7;
  d->k8 = cb->k8;
  d->e1 = cb->e[i1];
  d->e2 = cb->e[i2];
  d->h1 = cb->h[i1];
  d->h2 = cb->h[i2];
  d->edge1_effects = cb->edge_effects[i1];
  d->edge2_effects = cb->edge_effects[i2];
  d->tilt = cb->tilt;
  d->hgap = cb->hgap;
  d->fint = f1;
  d->dx = cb->dx;
  d->dy = cb->dy;
  d->dz = cb->dz;
  d->xKick = cb->xKick;
  d->yKick = cb->yKick;
  d->fse = cb->fse;
  d->fseDipole = cb->fseDipole;
  d->fseQuadrupole = cb->fseQuadrupole;
  d->etilt = cb->etilt;
  d->etiltSign = cb->etiltSign;
  d->b1 = cb->b1;
  d->b2 = cb->b2;
  d->b3 = cb->b3;
  d->b4 = cb->b4;
  d->b5 = cb->b5;
  d->b6 = cb->b6;
  d->b7 = cb->b7;
  d->b8 = cb->b8;
  d->nSlices = (int32_t)cb->nSlices;
  d->nonlinear = cb->nonlinear;
  d->edge_order = cb->edge_order;
  d->integration_order = cb->integration_order;
  d->expandHamiltonian = cb->expandHamiltonian;
  d->use_bn = cb->use_bn;
  d->expansionOrder = cb->expansionOrder;
  d->malignMethod = cb->malignMethod;
  d->edgeFlags = cb->edgeFlags;
  d->synch_rad = cb->synch_rad;
  d->kick_limit_scaling = (int32_t)(cb->kick_limit_scaling != 0);
  d->edge1_kick_limit = cb->edge_kick_limit[i1];
  d->edge2_kick_limit = cb->edge_kick_limit[i2];
  return 1;
}

static const char *const rfca_untracked[] __attribute__((unused)) = {"backtrack", NULL};

/* RFCA, src/track.h:1494-1508, as trackRfCavityWithWakes reads it
 * (src/simple_rfca.c:295-985).  The fiducial state (fiducial_seen,
 * phase_fiducial, t_fiducial) is copied in and written back after the call,
 * as the reference mutates the element; PHASE_REFERENCE groups are kept in
 * elegant's own table (gpu_trackRfCavityWithWakes).  backtrack is set only by
 * elegant's backtracking of RF cavities, which the library is not called for. */
static int fill_rfca(const RFCA *rf, EH_RFCA *d) {
  unsigned long mode;
  int32_t fid;
  if (rf->bodyFocusModel && rf->bodyFocusModel[0] && strcmp(rf->bodyFocusModel, "none") != 0)
    return 0;
  mode = parseFiducialMode(rf->fiducial);
  if (mode == FID_MODE_TMEAN)
    fid = EH_FID_TMEAN;
  else if (mode == FID_MODE_LIGHT)
    fid = EH_FID_LIGHT;
  else if (mode == FID_MODE_FIRST)
    fid = EH_FID_FIRST;
  else if (mode == FID_MODE_PMAX)
    fid = EH_FID_PMAX;
  else
    return 0; /* an invalid mode: elegant's routine reports it */
  memset(d, 0, sizeof *d);
  d->length = rf->length;
  d->volt = rf->volt;
  d->phase = rf->phase;
  d->freq = rf->freq;
  d->Q = rf->Q;
  d->dx = rf->dx;
  d->dy = rf->dy;
  d->tReference = rf->tReference;
  d->nKicks = (int32_t)rf->nKicks;
  d->change_p0 = rf->change_p0;
  d->change_t = rf->change_t;
  d->end1Focus = rf->end1Focus;
  d->end2Focus = rf->end2Focus;
  d->standingWave = rf->standingWave;
  d->linearize = rf->linearize;
  d->lockPhase = rf->lockPhase;
  d->fiducial_mode = fid;
  d->phase_reference = (int64_t)rf->phase_reference;
  d->fiducial_seen = rf->fiducial_seen;
  d->phase_fiducial = rf->phase_fiducial;
  d->t_fiducial = rf->t_fiducial;
  return 1;
}

static void rfca_write_back(RFCA *rf, const EH_RFCA *d) {
  rf->fiducial_seen = (short)d->fiducial_seen;
  rf->phase_fiducial = d->phase_fiducial;
  rf->t_fiducial = d->t_fiducial;
}

static const char *const malign_untracked[] __attribute__((unused)) = {"forceModifyMatrix", "floor", NULL};

/* MALIGN, src/track.h:1953-1957, offset_beam (src/do_tracking.c:3311-3372).
 * do_tracking applies on_pass and excludeOrbit before calling offset_beam
 * (:1853-1857), so the library applies the offset on every call;
 * forceModifyMatrix and floor concern matrices and floor coordinates. */
static int fill_malign(const MALIGN *m, EH_MALIGN *d) {
  if (m->startPID != -1 || m->endPID != -1)
    return 0;
  memset(d, 0, sizeof *d);
  d->dxp = m->dxp;
  d->dyp = m->dyp;
  d->dx = m->dx;
  d->dy = m->dy;
  d->dz = m->dz;
  d->dt = m->dt;
  d->dp = m->dp;
  d->de = m->de;
  d->on_pass = -1;
  d->startPID = (int32_t)m->startPID;
  d->endPID = (int32_t)m->endPID;
  return 1;
}

static const char *const maxamp_untracked[] __attribute__((unused)) = {NULL};

/* MAXAMP, src/track.h:2008-2013, as the in-element apertures and the
 * post-element sweeps read it (setupMultApertureData, src/multipole.c:1690-1749;
 * src/do_tracking.c:2506-2517): rectangular and (super-)elliptical, no open
 * side; exponents that do not fit the library's int32 stay on the host */
static int fill_maxamp(const MAXAMP *m, EH_MAXAMP *d) {
  if (!no_file(m->openSide))
    return 0;
  if (m->exponent != (int32_t)m->exponent || m->yExponent != (int32_t)m->yExponent)
    return 0;
  memset(d, 0, sizeof *d);
  d->x_max = m->x_max;
  d->y_max = m->y_max;
  d->elliptical = (int32_t)(m->elliptical != 0);
  d->exponent = (int32_t)m->exponent;
  d->y_exponent = (int32_t)m->yExponent;
  return 1;
}

/* elements whose fields do_tracking reads itself and hands to a gpu_ routine
 * as arguments: DRIF through its matrix (eptr->matrix, order), EDRIFT's length
 * to exactDrift, ROTATE's rotation_matrix and pass selection
 * (src/do_tracking.c:1836-1852), ENERGY's target momentum and matching mode
 * (:1861-1881); MALIGN's pass selection and orbit exclusion (:1853-1857) */
static const char *const drift_host_read[] __attribute__((unused)) = {"length", "order", NULL};
static const char *const edrift_host_read[] __attribute__((unused)) = {"length", NULL};
static const char *const rotate_host_read[] __attribute__((unused)) = {"tilt", "on_pass", "excludeFloor",
                                                                      "excludeOptics", NULL};
static const char *const energy_host_read[] __attribute__((unused)) = {"central_energy", "central_momentum",
                                                                      "match_beamline", "match_particles", NULL};
static const char *const malign_host_read[] __attribute__((unused)) = {"on_pass", "excludeOrbit", NULL};

/* a transport map of order 1 or 2 (src/track.h:216-221) in the library's
 * packed form; order 3 stays on the CPU */
static int fill_vmatrix(const VMATRIX *M, EH_VMATRIX *d) {
  int i, j, k;
  if (!M || M->order < 1 || M->order > 2)
    return 0;
  memset(d, 0, sizeof *d);
  d->order = (int32_t)M->order;
  for (i = 0; i < 6; i++) {
    d->C[i] = M->C[i];
    for (j = 0; j < 6; j++)
      d->R[i][j] = M->R[i][j];
    if (M->order == 2)
      for (j = 0; j < 6; j++)
        for (k = 0; k <= j; k++)
          d->T[i][j * (j + 1) / 2 + k] = M->T[i][j][k];
  }
  return 1;
}

/* ------------------------------------------------------------------------- */
/* which elements run on the GPU */

static int element_on_gpu(ELEMENT_LIST *e) {
  if (globalLossCoordOffset > 0 || totalPropertiesPerParticle != EH_NPROP)
    return 0;
  union {
    EH_KQUAD q;
    EH_KSEXT s;
    EH_KOCT o;
    EH_KQUSE u;
    EH_CSBEND b;
    EH_RFCA r;
    EH_MALIGN m;
    EH_MAXAMP a;
  } u;
  switch (e->type) {
  case T_MARK:
    return 1; /* nothing to track */
  case T_MAXAMP:
    return fill_maxamp((MAXAMP *)e->p_elem, &u.a);
  case T_EDRIFT:
    return 1;
  case T_KQUAD:
    return fill_kquad((KQUAD *)e->p_elem, &u.q);
  case T_KSEXT:
    return fill_ksext((KSEXT *)e->p_elem, &u.s);
  case T_KOCT:
    return fill_koct((KOCT *)e->p_elem, &u.o);
  case T_KQUSE:
    return fill_kquse((KQUSE *)e->p_elem, &u.u);
  case T_CSBEND:
    return fill_csbend((CSBEND *)e->p_elem, &u.b);
  case T_RFCA:
    return fill_rfca((RFCA *)e->p_elem, &u.r);
  case T_MALIGN:
    return fill_malign((MALIGN *)e->p_elem, &u.m);
  case T_ROTATE:
    return 1; /* track_particles with rotation_matrix (src/do_tracking.c:1836-1852) */
  case T_ENERGY:
    return 1; /* set_central_momentum; the matching modes run on the host copy */
  default:
    break;
  }
  /* matrix-tracked elements (src/do_tracking.c:1239-1311; DRIF among them):
   * track_particles on the element's matrix, if it is first or second order
   * and not a monitor (monitors read the host particles) */
  if ((entity_description[e->type].flags & MATRIX_TRACKING) && !IS_MONITOR(e->type) && e->matrix &&
      e->matrix->order >= 1 && e->matrix->order <= 2 && e->matrix->C[5] == 0)
    return 1;
  return 0;
}

/* a new pass starts when an element comes round again */
static void note_element(void *eptr) {
  long i;
  for (i = 0; i < G.n_seen; i++)
    if (G.seen[i] == eptr) {
      G.pass++;
      G.n_seen = 0;
      break;
    }
  if (G.n_seen == G.cap_seen) {
    long cap = G.cap_seen ? 2 * G.cap_seen : 1024;
    void **s = (void **)realloc(G.seen, sizeof(void *) * cap);
    if (!s)
      ad_fatal("out of memory");
    G.seen = s;
    G.cap_seen = cap;
  }
  G.seen[G.n_seen++] = eptr;
}

/* ------------------------------------------------------------------------- */
/* residency (src/do_tracking.c:411-414, :924-933, :3172-3179) */

void gpuBaseInit(double **coord, long nOriginal, double **accepted, double **lostParticles, long isMaster) {
  (void)lostParticles;
  if (!G.lib_ready) {
    if (eh_init(0) != EH_OK)
      ad_fatal("no HIP device");
    G.lib_ready = 1;
  }
  G.pub.coord = coord;
  G.pub.nOriginal = nOriginal;
  G.pub.isMaster = isMaster;
  G.pub.elementOnGpu = 0;
  G.accepted = accepted;
  G.host_ok = 1;
  G.dev_ok = 0;
  G.n_dev = 0;
  G.pass = 0;
  G.n_seen = 0;
}

void gpuBaseDealloc(void) {
  if (G.dev_ok && !G.host_ok)
    ad_to_host();
  G.dev_ok = 0;
  G.pub.elementOnGpu = 0;
  free(G.acc0);
  G.acc0 = NULL;
  G.accepted = NULL;
}

GPUBASE *getGpuBase(void) { return &G.pub; }

void setElementGpuData(void *eptr) {
  ELEMENT_LIST *e = (ELEMENT_LIST *)eptr;
  note_element(eptr);
  G.pub.elementOnGpu = (short)element_on_gpu(e);
  if (!G.pub.elementOnGpu) {
    /* elegant runs its CPU routine on the host copy, which may change it */
    ad_to_host();
    G.dev_ok = 0;
  }
}

long getElementOnGpu(void) { return G.pub.elementOnGpu; }

double **forceParticlesToCpu(char *routine) {
  (void)routine;
  ad_to_host();
  G.dev_ok = 0; /* the caller works on the host copy */
  G.pub.elementOnGpu = 0;
  return G.pub.coord;
}

/* the ID sort of do_tracking (src/do_tracking.c:938-948), with accepted sorted
 * alike; comp_IDs order (src/do_tracking.c:5256-5263) */
static int ad_comp_ids(const void *a, const void *b) {
  double ia = ((const double *)a)[6], ib = ((const double *)b)[6];
  return ia < ib ? -1 : (ia > ib ? 1 : 0);
}

void sortByPID(long np) {
  ad_to_host();
  qsort(G.pub.coord[0], np, EH_NPROP * sizeof(double), ad_comp_ids);
  if (G.accepted)
    qsort(G.accepted[0], np, EH_NPROP * sizeof(double), ad_comp_ids);
  G.dev_ok = 0;
}

void startGpuTimer(void) {}
void startCpuTimer(void) {}
void displayTimings(void) {}

/* ------------------------------------------------------------------------- */
/* element routines on the GPU */

static int apertures_plain(APCONTOUR *apcontour, APERTURE_DATA *apFileData) {
  return apcontour == NULL && (apFileData == NULL || !apFileData->initialized);
}

long gpu_multipole_tracking2(long n_part, ELEMENT_LIST *elem, double p_error, double Po, double **accepted,
                             double z_start, MAXAMP *maxamp, APCONTOUR *apcontour, APERTURE_DATA *apFileData,
                             double *sigmaDelta2, long iSlice) {
  union {
    EH_KQUAD q;
    EH_KSEXT s;
    EH_KOCT o;
    EH_KQUSE u;
  } p;
  EH_MAXAMP amp;
  EH_ELEMENT_LIST el;
  int ok = 0;
  long r;
  switch (elem->type) {
  case T_KQUAD:
    ok = fill_kquad((KQUAD *)elem->p_elem, &p.q);
    el.type = EH_T_KQUAD;
    break;
  case T_KSEXT:
    ok = fill_ksext((KSEXT *)elem->p_elem, &p.s);
    el.type = EH_T_KSEXT;
    break;
  case T_KOCT:
    ok = fill_koct((KOCT *)elem->p_elem, &p.o);
    el.type = EH_T_KOCT;
    break;
  case T_KQUSE:
    ok = fill_kquse((KQUSE *)elem->p_elem, &p.u);
    el.type = EH_T_KQUSE;
    break;
  default:
    break;
  }
  if (ok && (maxamp == NULL || fill_maxamp(maxamp, &amp)) && apertures_plain(apcontour, apFileData) &&
      sigmaDelta2 == NULL && iSlice < 0 && (accepted == NULL || accepted == G.accepted)) {
    el.occurence = (int32_t)elem->occurence;
    el.name = elem->name;
    el.p_elem = &p;
    ad_to_device(n_part);
    r = eh_multipole_tracking2(G.beam, n_part, &el, p_error, Po, z_start, maxamp ? &amp : NULL, iSlice);
    if (r >= 0) {
      ad_device_changed();
      return r;
    }
    if (r != EH_ERR_UNSUPPORTED)
      ad_fatal("multipole_tracking2");
  }
  {
    double **coord = ad_host_begin();
    r = multipole_tracking2(coord, n_part, elem, p_error, Po, accepted, z_start, maxamp, apcontour, apFileData,
                            sigmaDelta2, iSlice);
    ad_host_end(1);
  }
  return r;
}

long gpu_track_through_csbend(long n_part, CSBEND *csbend, double p_error, double Po, double **accepted,
                              double z_start, double *sigmaDelta2, char *rootname, MAXAMP *maxamp,
                              APCONTOUR *apContour, APERTURE_DATA *apFileData, long iSlice, ELEMENT_LIST *eptr) {
  EH_CSBEND b;
  EH_MAXAMP amp;
  long r;
  if (fill_csbend(csbend, &b) && (maxamp == NULL || fill_maxamp(maxamp, &amp)) &&
      apertures_plain(apContour, apFileData) && sigmaDelta2 == NULL && iSlice < 0 &&
      (accepted == NULL || accepted == G.accepted)) {
    ad_to_device(n_part);
    r = eh_track_through_csbend(G.beam, n_part, &b, p_error, Po, z_start, maxamp ? &amp : NULL, iSlice);
    if (r >= 0) {
      ad_device_changed();
      return r;
    }
    if (r != EH_ERR_UNSUPPORTED)
      ad_fatal("track_through_csbend");
  }
  {
    double **coord = ad_host_begin();
    r = track_through_csbend(coord, n_part, csbend, p_error, Po, accepted, z_start, sigmaDelta2, rootname, maxamp,
                             apContour, apFileData, iSlice, eptr);
    ad_host_end(1);
  }
  return r;
}

long gpu_trackRfCavityWithWakes(long np, RFCA *rfca, double **accepted, double *P_central, double zEnd,
                                long iPass, RUN *run, CHARGE *charge, WAKE *wake, TRWAKE *trwake,
                                LSCKICK *LSCKick, long wakesAtEnd) {
  EH_RFCA r_;
  long r;
  if (fill_rfca(rfca, &r_) && charge == NULL && wake == NULL && trwake == NULL && LSCKick == NULL &&
      (accepted == NULL || accepted == G.accepted)) {
    double ph;
    /* the cavity's phase-reference group as elegant's table holds it
     * (src/phase_reference.c:28-110), mirrored into the library's for the call
     * and back after it */
    if (r_.phase_reference) {
      eh_delete_phase_references();
      if (get_phase_reference(&ph, rfca->phase_reference) == REF_PHASE_RETURNED)
        eh_set_phase_reference(r_.phase_reference, ph);
    }
    ad_to_device(np);
    r = eh_simple_rf_cavity(G.beam, np, &r_, P_central, zEnd);
    if (r >= 0) {
      rfca_write_back(rfca, &r_);
      if (r_.phase_reference && eh_get_phase_reference(r_.phase_reference, &ph) == EH_REF_PHASE_RETURNED)
        set_phase_reference(rfca->phase_reference, ph);
      ad_device_changed();
      return r;
    }
    /* options the library does not carry, and fiducial failures, which
     * elegant's routine reports itself */
    if (r != EH_ERR_UNSUPPORTED && r != EH_ERR_ARG)
      ad_fatal("trackRfCavityWithWakes");
  }
  {
    double **coord = ad_host_begin();
    r = trackRfCavityWithWakes(coord, np, rfca, accepted, P_central, zEnd, iPass, run, charge, wake, trwake, LSCKick,
                               wakesAtEnd);
    ad_host_end(1);
  }
  return r;
}

void gpu_exactDrift(long np, double length) {
  ad_to_device(np);
  ad_check(eh_exactDrift(G.beam, np, length), "exactDrift");
  ad_device_changed();
}

void gpu_track_particles(VMATRIX *M, long n_part) {
  EH_VMATRIX m;
  if (fill_vmatrix(M, &m)) {
    ad_to_device(n_part);
    ad_check(eh_track_particles(G.beam, n_part, &m), "track_particles");
    ad_device_changed();
    return;
  }
  {
    double **coord = ad_host_begin();
    track_particles(coord, M, coord, n_part);
    ad_host_end(1);
  }
}

long gpu_limit_amplitudes(double xmax, double ymax, long np, double **accepted, double z, double Po,
                          long extrapolate_z, long openCode) {
  long r;
  if (openCode == 0 && (accepted == NULL || accepted == G.accepted)) {
    ad_to_device(np);
    r = eh_limit_amplitudes(G.beam, xmax, ymax, np, z, Po, extrapolate_z, openCode);
    if (r >= 0) {
      ad_device_changed();
      return r;
    }
    if (r != EH_ERR_UNSUPPORTED)
      ad_fatal("limit_amplitudes");
  }
  {
    double **coord = ad_host_begin();
    r = limit_amplitudes(coord, xmax, ymax, np, accepted, z, Po, extrapolate_z, openCode, NULL);
    ad_host_end(1);
  }
  return r;
}

/* the elliptical sweep (src/limit_amplitudes.c:487-620); open sides and a
 * foreign accepted array on the host copy */
long gpu_elimit_amplitudes(double xmax, double ymax, long np, double **accepted, double z, double Po,
                           long extrapolate_z, long openCode, long exponent, long yexponent, ELEMENT_LIST *eptr) {
  long r;
  if (openCode == 0 && (accepted == NULL || accepted == G.accepted)) {
    ad_to_device(np);
    r = eh_elimit_amplitudes(G.beam, xmax, ymax, np, z, Po, extrapolate_z, openCode, exponent, yexponent);
    if (r >= 0) {
      ad_device_changed();
      return r;
    }
    if (r != EH_ERR_UNSUPPORTED && r != EH_ERR_ARG)
      ad_fatal("elimit_amplitudes");
  }
  /* elegant's own routine: the exponent errors are its to report */
  {
    double **coord = ad_host_begin();
    r = elimit_amplitudes(coord, xmax, ymax, np, accepted, z, Po, extrapolate_z, openCode, exponent, yexponent,
                          eptr);
    ad_host_end(1);
  }
  return r;
}

long gpu_removeInvalidParticles(long np, double **accepted, double z, double Po) {
  long r;
  if (accepted == NULL || accepted == G.accepted) {
    ad_to_device(np);
    r = eh_removeInvalidParticles(G.beam, np, z, Po);
    ad_check(r, "removeInvalidParticles");
    ad_device_changed();
    return r;
  }
  {
    double **coord = ad_host_begin();
    r = removeInvalidParticles(coord, np, accepted, z, Po);
    ad_host_end(1);
  }
  return r;
}

void gpu_offset_beam(long nToTrack, MALIGN *offset, double P_central) {
  EH_MALIGN m;
  if (fill_malign(offset, &m)) {
    ad_to_device(nToTrack);
    ad_check(eh_offset_beam(G.beam, nToTrack, &m, P_central), "offset_beam");
    ad_device_changed();
    return;
  }
  {
    double **coord = ad_host_begin();
    offset_beam(coord, nToTrack, offset, P_central);
    ad_host_end(1);
  }
}

void gpu_set_central_momentum(long np, double P_new, double *P_central) {
  ad_to_device(np);
  ad_check(eh_set_central_momentum(G.beam, np, P_new, P_central), "set_central_momentum");
  ad_device_changed();
}

/* ------------------------------------------------------------------------- */
/* routines run on the host copy */

double gpu_findFiducialTime(long np, double s0, double sOffset, double p0, unsigned long mode) {
  double t;
  double **coord = ad_host_begin();
  t = findFiducialTime(coord, np, s0, sOffset, p0, mode);
  ad_host_end(0);
  return t;
}

/* do_match_energy on the device copy (src/do_tracking.c:3398-3411: elegant
 * calls it for RFCA CHANGE_P0, ENERGY matching and ALWAYS_CHANGE_P0) */
void gpu_do_match_energy(long np, double *P_central, long change_beam) {
  ad_to_device(np);
  ad_check(eh_do_match_energy(G.beam, np, P_central, change_beam), "do_match_energy");
  ad_device_changed();
}

void gpu_compute_centroids(double *centroid, long n_part) {
  double **coord = ad_host_begin();
  compute_centroids(centroid, coord, n_part);
  ad_host_end(0);
}

/* the trajectory centroids of do_tracking's host branch (src/do_tracking.c:2601-2605) */
void gpu_collect_trajectory_data(double *centroid, long n_part) {
  long i, j;
  double sum;
  double **coord = ad_host_begin();
  for (i = 0; i < 6; i++) {
    for (j = 0, sum = 0; j < n_part; j++)
      sum += coord[j][i];
    centroid[i] = n_part ? sum / n_part : 0;
  }
  ad_host_end(0);
}

void gpu_center_beam(CENTER *center, long np, long iPass, double p0) {
  double **coord = ad_host_begin();
  center_beam(coord, center, np, iPass, p0);
  ad_host_end(1);
}

void *getGpuBeamSums(void *sums) { return sums; }

void gpu_accumulate_beam_sums(BEAM_SUMS *sums, long n_part, double p_central, double mp_charge, double *timeValue,
                              double tMin, double tMax, long startPID, long endPID, unsigned long flags) {
  double **coord = ad_host_begin();
  accumulate_beam_sums(sums, coord, n_part, p_central, mp_charge, timeValue, tMin, tMax, startPID, endPID, flags);
  ad_host_end(0);
}

long gpu_imposeApertureData(long np, double **accepted, double z, double Po, APERTURE_DATA *apData) {
  long r;
  double **coord = ad_host_begin();
  r = imposeApertureData(coord, np, accepted, z, Po, apData, NULL);
  ad_host_end(1);
  return r;
}

/* ------------------------------------------------------------------------- */
/* routines of elements element_on_gpu() never selects */

static void ad_unreachable(const char *name) {
  char buf[256];
  snprintf(buf, sizeof buf, "%s called for an element the gfx950 adapter keeps on the CPU", name);
  bombElegant(buf, NULL);
}

void gpu_track_through_wake(long np0, WAKE *wakeData, double *PoInput, RUN *run, long i_pass, CHARGE *charge) {
  ad_unreachable("gpu_track_through_wake");
}
void gpu_track_through_trwake(long np0, TRWAKE *wakeData, double Po, RUN *run, long i_pass, CHARGE *charge) {
  ad_unreachable("gpu_track_through_trwake");
}
long gpu_track_through_rfcw(long np, RFCW *rfcw, double **accepted, double *P_central, double zEnd, RUN *run,
                            long i_pass, CHARGE *charge) {
  ad_unreachable("gpu_track_through_rfcw");
  return 0;
}
long gpu_track_through_matter(long np, long iPass, MATTER *matter, double Po, double **accepted, double z0) {
  ad_unreachable("gpu_track_through_matter");
  return 0;
}
void gpu_track_through_lscdrift(long np, LSCDRIFT *LSC, double Po, CHARGE *charge) {
  ad_unreachable("gpu_track_through_lscdrift");
}
long gpu_track_through_driftCSR(long np, CSRDRIFT *csrDrift, double Po, double **accepted, double zStart,
                                double revolutionLength, CHARGE *charge, char *rootname) {
  ad_unreachable("gpu_track_through_driftCSR");
  return 0;
}
long gpu_track_through_csbendCSR(long n_part, CSRCSBEND *csbend, double p_error, double Po, double **accepted,
                                 double z_start, double z_end, CHARGE *charge, char *rootname, MAXAMP *maxamp,
                                 APCONTOUR *apContour, APERTURE_DATA *apFileData, ELEMENT_LIST *eptr) {
  ad_unreachable("gpu_track_through_csbendCSR");
  return 0;
}
long gpu_rectangular_collimator(RCOL *rcol, long np, double **accepted, double z, double Po) {
  ad_unreachable("gpu_rectangular_collimator");
  return 0;
}
long gpu_elliptical_collimator(ECOL *ecol, long np, double **accepted, double z, double Po) {
  ad_unreachable("gpu_elliptical_collimator");
  return 0;
}
long gpu_beam_scraper(SCRAPER *scraper, long np, double **accepted, double z, double Po) {
  ad_unreachable("gpu_beam_scraper");
  return 0;
}
void gpu_matr_element_tracking(VMATRIX *M, MATR *matr, long np, double z) {
  ad_unreachable("gpu_matr_element_tracking");
}
void gpu_ematrix_element_tracking(VMATRIX *M, EMATRIX *matr, long np, double z, double *P_central) {
  ad_unreachable("gpu_ematrix_element_tracking");
}
void gpu_addCorrectorRadiationKick(long np, ELEMENT_LIST *elem, long type, double Po, double *sigmaDelta2,
                                   long disableISR) {
  ad_unreachable("gpu_addCorrectorRadiationKick");
}
