/* elegant_hip.h -- C ABI of the MI355X (gfx950) element-tracking library.
 *
 * The library replaces the per-particle element tracking that elegant's
 * do_tracking() loop performs on the CPU for DRIF, EDRIFT, KQUAD, KSEXT, KOCT,
 * CSBEND and RFCA, plus the MAXAMP aperture and NaN sweeps that follow every
 * element.  Particles live on the GPU as structure-of-arrays between calls
 * (the residency model of elegant's HAVE_GPU hooks: gpuBaseInit at
 * src/do_tracking.c:411-414, forceParticlesToCpu, gpuBaseDealloc at
 * src/do_tracking.c:3172-3179).  Every element entry point keeps the argument
 * meaning and order of the CPU routine it replaces, minus the particle array,
 * exactly like the gpu_* hooks elegant calls (e.g. gpu_multipole_tracking2 at
 * src/multipole.c:698-709).
 *
 * Element parameter structs below carry the PARAMETER-table names of
 * src/track_data.c (KQUAD :1227, KSEXT :1125, KOCT :1169, CSBEND :1898,
 * RFCA :579, MAXAMP :994, DRIF/EDRIFT) for the subset of fields the path
 * honours; a value outside that subset makes the call return EH_ERR_UNSUPPORTED
 * so that the caller can stay on its CPU routine.
 *
 * Coordinates follow src/track.h:81-98: rows of EH_NPROP doubles,
 * (x, x', y, y', s, delta, particleID, lossPass, bunch, weight).
 * All arithmetic is IEEE-754 double without contraction.
 *
 * Error convention: functions returning long return a particle count >= 0 or a
 * negative EH_ERR_* code; eh_last_error() gives a message.
 */
#ifndef ELEGANT_HIP_H
#define ELEGANT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EH_NPROP 10          /* totalPropertiesPerParticle, src/track_data.c:62 */
#define EH_ID_INDEX 6        /* particleIDIndex, src/track.h:88 */
#define EH_LOSS_PASS_INDEX 7 /* lossPassIndex */

enum eh_error {
  EH_OK = 0,
  EH_ERR_UNSUPPORTED = -1, /* parameter value outside the supported subset */
  EH_ERR_ARG = -2,         /* inconsistent argument (e.g. n_part != resident count) */
  EH_ERR_HIP = -3,         /* HIP runtime failure */
  EH_ERR_NOMEM = -4,
  EH_ERR_STATE = -5 /* call sequence error (e.g. no beam resident) */
};

/* element type codes (subset of elegant's T_* codes, by meaning only) */
enum eh_type {
  EH_T_MARK = 0,
  EH_T_DRIF = 1,
  EH_T_EDRIFT = 2,
  EH_T_KQUAD = 3,
  EH_T_KSEXT = 4,
  EH_T_KOCT = 5,
  EH_T_CSBEND = 6,
  EH_T_RFCA = 7,
  EH_T_MAXAMP = 8,
  EH_T_KQUSE = 9,
  EH_T_MULT = 10,
  EH_T_ROTATE = 11,
  EH_T_MALIGN = 12,
  EH_T_ENERGY = 13
};

/* DRIF: src/track.h:1364-1368 (length, order; order 0 means run_setup default_order=2) */
typedef struct {
  double length;
  int32_t order;
  int32_t pad_;
} EH_DRIFT;

/* EDRIFT: exact drift, src/csbend.c:3792 */
typedef struct {
  double length;
} EH_EDRIFT;

/* KQUAD: src/track.h:2204-2228, defaults src/track_data.c:1227 */
typedef struct {
  double length, k1, tilt, dx, dy, dz, fse;
  double xkick, ykick, xKickCalibration, yKickCalibration; /* HKICK, VKICK, H/VCALIBRATION */
  int32_t n_kicks, nSlices;                               /* N_KICKS (deprecated alias path), N_SLICES */
  int32_t integration_order;                              /* 2, 4 or 6 */
  int32_t expandHamiltonian;
  int32_t synch_rad; /* SYNCH_RAD: classical radiation (rad_coef, src/multipole.c:755-763) */
  /* BORE and B (pole-tip field, T): with BORE != 0 the strength comes from them,
   * n! B / BORE^n e / (m c P_central), instead of K<n> (src/multipole.c:739-804) */
  double bore, B;
} EH_KQUAD;

/* KSEXT: src/track.h:2149-2166, defaults src/track_data.c:1125 */
typedef struct {
  double length, k2, k1, j1, tilt, dx, dy, dz, fse;
  double xkick, ykick, xKickCalibration, yKickCalibration;
  int32_t n_kicks, nSlices;
  int32_t integration_order;
  int32_t expandHamiltonian;
  int32_t synch_rad; /* SYNCH_RAD: classical radiation (rad_coef, src/multipole.c:755-763) */
  /* BORE and B (pole-tip field, T): with BORE != 0 the strength comes from them,
   * n! B / BORE^n e / (m c P_central), instead of K<n> (src/multipole.c:739-804) */
  double bore, B;
} EH_KSEXT;

/* KOCT: src/track.h:2171-2182, defaults src/track_data.c:1169 (N_KICKS=4) */
typedef struct {
  double length, k3, tilt, dx, dy, dz, fse;
  int32_t n_kicks, nSlices;
  int32_t integration_order;
  int32_t expandHamiltonian;
  int32_t synch_rad; /* SYNCH_RAD: classical radiation (rad_coef, src/multipole.c:755-763) */
  /* BORE and B (pole-tip field, T): with BORE != 0 the strength comes from them,
   * n! B / BORE^n e / (m c P_central), instead of K<n> (src/multipole.c:739-804) */
  double bore, B;
} EH_KOCT;

/* KQUSE (quadrupole with a sextupole as secondary multipole): src/track.h:3498-3503,
 * parameters src/track_data.c:3270-3288, tracking src/multipole.c:923-949 */
typedef struct {
  double length, k1, k2, tilt, dx, dy, dz, fse1, fse2;
  int32_t n_kicks, nSlices;
  int32_t integration_order;
  int32_t expandHamiltonian;
  int32_t synch_rad; /* SYNCH_RAD: classical radiation (rad_coef, src/multipole.c:755-763) */
} EH_KQUSE;

/* MULT (thin multipole kicks between drifts): src/track.h:1407-1412, parameters
 * src/track_data.c:486-502, tracking src/multipole.c:439-658, SYNCH_RAD
 * included.  BORE/BTIPL and ORDER > 8 are not supported (EH_ERR_UNSUPPORTED). */
typedef struct {
  double length, KnL, tilt, bore, BTipL, dx, dy, dz, factor;
  int32_t order, synch_rad, expandHamiltonian;
  int32_t nSlices;
} EH_MULT;

/* ROTATE (rotation of the coordinates about s): parameters
 * src/track_data.c:1005-1010, tracking src/do_tracking.c:1836-1852 through
 * rotation_matrix (src/tilt_matrices.c:81-117) and the order-1 matrix loop
 * (src/matrix.c:424-460).  on_pass < 0: every pass. */
typedef struct {
  double tilt;
  int32_t on_pass, exclude_floor, exclude_optics;
} EH_ROTATE;

/* MALIGN (beam offset): parameters src/track_data.c:929-944, tracking
 * offset_beam (src/do_tracking.c:3311-3366).  Only the default particle range
 * (START_PID = END_PID = -1) is supported.  on_pass < 0: every pass. */
typedef struct {
  double dxp, dyp, dx, dy, dz, dt, dp, de;
  int32_t on_pass, startPID, endPID;
} EH_MALIGN;

/* ENERGY (reference momentum change): src/track.h:1998-2004, parameters
 * src/track_data.c:985-990, tracking src/do_tracking.c:1861-1881 through
 * set_central_momentum (:3563-3596) and, for MATCH_BEAMLINE (the reference
 * momentum moved to the beam's mean; no fiducial-beam flags) and
 * MATCH_PARTICLES (the particles' momenta moved so that their mean is the
 * reference momentum), do_match_energy (:3376-3549).  Both matching flags at
 * once is EH_ERR_ARG, as the reference refuses it. */
typedef struct {
  double central_energy, central_momentum;
  int32_t match_beamline, match_particles;
} EH_ENERGY;

/* CSBEND: src/track.h:2462-2502, defaults src/track_data.c:1898.
 * N_KICKS is an alias of N_SLICES (src/track_data.c:1943-1944). */
typedef struct {
  double length, angle;
  double k1, k2, k3, k4, k5, k6, k7, k8;
  double e1, e2, tilt, h1, h2, hgap, fint; /* FINT (both edges) */
  double dx, dy, dz, xKick, yKick;
  double fse, fseDipole, fseQuadrupole, etilt;
  double b1, b2, b3, b4, b5, b6, b7, b8; /* used when use_bn */
  int32_t nSlices;
  int32_t nonlinear;          /* default 1 */
  int32_t edge1_effects;      /* default 1; 0, 1, 2 (K. Hwang's fringe map, HGAP/FINT/H), 3 (the
                                 symplectic kick on the canonical momenta) or 4 (Hwang's map in
                                 R. Lindberg's symplectic form) */
  int32_t edge2_effects;
  int32_t edge_order;         /* default 1; >= 2: second-order edge map (apply_edge_effects) */
  int32_t integration_order;  /* 2, 4, 6 */
  int32_t expandHamiltonian;  /* 0 supported */
  int32_t use_bn;
  int32_t expansionOrder;     /* 0 = auto */
  int32_t etiltSign;          /* default 1 */
  int32_t malignMethod;       /* 0 supported */
  int32_t edgeFlags;          /* EH_BEND_EDGE1/2 as determined for the beamline */
  int32_t synch_rad;          /* SYNCH_RAD: classical radiation (src/csbend.c:956-959) */
  int32_t kick_limit_scaling; /* KICK_LIMIT_SCALING: the limits below times rho0/rho_actual */
  double edge1_kick_limit, edge2_kick_limit; /* EDGE*_KICK_LIMIT on the edges' x' kick (<= 0: none) */
} EH_CSBEND;

#define EH_BEND_EDGE1_EFFECTS 1
#define EH_BEND_EDGE2_EFFECTS 2

/* fiducial modes, src/simple_rfca.c:23-48 (FIDUCIAL "tmean", "light", "first",
 * "pmaximum"; as in serial elegant, FIRST always fails and PMAXIMUM fails when
 * the first particle has the largest momentum, src/simple_rfca.c:101-146) */
enum eh_fid_mode { EH_FID_TMEAN = 0, EH_FID_LIGHT = 1, EH_FID_FIRST = 2, EH_FID_PMAX = 3 };

/* RFCA: src/track.h:1494-1508, defaults src/track_data.c:579 */
typedef struct {
  double length, volt, phase, freq, Q;
  double dx, dy;
  double tReference; /* default -1 */
  int32_t nKicks;    /* 0 = matrix method */
  /* CHANGE_P0: the reference momentum matched to the surviving beam after the
   * cavity (src/simple_rfca.c:977-982); CHANGE_T: whole RF periods taken off the
   * time coordinates, judged on the particle in row 0 of the reference's
   * arrangement (:488-498, :586-588, :864-865).  Not in batched tracking or the
   * adjoint (EH_ERR_UNSUPPORTED). */
  int32_t change_p0, change_t, end1Focus, end2Focus, standingWave;
  /* LINEARIZE, LOCK_PHASE: on the beam's mean arrival time at each visit
   * (src/simple_rfca.c:500-556); N_KICKS must be 1 (EH_ERR_ARG otherwise).  Not
   * in batched tracking or the adjoint. */
  int32_t linearize, lockPhase;
  int32_t fiducial_mode; /* eh_fid_mode */
  /* PHASE_REFERENCE: cavities with the same non-zero number share the phase
   * fiducial of the first of them tracked (src/simple_rfca.c:412-466); 0: the
   * cavity's own (the library does not write a number back, as the reference
   * does with unused_phase_reference) */
  int64_t phase_reference;
  /* state mutated by tracking, like the reference element (src/simple_rfca.c:460-475) */
  int32_t fiducial_seen, pad_;
  double phase_fiducial, t_fiducial;
} EH_RFCA;

/* MAXAMP: src/track_data.c:994 (no open side).  ELLIPTICAL selects the
 * (super-)elliptical boundary |x/x_max|^EXPONENT + |y/y_max|^YEXPONENT = 1
 * (src/limit_amplitudes.c:487-620, src/multipole.c:1784-1792); y_exponent 0
 * means EXPONENT.  Exponents must be even and >= 2 (EH_ERR_ARG otherwise). */
typedef struct {
  double x_max, y_max;
  int32_t elliptical, exponent;
  int32_t y_exponent, pad_;
} EH_MAXAMP;

/* a first- or second-order transport map as track_particles applies it
 * (VMATRIX, src/track.h:205-221): final_i = C[i] + sum_j (R[i][j] + sum_{k<=j}
 * T[i][j][k] x_k) x_j, summed j = 5 down to 0 and k = j down to 0
 * (src/matrix.c:373-385; order 1: C[i] + sum_j R[i][j] x_j, src/matrix.c:438-445).
 * T is packed by rows: T[i][j*(j+1)/2 + k] holds elegant's T[i][j][k], k <= j. */
typedef struct {
  int32_t order; /* 1 or 2 */
  int32_t pad_;
  double C[6];
  double R[6][6];
  double T[6][21];
} EH_VMATRIX;

/* one beamline entry; mirrors ELEMENT_LIST's (name, type, p_elem) */
typedef struct {
  int32_t type; /* eh_type */
  int32_t occurence;
  const char *name;
  void *p_elem; /* EH_DRIFT*, EH_KQUAD*, ... according to type */
} EH_ELEMENT_LIST;

/* particle charge/mass constants used on the path (electrons, CODATA 2018) */
#define EH_C_MKS 2.99792458e8
#define EH_ME_MEV 0.51099895000
#define EH_E_MKS 1.602176634e-19      /* particleCharge */
#define EH_ME_MKS 9.1093837015e-31    /* particleMass */
#define EH_EPSILON_O 8.8541878128e-12 /* epsilon_o */

const char *eh_last_error(void);

/* ---------------- device / residency --------------------------------------- */
typedef struct eh_beam eh_beam;

/* select the HIP device; replaces gpuBaseInit's device setup */
int eh_init(int device);

/* allocate resident storage for up to n_capacity particles */
eh_beam *eh_beam_create(long n_capacity);
void eh_beam_free(eh_beam *beam);

/* copy rows [0, n_part) of an AoS array (EH_NPROP doubles per row) to the GPU;
 * the row order becomes the reference order for later loss replay */
long eh_beam_upload(eh_beam *beam, double **coord, long n_part);

/* copy the beam back in the exact row order the CPU routines would have
 * produced: survivors in [0, n_alive), lost particles above, with the loss
 * pass in column EH_LOSS_PASS_INDEX.  Returns n_alive. */
long eh_beam_download(eh_beam *beam, double **coord, long n_part);

/* the same for a contiguous array of n_part rows of EH_NPROP doubles (the layout
 * czarray_2d gives the reference's particle arrays, e.g. src/elegant.c:3183) */
long eh_beam_upload_flat(eh_beam *beam, const double *flat, long n_part);
long eh_beam_download_flat(eh_beam *beam, double *flat, long n_part);

/* the row arrangement a download produces: upload_row[r] = the row (of the
 * last upload) whose particle a download writes to row r.  The reference's
 * element routines move the rows of a caller's `accepted` array in lockstep
 * with the particle rows (swapParticles, src/multipole.c:1045-1046,
 * src/limit_amplitudes.c:270-271); a caller permutes its copy with this.
 * n_part = eh_beam_n_total. */
long eh_beam_row_order(eh_beam *beam, int32_t *upload_row, long n_part);

/* survivors, and all particles held (survivors + lost), i.e. the n_part of a download */
long eh_beam_n_alive(const eh_beam *beam);
long eh_beam_n_total(const eh_beam *beam);

/* pass number that subsequent losses are recorded with (do_tracking's i_pass) */
void eh_beam_set_pass(eh_beam *beam, long pass);

/* ---------------- per-element entry points (the plain drop-in path) -------- */
/* Each returns the number of surviving particles, like the CPU routine. */

/* replaces multipole_tracking, src/multipole.c:439-446 (MULT) */
long eh_multipole_tracking(eh_beam *beam, long n_part, const EH_MULT *mult, double p_error, double Po,
                           double z_start);

/* replaces offset_beam, src/do_tracking.c:3311-3315 (MALIGN, for the pass given
 * by eh_beam_set_pass) */
long eh_offset_beam(eh_beam *beam, long n_part, const EH_MALIGN *offset, double P_central);

/* replaces set_central_momentum, src/do_tracking.c:3563-3596: every resident
 * particle's delta is re-referenced to P_new and *P_central becomes P_new */
long eh_set_central_momentum(eh_beam *beam, long np, double P_new, double *P_central);

/* replaces do_match_energy, src/do_tracking.c:3376-3549, called by elegant
 * through gpu_do_match_energy (:3398-3411): the mean momentum of the resident
 * particles as a Kahan sum in the reference's row order; change_beam 0 moves
 * *P_central to it (when they differ by more than 1e-14 relative) and
 * re-expresses every delta, change_beam 1 moves every particle's momentum so
 * that the mean is *P_central, keeping each time of flight (a non-finite time
 * coordinate is EH_ERR_STATE, where the reference aborts) */
long eh_do_match_energy(eh_beam *beam, long np, double *P_central, long change_beam);

/* replaces track_particles with rotation_matrix(tilt) for ROTATE,
 * src/do_tracking.c:1836-1852 (for the pass given by eh_beam_set_pass) */
long eh_rotate_beam(eh_beam *beam, long n_part, const EH_ROTATE *rotate);

/* replaces multipole_tracking2, src/multipole.c:660-677 (KQUAD/KSEXT/KOCT/KQUSE) */
long eh_multipole_tracking2(eh_beam *beam, long n_part, const EH_ELEMENT_LIST *elem, double p_error, double Po,
                            double z_start, const EH_MAXAMP *maxamp, long iSlice);

/* replaces track_through_csbend, src/csbend.c:654-663 */
long eh_track_through_csbend(eh_beam *beam, long n_part, const EH_CSBEND *csbend, double p_error, double Po,
                             double z_start, const EH_MAXAMP *maxamp, long iSlice);

/* replaces simple_rf_cavity, src/simple_rfca.c:264-293 (mutates rfca state;
 * with CHANGE_P0 also *P_central) */
long eh_simple_rf_cavity(eh_beam *beam, long np, EH_RFCA *rfca, double *P_central, double zEnd);

/* replaces exactDrift, src/csbend.c:3792 */
long eh_exactDrift(eh_beam *beam, long np, double length);

/* replaces drift_beam / track_particles with drift_matrix, src/do_tracking.c:3754,
 * src/matrix.c:212, src/compute_matrices.c:391 */
long eh_drift_beam(eh_beam *beam, long np, double length, long order);

/* replaces track_particles, src/matrix.c:212-470, for first- and second-order
 * maps (order 3 returns EH_ERR_UNSUPPORTED): every matrix-tracked element
 * (src/do_tracking.c:1311), DRIF through drift_matrix, ROTATE through
 * rotation_matrix (src/do_tracking.c:1836-1852).  No particle is lost. */
long eh_track_particles(eh_beam *beam, long n_part, const EH_VMATRIX *M);

/* replaces limit_amplitudes, src/limit_amplitudes.c:197-199 (openCode must be 0) */
long eh_limit_amplitudes(eh_beam *beam, double xmax, double ymax, long np, double z, double Po, long extrapolate_z,
                         long openCode);

/* replaces elimit_amplitudes, src/limit_amplitudes.c:498-500 (openCode must be
 * 0): the (super-)elliptical sweep |x/xmax|^exponent + |y/ymax|^yexponent > 1
 * (yexponent 0 means exponent; both even and >= 2, else EH_ERR_ARG), a lost
 * particle's dp becoming sqrt((Po (1 + dp))^2 + 1) */
long eh_elimit_amplitudes(eh_beam *beam, double xmax, double ymax, long np, double z, double Po, long extrapolate_z,
                          long openCode, long exponent, long yexponent);

/* phase references (src/phase_reference.c:28-143): the library's table of
 * PHASE_REFERENCE groups, which eh_simple_rf_cavity and eh_do_tracking read and
 * set.  eh_get_phase_reference replaces get_phase_reference (:28) and returns
 * EH_REF_PHASE_RETURNED with *phase, EH_REF_PHASE_NOT_SET or
 * EH_REF_PHASE_NONEXISTENT (0 for group 0); eh_set_phase_reference replaces
 * set_phase_reference (:73); eh_delete_phase_references replaces
 * delete_phase_references (:115), which elegant calls between steps. */
enum { EH_REF_PHASE_RETURNED = 1, EH_REF_PHASE_NOT_SET = 2, EH_REF_PHASE_NONEXISTENT = 3 };
long eh_get_phase_reference(int64_t ref, double *phase);
long eh_set_phase_reference(int64_t ref, double phase);
long eh_delete_phase_references(void);

/* replaces removeInvalidParticles, src/limit_amplitudes.c:289-291 */
long eh_removeInvalidParticles(eh_beam *beam, long np, double z, double Po);

/* ---------------- whole tracking loop -------------------------------------- */
enum eh_track_mode {
  EH_MODE_PLAIN = 0, /* one kernel launch per element call, like the per-element hooks */
  EH_MODE_FUSED = 1  /* runs of elements and turns fused into persistent launches */
};

typedef struct {
  double particle_element_turns; /* surviving particles x counted elements, summed over passes */
  double kernel_ms;              /* GPU time inside all tracking kernels (incl. sweeps, compaction) */
  long launches;
  long n_alive;
  /* the element-tracking kernel alone (k_fused, k_plain element calls, k_batch spans):
   * stream time, launches and the particle-element units those launches processed */
  double main_ms;
  long main_launches;
  double main_units;
} eh_track_stats;

/* do_tracking over a beamline for n_passes (src/do_tracking.c:144, pass loop :498,
 * element loop :829, post-element sweeps :2506-2524, :2675-2684).  The beam must be
 * resident (eh_beam_upload).  Returns the number of surviving particles. */
long eh_do_tracking(eh_beam *beam, EH_ELEMENT_LIST *beamline, long n_elements, long n_passes, double *P_central,
                    int mode, int turns_per_launch, eh_track_stats *stats);

/* No counterpart in the reference: eh_do_tracking in EH_MODE_FUSED for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The element descriptors,
 * CSBEND field terms and launch counters live in a static scratch array of
 * about 9 MiB per GPU instead of per-beam allocations; the caller still creates
 * the beam (its particles are the inputs and outputs).  Limits: at most 4096
 * elements and 421 CSBENDs (the field terms of each at the worst case); larger
 * lines return EH_ERR_ARG before anything runs.  eh_init must have run on the device.  Calls
 * that overlap on one GPU share the scratch and race. */
long eh_do_tracking_noalloc(eh_beam *beam, EH_ELEMENT_LIST *beamline, long n_elements, long n_passes,
                            double *P_central, int turns_per_launch, eh_track_stats *stats);

/* host utility: CSBEND field-expansion table (computeCSBENDFieldCoefficients,
 * src/csbend.c:175-652) for b[9], c[9] (index 0 = dipole term) and curvature h.
 * Writes 11x11 row-major Fx, Fy; returns expansionOrder+1 or a negative code. */
long eh_csbend_field_table(const double *b, const double *c, double h, int nonlinear, int expansionOrder, double *Fx,
                           double *Fy, int *hasNormal, int *hasSkew);

/* ---------------- batched lattice variants --------------------------------- */
/* Many variants of one beamline (same element types in the same order, e.g.
 * error seeds or strength scans), each with its own particles, tracked together:
 * a persistent kernel pulls one-variant wavefronts from a device work list, and
 * the spans of a run are captured in one HIP graph.  Each variant's result equals
 * eh_do_tracking of that variant alone (no reference counterpart: the reference
 * runs variants one after another, src/elegant.c:1177, src/vary.c:255). */
typedef struct eh_batch eh_batch;

typedef struct {
  int32_t block;         /* 64, 128 or 256 threads per block */
  int32_t span_elements; /* elements per fused launch (whole beamline = one turn) */
  int32_t waves_per_cu;  /* persistent grid size = CUs x waves_per_cu wavefronts */
  int32_t use_graph;     /* capture the spans of a run in one graph */
} eh_batch_tune;

eh_batch *eh_batch_create(long n_variants, const long *n_particles);
void eh_batch_free(eh_batch *batch);
/* variant v's particles: n rows of EH_NPROP doubles (n = n_particles[v]) */
long eh_batch_upload(eh_batch *batch, long v, const double *flat, long n);
long eh_batch_download(eh_batch *batch, long v, double *flat, long n);
long eh_batch_n_alive(const eh_batch *batch, long v);
/* beamlines[v] is variant v's beamline of n_elements entries */
long eh_batch_track(eh_batch *batch, EH_ELEMENT_LIST *const *beamlines, long n_elements, long n_passes,
                    double *P_central, const eh_batch_tune *tune, eh_track_stats *stats);

/* ---------------- reverse-mode adjoint --------------------------------------- */
/* No reference counterpart: gradients of multi-turn tracking with respect to
 * element strengths and offsets, for gradient-based tuning in place of the
 * derivative-free loops of optimize.c (src/optimize.c:912, :2002). */
typedef struct eh_adjoint eh_adjoint;

typedef struct {
  double forward_ms;      /* GPU time of the recorded forward pass */
  double backward_ms;     /* GPU time of the reverse sweeps */
  double retrack_ms;      /* GPU time re-tracking segments between checkpoints */
  double forward_units;   /* particle-element units of the forward pass */
  double max_recon_err;   /* largest |recovered - stored| state (x, xp, y, yp, dp) at an anchor */
  double max_recon_err_s; /* same for s */
  double anchor_bytes;    /* device memory for anchor states */
  double checkpoint_bytes;
  long n_survivors;
} eh_adjoint_stats;

eh_adjoint *eh_adjoint_create(void);
void eh_adjoint_free(eh_adjoint *adj);
/* Track n particles (rows of 10 doubles, elegant layout) n_turns through the
 * beamline, exactly as eh_do_tracking does for the survivors, keeping
 * checkpoints every checkpoint_every turns (1: keep every anchor, no
 * re-tracking).  param_slots[4*e + k] is the gradient slot of element e's
 * strength (k=0: K1/K2/K3/K1 of KQUAD/KSEXT/KOCT/KQUSE, K1 of CSBEND), DX (k=1)
 * or DY (k=2) of the kick multipoles, or FSE (k=3) of CSBEND, or -1.
 * RF cavity fiducials must already be fixed (track one pass first).  Writes the
 * final coordinates (n x 6) and survival flags; returns the number of survivors. */
long eh_adjoint_forward(eh_adjoint *adj, const double *coord, long n, EH_ELEMENT_LIST *beamline, long n_elements,
                        const int32_t *param_slots, long n_params, long n_turns, double P_central,
                        long checkpoint_every, double *final_coord, int32_t *alive);
/* The same for nvar variants of one beamline structure (e.g. error seeds):
 * coord holds the variants' particles one after another (n_per_variant[v] rows
 * each), beamlines[v] variant v's elements, param_slots the slots shared by all
 * variants, so the gradient is summed over every variant's particles (the
 * gradient of a mean over seeds, up to the caller's weights).  All variants run
 * in one launch per checkpoint segment, each wavefront on one variant.  Final
 * coordinates and alive flags come back in the order of coord; backward takes
 * and returns per-particle arrays in that order too.  No reference
 * counterpart (a batched form of the entry point above). */
long eh_adjoint_forward_batch(eh_adjoint *adj, const double *coord, const long *n_per_variant, long nvar,
                              EH_ELEMENT_LIST *const *beamlines, long n_elements, const int32_t *param_slots,
                              long n_params, long n_turns, double P_central, long checkpoint_every,
                              double *final_coord, int32_t *alive);
/* Given dL/d(final coordinates) (n x 6; rows of lost particles are ignored),
 * writes dL/d(parameters) and, if grad_initial is not NULL, dL/d(initial
 * coordinates) (n x 6, zero for lost particles).  Returns the survivor count. */
long eh_adjoint_backward(eh_adjoint *adj, const double *grad_final, double *grad_params, double *grad_initial,
                         eh_adjoint_stats *stats);

/* ---------------- diagnostics ---------------------------------------------- */
/* reference row order after the given loss events (no reference counterpart:
 * it reproduces the row permutation of the reference's swap-to-top compaction,
 * e.g. src/limit_amplitudes.c:262-283, for testing) */
long eh_replay_order(long n, const int32_t *lost_event, const int32_t *lost_id, long n_lost, int32_t *slot_at_row);

/* the sequence in which the reference's element routines process the particles
 * they lose, for the same loss events (no reference counterpart, for testing):
 * order[k] = the k-th lost identity processed; pred[id] = -2 if it was the
 * first particle of its routine call, -1 if a surviving particle was processed
 * just before it, else the lost identity processed just before it.  It settles
 * the loss position a kick multipole gives a particle lost at its entry
 * (src/multipole.c:1160-1232). */
long eh_replay_sequence(long n, const int32_t *lost_event, const int32_t *lost_id, long n_lost, int32_t *pred,
                        int32_t *order);

/* GPU evaluation of sin, cos, asin, log, sqrt, division and tan for checking
 * bit-identity with the host (no reference counterpart) */
long eh_selftest_math(const double *x, const double *y, long n, double *out);

/* n doubles copied device-to-device by a plain kernel, reps times: a launch of
 * known HBM traffic for calibrating memory counters (no reference counterpart) */
long eh_diag_copy(long n, int reps);

#ifdef __cplusplus
}
#endif

#endif
