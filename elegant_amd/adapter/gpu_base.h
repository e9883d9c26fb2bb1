/* gpu_base.h -- the GPU interface elegant's sources call when built with
 * HAVE_GPU, implemented by gpu_adapter.c on the gfx950 tracking library.
 *
 * elegant includes this header (and the per-routine headers beside it, which
 * only include this one) after track.h, e.g. src/do_tracking.c:22-26,
 * src/multipole.c:19, src/csbend.c:20-22.  Every declaration below is taken from
 * its call sites in elegant: the residency protocol of do_tracking
 * (gpuBaseInit :413, setElementGpuData/getGpuBase :925-932, forceParticlesToCpu,
 * gpuBaseDealloc :3178) and one gpu_<routine> per element routine with the CPU
 * routine's arguments minus the particle array (e.g. src/multipole.c:698-709).
 *
 * Elements and options the library carries run on the GPU; for everything
 * else getElementOnGpu() is 0 and elegant runs its CPU routine on the host
 * copy of the particles, which the adapter keeps current (see gpu_adapter.c).
 * GPU_VERIFY builds are not supported.
 */
#ifndef ELEGANT_GPU_BASE_H
#define ELEGANT_GPU_BASE_H

#ifdef GPU_VERIFY
#error "the gfx950 adapter does not implement GPU_VERIFY (compareGpuCpu)"
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* the part of the GPU state do_tracking reads (src/do_tracking.c:928-931) */
typedef struct {
  short elementOnGpu; /* the current element runs on the GPU */
  double **coord;     /* host particle array given to gpuBaseInit */
  long nOriginal;
  long isMaster;
} GPUBASE;

/* residency */
void gpuBaseInit(double **coord, long nOriginal, double **accepted, double **lostParticles, long isMaster);
void gpuBaseDealloc(void);
GPUBASE *getGpuBase(void);
void setElementGpuData(void *eptr);
long getElementOnGpu(void);
double **forceParticlesToCpu(char *routine);
void sortByPID(long np);
void startGpuTimer(void);
void startCpuTimer(void);
void displayTimings(void);

/* element routines on the GPU (src/multipole.c:701, src/csbend.c:692, :3799,
 * src/simple_rfca.c:275, src/limit_amplitudes.c:207, :298, :510,
 * src/matrix.c:233, src/do_tracking.c:3323, :3576) */
long gpu_multipole_tracking2(long n_part, ELEMENT_LIST *elem, double p_error, double Po, double **accepted,
                             double z_start, MAXAMP *maxamp, APCONTOUR *apcontour, APERTURE_DATA *apFileData,
                             double *sigmaDelta2, long iSlice);
long gpu_track_through_csbend(long n_part, CSBEND *csbend, double p_error, double Po, double **accepted,
                              double z_start, double *sigmaDelta2, char *rootname, MAXAMP *maxamp,
                              APCONTOUR *apContour, APERTURE_DATA *apFileData, long iSlice, ELEMENT_LIST *eptr);
long gpu_trackRfCavityWithWakes(long np, RFCA *rfca, double **accepted, double *P_central, double zEnd,
                                long iPass, RUN *run, CHARGE *charge, WAKE *wake, TRWAKE *trwake,
                                LSCKICK *LSCKick, long wakesAtEnd);
void gpu_exactDrift(long np, double length);
void gpu_track_particles(VMATRIX *M, long n_part);
long gpu_limit_amplitudes(double xmax, double ymax, long np, double **accepted, double z, double Po,
                          long extrapolate_z, long openCode);
long gpu_elimit_amplitudes(double xmax, double ymax, long np, double **accepted, double z, double Po,
                           long extrapolate_z, long openCode, long exponent, long yexponent, ELEMENT_LIST *eptr);
long gpu_removeInvalidParticles(long np, double **accepted, double z, double Po);
void gpu_offset_beam(long nToTrack, MALIGN *offset, double P_central);
void gpu_set_central_momentum(long np, double P_new, double *P_central);

/* routines run on the host copy of the particles (synchronised on entry) */
double gpu_findFiducialTime(long np, double s0, double sOffset, double p0, unsigned long mode);
void gpu_do_match_energy(long np, double *P_central, long change_beam);
void gpu_compute_centroids(double *centroid, long n_part);
void gpu_collect_trajectory_data(double *centroid, long n_part);
void gpu_center_beam(CENTER *center, long np, long iPass, double p0);
void *getGpuBeamSums(void *sums);
void gpu_accumulate_beam_sums(BEAM_SUMS *sums, long n_part, double p_central, double mp_charge, double *timeValue,
                              double tMin, double tMax, long startPID, long endPID, unsigned long flags);
long gpu_imposeApertureData(long np, double **accepted, double z, double Po, APERTURE_DATA *apData);

/* routines of elements setElementGpuData never puts on the GPU: defined so that
 * elegant links; reaching one is an error */
void gpu_track_through_wake(long np0, WAKE *wakeData, double *PoInput, RUN *run, long i_pass, CHARGE *charge);
void gpu_track_through_trwake(long np0, TRWAKE *wakeData, double Po, RUN *run, long i_pass, CHARGE *charge);
long gpu_track_through_rfcw(long np, RFCW *rfcw, double **accepted, double *P_central, double zEnd, RUN *run,
                            long i_pass, CHARGE *charge);
long gpu_track_through_matter(long np, long iPass, MATTER *matter, double Po, double **accepted, double z0);
void gpu_track_through_lscdrift(long np, LSCDRIFT *LSC, double Po, CHARGE *charge);
long gpu_track_through_driftCSR(long np, CSRDRIFT *csrDrift, double Po, double **accepted, double zStart,
                                double revolutionLength, CHARGE *charge, char *rootname);
long gpu_track_through_csbendCSR(long n_part, CSRCSBEND *csbend, double p_error, double Po, double **accepted,
                                 double z_start, double z_end, CHARGE *charge, char *rootname, MAXAMP *maxamp,
                                 APCONTOUR *apContour, APERTURE_DATA *apFileData, ELEMENT_LIST *eptr);
long gpu_rectangular_collimator(RCOL *rcol, long np, double **accepted, double z, double Po);
long gpu_elliptical_collimator(ECOL *ecol, long np, double **accepted, double z, double Po);
long gpu_beam_scraper(SCRAPER *scraper, long np, double **accepted, double z, double Po);
void gpu_matr_element_tracking(VMATRIX *M, MATR *matr, long np, double z);
void gpu_ematrix_element_tracking(VMATRIX *M, EMATRIX *matr, long np, double z, double *P_central);
void gpu_addCorrectorRadiationKick(long np, ELEMENT_LIST *elem, long type, double Po, double *sigmaDelta2,
                                   long disableISR);

#ifdef __cplusplus
}
#endif

#endif
