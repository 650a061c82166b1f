/*
 * opalsc.h -- C ABI of the MI355X (gfx950) space-charge PIC step.
 *
 * A drop-in for the OPAL-T space-charge path: cloud-in-cell (CIC) charge
 * scatter, Hockney open-boundary FFT Poisson solve with the integrated
 * Green's function (IGF) on the doubled grid, CIC field gather and the Boris
 * push of ParallelTTracker.  Reference paths below are relative to the OPAL
 * 2.0 source tree.
 *
 * Conventions (shared by every entry point):
 *  - Every function returns 0 on success or a negative OPALSC_E* code; the
 *    message is available from opalsc_last_error(ctx).  No C++ exception
 *    crosses this boundary.
 *  - Particle vectors are AoS double[N][3] (x,y,z), scalars double[N], as
 *    OPAL's ParticleAttrib<Vektor<double,3>> stores them.  Grids are
 *    x-fastest double[n2][n1][n0] over the interior cells (no guard layer).
 *  - Pointers are host pointers unless the name ends in _dev.  The caller
 *    owns every array it passes; the context owns all device memory.
 *  - One context per host thread.  Work is ordered on the context's HIP
 *    stream; every host-pointer entry point returns after its results are in
 *    host memory.
 *  - All physics is in OPAL's units: R in m, P = beta*gamma, Q in C, dt in s,
 *    species charge in units of e, species mass as rest energy in eV.
 */
#ifndef OPALSC_H
#define OPALSC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OPALSC_ABI_VERSION 2

enum {
    OPALSC_OK = 0,
    OPALSC_EINVAL = -1,  /* bad argument */
    OPALSC_EHIP = -2,    /* HIP runtime error */
    OPALSC_ENOMEM = -3,  /* device allocation failed */
    OPALSC_EMESH = -4,   /* "h<0, can not build a mesh" (PartBunchBase.hpp:870-872) */
    OPALSC_ESTATE = -5,  /* call out of order (e.g. track before bunch) */
    OPALSC_ECAP = -6     /* more particles than the context was created for */
};

enum { OPALSC_FP64 = 0, OPALSC_MIXED = 1 };
enum { OPALSC_GREENS_INTEGRATED = 0, OPALSC_GREENS_STANDARD = 1 };
enum { OPALSC_BC_OPEN = 0, OPALSC_BC_PERIODIC = 1 };
enum { OPALSC_GRID_RHO = 0, OPALSC_GRID_PHI = 1, OPALSC_GRID_EFIELD = 2 };

typedef struct opalsc_ctx opalsc_ctx;

/* FIELDSOLVER attributes (src/Structure/FieldSolver.cpp:123-200) plus sizing. */
typedef struct {
    int n[3];              /* MX, MY, MT: cells per dimension (each >= 4) */
    int greens;            /* GREENSF: OPALSC_GREENS_INTEGRATED or _STANDARD */
    int precision;         /* OPALSC_FP64, or OPALSC_MIXED (FP32 FFT/convolution/gather) */
    double bboxincr;       /* BBOXINCR in percent (OPAL default 2) */
    int64_t max_particles; /* capacity of the device particle buffers */
    int device;            /* HIP device ordinal */
    int sort_interval;     /* tracker: re-sort the bunch by cell every k steps (0: never) */
    int bcz;               /* BCFFTZ (FieldSolver.cpp:153,428-448): OPALSC_BC_OPEN, or
                              OPALSC_BC_PERIODIC (DC beams: the solver is periodic in z,
                              FFTPoissonSolver.cpp:60,130-155, the field's z faces
                              periodic and particles wrapped in z, PartBunch.cpp:1079-1098;
                              no energy bins; n[2] = 2^a, 3*2^a or 5*2^a with a >= 2) */
} opalsc_config;

/* DataSink statistics (src/Structure/DataSink.cpp:360-391) as computed by
 * PartBunchBase::calcBeamParameters (PartBunchBase.hpp:1529-1613). */
typedef struct {
    double rmean[3], pmean[3], rrms[3], prms[3], eps_norm[3], eps[3], rprms[3];
    double gamma_mean, ekin_MeV;
    double t, pathLength; /* tracker time and path length after the step */
    int64_t n;
    int64_t dropped; /* CIC deposits that fell outside the grid during the step */
} opalsc_stats;

/* Tracker set-up: ParallelTTracker state (src/Algorithms/ParallelTTracker.cpp:214-426)
 * and the analytic stand-ins for element fields used by the benchmark
 * configurations: B = (0,0,Bz), E = (0,0,Ez0*sin(2*pi*rf_freq*t + rf_phase)),
 * both expressed in the reference frame. */
typedef struct {
    double dT;
    double charge;   /* species charge / e */
    double mass_eV;  /* species rest energy */
    double coupling; /* 1/(4 pi eps0) (src/Track/TrackRun.cpp:542-543) */
    double Bz, Ez0, rf_freq, rf_phase;
    double t, pathLength;
    double RefPartR[3], RefPartP[3];
    int space_charge; /* 0 behaves as FSTYPE=NONE */
} opalsc_track_params;

/* Synthetic bunch generated on the device from a counter-based RNG:
 * x,y,z ~ N(0, sigma_r), p - (0,0,p0) ~ N(0, sigma_p) (absolute beta*gamma),
 * a fraction halo_frac replaced by points uniform in the ball of radius
 * halo_radius*sigma (positions only), uniform macro-charge q_total/N. */
typedef struct {
    int64_t n;
    uint64_t seed;
    double sigma_r[3], sigma_p[3];
    double p0;
    double q_total;
    double halo_frac, halo_radius;
} opalsc_gauss_params;

/* --- context ----------------------------------------------------------- */
int opalsc_abi_version(void);
int opalsc_create(opalsc_ctx **out, const opalsc_config *cfg);
int opalsc_destroy(opalsc_ctx *ctx);
/* Last error message of ctx (or of the most recent failed create when ctx is NULL). */
const char *opalsc_last_error(const opalsc_ctx *ctx);
int opalsc_synchronize(opalsc_ctx *ctx);

/* --- Poisson solver ------------------------------------------------------
 * FFTPoissonSolver::computePotential(Field_t& rho, Vector_t hr),
 * src/Solvers/FFTPoissonSolver.cpp:337-382: in place, rho in, phi out, exact
 * circular convolution of the zero-padded rho with the (mirrored) Green's
 * function on the doubled grid, without the h^3 and 1/(4 pi eps0) factors. */
int opalsc_compute_potential(opalsc_ctx *ctx, double *rho_inout, const double hr[3]);
/* The same on a device array (rho_dev: device memory, n0*n1*n2 doubles); returns
 * when phi is in rho_dev.  For callers whose grids already live in HBM. */
int opalsc_compute_potential_dev(opalsc_ctx *ctx, double *rho_dev, const double hr[3]);
/* With bcz = OPALSC_BC_PERIODIC the same call is the periodic-z solve: circular
 * convolution on the (2 n0, 2 n1, n2) grid with the Green's function mirrored in
 * x and y only and read one-sided in z (mirrorRhoField, FFTPoissonSolver.cpp:654-664),
 * normalised by 1/(2 n0 2 n1 n2). */

/* FFTPoissonSolver::computePotential(Field_t& rho, Vector_t hr, double zshift),
 * src/Solvers/FFTPoissonSolver.cpp:289-330: the image-charge potential.  In
 * place: rho in; out: minus the circular convolution with the shifted
 * integrated Green's function (shiftedIntGreensFunction,
 * FFTPoissonSolver.cpp:546-639: not mirrored in z), read back flipped in z
 * (phi[i][j][k] = conv[i][j][n2-1-k]).  The integrated Green's function is
 * used whatever the context's greens setting. */
int opalsc_compute_potential_shifted(opalsc_ctx *ctx, double *rho_inout, const double hr[3],
                                     double zshift);

/* --- mesh and cloud-in-cell ----------------------------------------------
 * PartBunchBase::boundp (src/Classic/Algorithms/PartBunchBase.hpp:802-883):
 * bounds of R, BBOXINCR padding, hr = (max-min)/(n-1), origin = min - hr/2.
 * In a periodic-z context (DC beam) z is not padded: its extent is the period
 * length, hr[2] = period/(n2-1) (PartBunchBase.hpp:814,840-843). */
int opalsc_boundp(opalsc_ctx *ctx, int64_t N, const double *R, double hr[3], double origin[3]);

/* PartBunchBase::setBeamFrequency (src/Classic/Algorithms/PartBunchBase.hpp:2534-2537):
 * the DC beam's period length becomes c/f (default c / 1 GHz).  Applies to
 * opalsc_boundp and to the tracker (also after opalsc_track_init). */
int opalsc_set_beam_frequency(opalsc_ctx *ctx, double f_hz);

/* The particle boundary condition of a DC beam, which boundp's update()
 * applies after the mesh is set (ippl/src/Particle/ParticleSpatialLayout.hpp:296-297,
 * ParticlePeriodicBCond, ippl/src/Particle/ParticleBConds.h:38-45): R[i][2]
 * (AoS, in place) wrapped into [origin[2], n2*hr[2] + origin[2]); x and y
 * unchanged.  The self-field entry points do not wrap: a caller follows
 * opalsc_boundp with this call as the reference's boundp does.  (The tracker
 * applies it itself.) */
int opalsc_wrap_periodic_z(opalsc_ctx *ctx, int64_t N, double *R, const double origin[3], const double hr[3]);

/* IntCICImpl<3> cell lookup (ippl/src/Particle/IntCIC.h:816-851;
 * ippl/src/Meshes/UniformCartesian.hpp:1025-1039,1084-1097): per particle the
 * lower CIC grid point lgp[3] and normalised offsets dpos[3]. */
int opalsc_cic_index(opalsc_ctx *ctx, int64_t N, const double *R, const double origin[3],
                     const double hr[3], int32_t *lgp, double *dpos);

/* Stable sort of particles by the cell of their lgp: perm[i] = original index
 * of the i-th particle in cell order.  Cell order is the 3D Morton (Z-order)
 * index of the lgp cell (bits of i, j, k interleaved, i lowest); particles
 * outside the grid go last.  No counterpart in the reference (its scatter is a
 * serial loop); it is the order the device scatter and gather run in. */
int opalsc_sort_by_cell(opalsc_ctx *ctx, int64_t N, const double *R, const double origin[3],
                        const double hr[3], uint32_t *perm);

/* --- self fields ---------------------------------------------------------
 * PartBunch::computeSelfFields(), src/Classic/Algorithms/PartBunch.cpp:482-648:
 * scatter Q*dt with CIC, rho /= dT, Lorentz-scaled mesh, Poisson solve,
 * E = -Grad(phi) scaled by (gz, gz, 1/gz), CIC gather into Ef (overwritten),
 * Bf += (-betaC*Ey, betaC*Ex, 0) (accumulated).  Q is updated in place to
 * (Q*dt)/dt exactly as the reference leaves it.  origin/hr are the mesh
 * set by boundp.  In a periodic-z context the solve is periodic in z and the
 * gather reads the periodic images of E beyond the z faces (PeriodicFace
 * guard layers); deposits beyond the faces are dropped as on an open mesh
 * (BareField::accumGuardCells has no neighbour on one rank). */
int opalsc_compute_self_fields(opalsc_ctx *ctx, int64_t N, const double *R, const double *P,
                               double *Q, const double *dt, double dT, const double origin[3],
                               const double hr[3], double coupling, double *Ef, double *Bf);
/* The same on device arrays (same layouts; every _dev pointer must be device
 * memory, checked); no host copies.  Returns when Ef, Bf and Q are written. */
int opalsc_compute_self_fields_dev(opalsc_ctx *ctx, int64_t N, const double *R_dev, const double *P_dev,
                                   double *Q_dev, const double *dt_dev, double dT, const double origin[3],
                                   const double hr[3], double coupling, double *Ef_dev, double *Bf_dev);

/* Energy-binned self fields: the bin loop of
 * ParallelTTracker::computeSpaceChargeFields (src/Algorithms/ParallelTTracker.cpp:567-579)
 * over PartBunch::computeSelfFields(int) (src/Classic/Algorithms/PartBunch.cpp:138-430).
 * Particle i belongs to bin[i] (0 <= bin < nbins <= 64; other values belong
 * to no bin).  Bin b is scattered with charge qi*dt per particle
 * (setBinCharge, PartBunchBase.hpp:716-724), rho /= dT, and solved on the
 * mesh (origin, hr) with its own Lorentz factor gamma_bin[b] (NULL: the mean
 * of sqrt(1+P.P) over the bin, calcGammas, PartBunchBase.hpp:639-682; an
 * empty bin has no field).  If image != 0 its image charge in the cathode
 * plane z = 0 (shifted integrated Green's function, PartBunch.cpp:252-303) is
 * solved too.  Ef (overwritten) = sum over bins of E_bin + E_image;
 * Bf (accumulated) += sum over bins of betaC_b * (-Ey, Ex, 0) of the bin and
 * betaC_b * (Ey, -Ex, 0) of its image. */
int opalsc_compute_self_fields_binned(opalsc_ctx *ctx, int64_t N, const double *R, const double *P,
                                      const int32_t *bin, int nbins, double qi, const double *dt,
                                      double dT, const double origin[3], const double hr[3],
                                      double coupling, const double *gamma_bin, int image,
                                      double *Ef, double *Bf);

/* Grids of the most recent compute_self_fields / track step: OPALSC_GRID_RHO
 * (charge density entering the solver, n0*n1*n2), OPALSC_GRID_PHI (potential
 * in V, n0*n1*n2), OPALSC_GRID_EFIELD (AoS [n0*n1*n2][3] after the Lorentz
 * scaling; after a tracker step with energy bins, the sum over the bins and
 * their images that the gather reads).  Mixed precision grids are widened to
 * double. */
int opalsc_get_grid(opalsc_ctx *ctx, int which, double *out);

/* --- pusher --------------------------------------------------------------
 * ParallelTTracker::pushParticles (src/Algorithms/ParallelTTracker.h:448-455)
 * = switchToUnitlessPositions + BorisPusher::push + switchOffUnitlessPositions. */
int opalsc_push_half(opalsc_ctx *ctx, int64_t N, double *R, const double *P, const double *dt);

/* ParallelTTracker::kickParticles (src/Algorithms/ParallelTTracker.h:442-446)
 * = BorisPusher::kick (src/Steppers/BorisPusher.h:56-113) per particle. */
int opalsc_kick(opalsc_ctx *ctx, int64_t N, double *P, const double *Ef, const double *Bf,
                const double *dt, double charge, double mass_eV);

/* PartBunchBase::calcBeamParameters (src/Classic/Algorithms/PartBunchBase.hpp:1529-1613). */
int opalsc_beam_parameters(opalsc_ctx *ctx, int64_t N, const double *R, const double *P,
                           double mass_eV, opalsc_stats *out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): opalsc_beam_parameters on device arrays R_dev, P_dev (AoS
 * double[N][3]) into out_dev (device memory, one opalsc_stats; t, pathLength
 * and dropped are 0), enqueued on `stream` (a hipStream_t, NULL: the null
 * stream) of the current device, without synchronising and without
 * allocating.  The reduction's partial sums live in a 4 MiB file-scope device
 * array (at most 8192 rows: every N up to 2^32-16 fits); each GPU holds its
 * copy until the process ends, and calls that overlap on one GPU race on it.
 * Errors are read with opalsc_last_error(NULL). */
int opalsc_beam_parameters_noalloc(int64_t N, const double *R_dev, const double *P_dev, double mass_eV,
                                   opalsc_stats *out_dev, void *stream);

/* --- device-resident tracker --------------------------------------------
 * The ParallelTTracker::execute step loop (ParallelTTracker.cpp:343-388) with
 * the bunch kept in HBM: push, beam-frame rotation, boundp, self fields,
 * external fields, kick, push, reference-frame update, statistics. */
int opalsc_bunch_set(opalsc_ctx *ctx, int64_t N, const double *R, const double *P, const double *Q);
int opalsc_bunch_generate(opalsc_ctx *ctx, const opalsc_gauss_params *g);
/* Download R, P, Q in the order the bunch was set (any may be NULL). */
int opalsc_bunch_get(opalsc_ctx *ctx, double *R, double *P, double *Q);
/* Energy bins of the resident bunch: the Bin particle attribute with
 * setEnergyBins and the macro charge qi_m (PartBunchBase.hpp:450-456,716-724).
 * bin[i] belongs to the i-th particle as given to bunch_set (values outside
 * [0, nbins) belong to no bin and carry no charge).  With bins the tracker's
 * space-charge step is ParallelTTracker::computeSpaceChargeFields' bin loop
 * (ParallelTTracker.cpp:567-579): calcGammas, then per bin the charge qi*dT
 * of its particles solved in the bin's Lorentz frame and, if image != 0, its
 * image charge (PartBunch::computeSelfFields(int), PartBunch.cpp:138-430).
 * nbins = 0 removes the bins.  Bins stay until the bunch is set again. */
int opalsc_bunch_set_bins(opalsc_ctx *ctx, const int32_t *bin, int nbins, double qi, int image);
/* Number of bins, particles per bin (counts[nbins]) and the Lorentz factors of
 * the most recent tracker step (gammas[nbins]); counts/gammas may be NULL. */
int opalsc_bunch_get_bins(opalsc_ctx *ctx, int *nbins, int64_t *counts, double *gammas);
int opalsc_track_init(opalsc_ctx *ctx, const opalsc_track_params *p);
/* Run nsteps steps; per_step (may be NULL) receives nsteps statistics. */
int opalsc_track(opalsc_ctx *ctx, int nsteps, opalsc_stats *per_step);
/* The same entry point under the name of the fused device-resident step. */
int opalsc_step(opalsc_ctx *ctx, int nsteps, opalsc_stats *per_step);

/* --- many independent bunches --------------------------------------------
 * No counterpart in the reference, which tracks one bunch per run
 * (ParallelTTracker::execute): a batch steps several independent bunches,
 * each a context that is already set up (bunch_set/bunch_generate,
 * optional set_bins, track_init), together.  The contexts must stay alive
 * while the batch exists and may differ in grid size and precision. */
/* --- external field maps -------------------------------------------------------
 * Elements on a straight beamline along z whose fields the tracker adds in
 * every step (ParallelTTracker::computeExternalFields, ParallelTTracker.cpp:595-640,
 * and updateReferenceParticle for the reference particle, 1134-1155): a
 * Solenoid with a 2D magnetostatic map (FM2DMagnetoStatic; scale = KS) or an
 * RFCavity with a 2D dynamic map (FM2DDynamic; scale = VOLT, freq_hz = FREQ,
 * phase = LAG; the map's frequency when freq_hz is 0 or more than 1 % away,
 * RFCavity.cpp:290-305).  The map's first z point sits at elemedge + zbegin in
 * the lab frame, which is the reference frame moved by (0, 0, pathLength).
 * Files are read as the reference reads them (Fieldmap.cpp:519-556,
 * FM2DMagnetoStatic.cpp:11-140, FM2DDynamic.cpp:13-150).  At most 8 maps;
 * opalsc_fieldmap_add returns the element's index.  A particle beyond a map's
 * radius gets no field from it or from later maps (the reference removes it). */
int opalsc_fieldmap_add(opalsc_ctx *ctx, const char *path, double elemedge, double scale, double freq_hz,
                        double phase);
int opalsc_fieldmap_clear(opalsc_ctx *ctx);
/* The maps' field at lab positions R (AoS) and time t into E, B (overwritten);
 * lost (may be NULL) counts particles beyond a map's radius. */
int opalsc_fieldmap_eval(opalsc_ctx *ctx, int64_t N, const double *R, double t, double *E, double *B,
                         int64_t *lost);

/* --- file formats -----------------------------------------------------------
 * FROMFILE distributions (Distribution::getNumberOfParticlesInFile and
 * createDistributionFromFile, src/Distribution/Distribution.cpp:1073-1100,1103-1250):
 * an ASCII file whose first line that is not empty once '#' comments are
 * removed starts with the particle count, followed by x px y py z pz per
 * particle, whitespace separated.  momenta_in_ev != 0 is INPUTMOUNITS=EV: each
 * momentum component converted with converteVToBetaGamma (Distribution.cpp:1047-1054)
 * for the rest energy mass_eV.  As in the reference, values the stream cannot
 * read leave zeros.  opalsc_fromfile_read fills R and P (AoS) for N particles,
 * N the count in the file.  No context: errors are read with
 * opalsc_last_error(NULL). */
int opalsc_fromfile_count(const char *path, int64_t *N);
int opalsc_fromfile_read(const char *path, int64_t N, int momenta_in_ev, double mass_eV, double *R, double *P);

/* One line of the DataSink statistics file (src/Structure/DataSink.cpp:281-430),
 * in its column order.  energy is the mean kinetic energy, max_r get_bounds'
 * maximum, corr the x-px correlations (rprms), Dx..DDy the dispersion moments
 * (calcBeamParameters, PartBunchBase.hpp:1529-1613), B_ref/E_ref (T, MV/m) the
 * external field at the reference particle, R0/P0 the first particle set. */
typedef struct {
    double t_ns, s;
    int64_t num_particles;
    double charge, energy_MeV;
    double rms_r[3], rms_p[3], emit[3], mean_r[3], ref_r[3], ref_p[3], max_r[3], corr[3];
    double Dx, DDx, Dy, DDy;
    double B_ref[3], E_ref[3];
    double dE_MeV, dt_ns;
    int64_t parts_outside; /* 0: no halo boundary (Options::beamHaloBoundary = 0) */
    int has_particle0;     /* the R0/P0 columns are written */
    double R0[3], P0[3];
} opalsc_stat_row;

/* The statistics line of the resident bunch after opalsc_track, as
 * ParallelTTracker::writePhaseSpace passes it to DataSink (ParallelTTracker.cpp:1018-1071):
 * external field at the reference particle at t - dT/2.  The reference
 * particle is the tracker's (its frame; OPAL writes the lab frame, which the
 * analytic element fields here do not distinguish). */
int opalsc_stat_row_get(opalsc_ctx *ctx, opalsc_stat_row *row);

/* DataSink::writeSDDSHeader + doWriteStatData (DataSink.cpp:281-430,585-925):
 * new_file != 0 creates the file and writes the SDDS header (input_fn in its
 * description), then the row is appended: each value followed by a tab padded
 * to width 10, doubles in scientific notation with 15 digits. */
int opalsc_stat_write(const char *path, const char *input_fn, int new_file, const opalsc_stat_row *row);

typedef struct opalsc_batch opalsc_batch;
int opalsc_batch_create(opalsc_batch **out, opalsc_ctx *const *ctxs, int nctx);
int opalsc_batch_destroy(opalsc_batch *batch);
const char *opalsc_batch_last_error(const opalsc_batch *batch);
/* nsteps steps of every bunch (the ParallelTTracker step loop,
 * ParallelTTracker.cpp:338-388, of many independent bunches at once).  The
 * fused form computes what opalsc_track(ctx, nsteps, ...) computes for every
 * bunch (the next step's first half push fused into the gather pass, a re-sort
 * by cell every sort_interval steps of the context); the per-bunch form takes
 * nsteps single steps.  Either equals single-step calls bit for bit.
 * Fused form (default, when no bunch has energy bins, field maps or a DC beam):
 * every particle pass is one launch over all bunches of a kernel class
 * (precision, equal charges, external field), every Poisson pass one launch
 * over all grids of one shape; use_graph != 0 captures one step as a HIP graph
 * (grid shapes on parallel branches) and replays it per step, steps with a
 * re-sort are launched directly.  Otherwise (or after
 * opalsc_batch_set_fused(batch, 0)) each bunch's kernels form a branch of the
 * graph (use_graph != 0) or run on the bunch's stream; the per-bunch form does
 * not re-sort.  The batch re-plans or re-captures when a context's bunch, bins
 * or tracker set-up changed.  per_step (may be NULL) receives [nsteps][nctx]
 * statistics. */
int opalsc_step_batched(opalsc_batch *batch, int nsteps, opalsc_stats *per_step, int use_graph);
/* enable != 0 (default): the fused form where every bunch allows it; 0: the
 * per-bunch form.  No counterpart in the reference (a batching choice). */
int opalsc_batch_set_fused(opalsc_batch *batch, int enable);
/* Autotuning (no counterpart in the reference: tile shapes of this library's
 * kernels; process-wide, set before creating the contexts or batches it should
 * apply to, not thread-safe).  src_amd/autotune.py times the candidates on the
 * device and writes the table the library is built with.
 * FFT tiles: E elements per thread for the transforms of length L of one
 * precision (a workgroup holds W = 512 E / L lines).  _candidates writes the E
 * values the kernels take (at most cap) and returns their count;
 * _set overrides the table (E = 0 removes the override; OPALSC_EINVAL for an E
 * the kernels cannot take); _get returns the E a grid created now would use. */
int opalsc_tune_fft_candidates(int L, int precision, int *E, int cap);
int opalsc_tune_fft_set(int L, int precision, int E);
int opalsc_tune_fft_get(int L, int precision);
/* Deposit: a bunch of at most max_particles particles deposits through the
 * larger LDS box (5632 cells, three blocks per CU), a larger one through the
 * smaller box (4864 cells, four blocks per CU). */
int opalsc_tune_deposit_set(int64_t max_particles);
int64_t opalsc_tune_deposit_get(void);
/* Tracker re-sort: order by the Hilbert-curve key of the cell without its
 * drop_bits lowest bits (0: by cell; 3: by blocks of 2x2x2 cells, 6: 4x4x4, ...),
 * particles keeping their relative order within a block.  Does not apply to
 * opalsc_sort_by_cell (Morton order). */
int opalsc_tune_sort_set(int drop_bits);
int opalsc_tune_sort_get(void);
int opalsc_track_get_params(opalsc_ctx *ctx, opalsc_track_params *out);
/* Mesh (hr, origin) of the most recent track step or boundp call. */
int opalsc_get_mesh(opalsc_ctx *ctx, double hr[3], double origin[3]);

/* --- correctly rounded FP64 atan/log used by the IGF --------------------
 * The host evaluation runs the same double-double code as the device. */
int opalsc_host_atan_cr(int64_t n, const double *x, double *y);
int opalsc_host_log_cr(int64_t n, const double *x, double *y);
int opalsc_device_atan_cr(opalsc_ctx *ctx, int64_t n, const double *x, double *y);
int opalsc_device_log_cr(opalsc_ctx *ctx, int64_t n, const double *x, double *y);
/* The IGF corner term F(v0, v1, v2) of FFTPoissonSolver::integratedGreensFunction
 * (src/Solvers/FFTPoissonSolver.cpp:479-489) at n points v[3i..3i+2], cell volume 1:
 * fast = 0 with the correctly rounded functions (FP64 solve), fast = 1 with the
 * reduced-instruction form of the mixed-precision solve.  Host and device
 * evaluations of either form return the same bits. */
int opalsc_host_igf_term(int64_t n, const double *v, int fast, double *y);
int opalsc_device_igf_term(opalsc_ctx *ctx, int64_t n, const double *v, int fast, double *y);
/* The real octant G[k][j][i], (n0+1)(n1+1)(n2+1) doubles x-fastest, of
 * FFTPoissonSolver::integratedGreensFunction (table 0, with G[0,0,0] := G[0,0,1],
 * src/Solvers/FFTPoissonSolver.cpp:459-522,657) or of the image charge's
 * shiftedIntGreensFunction tables (1: z corners + zshift, 2: also - n2*h2;
 * src/Solvers/FFTPoissonSolver.cpp:546-639) for spacing h.  Host: fast = 0 the
 * FP64 solve's corner table with correctly rounded atan/log, fast = 1 the
 * mixed solve's one-pass form (the 8-corner difference within 16 cells of the
 * origin, the order-4 cell average of 1/r beyond).  Device: the context's
 * own kernels (its precision picks the form) on its grid size. */
int opalsc_host_green_octant(const int n[3], const double h[3], double zshift, int table, int fast, double *out);
int opalsc_device_green_octant(opalsc_ctx *ctx, const double h[3], double zshift, int table, double *out);

/* --- timing --------------------------------------------------------------
 * Per-kernel HIP-event timing of the tracker, recorded on the context's
 * stream when enabled: names[i] and total milliseconds ms[i] over the steps
 * since the last reset, launches[i] the launch count.  Returns the number of
 * kernel classes (<= cap). */
int opalsc_profile_enable(opalsc_ctx *ctx, int enable);
int opalsc_profile_read(opalsc_ctx *ctx, int cap, const char **names, double *ms, int64_t *launches);

/* Achievable HBM bandwidth on this device: a plain 16-byte-vector copy kernel
 * (k_copy_calibration) over two buffers of `bytes` each, `reps` times; *gbs
 * = 2*bytes per launch / average launch time.  No counterpart in the
 * reference: it calibrates roofline fractions and hardware byte counters. */
int opalsc_copy_bandwidth(opalsc_ctx *ctx, size_t bytes, int reps, double *gbs);

#ifdef __cplusplus
}
#endif
#endif /* OPALSC_H */
