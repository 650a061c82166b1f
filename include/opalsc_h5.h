/* H5Part phase-space files of OPAL-T (separate library libopalsc_h5.so,
 * linked against the HDF5 C library, so that libopalsc.so needs no HDF5).
 *
 * Layout written and read (H5PartWrapperForPT, src/Structure/H5PartWrapperForPT.cpp,
 * through H5hut's H5Part layer):
 *  - file attributes (writeHeader, :193-274): OPAL_version and the unit
 *    strings, "dump frequency" (int64), "dPhiGlobal" (float64);
 *  - one group "Step#k" per dump, k = 0, 1, ... in order of writing, with the
 *    step attributes of writeStepHeader (:276-377) and one 1D dataset per
 *    particle quantity of writeStepData (:379-424): x y z px py pz q (float64),
 *    id (int64), ptype (int32);
 *  - strings as scalar fixed-length C strings (length + 1 bytes), numbers as
 *    1D arrays of native doubles / int64.
 * Reading follows readStepHeader / readStepData (:100-190): the attributes
 * TIME, SPOS, LocalTrackStep, GlobalTrackStep, RefPartR, RefPartP and
 * TaitBryantAngles, particles [first, last] of every dataset.
 *
 * Every function returns 0 or a negative code; opalsc_h5_last_error() gives
 * the message of the last failure on the calling thread. */
#ifndef OPALSC_H5_H
#define OPALSC_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The step attributes of H5PartWrapperForPT::writeStepHeader, in its units:
 * positions m, momenta beta*gamma, time s, energies MeV, mass GeV, charge C,
 * B-ref T, E-ref MV/m. */
typedef struct {
    double spos;              /* SPOS */
    double time;              /* TIME */
    double ref_r[3];          /* RefPartR */
    double ref_p[3];          /* RefPartP */
    double centroid[3];       /* centroid */
    double rms_x[3];          /* RMSX */
    double mean_p[3];         /* MEANP */
    double rms_p[3];          /* RMSP */
    double tait_bryant[3];    /* TaitBryantAngles (rad) */
    double eps_norm[3];       /* #varepsilon */
    double eps_geom[3];       /* #varepsilon-geom */
    double min_x[3], max_x[3];/* minX, maxX */
    double min_p[3], max_p[3];/* minP, maxP */
    int64_t step;             /* Step: index of the group (set by the writer) */
    int64_t local_track_step; /* LocalTrackStep */
    int64_t global_track_step;/* GlobalTrackStep */
    double sigma;             /* #sigma */
    double energy;            /* ENERGY: mean kinetic energy */
    double dE;                /* dE: energy spread */
    double mass;              /* MASS */
    double charge;            /* CHARGE: sum of the macro charges */
    int64_t num_bunch;        /* NumBunch */
    int64_t step_to_last_inj; /* SteptoLastInj */
    double B_ref[3];          /* B-ref */
    double E_ref[3];          /* E-ref */
} opalsc_h5_step;

const char *opalsc_h5_last_error(void);

/* H5PartWrapperForPT::writeHeader (H5PartWrapperForPT.cpp:193-274): creates
 * (truncates) path with the file attributes; opal_version may be NULL. */
int opalsc_h5_create(const char *path, const char *opal_version, int64_t dump_frequency, double dphi_global);

/* H5PartWrapperForPT::writeStep (H5PartWrapperForPT.cpp:267-274,276-424):
 * appends group Step#k, k the number of steps already in the file (written
 * to *step_out and to the Step attribute), with attrs and the N particles
 * (AoS R, P; Q; id; ptype may be NULL: zeros).  N = 0 writes nothing, as the
 * reference does. */
int opalsc_h5_write_step(const char *path, int64_t N, const double *R, const double *P, const double *Q,
                         const int64_t *id, const int32_t *ptype, const opalsc_h5_step *attrs,
                         int64_t *step_out);

/* H5GetNumSteps: the number of Step# groups. */
int opalsc_h5_num_steps(const char *path, int64_t *nsteps);

/* Particles in step (-1: the last step). */
int opalsc_h5_num_particles(const char *path, int64_t step, int64_t *N);

/* Integer file attribute "dump frequency" and float "dPhiGlobal" (readHeader, :37-89). */
int opalsc_h5_read_header(const char *path, int64_t *dump_frequency, double *dphi_global);

/* readStep (H5PartWrapperForPT.cpp:91-190) of step (-1: the last, as
 * readStep reads): the step attributes into attrs (every attribute above
 * that the file holds) and particles first..last (inclusive) into R, P (AoS),
 * Q and id (each may be NULL).  OPALSC_H5_ERANGE if last is beyond the step's
 * particles or first > last, as the reference throws. */
int opalsc_h5_read_step(const char *path, int64_t step, int64_t first, int64_t last, double *R, double *P,
                        double *Q, int64_t *id, opalsc_h5_step *attrs);

enum { OPALSC_H5_OK = 0, OPALSC_H5_EINVAL = -1, OPALSC_H5_EIO = -2, OPALSC_H5_ERANGE = -3 };

#ifdef __cplusplus
}
#endif

#endif /* OPALSC_H5_H */
