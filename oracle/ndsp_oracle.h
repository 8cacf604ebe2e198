/*
 * CPU oracle for the ndspmhd hot path: a plain-C, FP64, serial restatement of
 * the reference's derivs call graph (src/derivs.f90:27) and leapfrog step
 * (src/stepND_leapfrog_mhd.f90:31).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker
 * (or, for the baseline, as the thing timed on the host CPU).  The product
 * path (ndspmhd_amd, libndsp_hip.so) never calls it.
 *
 * Option and state layout: iopt[32] / dopt[32] as in include/ndsp.h;
 * per-particle arrays in the reference's column-major order
 * (x(ndim,N) -> x[i*ndim+d], vel(3,N) -> vel[i*3+d]).
 */
#ifndef NDSP_ORACLE_H
#define NDSP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

/* returns NULL and a message in err (>=256 bytes) on an unsupported option */
orc_sim *orc_new(const int32_t *iopt, const double *dopt, int32_t npart,
                 char *err);
void orc_free(orc_sim *s);

/* copy in the real-particle state (arrays in reference layout) */
int orc_set_state(orc_sim *s, const double *x, const double *vel,
                  const double *pmass, const double *hh, const double *en,
                  const double *Bevol, const double *alpha, const double *psi,
                  const double *dustevol, const int32_t *itype);

/* one derivative evaluation: boundary, ghosts, link list, h-rho iteration,
 * conservative->primitive, rates (src/derivs.f90:27-287).  0 = ok. */
int orc_derivs(orc_sim *s);

/* one leapfrog step with the current dt, then the new dt
 * (src/stepND_leapfrog_mhd.f90:31-275).  0 = ok. */
int orc_step(orc_sim *s);

/* the predictor half of a step alone (stepND_leapfrog_mhd.f90:70-165): the
 * state advanced by dt with the rates of the last evaluation, no evaluation
 * and no corrector.  0 = ok. */
int orc_predict(orc_sim *s);

/* copy a per-particle field of the real particles; width 1, 3 or ndim.
 * Names: x vel pmass hh en Bevol alpha psi dustevol rho gradh gradgradh
 * drhodt dhdt dens pr spsound uu Bfield force dudt dendt dBevoldt daldt dpsidt
 * divB curlB gradpsi dustfrac rhogas rhodust ddustevoldt deltav graddivv
 * numneigh(its as double) nits.  Returns the width, or -1 for an unknown name. */
int orc_get(const orc_sim *s, const char *name, double *out);

/* overwrite a per-particle field of the real particles (the state fields,
 * rho, drhodt and the rates a step starts from: vel hh en Bevol alpha psi
 * dustevol rho drhodt dhdt force dendt dBevoldt daldt dpsidt ddustevoldt,
 * and x).  Returns the width, or -1 for a name it does not take. */
int orc_set(orc_sim *s, const char *name, const double *in);

/* scalars: dt dtcourant dtforce dtdrag dtav vsig2max vsigmax hhmax
 * itsdensity ntotal stressmax time; NaN for an unknown name */
double orc_scalar(const orc_sim *s, const char *name);
int orc_set_scalar(orc_sim *s, const char *name, double v);
int orc_set_ireal(orc_sim *s, const int32_t *ireal);

/* message of the last failure */
const char *orc_message(const orc_sim *s);

/* neighbour sets of the real particles at the current h, through the cells
 * of the last link list: gather (q2_i < R^2) if sym=0, symmetric
 * (q2_i < R^2 or q2_j < R^2) if sym=1.  Each pair is (i, j_real, image code),
 * 0-based indices; returns the pair count (or -needed if cap is too small). */
int64_t orc_neighbours(orc_sim *s, int sym, int32_t *pairs, int64_t cap);

/* FP64 conserved quantities of the real particles
 * (src/evwrite_mhd.f90:124-248): out[0..16] = ekin etherm emag etot
 * momx momy momz momtot angx angy angz angtot rhomin rhomax rhomean
 * totmassgas totmassdust */
int orc_conserved(const orc_sim *s, double *out);

/* kernel table access (src/kernelND.f90:127-4291, 4426-4469) */
int orc_kernel_table(int ikernel, int ndim, double *w, double *grw,
                     double *grgrw, double *radkern);
void orc_interpolate(int ikernel, int ndim, double q2, double *w, double *grw,
                     double *grgrw);

#ifdef __cplusplus
}
#endif
#endif
