/*
 * CPU oracle for the ndspmhd hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A serial FP64 restatement in plain C of the reference's derivs call graph
 * and leapfrog step, kept structurally faithful so that it reproduces the
 * reference's arithmetic (ghost copies, link-list cells, symmetric first
 * density pass then gather passes, the Newton-Raphson h-rho update, the pair
 * force terms) as closely as possible.  It is the checker for the HIP
 * library, never part of it.  Every function names the reference lines it
 * restates (paths relative to the reference's src/).
 *
 * Parity pins (tests/test_oracle_*.py):
 *  - 3D: against golden vectors written by the compiled reference
 *    (oracle/ref/harness.f90 + the unmodified reference sources):
 *    periodic MHD box (cubic kernel) and periodic one-fluid dust box
 *    (quintic kernel), fixed h and full h-rho iteration, neighbour sets;
 *  - 1D/2D (the reference's 1D/2D builds do not compile with amdflang):
 *    kernel normalisation integrals, lattice density sums, brute-force
 *    neighbour sets, Sigma m a = 0, uniform-state zero forces, dispersion-free
 *    closed forms; these functions are "parity pinned by closed forms only"
 *    in 1D and 2D.
 *  - leapfrog (stepND_leapfrog_mhd.f90 needs the un-vendored SPLASH geometry
 *    module and cannot be built): closed-form constant-force trajectory and
 *    the predictor/corrector identities.
 *
 * Supported option branches (anything else is rejected by orc_new):
 *  ndim 1..3; ikernel 0 (M4 cubic) or 3 (M6 quintic); ikernav=3, ihvar=2,
 *  icty=0, iprterm=0; iener 0 or 2; imhd 0 or 11 with imagforce 2;
 *  idivbzero 0 or 2; iav 2; iavlim(1) in {0,1,2}, iavlim(2) in {0,1},
 *  iavlim(3) in {0,1,3}; idust 0 or 4 with idrag_nature 2 and idustevol 0;
 *  ibound 0, 2 (reflecting, one dimension only) or 3 (periodic).
 */
#include "ndsp_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define IKERN 4000
#define EPSD DBL_EPSILON   /* epsilon(1.0d0) */
#define TINYD DBL_MIN      /* tiny(1.0d0) */
#define HUGED DBL_MAX      /* huge(1.0d0) */
#define PI_K 3.141592653589 /* kernelND.f90:41 */

enum { ITYPEGAS = 0, ITYPEBND = 1, ITYPEDUST = 2, ITYPEBNDDUST = 12 };

struct orc_sim {
  /* options */
  int ndim, ikernel, iener, imhd, imagforce, idivbzero, iav, iavlim[3], idust;
  int idrag_nature, idustevol, ibound[3], maxdensits, onef_dust;
  double hfact, tolh, gamma, polyk, Kdrag, alphamin, alphaumin, alphabmin;
  double beta, avdecayconst, psidecayfact, C_cour, C_force, xmin[3], xmax[3];
  double psep, Bconst[3], dt, damp, avfact, dndim;
  /* kernel tables, kernelND.f90:39-45 */
  double radkern, radkern2, dq2table, ddq2table;
  double wij[IKERN + 1], grwij[IKERN + 1], grgrwij[IKERN + 1];
  /* particles: real 0..npart-1, ghosts npart..ntotal-1 */
  int npart, ntotal, cap;
  double *x, *vel, *pmass, *rho, *hh, *uu, *en, *pr, *spsound, *dens, *Bevol;
  double *Bfield, *alpha, *psi, *gradh, *gradgradh, *force, *drhodt, *dudt;
  double *dendt, *dBevoldt, *daldt, *dpsidt, *gradpsi, *dhdt, *divB, *curlB;
  double *dustfrac, *dustevol, *ddustevoldt, *deltav, *rhogas, *rhodust;
  double *h1, *del2u, *hhin_nr, *graddivv;
  int *itype, *ireal, *numneigh, *ll, *iamincell, *nits, *redolist, *redoprev;
  int *listneigh;
  /* cells, linkND.f90:45-161 */
  int ncells, ncellsx[3], *ifirstincell;
  double dxcell, hhmax;
  /* step in-copies */
  double *xin, *velin, *Bevolin, *rhoin, *hhin, *enin, *alphain, *psiin;
  double *forcein, *dBevoldtin, *drhodtin, *dhdtin, *dendtin, *daldtin;
  double *dpsidtin, *dustevolin, *ddustevoldtin;
  /* scalars */
  double dtcourant, dtforce, dtdrag, dtav, dtvisc, vsig2max, vsigmax;
  double stressmax, time;
  int itsdensity, nsteps;
  char msg[256];
};

/* ------------------------------------------------------------------------ */
/* small helpers                                                             */
/* ------------------------------------------------------------------------ */

/* x**n for the integer powers 1..3 the reference uses (ndim, 2) */
static inline double pown(double x, int n) {
  if (n == 1) return x;
  if (n == 2) return x * x;
  return x * x * x;
}

static int fail(orc_sim *s, int code, const char *m) {
  snprintf(s->msg, sizeof s->msg, "%s", m);
  return code;
}

/* per-particle double fields and their widths, for (re)allocation and
 * copy_particle */
#define NDFIELDS 38
static double **dfield(orc_sim *s, int k, int *w) {
  static const int widths[NDFIELDS] = {3, 3, 1, 1, 1, 1, 1, 1, 1, 1, 3, 3,
                                       3, 1, 1, 1, 3, 1, 1, 1, 3, 3, 1, 3,
                                       1, 1, 3, 1, 1, 1, 3, 1, 1, 1, 1, 1,
                                       3, 1};
  double **p[NDFIELDS] = {
      &s->x,      &s->vel,      &s->pmass,    &s->rho,     &s->hh,
      &s->uu,     &s->en,       &s->pr,       &s->spsound, &s->dens,
      &s->Bevol,  &s->Bfield,   &s->alpha,    &s->psi,     &s->gradh,
      &s->gradgradh, &s->force, &s->drhodt,   &s->dudt,    &s->dendt,
      &s->dBevoldt, &s->daldt,  &s->dpsidt,   &s->gradpsi, &s->dhdt,
      &s->divB,   &s->curlB,    &s->dustfrac, &s->dustevol, &s->ddustevoldt,
      &s->deltav, &s->rhogas,   &s->rhodust,  &s->h1,      &s->del2u,
      &s->hhin_nr, &s->graddivv, &s->hhin_nr};
  *w = widths[k];
  return p[k];
}
#define NDUNIQUE 37 /* the last entry above is padding */

static int grow(orc_sim *s, int newcap) {
  if (newcap <= s->cap) return 0;
  for (int k = 0; k < NDUNIQUE; k++) {
    int w;
    double **p = dfield(s, k, &w);
    double *q = (double *)realloc(*p, sizeof(double) * (size_t)w * newcap);
    if (!q) return -1;
    memset(q + (size_t)w * s->cap, 0, sizeof(double) * (size_t)w * (newcap - s->cap));
    *p = q;
  }
  int **ip[] = {&s->itype, &s->ireal, &s->numneigh, &s->ll, &s->iamincell,
                &s->nits, &s->listneigh};
  for (unsigned k = 0; k < sizeof ip / sizeof ip[0]; k++) {
    int *q = (int *)realloc(*ip[k], sizeof(int) * newcap);
    if (!q) return -1;
    memset(q + s->cap, 0, sizeof(int) * (newcap - s->cap));
    *ip[k] = q;
  }
  s->cap = newcap;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* kernel tables and interpolation                                           */
/* ------------------------------------------------------------------------ */

/* setkerntable, kernelND.f90:127-161 (set-up), :211-269 (M6 quintic),
 * :4143-4190 (M4 cubic), :4279-4289 (normalisation, ddq2table) */
static int setkerntable(int ikernel, int ndim, double *wkern, double *grwkern,
                        double *grgrwkern, double *radkern, double *dq2table) {
  double cnormk = 0.0;
  if (ikernel == 3) {
    *radkern = 3.0;
    double r2 = (*radkern) * (*radkern);
    *dq2table = r2 / (double)IKERN;
    if (ndim == 1) cnormk = 1. / 120.;
    else if (ndim == 2) cnormk = 7. / (478 * PI_K);
    else cnormk = 1. / (120. * PI_K);
    for (int i = 0; i <= IKERN; i++) {
      double q2 = i * (*dq2table);
      double q4 = q2 * q2;
      double q = sqrt(q2);
      double term1 = -5. * pow(3. - q, 4.);
      if (q < 1.0) {
        wkern[i] = 66. - 60. * q2 + 30. * q4 - 10. * q4 * q;
        grwkern[i] = term1 + 30 * pow(2. - q, 4) - 75. * pow(1. - q, 4);
        grgrwkern[i] = 20. * pow(3. - q, 3) - 120. * pow(2. - q, 3) + 300. * pow(1. - q, 3);
      } else if (q >= 1.0 && q < 2.0) {
        wkern[i] = pow(3. - q, 5) - 6. * pow(2. - q, 5);
        grwkern[i] = term1 + 30 * pow(2. - q, 4);
        grgrwkern[i] = 20. * pow(3. - q, 3) - 120. * pow(2. - q, 3);
      } else if (q >= 2.0 && q <= 3.0) {
        wkern[i] = pow(3. - q, 5);
        grwkern[i] = term1;
        grgrwkern[i] = 20. * pow(3. - q, 3);
      } else {
        wkern[i] = 0.0;
        grwkern[i] = 0.0;
        grgrwkern[i] = 0.;
      }
    }
  } else if (ikernel == 0) {
    *radkern = 2.0;
    *dq2table = (*radkern) * (*radkern) / (double)IKERN;
    if (ndim == 1) cnormk = 0.66666666666;
    else if (ndim == 2) cnormk = 10. / (7. * PI_K);
    else cnormk = 1. / PI_K;
    for (int i = 0; i <= IKERN; i++) {
      double q2 = i * (*dq2table);
      double q = sqrt(q2);
      if (q < 1.0) {
        wkern[i] = 1. - 1.5 * q2 + 0.75 * q * q2;
        grwkern[i] = -3. * q + 2.25 * q2;
        grgrwkern[i] = -3. + 4.5 * q;
      } else if (q >= 1.0 && q <= 2.0) {
        double t = 2. - q;
        wkern[i] = 0.25 * (t * t * t);
        grwkern[i] = -0.75 * (t * t);
        grgrwkern[i] = 1.5 * t;
      } else {
        wkern[i] = 0.0;
        grwkern[i] = 0.0;
        grgrwkern[i] = 0.;
      }
    }
  } else {
    return -1;
  }
  for (int i = 0; i <= IKERN; i++) {
    wkern[i] = cnormk * wkern[i];
    grwkern[i] = cnormk * grwkern[i];
    grgrwkern[i] = cnormk * grgrwkern[i];
  }
  return 0;
}

/* interpolate_kernels_dens / interpolate_kernels, kernelND.f90:4599-4638,
 * 4426-4469: linear interpolation in q^2 with the index clamped to ikern */
static inline void kinterp(const orc_sim *s, double q2, double *w, double *grw,
                           double *grgrw) {
  double t = q2 * s->ddq2table;
  int index = (t < (double)(IKERN + 1)) ? (int)t : IKERN + 1;
  int index1 = index + 1;
  if (index > IKERN || index < 0) index = IKERN;
  if (index1 > IKERN || index1 < 0) index1 = IKERN;
  double dxx = q2 - index * s->dq2table;
  double dwdx = (s->wij[index1] - s->wij[index]) * s->ddq2table;
  *w = (s->wij[index] + dwdx * dxx);
  double dgrwdx = (s->grwij[index1] - s->grwij[index]) * s->ddq2table;
  *grw = (s->grwij[index] + dgrwdx * dxx);
  if (grgrw) {
    double dgrgrwdx = (s->grgrwij[index1] - s->grgrwij[index]) * s->ddq2table;
    *grgrw = (s->grgrwij[index] + dgrgrwdx * dxx);
  }
}

/* ------------------------------------------------------------------------ */
/* boundaries and ghosts                                                     */
/* ------------------------------------------------------------------------ */

/* fixed boundary particles in use (any ibound = 1, set_fixedbound.f90) */
static int anyfixed(const orc_sim *s) {
  for (int d = 0; d < s->ndim; d++)
    if (s->ibound[d] == 1) return 1;
  return 0;
}

/* copy_particle, copy_particle.f90:28-115 (the fields this oracle keeps) */
static void copy_particle(orc_sim *s, int i, int j) {
  for (int k = 0; k < NDUNIQUE; k++) {
    int w;
    double **p = dfield(s, k, &w);
    if (*p == s->x || *p == s->vel || *p == s->h1 || *p == s->del2u ||
        *p == s->hhin_nr)
      continue;
    for (int c = 0; c < w; c++) (*p)[(size_t)i * w + c] = (*p)[(size_t)j * w + c];
  }
}

/* makeghost, ghostND_mhd.f90:363-420 */
static int makeghost(orc_sim *s, int jpart, const double *xg, const double *vg) {
  int ipart = s->ntotal;
  if (ipart >= s->cap && grow(s, s->cap * 2 + 16)) return -1;
  s->ntotal = ipart + 1;
  for (int d = 0; d < s->ndim; d++) s->x[ipart * 3 + d] = xg[d];
  copy_particle(s, ipart, jpart);
  for (int d = 0; d < 3; d++) s->vel[ipart * 3 + d] = vg[d];
  s->ireal[ipart] = jpart;
  return 0;
}

static inline int isreflect(int ib) { return ib == 2 || ib == 4 || ib == 6; }

/* set_ghost_particles, ghostND_mhd.f90:33-360 (non-shearing boundaries) */
static int set_ghost_particles(orc_sim *s) {
  const int ndim = s->ndim;
  s->ntotal = s->npart;
  double hhmax = s->hh[0];
  for (int i = 1; i < s->npart; i++) hhmax = fmax(hhmax, s->hh[i]);
  s->hhmax = hhmax;
  double dxbound[3];
  for (int d = 0; d < 3; d++) dxbound[d] = s->radkern * hhmax;
  int imakeghost[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  double xnew[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  for (int jpart = 0; jpart < s->npart; jpart++) {
    /* local copies: makeghost may reallocate the particle arrays */
    double xj[3], vj[3];
    for (int d = 0; d < 3; d++) { xj[d] = s->x[jpart * 3 + d]; vj[d] = s->vel[jpart * 3 + d]; }
    for (int d = 0; d < ndim; d++)
      if (isreflect(s->ibound[d])) dxbound[d] = s->radkern * s->hh[jpart];
    for (int idimen = 0; idimen < ndim; idimen++) {
      if (s->ibound[idimen] <= 1) {
        imakeghost[idimen][0] = imakeghost[idimen][1] = 0;
        continue;
      }
      for (int imaxmin = 0; imaxmin < 2; imaxmin++) {
        double xpart[3], vpart[3], xbound, xperbound, dx;
        for (int d = 0; d < 3; d++) { xpart[d] = xj[d]; vpart[d] = vj[d]; }
        if (imaxmin == 0) {
          xbound = s->xmax[idimen]; xperbound = s->xmin[idimen];
          dx = s->xmax[idimen] - xj[idimen];
        } else {
          xbound = s->xmin[idimen]; xperbound = s->xmax[idimen];
          dx = xj[idimen] - s->xmin[idimen];
        }
        imakeghost[idimen][imaxmin] = (dx < dxbound[idimen]) && (dx > 0);
        if (!imakeghost[idimen][imaxmin]) continue;
        double dxshift = xj[idimen] - xbound;
        if (s->ibound[idimen] == 3) {
          xnew[idimen][imaxmin] = xperbound + dxshift;
        } else {
          xnew[idimen][imaxmin] = xbound - dxshift;
          vpart[idimen] = -vj[idimen];
        }
        xpart[idimen] = xnew[idimen][imaxmin];
        if (makeghost(s, jpart, xpart, vpart)) return -1;
        for (int ip = 0; ip < idimen; ip++) {
          for (int mp = 0; mp < 2; mp++) {
            if (!imakeghost[ip][mp]) continue;
            xpart[ip] = xnew[ip][mp];
            if (isreflect(s->ibound[ip]))
              for (int d = 0; d < 3; d++) vpart[d] = vj[d];
            if (makeghost(s, jpart, xpart, vpart)) return -1;
            if (ip >= 1) {
              int pp = ip - 1;
              for (int mpp = 0; mpp < 2; mpp++) {
                if (!imakeghost[pp][mpp]) continue;
                xpart[pp] = xnew[pp][mpp];
                if (isreflect(s->ibound[pp]))
                  for (int d = 0; d < 3; d++) vpart[d] = -vj[d];
                if (makeghost(s, jpart, xpart, vpart)) return -1;
                xpart[pp] = xj[pp];
              }
            }
            xpart[ip] = xj[ip];
          }
        }
      }
    }
  }
  for (int i = s->npart; i < s->ntotal; i++) s->itype[i] = s->itype[s->ireal[i]];
  return 0;
}

/* boundary, boundaryND.f90:30-164 (periodic wrap, then reflecting walls) */
static void boundary(orc_sim *s) {
  const int ndim = s->ndim;
  int anyper = 0, anyghost = 0;
  for (int d = 0; d < ndim; d++) {
    if (s->ibound[d] == 3) anyper = 1;
    if (s->ibound[d] > 1) anyghost = 1;
  }
  if (anyper) {
    for (int i = 0; i < s->npart; i++)
      for (int d = 0; d < ndim; d++) {
        if (s->ibound[d] != 3) continue;
        double *xi = &s->x[i * 3 + d];
        if (*xi > s->xmax[d]) *xi = s->xmin[d] + *xi - s->xmax[d];
        else if (*xi < s->xmin[d]) *xi = s->xmax[d] - (s->xmin[d] - *xi);
      }
  }
  if (anyghost) {
    for (int i = 0; i < s->npart; i++)
      for (int d = 0; d < ndim; d++) {
        if (s->ibound[d] <= 1) continue;
        double *xi = &s->x[i * 3 + d];
        if (*xi > s->xmax[d]) {
          *xi = s->xmax[d] - (*xi - s->xmax[d]);
          if (s->xin) s->xin[i * 3 + d] = *xi;
          s->vel[i * 3 + d] = -s->vel[i * 3 + d];
          if (s->velin) s->velin[i * 3 + d] = s->vel[i * 3 + d];
          s->force[i * 3 + d] = 0.;
        } else if (*xi < s->xmin[d]) {
          *xi = s->xmin[d] + (s->xmin[d] - *xi);
          if (s->xin) s->xin[i * 3 + d] = *xi;
          s->vel[i * 3 + d] = -s->vel[i * 3 + d];
          if (s->velin) s->velin[i * 3 + d] = s->vel[i * 3 + d];
          s->force[i * 3 + d] = 0.;
        }
      }
  }
}

/* ------------------------------------------------------------------------ */
/* link list and neighbour lists                                             */
/* ------------------------------------------------------------------------ */

/* set_linklist, linkND.f90:45-161 (cells are 0-based here) */
static int set_linklist(orc_sim *s) {
  const int ndim = s->ndim;
  int allfree = 1;
  for (int d = 0; d < ndim; d++) if (s->ibound[d] > 1) allfree = 0;
  if (allfree) {
    double hm = s->hh[0];
    for (int i = 1; i < s->npart; i++) hm = fmax(hm, s->hh[i]);
    s->hhmax = hm;
  }
  s->dxcell = s->radkern * s->hhmax;
  if (s->dxcell <= 0) return fail(s, 5, "link: max h <= 0");
  double xminpart[3], xmaxpart[3];
  for (int d = 0; d < ndim; d++) {
    double lo = s->x[d], hi = s->x[d];
    for (int i = 1; i < s->ntotal; i++) {
      lo = fmin(lo, s->x[i * 3 + d]);
      hi = fmax(hi, s->x[i * 3 + d]);
    }
    xminpart[d] = lo - 0.00001;
    xmaxpart[d] = hi + 0.00001;
    xminpart[d] = xminpart[d] - s->dxcell - 0.00001;
    xmaxpart[d] = xmaxpart[d] + s->dxcell + 0.00001;
  }
  long nc = 1;
  for (int d = 0; d < 3; d++) s->ncellsx[d] = 1;
  for (int d = 0; d < ndim; d++) {
    s->ncellsx[d] = (int)((xmaxpart[d] - xminpart[d]) / s->dxcell) + 1;
    if (s->ncellsx[d] == 0) return fail(s, 5, "link: number of cells = 0");
    nc *= s->ncellsx[d];
  }
  if (nc > 400000000L) return fail(s, 5, "link: too many cells");
  s->ncells = (int)nc;
  free(s->ifirstincell);
  s->ifirstincell = (int *)malloc(sizeof(int) * (size_t)nc);
  for (long c = 0; c < nc; c++) s->ifirstincell[c] = -1;
  for (int i = 0; i < s->ntotal; i++) {
    s->ll[i] = -1;
    int icellx[3] = {1, 1, 1};
    for (int d = 0; d < ndim; d++) {
      icellx[d] = (int)((s->x[i * 3 + d] - xminpart[d]) / s->dxcell) + 1;
      if (icellx[d] < 0 || icellx[d] > s->ncellsx[d])
        return fail(s, 6, "link: particle crossed boundary");
    }
    long icell = icellx[0];
    if (ndim >= 2) icell += (long)(icellx[1] - 1) * s->ncellsx[0];
    if (ndim >= 3) icell += (long)(icellx[2] - 1) * s->ncellsx[0] * s->ncellsx[1];
    icell -= 1;
    if (icell < 0 || icell >= nc) icell = nc - 1;
    s->ll[i] = s->ifirstincell[icell];
    s->ifirstincell[icell] = i;
    s->iamincell[i] = (int)icell;
  }
  return 0;
}

static int walk_cells(const orc_sim *s, const long *cells, int ncell, int *list) {
  int j = 0;
  for (int k = 0; k < ncell; k++) {
    long c = cells[k];
    if (c < 0 || c >= s->ncells) continue;
    int ipart = s->ifirstincell[c];
    if (ipart != -1) {
      while (s->ll[ipart] != -1) {
        list[j++] = ipart;
        ipart = s->ll[ipart];
      }
      list[j++] = ipart;
    }
  }
  return j;
}

/* get_neighbour_list, get_neighbour_lists.f90:37-165 (half stencil) */
static int get_neighbour_list(const orc_sim *s, int icell, int *list) {
  if (s->ifirstincell[icell] < 0) return 0;
  long nx = s->ncellsx[0], nxy = (long)s->ncellsx[0] * s->ncellsx[1];
  long c[14];
  int n = 0;
  c[n++] = icell;
  c[n++] = icell + 1;
  if (s->ndim >= 2) {
    c[n++] = icell + nx - 1; c[n++] = icell + nx; c[n++] = icell + nx + 1;
  }
  if (s->ndim >= 3) {
    c[n++] = icell + nxy - 1;      c[n++] = icell + nxy;      c[n++] = icell + nxy + 1;
    c[n++] = icell + nxy + nx - 1; c[n++] = icell + nxy + nx; c[n++] = icell + nxy + nx + 1;
    c[n++] = icell + nxy - nx - 1; c[n++] = icell + nxy - nx; c[n++] = icell + nxy - nx + 1;
  }
  return walk_cells(s, c, n, list);
}

/* get_neighbour_list_partial, get_neighbour_lists.f90:176-316 (full) */
static int get_neighbour_list_partial(const orc_sim *s, int icell, int *list) {
  if (s->ifirstincell[icell] < 0) return 0;
  long nx = s->ncellsx[0], nxy = (long)s->ncellsx[0] * s->ncellsx[1];
  long c[27];
  int n = 0;
  c[n++] = icell - 1; c[n++] = icell; c[n++] = icell + 1;
  if (s->ndim >= 2) {
    c[n++] = icell + nx - 1; c[n++] = icell + nx; c[n++] = icell + nx + 1;
    c[n++] = icell - nx - 1; c[n++] = icell - nx; c[n++] = icell - nx + 1;
    if (s->ndim >= 3) {
      c[n++] = icell + nxy + nx - 1; c[n++] = icell + nxy + nx; c[n++] = icell + nxy + nx + 1;
      c[n++] = icell + nxy - 1;      c[n++] = icell + nxy;      c[n++] = icell + nxy + 1;
      c[n++] = icell + nxy - nx - 1; c[n++] = icell + nxy - nx; c[n++] = icell + nxy - nx + 1;
      c[n++] = icell - nxy + nx - 1; c[n++] = icell - nxy + nx; c[n++] = icell - nxy + nx + 1;
      c[n++] = icell - nxy - 1;      c[n++] = icell - nxy;      c[n++] = icell - nxy + 1;
      c[n++] = icell - nxy - nx - 1; c[n++] = icell - nxy - nx; c[n++] = icell - nxy - nx + 1;
    }
  }
  return walk_cells(s, c, n, list);
}

static inline int type_pair_ok(int ti, int tj) {
  return (tj == ti) || (ti == ITYPEGAS && tj == ITYPEBND) ||
         (tj == ITYPEGAS && ti == ITYPEBND) ||
         (ti == ITYPEDUST && tj == ITYPEBNDDUST) ||
         (tj == ITYPEDUST && ti == ITYPEBNDDUST);
}

/* ------------------------------------------------------------------------ */
/* density sums                                                              */
/* ------------------------------------------------------------------------ */

/* density, density_sums.f90:38-379: every pair once, self weight 1/2 */
static void density_sym(orc_sim *s) {
  const int ndim = s->ndim;
  const double R2 = s->radkern2;
  for (int i = 0; i < s->npart; i++) {
    if (s->itype[i] == ITYPEBND || s->itype[i] == ITYPEBNDDUST) continue;
    s->rho[i] = 0.; s->drhodt[i] = 0.; s->gradh[i] = 0.; s->gradgradh[i] = 0.;
    if (s->imhd == 0) s->psi[i] = 0.;
    if (s->onef_dust) { s->rhogas[i] = 0.; s->rhodust[i] = 0.; }
  }
  for (int i = 0; i < s->ntotal; i++) s->numneigh[i] = 0;
  for (int i = 0; i < s->ntotal; i++) s->h1[i] = 1. / s->hh[i];
  int *list = s->listneigh;
  for (int icell = 0; icell < s->ncells; icell++) {
    int nneigh = get_neighbour_list(s, icell, list);
    int i = s->ifirstincell[icell];
    int idone = -1;
    while (i != -1) {
      idone++;
      const double pmassi = s->pmass[i];
      const double dustfraci = s->onef_dust ? s->dustfrac[i] : 0.;
      const double *xi = &s->x[i * 3];
      const double *veli = &s->vel[i * 3];
      const double hi1 = s->h1[i];
      const double hfacwabi = pown(hi1, ndim);
      const double hi21 = hi1 * hi1;
      const int itypei = s->itype[i];
      for (int n = idone; n < nneigh; n++) {
        int j = list[n];
        int itypej = s->itype[j];
        if (!type_pair_ok(itypei, itypej)) continue;
        double dx[3] = {0, 0, 0};
        for (int d = 0; d < ndim; d++) dx[d] = xi[d] - s->x[j * 3 + d];
        double hj1 = s->h1[j];
        double rij2 = 0.;
        for (int d = 0; d < ndim; d++) rij2 = rij2 + dx[d] * dx[d];
        double q2i = rij2 * hi21;
        double q2j = rij2 * hj1 * hj1;
        if (!((q2i < R2) || ((q2j < R2) && (i < s->npart || j < s->npart)))) continue;
        if (i < s->npart) s->numneigh[i]++;
        if (j < s->npart && j != i) s->numneigh[j]++;
        double rij = sqrt(rij2);
        double dr[3] = {0, 0, 0};
        for (int d = 0; d < ndim; d++) dr[d] = dx[d] / (rij + EPSD);
        double hfacwabj = pown(hj1, ndim);
        double weight = (j == i) ? 0.5 : 1.0;
        double pmassj = s->pmass[j];
        double wabi, grkerni, grgrkerni, wabj, grkernj, grgrkernj;
        kinterp(s, q2i, &wabi, &grkerni, &grgrkerni);
        kinterp(s, q2j, &wabj, &grkernj, &grgrkernj);
        wabi = wabi * hfacwabi;
        grkerni = grkerni * hfacwabi * hi1;
        grgrkerni = grgrkerni * hfacwabi * hi1 * hi1;
        wabj = wabj * hfacwabj;
        grkernj = grkernj * hfacwabj * hj1;
        grgrkernj = grgrkernj * hfacwabj * hj1 * hj1;
        double dwdhi = -(rij * grkerni * hi1) - ndim * wabi * hi1;
        double dwdhj = -(rij * grkernj * hj1) - ndim * wabj * hj1;
        double hi1s = hi1 * hi1, hj1s = hj1 * hj1;
        double dwdhdhi = (double)(ndim * (ndim + 1)) * wabi * hi1s +
                         2. * (ndim + 1) * rij * hi1s * grkerni + rij * rij * hi1s * grgrkerni;
        double dwdhdhj = (double)(ndim * (ndim + 1)) * wabj * hj1s +
                         2. * (ndim + 1) * rij * hj1s * grkernj + rij * rij * hj1s * grgrkernj;
        if (itypei != ITYPEBND) {
          s->rho[i] = s->rho[i] + pmassj * wabi * weight;
          if (s->onef_dust) {
            double dustfracj = s->dustfrac[j];
            s->rhodust[i] = s->rhodust[i] + pmassj * s->dustfrac[j] * wabi * weight;
            s->rhogas[i] = s->rhogas[i] + pmassj * (1. - dustfracj) * wabi * weight;
          }
        }
        if (itypej != ITYPEBND) {
          s->rho[j] = s->rho[j] + pmassi * wabj * weight;
          if (s->onef_dust) {
            s->rhodust[j] = s->rhodust[j] + pmassi * s->dustfrac[i] * wabj * weight;
            s->rhogas[j] = s->rhogas[j] + pmassi * (1. - dustfraci) * wabj * weight;
          }
        }
        if (i != j) {
          double dvdotr = 0.;
          for (int d = 0; d < 3; d++) dvdotr = dvdotr + (veli[d] - s->vel[j * 3 + d]) * dr[d];
          s->drhodt[i] = s->drhodt[i] + pmassj * dvdotr * grkerni;
          s->drhodt[j] = s->drhodt[j] + pmassi * dvdotr * grkernj;
        }
        if (itypei != ITYPEBND) {
          s->gradh[i] = s->gradh[i] + weight * pmassj * dwdhi;
          s->gradgradh[i] = s->gradgradh[i] + weight * pmassj * dwdhdhi;
        }
        if (itypej != ITYPEBND) {
          s->gradh[j] = s->gradh[j] + weight * pmassi * dwdhj;
          s->gradgradh[j] = s->gradgradh[j] + weight * pmassi * dwdhdhj;
        }
      }
      i = s->ll[i];
    }
  }
}

/* density_partial, density_sums.f90:396-658: gather on a list of particles */
static void density_partial(orc_sim *s, int nlist, const int *ipartlist) {
  const int ndim = s->ndim;
  const double R2 = s->radkern2;
  for (int ip = 0; ip < nlist; ip++) {
    int i = ipartlist[ip];
    s->rho[i] = 0.; s->drhodt[i] = 0.; s->gradh[i] = 0.; s->gradgradh[i] = 0.;
    s->numneigh[i] = 0;
    if (s->onef_dust) { s->rhodust[i] = 0.; s->rhogas[i] = 0.; }
  }
  int *list = s->listneigh;
  int icellprev = -1, nneigh = 0;
  for (int ip = 0; ip < nlist; ip++) {
    int i = ipartlist[ip];
    int icell = s->iamincell[i];
    if (icell != icellprev) nneigh = get_neighbour_list_partial(s, icell, list);
    icellprev = icell;
    const double hi = s->hh[i];
    const double hi1 = 1. / hi;
    const double hi21 = hi1 * hi1;
    const double hfacwabi = pown(hi1, ndim);
    const double hfacgrkerni = hfacwabi * hi1;
    const double *xi = &s->x[i * 3];
    const double *veli = &s->vel[i * 3];
    const int itypei = s->itype[i];
    if (itypei == ITYPEBND) continue;
    for (int n = 0; n < nneigh; n++) {
      int j = list[n];
      if (s->itype[j] != itypei && s->itype[j] != ITYPEBND) continue;
      double dx[3] = {0, 0, 0};
      for (int d = 0; d < ndim; d++) dx[d] = xi[d] - s->x[j * 3 + d];
      double rij2 = 0.;
      for (int d = 0; d < ndim; d++) rij2 = rij2 + dx[d] * dx[d];
      double q2i = rij2 * hi21;
      if (!(q2i < R2)) continue;
      double rij = sqrt(rij2);
      double dr[3] = {0, 0, 0};
      for (int d = 0; d < ndim; d++) dr[d] = dx[d] / (rij + EPSD);
      s->numneigh[i]++;
      double pmassj = s->pmass[j];
      double wabi, grkerni, grgrkerni;
      kinterp(s, q2i, &wabi, &grkerni, &grgrkerni);
      wabi = wabi * hfacwabi;
      grkerni = grkerni * hfacgrkerni;
      grgrkerni = grgrkerni * hfacwabi * hi1 * hi1;
      double dwdhi = -(rij * grkerni * hi1) - ndim * wabi * hi1;
      double hi1s = hi1 * hi1;
      double dwdhdhi = (double)(ndim * (ndim + 1)) * wabi * hi1s +
                       2. * (ndim + 1) * rij * hi1s * grkerni + rij * rij * hi1s * grgrkerni;
      s->rho[i] = s->rho[i] + pmassj * wabi;
      if (s->onef_dust) {
        double dustfracj = s->dustfrac[j];
        s->rhodust[i] = s->rhodust[i] + pmassj * s->dustfrac[j] * wabi;
        s->rhogas[i] = s->rhogas[i] + pmassj * (1. - dustfracj) * wabi;
      }
      if (i != j) {
        double dvdotr = 0.;
        for (int d = 0; d < 3; d++) dvdotr = dvdotr + (veli[d] - s->vel[j * 3 + d]) * dr[d];
        s->drhodt[i] = s->drhodt[i] + pmassj * dvdotr * grkerni;
      }
      s->gradh[i] = s->gradh[i] + pmassj * dwdhi;
      s->gradgradh[i] = s->gradgradh[i] + pmassj * dwdhdhi;
    }
    /* density_sums.f90:630-635: psi is zeroed on every re-iterated particle
     * when MHD is on */
    if (s->imhd != 0) s->psi[i] = 0.;
  }
}

/* iterate_density, iterate_density.f90:43-360 (ikernav=3, usenumdens off) */
static int iterate_density(orc_sim *s) {
  const int ndim = s->ndim;
  const int npart = s->npart;
  const int itsmax = s->maxdensits;
  int ncalc = npart, redolink = 0;
  s->itsdensity = 0;
  for (int i = 0; i < s->ntotal; i++) {
    if (s->itype[i] == ITYPEBND) continue;
    s->gradh[i] = 0.; s->gradgradh[i] = 0.; s->drhodt[i] = 0.; s->dhdt[i] = 0.;
  }
  for (int i = 0; i < npart; i++) {
    s->hhin_nr[i] = s->hh[i];
    if (s->hh[i] <= TINYD) return fail(s, 1, "h <= 0 in density call");
    s->nits[i] = 0;
  }
  for (int j = 0; j < npart; j++) s->redolist[j] = j;
  int anyghost = 0;
  for (int d = 0; d < ndim; d++) if (s->ibound[d] > 1) anyghost = 1;
  while (ncalc > 0 && s->itsdensity <= itsmax) {
    s->itsdensity++;
    if (redolink) {
      if (anyghost && set_ghost_particles(s)) return fail(s, 7, "ghost allocation failed");
      int e = set_linklist(s);
      if (e) return e;
    }
    if (ncalc == npart) density_sym(s);
    else density_partial(s, ncalc, s->redolist);
    int ncalcprev = ncalc;
    memcpy(s->redoprev, s->redolist, sizeof(int) * ncalcprev);
    ncalc = 0;
    redolink = 0;
    for (int jj = 0; jj < ncalcprev; jj++) {
      int i = s->redoprev[jj];
      if (s->itype[i] == ITYPEBND || s->itype[i] == ITYPEBNDDUST) continue;
      s->nits[i] = s->itsdensity;
      if (s->rho[i] <= 1.e-6 && s->rho[i] <= 0.) return fail(s, 2, "rho <= 0");
      const double h_min = 0., rhomin = 0.;
      double rhoi = s->pmass[i] / pown((s->hh[i] - h_min) / s->hfact, ndim) - rhomin;
      double dhdrhoi = -(s->hh[i] - h_min) / (ndim * (s->rho[i] + rhomin));
      double dwdhsumi = s->gradh[i];
      double omegai = 1. - dhdrhoi * s->gradh[i];
      if (omegai < 1.e-5) {
        if (fabs(omegai) == 0.) omegai = 1.;
      }
      s->gradh[i] = 1. / omegai;
      double func = rhoi - s->rho[i];
      double dfdh = omegai / dhdrhoi;
      double rn = s->rho[i] * ndim;
      double d2hdrho2i = s->hh[i] * (ndim + 1) / (rn * rn);
      s->gradgradh[i] = s->rho[i] * (d2hdrho2i * dwdhsumi + dhdrhoi * dhdrhoi * s->gradgradh[i]);
      double hnew = s->hh[i] - func / dfdh;
      if (hnew > 1.2 * s->hh[i]) hnew = 1.2 * s->hh[i];
      else if (hnew < 0.8 * s->hh[i]) hnew = 0.8 * s->hh[i];
      if (hnew <= 0 || s->gradh[i] <= TINYD)
        hnew = s->hfact * pow(s->pmass[i] / (s->rho[i] + rhomin), s->dndim);
      else if (s->itsdensity > 100)
        hnew = s->hfact * pow(s->pmass[i] / (s->rho[i] + rhomin), s->dndim);
      if (s->numneigh[i] <= 1) {
        hnew = s->hh[i] + s->psep;
        redolink = 1;
      }
      int converged = (fabs((hnew - s->hh[i]) / s->hhin_nr[i]) < s->tolh && omegai > 0.) || itsmax == 0;
      if (!converged) {
        s->redolist[ncalc++] = i;
        if (s->itsdensity <= itsmax && s->itype[i] != ITYPEBND) s->hh[i] = hnew;
        if (hnew > s->hhmax) redolink = 1;
      } else {
        s->drhodt[i] = s->drhodt[i] * s->gradh[i];
        s->dhdt[i] = dhdrhoi * s->drhodt[i];
      }
    }
    if (anyfixed(s)) {
      /* write over boundary particles, iterate_density.f90:310-326 */
      for (int i = 0; i < npart; i++) {
        if (s->itype[i] != ITYPEBND) continue;
        int j = s->ireal[i];
        if (j < 0) continue;
        s->rho[i] = s->rho[j]; s->drhodt[i] = s->drhodt[j]; s->dhdt[i] = s->dhdt[j];
        s->hh[i] = s->hh[j]; s->gradh[i] = s->gradh[j];
      }
    }
    if (anyghost) {
      for (int i = npart; i < s->ntotal; i++) {
        int j = s->ireal[i];
        s->rho[i] = s->rho[j]; s->drhodt[i] = s->drhodt[j]; s->dhdt[i] = s->dhdt[j];
        s->hh[i] = s->hh[j]; s->gradh[i] = s->gradh[j];
      }
    }
  }
  if (s->itsdensity > itsmax && itsmax > 0) return fail(s, 3, "density not converged");
  return 0;
}

/* ------------------------------------------------------------------------ */
/* conservative -> primitive and equation of state                           */
/* ------------------------------------------------------------------------ */

/* conservative2primitive, conservative2primitive.f90:42-470, and
 * equation_of_state, eos.f90:40-106, for iener 0/2, imhd 0/11, idustevol 0 */
static int cons2prim(orc_sim *s) {
  const int npart = s->npart;
  if (s->onef_dust) {
    for (int i = 0; i < npart; i++) {
      s->dustfrac[i] = s->dustevol[i];
      if (s->dustfrac[i] > 1.) {
        s->dustfrac[i] = 1.;
        s->dustevol[i] = 1.;
        if (s->dustevolin) s->dustevolin[i] = 1.;
      }
    }
    for (int i = 0; i < npart; i++) s->dens[i] = s->rho[i] * (1. - s->dustfrac[i]);
  } else {
    for (int i = 0; i < s->ntotal; i++) s->dens[i] = s->rho[i];
  }
  if (s->imhd == 11)
    for (int i = 0; i < 3 * s->ntotal; i++) s->Bfield[i] = s->Bevol[i];
  for (int i = 0; i < s->ntotal; i++) s->uu[i] = s->en[i];
  const double *rhoe = s->onef_dust ? s->dens : s->rho;
  const double gamma1 = s->gamma - 1.;
  for (int i = 0; i < npart; i++) {
    if (s->iener == 0) {
      if (rhoe[i] > 0. && s->itype[i] == ITYPEGAS) {
        s->pr[i] = s->polyk * pow(rhoe[i], s->gamma);
        s->spsound[i] = sqrt(s->gamma * s->pr[i] / rhoe[i]);
      } else if (s->itype[i] != ITYPEBND) {
        s->pr[i] = 0.;
      }
      if (fabs(gamma1) > 1.e-3 && rhoe[i] > 0.) s->uu[i] = s->pr[i] / (gamma1 * rhoe[i]);
    } else {
      if (s->uu[i] < 0.) return fail(s, 4, "eos: u_therm -ve");
      if (rhoe[i] > 0.) {
        s->pr[i] = gamma1 * s->uu[i] * rhoe[i];
        s->spsound[i] = sqrt(s->gamma * s->pr[i] / rhoe[i]);
      }
    }
  }
  if (anyfixed(s)) {
    /* make fixed particles exact replicas of their closest particle,
       conservative2primitive.f90:425-432 */
    for (int i = 0; i < npart; i++) {
      if (s->itype[i] != ITYPEBND && s->itype[i] != ITYPEBNDDUST) continue;
      int j = s->ireal[i];
      if (j >= 0) copy_particle(s, i, j);
    }
  }
  int anyghost = 0, allper = 1;
  for (int d = 0; d < s->ndim; d++) {
    if (s->ibound[d] > 1) anyghost = 1;
    if (s->ibound[d] != 3) allper = 0;
  }
  if (anyghost) {
    for (int i = npart; i < s->ntotal; i++) {
      int j = s->ireal[i];
      for (int c = 0; c < 3; c++) s->curlB[i * 3 + c] = s->curlB[j * 3 + c];
      s->psi[i] = s->psi[j];
      s->dens[i] = s->dens[j];
      s->uu[i] = s->uu[j];
      s->spsound[i] = s->spsound[j];
      s->pr[i] = s->pr[j];
      for (int c = 0; c < 3; c++) s->Bfield[i * 3 + c] = s->Bfield[j * 3 + c];
      if (s->onef_dust) s->dustfrac[i] = s->dustfrac[j];
      if (allper) copy_particle(s, i, j);
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* rates                                                                     */
/* ------------------------------------------------------------------------ */

static inline double dot3(const double *a, const double *b) {
  double r = 0.;
  for (int d = 0; d < 3; d++) r = r + a[d] * b[d];
  return r;
}

/* Fortran norm2 of a 3-vector; flang evaluates it as sqrt of the sum of
 * squares (scaled forms differ only below/above double range) */
static inline double norm2_3(const double *a) { return sqrt(dot3(a, a)); }

/* get_rates, ratesND_mhd.f90:29-979, with rates_core :1175-1690,
 * artificial_dissipation :1700-1894,
 * artificial_dissipation_dust_diffusion :2154-2206,
 * mhd_terms (default tensor force) :2377-2720,
 * dust_derivs_diffusion (idustevol 0) :2813-2878 */
static int get_rates(orc_sim *s) {
  const int ndim = s->ndim, npart = s->npart, ntotal = s->ntotal;
  const double R2 = s->radkern2;
  const double zero = 1.e-10;
  const int mhd = s->imhd != 0;
  s->dtcourant = 1.e6;
  s->dtav = HUGED;
  s->dtvisc = HUGED;
  s->vsigmax = 0.;
  for (int i = 0; i < ntotal; i++) {
    for (int c = 0; c < 3; c++) {
      s->force[i * 3 + c] = 0.; s->dBevoldt[i * 3 + c] = 0.; s->daldt[i * 3 + c] = 0.;
      s->gradpsi[i * 3 + c] = 0.;
      if (s->imhd > 0) s->curlB[i * 3 + c] = 0.;
    }
    s->dudt[i] = 0.; s->dendt[i] = 0.; s->dpsidt[i] = 0.; s->divB[i] = 0.;
    s->del2u[i] = 0.;
    for (int c = 0; c < 3; c++) s->graddivv[i * 3 + c] = 0.;
    s->h1[i] = 1. / s->hh[i];
    if (s->onef_dust) { s->ddustevoldt[i] = 0.; }
  }
  /* stress correction pre-pass, :231-245 */
  double stressmax = 0.;
  if (mhd && s->imagforce == 2) {
    double bcmax = fmax(fmax(s->Bconst[0], s->Bconst[1]), s->Bconst[2]);
    for (int i = 0; i < ntotal; i++) {
      double B2i = dot3(&s->Bfield[i * 3], &s->Bfield[i * 3]);
      double stressterm = fmax(0.5 * B2i - s->pr[i], 0.);
      stressmax = fmax(fmax(stressterm, stressmax), bcmax);
    }
  }
  s->stressmax = stressmax;

  int *list = s->listneigh;
  for (int icell = 0; icell < s->ncells; icell++) {
    int nneigh = get_neighbour_list(s, icell, list);
    int i = s->ifirstincell[icell];
    int idone = -1;
    while (i != -1) {
      idone++;
      const double *xi = &s->x[i * 3];
      const int itypei = s->itype[i];
      const double rhoi = s->rho[i];
      const double rho1i = 1. / rhoi;
      const double rho21i = rho1i * rho1i;
      const double pri = fmax(s->pr[i] - 0., 0.);
      const double prneti = pri - 0.;
      const double pmassi = s->pmass[i];
      const double Prho2i = pri * rho21i;
      const double spsoundi = s->spsound[i];
      const double *veli = &s->vel[i * 3];
      const double alphai = s->alpha[i * 3 + 0];
      const double alphaui = s->alpha[i * 3 + 1];
      const double alphaBi = s->alpha[i * 3 + 2];
      double dustfraci = 0., rhogasi = rhoi, rhodusti = 0.;
      if (s->onef_dust) {
        dustfraci = s->dustfrac[i];
        rhodusti = s->rhodust[i];
        rhogasi = s->rhogas[i];
      }
      double Bi[3] = {0, 0, 0}, Brhoi[3] = {0, 0, 0}, B2i = 0., Brho2i = 0., valfven2i = 0.;
      if (mhd) {
        for (int c = 0; c < 3; c++) { Bi[c] = s->Bfield[i * 3 + c]; Brhoi[c] = Bi[c] * rho1i; }
        B2i = dot3(Bi, Bi);
        Brho2i = B2i * rho21i;
        valfven2i = B2i * rho1i;
      }
      const double gradhi = s->gradh[i];
      const double hi = s->hh[i];
      if (hi <= 0.) return fail(s, 1, "rates: h <= 0");
      const double hi1 = s->h1[i];
      const double hi21 = hi1 * hi1;
      const double hfacwabi = pown(hi1, ndim);
      const double hfacgrkerni = hfacwabi * hi1;
      double forcei[3] = {0, 0, 0}, dBevoldti[3] = {0, 0, 0};
      for (int n = idone; n < nneigh; n++) {
        int j = list[n];
        if (j == i || (j >= npart && i >= npart)) continue;
        double dx[3] = {0, 0, 0};
        for (int d = 0; d < ndim; d++) dx[d] = xi[d] - s->x[j * 3 + d];
        const double hj = s->hh[j];
        const double hj1 = s->h1[j];
        const double hj21 = hj1 * hj1;
        double rij2 = 0.;
        for (int d = 0; d < ndim; d++) rij2 = rij2 + dx[d] * dx[d];
        const double q2i = rij2 * hi21;
        const double q2j = rij2 * hj21;
        if (!((q2i < R2) || (q2j < R2))) continue;
        const double rij = sqrt(rij2);
        double dr[3] = {0, 0, 0};
        if (rij <= EPSD) {
          if (rij < TINYD && itypei != 2) return fail(s, 8, "rates: dx = 0");
        } else {
          for (int d = 0; d < ndim; d++) dr[d] = dx[d] / rij;
        }
        const int itypej = s->itype[j];
        if (!type_pair_ok(itypei, itypej)) continue;
        /* ---- rates_core ---- */
        const double pmassj = s->pmass[j];
        double wabi, grkerni, wabj, grkernj;
        kinterp(s, q2i, &wabi, &grkerni, NULL);
        wabi = wabi * hfacwabi;
        grkerni = grkerni * hfacgrkerni;
        const double hfacwabj = pown(hj1, ndim);
        const double hfacgrkernj = hfacwabj * hj1;
        kinterp(s, q2j, &wabj, &grkernj, NULL);
        wabj = wabj * hfacwabj;
        grkernj = grkernj * hfacgrkernj;
        grkerni = grkerni * gradhi;
        grkernj = grkernj * s->gradh[j];
        const double grkern = 0.5 * (grkerni + grkernj);
        const double *velj = &s->vel[j * 3];
        double dvel[3];
        for (int c = 0; c < 3; c++) dvel[c] = veli[c] - velj[c];
        const double dvdotr = dot3(dvel, dr);
        const double rhoj = s->rho[j];
        const double rho1j = 1. / rhoj;
        const double rho21j = rho1j * rho1j;
        const double rhoav1 = 0.5 * (rho1i + rho1j);
        double dustfracj = 0., rhogasj = rhoj, rhodustj = 0.;
        if (s->onef_dust) {
          dustfracj = s->dustfrac[j];
          rhodustj = s->rhodust[j];
          rhogasj = s->rhogas[j];
        }
        const double prj = fmax(s->pr[j] - 0., 0.);
        const double prnetj = prj - 0.;
        const double Prho2j = prj * rho21j;
        const double spsoundj = s->spsound[j];
        double Bj[3] = {0, 0, 0}, Brhoj[3] = {0, 0, 0}, dB[3] = {0, 0, 0};
        double projBi = 0., projBj = 0., projdB = 0., projBrhoi = 0., projBrhoj = 0.;
        double B2j = 0., valfven2j = 0., Brho2j = 0.;
        if (mhd) {
          for (int c = 0; c < 3; c++) {
            Bj[c] = s->Bfield[j * 3 + c];
            Brhoj[c] = Bj[c] * rho1j;
            dB[c] = Bi[c] - Bj[c];
          }
          projBi = dot3(Bi, dr);
          projBj = dot3(Bj, dr);
          projdB = dot3(dB, dr);
          projBrhoi = dot3(Brhoi, dr);
          projBrhoj = dot3(Brhoj, dr);
          B2j = dot3(Bj, Bj);
          valfven2j = B2j / s->dens[j];
          Brho2j = B2j * rho21j;
        }
        double forcej[3] = {0, 0, 0};
        /* signal velocities, :1417-1482 */
        double vsigi, vsigj, vsigB;
        if (mhd) {
          double vsig2i = spsoundi * spsoundi + valfven2i;
          double vsig2j = spsoundj * spsoundj + valfven2j;
          double ti = spsoundi * projBi, tj = spsoundj * projBj;
          double vsigproji = vsig2i * vsig2i - 4. * (ti * ti) * rho1i;
          double vsigprojj = vsig2j * vsig2j - 4. * (tj * tj) * rho1j;
          if (vsigproji < 0. || vsigprojj < 0.) return fail(s, 9, "rates: vsig det < 0");
          vsigi = sqrt(0.5 * (vsig2i + sqrt(vsigproji)));
          vsigj = sqrt(0.5 * (vsig2j + sqrt(vsigprojj)));
          if (s->iavlim[2] != 2) vsigB = norm2_3(dvel);
          else vsigB = 0.5 * (vsigi + vsigj) + fabs(dvdotr);
        } else {
          vsigi = spsoundi;
          vsigj = spsoundj;
          vsigB = 0.;
        }
        const double vsig = 0.5 * (fmax(vsigi + vsigj - s->beta * dvdotr, 0.0));
        const double vsigu = sqrt(fabs(prneti - prnetj) * rhoav1);
        double vsigdtc = fmax(fmax(vsig, 0.5 * (vsigi + vsigj + s->beta * fabs(dvdotr))), vsigB);
        double vsigav = 0.;
        {
          double dvsigdtc = 1. / vsigdtc;
          s->vsigmax = fmax(s->vsigmax, vsigdtc);
          if (vsigdtc > zero) s->dtcourant = fmin(s->dtcourant, fmin(hi * dvsigdtc, hj * dvsigdtc));
        }
        /* artificial dissipation */
        if (s->idust == 4) {
          /* artificial_dissipation_dust_diffusion, :2154-2206 */
          if (dvdotr < 0.) {
            double vsi = alphai * spsoundi - s->beta * dvdotr;
            double vsj = s->alpha[j * 3 + 0] * spsoundj - s->beta * dvdotr;
            double qi = -rhogasi * vsi * dvdotr;
            double qj = -rhogasj * vsj * dvdotr;
            double visc = 0.5 * (qi * rho21i * grkerni + qj * rho21j * grkernj);
            for (int c = 0; c < 3; c++) {
              forcei[c] = forcei[c] - pmassj * visc * dr[c];
              forcej[c] = forcej[c] + pmassi * visc * dr[c];
            }
            double du = s->uu[i] - s->uu[j];
            double cfaci = 0.5 * alphaui * rhoi * vsigu * du;
            double cfacj = 0.5 * s->alpha[j * 3 + 1] * rhoj * vsigu * du;
            double diffu = cfaci * grkerni * (rho1i * rho1i) + cfacj * grkernj * (rho1j * rho1j);
            s->dudt[i] = s->dudt[i] + 0.5 * qi * rho1i / rhogasi * pmassj * dvdotr * grkerni + pmassj * diffu / (1. - dustfraci);
            s->dudt[j] = s->dudt[j] + 0.5 * qj * rho1j / rhogasj * pmassi * dvdotr * grkernj - pmassi * diffu / (1. - dustfracj);
          }
        } else {
          /* artificial_dissipation, :1700-1894 */
          const double alphaav = 0.5 * (alphai + s->alpha[j * 3 + 0]);
          const double alphau = 0.5 * (alphaui + s->alpha[j * 3 + 1]);
          const double alphaB = 0.5 * (alphaBi + s->alpha[j * 3 + 2]);
          vsigav = fmax(fmax(alphaav, alphau), alphaB) * vsig;
          const double dpmomdotr = -dvdotr;
          const double term = vsig * rhoav1 * grkern;
          const double termv = term;
          const double termu = vsigu * rhoav1 * grkern;
          const double termB = vsigB * rhoav1 * grkern;
          if (dvdotr < 0) {
            double visc = alphaav * termv * dpmomdotr;
            for (int c = 0; c < 3; c++) {
              forcei[c] = forcei[c] - pmassj * visc * dr[c];
              forcej[c] = forcej[c] + pmassi * visc * dr[c];
            }
          }
          if (s->imhd > 0) {
            for (int c = 0; c < 3; c++) {
              double Bvisc = dB[c] * rhoav1;
              double dBdtvisc = alphaB * termB * Bvisc;
              dBevoldti[c] = dBevoldti[c] + rhoi * pmassj * dBdtvisc;
              s->dBevoldt[j * 3 + c] = s->dBevoldt[j * 3 + c] - rhoj * pmassi * dBdtvisc;
            }
          }
          if (s->iener > 0) {
            double vissv = 0.;
            if (dvdotr < 0) {
              double t = dot3(veli, dr) - dot3(velj, dr);
              vissv = -alphaav * 0.5 * (t * t);
            }
            double vissu = alphau * (s->uu[i] - s->uu[j]);
            double vissB = 0.;
            if (s->imhd > 0) vissB = -alphaB * 0.5 * (dot3(dB, dB)) * rhoav1;
            s->dudt[i] = s->dudt[i] + pmassj * (term * (vissv) + termu * vissu + termB * (vissB));
            s->dudt[j] = s->dudt[j] + pmassi * (term * (vissv) - termu * vissu + termB * (vissB));
          }
        }
        if (vsigav > zero) s->dtav = fmin(s->dtav, fmin(hi / vsigav, hj / vsigav));
        /* pressure term, :1538-1568 */
        {
          double prterm = Prho2i * grkerni + Prho2j * grkernj;
          for (int c = 0; c < 3; c++) {
            forcei[c] = forcei[c] - pmassj * prterm * dr[c];
            forcej[c] = forcej[c] + pmassi * prterm * dr[c];
          }
        }
        /* mhd_terms, :2377-2720 */
        if (mhd) {
          double fiso = 0.5 * (Brho2i * grkerni + Brho2j * grkernj);
          for (int c = 0; c < 3; c++) {
            double faniso = (Brhoi[c] * projBrhoi - stressmax * dr[c] * rho21i) * grkerni +
                            (Brhoj[c] * projBrhoj - stressmax * dr[c] * rho21j) * grkernj;
            double fmagi = faniso - fiso * dr[c];
            forcei[c] = forcei[c] + pmassj * (fmagi);
            forcej[c] = forcej[c] - pmassi * (fmagi);
          }
          s->divB[i] = s->divB[i] - pmassj * projdB * grkern;
          s->divB[j] = s->divB[j] - pmassi * projdB * grkern;
          if (s->imhd > 0) {
            double cb[3];
            cb[0] = dB[1] * dr[2] - dB[2] * dr[1];
            cb[1] = dB[2] * dr[0] - dB[0] * dr[2];
            cb[2] = dB[0] * dr[1] - dB[1] * dr[0];
            for (int c = 0; c < 3; c++) {
              s->curlB[i * 3 + c] = s->curlB[i * 3 + c] + pmassj * cb[c] * grkern;
              s->curlB[j * 3 + c] = s->curlB[j * 3 + c] + pmassi * cb[c] * grkern;
            }
          }
          /* induction, imhd = 11 (surface flux-conservative form) */
          for (int c = 0; c < 3; c++) {
            dBevoldti[c] = dBevoldti[c] - pmassj * ((dvel[c] * projBrhoi) * grkerni);
            s->dBevoldt[j * 3 + c] = s->dBevoldt[j * 3 + c] - pmassi * ((dvel[c] * projBrhoj) * grkernj);
          }
          if (s->idivbzero >= 2) {
            double gradpsiterm = s->psi[i] * rho21i * grkerni + s->psi[j] * rho21j * grkernj;
            for (int c = 0; c < 3; c++) {
              s->gradpsi[i * 3 + c] = s->gradpsi[i * 3 + c] - pmassj * gradpsiterm * dr[c];
              s->gradpsi[j * 3 + c] = s->gradpsi[j * 3 + c] + pmassi * gradpsiterm * dr[c];
            }
          }
        }
        /* dust_derivs_diffusion (idustevol 0, constant t_s), :2813-2878 */
        if (s->idust == 4) {
          double tstopi = s->Kdrag, tstopj = s->Kdrag;
          double Di = rhodusti * rho1i * tstopi;
          double Dj = rhodustj * rho1j * tstopj;
          double Dav = 0.5 * (Di + Dj);
          double grgrkern = -2. * grkern / rij;
          double diffterm = rho1i * rho1j * Dav * (pri - prj) * grgrkern;
          s->ddustevoldt[i] = s->ddustevoldt[i] + pmassj * diffterm;
          s->ddustevoldt[j] = s->ddustevoldt[j] - pmassi * diffterm;
          if (s->iener > 0) {
            double du = s->uu[i] - s->uu[j];
            double pdvtermi = pri * rho1i / rhogasi * pmassj * dvdotr * grkerni;
            double pdvtermj = prj * rho1j / rhogasj * pmassi * dvdotr * grkernj;
            s->dudt[i] = s->dudt[i] + pdvtermi - 0.5 * pmassj * du * diffterm / (1. - dustfraci);
            s->dudt[j] = s->dudt[j] + pdvtermj - 0.5 * pmassi * du * diffterm / (1. - dustfracj);
          }
        }
        for (int c = 0; c < 3; c++) s->force[j * 3 + c] = s->force[j * 3 + c] + forcej[c];
        /* switch source, :1639-1656 */
        if (s->iavlim[1] > 0) {
          double graduterm = (s->uu[i] - s->uu[j]) / rij;
          s->del2u[i] = s->del2u[i] + pmassj * rho1j * graduterm * grkerni;
          s->del2u[j] = s->del2u[j] - pmassi * rho1i * graduterm * grkernj;
        }
        if (s->iavlim[0] == 3) {
          for (int c = 0; c < 3; c++) {
            s->graddivv[i * 3 + c] = s->graddivv[i * 3 + c] + pmassj * rho1j / rij * dvdotr * grkerni * dr[c];
            s->graddivv[j * 3 + c] = s->graddivv[j * 3 + c] - pmassi * rho1i / rij * dvdotr * grkernj * dr[c];
          }
        } else {
          /* the "curl v" sum, :1653-1654 */
          for (int c = 0; c < 3; c++) {
            s->graddivv[i * 3 + c] = s->graddivv[i * 3 + c] + pmassj * (dvel[c] - dvdotr) * grkerni;
            s->graddivv[j * 3 + c] = s->graddivv[j * 3 + c] + pmassi * (-dvel[c] - dvdotr) * grkernj;
          }
        }
      }
      for (int c = 0; c < 3; c++) {
        s->force[i * 3 + c] = s->force[i * 3 + c] + forcei[c];
        s->dBevoldt[i * 3 + c] = s->dBevoldt[i * 3 + c] + dBevoldti[c];
      }
      i = s->ll[i];
    }
  }

  /* per-particle finalisation, :482-943 */
  if (mhd && s->idivbzero >= 2) s->vsig2max = s->vsigmax * s->vsigmax;
  double fhmax = 0.0;
  s->dtdrag = HUGED;
  s->dtforce = HUGED;
  for (int i = 0; i < npart; i++) {
    const double rhoi = s->rho[i];
    const double rho1i = 1. / rhoi;
    if (s->idust == 4) {
      double rhodusti = s->rhodust[i];
      double rhogasi = s->rhogas[i];
      double dustfraci = rhodusti * rho1i;
      double tstop = s->Kdrag;
      for (int c = 0; c < 3; c++)
        s->deltav[i * 3 + c] = (rhogasi > TINYD) ? -rhoi / rhogasi * s->force[i * 3 + c] * tstop : 0.;
      s->dtdrag = fmin(s->dtdrag, 0.5 * (s->hh[i] * s->hh[i]) /
                                     (dustfraci * tstop * (s->spsound[i] * s->spsound[i])));
      if (s->dtdrag < 0.) s->dtdrag = fabs(s->dtdrag);
    }
    if (mhd) {
      if (s->imhd > 0)
        for (int c = 0; c < 3; c++) s->curlB[i * 3 + c] = s->curlB[i * 3 + c] * rho1i;
      s->divB[i] = s->divB[i] * rho1i;
    }
    double forcemag = sqrt(dot3(&s->force[i * 3], &s->force[i * 3]));
    double fonh = forcemag / s->hh[i];
    if (fonh > fhmax && s->itype[i] != 1) fhmax = fonh;
    double valfven2i = 0.;
    if (mhd) valfven2i = dot3(&s->Bfield[i * 3], &s->Bfield[i * 3]) / s->dens[i];
    double vsig2 = s->spsound[i] * s->spsound[i] + valfven2i;
    double vsig = sqrt(vsig2);
    if (s->imhd == 11) {
      for (int c = 0; c < 3; c++)
        s->dBevoldt[i * 3 + c] = 1.0 * s->dBevoldt[i * 3 + c] + s->Bevol[i * 3 + c] * rho1i * s->drhodt[i];
      if (s->idivbzero >= 2) {
        for (int c = 0; c < 3; c++) {
          s->gradpsi[i * 3 + c] = s->gradpsi[i * 3 + c] * rhoi;
          s->dBevoldt[i * 3 + c] = s->dBevoldt[i * 3 + c] + s->gradpsi[i * 3 + c];
        }
      }
    }
    if (s->iener > 0 && s->idust != 4) {
      s->dudt[i] = s->dudt[i] + s->pr[i] * (rho1i * rho1i) * s->drhodt[i];
      s->dendt[i] = s->dudt[i];
    } else {
      s->dendt[i] = s->dudt[i];
    }
    for (int c = 0; c < 3; c++) s->daldt[i * 3 + c] = 0.;
    if (s->iavlim[0] || s->iavlim[1] || s->iavlim[2]) {
      double tdecay1 = (s->avdecayconst * vsig) / s->hh[i];
      const double *al = &s->alpha[i * 3];
      if (s->iavlim[0] == 1 || s->iavlim[0] == 2) {
        double source = fmax(s->drhodt[i] * rho1i, 0.0);
        if (s->iavlim[0] == 2) source = source * (2.0 - al[0]);
        s->daldt[i * 3 + 0] = (s->alphamin - al[0]) * tdecay1 + s->avfact * source;
      }
      if (s->iener > 0 && s->iavlim[1] > 0) {
        double sourceu = 0.;
        if (s->uu[i] > EPSD) sourceu = s->hh[i] * fabs(s->del2u[i]) / sqrt(s->uu[i]);
        s->daldt[i * 3 + 1] = (s->alphaumin - al[1]) * tdecay1 + sourceu;
      }
      if (s->iavlim[2] != 0 && mhd) {
        double sourceJ = sqrt(dot3(&s->curlB[i * 3], &s->curlB[i * 3]) * rho1i);
        double sourcedivB = 10. * fabs(s->divB[i]) * sqrt(rho1i);
        double sourceB = fmax(sourceJ, sourcedivB);
        if (s->iavlim[2] == 3) {
          double source = fmax(s->drhodt[i] * rho1i, 0.0) * (2. - al[2]);
          sourceB = sqrt(source * sourceB);
        }
        s->daldt[i * 3 + 2] = (s->alphabmin - al[2]) * tdecay1 + sourceB;
      }
    }
    if (s->idivbzero >= 2 && s->idivbzero <= 7)
      s->dpsidt[i] = -s->vsig2max * s->divB[i] - s->psidecayfact * s->psi[i] * s->vsigmax / s->hh[i];
    else
      s->dpsidt[i] = 0.;
  }
  if (fhmax > 0.) s->dtforce = fmin(s->dtforce, sqrt(1. / fhmax));
  /* rates zero on ghosts and fixed particles, ratesND_mhd.f90:949-966 */
  for (int i = 0; i < ntotal; i++) {
    if (i < npart && s->itype[i] != ITYPEBND && s->itype[i] != ITYPEBNDDUST) continue;
    for (int c = 0; c < 3; c++) {
      s->force[i * 3 + c] = 0.; s->dBevoldt[i * 3 + c] = 0.; s->daldt[i * 3 + c] = 0.;
      s->gradpsi[i * 3 + c] = 0.; s->curlB[i * 3 + c] = 0.;
    }
    s->drhodt[i] = 0.; s->dhdt[i] = 0.; s->dudt[i] = 0.; s->dendt[i] = 0.;
    s->dpsidt[i] = 0.; s->divB[i] = 0.;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* derivs, step                                                              */
/* ------------------------------------------------------------------------ */

/* derivs, derivs.f90:27-287 */
int orc_derivs(orc_sim *s) {
  int e;
  int anyb = 0, anyghost = 0;
  for (int d = 0; d < s->ndim; d++) {
    if (s->ibound[d] != 0) anyb = 1;
    if (s->ibound[d] >= 2) anyghost = 1;
  }
  s->msg[0] = 0;
  if (anyb) boundary(s);
  if (anyghost) {
    if (set_ghost_particles(s)) return fail(s, 7, "ghost allocation failed");
  } else {
    s->ntotal = s->npart;
  }
  if ((e = set_linklist(s))) return e;
  if ((e = iterate_density(s))) return e;
  if ((e = cons2prim(s))) return e;
  if ((e = get_rates(s))) return e;
  return 0;
}

/* predictor half of step, stepND_leapfrog_mhd.f90:70-165: starting values
 * kept, state advanced by dt with the starting rates */
static void predict(orc_sim *s) {
  const int npart = s->npart, ndim = s->ndim;
  const double dt = s->dt;
  for (int i = 0; i < npart; i++) {
    for (int c = 0; c < 3; c++) {
      s->xin[i * 3 + c] = s->x[i * 3 + c];
      s->velin[i * 3 + c] = s->vel[i * 3 + c];
      s->Bevolin[i * 3 + c] = s->Bevol[i * 3 + c];
      s->alphain[i * 3 + c] = s->alpha[i * 3 + c];
      s->forcein[i * 3 + c] = s->force[i * 3 + c];
      s->dBevoldtin[i * 3 + c] = s->dBevoldt[i * 3 + c];
      s->daldtin[i * 3 + c] = s->daldt[i * 3 + c];
    }
    s->rhoin[i] = s->rho[i];
    s->hhin[i] = s->hh[i];
    s->enin[i] = s->en[i];
    s->psiin[i] = s->psi[i];
    s->drhodtin[i] = s->drhodt[i];
    s->dhdtin[i] = s->dhdt[i];
    s->dendtin[i] = s->dendt[i];
    s->dpsidtin[i] = s->dpsidt[i];
    if (s->onef_dust) {
      s->dustevolin[i] = s->dustevol[i];
      s->ddustevoldtin[i] = s->ddustevoldt[i];
    }
  }
  /* predictor, :107-165; fixed particles (:110-143) move with the velocity
     and force of the particle they copy and hold everything else */
  for (int i = 0; i < npart; i++) {
    if (s->itype[i] == ITYPEBND || s->itype[i] == ITYPEBNDDUST) {
      const int j = s->ireal[i] >= 0 ? s->ireal[i] : i;
      for (int d = 0; d < ndim; d++)
        s->x[i * 3 + d] = s->xin[i * 3 + d] + dt * s->velin[j * 3 + d] + 0.5 * dt * dt * s->forcein[j * 3 + d];
      for (int c = 0; c < 3; c++) s->Bevol[i * 3 + c] = s->Bevolin[i * 3 + c];
      s->rho[i] = s->rhoin[i];
      s->hh[i] = s->hhin[i];
      s->en[i] = s->enin[i];
      for (int c = 0; c < 3; c++) s->alpha[i * 3 + c] = s->alphain[i * 3 + c];
      s->psi[i] = s->psiin[i];
      if (s->onef_dust) s->dustevol[i] = s->dustevolin[i];
      continue;
    }
    for (int d = 0; d < ndim; d++)
      s->x[i * 3 + d] = s->xin[i * 3 + d] + dt * s->velin[i * 3 + d] + 0.5 * dt * dt * s->forcein[i * 3 + d];
    for (int c = 0; c < 3; c++)
      s->vel[i * 3 + c] = (s->velin[i * 3 + c] + dt * s->forcein[i * 3 + c]) / (1. + s->damp);
    if (s->imhd != 0)
      for (int c = 0; c < 3; c++)
        s->Bevol[i * 3 + c] = s->Bevolin[i * 3 + c] + dt * s->dBevoldtin[i * 3 + c];
    s->hh[i] = s->hhin[i] + dt * s->dhdtin[i];
    if (s->iener != 0) s->en[i] = s->enin[i] + dt * s->dendtin[i];
    for (int c = 0; c < 3; c++)
      if (s->iavlim[c] != 0)
        s->alpha[i * 3 + c] = fmin(s->alphain[i * 3 + c] + dt * s->daldtin[i * 3 + c], 1.0);
    if (s->idivbzero >= 2) s->psi[i] = s->psiin[i] + dt * s->dpsidtin[i];
    if (s->onef_dust) s->dustevol[i] = s->dustevolin[i] + dt * s->ddustevoldtin[i];
  }
}

int orc_predict(orc_sim *s) {
  predict(s);
  return 0;
}

/* step, stepND_leapfrog_mhd.f90:31-275 (damp = 0) */
int orc_step(orc_sim *s) {
  const int npart = s->npart, ndim = s->ndim;
  const double dt = s->dt, hdt = 0.5 * dt;
  predict(s);
  int e = orc_derivs(s);
  if (e) return e;
  /* corrector, :169-218; fixed particles (:172-186) get back their own
     starting velocity and held values */
  for (int i = 0; i < npart; i++) {
    if (s->itype[i] == ITYPEBND || s->itype[i] == ITYPEBNDDUST) {
      for (int c = 0; c < 3; c++) {
        s->vel[i * 3 + c] = s->velin[i * 3 + c];
        s->Bevol[i * 3 + c] = s->Bevolin[i * 3 + c];
        s->alpha[i * 3 + c] = s->alphain[i * 3 + c];
      }
      s->rho[i] = s->rhoin[i];
      s->hh[i] = s->hhin[i];
      s->en[i] = s->enin[i];
      s->psi[i] = s->psiin[i];
      if (s->onef_dust) s->dustevol[i] = s->dustevolin[i];
      continue;
    }
    for (int c = 0; c < 3; c++)
      s->vel[i * 3 + c] = (s->velin[i * 3 + c] + hdt * (s->force[i * 3 + c] + s->forcein[i * 3 + c])) / (1. + s->damp);
    if (s->imhd != 0)
      for (int c = 0; c < 3; c++)
        s->Bevol[i * 3 + c] = s->Bevolin[i * 3 + c] + hdt * (s->dBevoldt[i * 3 + c] + s->dBevoldtin[i * 3 + c]);
    s->hh[i] = s->hhin[i] + hdt * (s->dhdt[i] + s->dhdtin[i]);
    if (s->hh[i] <= 0.) return fail(s, 1, "step: hh -ve");
    if (s->iener != 0) s->en[i] = s->enin[i] + hdt * (s->dendt[i] + s->dendtin[i]);
    for (int c = 0; c < 3; c++)
      if (s->iavlim[c] != 0)
        s->alpha[i * 3 + c] = fmin(s->alphain[i * 3 + c] + hdt * (s->daldt[i * 3 + c] + s->daldtin[i * 3 + c]), 1.0);
    if (s->idivbzero >= 2) s->psi[i] = s->psiin[i] + hdt * (s->dpsidt[i] + s->dpsidtin[i]);
    if (s->onef_dust)
      s->dustevol[i] = s->dustevolin[i] + hdt * (s->ddustevoldt[i] + s->ddustevoldtin[i]);
  }
  int anyb = 0;
  for (int d = 0; d < ndim; d++) if (s->ibound[d] != 0) anyb = 1;
  if (anyb) boundary(s);
  /* new timestep, :252 */
  s->dt = fmin(fmin(fmin(s->C_force * s->dtforce, s->C_cour * s->dtcourant), 0.9 * s->dtdrag),
               s->C_force * s->dtvisc);
  s->time += dt;
  s->nsteps++;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* public API                                                                */
/* ------------------------------------------------------------------------ */

orc_sim *orc_new(const int32_t *iopt, const double *dopt, int32_t npart, char *err) {
  orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
#define REJECT(m) do { if (err) snprintf(err, 256, "%s", m); orc_free(s); return NULL; } while (0)
  s->ikernel = iopt[0]; s->iener = iopt[1]; s->imhd = iopt[2]; s->imagforce = iopt[3];
  s->idivbzero = iopt[4]; s->iav = iopt[5];
  for (int k = 0; k < 3; k++) { s->iavlim[k] = iopt[6 + k]; s->ibound[k] = iopt[11 + k]; }
  s->idust = iopt[9]; s->idrag_nature = iopt[10]; s->maxdensits = iopt[14];
  s->idustevol = iopt[15]; s->ndim = iopt[16];
  s->hfact = dopt[0]; s->tolh = dopt[1]; s->gamma = dopt[2]; s->polyk = dopt[3];
  s->Kdrag = dopt[4]; s->alphamin = dopt[5]; s->alphaumin = dopt[6]; s->alphabmin = dopt[7];
  s->beta = dopt[8]; s->avdecayconst = dopt[9]; s->psidecayfact = dopt[10];
  s->C_cour = dopt[11]; s->C_force = dopt[12];
  for (int k = 0; k < 3; k++) { s->xmin[k] = dopt[13 + k]; s->xmax[k] = dopt[16 + k]; s->Bconst[k] = dopt[20 + k]; }
  s->psep = dopt[19]; s->dt = dopt[23]; s->damp = dopt[24];
  if (s->ndim < 1 || s->ndim > 3) REJECT("ndim must be 1, 2 or 3");
  if (npart < 1) REJECT("npart must be positive");
  if (s->ikernel != 0 && s->ikernel != 3) REJECT("only ikernel 0 (cubic) and 3 (quintic)");
  if (s->iener != 0 && s->iener != 2) REJECT("only iener 0 and 2");
  if (s->imhd != 0 && s->imhd != 11) REJECT("only imhd 0 and 11");
  if (s->imhd && s->imagforce != 2) REJECT("only imagforce 2");
  if (s->idivbzero != 0 && s->idivbzero != 2) REJECT("only idivbzero 0 and 2");
  if (s->iav != 2) REJECT("only iav 2");
  if (s->iavlim[0] < 0 || s->iavlim[0] > 2) REJECT("iavlim(1) must be 0, 1 or 2");
  if (s->iavlim[1] < 0 || s->iavlim[1] > 1) REJECT("iavlim(2) must be 0 or 1");
  if (!(s->iavlim[2] == 0 || s->iavlim[2] == 1 || s->iavlim[2] == 3)) REJECT("iavlim(3) must be 0, 1 or 3");
  if (s->idust != 0 && s->idust != 4) REJECT("only idust 0 and 4");
  if (s->idust == 4 && (s->idrag_nature != 2 || s->idustevol != 0)) REJECT("dust: only idrag_nature 2, idustevol 0");
  if (s->idust == 4 && s->imhd) REJECT("dust with MHD not supported");
  if (s->damp != 0.) REJECT("damp must be 0");
  int nrefl = 0;
  for (int d = 0; d < s->ndim; d++) {
    int b = s->ibound[d];
    if (!(b == 0 || b == 1 || b == 2 || b == 3)) REJECT("ibound must be 0, 1, 2 or 3");
    if (b == 2) nrefl++;
  }
  if (nrefl > 1) REJECT("reflecting walls in at most one dimension");
  s->onef_dust = (s->idust == 1 || s->idust == 3 || s->idust == 4);
  s->dndim = 1. / s->ndim;
  if (fabs(s->gamma - 1.) > 1.e-3) s->avfact = log(4.) / (log((s->gamma + 1.) / (s->gamma - 1.)));
  else s->avfact = 1.0;
  setkerntable(s->ikernel, s->ndim, s->wij, s->grwij, s->grgrwij, &s->radkern, &s->dq2table);
  s->radkern2 = s->radkern * s->radkern;
  s->ddq2table = 1. / s->dq2table;
  s->npart = npart;
  s->ntotal = npart;
  if (grow(s, npart + 16)) REJECT("out of memory");
  double **inarr[] = {&s->xin, &s->velin, &s->Bevolin, &s->alphain, &s->forcein,
                      &s->dBevoldtin, &s->daldtin};
  for (unsigned k = 0; k < sizeof inarr / sizeof inarr[0]; k++)
    *inarr[k] = (double *)calloc((size_t)npart * 3, sizeof(double));
  double **in1[] = {&s->rhoin, &s->hhin, &s->enin, &s->psiin, &s->drhodtin, &s->dhdtin,
                    &s->dendtin, &s->dpsidtin, &s->dustevolin, &s->ddustevoldtin};
  for (unsigned k = 0; k < sizeof in1 / sizeof in1[0]; k++)
    *in1[k] = (double *)calloc((size_t)npart, sizeof(double));
  s->redolist = (int *)calloc((size_t)npart, sizeof(int));
  s->redoprev = (int *)calloc((size_t)npart, sizeof(int));
  s->vsig2max = 0.;
  return s;
#undef REJECT
}

void orc_free(orc_sim *s) {
  if (!s) return;
  for (int k = 0; k < NDUNIQUE; k++) {
    int w;
    free(*dfield(s, k, &w));
  }
  free(s->itype); free(s->ireal); free(s->numneigh); free(s->ll); free(s->iamincell);
  free(s->nits); free(s->listneigh); free(s->ifirstincell);
  free(s->xin); free(s->velin); free(s->Bevolin); free(s->alphain); free(s->forcein);
  free(s->dBevoldtin); free(s->daldtin); free(s->rhoin); free(s->hhin); free(s->enin);
  free(s->psiin); free(s->drhodtin); free(s->dhdtin); free(s->dendtin); free(s->dpsidtin);
  free(s->dustevolin); free(s->ddustevoldtin); free(s->redolist); free(s->redoprev);
  free(s);
}

int orc_set_state(orc_sim *s, const double *x, const double *vel, const double *pmass,
                  const double *hh, const double *en, const double *Bevol,
                  const double *alpha, const double *psi, const double *dustevol,
                  const int32_t *itype) {
  const int n = s->npart, ndim = s->ndim;
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) s->x[i * 3 + d] = d < ndim ? x[i * ndim + d] : 0.;
    for (int c = 0; c < 3; c++) {
      s->vel[i * 3 + c] = vel[i * 3 + c];
      s->Bevol[i * 3 + c] = s->imhd ? Bevol[i * 3 + c] : 0.;
      s->Bfield[i * 3 + c] = s->Bevol[i * 3 + c];
      s->alpha[i * 3 + c] = alpha[i * 3 + c];
    }
    s->pmass[i] = pmass[i];
    s->hh[i] = hh[i];
    s->en[i] = en[i];
    s->uu[i] = en[i];
    s->psi[i] = psi[i];
    s->dustevol[i] = s->onef_dust ? dustevol[i] : 0.;
    s->dustfrac[i] = s->dustevol[i];
    s->itype[i] = itype[i];
    s->ireal[i] = -1;
    s->rho[i] = 0.;
  }
  return 0;
}

/* the particle each fixed particle copies (0-based, -1 none), as
   set_fixedbound.f90:32-48 sets ireal */
int orc_set_ireal(orc_sim *s, const int32_t *ireal) {
  for (int i = 0; i < s->npart; i++) {
    if (ireal[i] >= s->npart || (ireal[i] >= 0 && s->itype[ireal[i]] != ITYPEGAS)) return fail(s, 8, "bad ireal");
    s->ireal[i] = ireal[i];
  }
  return 0;
}

int orc_get(const orc_sim *s, const char *name, double *out) {
  const int n = s->npart;
  struct { const char *nm; const double *p; int w; } t[] = {
      {"vel", s->vel, 3}, {"pmass", s->pmass, 1}, {"hh", s->hh, 1}, {"en", s->en, 1},
      {"Bevol", s->Bevol, 3}, {"alpha", s->alpha, 3}, {"psi", s->psi, 1},
      {"dustevol", s->dustevol, 1}, {"rho", s->rho, 1}, {"gradh", s->gradh, 1},
      {"gradgradh", s->gradgradh, 1}, {"drhodt", s->drhodt, 1}, {"dhdt", s->dhdt, 1},
      {"dens", s->dens, 1}, {"pr", s->pr, 1}, {"spsound", s->spsound, 1}, {"uu", s->uu, 1},
      {"Bfield", s->Bfield, 3}, {"force", s->force, 3}, {"dudt", s->dudt, 1},
      {"dendt", s->dendt, 1}, {"dBevoldt", s->dBevoldt, 3}, {"daldt", s->daldt, 3},
      {"dpsidt", s->dpsidt, 1}, {"divB", s->divB, 1}, {"curlB", s->curlB, 3},
      {"gradpsi", s->gradpsi, 3}, {"dustfrac", s->dustfrac, 1}, {"rhogas", s->rhogas, 1},
      {"rhodust", s->rhodust, 1}, {"ddustevoldt", s->ddustevoldt, 1},
      {"deltav", s->deltav, 3}, {"del2u", s->del2u, 1}, {"graddivv", s->graddivv, 3}};
  if (!strcmp(name, "x")) {
    for (int i = 0; i < n; i++)
      for (int d = 0; d < s->ndim; d++) out[i * s->ndim + d] = s->x[i * 3 + d];
    return s->ndim;
  }
  if (!strcmp(name, "numneigh") || !strcmp(name, "nits") || !strcmp(name, "itype")) {
    const int *p = !strcmp(name, "numneigh") ? s->numneigh : !strcmp(name, "nits") ? s->nits : s->itype;
    for (int i = 0; i < n; i++) out[i] = p[i];
    return 1;
  }
  for (unsigned k = 0; k < sizeof t / sizeof t[0]; k++) {
    if (strcmp(name, t[k].nm)) continue;
    memcpy(out, t[k].p, sizeof(double) * (size_t)n * t[k].w);
    return t[k].w;
  }
  return -1;
}

int orc_set(orc_sim *s, const char *name, const double *in) {
  const int n = s->npart;
  struct { const char *nm; double *p; int w; } t[] = {
      {"vel", s->vel, 3}, {"hh", s->hh, 1}, {"en", s->en, 1}, {"Bevol", s->Bevol, 3},
      {"alpha", s->alpha, 3}, {"psi", s->psi, 1}, {"dustevol", s->dustevol, 1},
      {"rho", s->rho, 1}, {"drhodt", s->drhodt, 1}, {"dhdt", s->dhdt, 1},
      {"force", s->force, 3}, {"dendt", s->dendt, 1}, {"dBevoldt", s->dBevoldt, 3},
      {"daldt", s->daldt, 3}, {"dpsidt", s->dpsidt, 1}, {"ddustevoldt", s->ddustevoldt, 1}};
  if (!strcmp(name, "x")) {
    for (int i = 0; i < n; i++)
      for (int d = 0; d < s->ndim; d++) s->x[i * 3 + d] = in[i * s->ndim + d];
    return s->ndim;
  }
  for (unsigned k = 0; k < sizeof t / sizeof t[0]; k++) {
    if (strcmp(name, t[k].nm)) continue;
    memcpy(t[k].p, in, sizeof(double) * (size_t)n * t[k].w);
    return t[k].w;
  }
  return -1;
}

double orc_scalar(const orc_sim *s, const char *name) {
  if (!strcmp(name, "dt")) return s->dt;
  if (!strcmp(name, "dtcourant")) return s->dtcourant;
  if (!strcmp(name, "dtforce")) return s->dtforce;
  if (!strcmp(name, "dtdrag")) return s->dtdrag;
  if (!strcmp(name, "dtav")) return s->dtav;
  if (!strcmp(name, "vsig2max")) return s->vsig2max;
  if (!strcmp(name, "vsigmax")) return s->vsigmax;
  if (!strcmp(name, "hhmax")) return s->hhmax;
  if (!strcmp(name, "itsdensity")) return s->itsdensity;
  if (!strcmp(name, "ntotal")) return s->ntotal;
  if (!strcmp(name, "stressmax")) return s->stressmax;
  if (!strcmp(name, "time")) return s->time;
  return NAN;
}

int orc_set_scalar(orc_sim *s, const char *name, double v) {
  if (!strcmp(name, "dt")) { s->dt = v; return 0; }
  if (!strcmp(name, "time")) { s->time = v; return 0; }
  if (!strcmp(name, "maxdensits")) { s->maxdensits = (int)v; return 0; }
  if (!strcmp(name, "tolh")) { s->tolh = v; return 0; }
  return -1;
}

int64_t orc_neighbours(orc_sim *s, int sym, int32_t *pairs, int64_t cap) {
  const int ndim = s->ndim;
  int64_t np = 0;
  int *list = s->listneigh;
  for (int i = 0; i < s->npart; i++) {
    int nneigh = get_neighbour_list_partial(s, s->iamincell[i], list);
    double hi1 = 1. / s->hh[i];
    double hi21 = hi1 * hi1;
    for (int n = 0; n < nneigh; n++) {
      int j = list[n];
      double rij2 = 0.;
      for (int d = 0; d < ndim; d++) {
        double dx = s->x[i * 3 + d] - s->x[j * 3 + d];
        rij2 = rij2 + dx * dx;
      }
      double q2i = rij2 * hi21;
      double hj1 = 1. / s->hh[j];
      double q2j = rij2 * hj1 * hj1;
      int in = sym ? (q2i < s->radkern2 || q2j < s->radkern2) : (q2i < s->radkern2);
      if (!in) continue;
      int ip = j, code = 0;
      if (j >= s->npart) {
        ip = s->ireal[j];
        for (int k = ndim - 1; k >= 0; k--) {
          code = 3 * code;
          if (s->x[j * 3 + k] > s->x[ip * 3 + k]) code += 1;
          if (s->x[j * 3 + k] < s->x[ip * 3 + k]) code += 2;
        }
      }
      if (np < cap) {
        pairs[np * 3 + 0] = i;
        pairs[np * 3 + 1] = ip;
        pairs[np * 3 + 2] = code;
      }
      np++;
    }
  }
  return np <= cap ? np : -np;
}

int orc_conserved(const orc_sim *s, double *out) {
  double ekin = 0, etherm = 0, emag = 0, mom[3] = {0, 0, 0}, ang[3] = {0, 0, 0};
  double rmin = HUGED, rmax = -HUGED, rmean = 0, mgas = 0, mdust = 0;
  for (int i = 0; i < s->npart; i++) {
    double m = s->pmass[i];
    const double *v = &s->vel[i * 3];
    const double *x = &s->x[i * 3];
    for (int c = 0; c < 3; c++) mom[c] = mom[c] + m * v[c];
    if (s->ndim == 3) {
      ang[0] = ang[0] + m * (x[1] * v[2] - x[2] * v[1]);
      ang[1] = ang[1] + m * (x[2] * v[0] - x[0] * v[2]);
      ang[2] = ang[2] + m * (x[0] * v[1] - x[1] * v[0]);
    } else if (s->ndim == 2) {
      ang[2] = ang[2] + m * (x[0] * v[1] - x[1] * v[0]);
    }
    ekin = ekin + 0.5 * m * dot3(v, v);
    if (s->onef_dust) {
      double eps = s->dustfrac[i], dterm = 1. - eps;
      ekin = ekin + 0.5 * m * eps * dterm * dot3(&s->deltav[i * 3], &s->deltav[i * 3]);
      etherm = etherm + m * s->uu[i] * dterm;
      mgas = mgas + m * dterm;
      mdust = mdust + m * eps;
    } else {
      etherm = etherm + m * s->uu[i];
    }
    if (s->imhd) emag = emag + 0.5 * m * dot3(&s->Bfield[i * 3], &s->Bfield[i * 3]) / s->rho[i];
    rmean = rmean + s->rho[i];
    rmin = fmin(rmin, s->rho[i]);
    rmax = fmax(rmax, s->rho[i]);
  }
  out[0] = ekin; out[1] = etherm; out[2] = emag; out[3] = ekin + emag + etherm;
  out[4] = mom[0]; out[5] = mom[1]; out[6] = mom[2]; out[7] = sqrt(dot3(mom, mom));
  out[8] = ang[0]; out[9] = ang[1]; out[10] = ang[2]; out[11] = sqrt(dot3(ang, ang));
  out[12] = rmin; out[13] = rmax; out[14] = rmean / s->npart;
  out[15] = mgas; out[16] = mdust;
  return 0;
}

int orc_kernel_table(int ikernel, int ndim, double *w, double *grw, double *grgrw,
                     double *radkern) {
  double dq2;
  return setkerntable(ikernel, ndim, w, grw, grgrw, radkern, &dq2);
}

void orc_interpolate(int ikernel, int ndim, double q2, double *w, double *grw, double *grgrw) {
  static orc_sim t;
  static int tk = -1, tn = -1;
  if (tk != ikernel || tn != ndim) {
    setkerntable(ikernel, ndim, t.wij, t.grwij, t.grgrwij, &t.radkern, &t.dq2table);
    t.ddq2table = 1. / t.dq2table;
    tk = ikernel;
    tn = ndim;
  }
  kinterp(&t, q2, w, grw, grgrw);
}

const char *orc_message(const orc_sim *s) { return s->msg; }
