// Kernel-side data layout shared by ndsp.hip: field enumeration of the
// particle state held in HBM (structure of arrays, FP64) and the argument
// block every kernel receives by value.
#pragma once
#include "ndsp_dev.h"

enum Field {
  // evolved state
  F_X, F_Y, F_Z, F_VX, F_VY, F_VZ, F_BX, F_BY, F_BZ, F_M, F_H, F_EN, F_PSI, F_EPS,
  F_A0, F_A1, F_A2,
  // derived by the density pass
  F_RHO, F_GRADH, F_GRADGRADH, F_DRHODT, F_DHDT, F_DENS, F_PR, F_CS, F_UU,
  F_DUSTFRAC, F_RHOGAS, F_RHODUST,
  // rates
  F_FX, F_FY, F_FZ, F_DUDT, F_DENDT, F_DBX, F_DBY, F_DBZ, F_DA0, F_DA1, F_DA2,
  F_DPSIDT, F_DIVB, F_CBX, F_CBY, F_CBZ, F_GPX, F_GPY, F_GPZ, F_DEPS, F_DVX,
  F_DVY, F_DVZ, F_DEL2U,
  F_GDVX, F_GDVY, F_GDVZ,  // graddivv of the rates (hydro and dust; output only)
  // Bfield: B at the last evaluation (a fixed particle's is its partner's),
  // which the outputs report, as uu is en at the last evaluation (output only)
  F_BFX, F_BFY, F_BFZ,
  F_MLPCORR,  // last step's log(h_converged / h_predictor), an MLP feature
  F_HEVAL,    // h of the last step's evaluation (the corrector then advances h)
  F_NPERSIST,  // fields above are carried by a reorder
  // step in-copies (src/stepND_leapfrog_mhd.f90:70-100)
  F_XIN = F_NPERSIST, F_YIN, F_ZIN, F_VXIN, F_VYIN, F_VZIN, F_BXIN, F_BYIN, F_BZIN,
  F_HIN, F_ENIN, F_A0IN, F_A1IN, F_A2IN, F_PSIIN, F_EPSIN, F_FXIN, F_FYIN, F_FZIN,
  F_DBXIN, F_DBYIN, F_DBZIN, F_DHDTIN, F_DENDTIN, F_DA0IN, F_DA1IN, F_DA2IN,
  F_DPSIDTIN, F_DEPSIN,
  F_RHOIN,  // rho at the start of a step (held by fixed particles)
  // a fixed particle's mass and dissipation parameters before the copy from
  // its partner, as its periodic and mirror images carry them
  F_MIMG, F_A0IMG, F_A1IMG, F_A2IMG,
  F_HNR0,  // h at entry of the h-rho iteration (hhin of src/iterate_density.f90:95)
  F_NFIELDS
};

// physics switches resolved once at context creation
struct Phys {
  int ndim, ikernel, iener, imhd, idivbzero, idust, iavlim[3], maxits;
  double hfact, tolh, gamma, polyk, Kdrag, alphamin, alphaumin, alphabmin, beta;
  double avdecayconst, psidecayfact, C_cour, C_force, psep, avfact, dndim;
  double radkern, radkern2, dq2table, ddq2table;
  double Bconstmax;
  double boxmin;  // smallest wall-to-wall length (search radii must stay below it)
  int use_list;   // neighbour lists (no reflecting walls)
  double skin;    // relative h margin of the lists
  int cell_sub;   // cells per search radius radkern*hmax (autotuned)
  int fixed_on;   // some wall holds fixed particles (ibound 1)
};

// device scalars (one per context); doubles touched by atomics hold
// non-negative values
struct DScal {
  double hmax_bin;    // max h at binning (src/ghostND_mhd.f90:77)
  double hmax_nr;     // max h after the h-rho iteration
  double stressmax;   // src/ratesND_mhd.f90:231-245
  double vsigmax;     // src/ratesND_mhd.f90:1475
  double dtcourant;   // src/ratesND_mhd.f90:1481
  double fhmax;       // src/ratesND_mhd.f90:814-816
  double dtdrag;      // src/ratesND_mhd.f90:606
  double dtforce;
  double dt, time;
  double hmax_cur;     // max current h (search radius of a symmetric pass)
  double hhmax_ghost;  // hhmax of the last ghost set-up (relink test, iterate_density.f90:260)
  double bbox_lo[3], bbox_hi[3];  // of the real particles (open dimensions)
  unsigned long long pairs_density, pairs_force, ncalc_total;
  int err;
  int itsmax;
  int nmin, nmax;
  int ncalc_next;      // unconverged particles after a pass
  int relink;
  int list_invalid;    // some h grew past its list's margin: force pass scans cells
  int list_overflow;   // particles whose list did not fit (they scan cells)
  int list_on;         // lists built by the last binning (images one period away)
  int grid_clamped;    // cells were widened to fit the key-space capacity
  double dt_cur;       // dt of the step in progress / the next step
  // h-rho pass control, advanced on the device by k_pass_ctl
  int pc_nlist;        // particles in the next pass
  int pc_its;          // its of the next pass (1 = first)
  int pc_sym;          // next pass is the symmetric routine
  int pc_persistent;   // next pass iterates its particles to convergence
  int pc_done;         // the iteration has finished (or failed)
  int pc_relink;       // ghost extents were rebuilt before the next pass
  int pc_refresh;      // 1/h of the packed view must be refreshed
};

template <typename PT>
struct V4 { typedef float4 type; };
template <>
struct V4<double> { typedef double4 type; };

struct KArgs {
  double* f[F_NFIELDS];
  int* numneigh;
  int* nits;
  int* uid;           // caller index of each slot
  // fixed boundary particles (ibound 1, src/set_fixedbound.f90): itype 1
  // marks a fixed slot, ireal the caller index of the particle it copies
  // (-1 none), slot_of the slot of each caller index
  int* itype;
  int* ireal;
  int* slot_of;
  // binning
  int* cellid;        // per slot
  int* cstart;        // ncell + 1
  int* cfill;         // ncell
  int* nrflag;        // unconverged after a pass, by sorted position (compacted in order; null: appended)
  int* sidx;          // sorted position -> slot
  // packed neighbour views in sorted order
  double4* P4;        // x, y, z, 1/h
  double* dxb;        // reflecting ghost extent radkern*h at ghost-creation time
  // density: one record per sorted particle (dr_stride bytes): x y z 1/h
  // in FP64, then the reals (vx vy vz m) and, with dust, (eps 0 0 0)
  void* DR;
  int dr_stride;
  // force: one record per sorted particle (fr_bytes() long): x y z 1/h in
  // FP64, then quads of reals (vx vy vz m) (rho pr cs gradh) (uu a0 a1 a2)
  // and (Bx By Bz psi) with MHD or (h dens eps rhogas) (rhodust 0 0 0) with
  // dust
  void* FR;
  void* GF;           // force, images only: m a0 a1 a2 (fixed particles: before the copy)
  // kernel tables (FP32 and FP64 copies of W, dW, d2W)
  const float* tabf;  // 3 x (IKERN+1)
  const double* tabd;
  // neighbour lists, particle-major: entry m of sorted particle k at
  // nbl[k*maxnb + m] = (sorted index << 7) | image code (kinds kx + 5 ky + 25 kz)
  unsigned* nbl;
  int* nbc;           // gather-list length, -1 when the list overflowed
  int* nbx;           // transposed extras after the gather list
  double* hcap;       // the list holds every neighbour while h <= hcap
  int maxnb;
  // per-block partials of the global reductions: quantity q of block b at
  // red[q * rcap + b] (combined into sc by k_red_final)
  double* red;
  int rcap;
  DScal* sc;
  int n;
};

// one global reduction: op | type << 4, into the DScal field at byte offset off
enum { RED_SUM = 0, RED_MAX = 1, RED_MIN = 2 };
enum { RED_F64 = 0, RED_I32 = 1, RED_U64 = 2 };
#define RED_MAXQ 8
struct RedSpec {
  int nq;
  int op[RED_MAXQ];
  int off[RED_MAXQ];
};
