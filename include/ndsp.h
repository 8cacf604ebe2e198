/*
 * ndsp.h -- C-ABI of the MI355X hot-path library (libndsp_hip.so).
 *
 * The reference (ndspmhd, serial Fortran) has no plugin or FFI layer: its hot
 * path is a set of argument-less subroutines that share module globals and
 * are swapped at link time (src/Makefile:227,263).  This header is the
 * boundary a Fortran host binds with iso_c_binding (see
 * ndspmhd_amd/fortran/ndsp_host.f90): each entry point below replaces one
 * reference routine, cited beside it.  Plain pointers and sizes only; every
 * entry returns an ndsp_status (0 = ok), the text of the last error is
 * ndsp_last_error().  One context = one simulation resident in HBM of the
 * current HIP device; a context is not thread-safe.
 *
 * Option vectors (same slots as the reference harness and the oracle):
 *   iopt[32]: 0 ikernel, 1 iener, 2 imhd, 3 imagforce, 4 idivbzero, 5 iav,
 *             6-8 iavlim(1:3), 9 idust, 10 idrag_nature, 11-13 ibound(1:3),
 *             14 maxdensits, 15 idustevol, 16 ndim,
 *             17 cellscan (device-only: 1 = scan the cell grid in every
 *                pass instead of per-particle neighbour lists)
 *   dopt[32]: 0 hfact, 1 tolh, 2 gamma, 3 polyk, 4 Kdrag, 5 alphamin,
 *             6 alphaumin, 7 alphaBmin, 8 beta, 9 avdecayconst,
 *             10 psidecayfact, 11 C_cour, 12 C_force, 13-15 xmin(1:3),
 *             16-18 xmax(1:3), 19 psep, 20-22 Bconst(1:3), 23 dt, 24 damp
 *   (src/variablesND.f90:33-300 module globals; defaults src/defaults.f90)
 * Particle arrays use the reference's column-major layout
 * (x(ndim,N) -> x[i*ndim+d], vel(3,N) -> vel[i*3+d]; src/variablesND.f90:161-188).
 */
#ifndef NDSP_H
#define NDSP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ndsp_ctx ndsp_ctx;

typedef enum {
  NDSP_OK = 0,
  NDSP_ERR_H_NONPOSITIVE = 1,    /* h <= 0 (src/iterate_density.f90:99-102) */
  NDSP_ERR_RHO_NONPOSITIVE = 2,  /* rho <= 0 (src/iterate_density.f90:168-170) */
  NDSP_ERR_NOT_CONVERGED = 3,    /* h-rho iteration (src/iterate_density.f90:349-351) */
  NDSP_ERR_EOS = 4,              /* u < 0 (src/eos.f90:64-72) */
  NDSP_ERR_CELLS = 5,            /* cell grid impossible (src/linkND.f90:73-101) */
  NDSP_ERR_OUTSIDE = 6,          /* particle outside domain (src/linkND.f90:122-125) */
  NDSP_ERR_ALLOC = 7,
  NDSP_ERR_COINCIDENT = 8,       /* dx = 0 pair (src/ratesND_mhd.f90:384-395) */
  NDSP_ERR_VSIG = 9,             /* vsig determinant < 0 (src/ratesND_mhd.f90:1422-1429) */
  NDSP_ERR_BAD_OPTION = 10,      /* option combination not implemented */
  NDSP_ERR_BAD_ARG = 11,
  NDSP_ERR_HIP = 12,             /* HIP runtime error */
  NDSP_ERR_NO_DEVICE = 13
} ndsp_status;

/* arithmetic of the pair sums */
enum {
  NDSP_PREC_MIXED = 0, /* FP64 positions/neighbour tests/h-rho update/integrator/
                          reductions, FP32 pair math with compensated sums */
  NDSP_PREC_FP64 = 1   /* everything FP64, plain sums */
};

typedef struct {
  int32_t itsdensity;   /* max Newton-Raphson passes over particles (itsdensity) */
  int32_t nmin, nmax;   /* min/max neighbour counts (numneigh) */
  int32_t pad;
  double ncalc_mean;    /* mean passes per particle (ncalctotal/npart) */
  int64_t pairs_density;/* ordered gather pairs summed over all passes */
  int64_t pairs_force;  /* unordered symmetric pairs in the force pass */
} ndsp_stats;

typedef struct {
  double dt;            /* next timestep (src/stepND_leapfrog_mhd.f90:252) */
  double dtcourant, dtforce, dtdrag, vsig2max, hhmax, stressmax, time;
} ndsp_dt;

/* Context with capacity for npart particles; replaces the module-global
 * set-up of initialise (src/initialiseND_mhd.f90:140-300: kernel tables via
 * setkernels src/kernelND.f90:102, avfact :166-172) and alloc
 * (src/allocateND.f90:34).  precision: NDSP_PREC_*. */
int ndsp_create(const int32_t *iopt, const double *dopt, int64_t npart,
                int32_t precision, ndsp_ctx **out);
int ndsp_destroy(ndsp_ctx *ctx);

/* Host -> HBM copy of the evolved state (reference layout); itype may be
 * NULL (all gas); itype 1 marks fixed boundary particles (ibound 1 walls,
 * partners then given by ndsp_set_ireal).  Replaces nothing in the
 * reference (its state lives in
 * module arrays); the Fortran host calls it after setup and after any host
 * edit of the arrays. */
int ndsp_upload(ndsp_ctx *ctx, const double *x, const double *vel,
                const double *pmass, const double *hh, const double *en,
                const double *Bevol, const double *alpha, const double *psi,
                const double *dustevol, const int32_t *itype);

/* Partners of the fixed boundary particles (src/set_fixedbound.f90:32-48
 * sets ireal; used by iterate_density.f90:310-326, conservative2primitive
 * .f90:425-432 and stepND_leapfrog_mhd.f90:110-186): ireal[i] is the
 * 0-based index (caller order) of the gas particle that fixed particle i
 * copies, -1 for particles that are not fixed.  Required after an upload
 * with itype 1 particles (walls with ibound 1) before the first binning. */
int ndsp_set_ireal(ndsp_ctx *ctx, const int32_t *ireal);

/* HBM -> host copy of one per-particle field of the real particles, in the
 * caller's particle order and reference layout.  Names as the reference's
 * arrays: x vel pmass hh en Bevol alpha psi dustevol rho gradh gradgradh
 * drhodt dhdt dens pr spsound uu Bfield force dudt dendt dBevoldt daldt
 * dpsidt divB curlB gradpsi dustfrac rhogas rhodust ddustevoldt deltav
 * del2u graddivv (graddivv of hydro and dust runs only, the runs whose dumps
 * hold it; zero for MHD) numneigh nits, and slot (the particle's current
 * position in device memory, which periodic reorders change).  Bfield and
 * uu are B and en at the last evaluation (after a step, the predictor's, as
 * the reference's .ev file and dumps report them; a fixed particle's are
 * its partner's), Bevol and en the evolved values.  *width receives 1, 3 or
 * ndim. */
int ndsp_download(ndsp_ctx *ctx, const char *name, double *out, int32_t *width);

/* Replaces set_linklist (src/linkND.f90:45) together with the ghost set-up
 * it relies on (src/ghostND_mhd.f90:33, src/boundaryND.f90:30): periodic
 * wrap / reflection, then binning into cells of side >= radkern*hmax by a
 * counting sort on cell keys. */
int ndsp_set_linklist(ndsp_ctx *ctx);

/* Replaces iterate_density (src/iterate_density.f90:43) with density /
 * density_partial (src/density_sums.f90:38,396): per-particle h-rho
 * Newton-Raphson to tolh, then conservative2primitive
 * (src/conservative2primitive.f90:42) and the equation of state
 * (src/eos.f90:40).  Needs ndsp_set_linklist first. */
int ndsp_iterate_density(ndsp_ctx *ctx, ndsp_stats *stats);

/* Replaces get_rates (src/ratesND_mhd.f90:29): fused pair sums and
 * per-particle finalisation; fills the timestep criteria. */
int ndsp_get_rates(ndsp_ctx *ctx, ndsp_dt *dt);

/* Replaces derivs (src/derivs.f90:27): set_linklist, iterate_density,
 * get_rates in order. */
int ndsp_derivs(ndsp_ctx *ctx, ndsp_stats *stats, ndsp_dt *dt);

/* Replaces step (src/stepND_leapfrog_mhd.f90:31): predictor, derivs,
 * corrector, boundary, new dt, all on the device.  Uses the context's dt
 * (dopt[23] at creation, then the value the previous step chose). */
int ndsp_step(ndsp_ctx *ctx, ndsp_stats *stats, ndsp_dt *dt);

/* Sets the timestep and time (evolve.f90:52-100 keeps both in module
 * timestep). */
int ndsp_set_dt(ndsp_ctx *ctx, double dt, double time);

/* FP64 conserved quantities of the real particles (src/evwrite_mhd.f90:124-248):
 * out[17] = ekin etherm emag etot momx momy momz momtot angx angy angz angtot
 * rhomin rhomax rhomean totmassgas totmassdust. */
int ndsp_conserved(ndsp_ctx *ctx, double *out);

/* The numbers of one .ev line, as evwrite computes them
 * (src/evwrite_mhd.f90:27; loop over the real particles :124-245, columns
 * :247-308), reduced on the device in FP64.  cols (25 doubles) receives, in
 * the order evwrite writes them: t ekin etherm emag epot etot momtot angtot
 * rhomax rhomean dt, then for MHD emagp crosshel betamhdmin betamhdav divBav
 * divBmax divBtot fdotBav fdotBmax force_err_av force_err_max omegamhdav
 * omegamhdmax fracdivBok (25 columns), for one-fluid dust ekiny totmassgas
 * totmassdust (14), otherwise ekiny dmomtot (13); *ncols receives the count.
 * t and dt are the caller's, as evwrite takes its argument t and the module
 * dt.  epot is 0 (no gravity or external forces on this path); the fdotB and
 * force_err columns are 0 because the tensor-form Lorentz force
 * (imagforce 2) leaves fmag zero (src/ratesND_mhd.f90:202,2621-2626). */
int ndsp_ev_columns(ndsp_ctx *ctx, double t, double dt, double *cols, int32_t *ncols);

/* Neighbour sets at the current h (test support; the reference's
 * equivalent is the brute-force check_neighbourlist,
 * src/check_neighbourlist.f90:27): gather (q2_i < R^2) if sym=0, symmetric
 * (q2_i < R^2 or q2_j < R^2) if sym=1.  pairs[3*k..] = (i, j, image code),
 * caller's 0-based indices; code digits per dimension (x least significant,
 * base 3): 0 none, 1 image at larger coordinate, 2 at smaller.  Returns
 * NDSP_ERR_BAD_ARG with *npairs = needed count when cap is too small. */
int ndsp_neighbours(ndsp_ctx *ctx, int32_t sym, int32_t *pairs, int64_t cap,
                    int64_t *npairs);

/* Average duration (ms) of each kernel of the last ndsp_step / ndsp_derivs,
 * measured with HIP events on the context's stream; names[] receives
 * comma-separated kernel labels.  Timing is off unless enabled. */
int ndsp_set_timing(ndsp_ctx *ctx, int32_t on);
int ndsp_kernel_times(ndsp_ctx *ctx, double *ms, int32_t cap, int32_t *n,
                      char *names, int32_t names_cap);

/* Autotuner (no counterpart in the reference): times `reps` derivs
 * evaluations for each tile shape -- workgroup size of the force and h-rho
 * kernels (64 or 256) and cells per search radius of the binning (1 or 2) --
 * and keeps the fastest; choice[3] (may be NULL) receives them.  The evolved
 * state is restored afterwards, so results do not change. */
int ndsp_autotune(ndsp_ctx *ctx, int32_t reps, int32_t *choice);

/* Sets the tile shapes the autotuner chooses among (no counterpart in the
 * reference): force and h-rho workgroup sizes (64 or 256) and cells per
 * search radius (1 or 2), e.g. a choice kept from an earlier autotune. */
int ndsp_set_tile_shapes(ndsp_ctx *ctx, int32_t force_wg, int32_t density_wg,
                         int32_t cells_per_radius);

/* Test support (no counterpart in the reference): copies nbytes between two
 * fresh device buffers with the library's plain copy kernel (k_copy_d), then
 * reads the largest power-of-two number of 64-byte lines that fits in nbytes
 * once each, in a scattered order (k_gather_probe, 4 bytes written per
 * line): references of known traffic for calibrating hardware memory
 * counters on streaming and on gather-shaped access. */
int ndsp_copy_probe(ndsp_ctx *ctx, int64_t nbytes);

/* Step graphs (no counterpart in the reference, whose step is a host loop):
 * with on != 0, ndsp_step records its two halves (predictor + binning +
 * first h-rho passes; rates + corrector + next dt) into HIP graphs once and
 * replays them, completing the h-rho iteration on the host path in between
 * when it needs more passes than recorded.  Results are identical to the
 * launch-by-launch path.  Not used while timing or the h predictor is on. */
int ndsp_set_graph(ndsp_ctx *ctx, int32_t on);

/* Interval of the periodic reorder (no counterpart in the reference, which
 * keeps its particles unsorted): every `every` steps (0 = never; default 10)
 * the evolved state is permuted into the cell order of the last binning, so
 * that per-particle reads and writes of the pair kernels stay coalesced.
 * Results differ from another interval only in summation order.  The first
 * reorder of a context outside a batch allocates a second set of the
 * persistent fields (62 doubles per particle) that the state then moves
 * between; when that allocation fails the state is copied back instead. */
int ndsp_set_reorder(ndsp_ctx *ctx, int32_t every);

/* ---- batches of independent simulations (a parameter sweep; the reference
 * runs them one after another from multi/ sweep descriptions).  A batch
 * references existing contexts (it does not own them); ndsp_batch_step takes
 * one leapfrog step of every member, exactly as ndsp_step would, replaying
 * both halves of all members as two recorded HIP graphs in which each phase
 * is one launch over all members.  stats and dt receive one entry per member (may be
 * NULL).  Members may not time kernels or use the h predictor. */
typedef struct ndsp_batch ndsp_batch;
int ndsp_batch_create(ndsp_ctx *const *ctxs, int32_t nsim, ndsp_batch **out);
int ndsp_batch_destroy(ndsp_batch *batch);
int ndsp_batch_step(ndsp_batch *batch, ndsp_stats *stats, ndsp_dt *dt);
/* Tile shapes of a batch (no counterpart in the reference): the workgroup
 * size of the merged pair kernels (64 or 256) and the cells per search radius
 * of every member (1 or 2), each timed over reps batch steps; every member's
 * state is restored after each trial.  choice[0] = workgroup size,
 * choice[1] = cells per radius. */
int ndsp_batch_autotune(ndsp_batch *batch, int32_t reps, int32_t *choice);

/* ---- h-predictor MLP (no counterpart in the reference; its seed for the
 * h-rho iteration is h = h_in + dt dh/dt, src/stepND_leapfrog_mhd.f90:153).
 * A 6 -> 32 -> 32 (GELU) -> 1 network, BF16 weights, FP32 accumulation,
 * predicts 1000 log(h_converged/h_predictor) from per-particle features; when
 * enabled, ndsp_step seeds the iteration with it (the iteration still
 * converges every particle to tolh) and, with train != 0, takes `iters` Adam
 * steps per step on `batch` of that step's converged (features, h) pairs.
 * Parameters: float[1313] = W1[32][6] b1[32] W2[32][32] b2[32] W3[32] b3. */
int ndsp_mlp_config(ndsp_ctx *ctx, int32_t enable, int32_t train, double lr,
                    int32_t batch, int32_t iters, int32_t seed_after_steps);
int ndsp_mlp_set_params(ndsp_ctx *ctx, const float *params, int32_t n);
int ndsp_mlp_get_params(ndsp_ctx *ctx, float *params, int32_t n, int32_t *adam_steps);
/* forward of B feature rows X[B][6] with the BF16 weights */
int ndsp_mlp_forward(ndsp_ctx *ctx, const float *X, int32_t B, float *y);
/* the same forward on the matrix cores (BF16 inputs and activations, FP32
 * accumulation), the arithmetic of the seed the steps use; no counterpart in
 * the reference */
int ndsp_mlp_infer(ndsp_ctx *ctx, const float *X, int32_t B, float *y);
/* ndsp_mlp_forward that allocates nothing: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  X and y
 * (host or device memory) are staged through a fixed device array of
 * 65536 rows per GPU, held until the process ends; B > 65536 is rejected
 * (NDSP_ERR_BAD_ARG), as is a context whose weights were never set
 * (ndsp_mlp_set_params or ndsp_mlp_config first).  Calls that overlap on one
 * GPU share that array and race. */
int ndsp_mlp_forward_noalloc(ndsp_ctx *ctx, const float *X, int32_t B, float *y);
/* gradient of sum (y - T)^2 / B over a caller batch (FP32 accumulation) */
int ndsp_mlp_grad(ndsp_ctx *ctx, const float *X, const float *T, int32_t B, float *grad,
                  double *loss);
/* iters Adam steps on a caller batch; loss of the last step */
int ndsp_mlp_train(ndsp_ctx *ctx, const float *X, const float *T, int32_t B, int32_t iters,
                   double lr, double *loss);

const char *ndsp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
