/*
 * skg.h -- C ABI of the MI355X-native SKID group-finding library (libskg.so).
 *
 * The library replaces the stage functions that SKID's main.c calls between
 * reading the tipsy snapshot and writing the group files.  Each entry point
 * below names the reference interface it replaces.  Conventions:
 *
 *   - Every call returns an int status: SKG_OK (0) or a negative SKG_E_* code
 *     (skg_strerror() describes it).  The library never aborts, exits or
 *     enables floating-point traps.
 *   - Host arrays belong to the caller; the library copies in and out.  Per-
 *     particle host arrays are indexed by iOrder (tipsy input order: gas, dark,
 *     star), as SKID's outputs are.
 *   - Device memory belongs to the context.  Calls are synchronous at return;
 *     work inside a call is stream-ordered on the context's HIP stream.
 *   - One host thread per context; separate contexts are independent.
 *
 * Plain C: pointers and sizes only.
 */
#ifndef SKG_H
#define SKG_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define SKG_OK           0
#define SKG_E_NODEVICE  -1   /* no usable GPU / HIP runtime error at init */
#define SKG_E_PARAM     -2   /* invalid parameter */
#define SKG_E_NOMEM     -3   /* device or host allocation failed */
#define SKG_E_STATE     -4   /* stage called out of order */
#define SKG_E_HIP       -5   /* HIP runtime error during a stage */
#define SKG_E_NSMOOTH   -6   /* nSmooth > number of scatterers (smooth1.c:12) */

#define SKG_DARK 1
#define SKG_GAS  2
#define SKG_STAR 4
#define SKG_PLUMMER 1
#define SKG_SPLINE  2

/* Accumulation modes of the density / gradient sums. */
#define SKG_ACC_PLAIN        0   /* fp32 running sums (the reference's arithmetic) */
#define SKG_ACC_COMPENSATED  1   /* fp32 terms summed in FP64 (default) */

typedef struct skg_params {
	int nSmooth;            /* -s, default 64 (main.c:103) */
	int nBucket;            /* leaf size, 16 in SKID (main.c:82) */
	float fPeriod[3];       /* -p; FLT_MAX when not periodic (main.c:125-128) */
	float fCenter[3];       /* -c/-cx/-cy/-cz */
	int bPeriodic;
	float fTau;             /* -tau */
	float fCvg;             /* -cvg, default tau/2 (main.c:343) */
	float fStep;            /* fCvg/2 (main.c:345) */
	float fScoop;           /* -scoop, default 2*tau (main.c:344) */
	float fDensMin, fTempMax, fMassMax;   /* -d -t -M */
	int nMembers;           /* -m */
	int nMaxMembers;        /* -maxgroup: groups of nMaxMembers or more are not unbound
	                           (kd.c:1330); default INT_MAX (main.c:110) */
	int bGasAndDark, bGasOnly, bForceInitialCut, bNoUnbind;   /* -gd -go -fic -nu */
	int iSoftType;          /* SKG_SPLINE or SKG_PLUMMER */
	float G;                /* -G */
	float z;                /* -z */
	float fShift;           /* 1/(1+z), from the host (kd.c:1317) */
	float fCosmo;           /* fShift*H(fShift), host FP64 cosmology (kd.c:1318) */
	int accumulate;         /* SKG_ACC_* */
	int device;             /* HIP device ordinal */
} skg_params;

typedef struct skg_ctx skg_ctx;

/* Fill *P with SKID's defaults (main.c:82-136); the caller then sets fTau etc. */
void skg_default_params(skg_params *P);

/* Replaces kdInit (kd.c:46-81) + smInit (smooth1.c:8-29). */
int skg_create(const skg_params *P, skg_ctx **out);
/* Replaces kdFinish (kd.c:1842-1850) + smFinish (smooth1.c:32-38). */
int skg_destroy(skg_ctx *ctx);
const char *skg_strerror(int status);
/* After SKG_E_HIP: the failing HIP call, its error string and source line, for
   the calling host thread (empty before any failure).  No reference
   counterpart: SKID's CPU code has no runtime errors of this kind. */
const char *skg_last_error(void);

/* Device memory held by the library in this process, over all contexts: bytes
   allocated now and the high-water mark since load (or since the last call with
   reset_peak nonzero, which sets the mark to the current bytes).  No reference
   counterpart: SKID keeps its arrays in host memory (kd.c malloc); this is for
   sizing runs against one GPU's HBM. */
int skg_device_memory(long long *current, long long *peak, int reset_peak);

/* Page-locked host memory (hipHostMalloc) for the host arrays the stage calls
   read and write: their copies then run as DMA at the link's rate, with no
   staging through a bounce buffer and no page faults on first touch.  Any host
   pointer works for those calls; these are for callers that repeat a pass and
   keep their output arrays.  No reference counterpart (SKID's arrays are
   malloc'd host memory, kd.c:16-28). */
int skg_host_alloc(void **p, long long bytes);
int skg_host_free(void *p);

/* Particle input in tipsy species order, n = nGas + nDark + nStar; replaces the
   array fill of kdReadTipsy (kd.c:122-222).  pos/vel are n*3 floats; vel, temp
   may be NULL. */
int skg_load(skg_ctx *ctx, int nGas, int nDark, int nStar, const float *pos,
		const float *vel, const float *mass, const float *soft, const float *temp);

/* Batched sub-volumes (no reference counterpart: the reference runs one box per
   process, so this replaces nBox separate kdReadTipsy + main.c:374-471 runs).
   Box b holds nGas[b] + nDark[b] + nStar[b] particles in tipsy species order;
   the boxes are stored one after another in pos/vel/mass/soft/temp, and iOrder
   runs over that concatenation.  All boxes share the parameter block (period,
   centre, nSmooth, ...) and one species mix (SKG_E_PARAM otherwise); fTau
   (optional, per box) gives each box its own tau and with it fCvg = tau/2,
   fStep = fCvg/2, fScoop = 2 tau (main.c:343-345).  NULL uses the parameter
   block's fTau/fCvg/fStep/fScoop for every box.  Every later stage runs all
   boxes in the same launches; no pair, link or scoop crosses a box.  Group ids
   stay numbered by smallest iOrder, so box b's groups are one contiguous range.
   skg_field_gradient is not available on a batched context. */
int skg_load_batch(skg_ctx *ctx, int nBox, const int *nGas, const int *nDark, const int *nStar,
		const float *fTau, const float *pos, const float *vel, const float *mass,
		const float *soft, const float *temp);

/* Per-box counters after the last move call (the numbers main.c prints on each
   Ittr: line, main.c:417, for every box): nActive[b] active movers, nScatter[b]
   live scatterers + replicas.  Either pointer may be NULL. */
int skg_box_counts(skg_ctx *ctx, int *nActive, int *nScatter);

/* Scatterer selection, tree build, kNN radii, symmetric M4 density and
   periodic replicas: kdScatterActive + kdBuildTree + smDensityInit
   (main.c:374-378; kd.c:669-693, 371-460; smooth1.c:132-335).  Outputs by
   iOrder (NULL to skip): fBall2 (-2 for non-scatterers), density (0 for
   non-scatterers). */
int skg_density(skg_ctx *ctx, float *fBall2, float *rho, int *nExtraScat);

/* k-nearest neighbour lists of the last skg_density, for inspection: nbr is
   n*nSmooth iOrders, each row sorted by (d2, iOrder); rows of non-scatterers
   are -1.  d2 (optional) holds the matching squared distances.  No reference
   counterpart: the reference keeps these lists only inside smDensityInit. */
int skg_neighbors(skg_ctx *ctx, int *nbr, float *d2);

/* Moving-particle selection: kdInitMove (kd.c:630-666). */
int skg_init_move(skg_ctx *ctx, int *nMove);

/* One Denmax step at iteration 0 (main.c:395-404): kdBuildMoveTree +
   smAccDensity(bInitial) + kdMoveParticles. */
int skg_move_first(skg_ctx *ctx, int *nActive, int *nScatter);
/* PRUNE_STEPS=5 steps + kdPruneInactive (main.c:410-416). */
int skg_move_block(skg_ctx *ctx, int *nActive, int *nScatter);

/* Friends-of-friends of the converged positions: kdFoF (kd.c:802-917).
   grp (by iOrder, optional) receives group ids 1..nGroup-1 numbered in order
   of each group's smallest iOrder; particles that do not move get 0. */
int skg_fof(skg_ctx *ctx, int *grp, int *nGroup);

/* The 5 micro-steps at 0.1*fStep (main.c:431-438).  nScatter (optional)
   receives 5 counts, the nScatter of each Microstep line (main.c:436): live
   scatterers + replicas after that step, summed over the boxes of a batch. */
int skg_microsteps(skg_ctx *ctx, int *nScatter);

/* Positions of the moving particles (3n by iOrder; NaN for non-moving). */
int skg_get_moved(skg_ctx *ctx, float *r);

/* kdInitpGroup + kdCalcCenter (kd.c:969-1075). */
int skg_centers(skg_ctx *ctx);

/* Groups from outside instead of the move loop + FOF: the -unbind re-entry,
   kdInGroup + kdInitpGroup + kdReadCenter (kd.c:920-1196; main.c:349-373).
   grp by iOrder (0 = no group; ids 1..max dense or sparse), after skg_load
   only.  center and vcm (3 per group 1..max, the .gtp star positions and
   velocities) or both NULL: then each group's centre is its mass-weighted
   centre relative to its member with the largest iOrder, and vcm its
   mass-weighted velocity (kd.c:1160-1193).  Batched contexts: a group must lie
   in one box.  Next calls: skg_group_energies, skg_unbind, skg_too_small. */
int skg_set_groups(skg_ctx *ctx, const int *grp, const float *center, const float *vcm);

/* Per-group energies before unbinding (no reference counterpart; the first
   pass of kdUnbind's energy test, kd.c:1376-1399): W[g] = -1/2 sum m_i phi_i
   from the fp32 compensated member potentials (kdCellPot, members only) and
   T[g] = 1/2 sum m_i |fShift (v_i - vcm) + fCosmo (r_i - rcm)|^2, both reduced
   in FP64, for g = 0..nGroup-1 (group 0 gets 0). */
int skg_group_energies(skg_ctx *ctx, double *W, double *T);

/* Iterative unbinding: kdUnbind (kd.c:1299-1466) with grav.c:8-135. */
int skg_unbind(skg_ctx *ctx, int *nUnbound);

/* kdTooSmall (kd.c:1251-1293). */
int skg_too_small(skg_ctx *ctx);
/* Per-group statistics of the .stat file (kdOutStats, kd.c:1703-1839), after
   skg_too_small: SKG_STAT_N floats per group for groups 0..nGroup-1 (row 0 is
   zero): total mass, gas mass (rho >= fDensMin, T <= fTempMax), star mass,
   Vmax^2 = max G M(<r)/r over members outside twice their softening (else the
   outer value), V(r_half)^2, V(r_out)^2, r at Vmax, r_half (first member past
   half the mass), r_out^2, 1-D dispersion of fExp (v - vcm) + fExpHub r,
   rCenter (3), vcm (3), rBound (3).  Running masses are FP64 (the reference
   sums in fp32).  The driver prints square roots where kdOutStats does. */
#define SKG_STAT_N 19
int skg_group_stats(skg_ctx *ctx, float *stats);

/* Diagnostics of the last skg_unbind(): groups unbound by the cooperative
   all-GPU kernel, cooperative launches the device refused (those groups then
   went to the one-workgroup kernel, with the same result), and full passes of
   the bounded-candidate kernels.  No reference counterpart. */
int skg_unbind_info(skg_ctx *ctx, int *nBig, int *nCoopRefused, long long *nFullPasses);

/* Group ids by iOrder and the group count (groups + 1, as kd->nGroup). */
int skg_get_groups(skg_ctx *ctx, int *grp, int *nGroup);

/* Per-group catalogue for groups 0..nGroup-1 (what kdWriteGroup writes,
   kd.c:1612-1687): mass, rCenter (3), vcm (3), rBound (3), radius, nMembers.
   Any pointer may be NULL. */
int skg_get_catalogue(skg_ctx *ctx, float *mass, float *center, float *vcm,
		float *rbound, float *radius, int *nMembers);

/* Vector-Jacobian product of the M4 density with frozen neighbour sets and
   smoothing lengths h_i = sqrt(fBall2_i)/2 (no reference counterpart: the
   reference only forms -grad of the frozen scatter field, smooth1.c:447-459).
   gbar (by iOrder) weights each density; outputs by iOrder:
     dx[3i+c] = sum_j gbar_j d rho_j / d x_i,c     (targets and neighbours)
     dh[i]    = sum_j gbar_j d rho_j / d h_i
   parts selects the gather (1), scatter (2) or both (3) halves of rho. */
int skg_density_adjoint(skg_ctx *ctx, const float *gbar, int parts, float *dx, float *dh);

/* Sensitivity of each density to nSmooth (no reference counterpart; the
   derivative of the density of smooth1.c:249-263 along its smoothing lengths):
   with the neighbour sets frozen and every h_j = sqrt(fBall2_j)/2 scaling as
   nSmooth^(1/3), drho[i] = d rho_i / d ln nSmooth (by iOrder; 0 for
   non-scatterers).  Forward product of d rho/dh along dh_j = h_j/3.  Once it
   has been called on a context, every later skg_density forms both halves in
   its own kernels and this call only combines them. */
int skg_density_dlnk(skg_ctx *ctx, float *drho);

/* Frozen-field gradient a(x) = sum_s (r_s - x) m_s g(q)/(pi h_s^5) at nq
   arbitrary points (the quantity smAccDensity accumulates, smooth1.c:447-459),
   over all scatterers and replicas of the last skg_density. */
int skg_field_gradient(skg_ctx *ctx, int nq, const float *qpos, float *acc);

/* FOF of an arbitrary point set (the link rule of kd.c:860-897 on the given
   positions), labels numbered by smallest index.  Standalone: needs only a
   context for the device and stream. */
int skg_fof_points(skg_ctx *ctx, int n, const float *pos, float fTau, int *grp, int *nGroup);

/* FOF of n points already in device memory, allocating nothing.  No
   reference counterpart: it is for callers that cannot allocate device memory
   (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  It gives skg_fof_points's result on the same points: d_pos
   holds 3n floats (x, y, z) and d_grp receives n group ids (1-based, numbered
   by smallest index), both in device memory of the context's device; *nGroup
   (host, optional) = groups + 1.  Its temporaries live in one fixed device
   array of SKG_NOALLOC_SCRATCH_BYTES per GPU, which the library holds from
   the load of its code on that GPU to the end of the process.  n above
   SKG_NOALLOC_MAX_POINTS returns SKG_E_PARAM (so does a context whose
   nBucket is below 4, when the tree no longer fits the array).  Every call on
   a GPU shares that GPU's array: calls that overlap on one GPU (contexts used
   from several host threads at once) race.  Synchronous at return. */
#define SKG_NOALLOC_MAX_POINTS 524288
#define SKG_NOALLOC_SCRATCH_BYTES (64 << 20)
int skg_fof_points_noalloc(skg_ctx *ctx, int n, const float *d_pos, float fTau, int *d_grp, int *nGroup);

/* Autotuner (no reference counterpart): times launch shapes on the loaded
   particles and keeps the fastest for this context.  what is a mask of
     SKG_TUNE_BUCKET    leaf size nBucket in {8, 16, 32, 64} (skg_density plus
                        the whole move loop per candidate);
     SKG_TUNE_MAPPING   the mover-to-lane mapping: the active count at and
                        below which a block runs 16 lanes per mover instead of
                        one lane per mover with its tiles' lists, from
                        {2^14, 2^16, 2^18, 2^20}, each timed over the whole
                        move loop;
     SKG_TUNE_POT_TILE  the unbinding potential's workgroup / LDS tile in
                        {64, 128, 256}, timed on the groups of the next
                        skg_unbind;
     SKG_TUNE_KNN       the kNN's shape: its tree's leaf size in
                        {8, 16, 32, 64} (independent of nBucket) and 1 or 2
                        wavefronts per kNN workgroup (each wavefront with its
                        own 64 queries and LDS heaps; 2 only for nSmooth <= 64),
                        timed by the kNN stage of skg_density.
   Needs loaded particles; leaves the context loaded (stage of skg_load), so
   the caller runs the path again from skg_density.  Outputs (any may be
   NULL): the chosen nBucket, switch size and tile (the tile as chosen so
   far; the next skg_unbind updates it). */
#define SKG_TUNE_BUCKET   1
#define SKG_TUNE_MAPPING  2
#define SKG_TUNE_POT_TILE 4
#define SKG_TUNE_KNN      8
#define SKG_TUNE_ALL      15
int skg_autotune(skg_ctx *ctx, int what, int *nBucket, long long *subMax, int *potTile);
/* The kNN shape in use: its tree's leaf size and wavefronts per kNN
   workgroup (set by SKG_TUNE_KNN, or SKG_KNN_BUCKET / SKG_KNN_WAVES in the
   environment at skg_create; default nBucket and 1).  No reference
   counterpart. */
int skg_knn_shape(skg_ctx *ctx, int *knnBucket, int *knnWaves);

/* Per-stage device time in ms since skg_create, indexed by SKG_T_*. */
#define SKG_T_TREE 0
#define SKG_T_KNN 1
#define SKG_T_DENSITY 2
#define SKG_T_MOVE 3
#define SKG_T_FOF 4
#define SKG_T_GROUP 5
#define SKG_T_UNBIND 6
#define SKG_T_N 8
int skg_stage_times(skg_ctx *ctx, double *ms);

/* Kernel-level timing (no reference counterpart): when enabled, each launch of
   the instrumented kernels is bracketed by HIP events on the context's stream.
   skg_kernel_stats() returns, for kernel class `which` (SKG_K_*), the summed
   device time in ms, the launch count and the summed algorithmic work of those
   launches (bytes for the search, density, move and FOF classes, FP32 flops
   for SKG_K_UNB_POT). */
#define SKG_K_KNN 0          /* 16 B per query */
#define SKG_K_COVER_DENS 1   /* 24 B per scatterer */
#define SKG_K_COVER_MOVE 2   /* 24 B per active mover + 24 B per live scatterer, per step */
#define SKG_K_FOF 3          /* 16 B per mover */
#define SKG_K_UNB_POT 4      /* 20 flops per ordered pair */
#define SKG_K_UNB_LOOP 5     /* removal loops of groups < 32768 members: 20 flops per kdSubPot pair */
#define SKG_K_UNB_BIG 6      /* removal loops of larger groups (cooperative): 20 flops per kdSubPot pair */
#define SKG_K_COVER_BUILD 7  /* a block's first move step building its entity lists: as SKG_K_COVER_MOVE */
#define SKG_K_N 8
int skg_set_profiling(skg_ctx *ctx, int on);
/* Calibration copy (no reference counterpart): allocates two device buffers of
   bytes bytes (rounded down to 16) and copies one into the other reps times
   with a plain load/store kernel (k_copy_calib), so that rocprofv3's
   FETCH_SIZE / WRITE_SIZE counters can be scaled to bytes.  Synchronous. */
int skg_device_copy(skg_ctx *ctx, long long bytes, int reps);
int skg_kernel_stats(skg_ctx *ctx, int which, double *ms, long long *launches, double *work);

#ifdef __cplusplus
}
#endif
#endif
