/*
 * skid_oracle.h -- CPU oracle for the SKID group-finding path (TEST INFRASTRUCTURE).
 *
 * This is the checker, never the product: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  It restates, single-threaded
 * and in the reference's FP32 arithmetic, the stages of SKID v1.4.1 between
 * kdScatterActive and kdTooSmall (main.c:374-471), plus brute-force and FP64
 * checkers.  Each function cites the reference file:line it follows.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - every pipeline stage is compared bit-for-bit with the real reference
 *     (oracle/_ref/libskidref.so built from /root/reference) on config-1 boxes
 *     and dark.std-like inputs;
 *   - or_knn_brute / or_fof_brute are O(N^2) brute force with the same fp32
 *     d2 recipe, checked against closed forms on lattices;
 *   - the FP64 density / gradient / potential sums are checked against closed
 *     forms (isolated particle, lattice, Plummer/uniform sphere energies).
 */
#ifndef SKID_ORACLE_H
#define SKID_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

#define OR_DARK 1
#define OR_GAS  2
#define OR_STAR 4
#define OR_PLUMMER 1
#define OR_SPLINE  2

typedef struct or_params {
	int nSmooth, nBucket;
	float fPeriod;          /* box length (all axes); ignored unless bPeriodic */
	float fCenter[3];
	int bPeriodic;
	float fTau, fCvg, fScoop;   /* fCvg/fScoop <= 0: defaults tau/2 and 2*tau (main.c:343-344) */
	float fDensMin, fTempMax, fMassMax;
	int nMembers, nMaxMembers;  /* nMaxMembers: -maxgroup, groups of n >= it are not unbound */
	int bGasAndDark, bGasOnly, bForceInitialCut, bNoUnbind;
	int iSoftType;
	float G, Omega0, Lambda, fQuintess, H0, z;
	float fEps;                 /* > 0: override softening (main.c:464) */
	int bOutDens;               /* mimic -den's Order() before kdInitMove */
	int bQuiet;                 /* suppress the reference's stdout lines */
} or_params;

typedef struct or_ctx or_ctx;

void or_default_params(or_params *P);
or_ctx *or_create(const or_params *P);
void or_destroy(or_ctx *c);
/* Copy particles in (tipsy species order: gas, dark, star); iOrder = input index. */
int or_load(or_ctx *c, int nGas, int nDark, int nStar, const float *pos, const float *vel,
		const float *mass, const float *soft, const float *temp);

/* Stages (main.c call order). Outputs are indexed by iOrder; NULL skips. */
int or_density(or_ctx *c, float *fBall2, float *fDensity, int *nExtraScat);
int or_init_move(or_ctx *c, int *nMove);
int or_move_first(or_ctx *c, int *nActive, int *nScatter);      /* main.c:395-404 */
int or_move_block(or_ctx *c, int *nActive, int *nScatter);      /* main.c:410-416 */
/* The reference's own density gradient a of the initial iteration (smooth1.c:408-518
   called at main.c:396-401), by iOrder; replaces or_init_move + or_move_first. */
int or_first_gradient(or_ctx *c, float *acc);
int or_fof(or_ctx *c, int *iGroup, int *nGroup);                /* main.c:425 */
int or_microsteps(or_ctx *c, int *nScatter);   /* main.c:431-438; nScatter: 5 counts or NULL */
int or_get_moved(or_ctx *c, float *r);                           /* 3N by iOrder, NaN if not moving */
int or_centers(or_ctx *c);                                       /* main.c:452-453 */
int or_unbind(or_ctx *c, int *nUnbound);                         /* main.c:459-469 */
int or_too_small(or_ctx *c);                                     /* main.c:471 */
int or_get_groups(or_ctx *c, int *iGroup, int *nGroup);
/* Per-group catalogue (index 0..nGroup-1): mass, rCenter[3], vcm[3], rBound[3], radius. */
int or_get_catalogue(or_ctx *c, float *mass, float *center, float *vcm, float *rbound,
		float *radius, int *nMembers);
/* Whole run, main.c:374-471; trace receives (nActive, nScatter) per Ittr line. */
int or_run(or_ctx *c, int *trace, int maxTrace, int *nTrace);

/* ---- standalone stage checkers on arbitrary inputs ---- */
/* FOF of a point set (kd.c:802-917 on a fresh tree in the given order). */
int or_fof_points(int n, const float *pos, float fTau, float fPeriod, int bPeriodic,
		int nBucket, int *iGroup);
/* Brute force: k nearest by (fp32 d2 with query-side image shift, index). */
int or_knn_brute(int n, const float *pos, int k, float fPeriod, int bPeriodic,
		float *fBall2, int *nbr /* n*k or NULL, sorted by (d2, index) */);
/* Brute force FOF components (link iff d2 < tau^2 from either side), labels =
   1 + rank of the component's smallest index. */
int or_fof_brute(int n, const float *pos, float fTau, float fPeriod, int bPeriodic, int *iGroup);
/* FP64 M4 density from given radii: rho_i = sum_{j: d2_ij < b_i} m_j W_i/2 + sum_{j: d2_ji < b_j} m_j W_j/2. */
int or_density_f64(int n, const float *pos, const float *mass, const float *fBall2,
		float fPeriod, int bPeriodic, double *rho);
/* FP64 frozen-field gradient a(x) = sum_s d m_s g(q)/(pi h^5) over scatterer images
   whose ball holds x (the quantity smAccDensity accumulates). */
int or_gradient_f64(int ns, const float *spos, const float *smass, const float *sball2,
		int nq, const float *qpos, float fPeriod, int bPeriodic, double *acc);
/* FP64 softened potential phi_i = sum_j G m_j K(r_ij; eps_i+eps_j) (grav.c:8-36). */
int or_potential_f64(int n, const float *pos, const float *mass, const float *soft,
		float G, int iSoftType, double *phi);
/* The reference's own fp32/f64-mixed spline kernel (grav.h:11-31) for one pair. */
float or_spline_pot(float d2, float twoh);

#ifdef __cplusplus
}
#endif
#endif
