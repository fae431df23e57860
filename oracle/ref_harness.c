/*
 * ref_harness.c -- in-process driver for the REAL SKID stage functions.
 *
 * Test infrastructure only.  This file is compiled (by oracle/ref.mk) together
 * with the reference sources where they lie (kd.c, smooth1.c, grav.c, cosmo.c,
 * romberg.c, runge.c) into oracle/_ref/libskidref.so.  It replaces only the
 * parts of main.c that read files and print timings: it loads particles from
 * caller arrays (the same field mapping as kdReadTipsy, kd.c:122-222) and then
 * calls the stage functions in main.c's order (main.c:374-471), copying out the
 * intermediate state that SKID never writes to disk (fBall2, moved positions,
 * the FOF labels before unbinding, the Ittr trace).
 *
 * It exists so that the oracle's restatement (oracle/skid_oracle.c) can be
 * checked against the reference's own arithmetic on identical inputs.  Nothing
 * in the product links it.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#include <float.h>
#include "kd.h"
#include "smooth1.h"

#define REF_PRUNE_STEPS 5     /* main.c:13 */
#define REF_MICRO_STEP  0.1   /* main.c:14 */

typedef struct ref_params {
	int nSmooth, nBucket;
	float fPeriod, fCenter[3];
	int bPeriodic;
	float fTau, fCvg, fScoop;   /* fCvg/fScoop <= 0 -> main.c:343-344 defaults */
	float fDensMin, fTempMax, fMassMax;
	int nMembers, nMaxMembers;  /* nMaxMembers as -maxgroup (main.c:110 default INT_MAX) */
	int bGasAndDark, bGasOnly, bForceInitialCut, bNoUnbind;
	int iSoftType;
	float G, Omega0, Lambda, fQuintess, H0, z;
	float fEps;                 /* <= 0 -> keep particle softening (main.c:464) */
	int bOutDens;               /* mimic -den: Order() before kdInitMove (main.c:383-387) */
	} REF_PARAMS;

/* Outputs, all indexed by iOrder (input order). Any pointer may be NULL. */
typedef struct ref_out {
	float *fBall2;      /* after smDensityInit; -2 for non-scatterers */
	float *fDensity;    /* after smDensityInit (the .den values) */
	float *rMoved;      /* 3*N: positions after the move loop (before FOF); NaN if not moving */
	float *rMicro;      /* 3*N: positions after the micro-steps; NaN if not moving */
	int *iFofGroup;     /* kdFoF labels (before unbinding) */
	int *iGroup;        /* final labels after kdUnbind + kdTooSmall */
	int nExtraScat;
	int nMove;
	int nFofGroup;      /* kd->nGroup after kdFoF (groups + 1) */
	int nGroup;         /* final kd->nGroup */
	int nIttr;          /* number of Ittr lines printed */
	int *trace;         /* 2*maxTrace: (nActive, nScatter) per Ittr line */
	int maxTrace;
	float *grpMass;     /* per final group (index 0..nGroup-1) */
	float *grpCenter;   /* 3*nGroup: rCenter */
	float *grpVcm;      /* 3*nGroup */
	int maxGroup;
	int micro[5];       /* nScatter of each Microstep line (main.c:436) */
	const char *gtpPath;   /* if set: kdWriteGroup writes the .gtp there (native byte order) */
	} REF_OUT;

static void ref_load(KD kd, int nGas, int nDark, int nStar, const float *pos,
		const float *vel, const float *mass, const float *soft, const float *temp)
{
	int i, j;
	PINIT *p;
	kd->inType = 0;
	kd->nDark = nDark; if (nDark) kd->inType |= DARK;
	kd->nGas = nGas;   if (nGas) kd->inType |= GAS;
	kd->nStar = nStar; if (nStar) kd->inType |= STAR;
	kd->fTime = 1.0;
	kd->nParticles = nGas + nDark + nStar;
	kd->nInitActive = kd->nParticles;
	kd->pInit = (PINIT *)malloc(kd->nParticles*sizeof(PINIT));
	p = kd->pInit;
	for (i = 0; i < kd->nParticles; ++i) {
		p[i].iOrder = i;
		p[i].fDensity = 0.0;
		p[i].fBall2 = -2.0f;
		p[i].fMass = mass[i];
		p[i].fSoft = soft[i];
		p[i].fTemp = (i < nGas && temp) ? temp[i] : 0.0f;
		for (j = 0; j < 3; ++j) {
			p[i].r[j] = pos[3*i+j];
			p[i].v[j] = vel ? vel[3*i+j] : 0.0f;
			}
		}
	}

static void ref_copy_init(KD kd, float *fBall2, float *fDensity)
{
	int i;
	for (i = 0; i < kd->nParticles; ++i) {
		int io = kd->pInit[i].iOrder;
		if (fBall2) fBall2[io] = (i < kd->nInitActive) ? kd->pInit[i].fBall2 : -2.0f;
		if (fDensity) fDensity[io] = kd->pInit[i].fDensity;
		}
	}

static void ref_copy_move(KD kd, float *r)
{
	int i, j;
	if (!r) return;
	for (i = 0; i < 3*kd->nParticles; ++i) r[i] = __builtin_nanf("");
	for (i = 0; i < kd->nMove; ++i)
		for (j = 0; j < 3; ++j) r[3*kd->pMove[i].iOrder+j] = kd->pMove[i].r[j];
	}

/* Density only: main.c:374-379. */
int ref_density(const REF_PARAMS *P, int nGas, int nDark, int nStar,
		const float *pos, const float *mass, const float *soft, const float *temp,
		float *fBall2, float *fDensity, int *nExtraScat)
{
	KD kd;
	SMX smx;
	float fPeriod[3], fCenter[3];
	int j;
	for (j = 0; j < 3; ++j) { fPeriod[j] = P->bPeriodic ? P->fPeriod : FLT_MAX; fCenter[j] = P->fCenter[j]; }
	kdInit(&kd, P->nBucket, fPeriod, fCenter, 0);
	ref_load(kd, nGas, nDark, nStar, pos, NULL, mass, soft, temp);
	kdScatterActive(kd, P->bGasAndDark, P->bGasOnly);
	kdBuildTree(kd);
	smInit(&smx, kd, P->nSmooth);
	smDensityInit(smx, P->bPeriodic);
	ref_copy_init(kd, fBall2, fDensity);
	if (nExtraScat) *nExtraScat = smx->nExtraScat;
	smFinish(smx);
	kdFinish(kd);
	return 0;
	}

/* Full pipeline: main.c:374-471 (without file output). */
int ref_pipeline(const REF_PARAMS *P, int nGas, int nDark, int nStar,
		const float *pos, const float *vel, const float *mass, const float *soft,
		const float *temp, REF_OUT *o)
{
	KD kd;
	SMX smx;
	float fPeriod[3], fCenter[3], fCvg, fScoop, fStep;
	int i, j, nActive, nScat, nIttr, nt = 0;
	for (j = 0; j < 3; ++j) { fPeriod[j] = P->bPeriodic ? P->fPeriod : FLT_MAX; fCenter[j] = P->fCenter[j]; }
	fCvg = (P->fCvg > 0) ? P->fCvg : 0.5*P->fTau;
	fScoop = (P->fScoop > 0) ? P->fScoop : 2.0*P->fTau;
	fStep = 0.5*fCvg;
	kdInit(&kd, P->nBucket, fPeriod, fCenter, 0);
	ref_load(kd, nGas, nDark, nStar, pos, vel, mass, soft, temp);
	kdScatterActive(kd, P->bGasAndDark, P->bGasOnly);
	kdBuildTree(kd);
	smInit(&smx, kd, P->nSmooth);
	smDensityInit(smx, P->bPeriodic);
	ref_copy_init(kd, o->fBall2, o->fDensity);
	o->nExtraScat = smx->nExtraScat;
	if (P->bOutDens) {
		/* kdOutDensity() calls Order() (kd.c:1541); that reorder is what -den changes */
		extern void Order(KD);
		Order(kd);
		}
	nActive = kdInitMove(kd, P->fDensMin, P->fTempMax, P->fMassMax, fCvg,
			P->bGasAndDark, P->bGasOnly);
	o->nMove = kd->nMove;
	kdBuildMoveTree(kd);
	if (bAllowInitialCut(kd) || P->bForceInitialCut) nScat = smAccDensity(smx, 1);
	else nScat = smAccDensity(smx, 0);
	if (o->trace && nt < o->maxTrace) { o->trace[2*nt] = nActive; o->trace[2*nt+1] = nScat; }
	++nt;
	kdMoveParticles(kd, fStep);
	nIttr = 1;
	while (nActive) {
		for (i = 0; i < REF_PRUNE_STEPS; ++i) {
			kdBuildMoveTree(kd);
			nScat = smAccDensity(smx, 0);
			kdMoveParticles(kd, fStep);
			}
		nActive = kdPruneInactive(kd, fCvg);
		if (o->trace && nt < o->maxTrace) { o->trace[2*nt] = nActive; o->trace[2*nt+1] = nScat; }
		++nt;
		++nIttr;
		}
	o->nIttr = nt;
	ref_copy_move(kd, o->rMoved);
	kdFoF(kd, P->fTau);
	o->nFofGroup = kd->nGroup;
	if (o->iFofGroup) memcpy(o->iFofGroup, kd->piGroup, kd->nParticles*sizeof(int));
	kdReactivateMove(kd);
	for (i = 0; i < REF_PRUNE_STEPS; ++i) {
		kdBuildMoveTree(kd);
		nScat = smAccDensity(smx, 0);
		kdMoveParticles(kd, REF_MICRO_STEP*fStep);
		o->micro[i] = nScat;
		}
	ref_copy_move(kd, o->rMicro);
	smFinish(smx);
	kdInitpGroup(kd);
	kdCalcCenter(kd);
	kdSetUniverse(kd, P->G, P->Omega0, P->Lambda, P->fQuintess, P->H0, P->z);
	if (P->fEps > 0) kdSetSoft(kd, P->fEps);
	kdUnbind(kd, P->iSoftType, fScoop, P->bGasAndDark, P->bNoUnbind, P->nMaxMembers);
	kdTooSmall(kd, P->nMembers);
	o->nGroup = kd->nGroup;
	if (o->gtpPath) kdWriteGroup(kd, (char *)o->gtpPath, 0);
	if (o->iGroup) memcpy(o->iGroup, kd->piGroup, kd->nParticles*sizeof(int));
	for (i = 0; i < kd->nGroup && i < o->maxGroup; ++i) {
		if (o->grpMass) o->grpMass[i] = kd->pGroup[i].fMass;
		for (j = 0; j < 3; ++j) {
			if (o->grpCenter) o->grpCenter[3*i+j] = kd->pGroup[i].rCenter[j];
			if (o->grpVcm) o->grpVcm[3*i+j] = kd->pGroup[i].vcm[j];
			}
		}
	kdFinish(kd);
	return 0;
	}

/* The -unbind re-entry: main.c:349-373 then 457-471, with the group ids from
   the caller instead of a .grp file and no .gtp (so kdReadCenter takes the
   centres of mass, kd.c:1160-1193).  grpOut receives the final ids by iOrder. */
int ref_unbind(const REF_PARAMS *P, int nGas, int nDark, int nStar,
		const float *pos, const float *vel, const float *mass, const float *soft,
		const float *temp, const int *grpIn, int *grpOut, int *nGroupOut)
{
	KD kd;
	float fPeriod[3], fCenter[3], fScoop;
	int i, j, mx = 0;
	for (j = 0; j < 3; ++j) { fPeriod[j] = P->bPeriodic ? P->fPeriod : FLT_MAX; fCenter[j] = P->fCenter[j]; }
	fScoop = (P->fScoop > 0) ? P->fScoop : 2.0*P->fTau;
	kdInit(&kd, P->nBucket, fPeriod, fCenter, 0);
	ref_load(kd, nGas, nDark, nStar, pos, vel, mass, soft, temp);
	/* kdInGroup (kd.c:920-962) from memory */
	kd->piGroup = (int *)malloc(kd->nParticles*sizeof(int));
	for (i = 0; i < kd->nParticles; ++i) {
		kd->piGroup[i] = grpIn[i];
		if (grpIn[i] > mx) mx = grpIn[i];
		}
	kd->nGroup = mx + 1;
	kdInitpGroup(kd);
	kdReadCenter(kd, "", 0);
	kdSetUniverse(kd, P->G, P->Omega0, P->Lambda, P->fQuintess, P->H0, P->z);
	if (P->fEps > 0) kdSetSoft(kd, P->fEps);
	kdUnbind(kd, P->iSoftType, fScoop, P->bGasAndDark, P->bNoUnbind, P->nMaxMembers);
	kdTooSmall(kd, P->nMembers);
	memcpy(grpOut, kd->piGroup, kd->nParticles*sizeof(int));
	if (nGroupOut) *nGroupOut = kd->nGroup;
	kdFinish(kd);
	return 0;
	}
