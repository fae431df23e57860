/*
 * skid.c -- command-line driver: SKID's options, tipsy input and group files,
 * with every group-finding stage running on the GPU through libskg (skg.h).
 *
 * Follows main.c:44-497 of SKID v1.4.1 for option names, defaults, stage
 * order and stdout counters; tipsy reading/writing follows kd.c:122-222 and
 * kd.c:1502-1687 (native little-endian or, with -std, XDR big-endian, decoded
 * here without libtirpc).  Host cosmology H(a) stays FP64 (cosmo.c:46-58).
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "skg.h"
#include "textio.h"

#define PRUNE_STEPS 5       /* main.c:13 */

static void usage(void)
{
	fprintf(stderr, "USAGE:\n");
	fprintf(stderr, "skid -tau <fLinkLength> [OPTIONAL ARGUMENTS]\n");
	fprintf(stderr, "	 reads TIPSY BINARY input file from stdin\n");
	fprintf(stderr, "     [-std]\n");
	fprintf(stderr, "COSMOLOGY and UNITS arguments:\n");
	fprintf(stderr, "     [-z <fRedShift>] [-O <fOmega>]\n");
	fprintf(stderr, "     [-G <fGravConst>] [-H <fHubble>]\n");
	fprintf(stderr, "     [-Lambda <fLambda>] [-Q <fQuintessense]\n");
	fprintf(stderr, "GROUP FINDING arguments (see man page!):\n");
	fprintf(stderr, "     [-s <nSmooth>] [-d <fMinDensity>] [-t <fMaxTemp>]\n");
	fprintf(stderr, "     [-cvg <fConvergeRadius>] [-scoop <fScoopRadius>]\n");
	fprintf(stderr, "     [-m <nMinMembers>] [-nu] [-gd] [-unbind <GroupName>[.grp]]\n");
	fprintf(stderr, "     [-M <fMaxMass>] [-fic] [-go] [-maxgroup nMaxMembers]\n");
	fprintf(stderr, "GRAVITATIONAL SOFTENING arguments:\n");
	fprintf(stderr, "     [-spline] [-plummer] [-e <fSoft>]\n");
	fprintf(stderr, "PERIODIC BOX specification:\n");
	fprintf(stderr, "     [-p <xyzPeriod>]\n");
	fprintf(stderr, "     [-c <xyzCenter>]\n");
	fprintf(stderr, "     [-cx <xCenter>] [-cy <yCenter>] [-cz <zCenter>]\n");
	fprintf(stderr, "OUTPUT arguments:\n");
	fprintf(stderr, "     [-o <Output Name>] [-ray] [-den] [-stats] [-diag]\n");
	fprintf(stderr, "     [-plain] (fp32 running sums instead of compensated sums)\n");
	exit(1);
}

static void die_skg(const char *what, int rc)
{
	fprintf(stderr, "skid: %s failed: %s (%d)\n", what, skg_strerror(rc), rc);
	exit(2);
}

/* ---------------------------------------------------------------- tipsy I/O */
typedef TioSnap Snap;

static uint32_t bswap32(uint32_t x) { return (x >> 24) | ((x >> 8) & 0xff00u) | ((x << 8) & 0xff0000u) | (x << 24); }

static float rdf(const unsigned char *p, int xdr) { return tio_get_f32(p, xdr); }
static int rdi(const unsigned char *p, int xdr) { return tio_get_i32(p, xdr); }

static void wrf(FILE *fp, float f, int xdr)
{
	uint32_t u;
	memcpy(&u, &f, 4);
	if (xdr) u = bswap32(u);
	fwrite(&u, 4, 1, fp);
}

static void wri(FILE *fp, int v, int xdr)
{
	uint32_t u = (uint32_t)v;
	if (xdr) u = bswap32(u);
	fwrite(&u, 4, 1, fp);
}

/* ---------------------------------------------------------------- cosmology */
/* H(a) for the flat/curved Lambda + quintessence model (cosmo.c:46-58). */
static double exp2hub(double H0, double om, double lam, double q, double a)
{
	double rad = 0.0, curv = 1.0 - om - lam - rad - q;
	return H0*sqrt(om*a + curv*a*a + rad + q*a*a*sqrt(a) + lam*a*a*a*a)/(a*a);
}

static double now_s(void)
{
	struct timespec t;
	clock_gettime(CLOCK_MONOTONIC, &t);
	return t.tv_sec + 1e-9*t.tv_nsec;
}

/* kdInGroup (kd.c:920-962): an ASCII .grp, the particle count then one group
   id per particle in input order.  Parsed on nIo threads; a file that is not
   plain whitespace-separated ints goes through one fscanf per value instead, so
   its error (or its reading) is that of the serial loop. */
static int read_grp(const char *path, int n, int *grp, int *nGroup, int nIo)
{
	FILE *fp;
	int i, m = 0, mx = 0;
	long hdr = 0;
	const int rc = tio_read_ints(path, n, grp, &hdr, nIo);
	if (rc == 0) {
		for (i = 0; i < n; ++i) {
			if (grp[i] < 0) {
				fprintf(stderr, "ERROR: bad group id at line %d of %s\n", i + 2, path);
				return 1;
			}
			if (grp[i] > mx) mx = grp[i];
		}
		*nGroup = mx + 1;
		return 0;
	}
	if (rc == 1) {
		fprintf(stderr, "ERROR: Mismatched number of particles\n");
		fprintf(stderr, "Number in Group file %s: %ld\n", path, hdr);
		fprintf(stderr, "Number in TIPSY BINARY input file: %d\n", n);
		return 1;
	}
	fp = fopen(path, "r");
	if (!fp) { fprintf(stderr, "ERROR: Could not open file:%s\n", path); return 1; }
	if (fscanf(fp, "%d", &m) != 1 || m != n) {
		fprintf(stderr, "ERROR: Mismatched number of particles\n");
		fprintf(stderr, "Number in Group file %s: %d\n", path, m);
		fprintf(stderr, "Number in TIPSY BINARY input file: %d\n", n);
		fclose(fp);
		return 1;
	}
	for (i = 0; i < n; ++i) {
		if (fscanf(fp, "%d", &grp[i]) != 1 || grp[i] < 0) {
			fprintf(stderr, "ERROR: bad group id at line %d of %s\n", i + 2, path);
			fclose(fp);
			return 1;
		}
		if (grp[i] > mx) mx = grp[i];
	}
	fclose(fp);
	*nGroup = mx + 1;
	return 0;
}

/* kdReadCenter (kd.c:1083-1196): the .gtp's star records give each group's
   centre and velocity.  No file: *cen = *vcm = NULL (centre of mass). */
static int read_gtp(const char *path, int xdr, int nGroup, const Snap *S, const int *grp, float **cen, float **vcm)
{
	FILE *fp = fopen(path, "rb");
	unsigned char h[32], rec[44];
	int i, k, nsph, ndark, nstar;
	double m1 = 0.0;
	*cen = *vcm = NULL;
	if (!fp) return 0;
	if (fread(h, 32, 1, fp) != 1) { fprintf(stderr, "ERROR: cannot read %s\n", path); fclose(fp); return 1; }
	nsph = rdi(h + 16, xdr); ndark = rdi(h + 20, xdr); nstar = rdi(h + 24, xdr);
	if (nstar != nGroup - 1) {
		fprintf(stderr, "ERROR: grp and gtp files don't match: %d vs. %d.\n", nstar, nGroup - 1);
		fclose(fp);
		return 1;
	}
	/* skip gas (12 floats) and dark (9 floats) records */
	if (fseek(fp, (long)nsph * 48 + (long)ndark * 36, SEEK_CUR)) { fclose(fp); return 1; }
	*cen = malloc(12 * (size_t)(nstar ? nstar : 1));
	*vcm = malloc(12 * (size_t)(nstar ? nstar : 1));
	if (!*cen || !*vcm) { fprintf(stderr, "skid: out of memory\n"); fclose(fp); return 1; }
	for (i = 0; i < S->n; ++i) if (grp[i] == 1) m1 += S->mass[i];
	for (i = 0; i < nstar; ++i) {
		if (fread(rec, 44, 1, fp) != 1) { fprintf(stderr, "ERROR: short file %s\n", path); fclose(fp); return 1; }
		/* mass pos[3] vel[3] metals tform eps phi; the first group's mass is
		   checked against the .grp (kd.c:1141-1149) */
		if (i == 0 && rdf(rec, xdr) != (float)m1) {
			fprintf(stderr, "WARNING: Mismatch in group masses between\n");
			fprintf(stderr, "         the .grp and .gtp files.\n");
			fprintf(stderr, "         %g vs. %g\n", rdf(rec, xdr), (float)m1);
		}
		for (k = 0; k < 3; ++k) {
			(*cen)[3 * i + k] = rdf(rec + 4 + 4 * k, xdr);
			(*vcm)[3 * i + k] = rdf(rec + 16 + 4 * k, xdr);
		}
	}
	fclose(fp);
	return 0;
}

int main(int argc, char **argv)
{
	const int nIo = tio_threads();                         /* threads for the text outputs and input decode */
	skg_params P;
	skg_ctx *ctx;
	Snap S;
	int bTau = 0, bCvg = 0, bScoop = 0, bEps = 0, bOutRay = 0, bOutDens = 0, bOutStats = 0, bStandard = 0;
	int bUnbindOnly = 0, bPlain = 0, i, rc, nActive, nScat, nIttr, nGroup, nUnb, nExtra;
	int nMicro[PRUNE_STEPS];
	float fTau = 0, z = 0, Omega0 = 1, G = 1, H0 = 0, Lambda = 0, fQuintess = 0, fCvg = 0, fScoop = 0, fEps = 0;
	float fStep;
	char achName[256] = "skid", achFile[512], achGroup[256] = "";
	float *gcen = NULL, *gvcm = NULL;
	double t0, t1, tDen = 0, tMove = 0, tFof = 0, tMicro = 0, tUnb = 0;
	float *rho = NULL, *r = NULL;
	int *grp;

	printf("SKID v1.4.1 group finder (MI355X GPU path)\n");
	skg_default_params(&P);
	for (i = 1; i < argc; ) {
#define ARG_F(flag, var) else if (!strcmp(argv[i], flag)) { if (++i >= argc) usage(); var = atof(argv[i++]); }
#define ARG_I(flag, var) else if (!strcmp(argv[i], flag)) { if (++i >= argc) usage(); var = atoi(argv[i++]); }
		if (!strcmp(argv[i], "-tau")) { if (++i >= argc) usage(); fTau = atof(argv[i++]); bTau = 1; }
		ARG_F("-z", z) ARG_F("-O", Omega0) ARG_F("-Lambda", Lambda) ARG_F("-Q", fQuintess)
		ARG_F("-G", G) ARG_F("-H", H0) ARG_I("-s", P.nSmooth) ARG_F("-d", P.fDensMin)
		ARG_F("-t", P.fTempMax) ARG_F("-M", P.fMassMax) ARG_I("-m", P.nMembers)
		ARG_I("-maxgroup", P.nMaxMembers)
		else if (!strcmp(argv[i], "-fic")) { P.bForceInitialCut = 1; ++i; }
		else if (!strcmp(argv[i], "-cvg")) { if (++i >= argc) usage(); fCvg = atof(argv[i++]); bCvg = 1; }
		else if (!strcmp(argv[i], "-scoop")) { if (++i >= argc) usage(); fScoop = atof(argv[i++]); bScoop = 1; }
		else if (!strcmp(argv[i], "-nu")) { P.bNoUnbind = 1; ++i; }
		else if (!strcmp(argv[i], "-gd")) { P.bGasAndDark = 1; ++i; }
		else if (!strcmp(argv[i], "-go")) { P.bGasOnly = 1; ++i; }
		else if (!strcmp(argv[i], "-unbind")) {
			if (++i >= argc) usage();
			snprintf(achGroup, sizeof(achGroup), "%s", argv[i++]);
			bUnbindOnly = 1;
		}
		else if (!strcmp(argv[i], "-spline")) { P.iSoftType = SKG_SPLINE; ++i; }
		else if (!strcmp(argv[i], "-plummer")) { P.iSoftType = SKG_PLUMMER; ++i; }
		else if (!strcmp(argv[i], "-e")) { if (++i >= argc) usage(); fEps = atof(argv[i++]); bEps = 1; }
		else if (!strcmp(argv[i], "-p")) {
			if (++i >= argc) usage();
			P.fPeriod[0] = P.fPeriod[1] = P.fPeriod[2] = atof(argv[i++]);
			P.bPeriodic = 1;
		}
		else if (!strcmp(argv[i], "-c")) {
			if (++i >= argc) usage();
			P.fCenter[0] = P.fCenter[1] = P.fCenter[2] = atof(argv[i++]);
		}
		ARG_F("-cx", P.fCenter[0]) ARG_F("-cy", P.fCenter[1]) ARG_F("-cz", P.fCenter[2])
		else if (!strcmp(argv[i], "-o")) { if (++i >= argc) usage(); snprintf(achName, sizeof(achName), "%s", argv[i++]); }
		else if (!strcmp(argv[i], "-ray")) { bOutRay = 1; ++i; }
		else if (!strcmp(argv[i], "-den")) { bOutDens = 1; ++i; }
		else if (!strcmp(argv[i], "-stats")) { bOutStats = 1; ++i; }
		else if (!strcmp(argv[i], "-diag")) { ++i; }
		else if (!strcmp(argv[i], "-std")) { bStandard = 1; ++i; }
		else if (!strcmp(argv[i], "-plain")) { bPlain = 1; ++i; }
		else usage();
#undef ARG_F
#undef ARG_I
	}
	if (!bTau) usage();                                     /* main.c:339 */
	if (!bCvg) fCvg = 0.5*fTau;                             /* main.c:343-345 */
	if (!bScoop) fScoop = 2.0*fTau;
	fStep = 0.5*fCvg;
	P.fTau = fTau; P.fCvg = fCvg; P.fScoop = fScoop; P.fStep = fStep;
	P.G = G; P.z = z;
	P.fShift = 1.0/(1.0 + z);                               /* kd.c:1317-1318 */
	P.fCosmo = P.fShift*exp2hub(H0, Omega0, Lambda, fQuintess, P.fShift);
	P.accumulate = bPlain ? SKG_ACC_PLAIN : SKG_ACC_COMPENSATED;

	memset(&S, 0, sizeof(S));
	if (tio_read_tipsy(stdin, bStandard, &S, nIo)) { fprintf(stderr, "skid: cannot read tipsy input\n"); return 1; }
	printf("nDark:%d nGas:%d nStar:%d\n", S.ndark, S.nsph, S.nstar);
	fflush(stdout);
	if (bEps) for (i = 0; i < S.n; ++i) S.soft[i] = fEps;  /* kdSetSoft (main.c:464) */

	if ((rc = skg_create(&P, &ctx))) die_skg("skg_create", rc);
	if ((rc = skg_load(ctx, S.nsph, S.ndark, S.nstar, S.pos, S.vel, S.mass, S.soft, S.temp))) die_skg("skg_load", rc);
	rho = malloc(4*(size_t)S.n);
	grp = malloc(4*(size_t)S.n);
	if (!rho || !grp) { fprintf(stderr, "skid: out of memory\n"); return 1; }

	if (bUnbindOnly) {                                      /* main.c:349-373 */
		size_t l = strlen(achGroup);
		if (l >= 4 && !strcmp(achGroup + l - 4, ".grp")) achGroup[l - 4] = 0;
		snprintf(achFile, sizeof(achFile), "%s.grp", achGroup);
		if (read_grp(achFile, S.n, grp, &nGroup, nIo)) return 1;
		snprintf(achFile, sizeof(achFile), "%s.gtp", achGroup);
		if (read_gtp(achFile, bStandard, nGroup, &S, grp, &gcen, &gvcm)) return 1;
		if ((rc = skg_set_groups(ctx, grp, gcen, gvcm))) die_skg("skg_set_groups", rc);
		goto unbind_only;
	}
	t0 = now_s();
	if ((rc = skg_density(ctx, NULL, rho, &nExtra))) die_skg("skg_density", rc);
	t1 = now_s(); tDen = t1 - t0;
	if (P.bPeriodic) printf("nExtraScat:%d\n", nExtra);
	if (bOutDens) {                                         /* kdOutDensity (kd.c:1526-1547) */
		snprintf(achFile, sizeof(achFile), "%s.den", achName);
		if (tio_write_floats(achFile, S.n, S.n, rho, "%.10g", 0, nIo)) { perror(achFile); return 1; }
	}
	t0 = now_s();
	if ((rc = skg_init_move(ctx, &nActive))) die_skg("skg_init_move", rc);
	if ((rc = skg_move_first(ctx, &nActive, &nScat))) die_skg("skg_move_first", rc);
	printf("Ittr:%d nActive:%d nScatter:%d\n", 0, nActive, nScat);
	fflush(stdout);
	nIttr = 1;
	while (nActive) {
		if ((rc = skg_move_block(ctx, &nActive, &nScat))) die_skg("skg_move_block", rc);
		printf("Ittr:%d nActive:%d nScatter:%d\n", nIttr, nActive, nScat);
		fflush(stdout);
		++nIttr;
	}
	t1 = now_s(); tMove = t1 - t0;
	t0 = now_s();
	if ((rc = skg_fof(ctx, NULL, &nGroup))) die_skg("skg_fof", rc);
	t1 = now_s(); tFof = t1 - t0;
	t0 = now_s();
	if ((rc = skg_microsteps(ctx, nMicro))) die_skg("skg_microsteps", rc);
	t1 = now_s(); tMicro = t1 - t0;
	for (i = 0; i < PRUNE_STEPS; ++i) printf("Microstep:%d nScatter:%d\n", i + 1, nMicro[i]);
	if (bOutRay) {                                          /* kdOutVector (kd.c:1550-1606) */
		float *d3;
		r = malloc(12*(size_t)S.n);
		d3 = malloc(12*(size_t)S.n);
		if (!r || !d3) { fprintf(stderr, "skid: out of memory\n"); return 1; }
		if ((rc = skg_get_moved(ctx, r))) die_skg("skg_get_moved", rc);
		tio_ray_vectors(S.n, r, S.pos, P.fPeriod, d3, nIo);   /* NaN: did not move, printed as 0 */
		snprintf(achFile, sizeof(achFile), "%s.ray", achName);
		if (tio_write_floats(achFile, S.n, 3L*S.n, d3, "%g", 1, nIo)) { perror(achFile); return 1; }
		free(d3);
	}
	if ((rc = skg_centers(ctx))) die_skg("skg_centers", rc);
unbind_only:
	printf("Groups before Unbind:%d\n", nGroup - 1);
	t0 = now_s();
	if ((rc = skg_unbind(ctx, &nUnb))) die_skg("skg_unbind", rc);
	t1 = now_s(); tUnb = t1 - t0;
	printf("Number of particles Unbound:%d\n", nUnb);
	if ((rc = skg_too_small(ctx))) die_skg("skg_too_small", rc);
	if ((rc = skg_get_groups(ctx, grp, &nGroup))) die_skg("skg_get_groups", rc);
	printf("Number of Groups:%d\n", nGroup - 1);
	{                                                       /* kdOutGroup (kd.c:1502-1523) */
		snprintf(achFile, sizeof(achFile), "%s.grp", achName);
		if (tio_write_ints(achFile, S.n, grp, nIo)) { perror(achFile); return 1; }
	}
	{                                                       /* kdWriteGroup (kd.c:1609-1687) */
		FILE *fp;
		int ng = nGroup;
		float *m = malloc(4*(size_t)ng), *cen = malloc(12*(size_t)ng), *vcm = malloc(12*(size_t)ng),
		      *rad = malloc(4*(size_t)ng);
		unsigned char hdr[32];
		int k;
		if (!m || !cen || !vcm || !rad) { fprintf(stderr, "skid: out of memory\n"); return 1; }
		if ((rc = skg_get_catalogue(ctx, m, cen, vcm, NULL, rad, NULL))) die_skg("skg_get_catalogue", rc);
		snprintf(achFile, sizeof(achFile), "%s.gtp", achName);
		fp = fopen(achFile, "wb");
		if (!fp) { perror(achFile); return 1; }
		memset(hdr, 0, sizeof(hdr));
		if (bStandard) {
			uint64_t u;
			double t = (double)(float)S.time;           /* kd->fTime is a float (kd.c:155) */
			memcpy(&u, &t, 8);
			u = ((uint64_t)bswap32((uint32_t)u) << 32) | bswap32((uint32_t)(u >> 32));
			fwrite(&u, 8, 1, fp);
		} else { double t = (double)(float)S.time; fwrite(&t, 8, 1, fp); }
		wri(fp, ng - 1, bStandard); wri(fp, 3, bStandard); wri(fp, 0, bStandard); wri(fp, 0, bStandard);
		wri(fp, ng - 1, bStandard); wri(fp, 0, bStandard);
		for (i = 1; i < ng; ++i) {
			wrf(fp, m[i], bStandard);
			for (k = 0; k < 3; ++k) wrf(fp, cen[3*i+k], bStandard);
			for (k = 0; k < 3; ++k) wrf(fp, vcm[3*i+k], bStandard);
			wrf(fp, 0.0f, bStandard);                       /* metals */
			wrf(fp, (float)S.time, bStandard);              /* tform */
			wrf(fp, rad[i], bStandard);                     /* eps = group radius */
			wrf(fp, 0.0f, bStandard);                       /* phi */
		}
		fclose(fp);
		free(m); free(cen); free(vcm); free(rad);
	}
	if (bOutStats) {                                        /* kdOutStats (kd.c:1703-1839) */
		FILE *fp;
		int ng = nGroup, k;
		float *st = malloc(4*(size_t)ng*SKG_STAT_N);
		int *cnt = malloc(4*(size_t)ng);
		if (!st || !cnt) { fprintf(stderr, "skid: out of memory\n"); return 1; }
		if ((rc = skg_group_stats(ctx, st))) die_skg("skg_group_stats", rc);
		if ((rc = skg_get_catalogue(ctx, NULL, NULL, NULL, NULL, NULL, cnt))) die_skg("skg_get_catalogue", rc);
		snprintf(achFile, sizeof(achFile), "%s.stat", achName);
		fp = fopen(achFile, "w");
		if (!fp) { perror(achFile); return 1; }
		for (i = 1; i < ng; ++i) {
			const float *r = st + (size_t)i*SKG_STAT_N;
			/* id, n, masses, sqrt of the three V^2, r_Vmax, r_half, r_out, dispersion, centre, vcm, rBound */
			fprintf(fp, "%d %d %g %g %g %g %g %g %g %g %g %g", i, cnt[i], r[0], r[1], r[2], sqrt(r[3]),
					sqrt(r[4]), sqrt(r[5]), r[6], r[7], sqrt(r[8]), r[9]);
			for (k = 10; k < SKG_STAT_N; ++k) fprintf(fp, " %g", r[k]);
			fprintf(fp, "\n");
		}
		fclose(fp);
		free(st); free(cnt);
	}
	printf("SKID GPU wall time:\n");
	if (!bUnbindOnly) {
		printf("   Initial Density:    %.6f\n", tDen);
		printf("   Moving Particles:   %.6f\n", tMove);
		printf("   Friends of Friends: %.6f\n", tFof);
		printf("   Microstepping:      %.6f\n", tMicro);
	}
	if (!P.bNoUnbind) printf("   Unbinding:          %.6f\n", tUnb);
	fflush(stdout);
	skg_destroy(ctx);
	free(rho); free(grp); free(r); free(gcen); free(gvcm);
	free(S.pos); free(S.vel); free(S.mass); free(S.soft); free(S.temp);
	return 0;
}
