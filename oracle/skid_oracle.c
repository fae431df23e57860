/*
 * skid_oracle.c -- single-threaded FP32-faithful restatement of SKID's
 * group-finding path (TEST INFRASTRUCTURE: the checker and the CPU baseline,
 * never the product).
 *
 * Each stage restates the reference's algorithm in this file's own data
 * structures; the step order, comparison directions and the float/double
 * mix of every expression are kept because bit-parity with the reference
 * depends on them.  Build with -O3 -ffp-contract=off and without -march
 * (x86-64 SSE2: no FMA, no excess precision), as the reference is built.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#include <float.h>
#include "skid_oracle.h"

#define ROOTN 1
#define LO(i) ((i) << 1)
#define HI(i) (((i) << 1) + 1)
#define UP(i) ((i) >> 1)
#define SIB(i) (((i) & 1) ? (i) - 1 : (i) + 1)
/* next node in a stackless pre-order walk that skips i's subtree (kd.h:11-15) */
static inline int skip_subtree(int i) { while (i & 1) i >>= 1; return i + 1; }

typedef struct { float r[3], v[3], m, eps, T, ball2, rho; int ord; } SPart;   /* kd.h:27-36 */
typedef struct { float r[3], r0[3], a[3]; int ord; } MPart;                      /* kd.h:38-43 */
typedef struct { float lo[3], hi[3]; float cut; int axis; int first, last; } Node; /* kd.h:57-68 */
typedef struct {
	float rel[3], cen[3], rbound[3], vcm[3];
	float mass, radius;
	int n, start, cur;
} Grp;                                                                           /* kd.h:45-55 */

struct or_ctx {
	or_params P;
	float L[3], C[3];
	int nGas, nDark, nStar, n, inType;
	SPart *p; int nScat;          /* pInit, nInitActive */
	MPart *m; int nMove, nActive; /* pMove */
	Node *node; int nSplit, nNodes, treeOfMove;
	SPart *rep; int nRep;         /* periodic replica scatterers */
	unsigned char *mark;
	int *grp; int nGroup; Grp *g;
	float fCvg, fScoop, fStep;
	int nUnbound;
};

#define SAY(c, ...) do { if (!(c)->P.bQuiet) { printf(__VA_ARGS__); fflush(stdout); } } while (0)

void or_default_params(or_params *P)
{
	memset(P, 0, sizeof(*P));
	P->nSmooth = 64; P->nBucket = 16;      /* main.c:82, 103 */
	P->fPeriod = FLT_MAX;
	P->fTempMax = FLT_MAX; P->fMassMax = FLT_MAX;
	P->nMembers = 8;
	P->iSoftType = OR_SPLINE;
	P->G = 1.0f; P->Omega0 = 1.0f;
}

or_ctx *or_create(const or_params *P)
{
	or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
	int j;
	c->P = *P;
	for (j = 0; j < 3; ++j) {
		c->L[j] = P->bPeriodic ? P->fPeriod : FLT_MAX;   /* main.c:125-128, 273-281 */
		c->C[j] = P->fCenter[j];
	}
	c->fCvg = (P->fCvg > 0) ? P->fCvg : 0.5*P->fTau;      /* main.c:343-345 */
	c->fScoop = (P->fScoop > 0) ? P->fScoop : 2.0*P->fTau;
	c->fStep = 0.5*c->fCvg;
	return c;
}

void or_destroy(or_ctx *c)
{
	if (!c) return;
	free(c->p); free(c->m); free(c->node); free(c->rep); free(c->mark);
	free(c->grp); free(c->g); free(c);
}

static int ptype(const or_ctx *c, int ord)                 /* kd.c:113-119 */
{
	if (ord < c->nGas) return OR_GAS;
	if (ord < c->nGas + c->nDark) return OR_DARK;
	if (ord < c->n) return OR_STAR;
	return 0;
}

int or_load(or_ctx *c, int nGas, int nDark, int nStar, const float *pos, const float *vel,
		const float *mass, const float *soft, const float *temp)
{                                                          /* kd.c:122-222 field mapping */
	int i, j;
	c->nGas = nGas; c->nDark = nDark; c->nStar = nStar;
	c->n = nGas + nDark + nStar;
	c->inType = (nDark ? OR_DARK : 0) | (nGas ? OR_GAS : 0) | (nStar ? OR_STAR : 0);
	c->p = (SPart *)malloc((size_t)c->n * sizeof(SPart));
	if (!c->p) return -1;
	for (i = 0; i < c->n; ++i) {
		SPart *q = &c->p[i];
		for (j = 0; j < 3; ++j) { q->r[j] = pos[3*i+j]; q->v[j] = vel ? vel[3*i+j] : 0.0f; }
		q->m = mass[i]; q->eps = soft[i];
		q->T = (i < nGas && temp) ? temp[i] : 0.0f;
		q->ball2 = -2.0f; q->rho = 0.0f; q->ord = i;
	}
	c->nScat = c->n;
	return 0;
}

/* ------------------------------------------------------------------ */
/* Tree: median quickselect on the longest axis of the inherited box,  */
/* heap-indexed nodes, tight bounds by an up-pass (kd.c:225-552).       */
/* ------------------------------------------------------------------ */
#define SWAPT(T, a, b) do { T _t = (a); (a) = (b); (b) = _t; } while (0)

/* Hoare-style selection leaving the k-th smallest (by axis) at k (kd.c:225-284). */
#define DEFINE_SELECT(NAME, T)                                                  \
static void NAME(T *a, int ax, int k, int lo, int hi)                          \
{                                                                               \
	while (hi > lo) {                                                           \
		double piv = a[k].r[ax];                                                \
		int i = lo - 1, j = hi;                                                 \
		SWAPT(T, a[hi], a[k]);                                                  \
		for (;;) {                                                              \
			while (i < j) { ++i; if (a[i].r[ax] >= piv) break; }                \
			while (i < j) { --j; if (a[j].r[ax] <= piv) break; }                \
			SWAPT(T, a[i], a[j]);                                               \
			if (j <= i) break;                                                  \
		}                                                                       \
		SWAPT(T, a[i], a[j]);          /* undo the crossing exchange */         \
		SWAPT(T, a[i], a[hi]);         /* pivot into its final slot */          \
		if (i >= k) hi = i - 1;                                                 \
		if (i <= k) lo = i + 1;                                                 \
	}                                                                           \
}
DEFINE_SELECT(select_s, SPart)
DEFINE_SELECT(select_m, MPart)

#define DEFINE_UPPASS(NAME, T)                                                  \
static void NAME(Node *nd, const T *a, int i)                                   \
{                                                                               \
	int j, k;                                                                   \
	if (nd[i].axis != -1) {                                                     \
		Node *l, *u;                                                            \
		NAME(nd, a, LO(i)); NAME(nd, a, HI(i));                                 \
		l = &nd[LO(i)]; u = &nd[HI(i)];                                         \
		for (j = 0; j < 3; ++j) {                                               \
			nd[i].lo[j] = (u->lo[j] < l->lo[j]) ? u->lo[j] : l->lo[j];          \
			nd[i].hi[j] = (u->hi[j] > l->hi[j]) ? u->hi[j] : l->hi[j];          \
		}                                                                       \
	} else {                                                                    \
		int f = nd[i].first, t = nd[i].last;                                    \
		for (j = 0; j < 3; ++j) nd[i].lo[j] = nd[i].hi[j] = a[t].r[j];          \
		for (k = f; k < t; ++k)                                                 \
			for (j = 0; j < 3; ++j) {                                           \
				if (a[k].r[j] < nd[i].lo[j]) nd[i].lo[j] = a[k].r[j];           \
				if (a[k].r[j] > nd[i].hi[j]) nd[i].hi[j] = a[k].r[j];           \
			}                                                                   \
	}                                                                           \
}
DEFINE_UPPASS(uppass_s, SPart)
DEFINE_UPPASS(uppass_m, MPart)

/* Build over a[0,n); the root box scans a[0,nBox) (kd.c:371-460, 463-552;
   the move tree scans nMove, not nActive: kd.c:500). */
#define DEFINE_BUILD(NAME, T, SEL, UPP)                                         \
static int NAME(or_ctx *c, T *a, int n, int nBox)                              \
{                                                                               \
	int l = 1, cnt = n, i, j, d, mid;                                           \
	float blo[3], bhi[3];                                                       \
	free(c->node); c->node = NULL;                                              \
	if (n == 0) return 0;                                                       \
	while (cnt > c->P.nBucket) { cnt >>= 1; l <<= 1; }                          \
	c->nSplit = l; c->nNodes = l << 1;                                          \
	c->node = (Node *)malloc((size_t)c->nNodes * sizeof(Node));                 \
	if (!c->node) return -1;                                                    \
	for (j = 0; j < 3; ++j) blo[j] = bhi[j] = a[0].r[j];                        \
	for (i = 1; i < nBox; ++i)                                                  \
		for (j = 0; j < 3; ++j) {                                               \
			if (blo[j] > a[i].r[j]) blo[j] = a[i].r[j];                         \
			else if (bhi[j] < a[i].r[j]) bhi[j] = a[i].r[j];                    \
		}                                                                       \
	{                                                                           \
		Node *nd = c->node;                                                     \
		nd[ROOTN].first = 0; nd[ROOTN].last = n - 1;                            \
		for (j = 0; j < 3; ++j) { nd[ROOTN].lo[j] = blo[j]; nd[ROOTN].hi[j] = bhi[j]; } \
		i = ROOTN;                                                              \
		for (;;) {                                                              \
			if (i < c->nSplit && nd[i].last - nd[i].first > 0) {                \
				d = 0;                                                          \
				for (j = 1; j < 3; ++j)                                         \
					if (nd[i].hi[j] - nd[i].lo[j] > nd[i].hi[d] - nd[i].lo[d]) d = j; \
				nd[i].axis = d;                                                 \
				mid = (nd[i].first + nd[i].last) / 2;                           \
				SEL(a, d, mid, nd[i].first, nd[i].last);                        \
				nd[i].cut = a[mid].r[d];                                        \
				nd[LO(i)] = nd[i]; nd[HI(i)] = nd[i];                           \
				nd[LO(i)].hi[d] = nd[i].cut; nd[LO(i)].last = mid - 1;          \
				nd[HI(i)].lo[d] = nd[i].cut; nd[HI(i)].first = mid;             \
				i = LO(i);                                                      \
			} else {                                                            \
				nd[i].axis = -1;                                                \
				i = skip_subtree(i);                                            \
				if (i == ROOTN) break;                                          \
			}                                                                   \
		}                                                                       \
		UPP(nd, a, ROOTN);                                                      \
	}                                                                           \
	return 0;                                                                   \
}
DEFINE_BUILD(build_s, SPart, select_s, uppass_s)
DEFINE_BUILD(build_m, MPart, select_m, uppass_m)

/* ------------------------------------------------------------------ */
/* Box tests (kd.h:102-335).                                            */
/* ------------------------------------------------------------------ */
static inline float box_d2_np(const Node *nd, float x, float y, float z)   /* kd.h:102-119 */
{
	float dx = nd->lo[0] - x, dx1 = x - nd->hi[0];
	float dy = nd->lo[1] - y, dy1 = y - nd->hi[1];
	float dz = nd->lo[2] - z, dz1 = z - nd->hi[2];
	float d2;
	if (dx > 0.0) d2 = dx*dx; else if (dx1 > 0.0) d2 = dx1*dx1; else d2 = 0.0;
	if (dy > 0.0) d2 += dy*dy; else if (dy1 > 0.0) d2 += dy1*dy1;
	if (dz > 0.0) d2 += dz*dz; else if (dz1 > 0.0) d2 += dz1*dz1;
	return d2;
}

/* Periodic box test choosing the nearer image per axis and returning the
   shifted query (kd.h:122-214).  Returns 0 when the box lies beyond ball2. */
static inline int box_visit_per(const Node *nd, float ball2, const float *L,
		const float *q, float *s)
{
	float d2 = 0.0f;
	int j;
	for (j = 0; j < 3; ++j) {
		float a = nd->lo[j] - q[j], b = q[j] - nd->hi[j], t;
		if (a > 0.0) {
			b += L[j];
			if (b < a) { t = b*b; s[j] = q[j] + L[j]; }
			else { t = a*a; s[j] = q[j]; }
			d2 = (j == 0) ? t : d2 + t;
			if (d2 > ball2) return 0;
		} else if (b > 0.0) {
			a += L[j];
			if (a < b) { t = a*a; s[j] = q[j] - L[j]; }
			else { t = b*b; s[j] = q[j]; }
			d2 = (j == 0) ? t : d2 + t;
			if (d2 > ball2) return 0;
		} else {
			if (j == 0) d2 = 0.0f;
			s[j] = q[j];
		}
	}
	return 1;
}

/* As box_visit_per, also reporting whether the whole box lies inside the
   ball (kd.h:217-335).  Returns 0 prune, 1 visit, 2 contained. */
static inline int box_visit_cont(const Node *nd, float ball2, const float *L,
		const float *q, float *s)
{
	float d2 = 0.0f, mx2 = 0.0f;
	int j;
	for (j = 0; j < 3; ++j) {
		float a = nd->lo[j] - q[j], b = q[j] - nd->hi[j], t, u;
		if (a > 0.0) {
			if (b + L[j] < a) { b += L[j]; a -= L[j]; s[j] = q[j] + L[j]; t = b*b; u = a*a; }
			else { s[j] = q[j]; t = a*a; u = b*b; }
			if (j == 0) { d2 = t; mx2 = u; } else { d2 += t; mx2 += u; }
			if (d2 > ball2) return 0;
		} else if (b > 0.0) {
			if (a + L[j] < b) { a += L[j]; b -= L[j]; s[j] = q[j] - L[j]; t = a*a; u = b*b; }
			else { s[j] = q[j]; t = b*b; u = a*a; }
			if (j == 0) { d2 = t; mx2 = u; } else { d2 += t; mx2 += u; }
			if (d2 > ball2) return 0;
		} else {
			s[j] = q[j];
			u = (a < b) ? a*a : b*b;
			if (j == 0) { d2 = 0.0f; mx2 = u; } else mx2 += u;
		}
	}
	return (mx2 < ball2) ? 2 : 1;
}

/* ------------------------------------------------------------------ */
/* Bounded max-queue as a replacement-selection loser tree             */
/* (smooth1.h:9-90).  Slot s is both external leaf s and internal node s. */
/* ------------------------------------------------------------------ */
typedef struct {
	int k;
	float *key; int *who; float *ox, *oy, *oz;
	int *loser, *winner;
	int head;
} Queue;

static int q_alloc(Queue *Q, int k)
{
	Q->k = k;
	Q->key = (float *)calloc(k, sizeof(float));
	Q->who = (int *)calloc(k, sizeof(int));
	Q->ox = (float *)calloc(k, sizeof(float));
	Q->oy = (float *)calloc(k, sizeof(float));
	Q->oz = (float *)calloc(k, sizeof(float));
	Q->loser = (int *)calloc(k, sizeof(int));
	Q->winner = (int *)calloc(k, sizeof(int));
	Q->head = 0;
	return (Q->key && Q->who && Q->ox && Q->oy && Q->oz && Q->loser && Q->winner) ? 0 : -1;
}
static void q_free(Queue *Q)
{
	free(Q->key); free(Q->who); free(Q->ox); free(Q->oy); free(Q->oz);
	free(Q->loser); free(Q->winner);
}
static void q_build(Queue *Q)                               /* smooth1.h:54-75 */
{
	int j, k = Q->k;
	for (j = k - 1; j > 0; --j) {
		int c0 = j << 1, c1 = c0 + 1;
		int w0 = (c0 < k) ? Q->winner[c0] : c0 - k;
		int w1 = (c1 < k) ? Q->winner[c1] : c1 - k;
		if (Q->key[w0] < Q->key[w1]) { Q->loser[j] = w0; Q->winner[j] = w1; }
		else { Q->loser[j] = w1; Q->winner[j] = w0; }
	}
	Q->head = Q->winner[1];
}
static void q_replace(Queue *Q)                             /* smooth1.h:78-90 */
{
	int q = Q->head, t = (q + Q->k) >> 1;
	while (t) {
		if (Q->key[Q->loser[t]] > Q->key[q]) { int w = Q->loser[t]; Q->loser[t] = q; q = w; }
		t = (t >= 2) ? (t >> 1) : 0;
	}
	Q->head = q;
}

/* Exact kNN refinement of one query (smooth1.c:41-129). */
static void ball_search(or_ctx *c, Queue *Q, const float *ri)
{
	const Node *nd = c->node;
	const SPart *p = c->p;
	unsigned char *mark = c->mark;
	float ball2 = Q->key[Q->head];
	float x = ri[0], y = ri[1], z = ri[2], s[3];
	int cell = ROOTN, cp, ct, k;
	while (nd[cell].axis >= 0) cell = (ri[nd[cell].axis] < nd[cell].cut) ? LO(cell) : HI(cell);
	for (k = nd[cell].first; k <= nd[cell].last; ++k) {
		float dx = x - p[k].r[0], dy = y - p[k].r[1], dz = z - p[k].r[2];
		float d2 = dx*dx + dy*dy + dz*dz;
		if (d2 < ball2) {
			int h;
			if (mark[k]) continue;
			h = Q->head;
			mark[Q->who[h]] = 0; mark[k] = 1;
			Q->key[h] = d2; Q->who[h] = k; Q->ox[h] = Q->oy[h] = Q->oz[h] = 0.0f;
			q_replace(Q);
			ball2 = Q->key[Q->head];
		}
	}
	while (cell != ROOTN) {
		cp = SIB(cell);
		ct = skip_subtree(cp);
		for (;;) {
			if (box_visit_per(&nd[cp], ball2, c->L, ri, s)) {
				if (nd[cp].axis >= 0) { cp = LO(cp); continue; }
				for (k = nd[cp].first; k <= nd[cp].last; ++k) {
					float dx = s[0] - p[k].r[0], dy = s[1] - p[k].r[1], dz = s[2] - p[k].r[2];
					float d2 = dx*dx + dy*dy + dz*dz;
					if (d2 < ball2) {
						int h;
						if (mark[k]) continue;
						h = Q->head;
						mark[Q->who[h]] = 0; mark[k] = 1;
						Q->key[h] = d2; Q->who[h] = k;
						Q->ox[h] = s[0] - x; Q->oy[h] = s[1] - y; Q->oz[h] = s[2] - z;
						q_replace(Q);
						ball2 = Q->key[Q->head];
					}
				}
			}
			cp = skip_subtree(cp);
			if (cp == ct) break;
		}
		cell = UP(cell);
	}
}

/* kNN radii + symmetric M4 density + periodic replicas (smooth1.c:132-335). */
static int density_init(or_ctx *c)
{
	SPart *p = c->p;
	const Node *nd;
	Queue Q;
	int nS = c->nScat, k = c->P.nSmooth;
	int next = 0, leap = 1, pi = 0, pin = 0, cell, i, s;
	float ax = 0.0f, ay = 0.0f, az = 0.0f;

	for (i = 0; i < nS; ++i) { p[i].ball2 = -1.0f; p[i].rho = 0.0f; }
	free(c->mark);
	c->mark = (unsigned char *)calloc(nS > 0 ? nS : 1, 1);
	c->nRep = 0;
	if (!c->node) return 0;
	if (k < 2 || k > nS) return -2;                         /* smooth1.c:12 */
	if (q_alloc(&Q, k)) return -1;
	nd = c->node;
	for (;;) {
		if (leap) {
			float x, y, z;
			int pj;
			while (p[next].ball2 >= 0) { if (++next == nS) goto done; }
			pi = next;
			x = p[pi].r[0]; y = p[pi].r[1]; z = p[pi].r[2];
			cell = ROOTN;
			while (nd[cell].axis >= 0)
				cell = (p[pi].r[nd[cell].axis] < nd[cell].cut) ? LO(cell) : HI(cell);
			for (s = 0; s < k; ++s) c->mark[Q.who[s]] = 0;
			pj = nd[cell].first;
			if (pj > nS - k) pj = nS - k;
			for (s = 0; s < k; ++s, ++pj) {
				float dx = x - p[pj].r[0], dy = y - p[pj].r[1], dz = z - p[pj].r[2];
				c->mark[pj] = 1;
				Q.key[s] = dx*dx + dy*dy + dz*dz;
				Q.who[s] = pj; Q.ox[s] = Q.oy[s] = Q.oz[s] = 0.0f;
			}
			q_build(&Q);
		} else {
			float x, y, z;
			pi = pin;
			x = p[pi].r[0]; y = p[pi].r[1]; z = p[pi].r[2];
			for (s = 0; s < k; ++s) {
				float dx, dy, dz;
				Q.ox[s] -= ax; Q.oy[s] -= ay; Q.oz[s] -= az;
				dx = x + Q.ox[s] - p[Q.who[s]].r[0];
				dy = y + Q.oy[s] - p[Q.who[s]].r[1];
				dz = z + Q.oz[s] - p[Q.who[s]].r[2];
				Q.key[s] = dx*dx + dy*dy + dz*dz;
			}
			q_build(&Q);
			ax = ay = az = 0.0f;
		}
		ball_search(c, &Q, p[pi].r);
		p[pi].ball2 = Q.key[Q.head];
		leap = 1;
		{
			float h2 = Q.key[Q.head], ih2, fNorm;
			ih2 = 4.0/h2;
			fNorm = 0.5*M_1_PI*sqrt(ih2)*ih2;
			for (s = 0; s < k; ++s) {
				float r2, rs;
				int j;
				if (s == Q.head) continue;
				r2 = Q.key[s]*ih2;
				rs = 2.0 - sqrt(r2);
				if (r2 < 1.0) rs = (1.0 - 0.75*rs*r2);
				else rs = 0.25*rs*rs*rs;
				rs *= fNorm;
				j = Q.who[s];
				p[pi].rho += rs*p[j].m;
				p[j].rho += rs*p[pi].m;
				if (p[j].ball2 >= 0) continue;
				if (Q.key[s] < h2) {
					pin = j; leap = 0; h2 = Q.key[s];
					ax = Q.ox[s]; ay = Q.oy[s]; az = Q.oz[s];
				}
			}
		}
	}
done:
	q_free(&Q);
	/* Replica scatterers for every image whose ball reaches the box. */
	if (c->P.bPeriodic) {
		Node box;
		int pass, ix, iy, iz, nr = 0;
		for (i = 0; i < 3; ++i) {
			box.lo[i] = c->C[i] - 0.5*c->L[i];
			box.hi[i] = c->C[i] + 0.5*c->L[i];
		}
		for (pass = 0; pass < 2; ++pass) {
			if (pass == 1) {
				SAY(c, "nExtraScat:%d\n", nr);
				free(c->rep);
				c->rep = (SPart *)malloc((size_t)(nr > 0 ? nr : 1) * sizeof(SPart));
				if (!c->rep) return -1;
				c->nRep = nr;
				nr = 0;
			}
			for (i = 0; i < nS; ++i)
				for (ix = -1; ix <= 1; ++ix) {
					float x = p[i].r[0] + ix*c->L[0];
					for (iy = -1; iy <= 1; ++iy) {
						float y = p[i].r[1] + iy*c->L[1];
						for (iz = -1; iz <= 1; ++iz) {
							float z = p[i].r[2] + iz*c->L[2];
							if (!(ix || iy || iz)) continue;
							if (box_d2_np(&box, x, y, z) < p[i].ball2) {
								if (pass == 1) {
									c->rep[nr] = p[i];
									c->rep[nr].r[0] = x; c->rep[nr].r[1] = y; c->rep[nr].r[2] = z;
								}
								++nr;
							}
						}
					}
				}
		}
	}
	return 0;
}

/* ------------------------------------------------------------------ */
/* Scatterer / mover selection (kd.c:555-699).                          */
/* ------------------------------------------------------------------ */
static int scatter_ok(const or_ctx *c, int i)              /* kd.c:600-627 */
{
	int t = ptype(c, c->p[i].ord), gd = c->P.bGasAndDark, go = c->P.bGasOnly;
	switch (c->inType) {
	case OR_DARK: return 1;
	case OR_GAS: case OR_DARK|OR_GAS:
		return gd || t == OR_GAS;
	case OR_STAR: case OR_DARK|OR_STAR:
		return t == OR_STAR;
	case OR_GAS|OR_STAR: case OR_DARK|OR_GAS|OR_STAR:
		if (gd) return 1;
		return t == OR_GAS || (t == OR_STAR && !go);
	}
	return 0;
}

static int move_ok(const or_ctx *c, int i)                 /* kd.c:555-597 */
{
	const SPart *q = &c->p[i];
	int t = ptype(c, q->ord), gd = c->P.bGasAndDark, go = c->P.bGasOnly;
	if (q->m > c->P.fMassMax) return 0;
	switch (c->inType) {
	case OR_DARK: return q->rho >= c->P.fDensMin;
	case OR_GAS: case OR_DARK|OR_GAS:
		if (gd && t == OR_DARK && q->rho >= c->P.fDensMin) return 1;
		return t == OR_GAS && q->rho >= c->P.fDensMin && q->T <= c->P.fTempMax;
	case OR_STAR: case OR_DARK|OR_STAR:
		return t == OR_STAR;
	case OR_GAS|OR_STAR: case OR_DARK|OR_GAS|OR_STAR:
		if (gd && t == OR_DARK && q->rho >= c->P.fDensMin) return 1;
		if (t == OR_GAS) return q->rho >= c->P.fDensMin && q->T <= c->P.fTempMax;
		return t == OR_STAR && !go;
	}
	return 0;
}

static void scatter_partition(or_ctx *c)                    /* kd.c:669-693 */
{
	int i = 0, j = c->n - 1;
	for (;;) {
		while (scatter_ok(c, i)) if (++i > j) goto done;
		while (!scatter_ok(c, j)) if (i > --j) goto done;
		SWAPT(SPart, c->p[i], c->p[j]);
	}
done:
	c->nScat = i;
}

int or_density(or_ctx *c, float *fBall2, float *fDensity, int *nExtraScat)
{
	int i, rc;
	scatter_partition(c);                                   /* main.c:374 */
	if (c->nScat > 0) {                                     /* main.c:375 */
		if (build_s(c, c->p, c->nScat, c->nScat)) return -1;
	} else { free(c->node); c->node = NULL; }
	c->treeOfMove = 0;
	rc = density_init(c);                                   /* main.c:376-378 */
	if (rc) return rc;
	for (i = 0; i < c->n; ++i) {
		int o = c->p[i].ord;
		if (fBall2) fBall2[o] = (i < c->nScat) ? c->p[i].ball2 : -2.0f;
		if (fDensity) fDensity[o] = c->p[i].rho;
	}
	if (nExtraScat) *nExtraScat = c->nRep;
	return 0;
}

static int by_order_s(const void *a, const void *b)
{ return ((const SPart *)a)->ord - ((const SPart *)b)->ord; }

int or_init_move(or_ctx *c, int *nMove)                     /* kd.c:630-666 */
{
	int i, j, n = 0;
	if (c->P.bOutDens) qsort(c->p, c->n, sizeof(SPart), by_order_s);   /* kd.c:1487-1499 */
	for (i = 0; i < c->n; ++i) n += move_ok(c, i);
	free(c->m);
	c->m = (MPart *)malloc((size_t)(n > 0 ? n : 1) * sizeof(MPart));
	if (!c->m) return -1;
	c->nMove = c->nActive = n;
	n = 0;
	for (i = 0; i < c->n; ++i)
		if (move_ok(c, i)) {
			for (j = 0; j < 3; ++j) { c->m[n].r[j] = c->p[i].r[j]; c->m[n].r0[j] = c->p[i].r[j]; c->m[n].a[j] = 0.0f; }
			c->m[n].ord = c->p[i].ord;
			++n;
		}
	if (nMove) *nMove = c->nMove;
	return 0;
}

static int build_move_tree(or_ctx *c)                       /* kd.c:463-552 */
{
	c->treeOfMove = 1;
	if (c->nActive == 0) { free(c->node); c->node = NULL; return 0; }
	return build_m(c, c->m, c->nActive, c->nMove);
}

/* Moving particles strictly inside one scatterer's ball, in tree order (smooth1.c:338-384). */
static int ball_gather(or_ctx *c, float ball2, const float *ri, int *hit, float *hd)
{
	const Node *nd = c->node;
	const MPart *m = c->m;
	float x = ri[0], y = ri[1], z = ri[2];
	int cp = ROOTN, n = 0, k;
	if (!nd) return 0;
	for (;;) {
		if (box_d2_np(&nd[cp], x, y, z) < ball2) {
			if (nd[cp].axis >= 0) { cp = LO(cp); continue; }
			for (k = nd[cp].first; k <= nd[cp].last; ++k) {
				float dx = x - m[k].r[0], dy = y - m[k].r[1], dz = z - m[k].r[2];
				float d2 = dx*dx + dy*dy + dz*dz;
				if (d2 < ball2) {
					hit[n] = k; hd[4*n] = d2; hd[4*n+1] = dx; hd[4*n+2] = dy; hd[4*n+3] = dz;
					++n;
				}
			}
		}
		cp = skip_subtree(cp);
		if (cp == ROOTN) break;
	}
	return n;
}

static int scatter_cut(SPart *p, int n, float thr)          /* smooth1.c:387-405 */
{
	int i = 0, j = n - 1;
	if (n <= 0) return 0;
	for (;;) {
		while (p[i].rho >= thr) if (++i > j) return i;
		while (p[j].rho < thr) if (i > --j) return i;
		SWAPT(SPart, p[i], p[j]);
	}
}

/* Frozen-field gradient at the moving positions + scatterer removal (smooth1.c:408-518). */
static int acc_density(or_ctx *c, int bInitial)
{
	int *hit = (int *)malloc((size_t)(c->nActive > 0 ? c->nActive : 1) * sizeof(int));
	float *hd = (float *)malloc((size_t)(c->nActive > 0 ? c->nActive : 1) * 4 * sizeof(float));
	float thr = 0.0f;
	int pass, i, t;
	if (!hit || !hd) { free(hit); free(hd); return -1; }
	for (i = 0; i < c->nActive; ++i) c->m[i].a[0] = c->m[i].a[1] = c->m[i].a[2] = 0.0f;
	for (pass = 0; pass < 2; ++pass) {
		SPart *p = pass ? c->rep : c->p;
		int np = pass ? c->nRep : c->nScat;
		for (i = 0; i < np; ++i) {
			int nh = ball_gather(c, p[i].ball2, p[i].r, hit, hd);
			if (nh > 0) {
				float ih2 = 4.0/p[i].ball2;
				float fNorm = M_1_PI*ih2*ih2*sqrt(ih2)*p[i].m;
				for (t = 0; t < nh; ++t) {
					float r2 = hd[4*t]*ih2, rs = sqrt(r2);
					MPart *q = &c->m[hit[t]];
					if (r2 < 1.0) rs = -3.0 + 2.25*rs;
					else rs = -3.0/rs + 3.0 - 0.75*rs;
					rs *= fNorm;
					q->a[0] += hd[4*t+1]*rs;
					q->a[1] += hd[4*t+2]*rs;
					q->a[2] += hd[4*t+3]*rs;
				}
				if (thr == 0.0) thr = p[i].rho;
				else if (p[i].rho < thr) thr = p[i].rho;
			} else if (bInitial) {
				p[i].rho = 0.0f;
			}
		}
	}
	c->nScat = scatter_cut(c->p, c->nScat, thr);
	if (c->nRep > 0) c->nRep = scatter_cut(c->rep, c->nRep, thr);
	free(hit); free(hd);
	return c->nScat + c->nRep;
}

static void move_step(or_ctx *c, float fStep)              /* kd.c:702-732 */
{
	int i, j;
	for (i = 0; i < c->nActive; ++i) {
		MPart *q = &c->m[i];
		float ax = q->a[0], ay = q->a[1], az = q->a[2], ai;
		ai = sqrt(ax*ax + ay*ay + az*az);
		if (ai > 0.0) ai = fStep/sqrt(ax*ax + ay*ay + az*az);
		else ai = 0.0;
		q->r[0] -= ai*ax; q->r[1] -= ai*ay; q->r[2] -= ai*az;
		for (j = 0; j < 3; ++j) {
			if (q->r[j] > c->C[j] + 0.5*c->L[j]) q->r[j] -= c->L[j];
			if (q->r[j] <= c->C[j] - 0.5*c->L[j]) q->r[j] += c->L[j];
		}
	}
}

static float disp2(const or_ctx *c, const MPart *q)        /* kd.c:754-762 */
{
	float hx = 0.5*c->L[0], hy = 0.5*c->L[1], hz = 0.5*c->L[2];
	float dx = q->r[0] - q->r0[0], dy = q->r[1] - q->r0[1], dz = q->r[2] - q->r0[2];
	if (dx > hx) dx -= 2*hx;
	if (dx <= -hx) dx += 2*hx;
	if (dy > hy) dy -= 2*hy;
	if (dy <= -hy) dy += 2*hy;
	if (dz > hz) dz -= 2*hz;
	if (dz <= -hz) dz += 2*hz;
	return dx*dx + dy*dy + dz*dz;
}

static int prune_inactive(or_ctx *c)                        /* kd.c:735-793 */
{
	float cv2 = c->fCvg*c->fCvg;
	int i = 0, j = c->nActive - 1, k;
	if (c->nActive > 0) {
		for (;;) {
			for (;;) { if (disp2(c, &c->m[i]) < cv2) break; if (++i > j) goto done; }
			for (;;) { if (disp2(c, &c->m[j]) >= cv2) break; if (i > --j) goto done; }
			SWAPT(MPart, c->m[i], c->m[j]);
		}
	}
done:
	c->nActive = i;
	for (i = 0; i < c->nActive; ++i)
		for (k = 0; k < 3; ++k) c->m[i].r0[k] = c->m[i].r[k];
	return c->nActive;
}

int or_move_first(or_ctx *c, int *nActive, int *nScatter)
{
	int ns;
	if (build_move_tree(c)) return -1;
	ns = acc_density(c, (c->inType == OR_DARK) || c->P.bForceInitialCut);   /* kd.c:696-699 */
	if (ns < 0) return -1;
	move_step(c, c->fStep);
	if (nActive) *nActive = c->nActive;
	if (nScatter) *nScatter = ns;
	return 0;
}

int or_first_gradient(or_ctx *c, float *acc)
{
	/* smAccDensity's accumulation of the initial iteration (main.c:394-401),
	   before any move: a at every mover, by iOrder (NaN for non-movers). */
	int i, j, nm;
	if (or_init_move(c, &nm)) return -1;
	if (build_move_tree(c)) return -1;
	if (acc_density(c, (c->inType == OR_DARK) || c->P.bForceInitialCut) < 0) return -1;
	for (i = 0; i < 3*c->n; ++i) acc[i] = NAN;
	for (i = 0; i < c->nMove; ++i)
		for (j = 0; j < 3; ++j) acc[3*c->m[i].ord+j] = c->m[i].a[j];
	return 0;
}

int or_move_block(or_ctx *c, int *nActive, int *nScatter)
{
	int i, ns = 0;
	for (i = 0; i < 5; ++i) {                                 /* PRUNE_STEPS, main.c:13 */
		if (build_move_tree(c)) return -1;
		ns = acc_density(c, 0);
		if (ns < 0) return -1;
		move_step(c, c->fStep);
	}
	prune_inactive(c);
	if (nActive) *nActive = c->nActive;
	if (nScatter) *nScatter = ns;
	return 0;
}

int or_get_moved(or_ctx *c, float *r)
{
	int i, j;
	for (i = 0; i < 3*c->n; ++i) r[i] = NAN;
	for (i = 0; i < c->nMove; ++i)
		for (j = 0; j < 3; ++j) r[3*c->m[i].ord+j] = c->m[i].r[j];
	return 0;
}

/* Breadth-first FOF over the moving tree (kd.c:802-917). lab[] is per tree slot. */
static int fof_bfs(or_ctx *c, float fTau, int *lab)
{
	const Node *nd = c->node;
	const MPart *m = c->m;
	int n = c->nActive, head = 0, tail = 0, ng = 0, pn, k, cp, st;
	float tau2 = fTau*fTau, s[3];
	int *fifo = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
	if (!fifo) return -1;
	for (pn = 0; pn < n; ++pn) lab[pn] = 0;
	for (pn = 0; pn < n; ++pn) {
		if (lab[pn]) continue;
		lab[pn] = ++ng;
		fifo[tail++] = pn; if (tail == n) tail = 0;
		while (head != tail) {
			int pi = fifo[head++];
			if (head == n) head = 0;
			cp = ROOTN;
			for (;;) {
				st = box_visit_cont(&nd[cp], tau2, c->L, m[pi].r, s);
				if (st == 1 && nd[cp].axis >= 0) { cp = LO(cp); continue; }
				if (st == 1) {
					for (k = nd[cp].first; k <= nd[cp].last; ++k) {
						float dx, dy, dz;
						if (lab[k]) continue;
						dx = s[0] - m[k].r[0]; dy = s[1] - m[k].r[1]; dz = s[2] - m[k].r[2];
						if (dx*dx + dy*dy + dz*dz < tau2) {
							lab[k] = ng; fifo[tail++] = k; if (tail == n) tail = 0;
						}
					}
				} else if (st == 2) {
					for (k = nd[cp].first; k <= nd[cp].last; ++k) {
						if (lab[k]) continue;
						lab[k] = ng; fifo[tail++] = k; if (tail == n) tail = 0;
					}
				}
				cp = skip_subtree(cp);
				if (cp == ROOTN) break;
			}
		}
	}
	free(fifo);
	return ng;
}

int or_fof(or_ctx *c, int *iGroup, int *nGroup)
{
	int i, ng;
	int *lab;
	c->nActive = c->nMove;
	free(c->grp);
	c->grp = (int *)calloc((size_t)(c->n > 0 ? c->n : 1), sizeof(int));
	if (!c->grp) return -1;
	if (c->nActive == 0) { c->nGroup = 1; goto out; }
	if (build_move_tree(c)) return -1;
	lab = (int *)malloc((size_t)c->nActive * sizeof(int));
	if (!lab) return -1;
	ng = fof_bfs(c, c->P.fTau, lab);
	if (ng < 0) { free(lab); return -1; }
	c->nGroup = ng + 1;
	for (i = 0; i < c->nActive; ++i) c->grp[c->m[i].ord] = lab[i];
	free(lab);
out:
	if (iGroup) memcpy(iGroup, c->grp, (size_t)c->n * sizeof(int));
	if (nGroup) *nGroup = c->nGroup;
	return 0;
}

int or_microsteps(or_ctx *c, int *nScatter)                 /* main.c:431-438 */
{                                                          /* nScatter: 5 counts, one per micro-step */
	int i, ns = 0;
	c->nActive = c->nMove;
	for (i = 0; i < 5; ++i) {
		if (build_move_tree(c)) return -1;
		ns = acc_density(c, 0);
		if (ns < 0) return -1;
		move_step(c, 0.1*c->fStep);
		SAY(c, "Microstep:%d nScatter:%d\n", i + 1, ns);
		if (nScatter) nScatter[i] = ns;
	}
	return 0;
}

/* ------------------------------------------------------------------ */
/* Group bookkeeping (kd.c:969-1075, 1205-1293).                         */
/* ------------------------------------------------------------------ */
int or_centers(or_ctx *c)
{
	int i, j, gi;
	free(c->g);
	c->g = (Grp *)calloc((size_t)c->nGroup, sizeof(Grp));
	if (!c->g) return -1;
	for (i = 0; i < c->n; ++i) {                            /* kd.c:990-1001 */
		gi = c->grp[c->p[i].ord];
		c->g[gi].n += 1;
		if (!gi) continue;
		c->g[gi].mass += c->p[i].m;
		for (j = 0; j < 3; ++j) c->g[gi].rel[j] = c->p[i].r[j];
	}
	for (i = 0; i < c->nMove; ++i) {                        /* kd.c:1031-1041 */
		gi = c->grp[c->m[i].ord];
		if (!gi) continue;
		for (j = 0; j < 3; ++j) {
			float del = c->m[i].r[j] - c->g[gi].rel[j];
			if (del > 0.5*c->L[j]) del -= c->L[j];
			if (del <= -0.5*c->L[j]) del += c->L[j];
			c->g[gi].cen[j] += del;
		}
	}
	for (i = 0; i < c->n; ++i) {                            /* kd.c:1042-1048 */
		gi = c->grp[c->p[i].ord];
		if (!gi) continue;
		for (j = 0; j < 3; ++j) c->g[gi].vcm[j] += c->p[i].m*c->p[i].v[j];
	}
	for (i = 1; i < c->nGroup; ++i)                         /* kd.c:1049-1059 */
		for (j = 0; j < 3; ++j) {
			c->g[i].cen[j] /= c->g[i].n;
			c->g[i].cen[j] += c->g[i].rel[j];
			if (c->g[i].cen[j] > c->C[j] + 0.5*c->L[j]) c->g[i].cen[j] -= c->L[j];
			if (c->g[i].cen[j] <= c->C[j] - 0.5*c->L[j]) c->g[i].cen[j] += c->L[j];
			c->g[i].vcm[j] /= c->g[i].mass;
		}
	free(c->m); c->m = NULL; c->nMove = c->nActive = 0;
	free(c->node); c->node = NULL;
	return 0;
}

static void group_order(or_ctx *c)                          /* kd.c:1205-1245 */
{
	int gi, pi = 0, ig = 0, ind;
	c->g[0].start = c->g[0].cur = 0;
	for (gi = 1; gi < c->nGroup; ++gi) {
		c->g[gi].start = c->g[gi-1].start + c->g[gi-1].n;
		c->g[gi].cur = c->g[gi].start;
	}
	if (c->n == 0) return;
	for (;;) {
		int g = c->grp[c->p[pi].ord];
		ind = c->g[g].cur++;
		if (ind == pi) {
			while (c->g[ig].cur == c->g[ig].start + c->g[ig].n)
				if (++ig == c->nGroup) return;
			pi = c->g[ig].cur;
			if (pi == c->n) return;
		} else {
			SWAPT(SPart, c->p[ind], c->p[pi]);
		}
	}
}

float or_spline_pot(float d2, float twoh)                   /* grav.h:11-31 */
{
	float a;
	double r = sqrt(d2), u, dih, dir;
	if (r < twoh) {
		dih = 2.0/twoh;
		u = r*dih;
		if (u < 1.0) {
			a = dih*(7.0/5.0 - 2.0/3.0*u*u + 3.0/10.0*u*u*u*u
					- 1.0/10.0*u*u*u*u*u);
		} else {
			dir = 1.0/r;
			a = -1.0/15.0*dir + dih*(8.0/5.0 - 4.0/3.0*u*u + u*u*u
					- 3.0/10.0*u*u*u*u + 1.0/30.0*u*u*u*u*u);
		}
	} else {
		a = 1.0/r;
	}
	return a;
}

static inline float pair_kernel(int soft, float d2, float twoh)   /* grav.c:24-31 */
{
	float dir;
	if (soft == OR_PLUMMER) dir = 1.0/sqrt(d2 + 0.25*twoh*twoh);
	else dir = or_spline_pot(d2, twoh);
	return dir;
}

static double exp2hub(const or_ctx *c, double a)            /* cosmo.c:46-58 */
{
	double om = c->P.Omega0, lam = c->P.Lambda, q = c->P.fQuintess, rad = 0.0;
	double curv = 1.0 - om - lam - rad - q;
	return (double)c->P.H0*sqrt(om*a + curv*a*a + rad + q*a*a*sqrt(a) + lam*a*a*a*a)/(a*a);
}

/* Periodic scoop ball over the group-0 tree (grav.c:63-135). */
static void scoop_pot(or_ctx *c, SPart *q, int n, const float *cen, const float *rel,
		double *pot)
{
	const Node *nd = c->node;
	const SPart *p = c->p;
	float ball2 = c->fScoop*c->fScoop, s[3];
	int cp = ROOTN, k, i;
	if (!nd) return;
	for (;;) {
		if (box_visit_per(&nd[cp], ball2, c->L, cen, s)) {
			if (nd[cp].axis >= 0) { cp = LO(cp); continue; }
			for (k = nd[cp].first; k <= nd[cp].last; ++k) {
				float dx = s[0] - p[k].r[0], dy = s[1] - p[k].r[1], dz = s[2] - p[k].r[2];
				float nx, ny, nz;
				if (!(dx*dx + dy*dy + dz*dz < ball2)) continue;
				nx = p[k].r[0] - rel[0]; ny = p[k].r[1] - rel[1]; nz = p[k].r[2] - rel[2];
				if (nx > 0.5*c->L[0]) nx -= c->L[0];
				if (nx <= -0.5*c->L[0]) nx += c->L[0];
				if (ny > 0.5*c->L[1]) ny -= c->L[1];
				if (ny <= -0.5*c->L[1]) ny += c->L[1];
				if (nz > 0.5*c->L[2]) nz -= c->L[2];
				if (nz <= -0.5*c->L[2]) nz += c->L[2];
				for (i = 0; i < n; ++i) {
					float ex = nx - q[i].r[0], ey = ny - q[i].r[1], ez = nz - q[i].r[2];
					float d2 = ex*ex + ey*ey + ez*ez;
					float twoh = p[k].eps + q[i].eps;
					pot[i] += c->P.G*p[k].m*pair_kernel(c->P.iSoftType, d2, twoh);
				}
			}
		}
		cp = skip_subtree(cp);
		if (cp == ROOTN) break;
	}
}

int or_unbind(or_ctx *c, int *nUnbound)                      /* kd.c:1299-1466 */
{
	float hx = 0.5*c->L[0], hy = 0.5*c->L[1], hz = 0.5*c->L[2];
	float fShift, fCosmo;
	int gi, i, j, nUnb = 0, nMax = c->P.nMaxMembers;   /* kd.c:1330: groups of n >= nMax stay */
	group_order(c);
	fShift = 1.0/(1.0 + c->P.z);
	fCosmo = fShift*exp2hub(c, fShift);
	c->nScat = c->g[0].n;                                    /* group 0 leads after ordering */
	if (c->nScat > 0) { if (build_s(c, c->p, c->nScat, c->nScat)) return -1; }
	else { free(c->node); c->node = NULL; }
	SAY(c, "Groups before Unbind:%d\n", c->nGroup - 1);
	for (gi = 1; gi < c->nGroup; ++gi) {
		int n = c->g[gi].n, iBig, iMin;
		float fBig, fMin;
		double dMass, rcm[3], vcm[3], *pot;
		SPart *q;
		if (n >= nMax) continue;
		q = (SPart *)malloc((size_t)(n > 0 ? n : 1) * sizeof(SPart));
		pot = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
		if (!q || !pot) { free(q); free(pot); return -1; }
		for (i = 0; i < n; ++i) {
			const SPart *s = &c->p[c->g[gi].start + i];
			float dx = s->r[0] - c->g[gi].rel[0], dy = s->r[1] - c->g[gi].rel[1], dz = s->r[2] - c->g[gi].rel[2];
			if (dx > hx) dx -= 2*hx;
			if (dx <= -hx) dx += 2*hx;
			if (dy > hy) dy -= 2*hy;
			if (dy <= -hy) dy += 2*hy;
			if (dz > hz) dz -= 2*hz;
			if (dz <= -hz) dz += 2*hz;
			q[i] = *s; q[i].r[0] = dx; q[i].r[1] = dy; q[i].r[2] = dz;
		}
		dMass = 0.0;
		for (j = 0; j < 3; ++j) rcm[j] = vcm[j] = 0.0;
		for (i = 0; i < n; ++i) {
			dMass += q[i].m;
			for (j = 0; j < 3; ++j) { rcm[j] += q[i].m*q[i].r[j]; vcm[j] += q[i].m*q[i].v[j]; }
		}
		for (j = 0; j < 3; ++j) { rcm[j] /= dMass; vcm[j] /= dMass; }
		for (i = 0; i < n; ++i) pot[i] = 0.0;               /* grav.c:8-36 */
		for (i = 0; i < n - 1; ++i) {
			float x = q[i].r[0], y = q[i].r[1], z = q[i].r[2];
			int k;
			for (k = i + 1; k < n; ++k) {
				float dx = x - q[k].r[0], dy = y - q[k].r[1], dz = z - q[k].r[2];
				float d2 = dx*dx + dy*dy + dz*dz;
				float twoh = q[i].eps + q[k].eps;
				float dir = pair_kernel(c->P.iSoftType, d2, twoh);
				pot[i] += c->P.G*q[k].m*dir;
				pot[k] += c->P.G*q[i].m*dir;
			}
		}
		scoop_pot(c, q, n, c->g[gi].cen, c->g[gi].rel, pot);   /* always on: kd.c:1379-1384 */
		iMin = 0;
		for (;;) {
			iBig = 0; fBig = -1.0; iMin = 0; fMin = 1.0;
			for (i = 0; i < n; ++i) {
				float dv, dv2 = 0.0, fTot;
				for (j = 0; j < 3; ++j) {
					dv = fShift*(q[i].v[j] - vcm[j]) + fCosmo*(q[i].r[j] - rcm[j]);
					dv2 += dv*dv;
				}
				fTot = 0.5*dv2 - pot[i]*(1.0 + c->P.z);
				if (fTot > fBig) { fBig = fTot; iBig = i; }
				if (fTot < fMin) { fMin = fTot; iMin = i; }
			}
			if (fBig < 0 || c->P.bNoUnbind) break;
			c->grp[q[iBig].ord] = 0;
			--c->g[gi].n; --n; ++c->g[0].n; ++nUnb;
			if (n == 0) { dMass = 0.0; for (j = 0; j < 3; ++j) vcm[j] = 0.0; break; }
			dMass -= q[iBig].m;
			for (j = 0; j < 3; ++j) {
				rcm[j] += q[iBig].m/dMass*(rcm[j] - q[iBig].r[j]);
				vcm[j] += q[iBig].m/dMass*(vcm[j] - q[iBig].v[j]);
			}
			SWAPT(SPart, q[iBig], q[n]);
			{ double t = pot[iBig]; pot[iBig] = pot[n]; pot[n] = t; }
			if (c->inType == OR_DARK || c->inType == OR_STAR) {   /* grav.c:39-60 */
				const SPart *s = &q[n];
				for (i = 0; i < n; ++i) {
					float dx = s->r[0] - q[i].r[0], dy = s->r[1] - q[i].r[1], dz = s->r[2] - q[i].r[2];
					float d2 = dx*dx + dy*dy + dz*dz;
					float twoh = s->eps + q[i].eps;
					pot[i] -= c->P.G*s->m*pair_kernel(c->P.iSoftType, d2, twoh);
				}
			}
		}
		c->g[gi].mass = dMass;
		for (j = 0; j < 3; ++j) c->g[gi].vcm[j] = vcm[j];
		for (j = 0; j < 3; ++j) {                           /* hx on every axis: kd.c:1452-1456 */
			float dx = q[iMin].r[j] + c->g[gi].rel[j];
			if (dx > hx) dx -= 2*hx;
			if (dx <= -hx) dx += 2*hx;
			c->g[gi].rbound[j] = dx;
		}
		free(pot); free(q);
	}
	SAY(c, "Number of particles Unbound:%d\n", nUnb);
	group_order(c);
	c->nUnbound = nUnb;
	if (nUnbound) *nUnbound = nUnb;
	return 0;
}

int or_too_small(or_ctx *c)                                 /* kd.c:1251-1293 */
{
	int *map, i, ng = 1;
	group_order(c);
	map = (int *)malloc((size_t)c->nGroup * sizeof(int));
	if (!map) return -1;
	for (i = 1; i < c->nGroup; ++i)
		if (c->g[i].n < c->P.nMembers) { c->g[0].n += c->g[i].n; c->g[i].n = 0; }
	map[0] = 0;
	for (i = 1; i < c->nGroup; ++i) {
		map[i] = ng;
		c->g[ng] = c->g[i];
		if (c->g[i].n == 0) map[i] = 0; else ++ng;
	}
	for (i = 0; i < c->n; ++i) c->grp[i] = map[c->grp[i]];
	free(map);
	c->nGroup = ng;
	group_order(c);
	return 0;
}

int or_get_groups(or_ctx *c, int *iGroup, int *nGroup)
{
	if (iGroup && c->grp) memcpy(iGroup, c->grp, (size_t)c->n * sizeof(int));
	if (nGroup) *nGroup = c->nGroup;
	return 0;
}

int or_get_catalogue(or_ctx *c, float *mass, float *center, float *vcm, float *rbound,
		float *radius, int *nMembers)
{                                                          /* kd.c:1612-1644 radius */
	int i, j;
	for (i = 1; i < c->nGroup; ++i) c->g[i].radius = 0.0f;
	for (i = 0; i < c->n; ++i) {
		int gi = c->grp[c->p[i].ord];
		float rad = 0.0f;
		if (!gi) continue;
		for (j = 0; j < 3; ++j) {
			float del = c->p[i].r[j] - c->g[gi].cen[j];
			if (del > 0.5*c->L[j]) del -= c->L[j];
			if (del <= -0.5*c->L[j]) del += c->L[j];
			rad += del*del;
		}
		rad = sqrt(rad);
		if (rad > c->g[gi].radius) c->g[gi].radius = rad;
	}
	for (i = 0; i < c->nGroup; ++i) {
		if (mass) mass[i] = c->g[i].mass;
		if (radius) radius[i] = c->g[i].radius;
		if (nMembers) nMembers[i] = c->g[i].n;
		for (j = 0; j < 3; ++j) {
			if (center) center[3*i+j] = c->g[i].cen[j];
			if (vcm) vcm[3*i+j] = c->g[i].vcm[j];
			if (rbound) rbound[3*i+j] = c->g[i].rbound[j];
		}
	}
	return 0;
}

int or_run(or_ctx *c, int *trace, int maxTrace, int *nTrace)
{
	int nA, nS, nt = 0, rc;
	if ((rc = or_density(c, NULL, NULL, NULL))) return rc;
	if (or_init_move(c, &nA)) return -1;
	if (or_move_first(c, &nA, &nS)) return -1;
	SAY(c, "Ittr:%d nActive:%d nScatter:%d\n", 0, nA, nS);
	if (trace && nt < maxTrace) { trace[2*nt] = nA; trace[2*nt+1] = nS; }
	++nt;
	while (nA) {
		if (or_move_block(c, &nA, &nS)) return -1;
		SAY(c, "Ittr:%d nActive:%d nScatter:%d\n", nt, nA, nS);
		if (trace && nt < maxTrace) { trace[2*nt] = nA; trace[2*nt+1] = nS; }
		++nt;
	}
	if (nTrace) *nTrace = nt;
	if (or_fof(c, NULL, NULL)) return -1;
	if (or_microsteps(c, NULL)) return -1;
	if (or_centers(c)) return -1;
	if (c->P.fEps > 0) { int i; for (i = 0; i < c->n; ++i) c->p[i].eps = c->P.fEps; }
	if (or_unbind(c, NULL)) return -1;
	if (or_too_small(c)) return -1;
	return 0;
}

/* ------------------------------------------------------------------ */
/* Standalone checkers.                                                 */
/* ------------------------------------------------------------------ */
int or_fof_points(int n, const float *pos, float fTau, float fPeriod, int bPeriodic,
		int nBucket, int *iGroup)
{
	or_params P;
	or_ctx *c;
	int i, j, ng;
	int *lab;
	or_default_params(&P);
	P.nBucket = nBucket; P.fPeriod = fPeriod; P.bPeriodic = bPeriodic; P.fTau = fTau;
	c = or_create(&P);
	c->m = (MPart *)malloc((size_t)(n > 0 ? n : 1) * sizeof(MPart));
	lab = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
	for (i = 0; i < n; ++i) {
		for (j = 0; j < 3; ++j) c->m[i].r[j] = c->m[i].r0[j] = pos[3*i+j];
		c->m[i].ord = i;
	}
	c->nMove = c->nActive = n;
	if (n > 0) build_m(c, c->m, n, n);
	ng = n > 0 ? fof_bfs(c, fTau, lab) : 0;
	for (i = 0; i < n; ++i) iGroup[c->m[i].ord] = lab[i];
	free(lab);
	or_destroy(c);
	return ng;
}

static inline float img_delta(float q, float p, float L, int per)
{                                     /* query-side image shift rounded first (Appendix B) */
	float d = q - p;
	if (per) {
		if (d > 0.5f*L) d = (q - L) - p;
		else if (d < -0.5f*L) d = (q + L) - p;
	}
	return d;
}

static inline float pair_d2(const float *q, const float *p, float L, int per)
{
	float dx = img_delta(q[0], p[0], L, per), dy = img_delta(q[1], p[1], L, per),
	      dz = img_delta(q[2], p[2], L, per);
	return dx*dx + dy*dy + dz*dz;
}

typedef struct { float d2; int j; } KD2;
static int kd2_cmp(const void *a, const void *b)
{
	const KD2 *x = (const KD2 *)a, *y = (const KD2 *)b;
	if (x->d2 < y->d2) return -1;
	if (x->d2 > y->d2) return 1;
	return x->j - y->j;
}

static inline int kd2_less(float a, int ia, float b, int ib)
{
	return a < b || (a == b && ia < ib);
}

int or_knn_brute(int n, const float *pos, int k, float fPeriod, int bPeriodic,
		float *fBall2, int *nbr)
{
	/* per query: bounded max-heap of the k smallest (d2, index), then sorted */
	KD2 *h = (KD2 *)malloc((size_t)(k > 0 ? k : 1) * sizeof(KD2));
	int i, j;
	if (!h) return -1;
	if (k > n || k < 1) { free(h); return -2; }
	for (i = 0; i < n; ++i) {
		int m = 0;
		for (j = 0; j < n; ++j) {
			float d2 = pair_d2(&pos[3*i], &pos[3*j], fPeriod, bPeriodic);
			int c, p;
			if (m < k) {                       /* sift up */
				c = m++;
				while (c > 0) {
					p = (c - 1) >> 1;
					if (!kd2_less(h[p].d2, h[p].j, d2, j)) break;
					h[c] = h[p]; c = p;
				}
				h[c].d2 = d2; h[c].j = j;
			} else if (kd2_less(d2, j, h[0].d2, h[0].j)) {   /* replace root, sift down */
				c = 0;
				for (;;) {
					int l = 2*c + 1, r = l + 1, b = l;
					if (l >= k) break;
					if (r < k && kd2_less(h[l].d2, h[l].j, h[r].d2, h[r].j)) b = r;
					if (!kd2_less(d2, j, h[b].d2, h[b].j)) break;
					h[c] = h[b]; c = b;
				}
				h[c].d2 = d2; h[c].j = j;
			}
		}
		qsort(h, k, sizeof(KD2), kd2_cmp);
		if (fBall2) fBall2[i] = h[k-1].d2;
		if (nbr) for (j = 0; j < k; ++j) nbr[(size_t)i*k + j] = h[j].j;
	}
	free(h);
	return 0;
}

static int uf_find(int *par, int x)
{
	while (par[x] != x) { par[x] = par[par[x]]; x = par[x]; }
	return x;
}

int or_fof_brute(int n, const float *pos, float fTau, float fPeriod, int bPeriodic, int *iGroup)
{
	int *par = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
	int *rank = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
	float tau2 = fTau*fTau;
	int i, j, ng = 0;
	if (!par || !rank) { free(par); free(rank); return -1; }
	for (i = 0; i < n; ++i) par[i] = i;
	for (i = 0; i < n; ++i)
		for (j = i + 1; j < n; ++j)
			if (pair_d2(&pos[3*i], &pos[3*j], fPeriod, bPeriodic) < tau2 ||
			    pair_d2(&pos[3*j], &pos[3*i], fPeriod, bPeriodic) < tau2) {
				int a = uf_find(par, i), b = uf_find(par, j);
				if (a < b) par[b] = a; else if (b < a) par[a] = b;
			}
	for (i = 0; i < n; ++i) rank[i] = 0;
	for (i = 0; i < n; ++i) { int r = uf_find(par, i); if (r == i) rank[i] = ++ng; }
	for (i = 0; i < n; ++i) iGroup[i] = rank[uf_find(par, i)];
	free(par); free(rank);
	return ng;
}

static inline double m4(double q)   /* W*pi*h^3 (Monaghan & Lattanzio 1985) */
{
	if (q < 1.0) return 1.0 - 1.5*q*q + 0.75*q*q*q;
	if (q < 2.0) { double t = 2.0 - q; return 0.25*t*t*t; }
	return 0.0;
}

int or_density_f64(int n, const float *pos, const float *mass, const float *fBall2,
		float fPeriod, int bPeriodic, double *rho)
{
	int i, j;
	for (i = 0; i < n; ++i) rho[i] = 0.0;
	for (i = 0; i < n; ++i) {
		double h = 0.5*sqrt((double)fBall2[i]), w0 = 1.0/(M_PI*h*h*h);
		for (j = 0; j < n; ++j) {
			double dx = (double)pos[3*i] - pos[3*j], dy = (double)pos[3*i+1] - pos[3*j+1],
			       dz = (double)pos[3*i+2] - pos[3*j+2], q;
			if (pair_d2(&pos[3*i], &pos[3*j], fPeriod, bPeriodic) >= fBall2[i]) continue;
			if (bPeriodic) {
				dx -= fPeriod*floor(dx/fPeriod + 0.5);
				dy -= fPeriod*floor(dy/fPeriod + 0.5);
				dz -= fPeriod*floor(dz/fPeriod + 0.5);
			}
			q = sqrt(dx*dx + dy*dy + dz*dz)/h;
			rho[i] += 0.5*mass[j]*w0*m4(q);
			rho[j] += 0.5*mass[i]*w0*m4(q);
		}
	}
	return 0;
}

int or_gradient_f64(int ns, const float *spos, const float *smass, const float *sball2,
		int nq, const float *qpos, float fPeriod, int bPeriodic, double *acc)
{
	int i, s;
	for (i = 0; i < 3*nq; ++i) acc[i] = 0.0;
	for (i = 0; i < nq; ++i)
		for (s = 0; s < ns; ++s) {
			double d[3], r, h, q, g, fn;
			int j;
			if (pair_d2(&spos[3*s], &qpos[3*i], fPeriod, bPeriodic) >= sball2[s]) continue;
			for (j = 0; j < 3; ++j) {
				d[j] = (double)spos[3*s+j] - qpos[3*i+j];
				if (bPeriodic) d[j] -= fPeriod*floor(d[j]/fPeriod + 0.5);
			}
			r = sqrt(d[0]*d[0] + d[1]*d[1] + d[2]*d[2]);
			h = 0.5*sqrt((double)sball2[s]);
			q = r/h;
			if (q <= 0.0) continue;
			g = (q < 1.0) ? (-3.0 + 2.25*q) : (-3.0/q + 3.0 - 0.75*q);
			fn = smass[s]/(M_PI*h*h*h*h*h);
			for (j = 0; j < 3; ++j) acc[3*i+j] += d[j]*g*fn;
		}
	return 0;
}

int or_potential_f64(int n, const float *pos, const float *mass, const float *soft,
		float G, int iSoftType, double *phi)
{
	int i, j;
	for (i = 0; i < n; ++i) phi[i] = 0.0;
	for (i = 0; i < n; ++i)
		for (j = 0; j < n; ++j) {
			double dx, dy, dz, r, twoh, k;
			if (i == j) continue;
			dx = (double)pos[3*i] - pos[3*j]; dy = (double)pos[3*i+1] - pos[3*j+1];
			dz = (double)pos[3*i+2] - pos[3*j+2];
			r = sqrt(dx*dx + dy*dy + dz*dz);
			twoh = (double)soft[i] + soft[j];
			if (iSoftType == OR_PLUMMER) k = 1.0/sqrt(r*r + 0.25*twoh*twoh);
			else if (r < twoh) {
				double dih = 2.0/twoh, u = r*dih;
				if (u < 1.0) k = dih*(1.4 - (2.0/3.0)*u*u + 0.3*u*u*u*u - 0.1*u*u*u*u*u);
				else k = -1.0/(15.0*r) + dih*(1.6 - (4.0/3.0)*u*u + u*u*u - 0.3*u*u*u*u + u*u*u*u*u/30.0);
			} else k = 1.0/r;
			phi[i] += (double)G*mass[j]*k;
		}
	return 0;
}
