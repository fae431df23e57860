// skg_group.hip -- group bookkeeping and iterative unbinding.
//
//   skg_centers    kdInitpGroup + kdCalcCenter (kd.c:969-1075)
//   skg_unbind     kdUnbind (kd.c:1299-1466) with kdCellPot / kdSubPot /
//                  kdAddScoopPot (grav.c:8-135)
//   skg_too_small  kdTooSmall (kd.c:1251-1293)
//
// Members are kept group-contiguous by one radix sort on (group, iOrder), the
// analogue of kdGroupOrder (kd.c:1205-1245).  The pairwise potential runs in
// fp32 with compensated sums (one wavefront per 64 targets, sources streamed
// as uniform loads); the energies, the per-group centre-of-mass updates and
// the bound/unbound sign test run in fp64.
#include "skg_common.h"
#include <algorithm>
#include <cfloat>
#include <climits>
#include <vector>

static inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

struct GroupState {
	int nG = 0;                       // groups including group 0
	int *gstart = nullptr, *gend = nullptr;   // member segment of each group in mem[]
	int *mem = nullptr;               // iOrder of members, sorted by (group, iOrder)
	float *mass = nullptr;            // per group
	float4 *rel = nullptr, *cen = nullptr, *vcm = nullptr, *rbound = nullptr;
	float *radius = nullptr;
	int *count = nullptr;             // current member counts
	float4 *mpos_ord = nullptr;       // moved positions by iOrder (original for non-movers)
	// unbinding workspace, aligned with mem[]
	float4 *uq = nullptr;             // relative position, mass
	float4 *uv = nullptr;             // velocity, softening
	float2 *uphi = nullptr;           // compensated potential
	int *uord = nullptr;
	int *d_nunb = nullptr;
	int unbound = 0;
};

void skg_group_free(skg_ctx *c) {
	GroupState *g = c->gs;
	if (!g) return;
	dfree(&g->gstart); dfree(&g->gend); dfree(&g->mem); dfree(&g->mass); dfree(&g->rel); dfree(&g->cen);
	dfree(&g->vcm); dfree(&g->rbound); dfree(&g->radius); dfree(&g->count); dfree(&g->mpos_ord);
	dfree(&g->uq); dfree(&g->uv); dfree(&g->uphi); dfree(&g->uord); dfree(&g->d_nunb);
	delete g;
	c->gs = nullptr;
}

// ---------------------------------------------------------------------------
__global__ void k_gkeys(const int *__restrict__ grp, int n, unsigned long long *key, int *val) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	key[i] = ((unsigned long long)(unsigned)grp[i] << 32) | (unsigned)i;
	val[i] = i;
}
__global__ void k_izero(int *a, int n) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) a[i] = 0;
}
__global__ void k_seg(const unsigned long long *__restrict__ key, int n, int *gs, int *ge) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int g = (int)(key[i] >> 32);
	if (i == 0 || (int)(key[i - 1] >> 32) != g) gs[g] = i;
	if (i == n - 1 || (int)(key[i + 1] >> 32) != g) ge[g] = i + 1;
}
__global__ void k_count(const int *gs, const int *ge, int nG, int *cnt) {
	int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g < nG) cnt[g] = ge[g] - gs[g];
}
__global__ void k_mpos_ord(const float *__restrict__ pos, int n, const int *__restrict__ mord, int nm,
		const float4 *__restrict__ mpos, float4 *out) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) out[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.f);
	(void)mord; (void)nm; (void)mpos;
}
__global__ void k_mpos_scatter(const int *__restrict__ mord, int nm, const float4 *__restrict__ mpos, float4 *out) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < nm) out[mord[i]] = mpos[i];
}

__device__ __forceinline__ float wrap_rel(float d, float L) {   // kd.c:1033-1035 (double compares)
	if ((double)d > 0.5 * (double)L) d = fsub(d, L);
	if ((double)d <= -0.5 * (double)L) d = fadd(d, L);
	return d;
}

// One wavefront per group g >= 1: mass, reference point (the member with the
// largest iOrder), density centre of the moved positions and mass-weighted
// velocity (kd.c:990-1059).
__global__ __launch_bounds__(64) void k_centers(const int *__restrict__ gs, const int *__restrict__ ge, int nG,
		const int *__restrict__ mem, const float *__restrict__ pos, const float *__restrict__ vel,
		const float *__restrict__ mass, const float4 *__restrict__ mpos, float Lx, float Ly, float Lz,
		float cx, float cy, float cz, float *gmass, float4 *grel, float4 *gcen, float4 *gvcm) {
	const int g = blockIdx.x + 1;
	if (g >= nG) return;
	const int lane = threadIdx.x, s = gs[g], e = ge[g], n = e - s;
	if (n <= 0) return;
	const int last = mem[e - 1];
	const float rx = pos[3 * last], ry = pos[3 * last + 1], rz = pos[3 * last + 2];
	Acc am, ax, ay, az, vx, vy, vz;
	am.init(); ax.init(); ay.init(); az.init(); vx.init(); vy.init(); vz.init();
	for (int t = s + lane; t < e; t += 64) {
		const int o = mem[t];
		const float m = mass[o];
		const float4 p = mpos[o];
		am.add(m);
		ax.add(wrap_rel(fsub(p.x, rx), Lx));
		ay.add(wrap_rel(fsub(p.y, ry), Ly));
		az.add(wrap_rel(fsub(p.z, rz), Lz));
		vx.add(fmul(m, vel[3 * o])); vy.add(fmul(m, vel[3 * o + 1])); vz.add(fmul(m, vel[3 * o + 2]));
	}
	float v[7] = {am.s, ax.s, ay.s, az.s, vx.s, vy.s, vz.s};
	float w[7] = {am.c, ax.c, ay.c, az.c, vx.c, vy.c, vz.c};
	for (int k = 0; k < 7; ++k) {
		// exact-ish wave sum of (s, c) pairs through double
		double d = (double)v[k] + (double)w[k];
		for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o, 64);
		v[k] = (float)d;
	}
	if (lane != 0) return;
	const float M = v[0];
	float c3[3] = {v[1], v[2], v[3]}, r3[3] = {rx, ry, rz};
	const float C[3] = {cx, cy, cz}, L[3] = {Lx, Ly, Lz};
	for (int j = 0; j < 3; ++j) {
		float q = c3[j] / (float)n;
		q = fadd(q, r3[j]);
		if ((double)q > (double)C[j] + 0.5 * (double)L[j]) q = fsub(q, L[j]);
		if ((double)q <= (double)C[j] - 0.5 * (double)L[j]) q = fadd(q, L[j]);
		c3[j] = q;
	}
	gmass[g] = M;
	grel[g] = make_float4(rx, ry, rz, 0.f);
	gcen[g] = make_float4(c3[0], c3[1], c3[2], 0.f);
	gvcm[g] = make_float4(v[4] / M, v[5] / M, v[6] / M, 0.f);
}

// ---------------------------------------------------------------------------
// Softened pair potentials (grav.h:11-31, grav.c:25-27) in fp32.
__device__ __forceinline__ float pot_kernel(int soft, float d2, float twoh) {
	if (soft == SKG_PLUMMER) return 1.0f / __fsqrt_rn(fadd(d2, fmul(fmul(0.25f, twoh), twoh)));
	const float r = __fsqrt_rn(d2);
	if (r < twoh) {
		const float dih = 2.0f / twoh, u = fmul(r, dih), u2 = fmul(u, u);
		if (u < 1.0f) {
			float p = fmaf(-0.1f, u, 0.3f);
			p = fmaf(p, u2, -2.0f / 3.0f);
			p = fmaf(p, u2, 1.4f);
			return fmul(dih, p);
		}
		float p = fmaf(1.0f / 30.0f, u, -0.3f);
		p = fmaf(p, u, 1.0f);
		p = fmaf(p, u, -4.0f / 3.0f);
		p = fmaf(p, u2, 1.6f);
		return fsub(fmul(dih, p), 1.0f / (15.0f * r));
	}
	return 1.0f / r;
}

// Relative coordinates and unbinding workspace for groups to unbind (kd.c:1341-1355).
__global__ void k_unb_setup(const int *__restrict__ gs, const int *__restrict__ ge, int nG,
		const int *__restrict__ mem, const float *__restrict__ pos, const float *__restrict__ vel,
		const float *__restrict__ mass, const float *__restrict__ soft, const float4 *__restrict__ grel,
		float hx, float hy, float hz, int nMax, float4 *uq, float4 *uv, float2 *uphi, int *uord, int nTot) {
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= nTot) return;
	// locate the group by binary search over segment starts
	int lo = 1, hi = nG - 1;
	while (lo < hi) { int mid = (lo + hi + 1) >> 1; if (gs[mid] <= t) lo = mid; else hi = mid - 1; }
	const int g = lo;
	if (g < 1 || t < gs[g] || t >= ge[g]) return;
	const int o = mem[t];
	const float4 r = grel[g];
	float dx = fsub(pos[3 * o], r.x), dy = fsub(pos[3 * o + 1], r.y), dz = fsub(pos[3 * o + 2], r.z);
	if (dx > hx) dx = fsub(dx, 2 * hx);
	if (dx <= -hx) dx = fadd(dx, 2 * hx);
	if (dy > hy) dy = fsub(dy, 2 * hy);
	if (dy <= -hy) dy = fadd(dy, 2 * hy);
	if (dz > hz) dz = fsub(dz, 2 * hz);
	if (dz <= -hz) dz = fadd(dz, 2 * hz);
	uq[t] = make_float4(dx, dy, dz, mass[o]);
	uv[t] = make_float4(vel[3 * o], vel[3 * o + 1], vel[3 * o + 2], soft[o]);
	uphi[t] = make_float2(0.f, 0.f);
	uord[t] = o;
	(void)nMax;
}

// phi_i = sum_{j != i} G m_j K(r_ij; eps_i + eps_j) over the member's group
// (kdCellPot, grav.c:8-36; both orders of each pair evaluated, no atomics).
__global__ __launch_bounds__(64) void k_unb_pot(const int2 *__restrict__ tiles, const int *__restrict__ gs,
		const int *__restrict__ ge, const float4 *__restrict__ uq, const float4 *__restrict__ uv,
		float2 *uphi, float G, int soft) {
	const int2 tl = tiles[blockIdx.x];   // (group, first target)
	const int g = tl.x, s = gs[g], e = ge[g];
	const int i = tl.y + threadIdx.x;
	const bool valid = i < e;
	const float4 qi = uq[valid ? i : s], vi = uv[valid ? i : s];
	Acc acc;
	acc.init();
	for (int j = s; j < e; ++j) {
		const float4 qj = uq[j];
		const float ej = uv[j].w;
		const float dx = fsub(qi.x, qj.x), dy = fsub(qi.y, qj.y), dz = fsub(qi.z, qj.z);
		const float k = pot_kernel(soft, sq3(dx, dy, dz), fadd(vi.w, ej));
		if (j != i) acc.add(fmul(fmul(G, qj.w), k));
	}
	if (valid) uphi[i] = make_float2(acc.s, acc.c);
}

// Scoop: group-0 particles within fScoop of the group centre add their
// potential to every member (kdAddScoopPot, grav.c:63-135).  One wavefront
// per group walks the group-0 tree; lanes split each bucket, then the members.
__global__ __launch_bounds__(64) void k_unb_scoop(const int *__restrict__ gs, const int *__restrict__ ge,
		int nG, const float4 *__restrict__ gcen, const float4 *__restrict__ grel, const float4 *__restrict__ zp,
		const float *__restrict__ zeps, int nz, const float4 *__restrict__ nlo, const float4 *__restrict__ nhi, int nLeaf2, int B,
		float L, int per, float scoop2, const float4 *__restrict__ uq, const float4 *__restrict__ uv,
		float2 *uphi, float G, int soft, int nMax) {
	const int g = blockIdx.x + 1;
	if (g >= nG) return;
	const int s = gs[g], e = ge[g];
	if (e - s <= 0 || e - s >= nMax) return;
	const int lane = threadIdx.x;
	const float4 c = gcen[g], rl = grel[g];
	const float hL = 0.5f * L;
	const float cxm = fsub(c.x, L), cxp = fadd(c.x, L), cym = fsub(c.y, L), cyp = fadd(c.y, L);
	const float czm = fsub(c.z, L), czp = fadd(c.z, L);
	int node = 1;
	for (;;) {
		const float4 lo = nlo[node], hi = nhi[node];
		float gx = per ? gap_axis<true>(lo.x, hi.x, c.x, c.x, L) : gap_axis<false>(lo.x, hi.x, c.x, c.x, L);
		float gy = per ? gap_axis<true>(lo.y, hi.y, c.y, c.y, L) : gap_axis<false>(lo.y, hi.y, c.y, c.y, L);
		float gz = per ? gap_axis<true>(lo.z, hi.z, c.z, c.z, L) : gap_axis<false>(lo.z, hi.z, c.z, c.z, L);
		const bool visit = sq3(gx, gy, gz) < scoop2;
		if (visit && node < nLeaf2) { node = 2 * node; continue; }
		if (visit) {
			const int st = (node - nLeaf2) * B;
			for (int t0 = st; t0 < min(nz, st + B); t0 += 64) {
				const int j = t0 + lane;
				bool hit = false;
				float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
				float pe = 0.f;
				if (j < min(nz, st + B)) {
					p = zp[j];
					pe = zeps[j];
					float dx = per ? img_delta<true>(c.x, cxm, cxp, p.x, hL) : fsub(c.x, p.x);
					float dy = per ? img_delta<true>(c.y, cym, cyp, p.y, hL) : fsub(c.y, p.y);
					float dz = per ? img_delta<true>(c.z, czm, czp, p.z, hL) : fsub(c.z, p.z);
					hit = sq3(dx, dy, dz) < scoop2;
				}
				unsigned long long b = __ballot(hit);
				while (b) {
					const int src = __ffsll((long long)b) - 1;
					b &= b - 1;
					const float px = __shfl(p.x, src, 64), py = __shfl(p.y, src, 64), pz = __shfl(p.z, src, 64);
					const float pm = __shfl(p.w, src, 64);
					const float peps = __shfl(pe, src, 64);
					float nx = wrap_rel(fsub(px, rl.x), L), ny = wrap_rel(fsub(py, rl.y), L), nz2 = wrap_rel(fsub(pz, rl.z), L);
					for (int i = s + lane; i < e; i += 64) {
						const float4 q = uq[i];
						const float dx = fsub(nx, q.x), dy = fsub(ny, q.y), dz = fsub(nz2, q.z);
						const float k = pot_kernel(soft, sq3(dx, dy, dz), fadd(peps, uv[i].w));
						float2 ph = uphi[i];
						Acc a; a.s = ph.x; a.c = ph.y;
						a.add(fmul(fmul(G, pm), k));
						uphi[i] = make_float2(a.s, a.c);
					}
				}
			}
		}
		int nx = node;
		while (nx & 1) nx >>= 1;
		node = nx + 1;
		if (node == 1) break;
	}
}

// Iterative removal of the most energetic member (kd.c:1385-1447), one
// workgroup per group.  Energies E = |fShift (v - vcm) + fCosmo (r - rcm)|^2/2
// - phi (1+z) and the sign test are fp64.
#define UNB_T 256
__global__ __launch_bounds__(UNB_T) void k_unb_loop(const int *__restrict__ glist, const int *__restrict__ gs,
		const int *__restrict__ ge, float4 *uq, float4 *uv, float2 *uphi, int *uord, int *grp,
		int *gcount, float *gmass, float4 *gvcm, float4 *grbound, const float4 *__restrict__ grel,
		float fShift, float fCosmo, float z, float G, int soft, int bSub, int bNoUnbind, float hx, int *nunb) {
	__shared__ double sd[UNB_T];
	__shared__ int si[UNB_T];
	__shared__ double smin[UNB_T];
	__shared__ int simin[UNB_T];
	__shared__ double cm[7];   // dMass, rcm[3], vcm[3]
	__shared__ int sbig, smn;
	__shared__ double sbigE;
	const int g = glist[blockIdx.x], tid = threadIdx.x;
	const int s = gs[g];
	int n = ge[g] - s;
	// centre of mass and velocity in fp64 (kd.c:1360-1375)
	double a[7] = {0, 0, 0, 0, 0, 0, 0};
	for (int i = tid; i < n; i += UNB_T) {
		const float4 q = uq[s + i], v = uv[s + i];
		a[0] += (double)q.w;
		a[1] += (double)fmul(q.w, q.x); a[2] += (double)fmul(q.w, q.y); a[3] += (double)fmul(q.w, q.z);
		a[4] += (double)fmul(q.w, v.x); a[5] += (double)fmul(q.w, v.y); a[6] += (double)fmul(q.w, v.z);
	}
	for (int k = 0; k < 7; ++k) {
		sd[tid] = a[k];
		__syncthreads();
		for (int o = UNB_T / 2; o > 0; o >>= 1) {
			if (tid < o) sd[tid] += sd[tid + o];
			__syncthreads();
		}
		if (tid == 0) cm[k] = sd[0];
		__syncthreads();
	}
	if (tid == 0) for (int k = 1; k < 7; ++k) cm[k] /= cm[0];
	__syncthreads();
	int removed = 0;
	const double zf = 1.0 + (double)z;
	for (;;) {
		double bigE = -INFINITY, minE = INFINITY;
		int bi = INT_MAX, mi = INT_MAX;
		const double rc0 = cm[1], rc1 = cm[2], rc2 = cm[3], vc0 = cm[4], vc1 = cm[5], vc2 = cm[6];
		for (int i = tid; i < n; i += UNB_T) {
			const float4 q = uq[s + i], v = uv[s + i];
			const float2 ph = uphi[s + i];
			const double d0 = (double)fShift * ((double)v.x - vc0) + (double)fCosmo * ((double)q.x - rc0);
			const double d1 = (double)fShift * ((double)v.y - vc1) + (double)fCosmo * ((double)q.y - rc1);
			const double d2 = (double)fShift * ((double)v.z - vc2) + (double)fCosmo * ((double)q.z - rc2);
			const double E = 0.5 * (d0 * d0 + d1 * d1 + d2 * d2) - ((double)ph.x + (double)ph.y) * zf;
			if (E > bigE || (E == bigE && i < bi)) { bigE = E; bi = i; }
			if (E < minE || (E == minE && i < mi)) { minE = E; mi = i; }
		}
		sd[tid] = bigE; si[tid] = bi; smin[tid] = minE; simin[tid] = mi;
		__syncthreads();
		for (int o = UNB_T / 2; o > 0; o >>= 1) {
			if (tid < o) {
				if (sd[tid + o] > sd[tid] || (sd[tid + o] == sd[tid] && si[tid + o] < si[tid])) {
					sd[tid] = sd[tid + o]; si[tid] = si[tid + o];
				}
				if (smin[tid + o] < smin[tid] || (smin[tid + o] == smin[tid] && simin[tid + o] < simin[tid])) {
					smin[tid] = smin[tid + o]; simin[tid] = simin[tid + o];
				}
			}
			__syncthreads();
		}
		if (tid == 0) { sbig = si[0]; sbigE = sd[0]; smn = simin[0]; }
		__syncthreads();
		const int ib = sbig;
		if (!(sbigE >= 0.0) || bNoUnbind || n == 0) break;
		// unbind member ib (kd.c:1413-1446)
		const float4 qb = uq[s + ib], vb = uv[s + ib];
		const float2 pb = uphi[s + ib];
		const int ob = uord[s + ib];
		__syncthreads();
		--n;
		++removed;
		if (tid == 0) {
			grp[ob] = 0;
			if (n > 0) {
				cm[0] -= (double)qb.w;
				const double f = (double)qb.w / cm[0];
				cm[1] += f * (cm[1] - (double)qb.x); cm[2] += f * (cm[2] - (double)qb.y); cm[3] += f * (cm[3] - (double)qb.z);
				cm[4] += f * (cm[4] - (double)vb.x); cm[5] += f * (cm[5] - (double)vb.y); cm[6] += f * (cm[6] - (double)vb.z);
				// swap the removed member with the last (kd.c:1435-1440)
				uq[s + ib] = uq[s + n]; uv[s + ib] = uv[s + n]; uphi[s + ib] = uphi[s + n]; uord[s + ib] = uord[s + n];
				uq[s + n] = qb; uv[s + n] = vb; uphi[s + n] = pb; uord[s + n] = ob;
			} else {
				cm[0] = 0.0; cm[4] = cm[5] = cm[6] = 0.0;
			}
		}
		__syncthreads();
		if (n == 0) { if (tid == 0) smn = 0; break; }
		if (bSub) {                                                    // kdSubPot (grav.c:39-60)
			for (int i = tid; i < n; i += UNB_T) {
				const float4 q = uq[s + i];
				const float dx = fsub(qb.x, q.x), dy = fsub(qb.y, q.y), dz = fsub(qb.z, q.z);
				const float k = pot_kernel(soft, sq3(dx, dy, dz), fadd(vb.w, uv[s + i].w));
				float2 ph = uphi[s + i];
				Acc ac; ac.s = ph.x; ac.c = ph.y;
				ac.add(-fmul(fmul(G, qb.w), k));
				uphi[s + i] = make_float2(ac.s, ac.c);
			}
		}
		__syncthreads();
	}
	if (tid == 0) {
		gcount[g] = n;
		gmass[g] = (float)cm[0];
		gvcm[g] = make_float4((float)cm[4], (float)cm[5], (float)cm[6], 0.f);
		const float4 q = uq[s + smn], r = grel[g];
		float b[3] = {fadd(q.x, r.x), fadd(q.y, r.y), fadd(q.z, r.z)};
		for (int j = 0; j < 3; ++j) {                                    // hx on all axes (kd.c:1452-1456)
			if (b[j] > hx) b[j] = fsub(b[j], 2 * hx);
			if (b[j] <= -hx) b[j] = fadd(b[j], 2 * hx);
		}
		grbound[g] = make_float4(b[0], b[1], b[2], 0.f);
		atomicAdd(nunb, removed);
	}
}

// ---------------------------------------------------------------------------
__global__ void k_keep_flags(const int *__restrict__ cnt, int nG, int nMin, int *keep) {
	int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= nG) return;
	keep[g] = (g > 0 && cnt[g] >= nMin) ? 1 : 0;
}
__global__ void k_remap(int *grp, int n, const int *__restrict__ keep, const int *__restrict__ off) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int g = grp[i];
	grp[i] = (g > 0 && keep[g]) ? off[g] + 1 : 0;
}
template <typename T>
__global__ void k_compact_rec(const T *__restrict__ in, T *out, int nG, const int *__restrict__ keep,
		const int *__restrict__ off) {
	int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g < nG && g > 0 && keep[g]) out[off[g] + 1] = in[g];
}
__global__ void k_radius(const int *__restrict__ grp, int n, const float *__restrict__ pos,
		const float4 *__restrict__ gcen, float Lx, float Ly, float Lz, unsigned *rad) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int g = grp[i];
	if (!g) return;
	float4 c = gcen[g];
	float dx = wrap_rel(fsub(pos[3 * i], c.x), Lx), dy = wrap_rel(fsub(pos[3 * i + 1], c.y), Ly);
	float dz = wrap_rel(fsub(pos[3 * i + 2], c.z), Lz);
	float r = (float)sqrt((double)sq3(dx, dy, dz));
	atomicMax(&rad[g], __float_as_uint(r));
}

__global__ void k_gather_pos_members(const int *__restrict__ mem, int n, const float *__restrict__ pos,
		const float *__restrict__ mass, const float *__restrict__ soft, float4 *out, float *eps) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int o = mem[i];
	out[i] = make_float4(pos[3 * o], pos[3 * o + 1], pos[3 * o + 2], mass[o]);
	eps[i] = soft[o];
}
__global__ void k_gather4g(const float4 *__restrict__ in, const float *__restrict__ ein,
		const int *__restrict__ perm, int n, float4 *out, float *eout) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) { out[i] = in[perm[i]]; eout[i] = ein[perm[i]]; }
}


// ---------------------------------------------------------------------------
static int group_sort(skg_ctx *c) {
	GroupState *G = c->gs;
	int rc;
	const int n = c->n;
	if (skg_ensure_keys(c, n)) return SKG_E_NOMEM;
	k_gkeys<<<nblk(n, 256), 256, 0, c->st>>>(c->d_grp, n, c->k64a, c->v32a);
	if ((rc = skg_sort_pairs(c, c->k64a, c->k64b, c->v32a, G->mem, n, 64))) return rc;
	k_izero<<<nblk(G->nG, 256), 256, 0, c->st>>>(G->gstart, G->nG);
	k_izero<<<nblk(G->nG, 256), 256, 0, c->st>>>(G->gend, G->nG);
	k_seg<<<nblk(n, 256), 256, 0, c->st>>>(c->k64b, n, G->gstart, G->gend);
	k_count<<<nblk(G->nG, 256), 256, 0, c->st>>>(G->gstart, G->gend, G->nG, G->count);
	SKG_HIP(hipGetLastError());
	return 0;
}

extern "C" int skg_centers(skg_ctx *c) {
	if (!c) return SKG_E_PARAM;
	if (c->stage != 6) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	skg_group_free(c);
	GroupState *G = c->gs = new (std::nothrow) GroupState();
	if (!G) return SKG_E_NOMEM;
	int rc;
	const int nG = std::max(1, c->nGroup), n = c->n;
	G->nG = nG;
	if (dalloc(&G->gstart, nG) || dalloc(&G->gend, nG) || dalloc(&G->mem, n) || dalloc(&G->mass, nG) ||
	    dalloc(&G->rel, nG) || dalloc(&G->cen, nG) || dalloc(&G->vcm, nG) || dalloc(&G->rbound, nG) ||
	    dalloc(&G->radius, nG) || dalloc(&G->count, nG) || dalloc(&G->mpos_ord, n) || dalloc(&G->d_nunb, 1))
		return SKG_E_NOMEM;
	skg_timer_start(c);
	SKG_HIP(hipMemsetAsync(G->mass, 0, 4 * (size_t)nG, c->st));
	SKG_HIP(hipMemsetAsync(G->rel, 0, 16 * (size_t)nG, c->st));
	SKG_HIP(hipMemsetAsync(G->cen, 0, 16 * (size_t)nG, c->st));
	SKG_HIP(hipMemsetAsync(G->vcm, 0, 16 * (size_t)nG, c->st));
	SKG_HIP(hipMemsetAsync(G->rbound, 0, 16 * (size_t)nG, c->st));
	if ((rc = group_sort(c))) return rc;
	k_mpos_ord<<<nblk(n, 256), 256, 0, c->st>>>(c->d_pos, n, c->m_ord, c->nMove, c->m_pos, G->mpos_ord);
	if (c->nMove > 0)
		k_mpos_scatter<<<nblk(c->nMove, 256), 256, 0, c->st>>>(c->m_ord, c->nMove, c->m_pos, G->mpos_ord);
	const float *L = c->P.fPeriod, *C = c->P.fCenter;
	if (nG > 1)
		k_centers<<<nG - 1, 64, 0, c->st>>>(G->gstart, G->gend, nG, G->mem, c->d_pos, c->d_vel, c->d_mass,
				G->mpos_ord, L[0], L[1], L[2], C[0], C[1], C[2], G->mass, G->rel, G->cen, G->vcm);
	SKG_HIP(hipGetLastError());
	SKG_HIP(hipStreamSynchronize(c->st));
	skg_timer_stop(c, SKG_T_GROUP);
	c->stage = 7;
	return SKG_OK;
}

extern "C" int skg_unbind(skg_ctx *c, int *nUnbound) {
	if (!c) return SKG_E_PARAM;
	if (c->stage != 7 || !c->gs) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	GroupState *G = c->gs;
	const int n = c->n, nG = G->nG;
	const int nMax = c->P.nMaxMembers > 0 ? c->P.nMaxMembers : INT_MAX;
	int rc;
	skg_timer_start(c);
	std::vector<int> hs(nG), he(nG);
	SKG_HIP(hipMemcpyAsync(hs.data(), G->gstart, 4 * (size_t)nG, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipMemcpyAsync(he.data(), G->gend, 4 * (size_t)nG, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	const int n0 = he[0] - hs[0];
	// members of groups 1.. occupy mem[n0, n)
	const int nTot = n;
	if (dalloc(&G->uq, nTot) || dalloc(&G->uv, nTot) || dalloc(&G->uphi, nTot) || dalloc(&G->uord, nTot))
		return SKG_E_NOMEM;
	const float *L = c->P.fPeriod;
	const float hx = 0.5f * L[0], hy = 0.5f * L[1], hz = 0.5f * L[2];
	std::vector<int2> tiles;
	std::vector<int> glist;
	for (int g = 1; g < nG; ++g) {
		const int m = he[g] - hs[g];
		if (m <= 0 || m >= nMax) continue;
		glist.push_back(g);
		for (int t = hs[g]; t < he[g]; t += 64) tiles.push_back(make_int2(g, t));
	}
	SKG_HIP(hipMemsetAsync(G->d_nunb, 0, 4, c->st));
	if (!glist.empty()) {
		int2 *dt = nullptr;
		int *dg = nullptr;
		if (dalloc(&dt, tiles.size()) || dalloc(&dg, glist.size())) return SKG_E_NOMEM;
		SKG_HIP(hipMemcpyAsync(dt, tiles.data(), 8 * tiles.size(), hipMemcpyHostToDevice, c->st));
		SKG_HIP(hipMemcpyAsync(dg, glist.data(), 4 * glist.size(), hipMemcpyHostToDevice, c->st));
		k_unb_setup<<<nblk(nTot, 256), 256, 0, c->st>>>(G->gstart, G->gend, nG, G->mem, c->d_pos, c->d_vel,
				c->d_mass, c->d_soft, G->rel, hx, hy, hz, nMax, G->uq, G->uv, G->uphi, G->uord, nTot);
		k_unb_pot<<<(int)tiles.size(), 64, 0, c->st>>>(dt, G->gstart, G->gend, G->uq, G->uv, G->uphi, c->P.G,
				c->P.iSoftType);
		// scoop potential from group 0 (always applied by the reference: kd.c:1379-1384)
		if (n0 > 0) {
			float4 *zp = nullptr, *zs = nullptr;
			float *ze = nullptr, *zes = nullptr;
			int *perm = nullptr;
			DevTree t;
			if (dalloc(&zp, n0) || dalloc(&zs, n0) || dalloc(&perm, n0) || dalloc(&ze, n0) || dalloc(&zes, n0))
				return SKG_E_NOMEM;
			k_gather_pos_members<<<nblk(n0, 256), 256, 0, c->st>>>(G->mem + hs[0], n0, c->d_pos, c->d_mass,
					c->d_soft, zp, ze);
			if ((rc = skg_morton_sort(c, zp, nullptr, n0, perm))) return rc;
			k_gather4g<<<nblk(n0, 256), 256, 0, c->st>>>(zp, ze, perm, n0, zs, zes);
			if ((rc = skg_build_tree(c, &t, zs, n0, c->P.nBucket, nullptr))) return rc;
			const float sc2 = c->P.fScoop * c->P.fScoop;
			k_unb_scoop<<<nG - 1, 64, 0, c->st>>>(G->gstart, G->gend, nG, G->cen, G->rel, zs, zes, n0, t.lo, t.hi,
					t.nLeaf2, t.B, c->P.bPeriodic ? L[0] : FLT_MAX, c->P.bPeriodic, sc2, G->uq, G->uv, G->uphi,
					c->P.G, c->P.iSoftType, nMax);
			SKG_HIP(hipStreamSynchronize(c->st));
			dfree(&zp); dfree(&zs); dfree(&perm); dfree(&ze); dfree(&zes);
			skg_free_tree(&t);
		}
		const int bSub = (c->inType == SKG_DARK || c->inType == SKG_STAR);   // kd.c:1441
		k_unb_loop<<<(int)glist.size(), UNB_T, 0, c->st>>>(dg, G->gstart, G->gend, G->uq, G->uv, G->uphi, G->uord,
				c->d_grp, G->count, G->mass, G->vcm, G->rbound, G->rel, c->P.fShift, c->P.fCosmo, c->P.z,
				c->P.G, c->P.iSoftType, bSub, c->P.bNoUnbind, hx, G->d_nunb);
		SKG_HIP(hipGetLastError());
		SKG_HIP(hipStreamSynchronize(c->st));
		dfree(&dt); dfree(&dg);
	}
	int nu = 0;
	SKG_HIP(hipMemcpy(&nu, G->d_nunb, 4, hipMemcpyDeviceToHost));
	G->unbound = nu;
	// group 0 gains the unbound members
	int c0 = 0;
	SKG_HIP(hipMemcpy(&c0, G->count, 4, hipMemcpyDeviceToHost));
	c0 += nu;
	SKG_HIP(hipMemcpy(G->count, &c0, 4, hipMemcpyHostToDevice));
	skg_timer_stop(c, SKG_T_UNBIND);
	if (nUnbound) *nUnbound = nu;
	c->stage = 8;
	return SKG_OK;
}

extern "C" int skg_too_small(skg_ctx *c) {
	if (!c) return SKG_E_PARAM;
	if (c->stage != 8 || !c->gs) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	GroupState *G = c->gs;
	const int nG = G->nG, n = c->n;
	int rc, *keep = nullptr, *off = nullptr;
	if (dalloc(&keep, nG) || dalloc(&off, nG)) return SKG_E_NOMEM;
	k_keep_flags<<<nblk(nG, 256), 256, 0, c->st>>>(G->count, nG, c->P.nMembers, keep);
	if ((rc = skg_exclusive_scan(c, keep, off, nG))) return rc;
	int last[2];
	SKG_HIP(hipMemcpyAsync(&last[0], off + nG - 1, 4, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipMemcpyAsync(&last[1], keep + nG - 1, 4, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	const int nKeep = last[0] + last[1];
	k_remap<<<nblk(n, 256), 256, 0, c->st>>>(c->d_grp, n, keep, off);
	float *m2 = nullptr;
	float4 *r2 = nullptr, *c2 = nullptr, *v2 = nullptr, *b2 = nullptr;
	const int nN = nKeep + 1;
	if (dalloc(&m2, nN) || dalloc(&r2, nN) || dalloc(&c2, nN) || dalloc(&v2, nN) || dalloc(&b2, nN))
		return SKG_E_NOMEM;
	SKG_HIP(hipMemsetAsync(m2, 0, 4 * (size_t)nN, c->st));
	SKG_HIP(hipMemsetAsync(r2, 0, 16 * (size_t)nN, c->st));
	SKG_HIP(hipMemsetAsync(c2, 0, 16 * (size_t)nN, c->st));
	SKG_HIP(hipMemsetAsync(v2, 0, 16 * (size_t)nN, c->st));
	SKG_HIP(hipMemsetAsync(b2, 0, 16 * (size_t)nN, c->st));
	k_compact_rec<float><<<nblk(nG, 256), 256, 0, c->st>>>(G->mass, m2, nG, keep, off);
	k_compact_rec<float4><<<nblk(nG, 256), 256, 0, c->st>>>(G->rel, r2, nG, keep, off);
	k_compact_rec<float4><<<nblk(nG, 256), 256, 0, c->st>>>(G->cen, c2, nG, keep, off);
	k_compact_rec<float4><<<nblk(nG, 256), 256, 0, c->st>>>(G->vcm, v2, nG, keep, off);
	k_compact_rec<float4><<<nblk(nG, 256), 256, 0, c->st>>>(G->rbound, b2, nG, keep, off);
	SKG_HIP(hipStreamSynchronize(c->st));
	std::swap(G->mass, m2); std::swap(G->rel, r2); std::swap(G->cen, c2); std::swap(G->vcm, v2);
	std::swap(G->rbound, b2);
	dfree(&m2); dfree(&r2); dfree(&c2); dfree(&v2); dfree(&b2); dfree(&keep); dfree(&off);
	G->nG = nN;
	c->nGroup = nN;
	dfree(&G->gstart); dfree(&G->gend); dfree(&G->count);
	if (dalloc(&G->gstart, nN) || dalloc(&G->gend, nN) || dalloc(&G->count, nN)) return SKG_E_NOMEM;
	if ((rc = group_sort(c))) return rc;                        // kdGroupOrder of the final ids
	SKG_HIP(hipStreamSynchronize(c->st));
	c->stage = 9;
	return SKG_OK;
}

extern "C" int skg_get_groups(skg_ctx *c, int *grp, int *nGroup) {
	if (!c) return SKG_E_PARAM;
	if (c->stage < 5) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	if (grp) {
		SKG_HIP(hipMemcpyAsync(grp, c->d_grp, 4 * (size_t)c->n, hipMemcpyDeviceToHost, c->st));
		SKG_HIP(hipStreamSynchronize(c->st));
	}
	if (nGroup) *nGroup = c->nGroup;
	return SKG_OK;
}

extern "C" int skg_get_catalogue(skg_ctx *c, float *mass, float *center, float *vcm, float *rbound,
		float *radius, int *nMembers) {
	if (!c) return SKG_E_PARAM;
	if (c->stage < 7 || !c->gs) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	GroupState *G = c->gs;
	const int nG = G->nG;
	const float *L = c->P.fPeriod;
	SKG_HIP(hipMemsetAsync(G->radius, 0, 4 * (size_t)nG, c->st));
	k_radius<<<nblk(c->n, 256), 256, 0, c->st>>>(c->d_grp, c->n, c->d_pos, G->cen, L[0], L[1], L[2],
			(unsigned *)G->radius);
	std::vector<float4> h4(nG);
	auto get4 = [&](const float4 *d, float *out) -> int {
		if (!out) return 0;
		SKG_HIP(hipMemcpyAsync(h4.data(), d, 16 * (size_t)nG, hipMemcpyDeviceToHost, c->st));
		SKG_HIP(hipStreamSynchronize(c->st));
		for (int g = 0; g < nG; ++g) { out[3 * g] = h4[g].x; out[3 * g + 1] = h4[g].y; out[3 * g + 2] = h4[g].z; }
		return 0;
	};
	int rc;
	if ((rc = get4(G->cen, center)) || (rc = get4(G->vcm, vcm)) || (rc = get4(G->rbound, rbound))) return rc;
	if (mass) SKG_HIP(hipMemcpyAsync(mass, G->mass, 4 * (size_t)nG, hipMemcpyDeviceToHost, c->st));
	if (radius) SKG_HIP(hipMemcpyAsync(radius, G->radius, 4 * (size_t)nG, hipMemcpyDeviceToHost, c->st));
	if (nMembers) SKG_HIP(hipMemcpyAsync(nMembers, G->count, 4 * (size_t)nG, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	return SKG_OK;
}

