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
#include <hip/hip_cooperative_groups.h>
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cstdlib>
#include <string>
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
	float4 *psrc = nullptr;           // potential sources: x, y, z, G m (aligned with mem[])
	float *peps = nullptr;            // and softening
	long nsrc = 0;
	int *d_epsMax = nullptr;          // largest softening over the workspace (float bits)
	int *d_nunb = nullptr;
	int unbound = 0;
};

void skg_group_free(skg_ctx *c) {
	GroupState *g = c->gs;
	if (!g) return;
	dfree(&g->gstart); dfree(&g->gend); dfree(&g->mem); dfree(&g->mass); dfree(&g->rel); dfree(&g->cen);
	dfree(&g->vcm); dfree(&g->rbound); dfree(&g->radius); dfree(&g->count); dfree(&g->mpos_ord);
	dfree(&g->uq); dfree(&g->uv); dfree(&g->uphi); dfree(&g->uord); dfree(&g->d_nunb);
	dfree(&g->psrc); dfree(&g->peps); dfree(&g->d_epsMax);
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
// Softened pair potentials (grav.h:11-31, grav.c:25-27) in fp32.  1/r comes
// from the hardware reciprocal square root (about 1 ulp; the potentials are
// checked against the FP64 oracle, not for bit parity), so the far field
// r >= eps_i + eps_j -- nearly every pair of a large group -- costs one
// transcendental.  Coincident points (d2 = 0) take the r = 0 branch.
// Spline softening inside the kernel (r < 2h = twoh), given ir = rsq(d2).
__device__ __forceinline__ float pot_spline_in(float d2, float ir, float twoh) {
	const float r = d2 > 0.f ? d2 * ir : 0.f;
	const float dih = 2.0f * __builtin_amdgcn_rcpf(twoh), u = r * dih, u2 = u * u;
	if (u < 1.0f) {
		float p = fmaf(-0.1f, u, 0.3f);
		p = fmaf(p, u2, -2.0f / 3.0f);
		p = fmaf(p, u2, 1.4f);
		return dih * p;
	}
	float p = fmaf(1.0f / 30.0f, u, -0.3f);
	p = fmaf(p, u, 1.0f);
	p = fmaf(p, u, -4.0f / 3.0f);
	p = fmaf(p, u2, 1.6f);
	return fmaf(dih, p, (-1.0f / 15.0f) * ir);
}
__device__ __forceinline__ float pot_kernel(int soft, float d2, float twoh) {
	if (soft == SKG_PLUMMER) return __builtin_amdgcn_rsqf(fmaf(0.25f * twoh, twoh, d2));
	const float ir = __builtin_amdgcn_rsqf(d2);
	if (d2 >= twoh * twoh) return ir;            // outside the kernel (r >= 2h): 1/r
	return pot_spline_in(d2, ir, twoh);
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

// phi_i += sum_{j != i} G m_j K(r_ij; eps_i + eps_j) over the member's group
// (kdCellPot, grav.c:8-36).  Both orders of each pair are evaluated; no
// atomics, fixed summation order.  Summation: the terms are added in fp32
// partials of BLK consecutive sources (by fma) and each partial enters a
// two-sum accumulator, so the rounding of the whole sum stays that of one
// BLK-term partial (at most BLK ulp relative, the terms being positive)
// whatever the group size; BLK = 1 is a two-sum per pair.
//
// Far test: with cfar = fl(fl(eps_i + eps_max)^2), d2 >= cfar implies
// d2 >= fl(fl(eps_i + eps_j)^2) (rounding is monotone), so such a pair is 1/r
// exactly as in pot_kernel, at one compare; the others take pot_kernel's own
// test and spline, sharing the rsq.
template <bool PLUM>
__device__ __forceinline__ float pot_k(float d2, float ei, float ej, float cfar) {
	if (PLUM) return pot_kernel(SKG_PLUMMER, d2, ei + ej);
	const float ir = __builtin_amdgcn_rsqf(d2);
	if (d2 >= cfar) return ir;
	const float twoh = ei + ej;
	if (d2 >= twoh * twoh) return ir;
	return pot_spline_in(d2, ir, twoh);
}

// BLK-blocked sum over sources src(0..n-1) for the lane's target (qi, ei);
// self: the source index of the lane's own target, tested only when SELF.
template <int BLK, bool SELF, bool PLUM, typename Src>
__device__ __forceinline__ void pot_sum(Acc &acc, Src src, int n, float4 qi, float ei, float cfar, int self) {
	auto one = [&](int t, float p, bool first) {
		float4 qj; float ej;
		src(t, qj, ej);
		const float dx = qi.x - qj.x, dy = qi.y - qj.y, dz = qi.z - qj.z;
		const float k = pot_k<PLUM>(fmaf(dx, dx, fmaf(dy, dy, dz * dz)), ei, ej, cfar);
		const float gm = SELF && t == self ? 0.f : qj.w;
		return first ? fmul(gm, k) : fmaf(gm, k, p);
	};
	int t = 0;
	for (; t + BLK <= n; t += BLK) {
		float p = one(t, 0.f, true);
#pragma unroll
		for (int u = 1; u < BLK; ++u) p = one(t + u, p, false);
		acc.add(p);
	}
	if (t < n) {
		float p = one(t, 0.f, true);
		for (++t; t < n; ++t) p = one(t, p, false);
		acc.add(p);
	}
}

// A zero the compiler cannot prove uniform: LDS tile addresses offset by it
// stay in one VGPR per block of sources (with immediate offsets per source)
// instead of being rematerialised from scalar registers for every source.
__device__ __forceinline__ int vgpr_zero() {
	int z;
	asm volatile("v_mov_b32 %0, 0" : "=v"(z));
	return z;
}

// A workgroup holds POT_T targets in registers and streams the group's sources
// (x, y, z, G m; eps) through LDS in tiles of POT_T.  Target and source tiles
// both start at the group's first member, so the self pair lies in the tile
// with j0 == first target.  POT_T (64, 128 or 256) is chosen by the autotuner.
// Far pairs cost 9 VALU operations (3 sub, mul, 2 fma for d2; rsq; compare;
// fma into the partial) and one 16-byte LDS broadcast; eps_j is read only for
// pairs the far test does not settle.
template <int POT_T, int BLK>
__global__ __launch_bounds__(POT_T) void k_unb_pot(const int2 *__restrict__ tiles, const int *__restrict__ gs,
		const int *__restrict__ ge, const float4 *__restrict__ uq, const float4 *__restrict__ uv,
		const float4 *__restrict__ ps, const float *__restrict__ pe, const float *__restrict__ epsMax,
		float2 *__restrict__ uphi, int soft) {
	__shared__ float4 sq[POT_T];
	__shared__ float se[POT_T];
	const int2 tl = tiles[blockIdx.x];   // (group, first target)
	const int g = tl.x, s = gs[g], e = ge[g];
	const int i = tl.y + threadIdx.x;
	const bool valid = i < e;
	const float4 qi = uq[valid ? i : s];
	const float ei = uv[valid ? i : s].w;
	const float cf = fadd(ei, *epsMax), cfar = fmul(cf, cf);
	const int vz = vgpr_zero();
	Acc acc;
	acc.init();
	for (int j0 = s; j0 < e; j0 += POT_T) {
		const int nt = min(POT_T, e - j0);
		__syncthreads();
		if (threadIdx.x < nt) {
			sq[threadIdx.x] = ps[j0 + threadIdx.x];
			se[threadIdx.x] = pe[j0 + threadIdx.x];
		}
		__syncthreads();
		auto src = [&](int t, float4 &q, float &ee) { q = sq[t + vz]; ee = se[t + vz]; };
		const bool diag = j0 == tl.y;
		if (soft == SKG_PLUMMER) {
			if (diag) pot_sum<BLK, true, true>(acc, src, nt, qi, ei, cfar, (int)threadIdx.x);
			else pot_sum<BLK, false, true>(acc, src, nt, qi, ei, cfar, -1);
		} else {
			if (diag) pot_sum<BLK, true, false>(acc, src, nt, qi, ei, cfar, (int)threadIdx.x);
			else pot_sum<BLK, false, false>(acc, src, nt, qi, ei, cfar, -1);
		}
	}
	if (valid) {
		float2 ph = uphi[i];
		Acc a; a.s = ph.x; a.c = ph.y;
		a.add(acc.s); a.add(acc.c);
		uphi[i] = make_float2(a.s, a.c);
	}
}

// Source records of the potential: (x, y, z, G m) and eps of every member, and
// the largest eps (eps >= 0, so its bits order as integers; *epsMax starts at
// 0).  Grid-stride with one atomic per workgroup: an atomic per wavefront to
// the one address serialises across the chip.
__global__ __launch_bounds__(256) void k_pot_src(const float4 *__restrict__ uq, const float4 *__restrict__ uv,
		int n, float G, float4 *ps, float *pe, int *epsMax) {
	__shared__ float sm[4];
	float ee = 0.f;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		const float4 q = uq[i];
		const float e = uv[i].w;
		ps[i] = make_float4(q.x, q.y, q.z, fmul(G, q.w));
		pe[i] = e;
		ee = fmaxf(ee, e);
	}
	for (int o = 32; o > 0; o >>= 1) ee = fmaxf(ee, __shfl_xor(ee, o, 64));
	if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = ee;
	__syncthreads();
	if (threadIdx.x == 0) {
		ee = fmaxf(fmaxf(sm[0], sm[1]), fmaxf(sm[2], sm[3]));
		if (ee > 0.f) atomicMax(epsMax, __float_as_int(ee));
	}
}

// Potential launch over (group, first target) tiles of T targets.
// SKG_POT_BLOCK=1: a two-sum per pair instead of per 8-term partial (error studies).
template <int T>
static void launch_unb_pot_t(skg_ctx *c, const GroupState *G, const int2 *tiles, int ntl, int blk) {
	const float *em = (const float *)G->d_epsMax;
#define SKG_POT_ARGS tiles, G->gstart, G->gend, G->uq, G->uv, G->psrc, G->peps, em, G->uphi, c->P.iSoftType
	if (blk == 1) k_unb_pot<T, 1><<<ntl, T, 0, c->st>>>(SKG_POT_ARGS);
	else k_unb_pot<T, 8><<<ntl, T, 0, c->st>>>(SKG_POT_ARGS);
#undef SKG_POT_ARGS
}
// n: length of the member workspace (uq, uv) the tiles index into.
static int launch_unb_pot(skg_ctx *c, GroupState *G, int T, const int2 *tiles, int ntl, int n) {
	if (ntl <= 0) return 0;
	const char *e = getenv("SKG_POT_BLOCK");
	const int blk = (e && atoi(e) == 1) ? 1 : 8;
	if (G->nsrc < n || !G->d_epsMax) {
		if (dalloc(&G->psrc, n) || dalloc(&G->peps, n) || dalloc(&G->d_epsMax, 1)) return SKG_E_NOMEM;
		G->nsrc = n;
	}
	k_izero<<<1, 64, 0, c->st>>>(G->d_epsMax, 1);
	k_pot_src<<<std::max(1, std::min(nblk(n, 256), 2048)), 256, 0, c->st>>>(G->uq, G->uv, n, c->P.G, G->psrc,
			G->peps, G->d_epsMax);
	if (T == 64) launch_unb_pot_t<64>(c, G, tiles, ntl, blk);
	else if (T == 128) launch_unb_pot_t<128>(c, G, tiles, ntl, blk);
	else launch_unb_pot_t<256>(c, G, tiles, ntl, blk);
	return 0;
}

// Groups of at most 64 members: one wavefront per group, member i in lane i.
// Potential by shuffles, then the removal loop entirely in registers
// (kd.c:1360-1457); energies, centre of mass and the sign test in fp64.
__global__ __launch_bounds__(64) void k_unb_small(const int *__restrict__ glist, const int *__restrict__ gs,
		const int *__restrict__ ge, const float4 *__restrict__ uq, const float4 *__restrict__ uv,
		const float2 *__restrict__ uphi, const int *__restrict__ uord, int *grp, int *gcount, float *gmass,
		float4 *gvcm, float4 *grbound, const float4 *__restrict__ grel, float fShift, float fCosmo, float z,
		float G, int soft, int bSub, int bNoUnbind, float hx, int *nunb, unsigned long long *npairs) {
	const int g = glist[blockIdx.x], lane = threadIdx.x;
	unsigned long long pairs = 0;
	const int s = gs[g], n0 = ge[g] - s;
	const bool mine = lane < n0;
	const float4 q = uq[s + (mine ? lane : 0)], v = uv[s + (mine ? lane : 0)];
	const float2 ph0 = uphi[s + (mine ? lane : 0)];
	Acc phi; phi.s = ph0.x; phi.c = ph0.y;
	for (int j = 0; j < n0; ++j) {
		const float qx = __shfl(q.x, j, 64), qy = __shfl(q.y, j, 64), qz = __shfl(q.z, j, 64);
		const float qm = __shfl(q.w, j, 64), qe = __shfl(v.w, j, 64);
		const float dx = q.x - qx, dy = q.y - qy, dz = q.z - qz;
		const float k = pot_kernel(soft, fmaf(dx, dx, fmaf(dy, dy, dz * dz)), v.w + qe);
		if (mine && j != lane) phi.add(fmul(G, qm) * k);
	}
	double cm[7];
	{
		double a[7] = {mine ? (double)q.w : 0.0,
			mine ? (double)fmul(q.w, q.x) : 0.0, mine ? (double)fmul(q.w, q.y) : 0.0, mine ? (double)fmul(q.w, q.z) : 0.0,
			mine ? (double)fmul(q.w, v.x) : 0.0, mine ? (double)fmul(q.w, v.y) : 0.0, mine ? (double)fmul(q.w, v.z) : 0.0};
		for (int k = 0; k < 7; ++k) {
			for (int o = 32; o > 0; o >>= 1) a[k] += __shfl_xor(a[k], o, 64);
			cm[k] = a[k];
		}
		for (int k = 1; k < 7; ++k) cm[k] /= cm[0];
	}
	bool alive = mine;
	int n = n0, removed = 0, imin = 0;
	const double zf = 1.0 + (double)z;
	for (;;) {
		double E = -INFINITY;
		if (alive) {
			const double d0 = (double)fShift * ((double)v.x - cm[4]) + (double)fCosmo * ((double)q.x - cm[1]);
			const double d1 = (double)fShift * ((double)v.y - cm[5]) + (double)fCosmo * ((double)q.y - cm[2]);
			const double d2 = (double)fShift * ((double)v.z - cm[6]) + (double)fCosmo * ((double)q.z - cm[3]);
			E = 0.5 * (d0 * d0 + d1 * d1 + d2 * d2) - ((double)phi.s + (double)phi.c) * zf;
		}
		double bE = E, mE = alive ? E : INFINITY;
		int bi = alive ? lane : 64, mi = alive ? lane : 64;
		for (int o = 32; o > 0; o >>= 1) {
			const double oE = __shfl_xor(bE, o, 64), omE = __shfl_xor(mE, o, 64);
			const int oi = __shfl_xor(bi, o, 64), omi = __shfl_xor(mi, o, 64);
			if (oE > bE || (oE == bE && oi < bi)) { bE = oE; bi = oi; }
			if (omE < mE || (omE == mE && omi < mi)) { mE = omE; mi = omi; }
		}
		imin = mi;
		if (!(bE >= 0.0) || bNoUnbind || n == 0) break;
		const float bx = __shfl(q.x, bi, 64), by = __shfl(q.y, bi, 64), bz = __shfl(q.z, bi, 64);
		const float bm = __shfl(q.w, bi, 64), be = __shfl(v.w, bi, 64);
		const float bvx = __shfl(v.x, bi, 64), bvy = __shfl(v.y, bi, 64), bvz = __shfl(v.z, bi, 64);
		if (lane == bi) { alive = false; grp[uord[s + lane]] = 0; }
		--n; ++removed;
		if (n == 0) { cm[0] = 0.0; cm[4] = cm[5] = cm[6] = 0.0; break; }
		cm[0] -= (double)bm;
		const double f = (double)bm / cm[0];
		cm[1] += f * (cm[1] - (double)bx); cm[2] += f * (cm[2] - (double)by); cm[3] += f * (cm[3] - (double)bz);
		cm[4] += f * (cm[4] - (double)bvx); cm[5] += f * (cm[5] - (double)bvy); cm[6] += f * (cm[6] - (double)bvz);
		if (bSub) pairs += n;
		if (bSub && alive) {
			const float dx = fsub(bx, q.x), dy = fsub(by, q.y), dz = fsub(bz, q.z);
			phi.add(-fmul(fmul(G, bm), pot_kernel(soft, fmaf(dx, dx, fmaf(dy, dy, dz * dz)), fadd(be, v.w))));
		}
	}
	const float rbx = __shfl(q.x, imin < 64 ? imin : 0, 64), rby = __shfl(q.y, imin < 64 ? imin : 0, 64);
	const float rbz = __shfl(q.z, imin < 64 ? imin : 0, 64);
	if (lane == 0) {
		gcount[g] = n;
		gmass[g] = (float)cm[0];
		gvcm[g] = make_float4((float)cm[4], (float)cm[5], (float)cm[6], 0.f);
		const float4 r = grel[g];
		float b[3] = {fadd(rbx, r.x), fadd(rby, r.y), fadd(rbz, r.z)};
		for (int j = 0; j < 3; ++j) {                                    // hx on all axes (kd.c:1452-1456)
			if (b[j] > hx) b[j] = fsub(b[j], 2 * hx);
			if (b[j] <= -hx) b[j] = fadd(b[j], 2 * hx);
		}
		grbound[g] = make_float4(b[0], b[1], b[2], 0.f);
		atomicAdd(nunb, removed);
		atomicAdd(npairs, pairs);
	}
}

// All large groups at once (one cooperative launch): group gi owns
// workgroups [wg0[gi], wg0[gi]+nwgG[gi]), each workgroup a fixed slice of the
// group's members.  Per removal round, one grid barrier publishes the slice
// maxima (double-buffered per round); each workgroup of a group reduces its
// group's maxima identically and applies the removal to its own slice
// (kd.c:1385-1447).  A group that has finished idles; all workgroups leave
// together once the count of unfinished groups, read right after a barrier,
// is zero.
#define COOP_T 256
__global__ __launch_bounds__(COOP_T) void k_unb_coop(const int *__restrict__ bg, const int *__restrict__ wg0,
		const int *__restrict__ nwgG, const int *__restrict__ wgGroup, const int *__restrict__ gs,
		const int *__restrict__ ge, const float4 *__restrict__ uq, const float4 *__restrict__ uv, float2 *uphi,
		int *uord, int *grp, int *gcount, float *gmass, float4 *gvcm, float4 *grbound,
		const float4 *__restrict__ grel, float fShift, float fCosmo, float z, float G, int soft, int bSub,
		int bNoUnbind, float hx, int *nunb, double *part /* 2 * gridDim * 4 */,
		const double *__restrict__ cmInit /* 8 per big group */, int *active, unsigned long long *npairs) {
	namespace cg = cooperative_groups;
	cg::grid_group grid = cg::this_grid();
	__shared__ double rE[COOP_T], rM[COOP_T];
	__shared__ int rI[COOP_T], rJ[COOP_T];
	__shared__ int sBig, sMin;
	__shared__ double sBigE;
	const int tid = threadIdx.x, nwgAll = gridDim.x, wg = blockIdx.x;
	const int gi = wgGroup[wg], g = bg[gi], lw = wg - wg0[gi], nwg = nwgG[gi];
	const int s = gs[g], n0 = ge[g] - s;
	const int chunk = (n0 + nwg - 1) / nwg;
	const int a0 = min(n0, lw * chunk), a1 = min(n0, a0 + chunk);
	double cm[7];
	for (int k = 0; k < 7; ++k) cm[k] = cmInit[8 * gi + k];
	const double zf = 1.0 + (double)z;
	int n = n0, removed = 0, prev = -1, imin = 0;
	bool done = false;
	unsigned long long pairs = 0;
	for (int iter = 0;; ++iter) {
		double *P = part + (size_t)(iter & 1) * nwgAll * 4;
		if (!done) {
			float4 pq = make_float4(0.f, 0.f, 0.f, 0.f);
			float pe = 0.f;
			if (prev >= 0) { pq = uq[s + prev]; pe = uv[s + prev].w; }
			double bE = -INFINITY, mE = INFINITY;
			int bi = INT_MAX, mi = INT_MAX;
			for (int i = a0 + tid; i < a1; i += COOP_T) {
				if (uord[s + i] < 0) continue;
				const float4 q = uq[s + i], v = uv[s + i];
				float2 ph = uphi[s + i];
				if (prev >= 0 && bSub) {                              // kdSubPot (grav.c:39-60)
					const float dx = fsub(pq.x, q.x), dy = fsub(pq.y, q.y), dz = fsub(pq.z, q.z);
					Acc ac; ac.s = ph.x; ac.c = ph.y;
					ac.add(-fmul(fmul(G, pq.w), pot_kernel(soft, fmaf(dx, dx, fmaf(dy, dy, dz * dz)), fadd(pe, v.w))));
					ph = make_float2(ac.s, ac.c);
					uphi[s + i] = ph;
				}
				const double d0 = (double)fShift * ((double)v.x - cm[4]) + (double)fCosmo * ((double)q.x - cm[1]);
				const double d1 = (double)fShift * ((double)v.y - cm[5]) + (double)fCosmo * ((double)q.y - cm[2]);
				const double d2 = (double)fShift * ((double)v.z - cm[6]) + (double)fCosmo * ((double)q.z - cm[3]);
				const double E = 0.5 * (d0 * d0 + d1 * d1 + d2 * d2) - ((double)ph.x + (double)ph.y) * zf;
				if (E > bE || (E == bE && i < bi)) { bE = E; bi = i; }
				if (E < mE || (E == mE && i < mi)) { mE = E; mi = i; }
			}
			rE[tid] = bE; rI[tid] = bi; rM[tid] = mE; rJ[tid] = mi;
			__syncthreads();
			for (int o = COOP_T / 2; o > 0; o >>= 1) {
				if (tid < o) {
					if (rE[tid + o] > rE[tid] || (rE[tid + o] == rE[tid] && rI[tid + o] < rI[tid])) { rE[tid] = rE[tid + o]; rI[tid] = rI[tid + o]; }
					if (rM[tid + o] < rM[tid] || (rM[tid + o] == rM[tid] && rJ[tid + o] < rJ[tid])) { rM[tid] = rM[tid + o]; rJ[tid] = rJ[tid + o]; }
				}
				__syncthreads();
			}
			if (tid == 0) {
				P[4 * wg] = rE[0]; P[4 * wg + 1] = (double)rI[0]; P[4 * wg + 2] = rM[0]; P[4 * wg + 3] = (double)rJ[0];
			}
			__threadfence();
		}
		grid.sync();
		if (!done) {
			// the group's workgroups reduce its slice maxima identically
			double BE = -INFINITY, ME = INFINITY;
			int BI = INT_MAX, MI = INT_MAX;
			for (int w = tid; w < nwg; w += COOP_T) {
				const int ww = wg0[gi] + w;
				const double e1 = __hip_atomic_load(&P[4 * ww], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				const double e2 = __hip_atomic_load(&P[4 * ww + 2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				const int i1 = (int)__hip_atomic_load(&P[4 * ww + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				const int i2 = (int)__hip_atomic_load(&P[4 * ww + 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				if (e1 > BE || (e1 == BE && i1 < BI)) { BE = e1; BI = i1; }
				if (e2 < ME || (e2 == ME && i2 < MI)) { ME = e2; MI = i2; }
			}
			rE[tid] = BE; rI[tid] = BI; rM[tid] = ME; rJ[tid] = MI;
			__syncthreads();
			for (int o = COOP_T / 2; o > 0; o >>= 1) {
				if (tid < o) {
					if (rE[tid + o] > rE[tid] || (rE[tid + o] == rE[tid] && rI[tid + o] < rI[tid])) { rE[tid] = rE[tid + o]; rI[tid] = rI[tid + o]; }
					if (rM[tid + o] < rM[tid] || (rM[tid + o] == rM[tid] && rJ[tid + o] < rJ[tid])) { rM[tid] = rM[tid + o]; rJ[tid] = rJ[tid + o]; }
				}
				__syncthreads();
			}
			if (tid == 0) { sBig = rI[0]; sBigE = rE[0]; sMin = rJ[0]; }
			__syncthreads();
		}
		const int ib = sBig;
		bool stop = done;
		if (!done) {
		imin = sMin;
		stop = !(sBigE >= 0.0) || bNoUnbind || n == 0 || ib == INT_MAX;
		if (!stop) {
			const float4 qb = uq[s + ib], vb = uv[s + ib];
			if (ib >= a0 && ib < a1 && tid == 0) {
				const int ob = uord[s + ib];
				grp[ob] = 0;
				uord[s + ib] = ~ob;
			}
			--n; ++removed;
			if (n == 0) {
				cm[0] = 0.0; cm[4] = cm[5] = cm[6] = 0.0;
				stop = true;
			} else {
				cm[0] -= (double)qb.w;
				const double f = (double)qb.w / cm[0];
				cm[1] += f * (cm[1] - (double)qb.x); cm[2] += f * (cm[2] - (double)qb.y); cm[3] += f * (cm[3] - (double)qb.z);
				cm[4] += f * (cm[4] - (double)vb.x); cm[5] += f * (cm[5] - (double)vb.y); cm[6] += f * (cm[6] - (double)vb.z);
				prev = ib;
				if (bSub) pairs += n;
			}
		}
		}
		if (stop && !done) {
			done = true;
			if (lw == 0 && tid == 0) {
				gcount[g] = n;
				gmass[g] = (float)cm[0];
				gvcm[g] = make_float4((float)cm[4], (float)cm[5], (float)cm[6], 0.f);
				const float4 q = uq[s + (imin < n0 ? imin : 0)], r = grel[g];
				float b[3] = {fadd(q.x, r.x), fadd(q.y, r.y), fadd(q.z, r.z)};
				for (int j = 0; j < 3; ++j) {
					if (b[j] > hx) b[j] = fsub(b[j], 2 * hx);
					if (b[j] <= -hx) b[j] = fadd(b[j], 2 * hx);
				}
				grbound[g] = make_float4(b[0], b[1], b[2], 0.f);
				atomicAdd(nunb, removed);
				atomicAdd(npairs, pairs);
				__threadfence();
				atomicSub(active, 1);
			}
		}
		__syncthreads();
		// Termination is read only between two barriers: no group finishes
		// (decrements) between this barrier and the next round's, so every
		// workgroup sees the same count and they all leave on the same round.
		if ((iter & 15) == 15) {
			grid.sync();
			if (__hip_atomic_load(active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
		}
	}
}

// Exact sequential removal with a bounded candidate set, one workgroup per
// group (kd.c:1385-1447).  With dv_i = u_i - w (u_i = fShift v_i + fCosmo r_i,
// w = fShift vcm + fCosmo rcm), E_i = |u_i - w|^2/2 - phi_i (1+z).  A full pass
// at w0 keeps each thread's CT most energetic members as candidates and
// records, over all other members, the largest energy E_cut and the largest
// |u_i - w0| = R.  Until the next full pass every non-candidate satisfies
//   E_i(w) <= E_cut + R |w - w0| + |w - w0|^2/2 + growth,
// where growth sums the largest possible potential drop (1+z) G m_b K_max of
// each removal (kdSubPot).  While the best candidate beats that bound strictly
// it is the global argmax and is removed without touching other members; the
// potentials of non-candidates catch up at the next full pass.  The sequence
// of removals is therefore the reference's (ties aside).
#define CAND_T 256
#define CT 8
#define CAND_WAVE 2048        // groups below this many members: one wavefront each
__device__ __forceinline__ float kernel_max(int soft, float twoh) {
	return soft == SKG_PLUMMER ? 2.0f / twoh : 2.8f / twoh;     // K(0): spline 1.4*2/twoh, Plummer 2/twoh
}
// Block-wide reductions of k_unb_cand: a butterfly within each wavefront, then
// the CAND_T / 64 wavefront results through LDS.  Larger E wins, ties to the
// smaller index (argmax); argmin the same with smaller E.
__device__ __forceinline__ void wave_argmax_d(double &E, int &i) {
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) {
		const double oE = __shfl_xor(E, o, 64);
		const int oi = __shfl_xor(i, o, 64);
		if (oE > E || (oE == E && oi < i)) { E = oE; i = oi; }
	}
}
__device__ __forceinline__ void wave_argmin_d(double &E, int &i) {
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) {
		const double oE = __shfl_xor(E, o, 64);
		const int oi = __shfl_xor(i, o, 64);
		if (oE < E || (oE == E && oi < i)) { E = oE; i = oi; }
	}
}
// The same argmax with DPP: row steps within quads, half-rows and rows, then
// the two row broadcasts, after which lane 63 holds the wavefront's (E, i);
// every lane takes lane 63's.  A lane whose source is not written keeps its
// own value.  Needs every lane of the wavefront active.
template <int CTRL, int RMASK>
__device__ __forceinline__ void dpp_argmax_step(double &E, int &i) {
	const int lo = __double2loint(E), hi = __double2hiint(E);
	const int olo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, RMASK, 0xF, false);
	const int ohi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, RMASK, 0xF, false);
	const int oi = __builtin_amdgcn_update_dpp(i, i, CTRL, RMASK, 0xF, false);
	const double oE = __hiloint2double(ohi, olo);
	if (oE > E || (oE == E && oi < i)) { E = oE; i = oi; }
}
__device__ __forceinline__ void wave_argmax_dpp(double &E, int &i) {
	dpp_argmax_step<0xB1, 0xF>(E, i);            // quad_perm [1,0,3,2]
	dpp_argmax_step<0x4E, 0xF>(E, i);            // quad_perm [2,3,0,1]
	dpp_argmax_step<0x141, 0xF>(E, i);           // row_half_mirror
	dpp_argmax_step<0x140, 0xF>(E, i);           // row_mirror
	dpp_argmax_step<0x142, 0xA>(E, i);           // row_bcast:15 into rows 1 and 3
	dpp_argmax_step<0x143, 0xC>(E, i);           // row_bcast:31 into rows 2 and 3
	E = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(E), 63),
			__builtin_amdgcn_readlane(__double2loint(E), 63));
	i = __builtin_amdgcn_readlane(i, 63);
}
__device__ __forceinline__ double wave_max_d(double v) {
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
	return v;
}

template <int T>
__global__ __launch_bounds__(T) void k_unb_cand(const int *__restrict__ glist, const int *__restrict__ gs,
		const int *__restrict__ ge, const float4 *__restrict__ uq, const float4 *__restrict__ uv, float2 *uphi,
		int *uord, int *grp, int *gcount, float *gmass, float4 *gvcm, float4 *grbound,
		const float4 *__restrict__ grel, float fShift, float fCosmo, float z, float G, int soft, int bSub,
		int bNoUnbind, float hx, int *nunb, unsigned long long *npairs, unsigned long long *nfull) {
	constexpr int NW = T / 64;
	__shared__ double rA[T];
	__shared__ int pend[T * CT];
	__shared__ double cmS[7];
	// full-pass results per wavefront
	__shared__ double fCut[NW], fMin[NW], fR2[NW];
	__shared__ int fIdx[NW], fEps[NW];
	// removal rounds: each wavefront's best candidate, two buffers used in turn
	// (a round's barrier keeps the next round's writes off the slots still read)
	__shared__ double bE_s[2][NW];
	__shared__ int bI_s[2][NW];
	__shared__ float4 bQ_s[2][NW], bV_s[2][NW];
	const int g = glist[blockIdx.x], tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
	const int s = gs[g], n0 = ge[g] - s;
	const double zf = 1.0 + (double)z;
	// fp64 mass, centre of mass and velocity (kd.c:1360-1375)
	{
		double a[7] = {0, 0, 0, 0, 0, 0, 0};
		for (int i = tid; i < n0; i += T) {
			const float4 q = uq[s + i], v = uv[s + i];
			a[0] += (double)q.w;
			a[1] += (double)fmul(q.w, q.x); a[2] += (double)fmul(q.w, q.y); a[3] += (double)fmul(q.w, q.z);
			a[4] += (double)fmul(q.w, v.x); a[5] += (double)fmul(q.w, v.y); a[6] += (double)fmul(q.w, v.z);
		}
		for (int k = 0; k < 7; ++k) {
			rA[tid] = a[k];
			__syncthreads();
			for (int o = T / 2; o > 0; o >>= 1) { if (tid < o) rA[tid] += rA[tid + o]; __syncthreads(); }
			if (tid == 0) cmS[k] = rA[0];
			__syncthreads();
		}
		if (tid == 0) for (int k = 1; k < 7; ++k) cmS[k] /= cmS[0];
		__syncthreads();
	}
	double cm[7];
	for (int k = 0; k < 7; ++k) cm[k] = cmS[k];
	int cI[CT];
	float4 cQ[CT], cV[CT];
	double cP[CT];
	int n = n0, removed = 0, npend = 0, imin = 0, rb = 0;
	unsigned long long pairs = 0, fulls = 0;
	double w0x = 0, w0y = 0, w0z = 0, Ecut = -INFINITY, Rmax = 0, grow = 0;
	float epsMin = 0.f;
	bool needFull = true, stop = false, lastWasFull = false;
	while (!stop) {
		if (needFull) {
			++fulls;
			__syncthreads();                          // the last removal's pend entry and member marks
			// potentials catch up with the removals since the last pass (kdSubPot, grav.c:39-60)
			const double wx = (double)fShift * cm[4] + (double)fCosmo * cm[1];
			const double wy = (double)fShift * cm[5] + (double)fCosmo * cm[2];
			const double wz = (double)fShift * cm[6] + (double)fCosmo * cm[3];
			double eCut = -INFINITY, r2Max = 0.0, mE = INFINITY;
			int mi = INT_MAX;
			float eMin = INFINITY;
			double cE[CT];
			#pragma unroll
			for (int c = 0; c < CT; ++c) { cI[c] = -1; cE[c] = -INFINITY; }
			int lo = 0;                               // the slot holding the smallest kept energy
			double loE = -INFINITY;
			// one member (its loads issued with the next three members')
			auto visit = [&](int i, const float4 &q, const float4 &v, float2 ph) {
				if (bSub && npend) {
					Acc ac; ac.s = ph.x; ac.c = ph.y;
					for (int k = 0; k < npend; ++k) {
						const int b = pend[k];
						const float4 qb2 = uq[s + b];
						const float eb = uv[s + b].w;
						const float dx = fsub(qb2.x, q.x), dy = fsub(qb2.y, q.y), dz = fsub(qb2.z, q.z);
						ac.add(-fmul(fmul(G, qb2.w), pot_kernel(soft, fmaf(dx, dx, fmaf(dy, dy, dz * dz)), fadd(eb, v.w))));
					}
					ph = make_float2(ac.s, ac.c);
					uphi[s + i] = ph;
				}
				const double d0 = (double)fShift * ((double)v.x - cm[4]) + (double)fCosmo * ((double)q.x - cm[1]);
				const double d1 = (double)fShift * ((double)v.y - cm[5]) + (double)fCosmo * ((double)q.y - cm[2]);
				const double d2 = (double)fShift * ((double)v.z - cm[6]) + (double)fCosmo * ((double)q.z - cm[3]);
				const double dd = d0 * d0 + d1 * d1 + d2 * d2;
				const double phi = (double)ph.x + (double)ph.y;
				const double E = 0.5 * dd - phi * zf;
				if (E < mE || (E == mE && i < mi)) { mE = E; mi = i; }
				eMin = fminf(eMin, v.w);
				// keep the CT largest; what is displaced feeds the bound (R as
				// its square: the square root of the largest square is the
				// largest root)
				double outE = E, outR2 = dd;
				if (E > loE) {
					#pragma unroll
					for (int c = 0; c < CT; ++c) {
						if (c == lo) {
							if (cI[c] >= 0) {
								const double ox = (double)fShift * (double)cV[c].x + (double)fCosmo * (double)cQ[c].x - wx;
								const double oy = (double)fShift * (double)cV[c].y + (double)fCosmo * (double)cQ[c].y - wy;
								const double oz = (double)fShift * (double)cV[c].z + (double)fCosmo * (double)cQ[c].z - wz;
								outE = cE[c]; outR2 = ox * ox + oy * oy + oz * oz;
							} else { outE = -INFINITY; outR2 = 0.0; }
							cI[c] = i; cQ[c] = q; cV[c] = v; cP[c] = phi; cE[c] = E;
						}
					}
					lo = 0; loE = cE[0];
					#pragma unroll
					for (int c = 1; c < CT; ++c) if (cE[c] < loE) { loE = cE[c]; lo = c; }
				}
				if (outE > eCut) eCut = outE;
				if (outR2 > r2Max) r2Max = outR2;
			};
			for (int i0 = tid; i0 < n0; i0 += 4 * T) {
				int o4[4];
				float4 q4[4], v4[4];
				float2 p4[4];
				#pragma unroll
				for (int u = 0; u < 4; ++u) {
					const int i = i0 + u * T;
					o4[u] = -1;
					if (i < n0) { o4[u] = uord[s + i]; q4[u] = uq[s + i]; v4[u] = uv[s + i]; p4[u] = uphi[s + i]; }
				}
				#pragma unroll
				for (int u = 0; u < 4; ++u)
					if (o4[u] >= 0) visit(i0 + u * T, q4[u], v4[u], p4[u]);
			}
			// block reductions: E_cut, R^2, smallest softening, argmin E
			eCut = wave_max_d(eCut);
			r2Max = wave_max_d(r2Max);
			wave_argmin_d(mE, mi);
			eMin = wave_min(eMin);
			if (lane == 0) { fCut[wv] = eCut; fR2[wv] = r2Max; fMin[wv] = mE; fIdx[wv] = mi; fEps[wv] = __float_as_int(eMin); }
			__syncthreads();
			eCut = fCut[0]; r2Max = fR2[0]; mE = fMin[0]; mi = fIdx[0]; eMin = __int_as_float(fEps[0]);
			#pragma unroll
			for (int w = 1; w < NW; ++w) {
				eCut = fmax(eCut, fCut[w]);
				r2Max = fmax(r2Max, fR2[w]);
				if (fMin[w] < mE || (fMin[w] == mE && fIdx[w] < mi)) { mE = fMin[w]; mi = fIdx[w]; }
				eMin = fminf(eMin, __int_as_float(fEps[w]));
			}
			Ecut = eCut; imin = mi; epsMin = eMin; Rmax = sqrt(r2Max);
			w0x = wx; w0y = wy; w0z = wz;
			grow = 0.0; npend = 0; needFull = false; lastWasFull = true;
			if (bNoUnbind || n == 0) break;
		}
		// best candidate at the current centre of mass
		double bE = -INFINITY;
		int bi = INT_MAX, bc = -1;
		#pragma unroll
		for (int c = 0; c < CT; ++c) {
			if (cI[c] < 0) continue;
			const double d0 = (double)fShift * ((double)cV[c].x - cm[4]) + (double)fCosmo * ((double)cQ[c].x - cm[1]);
			const double d1 = (double)fShift * ((double)cV[c].y - cm[5]) + (double)fCosmo * ((double)cQ[c].y - cm[2]);
			const double d2 = (double)fShift * ((double)cV[c].z - cm[6]) + (double)fCosmo * ((double)cQ[c].z - cm[3]);
			const double E = 0.5 * (d0 * d0 + d1 * d1 + d2 * d2) - cP[c] * zf;
			if (E > bE || (E == bE && cI[c] < bi)) { bE = E; bi = cI[c]; bc = c; }
		}
		// the wavefront's best: its owner (or lane 0, with no candidate in the
		// wavefront) publishes it with the member's records
		double wE = bE;
		int wI = bi;
		wave_argmax_dpp(wE, wI);
		if (wI == INT_MAX ? lane == 0 : (bc >= 0 && bi == wI)) {
			bE_s[rb][wv] = wE; bI_s[rb][wv] = wI;
			#pragma unroll
			for (int c = 0; c < CT; ++c) if (c == bc) { bQ_s[rb][wv] = cQ[c]; bV_s[rb][wv] = cV[c]; }
		}
		__syncthreads();
		double BE = bE_s[rb][0];
		int BI = bI_s[rb][0], BW = 0;
		#pragma unroll
		for (int w = 1; w < NW; ++w) {
			const double e = bE_s[rb][w];
			const int j = bI_s[rb][w];
			if (e > BE || (e == BE && j < BI)) { BE = e; BI = j; BW = w; }
		}
		const int rbw = rb;
		rb ^= 1;
		const double wx = (double)fShift * cm[4] + (double)fCosmo * cm[1];
		const double wy = (double)fShift * cm[5] + (double)fCosmo * cm[2];
		const double wz = (double)fShift * cm[6] + (double)fCosmo * cm[3];
		const double dl = sqrt((wx - w0x) * (wx - w0x) + (wy - w0y) * (wy - w0y) + (wz - w0z) * (wz - w0z));
		// right after a full pass every energy is exact and BE >= E_cut holds;
		// later the bound carries a relative slack for fp64 rounding
		const double slack = 1e-12 * (fabs(Ecut) + Rmax * Rmax + 1.0);
		const double bound = lastWasFull ? Ecut : Ecut + Rmax * dl + 0.5 * dl * dl + grow + slack;
		if (!(lastWasFull ? BE >= bound : BE > bound)) {
			if (BE < 0.0 && bound < 0.0) { stop = true; break; }      // every member is bound
			needFull = true;
			continue;
		}
		lastWasFull = false;
		if (BE < 0.0) { stop = true; break; }                         // the global maximum is negative
		// remove BI (owned by one thread's candidate slot)
		if (bc >= 0 && bi == BI) {
			#pragma unroll
			for (int c = 0; c < CT; ++c) if (c == bc) cI[c] = -1;
			const int ob = uord[s + BI];
			grp[ob] = 0;
			uord[s + BI] = ~ob;
		}
		const float4 qb = bQ_s[rbw][BW], vb = bV_s[rbw][BW];
		--n; ++removed;
		if (n == 0) { cm[0] = 0.0; cm[4] = cm[5] = cm[6] = 0.0; stop = true; break; }
		cm[0] -= (double)qb.w;
		const double f = (double)qb.w / cm[0];
		cm[1] += f * (cm[1] - (double)qb.x); cm[2] += f * (cm[2] - (double)qb.y); cm[3] += f * (cm[3] - (double)qb.z);
		cm[4] += f * (cm[4] - (double)vb.x); cm[5] += f * (cm[5] - (double)vb.y); cm[6] += f * (cm[6] - (double)vb.z);
		if (bSub) {
			pairs += n;
			#pragma unroll
			for (int c = 0; c < CT; ++c) {
				if (cI[c] < 0) continue;
				const float dx = fsub(qb.x, cQ[c].x), dy = fsub(qb.y, cQ[c].y), dz = fsub(qb.z, cQ[c].z);
				cP[c] -= (double)fmul(fmul(G, qb.w), pot_kernel(soft, fmaf(dx, dx, fmaf(dy, dy, dz * dz)), fadd(vb.w, cV[c].w)));
			}
			grow += zf * (double)fmul(fmul(G, qb.w), kernel_max(soft, fadd(vb.w, epsMin)));
			if (tid == 0) pend[npend] = BI;
			++npend;
			if (npend >= T * CT) needFull = true;
		}
	}
	// iMin over all members in the final state (kd.c:1400-1408 of the last pass)
	if (!lastWasFull && n > 0) {
		__syncthreads();                              // the last removals' member marks
		double mE = INFINITY;
		int mi = INT_MAX;
		for (int i = tid; i < n0; i += T) {
			if (uord[s + i] < 0) continue;
			const float4 q = uq[s + i], v = uv[s + i];
			const float2 ph = uphi[s + i];
			double phi = (double)ph.x + (double)ph.y;
			if (bSub)
				for (int k = 0; k < npend; ++k) {
					const int b = pend[k];
					const float4 qb2 = uq[s + b];
					const float dx = fsub(qb2.x, q.x), dy = fsub(qb2.y, q.y), dz = fsub(qb2.z, q.z);
					phi -= (double)fmul(fmul(G, qb2.w), pot_kernel(soft, fmaf(dx, dx, fmaf(dy, dy, dz * dz)), fadd(uv[s + b].w, v.w)));
				}
			const double d0 = (double)fShift * ((double)v.x - cm[4]) + (double)fCosmo * ((double)q.x - cm[1]);
			const double d1 = (double)fShift * ((double)v.y - cm[5]) + (double)fCosmo * ((double)q.y - cm[2]);
			const double d2 = (double)fShift * ((double)v.z - cm[6]) + (double)fCosmo * ((double)q.z - cm[3]);
			const double E = 0.5 * (d0 * d0 + d1 * d1 + d2 * d2) - phi * zf;
			if (E < mE || (E == mE && i < mi)) { mE = E; mi = i; }
		}
		wave_argmin_d(mE, mi);
		if (lane == 0) { fMin[wv] = mE; fIdx[wv] = mi; }
		__syncthreads();
		mE = fMin[0]; mi = fIdx[0];
		#pragma unroll
		for (int w = 1; w < NW; ++w)
			if (fMin[w] < mE || (fMin[w] == mE && fIdx[w] < mi)) { mE = fMin[w]; mi = fIdx[w]; }
		imin = mi;
	}
	if (tid == 0) {
		gcount[g] = n;
		gmass[g] = (float)cm[0];
		gvcm[g] = make_float4((float)cm[4], (float)cm[5], (float)cm[6], 0.f);
		const float4 q = uq[s + (imin < n0 ? imin : 0)], r = grel[g];
		float b[3] = {fadd(q.x, r.x), fadd(q.y, r.y), fadd(q.z, r.z)};
		for (int j = 0; j < 3; ++j) {
			if (b[j] > hx) b[j] = fsub(b[j], 2 * hx);
			if (b[j] <= -hx) b[j] = fadd(b[j], 2 * hx);
		}
		grbound[g] = make_float4(b[0], b[1], b[2], 0.f);
		atomicAdd(nunb, removed);
		atomicAdd(npairs, pairs);
		atomicAdd(nfull, fulls);
	}
}

// Large groups, all at once in one cooperative launch, with the bounded
// candidates of k_unb_cand spread over many workgroups.  Group gi owns
// workgroups [wg0, wg0 + nwg), each a fixed slice of its members.  A full
// pass (every workgroup over its slice) applies the kdSubPot subtractions of
// the removals since the last pass, keeps the slice's cw most energetic
// members and bounds the rest (E_cut, R = max |u - w0|, as in k_unb_cand).
// After a grid barrier every workgroup of the group holds all nwg*cw
// candidates and runs the same removal sequence on them -- the same inputs,
// reductions and decisions in each -- so a removal needs no barrier until the
// bound fails and the next full pass is due.  A round is [removals | full
// pass] barrier [read the candidates | finish] barrier [leave when no group is
// unfinished]; every workgroup passes both barriers of every round.
#define CPC 8            // candidate slots per thread (COOP_T * CPC per group)
#define CPEND 2048       // removals held between full passes
struct CoopCand { float4 q, v; double phi; int idx, pad; };
struct CoopHdr { double ecut, dmax, emin; int imin; float epsmin; };

__device__ __forceinline__ void blk_argmax(double &e, int &i, double *sE, int *sI) {
	for (int o = 32; o > 0; o >>= 1) {
		const double oe = __shfl_xor(e, o, 64);
		const int oi = __shfl_xor(i, o, 64);
		if (oe > e || (oe == e && oi < i)) { e = oe; i = oi; }
	}
	__syncthreads();                       // readers of the previous call are done
	if ((threadIdx.x & 63) == 0) { sE[threadIdx.x >> 6] = e; sI[threadIdx.x >> 6] = i; }
	__syncthreads();
	e = sE[0]; i = sI[0];
	for (int k = 1; k < COOP_T / 64; ++k)
		if (sE[k] > e || (sE[k] == e && sI[k] < i)) { e = sE[k]; i = sI[k]; }
}
__device__ __forceinline__ void blk_argmin(double &e, int &i, double *sE, int *sI) {
	e = -e;
	blk_argmax(e, i, sE, sI);
	e = -e;
}
__device__ __forceinline__ double blk_max(double e, double *sE) {
	int i = 0;
	blk_argmax(e, i, sE, (int *)(sE + COOP_T / 64));
	return e;
}

__global__ __launch_bounds__(COOP_T) void k_unb_coopc(const int *__restrict__ bg, const int *__restrict__ wg0,
		const int *__restrict__ nwgG, const int *__restrict__ wgGroup, const int *__restrict__ cwG,
		const int *__restrict__ gs, const int *__restrict__ ge, const float4 *__restrict__ uq,
		const float4 *__restrict__ uv, float2 *uphi, int *uord, int *grp, int *gcount, float *gmass, float4 *gvcm,
		float4 *grbound, const float4 *__restrict__ grel, float fShift, float fCosmo, float z, float G, int soft,
		int bSub, int bNoUnbind, float hx, int *nunb, CoopCand *pub, CoopHdr *hdr,
		const double *__restrict__ cmInit, int *active, unsigned long long *npairs, unsigned long long *nfull) {
	namespace cg = cooperative_groups;
	cg::grid_group grid = cg::this_grid();
	__shared__ float4 pq[CPEND];           // pending removals: x, y, z, m
	__shared__ float pe[CPEND];            // and softening
	__shared__ double sE[2 * COOP_T / 64];
	__shared__ int sI[COOP_T / 64];
	__shared__ float4 bq, bv;
	const int tid = threadIdx.x, wg = blockIdx.x;
	const int gi = wgGroup[wg], g = bg[gi], lw = wg - wg0[gi], nwg = nwgG[gi], cw = cwG[gi];
	const int s = gs[g], n0 = ge[g] - s;
	const int chunk = (n0 + nwg - 1) / nwg;
	const int a0 = min(n0, lw * chunk), a1 = min(n0, a0 + chunk);
	const int K = nwg * cw;                      // <= COOP_T * CPC (host)
	CoopCand *gp = pub + (size_t)gi * COOP_T * CPC;
	const double zf = 1.0 + (double)z;
	double cm[7];
	for (int k = 0; k < 7; ++k) cm[k] = cmInit[8 * gi + k];
	int cI[CPC];
	float4 cQ[CPC], cV[CPC];
	double cP[CPC];
#pragma unroll
	for (int c = 0; c < CPC; ++c) { cI[c] = -1; cQ[c] = cV[c] = make_float4(0.f, 0.f, 0.f, 0.f); cP[c] = 0.0; }
	int n = n0, removed = 0, npend = 0, imin = 0;
	unsigned long long pairs = 0, fulls = 0;
	double w0x = 0, w0y = 0, w0z = 0, Ecut = -INFINITY, Rmax = 0, grow = 0;
	float epsMin = 0.f;
	bool done = false, needFull = true, lastWasFull = false, fin = false, finishNow = false;
	for (;;) {
		// ---- removals from the candidates (block-uniform control flow)
		if (!done && !needFull) {
			for (;;) {
				double bE = -INFINITY;
				int bi = INT_MAX;
#pragma unroll
				for (int c = 0; c < CPC; ++c) {
					if (cI[c] < 0) continue;
					const double d0 = (double)fShift * ((double)cV[c].x - cm[4]) + (double)fCosmo * ((double)cQ[c].x - cm[1]);
					const double d1 = (double)fShift * ((double)cV[c].y - cm[5]) + (double)fCosmo * ((double)cQ[c].y - cm[2]);
					const double d2 = (double)fShift * ((double)cV[c].z - cm[6]) + (double)fCosmo * ((double)cQ[c].z - cm[3]);
					const double E = 0.5 * (d0 * d0 + d1 * d1 + d2 * d2) - cP[c] * zf;
					if (E > bE || (E == bE && cI[c] < bi)) { bE = E; bi = cI[c]; }
				}
				blk_argmax(bE, bi, sE, sI);
				const double BE = bE;
				const int BI = bi;
				const double wx = (double)fShift * cm[4] + (double)fCosmo * cm[1];
				const double wy = (double)fShift * cm[5] + (double)fCosmo * cm[2];
				const double wz = (double)fShift * cm[6] + (double)fCosmo * cm[3];
				const double dl = sqrt((wx - w0x) * (wx - w0x) + (wy - w0y) * (wy - w0y) + (wz - w0z) * (wz - w0z));
				const double slack = 1e-12 * (fabs(Ecut) + Rmax * Rmax + 1.0);
				const double bound = lastWasFull ? Ecut : Ecut + Rmax * dl + 0.5 * dl * dl + grow + slack;
				if (!(lastWasFull ? BE >= bound : BE > bound)) {
					if (BE < 0.0 && bound < 0.0) { fin = true; if (lastWasFull) finishNow = true; }   // every member is bound
					needFull = !finishNow;
					break;
				}
				if (BE < 0.0) {                                       // the global maximum is negative
					fin = true;
					if (lastWasFull) finishNow = true;
					needFull = !finishNow;
					break;
				}
				lastWasFull = false;
				// remove BI: its slot publishes the record; the workgroup whose
				// slice holds BI marks it (it is the one that scans it next)
#pragma unroll
				for (int c = 0; c < CPC; ++c)
					if (cI[c] == BI) { bq = cQ[c]; bv = cV[c]; cI[c] = -1; }
				if (tid == 0 && BI >= a0 && BI < a1) {
					const int ob = uord[s + BI];
					grp[ob] = 0;
					uord[s + BI] = ~ob;
				}
				__syncthreads();
				const float4 qb = bq, vb = bv;
				--n; ++removed;
				if (n == 0) { cm[0] = 0.0; cm[4] = cm[5] = cm[6] = 0.0; finishNow = true; break; }
				cm[0] -= (double)qb.w;
				const double f = (double)qb.w / cm[0];
				cm[1] += f * (cm[1] - (double)qb.x); cm[2] += f * (cm[2] - (double)qb.y); cm[3] += f * (cm[3] - (double)qb.z);
				cm[4] += f * (cm[4] - (double)vb.x); cm[5] += f * (cm[5] - (double)vb.y); cm[6] += f * (cm[6] - (double)vb.z);
				if (bSub) {
					pairs += n;
#pragma unroll
					for (int c = 0; c < CPC; ++c) {
						if (cI[c] < 0) continue;
						const float dx = fsub(qb.x, cQ[c].x), dy = fsub(qb.y, cQ[c].y), dz = fsub(qb.z, cQ[c].z);
						cP[c] -= (double)fmul(fmul(G, qb.w), pot_kernel(soft, fmaf(dx, dx, fmaf(dy, dy, dz * dz)), fadd(vb.w, cV[c].w)));
					}
					grow += zf * (double)fmul(fmul(G, qb.w), kernel_max(soft, fadd(vb.w, epsMin)));
					if (tid == 0) { pq[npend] = qb; pe[npend] = vb.w; }
					++npend;
					if (npend >= CPEND) { needFull = true; break; }
				}
			}
		}
		// ---- full pass over this workgroup's slice
		if (!done && needFull) {
			__syncthreads();                                       // pq / pe complete
			const double wx = (double)fShift * cm[4] + (double)fCosmo * cm[1];
			const double wy = (double)fShift * cm[5] + (double)fCosmo * cm[2];
			const double wz = (double)fShift * cm[6] + (double)fCosmo * cm[3];
			double tE[CPC], tD[CPC];
			int tI[CPC];
#pragma unroll
			for (int c = 0; c < CPC; ++c) { tE[c] = -INFINITY; tD[c] = 0.0; tI[c] = -1; }
			double dispE = -INFINITY, dispD = 0.0, mE = INFINITY;
			int mi = INT_MAX;
			float eMin = INFINITY;
			for (int i = a0 + tid; i < a1; i += COOP_T) {
				if (uord[s + i] < 0) continue;
				const float4 q = uq[s + i], v = uv[s + i];
				float2 ph = uphi[s + i];
				if (bSub && npend) {                                   // kdSubPot (grav.c:39-60)
					Acc ac; ac.s = ph.x; ac.c = ph.y;
					for (int k = 0; k < npend; ++k) {
						const float4 b = pq[k];
						const float dx = fsub(b.x, q.x), dy = fsub(b.y, q.y), dz = fsub(b.z, q.z);
						ac.add(-fmul(fmul(G, b.w), pot_kernel(soft, fmaf(dx, dx, fmaf(dy, dy, dz * dz)), fadd(pe[k], v.w))));
					}
					ph = make_float2(ac.s, ac.c);
					uphi[s + i] = ph;
				}
				const double d0 = (double)fShift * ((double)v.x - cm[4]) + (double)fCosmo * ((double)q.x - cm[1]);
				const double d1 = (double)fShift * ((double)v.y - cm[5]) + (double)fCosmo * ((double)q.y - cm[2]);
				const double d2 = (double)fShift * ((double)v.z - cm[6]) + (double)fCosmo * ((double)q.z - cm[3]);
				const double dd = d0 * d0 + d1 * d1 + d2 * d2;
				const double E = 0.5 * dd - ((double)ph.x + (double)ph.y) * zf;
				if (E < mE || (E == mE && i < mi)) { mE = E; mi = i; }
				eMin = fminf(eMin, v.w);
				int lo = 0;
				double loE = tE[0];
#pragma unroll
				for (int c = 1; c < CPC; ++c) if (tE[c] < loE) { loE = tE[c]; lo = c; }
				double outE = E, outD = dd;
				if (E > loE) {
#pragma unroll
					for (int c = 0; c < CPC; ++c)
						if (c == lo) {
							outE = tI[c] >= 0 ? tE[c] : -INFINITY;
							outD = tI[c] >= 0 ? tD[c] : 0.0;
							tE[c] = E; tD[c] = dd; tI[c] = i;
						}
				}
				dispE = fmax(dispE, outE);
				dispD = fmax(dispD, outD);
			}
			// the slice's cw most energetic members become candidates
			for (int r = 0; r < cw; ++r) {
				double e = -INFINITY;
				int i = INT_MAX;
#pragma unroll
				for (int c = 0; c < CPC; ++c)
					if (tI[c] >= 0 && (tE[c] > e || (tE[c] == e && tI[c] < i))) { e = tE[c]; i = tI[c]; }
				blk_argmax(e, i, sE, sI);
				CoopCand *slot = gp + lw * cw + r;
				if (i == INT_MAX) {
					if (tid == 0) slot->idx = -1;
					continue;
				}
#pragma unroll
				for (int c = 0; c < CPC; ++c)
					if (tI[c] == i) {
						tI[c] = -1;
						const float2 ph = uphi[s + i];
						slot->q = uq[s + i];
						slot->v = uv[s + i];
						slot->phi = (double)ph.x + (double)ph.y;
						slot->idx = i;
					}
			}
			// what stays behind is bounded by E_cut and R
#pragma unroll
			for (int c = 0; c < CPC; ++c)
				if (tI[c] >= 0) { dispE = fmax(dispE, tE[c]); dispD = fmax(dispD, tD[c]); }
			const double ecut = blk_max(dispE, sE);
			const double dmax = blk_max(dispD, sE);
			const double emin = -blk_max(-(double)eMin, sE);
			blk_argmin(mE, mi, sE, sI);
			if (tid == 0) {
				CoopHdr h;
				h.ecut = ecut; h.dmax = dmax; h.emin = mE; h.imin = mi; h.epsmin = (float)emin;
				hdr[wg] = h;
			}
			w0x = wx; w0y = wy; w0z = wz;
		}
		grid.sync();
		__threadfence();
		// ---- every workgroup of the group takes all its candidates and bounds
		if (!done && needFull) {
#pragma unroll
			for (int c = 0; c < CPC; ++c) {
				const int j = tid + c * COOP_T;
				if (j < K) {
					const CoopCand r = gp[j];
					cI[c] = r.idx; cQ[c] = r.q; cV[c] = r.v; cP[c] = r.phi;
				} else {
					cI[c] = -1;
				}
			}
			double ec = -INFINITY, dm = 0.0, em = -INFINITY, me = INFINITY;
			int mi = INT_MAX;
			for (int w = tid; w < nwg; w += COOP_T) {
				const CoopHdr h = hdr[wg0[gi] + w];
				ec = fmax(ec, h.ecut);
				dm = fmax(dm, h.dmax);
				em = fmax(em, -(double)h.epsmin);
				if (h.emin < me || (h.emin == me && h.imin < mi)) { me = h.emin; mi = h.imin; }
			}
			Ecut = blk_max(ec, sE);
			Rmax = sqrt(blk_max(dm, sE));
			epsMin = (float)(-blk_max(em, sE));
			blk_argmin(me, mi, sE, sI);
			imin = mi;
			grow = 0.0; npend = 0; needFull = false; lastWasFull = true; ++fulls;
			if (fin || bNoUnbind || n == 0) finishNow = true;
		}
		if (!done && finishNow) {
			done = true;
			if (lw == 0 && tid == 0) {
				gcount[g] = n;
				gmass[g] = (float)cm[0];
				gvcm[g] = make_float4((float)cm[4], (float)cm[5], (float)cm[6], 0.f);
				const float4 q = uq[s + (imin >= 0 && imin < n0 ? imin : 0)], r = grel[g];
				float b[3] = {fadd(q.x, r.x), fadd(q.y, r.y), fadd(q.z, r.z)};
				for (int j = 0; j < 3; ++j) {                            // hx on all axes (kd.c:1452-1456)
					if (b[j] > hx) b[j] = fsub(b[j], 2 * hx);
					if (b[j] <= -hx) b[j] = fadd(b[j], 2 * hx);
				}
				grbound[g] = make_float4(b[0], b[1], b[2], 0.f);
				atomicAdd(nunb, removed);
				atomicAdd(npairs, pairs);
				atomicAdd(nfull, fulls);
				__threadfence();
				atomicSub(active, 1);
			}
		}
		// groups finish only between the two barriers, so every workgroup
		// reads the same count here and all leave on the same round
		grid.sync();
		if (__hip_atomic_load(active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
	}
}

// Mass, centre of mass and velocity of one group in fp64 (kd.c:1360-1375).
__global__ __launch_bounds__(256) void k_group_cm(const int *__restrict__ bg, const int *__restrict__ gs,
		const int *__restrict__ ge, const float4 *__restrict__ uq, const float4 *__restrict__ uv, double *outAll) {
	__shared__ double sd[256];
	const int g = bg[blockIdx.x];
	double *out = outAll + 8 * blockIdx.x;
	const int s = gs[g], n = ge[g] - s, tid = threadIdx.x;
	double a[7] = {0, 0, 0, 0, 0, 0, 0};
	for (int i = tid; i < n; i += 256) {
		const float4 q = uq[s + i], v = uv[s + i];
		a[0] += (double)q.w;
		a[1] += (double)fmul(q.w, q.x); a[2] += (double)fmul(q.w, q.y); a[3] += (double)fmul(q.w, q.z);
		a[4] += (double)fmul(q.w, v.x); a[5] += (double)fmul(q.w, v.y); a[6] += (double)fmul(q.w, v.z);
	}
	for (int k = 0; k < 7; ++k) {
		sd[tid] = a[k];
		__syncthreads();
		for (int o = 128; o > 0; o >>= 1) { if (tid < o) sd[tid] += sd[tid + o]; __syncthreads(); }
		if (tid == 0) out[k] = k == 0 ? sd[0] : sd[0];
		__syncthreads();
	}
	if (tid == 0) for (int k = 1; k < 7; ++k) out[k] /= out[0];
}

// Scoop: group-0 particles within fScoop of the group centre add their
// potential to every member (kdAddScoopPot, grav.c:63-135).  One wavefront
// per group walks the group-0 tree; lanes split each bucket, then the members.
// Batched boxes (zbox given): a group scoops only group-0 particles of its own
// box, within its box's fScoop.
__global__ __launch_bounds__(64) void k_unb_scoop(const int *__restrict__ gs, const int *__restrict__ ge,
		int nG, const float4 *__restrict__ gcen, const float4 *__restrict__ grel, const float4 *__restrict__ zp,
		const float *__restrict__ zeps, int nz, const float4 *__restrict__ nlo, const float4 *__restrict__ nhi, int nLeaf2, int B,
		float L, int per, float scoop2c, const float4 *__restrict__ uq, const float4 *__restrict__ uv,
		float2 *uphi, float G, int soft, int nMax, const int *__restrict__ zbox, const int *__restrict__ mem,
		const int *__restrict__ boxOf, const DevBox *__restrict__ bt) {
	const int g = blockIdx.x + 1;
	if (g >= nG) return;
	const int s = gs[g], e = ge[g];
	if (e - s <= 0 || e - s >= nMax) return;
	const int gb = zbox ? boxOf[mem[s]] : 0;
	const float scoop2 = zbox ? bt[gb].scoop2 : scoop2c;
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
		const bool visit = brange_has(hi.w, gb) && sq3(gx, gy, gz) < scoop2;
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
					hit = sq3(dx, dy, dz) < scoop2 && (!zbox || zbox[j] == gb);
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
						const float k = pot_kernel(soft, fmaf(dx, dx, fmaf(dy, dy, dz * dz)), fadd(peps, uv[i].w));
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
		float fShift, float fCosmo, float z, float G, int soft, int bSub, int bNoUnbind, float hx, int *nunb,
		unsigned long long *npairs) {
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
		if (bSub && tid == 0) atomicAdd(npairs, (unsigned long long)n);
		if (bSub) {                                                    // kdSubPot (grav.c:39-60)
			for (int i = tid; i < n; i += UNB_T) {
				const float4 q = uq[s + i];
				const float dx = fsub(qb.x, q.x), dy = fsub(qb.y, q.y), dz = fsub(qb.z, q.z);
				const float k = pot_kernel(soft, fmaf(dx, dx, fmaf(dy, dy, dz * dz)), fadd(vb.w, uv[s + i].w));
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
		const float *__restrict__ mass, const float *__restrict__ soft, const int *__restrict__ boxOf,
		float4 *out, float *eps, int *obox) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int o = mem[i];
	out[i] = make_float4(pos[3 * o], pos[3 * o + 1], pos[3 * o + 2], mass[o]);
	eps[i] = soft[o];
	obox[i] = boxOf[o];
}
__global__ void k_gather4g(const float4 *__restrict__ in, const float *__restrict__ ein,
		const int *__restrict__ bin, const int *__restrict__ perm, int n, float4 *out, float *eout, int *bout) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) { out[i] = in[perm[i]]; eout[i] = ein[perm[i]]; bout[i] = bin[perm[i]]; }
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
	const int nMax = c->P.nMaxMembers;                      // kd.c:1330: groups of n >= nMax stay
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
	// groups by size class: <= 64 members one wavefront each, up to BIG one
	// workgroup each, larger ones one cooperative launch each
	// SKG_UNBIND_COOP_MIN overrides the size from which a group is unbound by
	// the cooperative kernel (tests use it to drive that path on small boxes).
	// Groups of at most SMALL members: one wavefront each.  Larger: the bounded
	// candidate kernel, one workgroup each.  SKG_UNBIND_COOP_MIN routes groups
	// from that size to the cooperative all-GPU kernel instead, and
	// SKG_UNBIND_SMALL_MAX lowers SMALL (tests drive both paths with these).
	// Groups of at least BIG members go to the cooperative all-GPU removal when
	// kdSubPot runs (DARK or STAR input): each removal then lowers every
	// potential, the candidate bound loosens with every removal, and one
	// workgroup took 12.9 s on a 10^6-member group (0.7 s cooperative).
	// Without kdSubPot the bound stays tight and the candidate kernel is the
	// faster one at every size (a 2^25 gas+dark box: 0.09 s vs 11 s).
	const int bSub0 = (c->inType == SKG_DARK || c->inType == SKG_STAR);   // kd.c:1441
	int BIG = bSub0 ? 100000 : INT_MAX, SMALL = 64;
	if (const char *e = getenv("SKG_UNBIND_COOP_MIN")) BIG = std::max(65, atoi(e));
	if (const char *e = getenv("SKG_UNBIND_SMALL_MAX")) SMALL = std::min(64, std::max(1, atoi(e)));
	std::vector<int2> tiles;
	std::vector<int> gsmall, gmed, gbig;
	for (int g = 1; g < nG; ++g) {
		const int m = he[g] - hs[g];
		if (m <= 0 || m >= nMax) continue;
		if (m <= SMALL) { gsmall.push_back(g); continue; }
		(m < BIG ? gmed : gbig).push_back(g);
	}
	// potential tiles (group, first target) for a workgroup of T targets
	auto make_tiles = [&](int T, std::vector<int2> &v) {
		v.clear();
		for (const std::vector<int> *L : {&gmed, &gbig})
			for (int g : *L)
				for (int t = hs[g]; t < he[g]; t += T) v.push_back(make_int2(g, t));
	};
	auto launch_pot = [&](int T, const int2 *dtl, int ntl) { return launch_unb_pot(c, G, T, dtl, ntl, nTot); };
	make_tiles(c->tunePot ? 64 : c->potTile, tiles);   // the most tiles any candidate needs
	if (gbig.size() > 128) {   // keep the 128 largest cooperative; the rest one workgroup each
		std::sort(gbig.begin(), gbig.end(), [&](int a, int b) { return he[a] - hs[a] > he[b] - hs[b]; });
		gmed.insert(gmed.end(), gbig.begin() + 128, gbig.end());
		gbig.resize(128);
	}
	// groups below CAND_WAVE members run one wavefront each (512 candidates,
	// no workgroup barriers), the rest 256 threads (2048 candidates); the
	// small ones first in the list
	const int nMedW = (int)(std::stable_partition(gmed.begin(), gmed.end(),
			[&](int g) { return he[g] - hs[g] < CAND_WAVE; }) - gmed.begin());
	SKG_HIP(hipMemsetAsync(G->d_nunb, 0, 4, c->st));
	const bool any = !gsmall.empty() || !gmed.empty() || !gbig.empty();
	if (any) {
		int2 *dt = nullptr;
		int *dgs = nullptr, *dgm = nullptr;
		double *part = nullptr, *cmv = nullptr;
		unsigned long long *npair = nullptr;
		// temporaries in the context's scratch slots (kept between calls, freed with the context)
		if (sget(c, 0, &dt, std::max<size_t>(1, tiles.size())) || sget(c, 1, &dgs, std::max<size_t>(1, gsmall.size())) ||
		    sget(c, 2, &dgm, std::max<size_t>(1, gmed.size())) || sget(c, 3, &cmv, 8) || sget(c, 4, &npair, 3))
			return SKG_E_NOMEM;
		SKG_HIP(hipMemsetAsync(npair, 0, 24, c->st));
		if (!tiles.empty()) SKG_HIP(hipMemcpyAsync(dt, tiles.data(), 8 * tiles.size(), hipMemcpyHostToDevice, c->st));
		if (!gsmall.empty()) SKG_HIP(hipMemcpyAsync(dgs, gsmall.data(), 4 * gsmall.size(), hipMemcpyHostToDevice, c->st));
		if (!gmed.empty()) SKG_HIP(hipMemcpyAsync(dgm, gmed.data(), 4 * gmed.size(), hipMemcpyHostToDevice, c->st));
		if (c->tunePot && !tiles.empty()) {
			// autotuner: time each LDS tile on these groups (setup resets the
			// potentials), with events of its own (the context's bracket this
			// call's stage time)
			hipEvent_t t0, t1;
			SKG_HIP(hipEventCreate(&t0));
			SKG_HIP(hipEventCreate(&t1));
			float best = FLT_MAX;
			for (int T : {64, 128, 256}) {
				std::vector<int2> tt;
				make_tiles(T, tt);
				SKG_HIP(hipMemcpyAsync(dt, tt.data(), 8 * tt.size(), hipMemcpyHostToDevice, c->st));
				k_unb_setup<<<nblk(nTot, 256), 256, 0, c->st>>>(G->gstart, G->gend, nG, G->mem, c->d_pos, c->d_vel,
						c->d_mass, c->d_soft, G->rel, hx, hy, hz, nMax, G->uq, G->uv, G->uphi, G->uord, nTot);
				SKG_HIP(hipEventRecord(t0, c->st));
				if (launch_pot(T, dt, (int)tt.size())) {
					(void)hipEventDestroy(t0);
					(void)hipEventDestroy(t1);
					return SKG_E_NOMEM;
				}
				SKG_HIP(hipEventRecord(t1, c->st));
				SKG_HIP(hipEventSynchronize(t1));
				float ms = 0.f;
				SKG_HIP(hipEventElapsedTime(&ms, t0, t1));
				if (ms < best) { best = ms; c->potTile = T; }
			}
			(void)hipEventDestroy(t0);
			(void)hipEventDestroy(t1);
			c->tunePot = 0;
			make_tiles(c->potTile, tiles);
			SKG_HIP(hipMemcpyAsync(dt, tiles.data(), 8 * tiles.size(), hipMemcpyHostToDevice, c->st));
		} else if (c->tunePot) {
			c->tunePot = 0;
		}
		k_unb_setup<<<nblk(nTot, 256), 256, 0, c->st>>>(G->gstart, G->gend, nG, G->mem, c->d_pos, c->d_vel,
				c->d_mass, c->d_soft, G->rel, hx, hy, hz, nMax, G->uq, G->uv, G->uphi, G->uord, nTot);
		// scoop potential from group 0 (always applied by the reference: kd.c:1379-1384)
		if (n0 > 0) {
			float4 *zp = nullptr, *zs = nullptr;
			float *ze = nullptr, *zes = nullptr;
			int *perm = nullptr, *zb = nullptr, *zbs = nullptr;
			DevTree &t = c->tfof;                  // FOF is done: its tree's arrays serve the scoop tree
			const bool batch = c->nBox > 1;
			if (sget(c, 6, &zp, n0) || sget(c, 7, &zs, n0) || sget(c, 8, &perm, n0) || sget(c, 9, &ze, n0) ||
			    sget(c, 10, &zes, n0) || sget(c, 11, &zb, n0) || sget(c, 12, &zbs, n0))
				return SKG_E_NOMEM;
			k_gather_pos_members<<<nblk(n0, 256), 256, 0, c->st>>>(G->mem + hs[0], n0, c->d_pos, c->d_mass,
					c->d_soft, c->d_boxOf, zp, ze, zb);
			if ((rc = skg_morton_sort(c, zp, nullptr, n0, perm, batch ? zb : nullptr))) return rc;
			k_gather4g<<<nblk(n0, 256), 256, 0, c->st>>>(zp, ze, zb, perm, n0, zs, zes, zbs);
			if ((rc = skg_build_tree(c, &t, zs, n0, c->P.nBucket, nullptr, batch ? zbs : nullptr))) return rc;
			const float sc2 = c->P.fScoop * c->P.fScoop;
			k_unb_scoop<<<nG - 1, 64, 0, c->st>>>(G->gstart, G->gend, nG, G->cen, G->rel, zs, zes, n0, t.lo, t.hi,
					t.nLeaf2, t.B, c->P.bPeriodic ? L[0] : FLT_MAX, c->P.bPeriodic, sc2, G->uq, G->uv, G->uphi,
					c->P.G, c->P.iSoftType, nMax, batch ? zbs : nullptr, G->mem, c->d_boxOf, c->d_bt);
		}
		double pairs = 0;
		for (int gg : gmed) pairs += (double)(he[gg] - hs[gg]) * (he[gg] - hs[gg] - 1);
		for (int gg : gbig) pairs += (double)(he[gg] - hs[gg]) * (he[gg] - hs[gg] - 1);
		const int pslot = skg_prof_begin(c);
		if (launch_pot(c->potTile, dt, (int)tiles.size())) return SKG_E_NOMEM;
		skg_prof_end(c, pslot, SKG_K_UNB_POT, 20.0 * pairs);
		const int bSub = (c->inType == SKG_DARK || c->inType == SKG_STAR);   // kd.c:1441
		const int lslot = skg_prof_begin(c);
		// the sizes run side by side: groups of at most 64 members and the
		// wavefront-sized ones on the context's second stream, forked from and
		// joined back to the main one, the 256-thread groups on the main one
		// (the groups are disjoint; the counters are atomic)
		const bool fork = (nMedW > 0 || !gsmall.empty()) && (int)gmed.size() > nMedW;
		if (fork) {
			SKG_HIP(hipEventRecord(c->evFork, c->st));
			SKG_HIP(hipStreamWaitEvent(c->st2, c->evFork, 0));
		}
		if ((int)gmed.size() > nMedW)
			k_unb_cand<CAND_T><<<(int)gmed.size() - nMedW, CAND_T, 0, c->st>>>(dgm + nMedW, G->gstart, G->gend,
					G->uq, G->uv, G->uphi, G->uord, c->d_grp, G->count, G->mass, G->vcm, G->rbound, G->rel,
					c->P.fShift, c->P.fCosmo, c->P.z, c->P.G, c->P.iSoftType, bSub, c->P.bNoUnbind, hx, G->d_nunb,
					npair, npair + 2);
		if (!gsmall.empty())
			k_unb_small<<<(int)gsmall.size(), 64, 0, fork ? c->st2 : c->st>>>(dgs, G->gstart, G->gend, G->uq, G->uv,
					G->uphi, G->uord, c->d_grp, G->count, G->mass, G->vcm, G->rbound, G->rel, c->P.fShift,
					c->P.fCosmo, c->P.z, c->P.G, c->P.iSoftType, bSub, c->P.bNoUnbind, hx, G->d_nunb, npair);
		if (nMedW > 0)
			k_unb_cand<64><<<nMedW, 64, 0, fork ? c->st2 : c->st>>>(dgm, G->gstart, G->gend, G->uq, G->uv, G->uphi,
					G->uord, c->d_grp, G->count, G->mass, G->vcm, G->rbound, G->rel, c->P.fShift, c->P.fCosmo, c->P.z,
					c->P.G, c->P.iSoftType, bSub, c->P.bNoUnbind, hx, G->d_nunb, npair, npair + 2);
		if (fork) {
			SKG_HIP(hipEventRecord(c->evJoin, c->st2));
			SKG_HIP(hipStreamWaitEvent(c->st, c->evJoin, 0));
		}
		unsigned long long hp[3] = {0, 0, 0};
		SKG_HIP(hipMemcpyAsync(&hp[0], npair, 8, hipMemcpyDeviceToHost, c->st));
		SKG_HIP(hipMemcpyAsync(&hp[2], npair + 2, 8, hipMemcpyDeviceToHost, c->st));
		SKG_HIP(hipStreamSynchronize(c->st));
		skg_prof_end(c, lslot, SKG_K_UNB_LOOP, 20.0 * (double)hp[0]);
		const int bslot = skg_prof_begin(c);
		if (!gbig.empty()) {
			// SKG_UNBIND_COOP=scan: one full scan per removal (k_unb_coop);
			// default: bounded candidates (k_unb_coopc)
			const char *ce = getenv("SKG_UNBIND_COOP");
			const bool scan = ce && std::string(ce) == "scan";
			int per = 0, ncu = 0;
			SKG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, scan ? (const void *)k_unb_coop :
					(const void *)k_unb_coopc, COOP_T, 0));
			SKG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->dev));
			const int nb = (int)gbig.size();
			int nwgAll = std::max(nb, std::max(1, std::min(per, 2)) * ncu);
			// workgroups per group in proportion to its members (at least one each)
			double tot = 0;
			for (int gg : gbig) tot += he[gg] - hs[gg];
			std::vector<int> w0(nb), wn(nb), wmap;
			int used = 0;
			for (int i = 0; i < nb; ++i) {
				const double m = he[gbig[i]] - hs[gbig[i]];
				wn[i] = std::max(1, (int)((nwgAll - nb) * m / tot) + 1);
				used += wn[i];
			}
			nwgAll = used;
			for (int i = 0, acc = 0; i < nb; ++i) { w0[i] = acc; acc += wn[i]; for (int j = 0; j < wn[i]; ++j) wmap.push_back(i); }
			int *dbg = nullptr, *dw0 = nullptr, *dwn = nullptr, *dwm = nullptr, *dact = nullptr;
			double *cmAll = nullptr;
			if (sget(c, 13, &dbg, nb) || sget(c, 14, &dw0, nb) || sget(c, 15, &dwn, nb) || sget(c, 16, &dwm, nwgAll) ||
			    sget(c, 17, &dact, 1) || sget(c, 18, &cmAll, 8 * (size_t)nb) || sget(c, 5, &part, 8 * (size_t)nwgAll))
				return SKG_E_NOMEM;
			SKG_HIP(hipMemcpyAsync(dbg, gbig.data(), 4 * nb, hipMemcpyHostToDevice, c->st));
			SKG_HIP(hipMemcpyAsync(dw0, w0.data(), 4 * nb, hipMemcpyHostToDevice, c->st));
			SKG_HIP(hipMemcpyAsync(dwn, wn.data(), 4 * nb, hipMemcpyHostToDevice, c->st));
			SKG_HIP(hipMemcpyAsync(dwm, wmap.data(), 4 * (size_t)nwgAll, hipMemcpyHostToDevice, c->st));
			SKG_HIP(hipMemcpyAsync(dact, &nb, 4, hipMemcpyHostToDevice, c->st));
			k_group_cm<<<nb, 256, 0, c->st>>>(dbg, G->gstart, G->gend, G->uq, G->uv, cmAll);
			SKG_HIP(hipStreamSynchronize(c->st));
			if (nwgAll > std::max(1, per) * ncu) return SKG_E_HIP;   // must be co-resident
			float fS = c->P.fShift, fC = c->P.fCosmo, zz = c->P.z, GG = c->P.G, hxx = hx;
			int sft = c->P.iSoftType, sub = bSub, nou = c->P.bNoUnbind;
			unsigned long long *npair1 = npair + 1, *npair2 = npair + 2;
			// A cooperative launch can be refused (another process on the device
			// may hold its cooperative resources): the same groups then go to the
			// one-workgroup candidate kernel, which gives the same removals.
			auto fallback = [&](hipError_t e) -> int {
				(void)hipGetLastError();
				c->nCoopFallbacks++;
				k_unb_cand<CAND_T><<<nb, CAND_T, 0, c->st>>>(dbg, G->gstart, G->gend, G->uq, G->uv, G->uphi, G->uord,
						c->d_grp, G->count, G->mass, G->vcm, G->rbound, G->rel, c->P.fShift, c->P.fCosmo, c->P.z,
						c->P.G, c->P.iSoftType, bSub, c->P.bNoUnbind, hx, G->d_nunb, npair1, npair + 2);
				(void)e;
				return 0;
			};
			if (scan) {
				void *args[] = {&dbg, &dw0, &dwn, &dwm, &G->gstart, &G->gend, &G->uq, &G->uv, &G->uphi, &G->uord,
						&c->d_grp, &G->count, &G->mass, &G->vcm, &G->rbound, &G->rel, &fS, &fC, &zz, &GG, &sft, &sub,
						&nou, &hxx, &G->d_nunb, &part, &cmAll, &dact, &npair1};
				const hipError_t le = hipLaunchCooperativeKernel((const void *)k_unb_coop, dim3(nwgAll), dim3(COOP_T),
						args, 0, c->st);
				if (le != hipSuccess) fallback(le);
			} else {
				// candidates per workgroup: the group's COOP_T * CPC slots shared out, at most 64
				std::vector<int> cwv(nb);
				for (int i = 0; i < nb; ++i) cwv[i] = std::max(1, std::min(64, COOP_T * CPC / wn[i]));
				int *dcw = nullptr;
				CoopCand *pubAll = nullptr;
				CoopHdr *hdrAll = nullptr;
				if (sget(c, 19, &dcw, nb) || sget(c, 20, &pubAll, (size_t)nb * COOP_T * CPC) || sget(c, 21, &hdrAll, nwgAll))
					return SKG_E_NOMEM;
				SKG_HIP(hipMemcpyAsync(dcw, cwv.data(), 4 * nb, hipMemcpyHostToDevice, c->st));
				void *args[] = {&dbg, &dw0, &dwn, &dwm, &dcw, &G->gstart, &G->gend, &G->uq, &G->uv, &G->uphi,
						&G->uord, &c->d_grp, &G->count, &G->mass, &G->vcm, &G->rbound, &G->rel, &fS, &fC, &zz, &GG,
						&sft, &sub, &nou, &hxx, &G->d_nunb, &pubAll, &hdrAll, &cmAll, &dact, &npair1, &npair2};
				const hipError_t le = hipLaunchCooperativeKernel((const void *)k_unb_coopc, dim3(nwgAll), dim3(COOP_T),
						args, 0, c->st);
				if (le != hipSuccess) fallback(le);
			}
		}
		SKG_HIP(hipGetLastError());
		SKG_HIP(hipMemcpyAsync(&hp[1], npair + 1, 8, hipMemcpyDeviceToHost, c->st));
		SKG_HIP(hipStreamSynchronize(c->st));
		skg_prof_end(c, bslot, SKG_K_UNB_BIG, 20.0 * (double)hp[1]);
		c->nBigGroups = (int)gbig.size();
		SKG_HIP(hipMemcpyAsync(&hp[2], npair + 2, 8, hipMemcpyDeviceToHost, c->st));
		SKG_HIP(hipStreamSynchronize(c->st));
		c->nFullPasses = (long long)hp[2];
	}
	// read on the context's stream: its (non-blocking) work is not ordered
	// with the null stream
	int nu = 0, c0 = 0;
	SKG_HIP(hipMemcpyAsync(&nu, G->d_nunb, 4, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipMemcpyAsync(&c0, G->count, 4, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	G->unbound = nu;
	// group 0 gains the unbound members
	c0 += nu;
	SKG_HIP(hipMemcpyAsync(G->count, &c0, 4, hipMemcpyHostToDevice, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	skg_timer_stop(c, SKG_T_UNBIND);
	if (nUnbound) *nUnbound = nu;
	c->stage = 8;
	return SKG_OK;
}

extern "C" int skg_unbind_info(skg_ctx *c, int *nBig, int *nCoopRefused, long long *nFullPasses) {
	if (!c) return SKG_E_PARAM;
	if (nBig) *nBig = c->nBigGroups;
	if (nCoopRefused) *nCoopRefused = c->nCoopFallbacks;
	if (nFullPasses) *nFullPasses = c->nFullPasses;
	return SKG_OK;
}

extern "C" int skg_too_small(skg_ctx *c) {
	if (!c) return SKG_E_PARAM;
	if (c->stage != 8 || !c->gs) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	GroupState *G = c->gs;
	const int nG = G->nG, n = c->n;
	int rc, *keep = nullptr, *off = nullptr;
	if (sget(c, 0, &keep, nG) || sget(c, 1, &off, nG)) return SKG_E_NOMEM;
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
	TmpFree tf;
	tf.add(&m2); tf.add(&r2); tf.add(&c2); tf.add(&v2); tf.add(&b2);
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
	dfree(&m2); dfree(&r2); dfree(&c2); dfree(&v2); dfree(&b2);
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


// ---------------------------------------------------------------------------
// Per-group statistics of the .stat file (kdOutStats, kd.c:1703-1839): members
// sorted by squared distance from rCenter (one radix sort of (group, r^2)
// keys), then one wavefront per group walks its members in that order with a
// running (FP64) mass: the largest G M(<r)/r outside twice the softening, the
// half-mass radius, the outer values, gas (rho >= fDensMin, T <= fTempMax) and
// star masses, and the 1-D dispersion of fExp (v - vcm) + fExpHub r.

__global__ void k_stat_keys(const int *__restrict__ grp, int n, const float *__restrict__ pos,
		const float4 *__restrict__ cen, float hx, float hy, float hz, unsigned long long *key, int *val) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int g = grp[i];
	float r2 = 0.f;
	if (g > 0) {
		const float4 c = cen[g];
		float d[3] = {fsub(pos[3 * i], c.x), fsub(pos[3 * i + 1], c.y), fsub(pos[3 * i + 2], c.z)};
		const float h[3] = {hx, hy, hz};
		for (int k = 0; k < 3; ++k) {
			if (d[k] > h[k]) d[k] = fsub(d[k], 2 * h[k]);
			if (d[k] <= -h[k]) d[k] = fadd(d[k], 2 * h[k]);
			r2 = fadd(r2, fmul(d[k], d[k]));                   // kd.c:1771-1775, in this order
		}
	}
	key[i] = ((unsigned long long)(unsigned)g << 32) | (unsigned long long)__float_as_uint(r2);
	val[i] = i;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	return v;
}

__global__ __launch_bounds__(64) void k_group_stats(const int *__restrict__ gs, const int *__restrict__ ge, int nG,
		const unsigned long long *__restrict__ key, const int *__restrict__ ord, const float *__restrict__ pos,
		const float *__restrict__ vel, const float *__restrict__ mass, const float *__restrict__ soft,
		const float *__restrict__ rho, const float *__restrict__ temp, const unsigned char *__restrict__ type,
		const float4 *__restrict__ cen, const float4 *__restrict__ vcm, const float4 *__restrict__ rbound, float G,
		float fExp, float fExpHub, float hx, float hy, float hz, float fDensMin, float fTempMax, float *out) {
	const int g = blockIdx.x + 1, lane = threadIdx.x;
	if (g >= nG) return;
	const int s = gs[g], e = ge[g], n = e - s;
	float *row = out + (size_t)g * SKG_STAT_N;
	if (n <= 0) { if (lane < SKG_STAT_N) row[lane] = 0.f; return; }
	const float4 c = cen[g], vc = vcm[g];
	double half = 0.0;
	for (int i = s + lane; i < e; i += 64) half += 0.5 * (double)mass[ord[i]];
	half = wave_sum_d(half);
	double carry = 0.0, vBest = 0.0, mBest = 0.0, mHalf = 0.0, gas = 0.0, star = 0.0, vd = 0.0;
	int iBest = INT_MAX, iHalf = INT_MAX;
	float r2Best = 0.f, r2Half = 0.f;
	for (int b = s; b < e; b += 64) {
		const int i = b + lane;
		const bool in = i < e;
		const int o = in ? ord[i] : ord[s];
		const float m = in ? mass[o] : 0.f;
		const float r2 = __uint_as_float((unsigned)key[in ? i : s]);
		double M = (double)m;
#pragma unroll
		for (int d = 1; d < 64; d <<= 1) {
			const double t = __shfl_up(M, d, 64);
			if (lane >= d) M += t;
		}
		M += carry;
		carry = __shfl(M, 63, 64);
		if (!in) continue;
		const double r = sqrt((double)r2);
		const double sf = (double)soft[o];
		if ((double)r2 > 4.0 * sf * sf) {                      // kd.c:1789-1793
			const double v = (double)G * M / r;
			if (v > vBest) { vBest = v; mBest = M; iBest = i; r2Best = r2; }
		}
		if (M > half && iHalf == INT_MAX) { iHalf = i; mHalf = M; r2Half = r2; }
		const int t = type[o];
		if (t == SKG_GAS && rho[o] >= fDensMin && temp[o] <= fTempMax) gas += (double)m;
		if (t == SKG_STAR) star += (double)m;
		float d[3] = {fsub(pos[3 * o], c.x), fsub(pos[3 * o + 1], c.y), fsub(pos[3 * o + 2], c.z)};
		const float h[3] = {hx, hy, hz}, w[3] = {vc.x, vc.y, vc.z};
		for (int k = 0; k < 3; ++k) {
			if (d[k] > h[k]) d[k] = fsub(d[k], 2 * h[k]);
			if (d[k] <= -h[k]) d[k] = fadd(d[k], 2 * h[k]);
			const double dv = (double)fExp * ((double)vel[3 * o + k] - (double)w[k]) + (double)fExpHub * (double)d[k];
			vd += dv * dv;
		}
	}
	// the first member reaching the largest circular velocity, the first past half the mass
	for (int o = 32; o > 0; o >>= 1) {
		const double ov = __shfl_xor(vBest, o, 64), om = __shfl_xor(mBest, o, 64);
		const int oi = __shfl_xor(iBest, o, 64);
		const float orr = __shfl_xor(r2Best, o, 64);
		if (ov > vBest || (ov == vBest && oi < iBest)) { vBest = ov; mBest = om; iBest = oi; r2Best = orr; }
		const int hi = __shfl_xor(iHalf, o, 64);
		const double hm = __shfl_xor(mHalf, o, 64);
		const float hr = __shfl_xor(r2Half, o, 64);
		if (hi < iHalf) { iHalf = hi; mHalf = hm; r2Half = hr; }
	}
	gas = wave_sum_d(gas);
	star = wave_sum_d(star);
	vd = wave_sum_d(vd);
	if (lane != 0) return;
	const float mTot = (float)carry;
	const float r2Out = __uint_as_float((unsigned)key[e - 1]);
	float vcirc = 0.f, rVmax = 0.f;
	if (iBest != INT_MAX) {
		rVmax = (float)sqrt((double)r2Best);
		vcirc = fmul(G, (float)mBest) / rVmax;
	}
	float rHalf = 0.f, vHalf = 0.f;
	if (iHalf != INT_MAX) {
		rHalf = (float)sqrt((double)r2Half);
		vHalf = fmul(G, (float)mHalf) / rHalf;
	}
	const float vOut = (float)((double)fmul(G, mTot) / sqrt((double)r2Out));
	if (vcirc == 0.f) { vcirc = vOut; rVmax = (float)sqrt((double)r2Out); }   // kd.c:1816-1819
	const float4 rb = rbound[g];
	const float v[SKG_STAT_N] = {mTot, (float)gas, (float)star, vcirc, vHalf, vOut, rVmax, rHalf, r2Out,
			(float)sqrt(vd / (3.0 * n)), c.x, c.y, c.z, vc.x, vc.y, vc.z, rb.x, rb.y, rb.z};
	for (int k = 0; k < SKG_STAT_N; ++k) row[k] = v[k];
}

extern "C" int skg_group_stats(skg_ctx *c, float *stats) {
	if (!c || !stats) return SKG_E_PARAM;
	if (c->stage < 9 || !c->gs) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	GroupState *G = c->gs;
	const int n = c->n, nG = G->nG;
	int rc;
	const float *L = c->P.fPeriod;
	const float hx = 0.5f * L[0], hy = 0.5f * L[1], hz = 0.5f * L[2];
	if (skg_ensure_keys(c, n)) return SKG_E_NOMEM;
	float *dst = nullptr;
	TmpFree tf;
	tf.add(&dst);
	if (dalloc(&dst, (size_t)nG * SKG_STAT_N)) return SKG_E_NOMEM;
	SKG_HIP(hipMemsetAsync(dst, 0, 4 * (size_t)SKG_STAT_N, c->st));        // group 0's row
	k_stat_keys<<<nblk(n, 256), 256, 0, c->st>>>(c->d_grp, n, c->d_pos, G->cen, hx, hy, hz, c->k64a, c->v32a);
	// (group, r^2) order: the group segments are those of group_sort's (group, iOrder) order
	if ((rc = skg_sort_pairs(c, c->k64a, c->k64b, c->v32a, c->v32b, n, 64))) return rc;
	if (nG > 1)
		k_group_stats<<<nG - 1, 64, 0, c->st>>>(G->gstart, G->gend, nG, c->k64b, c->v32b, c->d_pos, c->d_vel,
				c->d_mass, c->d_soft, c->d_rho, c->d_temp, c->d_type, G->cen, G->vcm, G->rbound, c->P.G,
				c->P.fShift, c->P.fCosmo, hx, hy, hz, c->P.fDensMin, c->P.fTempMax, dst);
	SKG_HIP(hipGetLastError());
	SKG_HIP(hipMemcpyAsync(stats, dst, 4 * (size_t)nG * SKG_STAT_N, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	dfree(&dst);
	return SKG_OK;
}

// ---------------------------------------------------------------------------
// Groups from outside: the -unbind re-entry (kdInGroup + kdInitpGroup +
// kdReadCenter, kd.c:920-1196; main.c:349-373).

// Per group g >= 1: mass, the reference point rel (member with the largest
// iOrder: kdInitpGroup's last member in input order), and, when no centres are
// given, the mass-weighted centre relative to rel and the mass-weighted
// velocity (kd.c:1160-1193).
__global__ __launch_bounds__(64) void k_group_com(const int *__restrict__ gs, const int *__restrict__ ge, int nG,
		const int *__restrict__ mem, const float *__restrict__ pos, const float *__restrict__ vel,
		const float *__restrict__ mass, float Lx, float Ly, float Lz, float cx, float cy, float cz, int withCentre,
		float *gmass, float4 *grel, float4 *gcen, float4 *gvcm) {
	const int g = blockIdx.x + 1;
	if (g >= nG) return;
	const int lane = threadIdx.x, s = gs[g], e = ge[g];
	if (e - s <= 0) return;
	const int last = mem[e - 1];
	const float rx = pos[3 * last], ry = pos[3 * last + 1], rz = pos[3 * last + 2];
	Acc am, ax, ay, az, vx, vy, vz;
	am.init(); ax.init(); ay.init(); az.init(); vx.init(); vy.init(); vz.init();
	for (int t = s + lane; t < e; t += 64) {
		const int o = mem[t];
		const float m = mass[o];
		am.add(m);
		ax.add(fmul(m, wrap_rel(fsub(pos[3 * o], rx), Lx)));
		ay.add(fmul(m, wrap_rel(fsub(pos[3 * o + 1], ry), Ly)));
		az.add(fmul(m, wrap_rel(fsub(pos[3 * o + 2], rz), Lz)));
		vx.add(fmul(m, vel[3 * o])); vy.add(fmul(m, vel[3 * o + 1])); vz.add(fmul(m, vel[3 * o + 2]));
	}
	float v[7] = {am.s, ax.s, ay.s, az.s, vx.s, vy.s, vz.s};
	float w[7] = {am.c, ax.c, ay.c, az.c, vx.c, vy.c, vz.c};
	for (int k = 0; k < 7; ++k) {
		double d = (double)v[k] + (double)w[k];
		for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o, 64);
		v[k] = (float)d;
	}
	if (lane != 0) return;
	const float M = v[0];
	gmass[g] = M;
	grel[g] = make_float4(rx, ry, rz, 0.f);
	if (!withCentre) return;
	float c3[3] = {v[1] / M, v[2] / M, v[3] / M};
	const float r3[3] = {rx, ry, rz}, C[3] = {cx, cy, cz}, L[3] = {Lx, Ly, Lz};
	for (int j = 0; j < 3; ++j) {
		float q = fadd(c3[j], r3[j]);
		if ((double)q > (double)C[j] + 0.5 * (double)L[j]) q = fsub(q, L[j]);
		if ((double)q <= (double)C[j] - 0.5 * (double)L[j]) q = fadd(q, L[j]);
		c3[j] = q;
	}
	gcen[g] = make_float4(c3[0], c3[1], c3[2], 0.f);
	gvcm[g] = make_float4(v[4] / M, v[5] / M, v[6] / M, 0.f);
}

extern "C" int skg_set_groups(skg_ctx *c, const int *grp, const float *center, const float *vcm) {
	if (!c || !grp || (center && !vcm) || (!center && vcm)) return SKG_E_PARAM;
	if (c->stage < 1) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	const int n = c->n;
	int mx = 0;
	for (int i = 0; i < n; ++i) {
		if (grp[i] < 0) return SKG_E_PARAM;
		mx = std::max(mx, grp[i]);
	}
	if (c->nBox > 1) {   // a group lives in one box
		std::vector<int> gbox(mx + 1, -1);
		for (int b = 0; b < c->nBox; ++b)
			for (int i = c->hbox[b].off; i < c->hbox[b].off + c->hbox[b].n; ++i) {
				const int g = grp[i];
				if (!g) continue;
				if (gbox[g] >= 0 && gbox[g] != b) return SKG_E_PARAM;
				gbox[g] = b;
			}
	}
	skg_group_free(c);
	GroupState *G = c->gs = new (std::nothrow) GroupState();
	if (!G) return SKG_E_NOMEM;
	int rc;
	const int nG = mx + 1;
	c->nGroup = nG;
	G->nG = nG;
	if (dalloc(&G->gstart, nG) || dalloc(&G->gend, nG) || dalloc(&G->mem, n) || dalloc(&G->mass, nG) ||
	    dalloc(&G->rel, nG) || dalloc(&G->cen, nG) || dalloc(&G->vcm, nG) || dalloc(&G->rbound, nG) ||
	    dalloc(&G->radius, nG) || dalloc(&G->count, nG) || dalloc(&G->d_nunb, 1))
		return SKG_E_NOMEM;
	skg_timer_start(c);
	SKG_HIP(hipMemcpyAsync(c->d_grp, grp, 4 * (size_t)n, hipMemcpyHostToDevice, c->st));
	SKG_HIP(hipMemsetAsync(G->mass, 0, 4 * (size_t)nG, c->st));
	SKG_HIP(hipMemsetAsync(G->rel, 0, 16 * (size_t)nG, c->st));
	SKG_HIP(hipMemsetAsync(G->cen, 0, 16 * (size_t)nG, c->st));
	SKG_HIP(hipMemsetAsync(G->vcm, 0, 16 * (size_t)nG, c->st));
	SKG_HIP(hipMemsetAsync(G->rbound, 0, 16 * (size_t)nG, c->st));
	if ((rc = group_sort(c))) return rc;
	if (center) {   // from a .gtp: rCenter = star position, vcm = star velocity (kd.c:1147-1150)
		std::vector<float4> hc(nG), hv(nG);
		hc[0] = hv[0] = make_float4(0.f, 0.f, 0.f, 0.f);
		for (int g = 1; g < nG; ++g) {
			hc[g] = make_float4(center[3 * (g - 1)], center[3 * (g - 1) + 1], center[3 * (g - 1) + 2], 0.f);
			hv[g] = make_float4(vcm[3 * (g - 1)], vcm[3 * (g - 1) + 1], vcm[3 * (g - 1) + 2], 0.f);
		}
		SKG_HIP(hipMemcpyAsync(G->cen, hc.data(), 16 * (size_t)nG, hipMemcpyHostToDevice, c->st));
		SKG_HIP(hipMemcpyAsync(G->vcm, hv.data(), 16 * (size_t)nG, hipMemcpyHostToDevice, c->st));
		SKG_HIP(hipStreamSynchronize(c->st));
	}
	const float *L = c->P.fPeriod, *C = c->P.fCenter;
	if (nG > 1)
		k_group_com<<<nG - 1, 64, 0, c->st>>>(G->gstart, G->gend, nG, G->mem, c->d_pos, c->d_vel, c->d_mass,
				L[0], L[1], L[2], C[0], C[1], C[2], center ? 0 : 1, G->mass, G->rel, G->cen, G->vcm);
	SKG_HIP(hipGetLastError());
	SKG_HIP(hipStreamSynchronize(c->st));
	skg_timer_stop(c, SKG_T_GROUP);
	c->stage = 7;
	return SKG_OK;
}

// FP64 per-group energies from the fp32 compensated member potentials:
// W_g = -1/2 sum m_i phi_i (phi from kdCellPot, members only) and
// T_g = 1/2 sum m_i |fShift (v_i - vcm) + fCosmo (r_i - rcm)|^2 with the
// FP64 centre of mass (the frame of kd.c:1394-1399).
__global__ __launch_bounds__(256) void k_group_energy(const int *__restrict__ gs, const int *__restrict__ ge, int nG,
		const float4 *__restrict__ uq, const float4 *__restrict__ uv, const float2 *__restrict__ uphi,
		float fShift, float fCosmo, double *W, double *T) {
	__shared__ double sd[256];
	const int g = blockIdx.x + 1;
	if (g >= nG) return;
	const int s = gs[g], e = ge[g], tid = threadIdx.x;
	double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	for (int i = s + tid; i < e; i += 256) {
		const float4 q = uq[i], v = uv[i];
		const double m = q.w;
		a[0] += m; a[1] += m * q.x; a[2] += m * q.y; a[3] += m * q.z;
		a[4] += m * v.x; a[5] += m * v.y; a[6] += m * v.z;
		a[7] += m * ((double)uphi[i].x + (double)uphi[i].y);
	}
	double r[8];
	for (int k = 0; k < 8; ++k) {
		sd[tid] = a[k];
		__syncthreads();
		for (int o = 128; o > 0; o >>= 1) { if (tid < o) sd[tid] += sd[tid + o]; __syncthreads(); }
		r[k] = sd[0];
		__syncthreads();
	}
	const double M = r[0] > 0 ? r[0] : 1.0;
	const double cm[7] = {r[0], r[1] / M, r[2] / M, r[3] / M, r[4] / M, r[5] / M, r[6] / M};
	double t = 0.0;
	for (int i = s + tid; i < e; i += 256) {
		const float4 q = uq[i], v = uv[i];
		const double d0 = (double)fShift * ((double)v.x - cm[4]) + (double)fCosmo * ((double)q.x - cm[1]);
		const double d1 = (double)fShift * ((double)v.y - cm[5]) + (double)fCosmo * ((double)q.y - cm[2]);
		const double d2 = (double)fShift * ((double)v.z - cm[6]) + (double)fCosmo * ((double)q.z - cm[3]);
		t += 0.5 * (double)q.w * (d0 * d0 + d1 * d1 + d2 * d2);
	}
	sd[tid] = t;
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1) { if (tid < o) sd[tid] += sd[tid + o]; __syncthreads(); }
	if (tid == 0) { W[g] = -0.5 * r[7]; T[g] = sd[0]; }
}

extern "C" int skg_group_energies(skg_ctx *c, double *W, double *T) {
	if (!c || !W || !T) return SKG_E_PARAM;
	if (c->stage != 7 || !c->gs) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	GroupState *G = c->gs;
	const int n = c->n, nG = G->nG;
	std::vector<int> hs(nG), he(nG);
	SKG_HIP(hipMemcpyAsync(hs.data(), G->gstart, 4 * (size_t)nG, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipMemcpyAsync(he.data(), G->gend, 4 * (size_t)nG, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	if (dalloc(&G->uq, n) || dalloc(&G->uv, n) || dalloc(&G->uphi, n) || dalloc(&G->uord, n)) return SKG_E_NOMEM;
	const int TT = c->potTile;
	std::vector<int2> tiles;
	for (int g = 1; g < nG; ++g)
		for (int t = hs[g]; t < he[g]; t += TT) tiles.push_back(make_int2(g, t));
	int2 *dt = nullptr;
	double *dW = nullptr, *dT = nullptr;
	TmpFree tf;
	tf.add(&dt); tf.add(&dW); tf.add(&dT);
	if (dalloc(&dt, std::max<size_t>(1, tiles.size())) || dalloc(&dW, nG) || dalloc(&dT, nG)) return SKG_E_NOMEM;
	if (!tiles.empty()) SKG_HIP(hipMemcpyAsync(dt, tiles.data(), 8 * tiles.size(), hipMemcpyHostToDevice, c->st));
	SKG_HIP(hipMemsetAsync(dW, 0, 8 * (size_t)nG, c->st));
	SKG_HIP(hipMemsetAsync(dT, 0, 8 * (size_t)nG, c->st));
	const float *L = c->P.fPeriod;
	k_unb_setup<<<nblk(n, 256), 256, 0, c->st>>>(G->gstart, G->gend, nG, G->mem, c->d_pos, c->d_vel, c->d_mass,
			c->d_soft, G->rel, 0.5f * L[0], 0.5f * L[1], 0.5f * L[2], INT_MAX, G->uq, G->uv, G->uphi, G->uord, n);
	const int nt = (int)tiles.size();
	if (nt > 0) {
		if (launch_unb_pot(c, G, TT, dt, nt, n)) return SKG_E_NOMEM;
	}
	if (nG > 1)
		k_group_energy<<<nG - 1, 256, 0, c->st>>>(G->gstart, G->gend, nG, G->uq, G->uv, G->uphi, c->P.fShift,
				c->P.fCosmo, dW, dT);
	SKG_HIP(hipGetLastError());
	SKG_HIP(hipMemcpyAsync(W, dW, 8 * (size_t)nG, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipMemcpyAsync(T, dT, 8 * (size_t)nG, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	dfree(&dt); dfree(&dW); dfree(&dT);
	return SKG_OK;
}
