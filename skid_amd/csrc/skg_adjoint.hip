// skg_adjoint.hip -- reverse-mode derivative of the M4 density and the frozen
// density field's gradient at arbitrary points.
//
// Density (smooth1.c:249-263): rho_j = sum_{i in N_j} m_i W(r_ij, h_j)/2
//                                    + sum_{i : j in N_i} m_i W(r_ij, h_i)/2,
// W(r, h) = w(r/h)/(pi h^3), h = sqrt(fBall2)/2, N_i the kNN set of i.  With
// the sets and radii frozen, every term is a pair (i, j in N_i) evaluated with
// h_i, so the vector-Jacobian product for weights gbar is
//   c_ij  = dW/dr(r_ij, h_i)/2 * (gbar_i m_j [gather] + gbar_j m_i [scatter])
//   dx_i += c_ij (r_i - r_j)/r,   dx_j -= c_ij (r_i - r_j)/r,
//   dh_i += dW/dh(r_ij, h_i)/2 * (gbar_i m_j + gbar_j m_i),
// with dW/dr = w'(q)/(pi h^4) and dW/dh = -(3 w(q) + q w'(q))/(pi h^4).  The
// first two sums run per query over its kNN list; the third (each point's
// share as a neighbour) runs per point over the balls that hold it, so
// neither pass needs atomics.  Since g(q) = w'(q)/q, the reference's gradient
// (smooth1.c:447-459) is a = -2 d(rho_scatter)/dx at a point.
#include "skg_common.h"
#include <cfloat>
#include <vector>

static inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

// exact periodic separation a - b under the nearest-image rule
__device__ __forceinline__ float per_delta(float a, float b, float L, int per) {
	float d = fsub(a, b);
	if (per) {
		const float hL = 0.5f * L;
		if (d > hL) d = exact_delta(a, b, -L);
		else if (d < -hL) d = exact_delta(a, b, L);
	}
	return d;
}

// Pass 1: per query i, over its kNN list (heap keys; slot 0 is the head).
// The lists hold iOrder; each neighbour's position, mass and weight are read in
// scatterer (curve) order through sinv (iOrder -> scatterer index), so that the
// reads of a wavefront's neighbours fall on nearby lines.
__global__ __launch_bounds__(256) void k_adj_gather(const unsigned long long *__restrict__ heap, int k,
		const int *__restrict__ sord, const int *__restrict__ sinv, const float4 *__restrict__ spos,
		const float *__restrict__ gs, const float *__restrict__ sb2, int nS, int parts, float L, int per,
		float *dx, float *dh) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nS) return;
	const int oi = sord[i];
	const float4 pi = spos[i];
	const float xi = pi.x, yi = pi.y, zi = pi.z, mi = pi.w, gi = gs[i];
	const float h = 0.5f * sqrtf(sb2[i]), ih = 1.0f / h, norm = (float)(M_1_PI) * ih * ih * ih * ih;
	Sum<true> ax, ay, az, ahh;
	ax.init(); ay.init(); az.init(); ahh.init();
	for (int s = 1; s < k; ++s) {
		const unsigned long long key = heap[(size_t)s * nS + i];     // slot-major: coalesced
		const int sj = sinv[(int)(unsigned)key];
		const float4 pj = spos[sj];
		float dxv = 0.f, dyv = 0.f, dzv = 0.f, r = 0.f;
		if (sj != i) {
			dxv = per_delta(xi, pj.x, L, per); dyv = per_delta(yi, pj.y, L, per);
			dzv = per_delta(zi, pj.z, L, per);
			r = sqrtf(sq3(dxv, dyv, dzv));
		}
		float w, wp;
		m4_derivs(r * ih, w, wp);
		const float coef = ((parts & 1) ? gi * pj.w : 0.f) + ((parts & 2) ? gs[sj] * mi : 0.f);
		// the self pair (r = 0, W = 1/(pi h^3)) moves with h but not with x
		ahh.add(-0.5f * (3.0f * w + r * ih * wp) * norm * coef);
		if (!(r > 0.f)) continue;
		const float c = 0.5f * wp * norm * coef / r;
		ax.add(c * dxv); ay.add(c * dyv); az.add(c * dzv);
	}
	dx[3 * oi] = ax.val(); dx[3 * oi + 1] = ay.val(); dx[3 * oi + 2] = az.val();
	dh[oi] = ahh.val();
}

// The weights in scatterer order.
__global__ void k_scat_weights(const int *__restrict__ sord, int nS, const float *__restrict__ g, float *gs) {
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s < nS) gs[s] = g[sord[s]];
}

// Sensitivity of the density to nSmooth with the neighbour sets frozen: all
// smoothing lengths scale as nSmooth^(1/3), so h dW/dh = -(3w + q w')/(pi h^3)
// gives d rho_i / d ln k = -(1/6) [sum_{j in N_i} m_j (3w + q w')(q_ij/h_i)/(pi h_i^3)
//                                 + sum_{j : i in N_j} m_j (3w + q w')(q_ij/h_j)/(pi h_j^3)].
// Gather half here (per query over its kNN list, self at q = 0 included);
// the scatter half is k_adj_cover MODE 2.
__global__ __launch_bounds__(256) void k_dlnk_gather(const unsigned long long *__restrict__ heap, int k,
		const int *__restrict__ sord, const float *__restrict__ sb2, int nS, const float *__restrict__ pos,
		const float *__restrict__ mass, float L, int per, float *out) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nS) return;
	const int oi = sord[i];
	const float xi = pos[3 * oi], yi = pos[3 * oi + 1], zi = pos[3 * oi + 2];
	const float h = 0.5f * sqrtf(sb2[i]), ih = 1.0f / h;
	Sum<true> a;
	a.init();
	for (int s = 1; s < k; ++s) {
		const int j = (int)(unsigned)heap[(size_t)s * nS + i];
		float r = 0.f;
		if (j != oi) {
			const float dx = per_delta(xi, pos[3 * j], L, per), dy = per_delta(yi, pos[3 * j + 1], L, per);
			const float dz = per_delta(zi, pos[3 * j + 2], L, per);
			r = sqrtf(sq3(dx, dy, dz));
		}
		const float q = r * ih;
		float w, wp;
		m4_derivs(q, w, wp);
		a.add(mass[j] * (3.0f * w + q * wp));
	}
	out[oi] = a.val() * (float)(-M_1_PI / 6.0) * ih * ih * ih;
}

// Pass 2 (and the field gradient): per point, every entity (scatterer or
// periodic replica) whose ball strictly holds it, over the entity tree with
// the radii as built.  MODE 0: dx_j -= c_ij r_hat; MODE 1: a(x); MODE 2 the
// scatter half of d rho / d ln k (out[ii], one value).  pbox
// (batched boxes) restricts each point to its own box's entities.
template <int MODE>
__global__ __launch_bounds__(SKG_TILE_T) void k_adj_cover(const float4 *__restrict__ pts, int np,
		const float4 *__restrict__ epos, const float4 *__restrict__ eaux, const float *__restrict__ eb2,
		const float4 *__restrict__ eorg, const int *__restrict__ ecode, const float *__restrict__ ew, int ne,
		const float4 *__restrict__ nlo, const float4 *__restrict__ nhi, int nLeaf2, int B,
		const float *__restrict__ mass, const float *__restrict__ gbar, const int *__restrict__ pord,
		int parts, float Lx, float Ly, float Lz, const int *__restrict__ pbox, const int *__restrict__ lstart,
		float *out) {
	const int lane = threadIdx.x & 63, base = xcd_tile(blockIdx.x, gridDim.x) * SKG_TILE_T + (threadIdx.x & ~63),
			ii = base + lane;
	// staging slots for an opened leaf's entity records, 64 per wavefront
	__shared__ float4 sA[SKG_TILE_T];
	__shared__ float sB[SKG_TILE_T];
	__shared__ int sK[SKG_TILE_T];
	const int wb = threadIdx.x & ~63;
	if (base >= np) return;                       // whole wavefront past the end
	const bool valid = ii < np;
	const float4 p = pts[valid ? ii : base];
	const int pb = pbox ? pbox[valid ? ii : base] : 0;
	float mj = 0.f, gj = 0.f;
	if (MODE == 0 && valid) { const int oj = pord[ii]; mj = mass[oj]; gj = gbar[oj]; }
	Sum<true> s0, s1, s2;
	s0.init(); s1.init(); s2.init();
	// One entity e (position and mass A, radius as built b2, image code K)
	// against the lane's point.
	auto hit = [&](int e, const float4 &A, float b2, int K) {
		float dx = fsub(A.x, p.x), dy = fsub(A.y, p.y), dz = fsub(A.z, p.z);
		if (!(sq3(dx, dy, dz) < b2)) return;
		if (K != 13) {                            // replica: source position plus the exact shift
			const float4 O = eorg[e];
			dx = exact_delta(O.x, p.x, (float)(K / 9 - 1) * Lx);
			dy = exact_delta(O.y, p.y, (float)((K / 3) % 3 - 1) * Ly);
			dz = exact_delta(O.z, p.z, (float)(K % 3 - 1) * Lz);
		}
		if (MODE == 1) {
			const float4 X = eaux[e];
			const float f = grad_factor(sq3(dx, dy, dz), X.y, X.z);
			s0.add(fmul(dx, f)); s1.add(fmul(dy, f)); s2.add(fmul(dz, f));
		} else if (MODE == 2) {
			const float ih = 2.0f / sqrtf(b2), q = sqrtf(sq3(dx, dy, dz)) * ih;
			float w, wp;
			m4_derivs(q, w, wp);
			s0.add(A.w * (3.0f * w + q * wp) * ih * ih * ih);
		} else {
			// hardware square root, reciprocal square root and reciprocal (about
			// 1 ulp each) for r, 1/h and 1/r
			const float r = __builtin_amdgcn_sqrtf(sq3(dx, dy, dz));
			if (!(r > 0.f)) return;
			const float ih = 2.0f * __builtin_amdgcn_rsqf(b2);
			float w, wp;
			m4_derivs(r * ih, w, wp);
			const float coef = ((parts & 1) ? ew[e] * mj : 0.f) + ((parts & 2) ? gj * A.w : 0.f);
			const float c = 0.5f * (float)M_1_PI * wp * ih * ih * ih * ih * coef * __builtin_amdgcn_rcpf(r);
			// dx_j -= c (r_i - r_j)/r with (r_i - r_j) = (dx, dy, dz)
			s0.add(-c * dx); s1.add(-c * dy); s2.add(-c * dz);
		}
	};
	int node = 1;
	for (;;) {
		node = uniform_int(node);
		const float4 lo = nlo[node], hi = nhi[node];
		const bool need = valid && brange_has(hi.w, pb) && point_box_d2_np(lo, hi, p) < lo.w;
		const bool visit = __any(need);
		if (visit && node < nLeaf2) { node = 2 * node; continue; }
		if (visit) {
			const int2 rg = leaf_range(lstart, node - nLeaf2, B, ne);
			const int start = uniform_int(rg.x), cnt = uniform_int(rg.y - rg.x);
			// the leaf's records staged in this wavefront's LDS slots, 64 at a
			// time (lane j reads record j: one round trip), then read by every
			// lane with need at one address; the wavefront reads back only what
			// it wrote, and its LDS operations complete in order
			for (int c0 = 0; c0 < cnt; c0 += 64) {
				const int m = min(64, cnt - c0);
				if (lane < m) {
					sA[threadIdx.x] = epos[start + c0 + lane];
					sB[threadIdx.x] = eb2[start + c0 + lane];
					sK[threadIdx.x] = ecode[start + c0 + lane];
				}
				__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
				__builtin_amdgcn_wave_barrier();
				__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
				if (need)
					for (int t = 0; t < m; ++t) hit(start + c0 + t, sA[wb + t], sB[wb + t], sK[wb + t]);
				// every lane has read these records before the next chunk overwrites them
				__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
				__builtin_amdgcn_wave_barrier();
				__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
			}
		}
		int nx = node;
		while (nx & 1) nx >>= 1;
		node = nx + 1;
		if (node == 1) break;
	}
	if (!valid) return;
	if (MODE == 2) { out[ii] = s0.val() * (float)(-M_1_PI / 6.0); return; }
	out[3 * ii] = s0.val(); out[3 * ii + 1] = s1.val(); out[3 * ii + 2] = s2.val();
}

// Each entity's weight for the dx adjoint: gbar of its source scatterer (0 for
// fillers).
__global__ void k_entity_weights(const int *__restrict__ esrc, const int *__restrict__ sord, int ne,
		const float *__restrict__ gbar, float *ew) {
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e < ne) ew[e] = esrc[e] >= 0 ? gbar[sord[esrc[e]]] : 0.f;
}

__global__ void k_full_b2(const float *__restrict__ eb2, int ne, float *b2) {
	int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e < ne) b2[e] = eb2[e];
}
__global__ void k_gather_pts(const float4 *__restrict__ in, const int *__restrict__ perm, int n, float4 *out) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) out[i] = in[perm[i]];
}
__global__ void k_unpermute3(const float *__restrict__ in, const int *__restrict__ perm, int n, float *out) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int o = perm[i];
	out[3 * o] = in[3 * i]; out[3 * o + 1] = in[3 * i + 1]; out[3 * o + 2] = in[3 * i + 2];
}
__global__ void k_add_ord(const float *__restrict__ in, const int *__restrict__ ord, int n, float *out) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) out[ord[i]] = fadd(out[ord[i]], in[i]);
}
__global__ void k_add3_ord(const float *__restrict__ in, const int *__restrict__ ord, int n, float *dx) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int o = ord[i];
	for (int c = 0; c < 3; ++c) dx[3 * o + c] = fadd(dx[3 * o + c], in[3 * i + c]);
}

__global__ void k_dlnk_combine(const float *__restrict__ g, const float *__restrict__ s, const int *__restrict__ ord,
		int n, float *out) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) out[ord[i]] = fadd(g[i], s[i]);
}

// The entity tree with every radius as built (independent of later ScatterCuts):
// the move tree's leaves, refitted with the radii as built.
static int full_tree(skg_ctx *c) {
	if (c->tfull.lo && c->tfull.n == c->nE) return 0;
	DevTree *t = &c->tfull;
	if (dalloc(&t->lo, 2 * (size_t)c->te.nLeaf2) || dalloc(&t->hi, 2 * (size_t)c->te.nLeaf2)) return SKG_E_NOMEM;
	int rc;
	if ((rc = skg_copy_tree_leaves(c, t, &c->te))) return rc;
	t->n = c->te.n; t->B = c->te.B; t->nLeaf = c->te.nLeaf; t->nLeaf2 = c->te.nLeaf2;
	return skg_refit_tree(c, t, c->e_pos, c->e_b2, c->e_box);
}

template <int MODE>
static void launch_cover(skg_ctx *c, const float4 *pts, int np, const int *pord, const float *gbar, int parts,
		const int *pbox, float *out, const float *ew = nullptr) {
	const float *L = c->P.fPeriod;
	const bool per = c->P.bPeriodic != 0;
	k_adj_cover<MODE><<<nblk(np, SKG_TILE_T), SKG_TILE_T, 0, c->st>>>(pts, np, c->e_pos, c->e_aux, c->e_b2, c->e_org,
			c->e_code, ew, c->nE, c->tfull.lo, c->tfull.hi, c->tfull.nLeaf2, c->tfull.B, c->d_mass, gbar, pord, parts,
			per ? L[0] : 0.f, per ? L[1] : 0.f, per ? L[2] : 0.f, c->nBox > 1 ? pbox : nullptr, c->tfull.lstart, out);
}

int skg_gradient_points(skg_ctx *c, const float4 *q, int nq, float *out) {
	int rc;
	if ((rc = full_tree(c))) return rc;
	int *perm = nullptr;
	float4 *qs = nullptr;
	float *tmp = nullptr;
	if (sget(c, 2, &perm, nq) || sget(c, 3, &qs, nq) || sget(c, 4, &tmp, 3 * (size_t)nq)) return SKG_E_NOMEM;
	if ((rc = skg_morton_sort(c, q, nullptr, nq, perm, nullptr))) return rc;
	k_gather_pts<<<nblk(nq, 256), 256, 0, c->st>>>(q, perm, nq, qs);
	launch_cover<1>(c, qs, nq, nullptr, nullptr, 0, nullptr, tmp);
	k_unpermute3<<<nblk(nq, 256), 256, 0, c->st>>>(tmp, perm, nq, out);
	SKG_HIP(hipGetLastError());
	return 0;
}

extern "C" int skg_density_adjoint(skg_ctx *c, const float *gbar, int parts, float *dx, float *dh) {
	if (!c || !gbar || !dx || !dh || parts < 1 || parts > 3) return SKG_E_PARAM;
	if (c->stage < 2 || !c->s_heap || c->nS == 0) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	int rc;
	const int n = c->n, nS = c->nS;
	float *dg = nullptr, *ddx = nullptr, *ddh = nullptr, *part2 = nullptr, *gs = nullptr, *ew = nullptr;
	if (sget(c, 0, &dg, n) || sget(c, 1, &ddx, 3 * (size_t)n) || sget(c, 2, &ddh, n) || sget(c, 3, &part2, 3 * (size_t)nS) ||
	    sget(c, 5, &gs, nS) || sget(c, 6, &ew, (size_t)c->nE + SKG_EPAD))
		return SKG_E_NOMEM;
	SKG_HIP(hipMemcpyAsync(dg, gbar, 4 * (size_t)n, hipMemcpyHostToDevice, c->st));
	SKG_HIP(hipMemsetAsync(ddx, 0, 12 * (size_t)n, c->st));
	SKG_HIP(hipMemsetAsync(ddh, 0, 4 * (size_t)n, c->st));
	const bool per = c->P.bPeriodic != 0;
	k_scat_weights<<<nblk(nS, 256), 256, 0, c->st>>>(c->s_ord, nS, dg, gs);
	k_adj_gather<<<nblk(nS, 256), 256, 0, c->st>>>(c->s_heap, c->P.nSmooth, c->s_ord, c->s_inv, c->s_pos, gs, c->s_ball2,
			nS, parts, per ? c->P.fPeriod[0] : 0.f, per, ddx, ddh);
	if ((rc = full_tree(c))) return rc;
	k_entity_weights<<<nblk(c->nE, 256), 256, 0, c->st>>>(c->e_src, c->s_ord, c->nE, dg, ew);
	launch_cover<0>(c, c->s_pos, nS, c->s_ord, dg, parts, c->s_box, part2, ew);
	k_add3_ord<<<nblk(nS, 256), 256, 0, c->st>>>(part2, c->s_ord, nS, ddx);
	SKG_HIP(hipGetLastError());
	SKG_HIP(hipMemcpyAsync(dx, ddx, 12 * (size_t)n, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipMemcpyAsync(dh, ddh, 4 * (size_t)n, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	return SKG_OK;
}

extern "C" int skg_density_dlnk(skg_ctx *c, float *drho) {
	if (!c || !drho) return SKG_E_PARAM;
	if (c->stage < 2 || !c->s_heap || c->nS == 0) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	int rc;
	const int n = c->n, nS = c->nS;
	float *d = nullptr, *part = nullptr;
	if (sget(c, 0, &d, n) || sget(c, 1, &part, nS)) return SKG_E_NOMEM;
	SKG_HIP(hipMemsetAsync(d, 0, 4 * (size_t)n, c->st));
	if (c->dlnkReady) {
		// both halves were formed by the last skg_density's kernels
		k_dlnk_combine<<<nblk(nS, 256), 256, 0, c->st>>>(c->s_dlnkG, c->s_dlnkS, c->s_ord, nS, d);
	} else {
		const bool per = c->P.bPeriodic != 0;
		k_dlnk_gather<<<nblk(nS, 256), 256, 0, c->st>>>(c->s_heap, c->P.nSmooth, c->s_ord, c->s_ball2, nS, c->d_pos,
				c->d_mass, per ? c->P.fPeriod[0] : 0.f, per, d);
		if ((rc = full_tree(c))) return rc;
		launch_cover<2>(c, c->s_pos, nS, c->s_ord, nullptr, 0, c->s_box, part);
		k_add_ord<<<nblk(nS, 256), 256, 0, c->st>>>(part, c->s_ord, nS, d);
		c->wantDlnk = 1;                          // later skg_density calls form the halves themselves
	}
	SKG_HIP(hipGetLastError());
	SKG_HIP(hipMemcpyAsync(drho, d, 4 * (size_t)n, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	return SKG_OK;
}
