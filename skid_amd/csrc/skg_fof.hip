// skg_fof.hip -- friends-of-friends of the converged positions (kd.c:802-917).
//
// The reference grows each group breadth-first from the first unvisited
// mover.  Here every mover links to every mover within tau in one parallel
// pass over a Morton-ordered tree, merging with a lock-free union-find
// (larger root hooked under the smaller, so roots are component minima).
// Link rule: f32 d2 < tau^2 evaluated with the query-side image shift
// (kd.h:217-335), from either end of the pair.  A tree node lying wholly
// inside the ball of a point (the reference's ContainedCell, kd.c:888-897)
// is merged without per-pair tests; the bound is computed from the same f32
// operands as the pair deltas, so it never admits a pair the test rejects.
// Group ids are numbered in order of each group's smallest input index.
// Batched boxes: every point links only within its own box, with its box's
// tau; the tree carries box ranges, so a lane opens only subtrees of its box.
#include "skg_common.h"
#include <climits>
#include <cfloat>
#include <mutex>

static inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

__device__ __forceinline__ int uf_find(int *par, int x) {
	for (;;) {
		int p = __hip_atomic_load(&par[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (p == x) return x;
		int gp = __hip_atomic_load(&par[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (gp != p) __hip_atomic_store(&par[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		x = gp;
	}
}

// Returns the root the two components share once merged (a root of that
// component at the time: other merges may hook it under a smaller one later).
__device__ __forceinline__ int uf_union(int *par, int a, int b) {
	for (;;) {
		a = uf_find(par, a);
		b = uf_find(par, b);
		if (a == b) return a;
		if (a > b) { int t = a; a = b; b = t; }
		if (atomicCAS(&par[b], b, a) == b) return a;
	}
}

// Largest |delta| over [lo,hi] on one axis under the query-side rule, or -1
// when points of the range would take different image shifts.
template <bool PER>
__device__ __forceinline__ float max_delta(float q, float qm, float qp, float lo, float hi, float hL) {
	float a = fsub(q, lo), b = fsub(q, hi);          // a >= b
	if (!PER) return fmaxf(fabsf(a), fabsf(b));
	if (a <= hL && b >= -hL) return fmaxf(fabsf(a), fabsf(b));
	if (b > hL) { float x = fsub(qm, lo), y = fsub(qm, hi); return fmaxf(fabsf(x), fabsf(y)); }
	if (a < -hL) { float x = fsub(qp, lo), y = fsub(qp, hi); return fmaxf(fabsf(x), fabsf(y)); }
	return -1.f;
}

template <bool PER>
__global__ __launch_bounds__(SKG_TILE_T) void k_fof_links(const float4 *__restrict__ fp, int n,
		const float4 *__restrict__ nlo, const float4 *__restrict__ nhi, int nLeaf2, int depth, int B,
		float L, float tau2c, const int *__restrict__ fbox, const DevBox *__restrict__ bt, int *par, int *united) {
	__shared__ float4 sP[SKG_TILE_T];               // staged leaf points, 64 per wavefront
	const int lane = threadIdx.x & 63, wb = threadIdx.x & ~63;
	const int base = xcd_tile(blockIdx.x, gridDim.x) * SKG_TILE_T + (threadIdx.x & ~63);
	if (base >= n) return;                        // whole wavefront past the end
	const int i = base + lane;
	const bool valid = i < n;
	const float4 q = fp[valid ? i : base];
	const int qb = fbox ? fbox[valid ? i : base] : 0;
	const float tau2 = fbox ? bt[qb].tau2 : tau2c;
	const float tau2w = wave_max(tau2);
	const float hL = 0.5f * L;
	const float qxm = fsub(q.x, L), qxp = fadd(q.x, L), qym = fsub(q.y, L), qyp = fadd(q.y, L);
	const float qzm = fsub(q.z, L), qzp = fadd(q.z, L);
	const float3 blo = make_float3(wave_min(q.x), wave_min(q.y), wave_min(q.z));
	const float3 bhi = make_float3(wave_max(q.x), wave_max(q.y), wave_max(q.z));
	int node = 1, lev = 0;
	// a node of the lane's own component (components only ever merge, so one
	// seen earlier stays one): a point whose parent is that node is already
	// linked to the lane's point, and its union is skipped
	int mine = i;
	for (;;) {
		node = uniform_int(node);
		lev = uniform_int(lev);
		const float4 lo = nlo[node], hi = nhi[node];
		bool descend = false;
		if (box_box_d2<PER>(lo, hi, blo, bhi, L) < tau2w) {
			// per-lane: contained / overlapping / disjoint
			const int first = ((node << (depth - lev)) - nLeaf2) * B;
			int last = (((node + 1) << (depth - lev)) - nLeaf2) * B;
			last = min(last, n);
			bool overlap = false, contained = false;
			if (valid && first < n && brange_has(hi.w, qb)) {
				float gx = gap_axis<PER>(lo.x, hi.x, q.x, q.x, L), gy = gap_axis<PER>(lo.y, hi.y, q.y, q.y, L);
				float gz = gap_axis<PER>(lo.z, hi.z, q.z, q.z, L);
				overlap = sq3(gx, gy, gz) < tau2;
				if (overlap && last - 1 > i && brange_is(hi.w, qb)) {
					float mx = max_delta<PER>(q.x, qxm, qxp, lo.x, hi.x, hL);
					float my = max_delta<PER>(q.y, qym, qyp, lo.y, hi.y, hL);
					float mz = max_delta<PER>(q.z, qzm, qzp, lo.z, hi.z, hL);
					contained = mx >= 0.f && my >= 0.f && mz >= 0.f && sq3(mx, my, mz) < tau2;
				}
			}
			if (contained) {
				// every member is within tau of q: merge q with the node, and the
				// node's members with each other once.
				mine = uf_union(par, i, first);
				if (atomicCAS(&united[node], 0, 1) == 0)
					for (int j = first + 1; j < last; ++j) uf_union(par, first, j);
			}
			const bool want = overlap && !contained && (last - 1 > i);
			if (__any(want)) {
				if (node < nLeaf2) descend = true;
				else {
					// the leaf's points staged 64 at a time in this wavefront's LDS
					// slots (lane t reads point c0 + t: one round trip), then each
					// lane with want reads its own range there; the wavefront reads
					// back only what it wrote, and its LDS operations complete in
					// order
					const bool oneBox = !fbox || brange_is(hi.w, qb);      // the leaf holds only the lane's box
					for (int c0 = first; c0 < last; c0 += 64) {
						const int c1 = min(c0 + 64, last);
						if (c0 + lane < c1) sP[threadIdx.x] = fp[c0 + lane];
						__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
						__builtin_amdgcn_wave_barrier();
						__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
						for (int j = max(c0, i + 1); want && j < c1; ++j) {
							if (!oneBox && fbox[j] != qb) continue;
							const float4 c = sP[wb + j - c0];
							float dx = fsub(q.x, c.x), dy = fsub(q.y, c.y), dz = fsub(q.z, c.z);
							// images only where a delta exceeds half a period (the values
							// img_delta gives); otherwise the pair seen from j is the same
							// d2 (c - q = -(q - c) exactly), so only then is it tested again
							const bool far = PER && fmaxf(fmaxf(fabsf(dx), fabsf(dy)), fabsf(dz)) > hL;
							if (far) {
								dx = img_delta<PER>(q.x, qxm, qxp, c.x, hL);
								dy = img_delta<PER>(q.y, qym, qyp, c.y, hL);
								dz = img_delta<PER>(q.z, qzm, qzp, c.z, hL);
							}
							bool link = sq3(dx, dy, dz) < tau2;
							if (far && !link) {
								// the pair seen from j: shift applied on j's side
								float ex = img_delta<PER>(c.x, fsub(c.x, L), fadd(c.x, L), q.x, hL);
								float ey = img_delta<PER>(c.y, fsub(c.y, L), fadd(c.y, L), q.y, hL);
								float ez = img_delta<PER>(c.z, fsub(c.z, L), fadd(c.z, L), q.z, hL);
								link = sq3(ex, ey, ez) < tau2;
							}
							if (link && __hip_atomic_load(&par[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != mine)
								mine = uf_union(par, mine, j);
					}
					// every lane has read these points before the next chunk overwrites them
					__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
					__builtin_amdgcn_wave_barrier();
					__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
					}
				}
			}
		}
		if (descend) { node = 2 * node; ++lev; continue; }
		while (node & 1) { node >>= 1; --lev; }
		node += 1;
		if (node == 1) break;
	}
}

__global__ void k_fof_init(int *par, int n) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) par[i] = i;
}
__global__ void k_zero(int *a, int n) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) a[i] = 0;
}
__global__ void k_fill(int *a, int n, int v) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) a[i] = v;
}

// The root of x, reading only.  Every value read is an ancestor of the node it
// was read from (parents only decrease; a root keeps itself), so the walk ends
// at x's root.
__device__ __forceinline__ int uf_root(const int *par, int x) {
	for (;;) {
		const int p = __hip_atomic_load(&par[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (p == x) return x;
		x = p;
	}
}

// After the links: par[i] = root of i, and cmin[root] = the component's smallest
// input index.  The walk here must not write: with uf_find's halving, another
// thread that had read par[i] before thread i stored the root could store a
// non-root ancestor over it afterwards, and k_fof_label would then index rank
// with cmin of a non-root (INT_MAX).  Thread i is the only writer of par[i].
__global__ void k_fof_root_min(int *par, int n, const int *__restrict__ orig, int *cmin) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int r = uf_root(par, i);
	__hip_atomic_store(&par[i], r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	atomicMin(&cmin[r], orig[i]);
}
__global__ void k_fof_first(const int *__restrict__ par, int n, const int *__restrict__ cmin, int *isFirst) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n && par[i] == i) isFirst[cmin[i]] = 1;
}
__global__ void k_fof_label(const int *__restrict__ par, int n, const int *__restrict__ orig,
		const int *__restrict__ cmin, const int *__restrict__ rank, const int *__restrict__ outmap,
		int *grp) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int o = orig[i];
	grp[outmap ? outmap[o] : o] = rank[cmin[par[i]]] + 1;
}
__global__ void k_gather4(const float4 *__restrict__ in, const int *__restrict__ perm, int n, float4 *out,
		const int *__restrict__ box, int *obox) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) {
		out[i] = in[perm[i]];
		if (box) obox[i] = box[perm[i]];
	}
}

// Storage of one FOF run: the points in curve order (fp, perm, fbox), the
// union-find arrays, the tree (tlo/thi: 2 nLeaf2 records each; null: the tree
// allocates its own) and the scan's temporary storage (null: the context's).
struct FofBufs {
	float4 *fp = nullptr;
	int *perm = nullptr, *fbox = nullptr;
	int *par = nullptr, *cmin = nullptr, *isFirst = nullptr, *rank = nullptr, *united = nullptr;
	float4 *tlo = nullptr, *thi = nullptr;
	void *scanTmp = nullptr;
	size_t scanBytes = 0;
};

// Links and labels of n points already in curve order (b.fp, with b.perm[i]
// the input index of sorted point i): writes grp[outmap ? outmap[o] : o] =
// 1-based id for every input point o, numbered by smallest o, and the number
// of groups to *ngOut.  With b.fbox (sorted like fp) points link only within
// a box, using its tau.
static int fof_label(skg_ctx *c, const FofBufs &b, int n, float tau, const int *outmap, int *grp, int *ngOut) {
	int rc;
	const bool per = c->P.bPeriodic != 0;
	const float L = per ? c->P.fPeriod[0] : FLT_MAX;
	DevTree t;
	if (b.tlo) rc = skg_build_tree_at(c, &t, b.fp, n, c->P.nBucket, b.tlo, b.thi);
	else rc = skg_build_tree(c, &t, b.fp, n, c->P.nBucket, nullptr, b.fbox);
	if (rc) { if (!b.tlo) skg_free_tree(&t); return rc; }
	int depth = 0;
	while ((1 << depth) < t.nLeaf2) ++depth;
	k_zero<<<nblk(2 * t.nLeaf2, 256), 256, 0, c->st>>>(b.united, 2 * t.nLeaf2);
	k_fof_init<<<nblk(n, 256), 256, 0, c->st>>>(b.par, n);
	const float tau2 = tau * tau;
	const int pslot = skg_prof_begin(c);
	if (per) k_fof_links<true><<<nblk(n, SKG_TILE_T), SKG_TILE_T, 0, c->st>>>(b.fp, n, t.lo, t.hi, t.nLeaf2, depth, t.B, L, tau2,
			b.fbox, c->d_bt, b.par, b.united);
	else k_fof_links<false><<<nblk(n, SKG_TILE_T), SKG_TILE_T, 0, c->st>>>(b.fp, n, t.lo, t.hi, t.nLeaf2, depth, t.B, L, tau2,
			b.fbox, c->d_bt, b.par, b.united);
	skg_prof_end(c, pslot, SKG_K_FOF, 16.0 * n);
	k_fill<<<nblk(n, 256), 256, 0, c->st>>>(b.cmin, n, INT_MAX);
	k_zero<<<nblk(n, 256), 256, 0, c->st>>>(b.isFirst, n);
	k_fof_root_min<<<nblk(n, 256), 256, 0, c->st>>>(b.par, n, b.perm, b.cmin);
	k_fof_first<<<nblk(n, 256), 256, 0, c->st>>>(b.par, n, b.cmin, b.isFirst);
	rc = b.scanTmp ? skg_exclusive_scan_at(c->st, b.scanTmp, b.scanBytes, nullptr, b.isFirst, b.rank, n)
	               : skg_exclusive_scan(c, b.isFirst, b.rank, n);
	if (!b.tlo) skg_free_tree(&t);
	if (rc) return rc;
	k_fof_label<<<nblk(n, 256), 256, 0, c->st>>>(b.par, n, b.perm, b.cmin, b.rank, outmap, grp);
	SKG_HIP(hipGetLastError());
	int last[2];
	SKG_HIP(hipMemcpyAsync(&last[0], b.rank + n - 1, 4, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipMemcpyAsync(&last[1], b.isFirst + n - 1, 4, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	*ngOut = last[0] + last[1];
	return 0;
}

// FOF of pos[0,n) (float4) on storage the library allocates for the call;
// see fof_label for the outputs.  With box (indexed like pos) points link
// only within a box, using its tau.
static int fof_run(skg_ctx *c, const float4 *pos, int n, float tau, const int *outmap, int *grp, int *ngOut,
		const int *box) {
	FofBufs b;
	auto done = [&](int status) {
		dfree(&b.perm); dfree(&b.par); dfree(&b.united); dfree(&b.cmin); dfree(&b.isFirst); dfree(&b.rank);
		dfree(&b.fp); dfree(&b.fbox);
		return status;
	};
	int rc;
	const int nLeaf2 = skg_tree_leaves2(n, c->P.nBucket, nullptr);
	if (dalloc(&b.perm, n) || dalloc(&b.par, n) || dalloc(&b.cmin, n) || dalloc(&b.isFirst, n) ||
	    dalloc(&b.rank, n) || dalloc(&b.fp, n) || dalloc(&b.united, 2 * (size_t)nLeaf2) ||
	    (box && dalloc(&b.fbox, n)))
		return done(SKG_E_NOMEM);
	if ((rc = skg_hilbert_sort(c, pos, nullptr, n, b.perm, box))) return done(rc);   // compact runs of 64 points
	k_gather4<<<nblk(n, 256), 256, 0, c->st>>>(pos, b.perm, n, b.fp, box, b.fbox);
	return done(fof_label(c, b, n, tau, outmap, grp, ngOut));
}

extern "C" int skg_fof(skg_ctx *c, int *grp, int *nGroup) {
	if (!c) return SKG_E_PARAM;
	if (c->stage != 4) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	int rc, ng = 0;
	skg_timer_start(c);
	k_zero<<<nblk(c->n, 256), 256, 0, c->st>>>(c->d_grp, c->n);
	if (c->nMove > 0) {
		if ((rc = fof_run(c, c->m_pos, c->nMove, c->P.fTau, c->m_ord, c->d_grp, &ng,
				c->nBox > 1 ? c->m_box : nullptr))) return rc;
	}
	skg_timer_stop(c, SKG_T_FOF);
	c->nGroup = ng + 1;                                     // kd.c:905
	if (grp) {
		SKG_HIP(hipMemcpyAsync(grp, c->d_grp, 4 * (size_t)c->n, hipMemcpyDeviceToHost, c->st));
		SKG_HIP(hipStreamSynchronize(c->st));
	}
	if (nGroup) *nGroup = c->nGroup;
	c->stage = 5;
	return SKG_OK;
}

extern "C" int skg_fof_points(skg_ctx *c, int n, const float *pos, float fTau, int *grp, int *nGroup) {
	if (!c || n < 0 || (n && (!pos || !grp)) || !(fTau > 0.f)) return SKG_E_PARAM;
	if (n == 0) { if (nGroup) *nGroup = 1; return SKG_OK; }
	(void)hipSetDevice(c->dev);
	float4 *p4 = nullptr;
	int *g = nullptr, rc, ng = 0;
	TmpFree tf;
	tf.add(&p4); tf.add(&g);
	if (dalloc(&p4, n) || dalloc(&g, n)) return SKG_E_NOMEM;
	float4 *h = (float4 *)malloc(16 * (size_t)n);
	if (!h) return SKG_E_NOMEM;
	for (int i = 0; i < n; ++i) h[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.f);
	hipError_t e = hipMemcpyAsync(p4, h, 16 * (size_t)n, hipMemcpyHostToDevice, c->st);
	if (e == hipSuccess) e = hipStreamSynchronize(c->st);
	free(h);
	SKG_HIP(e);
	if ((rc = fof_run(c, p4, n, fTau, nullptr, g, &ng, nullptr))) return rc;
	SKG_HIP(hipMemcpyAsync(grp, g, 4 * (size_t)n, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	dfree(&p4); dfree(&g);
	if (nGroup) *nGroup = ng + 1;
	return SKG_OK;
}

// ---------------------------------------------------------------------------
// FOF of device-resident points without allocating (skg_fof_points_noalloc).
// Every temporary of the call lives in `scratch`, one file-scope device array
// per GPU.  Host code never passes the array's name to a kernel: the name is a
// host-side stand-in, and the device address of each GPU's copy comes from
// hipGetSymbolAddress, looked up once per device by skg_noalloc_setup (run by
// skg_create with the context's device current) and kept in g_scr_ptr.
static __device__ __align__(256) unsigned char scratch[SKG_NOALLOC_SCRATCH_BYTES];

#define SKG_SCR_DEVICES 64
static std::mutex g_scr_mu;
static void *g_scr_ptr[SKG_SCR_DEVICES];
static int g_scr_err[SKG_SCR_DEVICES];       // hipError_t of a failed lookup (0: none)

int skg_noalloc_setup(int dev) {
	if (dev < 0 || dev >= SKG_SCR_DEVICES) return SKG_E_PARAM;
	std::lock_guard<std::mutex> lk(g_scr_mu);
	if (g_scr_ptr[dev]) return SKG_OK;
	void *p = nullptr;
	const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
	if (e != hipSuccess || !p) {
		g_scr_err[dev] = e != hipSuccess ? (int)e : (int)hipErrorInvalidSymbol;
		skg_note_hip((hipError_t)g_scr_err[dev], "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
		return SKG_E_HIP;
	}
	g_scr_ptr[dev] = p;
	g_scr_err[dev] = 0;
	return SKG_OK;
}

static unsigned char *scratch_of(int dev) {
	if (dev < 0 || dev >= SKG_SCR_DEVICES) return nullptr;
	std::lock_guard<std::mutex> lk(g_scr_mu);
	return (unsigned char *)g_scr_ptr[dev];
}

__global__ void k_pack4(const float *__restrict__ p3, int n, float4 *p4) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p4[i] = make_float4(p3[3 * i], p3[3 * i + 1], p3[3 * i + 2], 0.f);
}

// Byte offsets into the scratch array of one call's arrays (each 256-aligned).
struct ScrLayout {
	size_t p4, fp, perm, ka, kb, va, part, bb, tlo, thi, united, par, cmin, isFirst, rank, sortTmp, scanTmp, end;
	size_t sortBytes, scanBytes;
};

static int scr_layout(hipStream_t st, int n, int B, ScrLayout *L) {
	size_t off = 0;
	auto take = [&](size_t bytes) { const size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
	const size_t N = (size_t)n, nLeaf2 = (size_t)skg_tree_leaves2(n, B, nullptr);
	int rc;
	if ((rc = skg_sort_pairs_at(st, nullptr, 0, &L->sortBytes, nullptr, nullptr, nullptr, nullptr, n, 63)) ||
	    (rc = skg_exclusive_scan_at(st, nullptr, 0, &L->scanBytes, nullptr, nullptr, n)))
		return rc;
	L->p4 = take(16 * N); L->fp = take(16 * N); L->perm = take(4 * N);
	L->ka = take(8 * N); L->kb = take(8 * N); L->va = take(4 * N);
	L->part = take(24 * (size_t)nblk(n, 256)); L->bb = take(24);
	L->tlo = take(32 * nLeaf2); L->thi = take(32 * nLeaf2); L->united = take(8 * nLeaf2);
	L->par = take(4 * N); L->cmin = take(4 * N); L->isFirst = take(4 * N); L->rank = take(4 * N);
	L->sortTmp = take(L->sortBytes); L->scanTmp = take(L->scanBytes);
	L->end = off;
	return SKG_OK;
}

extern "C" int skg_fof_points_noalloc(skg_ctx *c, int n, const float *d_pos, float fTau, int *d_grp,
		int *nGroup) {
	if (!c || n < 0 || (n && (!d_pos || !d_grp)) || !(fTau > 0.f)) return SKG_E_PARAM;
	if (n > SKG_NOALLOC_MAX_POINTS) return SKG_E_PARAM;
	if (n == 0) { if (nGroup) *nGroup = 1; return SKG_OK; }
	(void)hipSetDevice(c->dev);
	unsigned char *s = scratch_of(c->dev);
	if (!s) {
		const int e = (c->dev >= 0 && c->dev < SKG_SCR_DEVICES) ? g_scr_err[c->dev] : 0;
		skg_note_hip((hipError_t)(e ? e : (int)hipErrorInvalidSymbol), "scratch address (skg_noalloc_setup)",
				__FILE__, __LINE__);
		return SKG_E_HIP;
	}
	ScrLayout L;
	int rc;
	if ((rc = scr_layout(c->st, n, c->P.nBucket, &L))) return rc;
	if (L.end > (size_t)SKG_NOALLOC_SCRATCH_BYTES) return SKG_E_PARAM;   // only with nBucket < 4
	float4 *p4 = (float4 *)(s + L.p4);
	unsigned long long *ka = (unsigned long long *)(s + L.ka), *kb = (unsigned long long *)(s + L.kb);
	int *va = (int *)(s + L.va);
	FofBufs b;
	b.fp = (float4 *)(s + L.fp); b.perm = (int *)(s + L.perm);
	b.par = (int *)(s + L.par); b.cmin = (int *)(s + L.cmin); b.isFirst = (int *)(s + L.isFirst);
	b.rank = (int *)(s + L.rank); b.united = (int *)(s + L.united);
	b.tlo = (float4 *)(s + L.tlo); b.thi = (float4 *)(s + L.thi);
	b.scanTmp = s + L.scanTmp; b.scanBytes = L.scanBytes;
	// the same kernels as skg_fof_points: pack, Hilbert order, gather, links, labels
	k_pack4<<<nblk(n, 256), 256, 0, c->st>>>(d_pos, n, p4);
	if ((rc = skg_curve_keys(c, p4, nullptr, n, nullptr, true, ka, va, (float *)(s + L.part), (unsigned *)(s + L.bb))))
		return rc;
	if ((rc = skg_sort_pairs_at(c->st, s + L.sortTmp, L.sortBytes, nullptr, ka, kb, va, b.perm, n, 63))) return rc;
	k_gather4<<<nblk(n, 256), 256, 0, c->st>>>(p4, b.perm, n, b.fp, nullptr, nullptr);
	int ng = 0;
	if ((rc = fof_label(c, b, n, fTau, nullptr, d_grp, &ng))) return rc;
	if (nGroup) *nGroup = ng + 1;
	return SKG_OK;
}
