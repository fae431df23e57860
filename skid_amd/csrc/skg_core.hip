// skg_core.hip -- context, trees, kNN + density, and the Denmax move loop.
//
// Stage map to the reference (SKID v1.4.1):
//   skg_density     kdScatterActive + kdBuildTree + smDensityInit (main.c:374-378)
//   skg_init_move   kdInitMove (kd.c:630-666)
//   skg_move_first  main.c:395-404;  skg_move_block  main.c:410-416
//   skg_microsteps  main.c:431-438
#include "skg_common.h"
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cstring>
#include <vector>
#include <chrono>
#include <mutex>
#include <type_traits>
#include <unordered_map>

static const unsigned long long KEY_SENT = 0x7F800000FFFFFFFFull;   // d2 = +inf, index = max

// =====================================================================
// scratch, timers, sorting
// =====================================================================
// The tracker lives on the heap and is never destroyed: a context released
// late in process exit (after this library's static destructors) still finds
// it intact.
static std::mutex &mem_mu() { static std::mutex *m = new std::mutex; return *m; }
static std::unordered_map<void *, size_t> &mem_live() {
	static std::unordered_map<void *, size_t> *m = new std::unordered_map<void *, size_t>;
	return *m;
}
static long long g_mem_cur = 0, g_mem_peak = 0;

// An array that already holds an allocation keeps it when that is large
// enough and at most twice the request (plus 64 MiB): the stages re-size their
// arrays on every call, and a hipFree synchronises the device.
int skg_dmalloc(void **p, size_t bytes) {
	if (*p) {
		std::lock_guard<std::mutex> lk(mem_mu());
		auto it = mem_live().find(*p);
		if (it != mem_live().end() && it->second >= bytes && it->second <= 2 * bytes + ((size_t)64 << 20)) return 0;
	}
	skg_dfree(p);
	if (hipMalloc(p, bytes) != hipSuccess) { *p = nullptr; return SKG_E_NOMEM; }
	std::lock_guard<std::mutex> lk(mem_mu());
	mem_live()[*p] = bytes;
	g_mem_cur += (long long)bytes;
	g_mem_peak = std::max(g_mem_peak, g_mem_cur);
	return 0;
}

// Per-context scratch slots: a stage's temporaries live in numbered slots that
// keep their allocation from call to call (freed by skg_destroy), so a stage
// entry point allocates only when an input outgrows the last call's.
int skg_slot(skg_ctx *c, int s, void **p, size_t bytes) {
	if (s < 0 || s >= SKG_NSLOT) return SKG_E_PARAM;
	if (bytes == 0) bytes = 1;
	if (c->slotBytes[s] < bytes || !c->slot[s]) {
		skg_dfree(&c->slot[s]);
		c->slotBytes[s] = 0;
		if (skg_dmalloc(&c->slot[s], bytes)) return SKG_E_NOMEM;
		c->slotBytes[s] = bytes;
	}
	*p = c->slot[s];
	return 0;
}

void skg_dfree(void **p) {
	if (!*p) return;
	{
		std::lock_guard<std::mutex> lk(mem_mu());
		auto it = mem_live().find(*p);
		if (it != mem_live().end()) { g_mem_cur -= (long long)it->second; mem_live().erase(it); }
	}
	(void)hipFree(*p);
	*p = nullptr;
}

extern "C" int skg_device_memory(long long *current, long long *peak, int reset_peak) {
	std::lock_guard<std::mutex> lk(mem_mu());
	if (current) *current = g_mem_cur;
	if (peak) *peak = g_mem_peak;
	if (reset_peak) g_mem_peak = g_mem_cur;
	return SKG_OK;
}

extern "C" int skg_host_alloc(void **p, long long bytes) {
	if (!p || bytes < 0) return SKG_E_PARAM;
	*p = nullptr;
	if (bytes == 0) return SKG_OK;
	const hipError_t e = hipHostMalloc(p, (size_t)bytes, hipHostMallocDefault);
	if (e != hipSuccess) { *p = nullptr; return e == hipErrorNoDevice ? SKG_E_NODEVICE : SKG_E_NOMEM; }
	return SKG_OK;
}

extern "C" int skg_host_free(void *p) {
	if (p) (void)hipHostFree(p);
	return SKG_OK;
}

int skg_ensure_tmp(skg_ctx *c, size_t bytes) {
	if (bytes <= c->tmpBytes && c->tmp) return 0;
	c->tmpBytes = 0;
	bytes = std::max<size_t>(bytes, 1 << 20);
	if (skg_dmalloc(&c->tmp, bytes)) return SKG_E_NOMEM;
	c->tmpBytes = bytes;
	return 0;
}

int skg_ensure_part(skg_ctx *c, size_t n) {
	if (n <= c->partCap && c->part) return 0;
	size_t cap = std::max<size_t>(n, 4096);
	if (dalloc(&c->part, cap)) return SKG_E_NOMEM;
	c->partCap = cap;
	return 0;
}

int skg_ensure_keys(skg_ctx *c, size_t n) {
	if (n <= c->kcap && c->k64a) return 0;
	size_t cap = std::max<size_t>(n, 1024);
	if (dalloc(&c->k64a, cap) || dalloc(&c->k64b, cap) || dalloc(&c->v32a, cap) || dalloc(&c->v32b, cap))
		return SKG_E_NOMEM;
	c->kcap = cap;
	c->vcap = cap;
	return 0;
}

int skg_sort_pairs(skg_ctx *c, unsigned long long *kin, unsigned long long *kout, int *vin, int *vout,
		int n, int endBit) {
	size_t bytes = 0;
	if (n <= 0) return 0;
	SKG_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)endBit, c->st));
	if (skg_ensure_tmp(c, bytes)) return SKG_E_NOMEM;
	bytes = c->tmpBytes;
	SKG_HIP(rocprim::radix_sort_pairs(c->tmp, bytes, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)endBit, c->st));
	return 0;
}

int skg_exclusive_scan(skg_ctx *c, const int *in, int *out, int n) {
	size_t bytes = 0;
	if (n <= 0) return 0;
	SKG_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, 0, (size_t)n, rocprim::plus<int>(), c->st));
	if (skg_ensure_tmp(c, bytes)) return SKG_E_NOMEM;
	bytes = c->tmpBytes;
	SKG_HIP(rocprim::exclusive_scan(c->tmp, bytes, in, out, 0, (size_t)n, rocprim::plus<int>(), c->st));
	return 0;
}

// The same two primitives on caller storage (tmp of tmpBytes bytes); with tmp
// null they only report the bytes they need in *need.  rocprim allocates
// nothing when handed its temporary storage.
int skg_sort_pairs_at(hipStream_t st, void *tmp, size_t tmpBytes, size_t *need, unsigned long long *kin,
		unsigned long long *kout, int *vin, int *vout, int n, int endBit) {
	size_t bytes = 0;
	SKG_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, (size_t)std::max(n, 1), 0u,
			(unsigned)endBit, st));
	if (need) *need = bytes;
	if (!tmp || n <= 0) return 0;
	if (bytes > tmpBytes) return SKG_E_PARAM;
	SKG_HIP(rocprim::radix_sort_pairs(tmp, tmpBytes, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)endBit, st));
	return 0;
}
int skg_exclusive_scan_at(hipStream_t st, void *tmp, size_t tmpBytes, size_t *need, const int *in, int *out,
		int n) {
	size_t bytes = 0;
	SKG_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, 0, (size_t)std::max(n, 1), rocprim::plus<int>(), st));
	if (need) *need = bytes;
	if (!tmp || n <= 0) return 0;
	if (bytes > tmpBytes) return SKG_E_PARAM;
	SKG_HIP(rocprim::exclusive_scan(tmp, tmpBytes, in, out, 0, (size_t)n, rocprim::plus<int>(), st));
	return 0;
}

void skg_timer_start(skg_ctx *c) { (void)hipEventRecord(c->e0, c->st); }
void skg_timer_stop(skg_ctx *c, int slot) {
	float ms = 0.f;
	(void)hipEventRecord(c->e1, c->st);
	(void)hipEventSynchronize(c->e1);
	(void)hipEventElapsedTime(&ms, c->e0, c->e1);
	c->tms[slot] += ms;
}

static inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

int skg_prof_begin(skg_ctx *c) {
	if (!c->prof) return -1;
	if (c->nrec == c->caprec) {
		int nc = c->caprec ? 2 * c->caprec : 256;
		skg_ctx::ProfRec *p = (skg_ctx::ProfRec *)realloc(c->prec, nc * sizeof(skg_ctx::ProfRec));
		if (!p) return -1;
		for (int i = c->caprec; i < nc; ++i) {
			if (hipEventCreate(&p[i].a) != hipSuccess || hipEventCreate(&p[i].b) != hipSuccess) { c->prec = p; return -1; }
		}
		c->prec = p; c->caprec = nc;
	}
	const int slot = c->nrec++;
	(void)hipEventRecord(c->prec[slot].a, c->st);
	return slot;
}
void skg_prof_end(skg_ctx *c, int slot, int which, double work) {
	if (slot < 0) return;
	(void)hipEventRecord(c->prec[slot].b, c->st);
	c->prec[slot].k = which;
	c->prec[slot].work = work;
}
void skg_prof_flush(skg_ctx *c) {
	if (!c->nrec) return;
	(void)hipStreamSynchronize(c->st);
	for (int i = 0; i < c->nrec; ++i) {
		float ms = 0.f;
		(void)hipEventElapsedTime(&ms, c->prec[i].a, c->prec[i].b);
		const int k = c->prec[i].k;
		c->kms[k] += ms; c->kwork[k] += c->prec[i].work; c->klaunch[k] += 1;
	}
	c->nrec = 0;
}

extern "C" int skg_set_profiling(skg_ctx *c, int on) {
	if (!c) return SKG_E_PARAM;
	skg_prof_flush(c);
	c->prof = on;
	for (int i = 0; i < SKG_K_N; ++i) { c->kms[i] = 0; c->kwork[i] = 0; c->klaunch[i] = 0; }
	return SKG_OK;
}

extern "C" int skg_kernel_stats(skg_ctx *c, int which, double *ms, long long *launches, double *work) {
	if (!c || which < 0 || which >= SKG_K_N) return SKG_E_PARAM;
	skg_prof_flush(c);
	if (ms) *ms = c->kms[which];
	if (launches) *launches = c->klaunch[which];
	if (work) *work = c->kwork[which];
	return SKG_OK;
}

// =====================================================================
// Morton ordering
// =====================================================================
__device__ __forceinline__ unsigned ord_f(float f) {
	unsigned u = __float_as_uint(f);
	return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord_f(unsigned u) {
	return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// Bounding box of the points: block minima/maxima into partials, then one
// block folds the partials (no contended atomics).
__global__ __launch_bounds__(256) void k_bbox(const float4 *__restrict__ pos, const int *__restrict__ idx, int n,
		float *part) {
	__shared__ float sm[6][4];
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	float4 p = make_float4(FLT_MAX, FLT_MAX, FLT_MAX, 0.f);
	float4 q = make_float4(-FLT_MAX, -FLT_MAX, -FLT_MAX, 0.f);
	if (i < n) { p = pos[idx ? idx[i] : i]; q = p; }
	float v[6] = {wave_min(p.x), wave_min(p.y), wave_min(p.z), wave_max(q.x), wave_max(q.y), wave_max(q.z)};
	const int w = threadIdx.x >> 6;
	if ((threadIdx.x & 63) == 0) for (int k = 0; k < 6; ++k) sm[k][w] = v[k];
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int k = 0; k < 6; ++k) {
			float r = sm[k][0];
			for (int j = 1; j < 4; ++j) r = k < 3 ? fminf(r, sm[k][j]) : fmaxf(r, sm[k][j]);
			part[6 * blockIdx.x + k] = r;
		}
	}
}
__global__ __launch_bounds__(256) void k_bbox_final(const float *__restrict__ part, int nb, unsigned *bb) {
	float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
	for (int b = threadIdx.x; b < nb; b += 256)
		for (int k = 0; k < 6; ++k) v[k] = k < 3 ? fminf(v[k], part[6 * b + k]) : fmaxf(v[k], part[6 * b + k]);
	__shared__ float sm[6][4];
	for (int k = 0; k < 6; ++k) v[k] = k < 3 ? wave_min(v[k]) : wave_max(v[k]);
	const int w = threadIdx.x >> 6;
	if ((threadIdx.x & 63) == 0) for (int k = 0; k < 6; ++k) sm[k][w] = v[k];
	__syncthreads();
	if (threadIdx.x == 0)
		for (int k = 0; k < 6; ++k) {
			float r = sm[k][0];
			for (int j = 1; j < 4; ++j) r = k < 3 ? fminf(r, sm[k][j]) : fmaxf(r, sm[k][j]);
			bb[k] = ord_f(r);
		}
}

__device__ __forceinline__ unsigned long long spread21(unsigned v) {
	unsigned long long x = v & 0x1fffffu;
	x = (x | x << 32) & 0x1f00000000ffffull;
	x = (x | x << 16) & 0x1f0000ff0000ffull;
	x = (x | x << 8) & 0x100f00f00f00f00full;
	x = (x | x << 4) & 0x10c30c30c30c30c3ull;
	x = (x | x << 2) & 0x1249249249249249ull;
	return x;
}

// Position on the 3-D Hilbert curve of 21-bit coordinates (Skilling's
// transpose method, AIP Conf. Proc. 707, 381 (2004)): undo the excess work of
// the reflected Gray code level by level, Gray-encode, and interleave.  Unlike
// Morton order the curve has no jumps, so runs of 64 consecutive points are
// more compact.
__device__ __forceinline__ unsigned long long hilbert21(unsigned x, unsigned y, unsigned z) {
	unsigned X[3] = {x, y, z};
	for (unsigned Q = 1u << 20; Q > 1; Q >>= 1) {
		const unsigned P = Q - 1;
		for (int i = 0; i < 3; ++i) {
			if (X[i] & Q) X[0] ^= P;
			else { const unsigned t = (X[0] ^ X[i]) & P; X[0] ^= t; X[i] ^= t; }
		}
	}
	X[1] ^= X[0];
	X[2] ^= X[1];
	unsigned t = 0;
	for (unsigned Q = 1u << 20; Q > 1; Q >>= 1)
		if (X[2] & Q) t ^= Q - 1;
	X[0] ^= t; X[1] ^= t; X[2] ^= t;
	return (spread21(X[0]) << 2) | (spread21(X[1]) << 1) | spread21(X[2]);
}

template <bool HILBERT>
__global__ void k_morton(const float4 *__restrict__ pos, const int *__restrict__ idx, int n,
		const unsigned *__restrict__ bb, unsigned long long *keys, int *vals, const int *__restrict__ box,
		int bbits) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int j = idx ? idx[i] : i;
	float4 p = pos[j];
	float lo[3] = {unord_f(bb[0]), unord_f(bb[1]), unord_f(bb[2])};
	float hi[3] = {unord_f(bb[3]), unord_f(bb[4]), unord_f(bb[5])};
	float v[3] = {p.x, p.y, p.z};
	unsigned q[3];
	for (int a = 0; a < 3; ++a) {
		float ext = hi[a] - lo[a];
		float t = ext > 0.f ? (v[a] - lo[a]) / ext : 0.f;
		t = fminf(fmaxf(t, 0.f), 1.f);
		q[a] = (unsigned)fminf(t * 2097152.0f, 2097151.0f);
	}
	unsigned long long key = HILBERT ? hilbert21(q[0], q[1], q[2]) :
			(spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
	// boxes first: the box id takes the top bits, the coarse Morton digits the rest
	if (box && bbits > 0) key = ((unsigned long long)box[j] << (63 - bbits)) | (key >> bbits);
	keys[i] = key;
	vals[i] = j;
}

int skg_box_bits(const skg_ctx *c) {
	int b = 0;
	while ((1 << b) < c->nBox) ++b;
	return b;
}

// Curve keys of n points (bounding box, then Hilbert or Morton keys) into
// keys/vals, on caller storage: part holds 6 floats per 256 points, bb 6
// words.  vals[i] = the point's index into pos (or its idx[] entry).
int skg_curve_keys(skg_ctx *c, const float4 *pos, const int *idx, int n, const int *box, bool hilbert,
		unsigned long long *keys, int *vals, float *part, unsigned *bb) {
	if (n <= 0) return 0;
	const int nb = nblk(n, 256);
	k_bbox<<<nb, 256, 0, c->st>>>(pos, idx, n, part);
	k_bbox_final<<<1, 256, 0, c->st>>>(part, nb, bb);
	if (hilbert)
		k_morton<true><<<nblk(n, 256), 256, 0, c->st>>>(pos, idx, n, bb, keys, vals, box, box ? skg_box_bits(c) : 0);
	else
		k_morton<false><<<nblk(n, 256), 256, 0, c->st>>>(pos, idx, n, bb, keys, vals, box, box ? skg_box_bits(c) : 0);
	SKG_HIP(hipGetLastError());
	return 0;
}

// perm_out[i] = index (into pos, or idx[] entry) of the i-th point in Morton
// order; with box (indexed like pos) the order is by box first.
static int curve_sort(skg_ctx *c, const float4 *pos, const int *idx, int n, int *perm_out, const int *box,
		bool hilbert) {
	if (n <= 0) return 0;
	if (skg_ensure_keys(c, n)) return SKG_E_NOMEM;
	if (skg_ensure_part(c, 6 * (size_t)nblk(n, 256))) return SKG_E_NOMEM;
	int rc = skg_curve_keys(c, pos, idx, n, box, hilbert, c->k64a, c->v32a, c->part, (unsigned *)(c->d_cnt + 8));
	if (rc) return rc;
	return skg_sort_pairs(c, c->k64a, c->k64b, c->v32a, perm_out, n, 63);
}
int skg_morton_sort(skg_ctx *c, const float4 *pos, const int *idx, int n, int *perm_out, const int *box) {
	return curve_sort(c, pos, idx, n, perm_out, box, false);
}
int skg_hilbert_sort(skg_ctx *c, const float4 *pos, const int *idx, int n, int *perm_out, const int *box) {
	return curve_sort(c, pos, idx, n, perm_out, box, true);
}
// Movers: tiles of 64 consecutive movers scan the union of their leaves, so
// the more compact Hilbert runs (SKG_MOVER_ORDER=morton: Morton).
static bool mover_morton() {
	static const bool morton = [] { const char *e = getenv("SKG_MOVER_ORDER"); return e && !strcmp(e, "morton"); }();
	return morton;
}
static int mover_sort(skg_ctx *c, int n) {
	c->lastSort = n;
	return curve_sort(c, c->m_pos, nullptr, n, c->m_list, c->nBox > 1 ? c->m_box : nullptr, !mover_morton());
}
// The active list (host count c->nActive) is re-sorted by current position
// once it has fallen below this fraction of its size at the last sort
// (SKG_MOVER_RESORT; 0: never).
static double mover_resort_frac() {
	static const double f = [] { const char *e = getenv("SKG_MOVER_RESORT"); return e ? atof(e) : 0.9; }();
	return f;
}
static int mover_resort(skg_ctx *c) {
	const int n = c->nActive;
	c->lastSort = n;
	int rc = curve_sort(c, c->m_pos, c->m_list, n, c->m_list2, c->nBox > 1 ? c->m_box : nullptr, !mover_morton());
	if (rc) return rc;
	SKG_HIP(hipMemcpyAsync(c->m_list, c->m_list2, 4 * (size_t)n, hipMemcpyDeviceToDevice, c->st));
	return 0;
}

// =====================================================================
// Trees: complete binary tree over leaves of B consecutive sorted items.
// =====================================================================
// Leaf bounds skip filler items (NaN position); hi.w receives the leaf's box
// range (ibox indexed like pos; none: box 0).
__global__ void k_tree_leaves(const float4 *__restrict__ pos, int n, int B, int nLeaf, int nLeaf2,
		const float *__restrict__ b2, const int *__restrict__ ibox, const int *__restrict__ lstart,
		float4 *lo, float4 *hi) {
	int l = blockIdx.x * blockDim.x + threadIdx.x;
	if (l >= nLeaf2) return;
	float4 a = make_float4(INFINITY, INFINITY, INFINITY, -INFINITY);
	float4 b = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
	int b0 = 0xFFFF, b1 = 0;
	if (l < nLeaf) {
		const int2 r = leaf_range(lstart, l, B, n);
		for (int i = r.x; i < r.y; ++i) {
			const int bx = ibox ? ibox[i] : 0;
			b0 = min(b0, bx); b1 = max(b1, bx);
			float4 p = pos[i];
			if (p.x != p.x) continue;
			a.x = fminf(a.x, p.x); a.y = fminf(a.y, p.y); a.z = fminf(a.z, p.z);
			b.x = fmaxf(b.x, p.x); b.y = fmaxf(b.y, p.y); b.z = fmaxf(b.z, p.z);
			a.w = fmaxf(a.w, b2 ? b2[i] : 0.f);
		}
	}
	b.w = brange_pack(b0, b1);
	lo[nLeaf2 + l] = a;
	hi[nLeaf2 + l] = b;
}

__global__ void k_tree_level(float4 *lo, float4 *hi, int first, int count) {
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= count) return;
	int i = first + t;
	float4 a0 = lo[2 * i], a1 = lo[2 * i + 1], b0 = hi[2 * i], b1 = hi[2 * i + 1];
	const unsigned r0 = __float_as_uint(b0.w), r1 = __float_as_uint(b1.w);
	const int x0 = min((int)(r0 & 0xFFFFu), (int)(r1 & 0xFFFFu)), x1 = max((int)(r0 >> 16), (int)(r1 >> 16));
	lo[i] = make_float4(fminf(a0.x, a1.x), fminf(a0.y, a1.y), fminf(a0.z, a1.z), fmaxf(a0.w, a1.w));
	hi[i] = make_float4(fmaxf(b0.x, b1.x), fmaxf(b0.y, b1.y), fmaxf(b0.z, b1.z), brange_pack(x0, x1));
}

// Refit of the entity tree's largest live ball alone: entities do not move and
// ScatterCut only removes balls, so the position bounds stay as built and only
// lo.w (the largest live fBall2 below; dead and filler entities carry -1)
// changes.
__global__ void k_leaf_ball(const float4 *__restrict__ eaux, int n, int B, int nLeaf, int nLeaf2,
		const int *__restrict__ lstart, float4 *lo) {
	const int l = blockIdx.x * blockDim.x + threadIdx.x;
	if (l >= nLeaf2) return;
	float w = -INFINITY;
	if (l < nLeaf) {
		const int2 r = leaf_range(lstart, l, B, n);
		for (int i = r.x; i < r.y; ++i) w = fmaxf(w, eaux[i].x);
	}
	lo[nLeaf2 + l].w = w;
}
// The same for the leaves flagged by the sorted cut only (the others' entities
// are unchanged since the last refit); clears the flags.
__global__ void k_leaf_ball_dirty(const float4 *__restrict__ eaux, int n, int B, int nLeaf, int nLeaf2,
		const int *__restrict__ lstart, float4 *lo, int *dirty) {
	const int l = blockIdx.x * blockDim.x + threadIdx.x;
	if (l >= nLeaf || !dirty[l]) return;
	dirty[l] = 0;
	float w = -INFINITY;
	const int2 r = leaf_range(lstart, l, B, n);
	for (int i = r.x; i < r.y; ++i) w = fmaxf(w, eaux[i].x);
	lo[nLeaf2 + l].w = w;
}
// entity -> leaf, and no leaf flagged
__global__ void k_entity_leaf(int n, int B, int nLeaf, int nLeaf2, const int *__restrict__ lstart, int *eleaf,
		int *dirty) {
	const int l = blockIdx.x * blockDim.x + threadIdx.x;
	if (l >= nLeaf2) return;
	dirty[l] = 0;
	if (l >= nLeaf) return;
	const int2 r = leaf_range(lstart, l, B, n);
	for (int i = r.x; i < r.y; ++i) eleaf[i] = l;
}
__global__ void k_level_ball(float4 *lo, int first, int count) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= count) return;
	const int i = first + t;
	lo[i].w = fmaxf(lo[2 * i].w, lo[2 * i + 1].w);
}

// The levels of at most mtop nodes in one workgroup, level by level (one
// launch instead of one per level).  A level reads what the one below wrote
// in the same launch: stores complete to L2 (agent fence) before the barrier,
// and the reads bypass the CU's L1.
__global__ __launch_bounds__(1024) void k_level_ball_top(float4 *lo, int mtop) {
	for (int m = mtop; m >= 1; m >>= 1) {
		for (int t = threadIdx.x; t < m; t += blockDim.x) {
			const int i = m + t;
			const float a = __hip_atomic_load(&lo[2 * i].w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			const float b = __hip_atomic_load(&lo[2 * i + 1].w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			lo[i].w = fmaxf(a, b);
		}
		__threadfence();
		__syncthreads();
	}
}

void skg_free_tree(DevTree *t) { dfree(&t->lo); dfree(&t->hi); dfree(&t->lstart); t->n = t->nLeaf = t->nLeaf2 = 0; }

static int tree_fill(skg_ctx *c, DevTree *t, const float4 *pos, const float *b2, const int *ibox) {
	k_tree_leaves<<<nblk(t->nLeaf2, 256), 256, 0, c->st>>>(pos, t->n, t->B, t->nLeaf, t->nLeaf2, b2, ibox,
			t->lstart, t->lo, t->hi);
	for (int m = t->nLeaf2 >> 1; m >= 1; m >>= 1)
		k_tree_level<<<nblk(m, 256), 256, 0, c->st>>>(t->lo, t->hi, m, m);
	SKG_HIP(hipGetLastError());
	return 0;
}

int skg_build_tree(skg_ctx *c, DevTree *t, const float4 *pos, int n, int B, const float *b2, const int *ibox) {
	int nLeaf = std::max(1, (n + B - 1) / B), nLeaf2 = 1;
	while (nLeaf2 < nLeaf) nLeaf2 <<= 1;
	if (nLeaf2 != t->nLeaf2 || !t->lo) {
		if (dalloc(&t->lo, 2 * (size_t)nLeaf2) || dalloc(&t->hi, 2 * (size_t)nLeaf2)) return SKG_E_NOMEM;
	}
	dfree(&t->lstart);
	t->n = n; t->B = B; t->nLeaf = nLeaf; t->nLeaf2 = nLeaf2;
	return tree_fill(c, t, pos, b2, ibox);
}

// Leaves and their power-of-two count for n items in leaves of B.
int skg_tree_leaves2(int n, int B, int *nLeaf) {
	const int nl = std::max(1, (n + B - 1) / B);
	int nl2 = 1;
	while (nl2 < nl) nl2 <<= 1;
	if (nLeaf) *nLeaf = nl;
	return nl2;
}

// The same tree on caller storage: lo and hi hold 2 * skg_tree_leaves2(n, B)
// records each and stay the caller's (the tree must not be freed).
int skg_build_tree_at(skg_ctx *c, DevTree *t, const float4 *pos, int n, int B, float4 *lo, float4 *hi) {
	int nLeaf = 0;
	const int nLeaf2 = skg_tree_leaves2(n, B, &nLeaf);
	t->lo = lo; t->hi = hi; t->lstart = nullptr;
	t->n = n; t->B = B; t->nLeaf = nLeaf; t->nLeaf2 = nLeaf2;
	return tree_fill(c, t, pos, nullptr, nullptr);
}

// Same structure, new bounds and radii (positions or live radii changed).
int skg_refit_tree(skg_ctx *c, DevTree *t, const float4 *pos, const float *b2, const int *ibox) {
	return tree_fill(c, t, pos, b2, ibox);
}

// ---------------------------------------------------------------------------
// Prefix leaves: each leaf is a maximal cell of the key hierarchy (a Morton
// prefix; the box id leads the key) holding at most B items, so leaves are
// compact octree-like cells instead of runs of B consecutive keys (which
// straddle cell boundaries and give long thin bounding boxes).  Equal keys are
// split further by index bits.  For item i the leaf depth is the smallest d
// with #{j : lcp(i, j) >= d} <= B; a cell of at most B items lies within
// [i-B, i+B], so counting over that window is exact.
__device__ __forceinline__ int key_lcp(const unsigned long long *__restrict__ k, int i, int j) {
	const unsigned long long a = k[i], b = k[j];
	return a != b ? __clzll((long long)(a ^ b)) : 64 + __clz(i ^ j);
}
// lcp of each item with the next (the last item's entry is not used).
__global__ void k_adj_lcp(const unsigned long long *__restrict__ keys, int n, int *adj) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) adj[i] = i + 1 < n ? key_lcp(keys, i, i + 1) : 0;
}
// The depth d of item i's leaf: the smallest d with at most B items (itself
// included) of lcp >= d with it in [i-B, i+B].  Over sorted keys lcp(i, j) is
// the minimum of the adjacent lcps between i and j, so it falls as j moves
// away from i on either side; the B-th largest v of the two runs, found by
// merging them, gives d = v + 1 (0 when the window holds fewer than B others).
__global__ void k_leaf_depth(const int *__restrict__ adj, int n, int B, int minDepth, int *depth) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int ra = min(B, n - 1 - i), la = min(B, i);
	if (ra + la < B) { depth[i] = max(0, minDepth); return; }
	int a = 0, b = 0, v = 0;
	int rv = ra > 0 ? adj[i] : -1, lv = la > 0 ? adj[i - 1] : -1;
	for (int t = 0; t < B; ++t) {
		if (a < ra && (b >= la || rv >= lv)) {
			v = rv;
			++a;
			rv = a < ra ? min(rv, adj[i + a]) : -1;
		} else {
			v = lv;
			++b;
			lv = b < la ? min(lv, adj[i - 1 - b]) : -1;
		}
	}
	depth[i] = max(v + 1, minDepth);
}
__global__ void k_leaf_flags(const unsigned long long *__restrict__ keys, const int *__restrict__ depth, int n,
		int *flag) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	flag[i] = (i == 0 || depth[i] != depth[i - 1] || key_lcp(keys, i, i - 1) < depth[i]) ? 1 : 0;
}
__global__ void k_leaf_starts(const int *__restrict__ flag, const int *__restrict__ off, int n, int nLeaf,
		int nLeaf2, int *lstart) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n && flag[i]) lstart[off[i]] = i;
	if (i >= nLeaf && i <= nLeaf2) lstart[i] = n;     // end marker and empty leaves
}

int skg_build_tree_prefix(skg_ctx *c, DevTree *t, const float4 *pos, const unsigned long long *keys, int n,
		int B, int minDepth, const float *b2, const int *ibox) {
	int rc;
	if (n <= 0) return skg_build_tree(c, t, pos, n, B, b2, ibox);
	int *depth = nullptr, *flag = nullptr, *off = nullptr;
	if (sget(c, 17, &depth, n) || sget(c, 18, &flag, n) || sget(c, 19, &off, n)) return SKG_E_NOMEM;
	k_adj_lcp<<<nblk(n, 256), 256, 0, c->st>>>(keys, n, off);          // off: free until the scan
	k_leaf_depth<<<nblk(n, 256), 256, 0, c->st>>>(off, n, B, minDepth, depth);
	k_leaf_flags<<<nblk(n, 256), 256, 0, c->st>>>(keys, depth, n, flag);
	if ((rc = skg_exclusive_scan(c, flag, off, n))) return rc;
	int last[2];
	SKG_HIP(hipMemcpyAsync(&last[0], off + n - 1, 4, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipMemcpyAsync(&last[1], flag + n - 1, 4, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	const int nLeaf = last[0] + last[1];
	int nLeaf2 = 1;
	while (nLeaf2 < nLeaf) nLeaf2 <<= 1;
	if (nLeaf2 != t->nLeaf2 || !t->lo) {
		if (dalloc(&t->lo, 2 * (size_t)nLeaf2) || dalloc(&t->hi, 2 * (size_t)nLeaf2)) return SKG_E_NOMEM;
	}
	if (dalloc(&t->lstart, (size_t)nLeaf2 + 1)) return SKG_E_NOMEM;
	k_leaf_starts<<<nblk(std::max(n, nLeaf2 + 1), 256), 256, 0, c->st>>>(flag, off, n, nLeaf, nLeaf2, t->lstart);
	t->n = n; t->B = B; t->nLeaf = nLeaf; t->nLeaf2 = nLeaf2;
	return tree_fill(c, t, pos, b2, ibox);
}

int skg_copy_tree_leaves(skg_ctx *c, DevTree *dst, const DevTree *src) {
	if (!src->lstart) { dfree(&dst->lstart); return 0; }
	if (dalloc(&dst->lstart, (size_t)src->nLeaf2 + 1)) return SKG_E_NOMEM;
	SKG_HIP(hipMemcpyAsync(dst->lstart, src->lstart, 4 * ((size_t)src->nLeaf2 + 1), hipMemcpyDeviceToDevice, c->st));
	return 0;
}


// =====================================================================
// kNN: one wavefront per block of 64 Morton-consecutive queries, one lane
// per query.  Each lane keeps a bounded max-heap of packed (d2, iOrder) keys
// in LDS (slot-major so a slot of all 64 lanes is one conflict-free row).
// The wave walks the tree once for the block (nearest child first), prunes
// with the block's largest current radius, and streams each visited bucket
// through the lanes as a uniform (scalar-load) candidate tile.
// =====================================================================
// Slot k holds 0, below every key, so a node's right child is always
// readable: both children come in one LDS read (rows 512 bytes apart).  The
// sift-down is unrolled to the heap's depth (no loop bookkeeping; a lane that
// has placed its key idles through the remaining levels), and the new top is
// known at the first level -- the larger of the key and the root's larger
// child -- so no LDS read follows the sift.
template <int KMAX>
__device__ __forceinline__ unsigned long long heap_replace(unsigned long long *heap, int lane, int k,
		unsigned long long key) {
	constexpr int LEVELS = KMAX <= 32 ? 5 : KMAX <= 64 ? 6 : KMAX <= 128 ? 7 : 8;
	unsigned long long *h = heap + lane;
	int i = 0;
	bool done = false;
	unsigned long long top = key;
#pragma unroll
	for (int d = 0; d < LEVELS; ++d) {
		const int l = 2 * i + 1;
		if (!done && l < k) {
			const unsigned long long *row = h + l * 64;
			unsigned long long cv = row[0];
			const unsigned long long cr = row[64];
			int ci = l;
			if (cr > cv) { cv = cr; ci = l + 1; }
			if (cv <= key) {
				done = true;
			} else {
				h[i * 64] = cv;
				if (d == 0) top = cv;
				i = ci;
			}
		} else {
			done = true;
		}
	}
	h[i * 64] = key;
	return top;
}

template <bool PER>
__device__ __forceinline__ float point_box_d2(float4 lo, float4 hi, float qx, float qy, float qz,
		float L) {
	float gx = gap_axis<PER>(lo.x, hi.x, qx, qx, L);
	float gy = gap_axis<PER>(lo.y, hi.y, qy, qy, L);
	float gz = gap_axis<PER>(lo.z, hi.z, qz, qz, L);
	return sq3(gx, gy, gz);
}

__device__ __forceinline__ float key_d2(unsigned long long k) { return __uint_as_float((unsigned)(k >> 32)); }

// BATCH: several boxes share the arrays; a lane accepts only candidates whose
// iOrder lies in its own box's range, and the walk skips subtrees holding none
// of the block's boxes.
#define SKG_KNN_SEED 64
template <int KMAX, bool PER, bool COMP, bool BATCH, int W>
__global__ __launch_bounds__(64 * W) void k_knn(
		const float4 *__restrict__ spos, const int *__restrict__ sord, int n,
		const float4 *__restrict__ nlo, const float4 *__restrict__ nhi, int nLeaf2, int B,
		float L, int k, const int *__restrict__ sinv,
		float *__restrict__ ball2_out, float *__restrict__ rhoG, unsigned long long *__restrict__ heap_out,
		const int *__restrict__ sbox, const DevBox *__restrict__ bt, float *__restrict__ dlnk_out) {
	// W wavefronts per workgroup, each with its own 64 queries, heaps and stack
	__shared__ unsigned long long heapW[W][(KMAX + 1) * 64];   // k slots + the zero sentinel row
	__shared__ int stkW[W][64];
	__shared__ float stkdW[W][64];
	const int wv = W > 1 ? (int)(threadIdx.x >> 6) : 0;
	unsigned long long *heap = heapW[wv];
	int *stk = stkW[wv];
	float *stkd = stkdW[wv];
	const int lane = threadIdx.x & 63;
	const int base = (xcd_tile(blockIdx.x, gridDim.x) * W + wv) * 64;
	if (base >= n) return;                        // a wavefront past the end (no workgroup barrier here)
	const int sd0 = max(0, base - SKG_KNN_SEED), sd1 = min(n, base + 64 + SKG_KNN_SEED);   // seeded range
	const int qi = base + lane;
	const bool valid = qi < n;
	const float4 q = spos[valid ? qi : base];
	int qb = 0;
	unsigned boff = 0u, bn = 0xFFFFFFFFu;
	if (BATCH) { qb = sbox[valid ? qi : base]; boff = (unsigned)bt[qb].off; bn = (unsigned)bt[qb].n; }
	const int wb0 = BATCH ? wave_min_i(qb) : 0, wb1 = BATCH ? wave_max_i(qb) : 0;
	const float hL = 0.5f * L;
	const float qxm = fsub(q.x, L), qxp = fadd(q.x, L);
	const float qym = fsub(q.y, L), qyp = fadd(q.y, L);
	const float qzm = fsub(q.z, L), qzp = fadd(q.z, L);
	const float3 blo = make_float3(wave_min(q.x), wave_min(q.y), wave_min(q.z));
	const float3 bhi = make_float3(wave_max(q.x), wave_max(q.y), wave_max(q.z));
	for (int s = 0; s < k; ++s) heap[s * 64 + lane] = KEY_SENT;
	heap[k * 64 + lane] = 0ull;
	unsigned long long top = KEY_SENT;

	// One candidate (position cc, iOrder oj) against the lane's query.  CB:
	// test that the candidate is of the lane's box; CI: look for a delta over
	// half a period (a leaf that is all the lane's box, or that no delta of
	// the wavefront's can take that far, skips the test).
	auto try_cand = [&](const float4 cc, unsigned oj, auto CB, auto CI) {
		if (!(BATCH && decltype(CB)::value) || oj - boff < bn) {
			float dx = fsub(q.x, cc.x), dy = fsub(q.y, cc.y), dz = fsub(q.z, cc.z);
			// the query-side image only where a delta exceeds half a period (the
			// same values img_delta gives; most wavefronts never take it)
			if (PER && decltype(CI)::value && fmaxf(fmaxf(fabsf(dx), fabsf(dy)), fabsf(dz)) > hL) {
				dx = img_delta<PER>(q.x, qxm, qxp, cc.x, hL);
				dy = img_delta<PER>(q.y, qym, qyp, cc.y, hL);
				dz = img_delta<PER>(q.z, qzm, qzp, cc.z, hL);
			}
			const unsigned long long key = ((unsigned long long)__float_as_uint(sq3(dx, dy, dz)) << 32) | oj;
			if (key < top) top = heap_replace<KMAX>(heap, lane, k, key);
		}
	};
	// The cnt candidates from start (a uniform range): records are read four
	// at a time through the scalar cache, one wait per four (the arrays carry
	// SKG_EPAD entries past the end, so the reads past a range's end stay in
	// the allocation; their values are never used).  skip: the seeded range.
	auto scan = [&](int start, int cnt, bool skip, auto CB, auto CI) {
		for (int t0 = 0; t0 < cnt; t0 += 4) {
			float4 C4[4];
			unsigned O4[4];
#pragma unroll
			for (int u = 0; u < 4; ++u) { C4[u] = spos[start + t0 + u]; O4[u] = (unsigned)sord[start + t0 + u]; }
#pragma unroll
			for (int u = 0; u < 4; ++u) {
				if (t0 + u >= cnt) break;
				const int j = start + t0 + u;
				if (skip && j >= sd0 && j < sd1) continue;
				try_cand(C4[u], O4[u], CB, CI);
			}
		}
	};
	using yes = std::true_type;
	using no = std::false_type;

	// Seed with the block's own 64 points and the SKG_KNN_SEED on each side
	// along the curve (the nearest candidates: the radii start small, so the
	// walk prunes from its first node).
	if (valid) scan(sd0, sd1 - sd0, false, yes(), yes());
	float R2 = wave_max(valid ? key_d2(top) : 0.f);

	int sp = 1;
	stk[0] = 1; stkd[0] = 0.f;
	while (sp > 0) {
		--sp;
		const int node = uniform_int(stk[sp]);
		if (stkd[sp] > R2) continue;
		const float4 lo = nlo[node], hi = nhi[node];
		if (node < nLeaf2) {
			const int c0 = 2 * node;
			const float4 h0 = nhi[c0], h1 = nhi[c0 + 1];
			const float d0 = box_box_d2<PER>(nlo[c0], h0, blo, bhi, L);
			const float d1 = box_box_d2<PER>(nlo[c0 + 1], h1, blo, bhi, L);
			const bool m0 = !BATCH || brange_meets(h0.w, wb0, wb1), m1 = !BATCH || brange_meets(h1.w, wb0, wb1);
			const bool near0 = d0 <= d1;
			const int nf = near0 ? c0 : c0 + 1;
			const float dn = near0 ? d0 : d1, df = near0 ? d1 : d0;
			const bool mn = near0 ? m0 : m1, mf = near0 ? m1 : m0;
			if (mf && df <= R2) { stk[sp] = nf ^ 1; stkd[sp] = df; ++sp; }
			if (mn && dn <= R2) { stk[sp] = nf; stkd[sp] = dn; ++sp; }
			continue;
		}
		const int start = (node - nLeaf2) * B;
		if (start >= n) continue;
		const int cnt = min(B, n - start);
		if (start >= sd0 && start + cnt <= sd1) continue;            // seeded range
		const bool need = valid && (!BATCH || brange_has(hi.w, qb)) &&
				point_box_d2<PER>(lo, hi, q.x, q.y, q.z, L) <= key_d2(top);
		if (!__any(need)) continue;
		// the leaf is all of every needing lane's box: no box test; every
		// delta of the wavefront's queries to the leaf's box under 0.999 of
		// half a period: no image test
		const bool cb = BATCH && __any(need && !brange_is(hi.w, qb));
		const bool ci = PER && uniform_int(!(fmaxf(bhi.x - lo.x, hi.x - blo.x) < 0.999f * hL &&
				fmaxf(bhi.y - lo.y, hi.y - blo.y) < 0.999f * hL &&
				fmaxf(bhi.z - lo.z, hi.z - blo.z) < 0.999f * hL)) != 0;
		if (need) {
			if (cb) { if (ci) scan(start, cnt, true, yes(), yes()); else scan(start, cnt, true, yes(), no()); }
			else { if (ci) scan(start, cnt, true, no(), yes()); else scan(start, cnt, true, no(), no()); }
		}
		R2 = wave_max(valid ? key_d2(top) : 0.f);
	}
	if (!valid) return;
	// fBall2 = the k-th smallest d2; gather half of the symmetric density over
	// all queue members but the head (smooth1.c:243-263).
	const float h2 = key_d2(top);
	const float ih2 = (float)(4.0 / (double)h2);
	const float fNorm = (float)(0.5 * M_1_PI * sqrt((double)ih2) * (double)ih2);
	// with dlnk_out, the gather half of d rho / d ln nSmooth over the same
	// pairs, from exact periodic separations, h = sqrt(fBall2)/2
	const bool dl = dlnk_out != nullptr;
	const float ih = dl ? 1.0f / (0.5f * sqrtf(h2)) : 0.f;
	Sum<COMP> acc;
	acc.init();
	Sum<true> dacc;
	dacc.init();
	for (int s = 1; s < k; ++s) {
		const unsigned long long key = heap[s * 64 + lane];
		// the neighbour's record in scatterer order (near the query's own)
		const float4 pj = spos[sinv[(unsigned)key]];
		float d2 = key_d2(key), d2x = d2;
		if (PER && (COMP || dl)) {
			// weights from exact periodic separations (membership stays the
			// reference's d2 test, so the neighbour set is unchanged)
			const float px = pj.x, py = pj.y, pz = pj.z;
			float dx = fsub(q.x, px), dy = fsub(q.y, py), dz = fsub(q.z, pz);
			if (fabsf(dx) > hL || fabsf(dy) > hL || fabsf(dz) > hL) {
				dx = exact_delta(q.x, px, dx > hL ? -L : (dx < -hL ? L : 0.f));
				dy = exact_delta(q.y, py, dy > hL ? -L : (dy < -hL ? L : 0.f));
				dz = exact_delta(q.z, pz, dz > hL ? -L : (dz < -hL ? L : 0.f));
				d2x = sq3(dx, dy, dz);
			}
			if (COMP) d2 = d2x;
		}
		const float mj = pj.w;
		acc.add(fmul(m4_w<COMP>(d2, ih2, fNorm), mj));
		if (dl) dacc.add(dlnk_term(sqrtf(d2x), ih, mj));
	}
	ball2_out[qi] = h2;
	rhoG[qi] = acc.val();
	if (dl) dlnk_out[qi] = dacc.val() * (float)(-M_1_PI / 6.0) * ih * ih * ih;
	if (heap_out)                                  // slot-major: one slot of 64 queries is one 512-byte store
		for (int s = 0; s < k; ++s) heap_out[(size_t)s * n + qi] = heap[s * 64 + lane];
}

// =====================================================================
// Ball cover: for each point, every entity whose ball strictly holds it
// (non-periodic; periodic images are explicit replica entities, as the
// reference's smx->pp).  MODE 0 accumulates the scatter half of the density,
// MODE 1 the density gradient and then moves the point (kdMoveParticles).
// =====================================================================
// One entity e (record A, X, image code K) against the lane's point when want.
// MODE 0 accumulates the scatter density (with dl also the scatter half of
// d rho / d ln nSmooth into s1, exact separations, the factor -1/(6 pi)
// applied by the caller); MODE 1 the gradient, and for ScatterCut either
// stamps the hit entity (ehit) or keeps the lane's smallest positive density
// among its hits (hmin; during the move loop the record's w holds the
// entity's density, +inf for 0).
template <int MODE, bool COMP, bool LAZYK = false>
__device__ __forceinline__ void cover_hit(int e, const float4 &A, const float4 &X, int K, bool want,
		const float4 &p, const float4 *__restrict__ eorg, int *__restrict__ ehit, int stamp, float &hmin,
		float Lx, float Ly, float Lz, Sum<COMP> &s0, Sum<COMP> &s1, Sum<COMP> &s2, bool dl = false) {
	float dx = fsub(A.x, p.x), dy = fsub(A.y, p.y), dz = fsub(A.z, p.z);
	float d2 = sq3(dx, dy, dz);
	if (want && d2 < X.x) {
		if (COMP || (MODE == 0 && dl)) {
			// replica: separation from the source position plus the exact shift;
			// LAZYK: the caller passes K < 0 and the code is read here, from the
			// source record's w
			float4 O = make_float4(0.f, 0.f, 0.f, 0.f);
			if (LAZYK && K < 0) { O = eorg[e]; K = (int)O.w; }
			if (K != 13) {                       // 9(ix+1) + 3(iy+1) + (iz+1)
				O = eorg[e];
				dx = exact_delta(O.x, p.x, (float)(K / 9 - 1) * Lx);
				dy = exact_delta(O.y, p.y, (float)((K / 3) % 3 - 1) * Ly);
				dz = exact_delta(O.z, p.z, (float)(K % 3 - 1) * Lz);
			}
		}
		if (MODE == 0) {
			s0.add(fmul(m4_w<COMP>(COMP ? sq3(dx, dy, dz) : d2, X.y, X.w), A.w));
			if (dl) {
				const float ih = __fsqrt_rn(X.y);      // X.y = 4 / fBall2 = 1 / h^2
				s1.add(dlnk_term(sqrtf(sq3(dx, dy, dz)), ih, A.w) * (ih * ih * ih));
			}
		} else {
			const float f = COMP ? grad_factor_f32(sq3(dx, dy, dz), X.y, X.z) : grad_factor(d2, X.y, X.z);
			s0.add(fmul(dx, f)); s1.add(fmul(dy, f)); s2.add(fmul(dz, f));
			if (ehit) ehit[e] = stamp;
			else hmin = __builtin_amdgcn_fmed3f(X.w, -INFINITY, hmin);   // min; w is +inf for a density of 0
		}
	}
}

// The same scan called by the whole wavefront (want per lane): the leaf's
// records are staged 64 at a time in the wavefront's LDS slots sA/sX/sK (lane
// j reads record j, one round trip), then every lane with want reads each
// staged record at one address.  The wavefront reads back only what it wrote,
// and its LDS operations complete in order.
template <int MODE, bool COMP>
__device__ __forceinline__ void cover_scan_staged(int start, int cnt, bool want, const float4 &p,
		const float4 *__restrict__ epos, const float4 *__restrict__ eaux, const float4 *__restrict__ eorg,
		const int *__restrict__ ecode, int *__restrict__ ehit, int stamp, float &hmin, float Lx, float Ly,
		float Lz, Sum<COMP> &s0, Sum<COMP> &s1, Sum<COMP> &s2, bool dl, float4 *sA, float4 *sX, int *sK,
		int lane) {
	for (int c0 = 0; c0 < cnt; c0 += 64) {
		const int m = min(64, cnt - c0);
		if (lane < m) {
			sA[lane] = epos[start + c0 + lane];
			sX[lane] = eaux[start + c0 + lane];
			sK[lane] = (COMP || (MODE == 0 && dl)) ? ecode[start + c0 + lane] : 13;
		}
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		if (want)
			for (int t = 0; t < m; ++t)
				cover_hit<MODE, COMP>(start + c0 + t, sA[t], sX[t], sK[t], true, p, eorg, ehit, stamp, hmin, Lx, Ly,
						Lz, s0, s1, s2, dl);
		// every lane has read these records before the next chunk overwrites them
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
	}
}

// ScatterCut threshold from the move step (smooth1.c:460-461): each box's
// smallest positive density over the entities its movers hit this step, as
// the minimum of the lanes' hmin, one atomic per box present in the
// wavefront (densities are >= 0, so their bits order as integers; the slots
// start at +inf).  A box has SKG_THR_SLOTS slots, picked by the wavefront's
// tile, so that the thousands of wavefronts of a large box do not all meet at
// one address (device-scope atomics from every XCD serialise there);
// k_thr_fold takes the minimum over the slots.  All 64 lanes must be active.
#define SKG_THR_SLOTS 64
__device__ __forceinline__ void box_min_atomic(float hmin, bool valid, int pb, int *bthr) {
	const int lane = threadIdx.x & 63;
	const bool has = valid && hmin < INFINITY;
	unsigned long long todo = __ballot(has);
	while (todo) {
		const int l = __ffsll((long long)todo) - 1;
		const int b = __shfl(pb, l, 64);
		const bool mine = has && pb == b;
		const float m = wave_min(mine ? hmin : INFINITY);
		const int slot = (int)((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (SKG_THR_SLOTS - 1));
		if (lane == l) atomicMin(bthr + (size_t)b * SKG_THR_SLOTS + slot, __float_as_int(m));
		todo &= ~__ballot(mine);
	}
}

// kdMoveParticles (kd.c:702-732) for one lane: a step of fStep against the
// gradient (ax, ay, az), then the periodic wrap into the box; a wrap sets the
// mover's flag (mwrap) that sends its wavefront back to the tree walk for the
// rest of the block.
#define MWRAP_OUT 8
__device__ __forceinline__ void move_point(float4 *__restrict__ ppos, int pi, const float4 &p, float ax, float ay,
		float az, float fStep, float cx, float cy, float cz, float Lx, float Ly, float Lz, int *__restrict__ mwrap) {
	const float ss = sq3(ax, ay, az);
	float ai = (float)sqrt((double)ss);
	if (ai > 0.0f) ai = (float)((double)fStep / sqrt((double)ss));
	else ai = 0.0f;
	float r[3] = {fsub(p.x, fmul(ai, ax)), fsub(p.y, fmul(ai, ay)), fsub(p.z, fmul(ai, az))};
	const float C[3] = {cx, cy, cz}, LL[3] = {Lx, Ly, Lz};
	int wrapped = 0;                                 // axes crossed, one bit each
	for (int j = 0; j < 3; ++j) {
		if ((double)r[j] > (double)C[j] + 0.5 * (double)LL[j]) { r[j] = fsub(r[j], LL[j]); wrapped |= 1 << j; }
		if ((double)r[j] <= (double)C[j] - 0.5 * (double)LL[j]) { r[j] = fadd(r[j], LL[j]); wrapped |= 1 << j; }
	}
	ppos[pi] = make_float4(r[0], r[1], r[2], p.w);
	// mwrap holds the axis whose far side the block's list covers (1 + axis,
	// 0 for none) and gains MWRAP_OUT when the mover crosses another face
	if (wrapped && mwrap) {
		const int code = mwrap[pi] & 7;
		if (code == 0 || code > 3 || (wrapped & ~(1 << (code - 1)))) mwrap[pi] = code | MWRAP_OUT;
	}
}

// Leaf lists for a block of move steps.  Entities do not move and ScatterCut
// only removes balls, so while the active list (and with it every tile of 64
// movers) is unchanged, the leaves any step can need are those whose ball
// reach, widened by the distance a mover can still travel (margin), holds a
// lane's point now.  One walk per block then replaces one walk per step.  A
// tile with more than cap entries gets count -1 and walks the tree.  Near a
// periodic face the listed entities include the replicas there; a mover that
// crosses the face during the block jumps to the far side, which this list
// does not cover, so the move step flags it (mwrap, cleared here: no face
// covered) and its wavefront walks the tree for the rest of the block.  The
// fused build (k_cover_build) also lists the far side of a lane's one near
// face.
//
// ELIST: the list holds entities instead of leaves -- every live entity whose
// ball, widened by the margin, holds a lane's point now, in tree order (leaf
// by leaf, entities in index order, as the walk meets them), so a lane's hits
// and its sums keep the walk's order.  A lane tests only leaves of its own
// box, so every listed entity is of some lane's box.
template <bool ELIST>
__global__ __launch_bounds__(SKG_TILE_T) void k_cover_lists(const float4 *__restrict__ ppos,
		const int *__restrict__ plist, int np, const float4 *__restrict__ nlo, const float4 *__restrict__ nhi,
		int nLeaf2, const int *__restrict__ pbox, const DevBox *__restrict__ bt, int micro, float msteps,
		float mabs, int per, float cx, float cy, float cz, float Lx, float Ly, float Lz, int cap,
		int *__restrict__ tlist, int *__restrict__ tcnt, const int *__restrict__ npDev,
		const float4 *__restrict__ epos, const float4 *__restrict__ eaux, const int *__restrict__ lstart, int B,
		int ne, int *__restrict__ mwrap) {
	if (npDev) np = *npDev;
	const int lane = threadIdx.x & 63;
	const int base = xcd_tile(blockIdx.x, gridDim.x) * SKG_TILE_T + (threadIdx.x & ~63);
	if (base >= np) return;
	const int tile = base >> 6;
	const int ii = base + lane;
	const bool valid = ii < np;
	const int pi = plist ? plist[valid ? ii : base] : (valid ? ii : base);
	const float4 p = ppos[pi];
	const int pb = pbox ? pbox[pi] : 0;
	// each lane's margin: msteps (1.02 x the block's steps) of its own box's step
	const float margin = fmaf(msteps, micro ? bt[pb].fMicro : bt[pb].fStep, mabs);
	if (valid && mwrap) mwrap[pi] = 0;
	int n = 0, node = 1;
	int *out = tlist + (size_t)tile * cap;
	// entries collect in one VGPR, lane n % 64 holding entry n, and go out as
	// one coalesced store per 64 (cap is a multiple of 64)
	int buf = 0;
	auto emit = [&](int idx) {
		if (lane == (n & 63)) buf = idx;
		++n;
		if ((n & 63) == 0 && n <= cap) out[n - 64 + lane] = buf;
	};
	while (node) {
		node = uniform_int(node);
		const float4 lo = nlo[node], hi = nhi[node];
		bool need = false;
		float c = 0.f;
		if (valid && lo.w >= 0.f && brange_has(hi.w, pb)) {
			// hardware square root (about 1 ulp): covered by the factor below
			c = __builtin_amdgcn_sqrtf(lo.w);
			const float r = c + margin;
			need = point_box_d2<false>(lo, hi, p.x, p.y, p.z, 0.f) < r * r * 1.000001f;
		}
		if (__any(need)) {
			if (node < nLeaf2) { node = 2 * node; continue; }
			if (!ELIST) {
				emit(node - nLeaf2);
			} else if (n <= cap) {
				// An entity of ball r is listed when its reach r + margin holds
				// the lane's point.  With c the leaf's largest live ball radius,
				// (r + m)^2 <= r^2 (1 + m/c) + m (c + m) for any c > 0 (since
				// r^2/c + c >= 2r), so one fma per entity replaces a square
				// root; the bound is exact for the leaf's largest ball.
				c = fmaxf(c, 1e-30f);
				const float k1 = fmaf(margin, 1.f / c, 1.f) * 1.000002f;
				const float k2 = margin * (c + margin) * 1.000002f;
				const int2 rg = leaf_range(lstart, node - nLeaf2, B, ne);
				const int start = uniform_int(rg.x), cnt = uniform_int(rg.y - rg.x);
				for (int t0 = 0; t0 < cnt; t0 += 4) {
					// four records per scalar load (the arrays carry SKG_EPAD spare records)
					float4 A4[4], X4[4];
#pragma unroll
					for (int u = 0; u < 4; ++u) { A4[u] = epos[start + t0 + u]; X4[u] = eaux[start + t0 + u]; }
#pragma unroll
					for (int u = 0; u < 4; ++u) {
						if (t0 + u >= cnt) break;
						if (X4[u].x < 0.f) continue;          // cut (or filler): never hit again
						const float dx = fsub(A4[u].x, p.x), dy = fsub(A4[u].y, p.y), dz = fsub(A4[u].z, p.z);
						const bool hit = need && sq3(dx, dy, dz) < fmaf(X4[u].x, k1, k2);
						if (__any(hit)) emit(start + t0 + u);
					}
				}
			}
		}
		node = (node >> __builtin_ctz(~node)) + 1;
		if (node == 1) node = 0;
	}
	const int rem = n & 63;
	if (n <= cap && rem && lane < rem) out[n - rem + lane] = buf;
	if (lane == 0) tcnt[tile] = n <= cap ? n : -1;
}

// pbox (indexed like ppos; none: box 0) gives each point's box: a lane opens
// only nodes holding its box, and the entity array starts every box on a new
// leaf, so every entity a lane tests is of its own box.  MODE 1 takes the
// box's step (micro: its micro-step).
template <int MODE, bool COMP>
__global__ __launch_bounds__(SKG_TILE_T) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_cover(
		float4 *__restrict__ ppos, const int *__restrict__ plist, int np,
		const float4 *__restrict__ epos, const float4 *__restrict__ eaux, const float4 *__restrict__ eorg,
		const int *__restrict__ ecode, int ne, const float4 *__restrict__ nlo, const float4 *__restrict__ nhi,
		int nLeaf2, int B, float *__restrict__ out_sum, int *__restrict__ ehit, int stamp,
		int *__restrict__ bthr, const int *__restrict__ pbox, const DevBox *__restrict__ bt, int micro,
		float cx, float cy, float cz, float Lx, float Ly, float Lz, const int *__restrict__ lstart,
		const int *__restrict__ tlist, const int *__restrict__ tcnt, int tcap, const int *__restrict__ npDev,
		const int *__restrict__ stampDev, int elist, const int *__restrict__ eblk, float *__restrict__ out_dl,
		int *__restrict__ mwrap) {
	if (npDev) np = *npDev;                       // device-resident active count and step stamp
	if (stampDev) stamp = *stampDev;
	// staging slots for the entity-list scan, 64 per wavefront
	__shared__ float4 s_eA[SKG_TILE_T], s_eX[SKG_TILE_T];
	__shared__ int s_eK[SKG_TILE_T], s_eE[SKG_TILE_T];
	const int lane = threadIdx.x & 63;
	const int base = xcd_tile(blockIdx.x, gridDim.x) * SKG_TILE_T + (threadIdx.x & ~63);
	if (base >= np) return;                       // whole wavefront past the end
	const int ii = base + lane;
	const bool valid = ii < np;
	const int pi = plist ? plist[valid ? ii : base] : (valid ? ii : base);
	const float4 p = ppos[pi];
	const int pb = pbox ? pbox[pi] : 0;
	Sum<COMP> s0, s1, s2;
	s0.init(); s1.init(); s2.init();
	float hmin = INFINITY;
	const bool dl = MODE == 0 && out_dl != nullptr;
	// With a leaf list for this tile (built at the start of the block, see
	// k_cover_lists) the wavefront scans those leaves and walks no tree.
	// Otherwise a node is opened when its largest live ball reaches at least
	// one lane's own point (so scattered blocks late in the move loop walk
	// only the union of their lanes' paths).  The node index is wave-uniform,
	// so node bounds and entity records are read through the scalar cache.
	const int tile = base >> 6;
	int nl = tcnt ? uniform_int(tcnt[tile]) : -1;
	// a lane that crossed a periodic face since the list was built is on the
	// far side, which the list does not cover: the wavefront walks the tree
	if (nl >= 0 && mwrap && __any(valid && (mwrap[pi] & MWRAP_OUT) != 0)) nl = -1;
	int node = nl >= 0 ? 0 : 1;
	if (elist && nl > 0) {
		// entity list (k_cover_lists<true>): the listed entities in tree
		// order; a tile that straddles two boxes lists entities of both, and a
		// lane takes only its own box's range of the entity array (that test
		// only in such a tile).  The list is staged 64 entries at a time in
		// this wavefront's slots of LDS: lane j gathers entry j's records
		// (one coalesced index read, record reads near one another in tree
		// order), then every lane reads each staged record at one address, a
		// broadcast.  The wavefront reads back only what it wrote itself, and
		// LDS operations of one wavefront complete in order.
		const int *L = tlist + (size_t)tile * tcap;
		int elo = 0, ehi = 0x7FFFFFFF;
		if (pbox) { elo = eblk[pb] * 256; ehi = eblk[pb + 1] * 256; }
		const int wb = threadIdx.x & ~63;
		auto scan_list = [&](auto straddle) {
			for (int c = 0; c < nl; c += 64) {
				const int m = min(64, nl - c);
				if (lane < m) {
					const int e = L[c + lane];
					s_eA[threadIdx.x] = epos[e];
					s_eX[threadIdx.x] = eaux[e];
					s_eK[threadIdx.x] = COMP ? ecode[e] : 13;
					s_eE[threadIdx.x] = e;
				}
				__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
				__builtin_amdgcn_wave_barrier();
				__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
				for (int j = 0; j < m; j += 2) {
					float4 A2[2], X2[2];
					int K2[2], E2[2];
#pragma unroll
					for (int u = 0; u < 2; ++u) {
						A2[u] = s_eA[wb + j + u]; X2[u] = s_eX[wb + j + u];
						K2[u] = s_eK[wb + j + u]; E2[u] = s_eE[wb + j + u];
					}
#pragma unroll
					for (int u = 0; u < 2; ++u) {
						if (j + u >= m) break;
						const int e = E2[u];
						const bool mine = decltype(straddle)::value ? (valid && e >= elo && e < ehi) : valid;
						cover_hit<MODE, COMP>(e, A2[u], X2[u], K2[u], mine, p, eorg, ehit, stamp, hmin, Lx, Ly, Lz,
								s0, s1, s2);
					}
				}
				// every lane has read this chunk before the next overwrites it
				__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
				__builtin_amdgcn_wave_barrier();
				__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
			}
		};
		const int pb0 = uniform_int(pb);
		if (__any(valid && pb != pb0)) scan_list(std::true_type());
		else scan_list(std::false_type());
	}
	for (int k = 0; !elist && k < nl; ++k) {
		const int leaf = uniform_int(tlist[(size_t)tile * tcap + k]);
		// the list holds what any of the block's steps can reach; this step
		// scans a leaf only where the walk below would open it (leaves hold
		// one box each, and a tile may straddle two boxes)
		const float4 lo = nlo[nLeaf2 + leaf], hi = nhi[nLeaf2 + leaf];
		const bool want = valid && (!pbox || brange_has(hi.w, pb)) &&
				point_box_d2<false>(lo, hi, p.x, p.y, p.z, 0.f) < lo.w;
		if (!__any(want)) continue;
		const int2 rg = leaf_range(lstart, leaf, B, ne);
		cover_scan_staged<MODE, COMP>(uniform_int(rg.x), uniform_int(rg.y - rg.x), want, p, epos, eaux, eorg, ecode,
				ehit, stamp, hmin, Lx, Ly, Lz, s0, s1, s2, dl, s_eA + (threadIdx.x & ~63), s_eX + (threadIdx.x & ~63),
				s_eK + (threadIdx.x & ~63), lane);
	}
	while (node) {
		node = uniform_int(node);
		const float4 lo = nlo[node], hi = nhi[node];
		const bool need = valid && brange_has(hi.w, pb) && point_box_d2<false>(lo, hi, p.x, p.y, p.z, 0.f) < lo.w;
		const bool visit = __any(need);
		if (visit && node < nLeaf2) { node = 2 * node; continue; }
		if (visit) {
			const int2 rg = leaf_range(lstart, node - nLeaf2, B, ne);
			cover_scan_staged<MODE, COMP>(uniform_int(rg.x), uniform_int(rg.y - rg.x), need, p, epos, eaux, eorg,
					ecode, ehit, stamp, hmin, Lx, Ly, Lz, s0, s1, s2, dl, s_eA + (threadIdx.x & ~63),
					s_eX + (threadIdx.x & ~63), s_eK + (threadIdx.x & ~63), lane);
		}
		// next node in pre-order after node's subtree (strip trailing ones, step)
		node = (node >> __builtin_ctz(~node)) + 1;
		if (node == 1) node = 0;
	}
	if (MODE == 1 && bthr) box_min_atomic(hmin, valid, pb, bthr);
	if (!valid) return;
	if (MODE == 0) {
		out_sum[ii] = s0.val();
		if (dl) out_dl[ii] = s1.val() * (float)(-M_1_PI / 6.0);
		return;
	}
	move_point(ppos, pi, p, s0.val(), s1.val(), s2.val(), micro ? bt[pb].fMicro : bt[pb].fStep, cx, cy, cz, Lx,
			Ly, Lz, mwrap);
}

// The first step of a 5-step block with its tile's entity list built on the
// way: one walk with the block's margin does what k_cover_lists<true> and
// k_cover<1> would do in two.  A leaf is opened when its largest live ball,
// widened by the lane's margin, reaches a lane's point; each live entity in it
// is listed when its own widened reach holds a lane's point (the same fma
// bound as k_cover_lists<true>), and it adds to a lane's gradient when its
// ball holds the lane's point.  The widened test opens every node the plain
// walk opens (and more, whose entities hold no lane's point), and both visit
// leaves in tree order, so the hits and the order of the sums are k_cover's.
// Then the lanes' points move.  The list and count go to tlist / tcnt as
// k_cover_lists<true> writes them.
template <bool COMP>
__global__ __launch_bounds__(SKG_TILE_T) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_cover_build(
		float4 *__restrict__ ppos, const int *__restrict__ plist, int np, const float4 *__restrict__ epos,
		const float4 *__restrict__ eaux, const float4 *__restrict__ eorg, const int *__restrict__ ecode, int ne,
		const float4 *__restrict__ nlo, const float4 *__restrict__ nhi, int nLeaf2, int B, int *__restrict__ ehit,
		int *__restrict__ bthr, const int *__restrict__ pbox, const DevBox *__restrict__ bt, int micro,
		float msteps, float mabs, int per, float wsteps, float cx, float cy, float cz, float Lx, float Ly,
		float Lz, const int *__restrict__ lstart, int cap, int *__restrict__ tlist, int *__restrict__ tcnt,
		const int *__restrict__ npDev, const int *__restrict__ stampDev, int *__restrict__ mwrap) {
	if (npDev) np = *npDev;
	const int stamp = stampDev ? *stampDev : 0;
	// staging slots for an opened leaf's entity records, 64 per wavefront
	__shared__ float4 s_bA[SKG_TILE_T], s_bX[SKG_TILE_T];
	__shared__ int s_bK[SKG_TILE_T];
	const int lane = threadIdx.x & 63;
	const int base = xcd_tile(blockIdx.x, gridDim.x) * SKG_TILE_T + (threadIdx.x & ~63);
	if (base >= np) return;                       // whole wavefront past the end
	const int tile = base >> 6;
	const int wb = threadIdx.x & ~63;
	const int ii = base + lane;
	const bool valid = ii < np;
	const int pi = plist ? plist[valid ? ii : base] : (valid ? ii : base);
	const float4 p = ppos[pi];
	const int pb = pbox ? pbox[pi] : 0;
	const float fStep = micro ? bt[pb].fMicro : bt[pb].fStep;
	const float margin = fmaf(msteps, fStep, mabs);
	// A lane within wsteps steps of exactly one face (periodic) may cross it
	// during the block: the list then also holds what reaches the lane's image
	// q across that face (with the full margin), and mwrap records the axis
	// (1 + axis), so that a crossing of that face leaves the list valid.  A
	// lane near two or more faces gets code 4, one near none 0: then any
	// crossing sends its wavefront to the tree walk.
	float4 q = p;
	int code = 0;
	if (per) {
		const float cc[3] = {cx, cy, cz}, LL[3] = {Lx, Ly, Lz}, pp[3] = {p.x, p.y, p.z};
		const float wm = fmaf(wsteps, fStep, mabs);
		float qq[3] = {p.x, p.y, p.z};
		int nf = 0;
#pragma unroll
		for (int j = 0; j < 3; ++j) {
			const float h = 0.5f * LL[j];
			if (pp[j] + wm > cc[j] + h) { qq[j] = pp[j] - LL[j]; code = 1 + j; ++nf; }
			if (pp[j] - wm <= cc[j] - h) { qq[j] = pp[j] + LL[j]; code = 1 + j; ++nf; }
		}
		if (nf > 1) code = 4;
		if (code >= 1 && code <= 3) q = make_float4(qq[0], qq[1], qq[2], p.w);
	}
	const bool img = valid && code >= 1 && code <= 3;
	const bool wimg = __any(img);
	if (valid && mwrap) mwrap[pi] = code;
	Sum<COMP> s0, s1, s2;
	s0.init(); s1.init(); s2.init();
	float hmin = INFINITY;
	int n = 0, node = 1;
	int *out = tlist + (size_t)tile * cap;
	int buf = 0;                                   // entry n in lane n % 64, stored 64 at a time
	auto emit = [&](int idx) {
		if (lane == (n & 63)) buf = idx;
		++n;
		if ((n & 63) == 0 && n <= cap) out[n - 64 + lane] = buf;
	};
	while (node) {
		node = uniform_int(node);
		const float4 lo = nlo[node], hi = nhi[node];
		bool need = false;
		float c = 0.f;
		if (valid && lo.w >= 0.f && brange_has(hi.w, pb)) {
			// hardware square root (about 1 ulp): covered by the factor below
			c = __builtin_amdgcn_sqrtf(lo.w);
			const float r = c + margin, r2 = r * r * 1.000001f;
			need = point_box_d2<false>(lo, hi, p.x, p.y, p.z, 0.f) < r2;
			if (wimg && img && !need) need = point_box_d2<false>(lo, hi, q.x, q.y, q.z, 0.f) < r2;
		}
		if (__any(need)) {
			if (node < nLeaf2) { node = 2 * node; continue; }
			c = fmaxf(c, 1e-30f);
			const float k1 = fmaf(margin, 1.f / c, 1.f) * 1.000002f;
			const float k2 = margin * (c + margin) * 1.000002f;
			const int2 rg = leaf_range(lstart, node - nLeaf2, B, ne);
			const int start = uniform_int(rg.x), cnt = uniform_int(rg.y - rg.x);
			// the leaf's records staged in this wavefront's LDS slots, 64 at a
			// time (lane j reads record j: one round trip), then read by every
			// lane at one address; the wavefront reads back only what it wrote,
			// and its LDS operations complete in order
			for (int c0 = 0; c0 < cnt; c0 += 64) {
				const int m = min(64, cnt - c0);
				if (lane < m) {
					s_bA[threadIdx.x] = epos[start + c0 + lane];
					s_bX[threadIdx.x] = eaux[start + c0 + lane];
					s_bK[threadIdx.x] = COMP ? ecode[start + c0 + lane] : 13;
				}
				__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
				__builtin_amdgcn_wave_barrier();
				__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
				for (int t = 0; t < m; ++t) {
					const float4 A = s_bA[wb + t], X = s_bX[wb + t];
					if (X.x < 0.f) continue;              // cut (or filler): neither listed nor hit
					const int e = start + c0 + t;
					const float dx = fsub(A.x, p.x), dy = fsub(A.y, p.y), dz = fsub(A.z, p.z);
					const float reach = fmaf(X.x, k1, k2);
					bool listed = need && sq3(dx, dy, dz) < reach;
					if (wimg && img && need && !listed)
						listed = sq3(fsub(A.x, q.x), fsub(A.y, q.y), fsub(A.z, q.z)) < reach;
					if (__any(listed)) emit(e);
					cover_hit<1, COMP>(e, A, X, s_bK[wb + t], need, p, eorg, ehit, stamp, hmin, Lx, Ly, Lz, s0, s1,
							s2);
				}
				// every lane has read these records before the next chunk overwrites them
				__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
				__builtin_amdgcn_wave_barrier();
				__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
			}
		}
		node = (node >> __builtin_ctz(~node)) + 1;
		if (node == 1) node = 0;
	}
	const int rem = n & 63;
	if (n <= cap && rem && lane < rem) out[n - rem + lane] = buf;
	if (lane == 0) tcnt[tile] = n <= cap ? n : -1;
	if (bthr) box_min_atomic(hmin, valid, pb, bthr);
	if (!valid) return;
	move_point(ppos, pi, p, s0.val(), s1.val(), s2.val(), fStep, cx, cy, cz, Lx, Ly, Lz, mwrap);
}

// Sub-wave variant: S lanes per point (64/S points per wavefront), each lane
// testing every S-th entity of a visited leaf, partial sums folded across the
// S lanes at the end.  A lane-per-point wavefront walks the union of its
// lanes' paths, which costs 64 paths when the active points are scattered
// (late in the move loop); here a wavefront walks S-lane-wide paths of 64/S
// points.  Same arithmetic per pair as k_cover; only the order of the sums
// differs.
template <int MODE, bool COMP, int S>
__global__ __launch_bounds__(SKG_TILE_T) void k_cover_sub(
		float4 *__restrict__ ppos, const int *__restrict__ plist, int np,
		const float4 *__restrict__ epos, const float4 *__restrict__ eaux, const float4 *__restrict__ eorg,
		int ne, const float4 *__restrict__ nlo, const float4 *__restrict__ nhi,
		int nLeaf2, int B, float *__restrict__ out_sum, int *__restrict__ ehit, int stamp, int *__restrict__ bthr,
		const int *__restrict__ pbox, const DevBox *__restrict__ bt, int micro,
		float cx, float cy, float cz, float Lx, float Ly, float Lz, const int *__restrict__ lstart,
		const int *__restrict__ npDev, const int *__restrict__ stampDev, float *__restrict__ out_dl) {
	if (npDev) np = *npDev;
	if (stampDev) stamp = *stampDev;
	const int blk = xcd_tile(blockIdx.x, gridDim.x);
	const int gl = blk * SKG_TILE_T + threadIdx.x;
	const int ii = gl / S, sub = gl % S;
	const int wbase = (blk * SKG_TILE_T + (threadIdx.x & ~63)) / S;
	if (wbase >= np) return;                      // whole wavefront past the end
	const bool valid = ii < np;
	const int pi = plist ? plist[valid ? ii : wbase] : (valid ? ii : wbase);
	const float4 p = ppos[pi];
	const int pb = pbox ? pbox[pi] : 0;
	Sum<COMP> s0, s1, s2;
	s0.init(); s1.init(); s2.init();
	float hmin = INFINITY;
	const bool dl = MODE == 0 && out_dl != nullptr;
	int node = valid ? 1 : 0;
	while (node) {                                // the S lanes of a point move in lockstep
		if (S == 64) node = uniform_int(node);    // one point per wavefront: scalar node reads
		const float4 lo = nlo[node], hi = nhi[node];
		const bool need = brange_has(hi.w, pb) && point_box_d2<false>(lo, hi, p.x, p.y, p.z, 0.f) < lo.w;
		if (need && node < nLeaf2) { node = 2 * node; continue; }
		if (need) {
			const int2 rg = leaf_range(lstart, node - nLeaf2, B, ne);
			for (int e = rg.x + sub; e < rg.y; e += S) {
				cover_hit<MODE, COMP, true>(e, epos[e], eaux[e], -1, true, p, eorg, ehit, stamp, hmin, Lx, Ly, Lz, s0, s1, s2,
						dl);
			}
		}
		int nx = node;
		while (nx & 1) nx >>= 1;
		node = nx + 1;
		if (node == 1) node = 0;
	}
	// fold the S partial sums of each point (xor partners stay within the group)
#pragma unroll
	for (int o = S / 2; o > 0; o >>= 1) {
		s0.xor_merge(o);
		if (MODE == 1 || dl) s1.xor_merge(o);
		if (MODE == 1) s2.xor_merge(o);
	}
	if (MODE == 1 && bthr) box_min_atomic(hmin, valid, pb, bthr);
	if (!valid || sub != 0) return;
	if (MODE == 0) {
		out_sum[ii] = s0.val();
		if (dl) out_dl[ii] = s1.val() * (float)(-M_1_PI / 6.0);
		return;
	}
	const float fStep = micro ? bt[pb].fMicro : bt[pb].fStep;
	const float ax = s0.val(), ay = s1.val(), az = s2.val();
	const float ss = sq3(ax, ay, az);
	float ai = (float)sqrt((double)ss);
	if (ai > 0.0f) ai = (float)((double)fStep / sqrt((double)ss));
	else ai = 0.0f;
	float r[3] = {fsub(p.x, fmul(ai, ax)), fsub(p.y, fmul(ai, ay)), fsub(p.z, fmul(ai, az))};
	const float C[3] = {cx, cy, cz}, LL[3] = {Lx, Ly, Lz};
	for (int j = 0; j < 3; ++j) {
		if ((double)r[j] > (double)C[j] + 0.5 * (double)LL[j]) r[j] = fsub(r[j], LL[j]);
		if ((double)r[j] <= (double)C[j] - 0.5 * (double)LL[j]) r[j] = fadd(r[j], LL[j]);
	}
	ppos[pi] = make_float4(r[0], r[1], r[2], p.w);
}

// =====================================================================
// Small kernels: gathers, replicas, entity setup, ScatterCut, prune.
// =====================================================================
// ScatterCriterion (kd.c:600-627) depends only on the species: flag = the
// particle's species is in the selected set.
__global__ void k_species_flags(const unsigned char *__restrict__ type, int n, int mask, int *flag) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) flag[i] = (type[i] & mask) != 0;
}

__global__ void k_gather_scat(const float *__restrict__ pos, const float *__restrict__ mass,
		const int *__restrict__ ord, const int *__restrict__ boxOf, int nS, float4 *out4, int *obox) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nS) return;
	const int o = ord[i];
	out4[i] = make_float4(pos[3 * o], pos[3 * o + 1], pos[3 * o + 2], mass[o]);
	obox[i] = boxOf[o];
}

// inv[ord[i]] = i: the inverse of an order over its n entries.
__global__ void k_inverse(const int *__restrict__ ord, int n, int *inv) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) inv[ord[i]] = i;
}

__global__ void k_permute_s(const float4 *__restrict__ in4, const int *__restrict__ inord,
		const int *__restrict__ inbox, const int *__restrict__ perm, int n, float4 *out4, int *outord,
		int *outbox) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int j = perm[i];
	out4[i] = in4[j];
	outord[i] = inord[j];
	outbox[i] = inbox[j];
}

// Box of each particle (binary search over box offsets) and its species from
// the box's tipsy order (gas, dark, star).  binfo: off, nGas, nDark per box.
__global__ void k_box_assign(int n, int nBox, const int *__restrict__ binfo, int *boxOf, unsigned char *type) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int lo = 0, hi = nBox - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (binfo[3 * mid] <= i) lo = mid; else hi = mid - 1;
	}
	const int l = i - binfo[3 * lo], ng = binfo[3 * lo + 1], nd = binfo[3 * lo + 2];
	boxOf[i] = lo;
	type[i] = (unsigned char)(l < ng ? SKG_GAS : (l < ng + nd ? SKG_DARK : SKG_STAR));
}
__global__ void k_temp_nongas(const unsigned char *__restrict__ type, int n, float *temp) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n && type[i] != SKG_GAS) temp[i] = 0.f;
}

// Replica scatterers: the 26 images whose ball reaches the box (smooth1.c:283-331),
// tested with INTERSECTNP arithmetic against the float box bounds.
__device__ __forceinline__ float box_np_d2(float lx, float ly, float lz, float hx, float hy, float hz,
		float x, float y, float z) {
	float dx = fsub(lx, x), dx1 = fsub(x, hx), dy = fsub(ly, y), dy1 = fsub(y, hy);
	float dz = fsub(lz, z), dz1 = fsub(z, hz);
	float gx = dx > 0.f ? dx : (dx1 > 0.f ? dx1 : 0.f);
	float gy = dy > 0.f ? dy : (dy1 > 0.f ? dy1 : 0.f);
	float gz = dz > 0.f ? dz : (dz1 > 0.f ? dz1 : 0.f);
	return sq3(gx, gy, gz);
}

__global__ void k_rep_mask(const float4 *__restrict__ spos, const float *__restrict__ b2, int nS,
		float lx, float ly, float lz, float hx, float hy, float hz, float Lx, float Ly, float Lz,
		unsigned *mask, int *cnt) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nS) return;
	float4 p = spos[i];
	unsigned m = 0;
	int bit = 0;
	for (int ix = -1; ix <= 1; ++ix) {
		float x = fadd(p.x, (float)ix * Lx);
		for (int iy = -1; iy <= 1; ++iy) {
			float y = fadd(p.y, (float)iy * Ly);
			for (int iz = -1; iz <= 1; ++iz, ++bit) {
				float z = fadd(p.z, (float)iz * Lz);
				if (!(ix || iy || iz)) continue;
				if (box_np_d2(lx, ly, lz, hx, hy, hz, x, y, z) < b2[i]) m |= 1u << bit;
			}
		}
	}
	mask[i] = m;
	cnt[i] = __popc(m);
}

// Unsorted entity list: [0,nS) scatterers, then replicas in (scatterer, image) order.
__global__ void k_rep_write(const float4 *__restrict__ spos, const float *__restrict__ b2, int nS,
		const unsigned *__restrict__ mask, const int *__restrict__ off, float Lx, float Ly, float Lz,
		float4 *e4, float *eb2, int *esrc) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nS) return;
	float4 p = spos[i];
	e4[i] = p; eb2[i] = b2[i]; esrc[i] = i;
	unsigned m = mask[i];
	int o = nS + off[i], bit = 0;
	for (int ix = -1; ix <= 1; ++ix)
		for (int iy = -1; iy <= 1; ++iy)
			for (int iz = -1; iz <= 1; ++iz, ++bit)
				if (m & (1u << bit)) {
					e4[o] = make_float4(fadd(p.x, (float)ix * Lx), fadd(p.y, (float)iy * Ly),
							fadd(p.z, (float)iz * Lz), p.w);
					eb2[o] = b2[i]; esrc[o] = i; ++o;
				}
}

// Box of every unsorted entity (from its source scatterer).
__global__ void k_entity_box(const int *__restrict__ esrcu, const int *__restrict__ sbox, int ne, int *ebu) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < ne) ebu[i] = sbox[esrcu[i]];
}
// First sorted entity of each box (entities are sorted by box first).
__global__ void k_box_first(const int *__restrict__ ebu, const int *__restrict__ perm, int ne, int *first) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= ne) return;
	const int b = ebu[perm[i]];
	if (i == 0 || ebu[perm[i - 1]] != b) first[b] = i;
}
// Filler entities of the padded array: dead, NaN position, never hit.
__global__ void k_entity_fill(int ne, const int *__restrict__ eblk, int nBox, float4 *e4, float4 *eaux,
		float4 *eorg, int *ecode, float *erho, int *esrc, int *elive, int *ehit, float *eb2s, int *ebox,
		unsigned long long *ekey, int bbits) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= ne) return;
	int lo = 0, hi = nBox - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (eblk[mid] * 256 <= i) lo = mid; else hi = mid - 1;
	}
	e4[i] = make_float4(NAN, NAN, NAN, 0.f);
	eaux[i] = make_float4(-1.f, 0.f, 0.f, 0.f);
	eorg[i] = make_float4(NAN, NAN, NAN, 13.f);
	ecode[i] = 13;
	erho[i] = 0.f; esrc[i] = -1; elive[i] = 0; ehit[i] = -1; eb2s[i] = -1.f; ebox[i] = lo;
	// after every real key of the box
	ekey[i] = ((unsigned long long)lo << (63 - bbits)) | ((1ull << (63 - bbits)) - 1);
}

// Entities in Morton order + per-entity constants (smooth1.c:447-448, 250-251),
// written to their box's padded range.
__global__ void k_entity_setup(const float4 *__restrict__ e4u, const float *__restrict__ eb2u,
		const int *__restrict__ esrcu, const int *__restrict__ perm, int ne,
		const int *__restrict__ ebu, const int *__restrict__ bfirst, const int *__restrict__ eblk,
		const float *__restrict__ srho, const float4 *__restrict__ spos, float Lx, float Ly, float Lz,
		float4 *e4, float4 *eaux, float4 *eorg, int *ecode, float *erho, int *esrc, int *elive, int *ehit,
		float *eb2s, const unsigned long long *__restrict__ ksorted, unsigned long long *ekey) {
	int ii = blockIdx.x * blockDim.x + threadIdx.x;
	if (ii >= ne) return;
	int j = perm[ii];
	const int bx = ebu[j];
	const int i = eblk[bx] * 256 + (ii - bfirst[bx]);
	float4 p = e4u[j];
	float b2 = eb2u[j];
	float ih2 = (float)(4.0 / (double)b2);
	double dih2 = ih2;
	float gnorm = (float)(M_1_PI * dih2 * dih2 * sqrt(dih2) * (double)p.w);
	float dnorm = (float)(0.5 * M_1_PI * sqrt(dih2) * dih2);
	e4[i] = p;
	const float4 o = spos[esrcu[j]];
	int ix = 0, iy = 0, iz = 0;
	if (Lx > 0.f) {
		ix = (int)rintf((p.x - o.x) / Lx); iy = (int)rintf((p.y - o.y) / Ly); iz = (int)rintf((p.z - o.z) / Lz);
	}
	eorg[i] = make_float4(o.x, o.y, o.z, (float)(9 * (ix + 1) + 3 * (iy + 1) + (iz + 1)));
	ecode[i] = 9 * (ix + 1) + 3 * (iy + 1) + (iz + 1);
	eaux[i] = make_float4(b2, ih2, gnorm, dnorm);
	eb2s[i] = b2;
	esrc[i] = esrcu[j];
	erho[i] = srho ? srho[esrcu[j]] : 0.f;
	elive[i] = 1;
	ehit[i] = -1;
	ekey[i] = ksorted[ii];
}

__global__ void k_entity_rho(const int *__restrict__ esrc, int ne, const float *__restrict__ srho, float *erho) {
	int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e < ne) erho[e] = esrc[e] >= 0 ? srho[esrc[e]] : 0.f;
}

__global__ void k_combine_rho(const float *__restrict__ g, const float *__restrict__ s,
		const float *__restrict__ b2, const int *__restrict__ ord, int nS, float *rho_s,
		float *rho_ord, float *b2_ord) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nS) return;
	float r = fadd(g[i], s[i]);
	rho_s[i] = r;
	rho_ord[ord[i]] = r;
	b2_ord[ord[i]] = b2[i];
}

__global__ void k_fill_f(float *a, int n, float v) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) a[i] = v;
}
__global__ void k_fill_i(int *a, int n, int v) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) a[i] = v;
}
__global__ void k_iota(int *a, int n) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) a[i] = i;
}

// ScatterCut threshold: smallest density among live entities that held an
// active mover this step; on the initial step live entities that held none
// get density 0 (smooth1.c:460-470, 497-507).  Block minima go to partials.
__global__ __launch_bounds__(256) void k_scat_thr(const int *__restrict__ elive, const int *__restrict__ ehit,
		float *erho, float4 *eaux, int ne, const int *__restrict__ stampDev, int bInitial, float *part) {
	const int stamp = *stampDev;
	__shared__ float sm[4];
	int e = blockIdx.x * blockDim.x + threadIdx.x;
	float v = INFINITY;
	if (e < ne && elive[e]) {
		if (ehit[e] == stamp) { float r = erho[e]; if (r > 0.f) v = r; }
		else if (bInitial) { erho[e] = 0.f; eaux[e].w = INFINITY; }
	}
	v = wave_min(v);
	if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
	__syncthreads();
	if (threadIdx.x == 0) part[blockIdx.x] = fminf(fminf(sm[0], sm[1]), fminf(sm[2], sm[3]));
}
__global__ __launch_bounds__(256) void k_min_final(const float *__restrict__ part, int nb, float *out) {
	__shared__ float sm[4];
	float v = INFINITY;
	for (int b = threadIdx.x; b < nb; b += 256) v = fminf(v, part[b]);
	v = wave_min(v);
	if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
	__syncthreads();
	if (threadIdx.x == 0) out[0] = fminf(fminf(sm[0], sm[1]), fminf(sm[2], sm[3]));
}

// One wavefront per box folds that box's block partials (a box's entities
// fill whole 256-entity blocks).
__global__ __launch_bounds__(64) void k_box_min(const float *__restrict__ part, const int *__restrict__ eblk,
		float *out, int *liveTotal) {
	const int b = blockIdx.x;
	if (b == 0 && threadIdx.x == 0) liveTotal[0] = 0;       // summed by k_box_sum after the cut
	float v = INFINITY;
	for (int i = eblk[b] + (int)threadIdx.x; i < eblk[b + 1]; i += 64) v = fminf(v, part[i]);
	v = wave_min(v);
	if (threadIdx.x == 0) out[b] = v;
}
// From the move loop on, an entity record's w (the scatter density's kernel
// normalisation, used only by skg_density, which rebuilds the records) holds
// the entity's density, which the move step reads with the record for the
// ScatterCut threshold (smooth1.c:460-461 takes the smallest positive one), or
// +inf for a density of 0, so that a hit takes a plain minimum.
__global__ void k_aux_rho(float4 *eaux, const float *__restrict__ erho, int ne) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < ne) eaux[i].w = erho[i] > 0.f ? erho[i] : INFINITY;
}

// Per-step reset for the move kernel's threshold atomics: every box's slots
// to +inf, and the live total k_box_sum adds into to 0.
// ml (optional): also advances the step stamp (ml[0]), one launch fewer per step
__global__ void k_thr_reset(float *slots, int n, int *liveTotal, int *ml) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) slots[i] = INFINITY;
	if (i == 0) { liveTotal[0] = 0; if (ml) ml[0] += 1; }
}
// thr[b] = the minimum over box b's slots (one wavefront per box).
__global__ __launch_bounds__(64) void k_thr_fold(const float *__restrict__ slots, float *thr) {
	const float v = wave_min(slots[(size_t)blockIdx.x * SKG_THR_SLOTS + threadIdx.x]);
	if (threadIdx.x == 0) thr[blockIdx.x] = v;
}
__global__ __launch_bounds__(64) void k_box_sum(const int *__restrict__ part, const int *__restrict__ eblk,
		int *out, int *liveTotal) {
	const int b = blockIdx.x;
	int v = 0;
	for (int i = eblk[b] + (int)threadIdx.x; i < eblk[b + 1]; i += 64) v += part[i];
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	if (threadIdx.x == 0) { out[b] = v; atomicAdd(liveTotal, v); }
}

// ScatterCut (smooth1.c:387-405): keep entities with density >= their box's
// threshold (no hit in the box: threshold 0, everything kept).  Block live
// counts go to partials.
__global__ __launch_bounds__(256) void k_scat_cut(int *elive, float4 *eaux, const float *__restrict__ erho, int ne,
		const int *__restrict__ ebox, const float *__restrict__ thr, int *part) {
	__shared__ int sm[4];
	int e = blockIdx.x * blockDim.x + threadIdx.x;
	int live = 0;
	if (e < ne) {
		const float t0 = thr[ebox[e]];
		const float t = (t0 == INFINITY) ? 0.f : t0;
		live = elive[e];
		if (live && !(erho[e] >= t)) { live = 0; elive[e] = 0; eaux[e].x = -1.f; }
	}
	const int c = __popcll(__ballot(live));
	if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = c;
	__syncthreads();
	if (threadIdx.x == 0) part[blockIdx.x] = sm[0] + sm[1] + sm[2] + sm[3];
}

// ScatterCut after the initial step, from each box's live entities sorted by
// density.  A box's threshold never falls: every live entity has density >=
// the last threshold, and the next one is a minimum over hit (live) entities;
// a box with no hit cuts nothing.  So the entities a cut removes are the next
// ones in the box's ascending order: one binary search per box, then marking
// only the newly removed (the same set as k_scat_cut's pass over all
// entities; densities are finite).  Keys: box in the high word, density bits
// (>= 0, so ordered as integers) in the low; dead entities sort last.
__global__ void k_cut_keys(const int *__restrict__ elive, const int *__restrict__ ebox,
		const float *__restrict__ erho, int ne, unsigned long long *key, int *val) {
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= ne) return;
	key[e] = elive[e] ? ((unsigned long long)(unsigned)ebox[e] << 32) | __float_as_uint(erho[e]) : ~0ull;
	val[e] = e;
}
// per box b <= nBox: first sorted position of box b (nBox: the live total)
__global__ void k_cut_bounds(const unsigned long long *__restrict__ key, int ne, int nBox, int *off, int *pos,
		float *srho, const int *__restrict__ sval, int *perm) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < ne) { srho[i] = __uint_as_float((unsigned)key[i]); perm[i] = sval[i]; }
	if (i <= nBox) {
		const unsigned long long k0 = (unsigned long long)(unsigned)i << 32;
		int a = 0, z = ne;
		while (a < z) { const int m = (a + z) >> 1; if (key[m] >= k0) z = m; else a = m + 1; }
		off[i] = a;
		if (i < nBox) pos[i] = a;
	}
}
__global__ __launch_bounds__(256) void k_cut_sorted(float *thr, const int *__restrict__ off,
		int *pos, const float *__restrict__ srho, const int *__restrict__ perm, int *elive, float4 *eaux,
		int *bscat, int *liveTotal, const int *__restrict__ eleaf, int *dirty, const float *__restrict__ slots) {
	__shared__ int snew;
	const int b = blockIdx.x;
	const int lo = pos[b], hi = off[b + 1];
	// slots (optional): the box's SKG_THR_SLOTS partial minima, folded here
	// (k_thr_fold's work, one launch fewer per step)
	float tf = 0.f;
	if (slots && threadIdx.x < 64) tf = wave_min(slots[(size_t)b * SKG_THR_SLOTS + threadIdx.x]);
	if (threadIdx.x == 0) {
		if (slots) thr[b] = tf;
		const float t0 = slots ? tf : thr[b];
		const float t = (t0 == INFINITY) ? 0.f : t0;
		int a = lo, z = hi;
		while (a < z) { const int m = (a + z) >> 1; if (srho[m] >= t) z = m; else a = m + 1; }
		snew = a;
	}
	__syncthreads();
	const int nw = snew;
	for (int i = lo + (int)threadIdx.x; i < nw; i += blockDim.x) {
		const int e = perm[i];
		elive[e] = 0;
		eaux[e].x = -1.f;
		dirty[eleaf[e]] = 1;
	}
	if (threadIdx.x == 0) {
		pos[b] = nw;
		bscat[b] = hi - nw;
		atomicAdd(liveTotal, hi - nw);
	}
}

// Device-side move-loop bookkeeping: ml[0] step stamp, ml[1] active count,
// ml[2] count after the prune.  Everything a block of steps launches reads
// these, so the block can be replayed as one graph with no host round trip.
__global__ void k_ml_set(int *ml, int i, int v) { ml[i] = v; }
__global__ void k_bump(int *ml) { ml[0] += 1; }
__global__ void k_ml_copy(int *ml) { ml[1] = ml[2]; }

// Stream compaction of the active list by the prune flags: per-256 counts, a
// one-workgroup scan of the counts, then ballot prefixes within each block.
// Order is kept, so a Morton-sorted active list stays sorted.
__global__ __launch_bounds__(256) void k_flag_counts(const int *__restrict__ flag, const int *__restrict__ nDev,
		int *bcnt) {
	__shared__ int sm[4];
	const int n = *nDev;
	const int i = blockIdx.x * 256 + threadIdx.x;
	const int f = i < n ? flag[i] : 0;
	const int c = __popcll(__ballot(f));
	if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = c;
	__syncthreads();
	if (threadIdx.x == 0) bcnt[blockIdx.x] = sm[0] + sm[1] + sm[2] + sm[3];
}
__global__ __launch_bounds__(1024) void k_scan_counts(const int *__restrict__ bcnt, int nb, int *boff, int *total) {
	__shared__ int sm[1024];
	const int t = threadIdx.x;
	const int per = (nb + 1023) / 1024;
	const int a = min(nb, t * per), b = min(nb, a + per);
	int sum = 0;
	for (int i = a; i < b; ++i) sum += bcnt[i];
	sm[t] = sum;
	__syncthreads();
	for (int o = 1; o < 1024; o <<= 1) {
		const int v = t >= o ? sm[t - o] : 0;
		__syncthreads();
		sm[t] += v;
		__syncthreads();
	}
	int run = sm[t] - sum;
	for (int i = a; i < b; ++i) { boff[i] = run; run += bcnt[i]; }
	if (t == 1023) *total = sm[1023];
}
__global__ __launch_bounds__(256) void k_compact_dev(const int *__restrict__ list, const int *__restrict__ flag,
		const int *__restrict__ nDev, const int *__restrict__ boff, int *out) {
	__shared__ int sm[4];
	const int n = *nDev;
	const int i = blockIdx.x * 256 + threadIdx.x;
	const int f = i < n ? flag[i] : 0;
	const unsigned long long m = __ballot(f);
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	if (lane == 0) sm[w] = __popcll(m);
	__syncthreads();
	int off = boff[blockIdx.x];
	for (int k = 0; k < w; ++k) off += sm[k];
	if (f) out[off + __popcll(m & ((1ull << lane) - 1ull))] = list[i];
}
__global__ void k_copy_list(const int *__restrict__ in, int *ml, int *out) {
	const int n = ml[2];
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) out[i] = in[i];
}


// kdPruneInactive test (kd.c:754-776): a mover stays active while it moved at
// least fCvg over the last block; survivors restart from here.
__global__ void k_prune(const int *__restrict__ list, const int *__restrict__ naDev, const float4 *__restrict__ mpos,
		float4 *mold, float hx, float hy, float hz, const int *__restrict__ mbox, const DevBox *__restrict__ bt,
		int *flag) {
	const int na = *naDev;
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= na) return;
	int m = list[i];
	const float cv2 = bt[mbox[m]].cvg2;
	float4 p = mpos[m], o = mold[m];
	float dx = fsub(p.x, o.x), dy = fsub(p.y, o.y), dz = fsub(p.z, o.z);
	if (dx > hx) dx = fsub(dx, 2 * hx);
	if (dx <= -hx) dx = fadd(dx, 2 * hx);
	if (dy > hy) dy = fsub(dy, 2 * hy);
	if (dy <= -hy) dy = fadd(dy, 2 * hy);
	if (dz > hz) dz = fsub(dz, 2 * hz);
	if (dz <= -hz) dz = fadd(dz, 2 * hz);
	int keep = sq3(dx, dy, dz) >= cv2;
	flag[i] = keep;
	if (keep) mold[m] = p;
}

__global__ void k_compact(const int *__restrict__ list, const int *__restrict__ flag,
		const int *__restrict__ off, int n, int *out) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n && flag[i]) out[off[i]] = list ? list[i] : i;
}

// Active movers per box (lanes of one box add once per wavefront).
// Active movers per box.  The active list is ordered by box (the movers are
// sorted by box, then along the curve, and every prune compacts the list in
// order), so the thread at a box's last mover finds the box's first by binary
// search and writes the count; no atomics.  Boxes with no mover keep 0.
__global__ void k_box_count(const int *__restrict__ list, const int *__restrict__ nDev, const int *__restrict__ mbox,
		int *cnt) {
	const int n = *nDev;
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int b = mbox[list[i]];
	if (i + 1 < n && mbox[list[i + 1]] == b) return;
	int lo = 0, hi = i;                            // the first index whose box is >= b
	while (lo < hi) {
		const int m = (lo + hi) >> 1;
		if (mbox[list[m]] < b) lo = m + 1;
		else hi = m;
	}
	cnt[b] = i + 1 - lo;
}

// CutCriterion (kd.c:555-597) per particle by iOrder.
__global__ void k_mover_flags(int n, const unsigned char *__restrict__ type, int inType, const float *__restrict__ rho,
		const float *__restrict__ mass, const float *__restrict__ temp, float fDensMin, float fTempMax,
		float fMassMax, int gd, int go, int *flag) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	int t = type[i];
	int ok = 0;
	float r = rho[i];
	if (!(mass[i] > fMassMax)) {
		switch (inType) {
		case SKG_DARK: ok = r >= fDensMin; break;
		case SKG_GAS: case SKG_DARK | SKG_GAS:
			ok = (gd && t == SKG_DARK && r >= fDensMin) ||
			     (t == SKG_GAS && r >= fDensMin && temp[i] <= fTempMax);
			break;
		case SKG_STAR: case SKG_DARK | SKG_STAR: ok = t == SKG_STAR; break;
		case SKG_GAS | SKG_STAR: case SKG_DARK | SKG_GAS | SKG_STAR:
			if (gd && t == SKG_DARK && r >= fDensMin) ok = 1;
			else if (t == SKG_GAS) ok = r >= fDensMin && temp[i] <= fTempMax;
			else ok = t == SKG_STAR && !go;
			break;
		}
	}
	flag[i] = ok;
}

__global__ void k_mover_init(const int *__restrict__ ord, int nm, const float *__restrict__ pos,
		const int *__restrict__ boxOf, float4 *mpos, float4 *mold, int *mbox) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nm) return;
	int o = ord[i];
	float4 p = make_float4(pos[3 * o], pos[3 * o + 1], pos[3 * o + 2], 0.f);
	mpos[i] = p; mold[i] = p;
	mbox[i] = boxOf[o];
}

__global__ void k_scatter_moved(const int *__restrict__ ord, int nm, const float4 *__restrict__ mpos, float *r) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nm) return;
	int o = ord[i];
	float4 p = mpos[i];
	r[3 * o] = p.x; r[3 * o + 1] = p.y; r[3 * o + 2] = p.z;
}

// =====================================================================
// Host API
// =====================================================================
extern "C" void skg_default_params(skg_params *P) {
	memset(P, 0, sizeof(*P));
	P->nSmooth = 64;
	P->nBucket = 16;
	for (int j = 0; j < 3; ++j) { P->fPeriod[j] = FLT_MAX; P->fCenter[j] = 0.f; }
	P->fTempMax = FLT_MAX;
	P->fMassMax = FLT_MAX;
	P->nMembers = 8;
	P->nMaxMembers = INT_MAX;                                // main.c:110
	P->iSoftType = SKG_SPLINE;
	P->G = 1.f;
	P->fShift = 1.f;
	P->accumulate = SKG_ACC_COMPENSATED;
}

static thread_local char g_last_err[512];
void skg_note_hip(hipError_t e, const char *call, const char *file, int line) {
	const char *f = strrchr(file, '/');
	snprintf(g_last_err, sizeof(g_last_err), "%s (%d) from %.200s at %s:%d", hipGetErrorString(e), (int)e, call,
			f ? f + 1 : file, line);
}
extern "C" const char *skg_last_error(void) { return g_last_err; }

extern "C" const char *skg_strerror(int s) {
	switch (s) {
	case SKG_OK: return "ok";
	case SKG_E_NODEVICE: return "no usable GPU";
	case SKG_E_PARAM: return "invalid parameter";
	case SKG_E_NOMEM: return "allocation failed";
	case SKG_E_STATE: return "stage called out of order";
	case SKG_E_HIP: return "HIP runtime error";
	case SKG_E_NSMOOTH: return "nSmooth exceeds the number of scatterers";
	}
	return "unknown status";
}

extern "C" int skg_create(const skg_params *P, skg_ctx **out) {
	if (!P || !out) return SKG_E_PARAM;
	*out = nullptr;
	if (P->nSmooth < 2 || P->nSmooth > 256 || P->nBucket < 1 || P->nBucket > 1024) return SKG_E_PARAM;
	if (!(P->fTau > 0.f)) return SKG_E_PARAM;
	int nd = 0;
	if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0) return SKG_E_NODEVICE;
	if (P->device < 0 || P->device >= nd) return SKG_E_PARAM;
	if (hipSetDevice(P->device) != hipSuccess) return SKG_E_NODEVICE;
	skg_ctx *c = new (std::nothrow) skg_ctx();
	if (!c) return SKG_E_NOMEM;
	c->P = *P;
	c->dev = P->device;
	c->subMax = 1 << 18;
	if (const char *e = getenv("SKG_COVER_SUB_MAX")) c->subMax = atol(e);
	c->s64Max = 1 << 14;
	if (const char *e = getenv("SKG_COVER_S64_MAX")) c->s64Max = atol(e);
	if (const char *e = getenv("SKG_KNN_BUCKET")) c->knnBucket = std::max(0, std::min(1024, atoi(e)));
	if (const char *e = getenv("SKG_KNN_WAVES")) c->knnWaves = atoi(e) == 2 ? 2 : 1;
	if (const char *e = getenv("SKG_FUSED_BUILD")) c->fusedBuild = atoi(e) != 0;
	if (const char *e = getenv("SKG_TILE_LISTS")) c->useLists = atoi(e);
	c->tlCap = c->useLists == 2 ? 512 : 1024;
	if (const char *e = getenv("SKG_TILE_CAP")) { c->tlCap = std::max(4, atoi(e) & ~3); c->tlCapFixed = 1; }
	if (const char *e = getenv("SKG_GRAPH")) c->useGraph = atoi(e);
	if (const char *e = getenv("SKG_GRAPH_SHRINK")) c->gShrink = std::max(1, atoi(e));
	if (const char *e = getenv("SKG_POT_TILE")) {
		const int t = atoi(e);
		if (t == 64 || t == 128 || t == 256) c->potTile = t;
	}
	if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess ||
	    hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking) != hipSuccess ||
	    hipEventCreateWithFlags(&c->evFork, hipEventDisableTiming) != hipSuccess ||
	    hipEventCreateWithFlags(&c->evJoin, hipEventDisableTiming) != hipSuccess ||
	    hipEventCreate(&c->e0) != hipSuccess || hipEventCreate(&c->e1) != hipSuccess) {
		delete c;
		return SKG_E_NODEVICE;
	}
	if (dalloc(&c->d_cnt, 64) || dalloc(&c->d_thr, 8)) { skg_destroy(c); return SKG_E_NOMEM; }
	// the device address of the no-allocation entry point's scratch array, once
	// per device (a failed lookup is kept and reported by that entry point)
	(void)skg_noalloc_setup(c->dev);
	*out = c;
	return SKG_OK;
}

void skg_group_free(skg_ctx *c);
void skg_drop_graph(skg_ctx *c);

extern "C" int skg_destroy(skg_ctx *c) {
	if (!c) return SKG_OK;
	(void)hipSetDevice(c->dev);
	if (c->st) (void)hipStreamSynchronize(c->st);
	skg_group_free(c);
	dfree(&c->d_pos); dfree(&c->d_vel); dfree(&c->d_mass); dfree(&c->d_soft); dfree(&c->d_temp);
	dfree(&c->d_rho); dfree(&c->d_ball2); dfree(&c->d_grp);
	dfree(&c->s_pos); dfree(&c->s_ord); dfree(&c->s_inv); dfree(&c->s_ball2); dfree(&c->s_rhoG); dfree(&c->s_rhoS);
	dfree(&c->s_heap); dfree(&c->s_box); dfree(&c->s_dlnkG); dfree(&c->s_dlnkS);
	skg_free_tree(&c->ts); skg_free_tree(&c->te); skg_free_tree(&c->tfull); skg_free_tree(&c->tfof);
	dfree(&c->e_b2); dfree(&c->e_pos); dfree(&c->e_aux); dfree(&c->e_org); dfree(&c->e_code); dfree(&c->e_rho); dfree(&c->e_src); dfree(&c->e_live); dfree(&c->e_hit);
	dfree(&c->e_box);
	dfree(&c->m_pos); dfree(&c->m_old); dfree(&c->m_ord); dfree(&c->m_list); dfree(&c->m_list2); dfree(&c->m_box); dfree(&c->m_wrap);
	dfree(&c->tl_list); dfree(&c->tl_cnt);
	skg_drop_graph(c);
	dfree(&c->d_ml); dfree(&c->m_flag); dfree(&c->m_bcnt);
	if (c->h_pin) { (void)hipHostFree(c->h_pin); c->h_pin = nullptr; }
	dfree(&c->d_bt); dfree(&c->d_boxOf); dfree(&c->d_type); dfree(&c->d_eblk); dfree(&c->d_bact); dfree(&c->d_bscat);
	dfree(&c->d_bthr); dfree(&c->d_bslot);
	dfree(&c->cs_perm); dfree(&c->cs_off); dfree(&c->cs_pos); dfree(&c->cs_rho);
	dfree(&c->e_leaf); dfree(&c->leafDirty);
	delete[] c->hbox; delete[] c->h_eblk;
	dfree(&c->d_thr); dfree(&c->d_cnt);
	skg_dfree(&c->tmp);
	dfree(&c->k64a); dfree(&c->k64b); dfree(&c->v32a); dfree(&c->v32b); dfree(&c->part);
	for (int i = 0; i < SKG_NSLOT; ++i) { skg_dfree(&c->slot[i]); c->slotBytes[i] = 0; }
	for (int i = 0; i < c->caprec; ++i) { (void)hipEventDestroy(c->prec[i].a); (void)hipEventDestroy(c->prec[i].b); }
	free(c->prec);
	if (c->e0) (void)hipEventDestroy(c->e0);
	if (c->e1) (void)hipEventDestroy(c->e1);
	if (c->evFork) (void)hipEventDestroy(c->evFork);
	if (c->evJoin) (void)hipEventDestroy(c->evJoin);
	if (c->st2) (void)hipStreamDestroy(c->st2);
	if (c->st) (void)hipStreamDestroy(c->st);
	delete c;
	return SKG_OK;
}

extern "C" int skg_load_batch(skg_ctx *c, int nBox, const int *nGas, const int *nDark, const int *nStar,
		const float *fTau, const float *pos, const float *vel, const float *mass, const float *soft,
		const float *temp) {
	if (!c || !pos || !mass || !soft || !nGas || !nDark || !nStar || nBox < 1 || nBox > SKG_MAX_BOX)
		return SKG_E_PARAM;
	long n = 0;
	int inType = -1, tg = 0, td = 0, ts = 0;
	for (int b = 0; b < nBox; ++b) {
		if (nGas[b] < 0 || nDark[b] < 0 || nStar[b] < 0) return SKG_E_PARAM;
		const long nb = (long)nGas[b] + nDark[b] + nStar[b];
		if (nb <= 0) return SKG_E_PARAM;
		if (fTau && !(fTau[b] > 0.f)) return SKG_E_PARAM;
		const int t = (nDark[b] ? SKG_DARK : 0) | (nGas[b] ? SKG_GAS : 0) | (nStar[b] ? SKG_STAR : 0);
		if (inType >= 0 && t != inType) return SKG_E_PARAM;   // one species mix per batch
		inType = t;
		n += nb; tg += nGas[b]; td += nDark[b]; ts += nStar[b];
	}
	if (n > 0x7FFFFFF0L) return SKG_E_PARAM;
	(void)hipSetDevice(c->dev);
	c->nBox = nBox; c->nGas = tg; c->nDark = td; c->nStar = ts; c->n = (int)n; c->inType = inType;
	delete[] c->hbox;
	c->hbox = new (std::nothrow) HostBox[nBox];
	if (!c->hbox) return SKG_E_NOMEM;
	std::vector<DevBox> db(nBox);
	std::vector<int> binfo(3 * (size_t)nBox);
	for (int b = 0, off = 0; b < nBox; ++b) {
		HostBox &h = c->hbox[b];
		h.off = off; h.nGas = nGas[b]; h.nDark = nDark[b]; h.nStar = nStar[b];
		h.n = nGas[b] + nDark[b] + nStar[b];
		if (fTau) {   // main.c:343-345 defaults from the box's tau
			h.tau = fTau[b]; h.cvg = (float)(0.5 * (double)h.tau); h.scoop = (float)(2.0 * (double)h.tau);
			h.step = (float)(0.5 * (double)h.cvg);
		} else {
			h.tau = c->P.fTau; h.cvg = c->P.fCvg; h.scoop = c->P.fScoop; h.step = c->P.fStep;
		}
		DevBox &d = db[b];
		d.off = off; d.n = h.n; d.fStep = h.step; d.fMicro = (float)(0.1 * (double)h.step);
		d.cvg2 = h.cvg * h.cvg; d.tau2 = h.tau * h.tau; d.scoop2 = h.scoop * h.scoop; d.pad = 0.f;
		binfo[3 * b] = off; binfo[3 * b + 1] = nGas[b]; binfo[3 * b + 2] = nDark[b];
		off += h.n;
	}
	int *dbi = nullptr;
	if (dalloc(&c->d_pos, 3 * n) || dalloc(&c->d_vel, 3 * n) || dalloc(&c->d_mass, n) ||
	    dalloc(&c->d_soft, n) || dalloc(&c->d_temp, n) || dalloc(&c->d_rho, n) || dalloc(&c->d_ball2, n) ||
	    dalloc(&c->d_grp, n) || dalloc(&c->d_boxOf, n) || dalloc(&c->d_type, n) || dalloc(&c->d_bt, nBox) ||
	    dalloc(&c->d_bact, nBox) || dalloc(&c->d_bscat, nBox) || dalloc(&c->d_bthr, nBox) || dalloc(&c->d_bslot, (size_t)nBox * SKG_THR_SLOTS) ||
	    dalloc(&dbi, 3 * (size_t)nBox))
		return SKG_E_NOMEM;
	SKG_HIP(hipMemcpyAsync(c->d_bt, db.data(), sizeof(DevBox) * nBox, hipMemcpyHostToDevice, c->st));
	SKG_HIP(hipMemcpyAsync(dbi, binfo.data(), 12 * (size_t)nBox, hipMemcpyHostToDevice, c->st));
	k_box_assign<<<nblk(n, 256), 256, 0, c->st>>>((int)n, nBox, dbi, c->d_boxOf, c->d_type);
	SKG_HIP(hipMemcpyAsync(c->d_pos, pos, 12 * n, hipMemcpyHostToDevice, c->st));
	if (vel) SKG_HIP(hipMemcpyAsync(c->d_vel, vel, 12 * n, hipMemcpyHostToDevice, c->st));
	else SKG_HIP(hipMemsetAsync(c->d_vel, 0, 12 * n, c->st));
	SKG_HIP(hipMemcpyAsync(c->d_mass, mass, 4 * n, hipMemcpyHostToDevice, c->st));
	SKG_HIP(hipMemcpyAsync(c->d_soft, soft, 4 * n, hipMemcpyHostToDevice, c->st));
	if (temp) {
		SKG_HIP(hipMemcpyAsync(c->d_temp, temp, 4 * n, hipMemcpyHostToDevice, c->st));
		k_temp_nongas<<<nblk(n, 256), 256, 0, c->st>>>(c->d_type, (int)n, c->d_temp);
	} else SKG_HIP(hipMemsetAsync(c->d_temp, 0, 4 * n, c->st));
	SKG_HIP(hipGetLastError());
	SKG_HIP(hipStreamSynchronize(c->st));
	dfree(&dbi);
	c->stage = 1;
	return SKG_OK;
}

extern "C" int skg_load(skg_ctx *c, int nGas, int nDark, int nStar, const float *pos, const float *vel,
		const float *mass, const float *soft, const float *temp) {
	return skg_load_batch(c, 1, &nGas, &nDark, &nStar, nullptr, pos, vel, mass, soft, temp);
}

extern "C" int skg_box_counts(skg_ctx *c, int *nActive, int *nScatter) {
	if (!c) return SKG_E_PARAM;
	if (c->stage < 3) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	if (nActive) SKG_HIP(hipMemcpyAsync(nActive, c->d_bact, 4 * (size_t)c->nBox, hipMemcpyDeviceToHost, c->st));
	if (nScatter) SKG_HIP(hipMemcpyAsync(nScatter, c->d_bscat, 4 * (size_t)c->nBox, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	return SKG_OK;
}

// ScatterCriterion (kd.c:600-627) selects whole species: the mask of scatterer
// species for the context's species mix.
static int scatter_species(const skg_ctx *c) {
	const int gd = c->P.bGasAndDark, go = c->P.bGasOnly;
	switch (c->inType) {
	case SKG_DARK: return SKG_DARK;
	case SKG_GAS: case SKG_DARK | SKG_GAS: return SKG_GAS | (gd ? SKG_DARK : 0);
	case SKG_STAR: case SKG_DARK | SKG_STAR: return SKG_STAR;
	case SKG_GAS | SKG_STAR: case SKG_DARK | SKG_GAS | SKG_STAR:
		return gd ? (SKG_GAS | SKG_DARK | SKG_STAR) : (SKG_GAS | (go ? 0 : SKG_STAR));
	}
	return 0;
}

template <int KMAX, bool BATCH, int W>
static void launch_knn(skg_ctx *c, bool per, bool comp) {
	const int nb = nblk(c->nS, 64 * W);
	const float L = per ? c->P.fPeriod[0] : FLT_MAX;
#define KL(P_, C_)                                                                           \
	k_knn<KMAX, P_, C_, BATCH, W><<<nb, 64 * W, 0, c->st>>>(c->s_pos, c->s_ord, c->nS, c->ts.lo, c->ts.hi, \
			c->ts.nLeaf2, c->ts.B, L, c->P.nSmooth, c->s_inv, c->s_ball2, c->s_rhoG, c->s_heap, \
			c->s_box, c->d_bt, c->wantDlnk ? c->s_dlnkG : nullptr)
	if (per) { if (comp) KL(true, true); else KL(true, false); }
	else { if (comp) KL(false, true); else KL(false, false); }
#undef KL
}
// knnWaves: wavefronts per kNN workgroup (1 or 2; 2 only up to k = 64, where
// two heaps fit a workgroup's LDS at useful occupancy)
template <int KMAX>
static void launch_knn(skg_ctx *c, bool per, bool comp) {
	const bool two = KMAX <= 64 && c->knnWaves == 2;
	if (c->nBox > 1) {
		if (two) launch_knn<KMAX, true, (KMAX <= 64 ? 2 : 1)>(c, per, comp);
		else launch_knn<KMAX, true, 1>(c, per, comp);
	} else {
		if (two) launch_knn<KMAX, false, (KMAX <= 64 ? 2 : 1)>(c, per, comp);
		else launch_knn<KMAX, false, 1>(c, per, comp);
	}
}

static int cover_launch(skg_ctx *c, int mode, float4 *ppos, const int *plist, int np, const int *pbox,
		float *out, int stamp, int micro, double work, float *out_dl = nullptr) {
	if (np <= 0) return 0;
	const int pslot = skg_prof_begin(c);
	const int nb = nblk(np, SKG_TILE_T);
	const bool comp = c->P.accumulate == SKG_ACC_COMPENSATED;
	const float *C = c->P.fCenter, *L = c->P.fPeriod;
#define CL(M_, C_)                                                                             \
	k_cover<M_, C_><<<nb, SKG_TILE_T, 0, c->st>>>(ppos, plist, np, c->e_pos, c->e_aux, c->e_org, c->e_code, c->nE, c->te.lo, \
			c->te.hi, c->te.nLeaf2, c->te.B, out, c->e_hit, stamp, nullptr, c->nBox > 1 ? pbox : nullptr, c->d_bt, micro, \
			C[0], C[1], C[2], L[0], L[1], L[2], c->te.lstart, nullptr, nullptr, c->tlCap, nullptr, nullptr, 0, \
			nullptr, out_dl, nullptr)
#define CS(M_, C_)                                                                             \
	k_cover_sub<M_, C_, 16><<<nblk(16L * np, SKG_TILE_T), SKG_TILE_T, 0, c->st>>>(ppos, plist, np, c->e_pos, \
			c->e_aux, c->e_org, c->nE, c->te.lo, c->te.hi, c->te.nLeaf2, c->te.B, out, c->e_hit, stamp,       \
			nullptr, c->nBox > 1 ? pbox : nullptr, c->d_bt, micro, C[0], C[1], C[2], L[0], L[1], L[2], c->te.lstart, \
			nullptr, nullptr, out_dl)
	if (np <= c->subMax) {
		if (mode == 0) { if (comp) CS(0, true); else CS(0, false); }
		else { if (comp) CS(1, true); else CS(1, false); }
	} else {
		if (mode == 0) { if (comp) CL(0, true); else CL(0, false); }
		else { if (comp) CL(1, true); else CL(1, false); }
	}
#undef CS
#undef CL
	skg_prof_end(c, pslot, mode == 0 ? SKG_K_COVER_DENS : SKG_K_COVER_MOVE, work);
	SKG_HIP(hipGetLastError());
	return 0;
}

extern "C" int skg_density(skg_ctx *c, float *fBall2, float *rho, int *nExtraScat) {
	if (!c) return SKG_E_PARAM;
	if (c->stage < 1) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	int rc;
	const int k = c->P.nSmooth, B = c->P.nBucket, n = c->n, nBox = c->nBox;
	const bool per = c->P.bPeriodic != 0;
	if (per && !(c->P.fPeriod[0] == c->P.fPeriod[1] && c->P.fPeriod[1] == c->P.fPeriod[2]))
		return SKG_E_PARAM;   // -p sets one length for all axes (main.c:273-281)
	if (256 % B != 0 && nBox > 1) return SKG_E_PARAM;   // boxes start on whole leaves
	const int smask = scatter_species(c);
	c->dlnkReady = 0;
	int nS = 0;
	for (int b = 0; b < nBox; ++b) {
		const HostBox &h = c->hbox[b];
		const int nb = ((smask & SKG_GAS) ? h.nGas : 0) + ((smask & SKG_DARK) ? h.nDark : 0) +
				((smask & SKG_STAR) ? h.nStar : 0);
		if (nb > 0 && k > nb) return SKG_E_NSMOOTH;   // smooth1.c:12, per box
		nS += nb;
	}
	c->nS = nS;
	SKG_HIP(hipMemsetAsync(c->d_rho, 0, 4 * (size_t)n, c->st));
	k_fill_f<<<nblk(n, 256), 256, 0, c->st>>>(c->d_ball2, n, -2.f);
	skg_timer_start(c);
	if (nS > 0) {
		// scatterers (kdScatterActive) in (box, Morton) order
		float4 *tmp4 = nullptr;
		int *tmpo = nullptr, *tmpb = nullptr, *perm = nullptr, *flag = nullptr, *off = nullptr;
		if ((c->wantDlnk && (dalloc(&c->s_dlnkG, nS) || dalloc(&c->s_dlnkS, nS))) ||
		    dalloc(&c->s_pos, nS + SKG_EPAD) || dalloc(&c->s_ord, nS + SKG_EPAD) || dalloc(&c->s_inv, n) ||
		    dalloc(&c->s_box, nS) || dalloc(&c->s_ball2, nS) ||
		    dalloc(&c->s_rhoG, nS) || dalloc(&c->s_rhoS, nS) || sget(c, 0, &tmp4, nS) || sget(c, 1, &tmpo, nS) ||
		    sget(c, 2, &tmpb, nS) || sget(c, 3, &perm, nS) || sget(c, 4, &flag, n) || sget(c, 5, &off, n))
			return SKG_E_NOMEM;
		k_species_flags<<<nblk(n, 256), 256, 0, c->st>>>(c->d_type, n, smask, flag);
		if ((rc = skg_exclusive_scan(c, flag, off, n))) return rc;
		k_compact<<<nblk(n, 256), 256, 0, c->st>>>(nullptr, flag, off, n, tmpo);
		k_gather_scat<<<nblk(nS, 256), 256, 0, c->st>>>(c->d_pos, c->d_mass, tmpo, c->d_boxOf, nS, tmp4, tmpb);
		// scatterers along the Hilbert curve: the kNN tree's leaves and the
		// wavefronts' runs of 64 queries are compact (no Morton jumps)
		if ((rc = skg_hilbert_sort(c, tmp4, nullptr, nS, perm, nBox > 1 ? tmpb : nullptr))) return rc;
		k_permute_s<<<nblk(nS, 256), 256, 0, c->st>>>(tmp4, tmpo, tmpb, perm, nS, c->s_pos, c->s_ord, c->s_box);
		k_inverse<<<nblk(nS, 256), 256, 0, c->st>>>(c->s_ord, nS, c->s_inv);
		// the kNN tree's own leaf size (knnBucket; 0: nBucket)
		const int Bk = c->knnBucket > 0 ? c->knnBucket : B;
		if ((rc = skg_build_tree(c, &c->ts, c->s_pos, nS, Bk, nullptr, nBox > 1 ? c->s_box : nullptr))) return rc;
		skg_timer_stop(c, SKG_T_TREE);

		// kNN radii + gather half of the density
		skg_timer_start(c);
		if (dalloc(&c->s_heap, (size_t)nS * k)) return SKG_E_NOMEM;
		const bool comp = c->P.accumulate == SKG_ACC_COMPENSATED;
		const int pslot = skg_prof_begin(c);
		if (k <= 32) launch_knn<32>(c, per, comp);
		else if (k <= 64) launch_knn<64>(c, per, comp);
		else if (k <= 128) launch_knn<128>(c, per, comp);
		else launch_knn<256>(c, per, comp);
		skg_prof_end(c, pslot, SKG_K_KNN, 16.0 * nS);
		SKG_HIP(hipGetLastError());
		skg_timer_stop(c, SKG_T_KNN);

		// replicas (periodic) and the entity set S'
		skg_timer_start(c);
		unsigned *mask = nullptr;
		int *cnt = nullptr, *roff = nullptr;
		if (sget(c, 6, &mask, nS) || sget(c, 7, &cnt, nS) || sget(c, 8, &roff, nS)) return SKG_E_NOMEM;
		int nRep = 0;
		if (per) {
			const float *L = c->P.fPeriod, *C = c->P.fCenter;
			float lo[3], hi[3];
			for (int j = 0; j < 3; ++j) { lo[j] = C[j] - 0.5 * L[j]; hi[j] = C[j] + 0.5 * L[j]; }
			k_rep_mask<<<nblk(nS, 256), 256, 0, c->st>>>(c->s_pos, c->s_ball2, nS, lo[0], lo[1], lo[2],
					hi[0], hi[1], hi[2], L[0], L[1], L[2], mask, cnt);
			if ((rc = skg_exclusive_scan(c, cnt, roff, nS))) return rc;
			int last[2];
			SKG_HIP(hipMemcpyAsync(&last[0], roff + nS - 1, 4, hipMemcpyDeviceToHost, c->st));
			SKG_HIP(hipMemcpyAsync(&last[1], cnt + nS - 1, 4, hipMemcpyDeviceToHost, c->st));
			SKG_HIP(hipStreamSynchronize(c->st));
			nRep = last[0] + last[1];
		} else {
			SKG_HIP(hipMemsetAsync(mask, 0, 4 * (size_t)nS, c->st));
			SKG_HIP(hipMemsetAsync(roff, 0, 4 * (size_t)nS, c->st));
		}
		c->nRep = nRep;
		c->tfull.n = -1;                        // rebuilt on demand (full_tree), arrays kept
		const int nEu = nS + nRep;
		c->nLive = nEu;
		float4 *e4u = nullptr;
		float *eb2u = nullptr;
		int *esrcu = nullptr, *perm2 = nullptr, *ebu = nullptr, *bfirst = nullptr;
		if (sget(c, 9, &e4u, nEu) || sget(c, 10, &eb2u, nEu) || sget(c, 11, &esrcu, nEu) || sget(c, 12, &perm2, nEu) ||
		    sget(c, 13, &ebu, nEu) || sget(c, 14, &bfirst, nBox))
			return SKG_E_NOMEM;
		const float *L = c->P.fPeriod;
		k_rep_write<<<nblk(nS, 256), 256, 0, c->st>>>(c->s_pos, c->s_ball2, nS, mask, roff,
				per ? L[0] : 0.f, per ? L[1] : 0.f, per ? L[2] : 0.f, e4u, eb2u, esrcu);
		k_entity_box<<<nblk(nEu, 256), 256, 0, c->st>>>(esrcu, c->s_box, nEu, ebu);
		// Hilbert order: its key prefixes are boxes too (octree cells and their
		// halves and quarters), and consecutive leaves are face neighbours
		if ((rc = skg_hilbert_sort(c, e4u, nullptr, nEu, perm2, nBox > 1 ? ebu : nullptr))) return rc;
		// each box's entities start on a 256-entity block (and so on a leaf)
		k_fill_i<<<nblk(nBox, 256), 256, 0, c->st>>>(bfirst, nBox, -1);
		k_box_first<<<nblk(nEu, 256), 256, 0, c->st>>>(ebu, perm2, nEu, bfirst);
		std::vector<int> hf(nBox);
		SKG_HIP(hipMemcpyAsync(hf.data(), bfirst, 4 * (size_t)nBox, hipMemcpyDeviceToHost, c->st));
		SKG_HIP(hipStreamSynchronize(c->st));
		delete[] c->h_eblk;
		c->h_eblk = new (std::nothrow) int[nBox + 1];
		if (!c->h_eblk) return SKG_E_NOMEM;
		{
			std::vector<int> cntb(nBox, 0);
			// boxes without entities have first = -1; counts from the next present box
			int nextStart = nEu;
			for (int b = nBox - 1; b >= 0; --b)
				if (hf[b] >= 0) { cntb[b] = nextStart - hf[b]; nextStart = hf[b]; }
			int blk = 0;
			for (int b = 0; b < nBox; ++b) { c->h_eblk[b] = blk; blk += (cntb[b] + 255) / 256; }
			c->h_eblk[nBox] = blk;
		}
		const int nE = 256 * c->h_eblk[nBox];
		c->nE = nE;
		unsigned long long *ekey = nullptr;
		if (sget(c, 15, &ekey, nE) || dalloc(&c->e_b2, nE + SKG_EPAD) || dalloc(&c->e_pos, nE + SKG_EPAD) || dalloc(&c->e_aux, nE + SKG_EPAD) || dalloc(&c->e_rho, nE) ||
		    dalloc(&c->e_org, nE) || dalloc(&c->e_code, nE + SKG_EPAD) || dalloc(&c->e_src, nE) || dalloc(&c->e_live, nE) || dalloc(&c->e_hit, nE) ||
		    dalloc(&c->e_box, nE) || dalloc(&c->d_eblk, nBox + 1))
			return SKG_E_NOMEM;
		SKG_HIP(hipMemcpyAsync(c->d_eblk, c->h_eblk, 4 * (size_t)(nBox + 1), hipMemcpyHostToDevice, c->st));
		const int bbits = nBox > 1 ? skg_box_bits(c) : 0;
		k_entity_fill<<<nblk(nE, 256), 256, 0, c->st>>>(nE, c->d_eblk, nBox, c->e_pos, c->e_aux, c->e_org, c->e_code, c->e_rho,
				c->e_src, c->e_live, c->e_hit, c->e_b2, c->e_box, ekey, bbits);
		k_entity_setup<<<nblk(nEu, 256), 256, 0, c->st>>>(e4u, eb2u, esrcu, perm2, nEu, ebu, bfirst, c->d_eblk,
				nullptr, c->s_pos, per ? L[0] : 0.f, per ? L[1] : 0.f, per ? L[2] : 0.f, c->e_pos, c->e_aux,
				c->e_org, c->e_code, c->e_rho, c->e_src, c->e_live, c->e_hit, c->e_b2, c->k64b, ekey);
		// leaves: maximal key cells of at most B entities, never across boxes
		if ((rc = skg_build_tree_prefix(c, &c->te, c->e_pos, ekey, nE, B, 1 + bbits, c->e_b2, c->e_box))) return rc;

		// scatter half of the density: balls of others holding each scatterer
		// (and of d rho / d ln nSmooth once that has been asked for)
		if ((rc = cover_launch(c, 0, c->s_pos, nullptr, nS, c->s_box, c->s_rhoS, 0, 0, 24.0 * nS,
				c->wantDlnk ? c->s_dlnkS : nullptr))) return rc;
		c->dlnkReady = c->wantDlnk;
		float *rhoS = nullptr;
		if (sget(c, 16, &rhoS, nS)) return SKG_E_NOMEM;
		k_combine_rho<<<nblk(nS, 256), 256, 0, c->st>>>(c->s_rhoG, c->s_rhoS, c->s_ball2, c->s_ord, nS, rhoS,
				c->d_rho, c->d_ball2);
		// entity densities for ScatterCut: each entity carries its source's density
		k_entity_rho<<<nblk(nE, 256), 256, 0, c->st>>>(c->e_src, nE, rhoS, c->e_rho);
		SKG_HIP(hipGetLastError());
		skg_timer_stop(c, SKG_T_DENSITY);
	}
	if (fBall2) SKG_HIP(hipMemcpyAsync(fBall2, c->d_ball2, 4 * (size_t)n, hipMemcpyDeviceToHost, c->st));
	if (rho) SKG_HIP(hipMemcpyAsync(rho, c->d_rho, 4 * (size_t)n, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	if (nExtraScat) *nExtraScat = c->nRep;
	c->stage = 2;
	return SKG_OK;
}


extern "C" int skg_neighbors(skg_ctx *c, int *nbr, float *d2) {
	if (!c || !nbr) return SKG_E_PARAM;
	if (c->stage < 2 || !c->s_heap) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	const int k = c->P.nSmooth;
	std::vector<unsigned long long> h((size_t)c->nS * k);
	std::vector<int> ord(c->nS);
	SKG_HIP(hipMemcpyAsync(h.data(), c->s_heap, h.size() * 8, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipMemcpyAsync(ord.data(), c->s_ord, 4 * (size_t)c->nS, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	for (size_t i = 0; i < (size_t)c->n * k; ++i) { nbr[i] = -1; if (d2) d2[i] = -1.f; }
	std::vector<unsigned long long> rowv(k);
	for (int i = 0; i < c->nS; ++i) {
		unsigned long long *row = rowv.data();
		for (int s = 0; s < k; ++s) row[s] = h[(size_t)s * c->nS + i];   // the heap is stored slot-major
		std::sort(row, row + k);
		for (int s = 0; s < k; ++s) {
			nbr[(size_t)ord[i] * k + s] = (int)(unsigned)row[s];
			if (d2) d2[(size_t)ord[i] * k + s] = __builtin_bit_cast(float, (unsigned)(row[s] >> 32));
		}
	}
	return SKG_OK;
}

// ---------------------------------------------------------------------------
// Move loop.  Every step reads the active count and step stamp from device
// memory (ml), and a block of steps (lists, 5 x [gradient+move, ScatterCut],
// prune, compaction, refit) launches with grids sized for a capacity, so the
// whole block can be captured once as a hipGraph (SKG_GRAPH=1) and replayed
// while the active set shrinks; it is recaptured when the count falls below
// gcap/gShrink (fewer idle waves) or crosses the sparse-kernel size.  By
// default, and with kernel profiling on, the same launches run directly: the
// launches here are large, the host waits once per block either way, and the
// capacity-sized grids of a replayed graph cost more than they save.

static void ml_set(skg_ctx *c, int i, int v) { k_ml_set<<<1, 1, 0, c->st>>>(c->d_ml, i, v); }

// Active movers per box (for the per-box Ittr counters).
static int count_active(skg_ctx *c, long cap) {
	k_fill_i<<<nblk(c->nBox, 256), 256, 0, c->st>>>(c->d_bact, c->nBox, 0);
	if (cap > 0)
		k_box_count<<<nblk(cap, 256), 256, 0, c->st>>>(c->m_list, c->d_ml + 1, c->m_box, c->d_bact);
	SKG_HIP(hipGetLastError());
	return 0;
}

// Leaf lists for the next nsteps steps (margin: the distance a mover can
// travel before the last of them, nsteps - 1 steps of its own box's step, a
// lane at a time: step i takes its gradient at the point i moves away).
static void enqueue_lists(skg_ctx *c, long cap, int nsteps, int micro) {
	const float mabs = 1e-6f * c->P.fPeriod[0] * (c->P.bPeriodic ? 1.f : 0.f);
	const float *C = c->P.fCenter, *L = c->P.fPeriod;
#define LL(E_) k_cover_lists<E_><<<nblk(cap, SKG_TILE_T), SKG_TILE_T, 0, c->st>>>(c->m_pos, c->m_list, (int)cap,    \
			c->te.lo, c->te.hi, c->te.nLeaf2, c->nBox > 1 ? c->m_box : nullptr, c->d_bt, micro,                    \
			1.02f * (nsteps - 1),                                                                                \
			mabs, c->P.bPeriodic, C[0], C[1], C[2], L[0], L[1], L[2], c->tlCap, c->tl_list, c->tl_cnt,          \
			c->d_ml + 1, c->e_pos, c->e_aux, c->te.lstart, c->te.B, c->nE, c->m_wrap)
	if (c->useLists == 2) LL(false);
	else LL(true);
#undef LL
}

// nsteps Denmax steps over the active list (at most cap movers).
// cntLog (device, optional): the live-entity total after each step, one int
// per step (the nScatter of each Microstep line).
__global__ void k_log_count(const int *__restrict__ cnt, int *dst) { *dst = cnt[0]; }

static int enqueue_steps(skg_ctx *c, long cap, bool sub, bool lists, int nsteps, int micro, int bInitial,
		bool capturing, int *cntLog = nullptr) {
	const bool comp = c->P.accumulate == SKG_ACC_COMPENSATED;
	const float *C = c->P.fCenter, *L = c->P.fPeriod;
	const int *pbox = c->nBox > 1 ? c->m_box : nullptr;
	const int nb = nblk(c->nE, 256);
	const float mabs = 1e-6f * c->P.fPeriod[0] * (c->P.bPeriodic ? 1.f : 0.f);
	const char *ft = getenv("SKG_FUSED_THR");
	const bool fusedThr = !(ft && atoi(ft) == 0);
	// entity lists: built by the block's first step itself (k_cover_build), or
	// by a pass of their own before it (SKG_FUSED_BUILD=0, and leaf lists)
	const bool fusedBuild = lists && c->useLists != 2 && c->fusedBuild;
	// lanes within this many steps of a face get their image listed too: they
	// can cross it in time for two or more of the block's list steps (closer
	// to the margin the build walks more than the list steps save)
	const float wsteps = 1.02f * (nsteps - 2);
	if (lists && !fusedBuild) enqueue_lists(c, cap, nsteps, micro);
	for (int i = 0; i < nsteps; ++i) {
		// the initial step stamps its hits (unhit entities get density 0);
		// later steps take the threshold from the move kernel itself
		// (SKG_FUSED_THR=0: stamps and a pass over the entities every step)
		const bool stamps = bInitial || !fusedThr;
		int *ehit = stamps ? c->e_hit : nullptr;
		int *bthr = stamps ? nullptr : (int *)c->d_bslot;
		if (!stamps)
			k_thr_reset<<<nblk((long)c->nBox * SKG_THR_SLOTS, 256), 256, 0, c->st>>>(c->d_bslot,
					c->nBox * SKG_THR_SLOTS, c->d_cnt, c->d_ml);
		else
			k_bump<<<1, 1, 0, c->st>>>(c->d_ml);
		const int pslot = capturing ? -1 : skg_prof_begin(c);
#define MV(C_)                                                                                                \
		if (sub && cap <= c->s64Max) k_cover_sub<1, C_, 64><<<nblk(64L * cap, SKG_TILE_T), SKG_TILE_T, 0, c->st>>>( \
				c->m_pos, c->m_list, (int)cap, c->e_pos, c->e_aux, c->e_org, c->nE, c->te.lo, c->te.hi,          \
				c->te.nLeaf2, c->te.B, nullptr, ehit, 0, bthr, pbox, c->d_bt, micro, C[0], C[1], C[2], L[0],     \
				L[1], L[2], c->te.lstart, c->d_ml + 1, c->d_ml, nullptr);                                        \
		else if (sub) k_cover_sub<1, C_, 16><<<nblk(16L * cap, SKG_TILE_T), SKG_TILE_T, 0, c->st>>>(c->m_pos,  \
				c->m_list, (int)cap, c->e_pos, c->e_aux, c->e_org, c->nE, c->te.lo, c->te.hi, c->te.nLeaf2,      \
				c->te.B, nullptr, ehit, 0, bthr, pbox, c->d_bt, micro, C[0], C[1], C[2], L[0], L[1],             \
				L[2], c->te.lstart, c->d_ml + 1, c->d_ml, nullptr);                                              \
		else k_cover<1, C_><<<nblk(cap, SKG_TILE_T), SKG_TILE_T, 0, c->st>>>(c->m_pos, c->m_list, (int)cap,    \
				c->e_pos, c->e_aux, c->e_org, c->e_code, c->nE, c->te.lo, c->te.hi, c->te.nLeaf2, c->te.B, nullptr, \
				ehit, 0, bthr, pbox, c->d_bt, micro, C[0], C[1], C[2], L[0], L[1], L[2], c->te.lstart,           \
				lists ? c->tl_list : nullptr, lists ? c->tl_cnt : nullptr, c->tlCap, c->d_ml + 1, c->d_ml,  \
				c->useLists != 2, c->d_eblk, nullptr, lists ? c->m_wrap : nullptr)
#define MB(C_)                                                                                                \
		k_cover_build<C_><<<nblk(cap, SKG_TILE_T), SKG_TILE_T, 0, c->st>>>(c->m_pos, c->m_list, (int)cap, c->e_pos, \
				c->e_aux, c->e_org, c->e_code, c->nE, c->te.lo, c->te.hi, c->te.nLeaf2, c->te.B, ehit, bthr, pbox,   \
				c->d_bt, micro, 1.02f * (nsteps - 1), mabs, c->P.bPeriodic, wsteps, C[0], C[1], C[2], L[0], L[1], L[2], \
				c->te.lstart,                                                                                        \
				c->tlCap, c->tl_list, c->tl_cnt, c->d_ml + 1, c->d_ml, c->m_wrap)
		if (fusedBuild && i == 0) { if (comp) { MB(true); } else { MB(false); } }
		else if (comp) { MV(true); } else { MV(false); }
#undef MB
#undef MV
		if (pslot >= 0)
			skg_prof_end(c, pslot, fusedBuild && i == 0 ? SKG_K_COVER_BUILD : SKG_K_COVER_MOVE,
					24.0 * c->nActive + 24.0 * c->nLive);
		if (stamps) {
			k_scat_thr<<<nb, 256, 0, c->st>>>(c->e_live, c->e_hit, c->e_rho, c->e_aux, c->nE, c->d_ml, bInitial,
					c->part);
			k_box_min<<<c->nBox, 64, 0, c->st>>>(c->part, c->d_eblk, c->d_bthr, c->d_cnt);
		} else if (!(c->cutReady && !bInitial)) {
			k_thr_fold<<<c->nBox, SKG_THR_SLOTS, 0, c->st>>>(c->d_bslot, c->d_bthr);
		}
		if (c->cutReady && !bInitial) {
			k_cut_sorted<<<c->nBox, 256, 0, c->st>>>(c->d_bthr, c->cs_off, c->cs_pos, c->cs_rho, c->cs_perm,
					c->e_live, c->e_aux, c->d_bscat, c->d_cnt, c->e_leaf, c->leafDirty,
					stamps ? nullptr : c->d_bslot);
		} else {
			k_scat_cut<<<nb, 256, 0, c->st>>>(c->e_live, c->e_aux, c->e_rho, c->nE, c->e_box, c->d_bthr,
					(int *)c->part);
			k_box_sum<<<c->nBox, 64, 0, c->st>>>((const int *)c->part, c->d_eblk, c->d_bscat, c->d_cnt);
		}
		if (cntLog) k_log_count<<<1, 1, 0, c->st>>>(c->d_cnt, cntLog + i);
	}
	SKG_HIP(hipGetLastError());
	return 0;
}

static int refit_entities(skg_ctx *c);

// kdPruneInactive (kd.c:735-793) + compaction + the entity tree refitted to
// the live balls.
static int enqueue_prune(skg_ctx *c, long cap) {
	const float *L = c->P.fPeriod;
	const float hx = 0.5f * L[0], hy = 0.5f * L[1], hz = 0.5f * L[2];
	const int nb = nblk(cap, 256);
	k_prune<<<nb, 256, 0, c->st>>>(c->m_list, c->d_ml + 1, c->m_pos, c->m_old, hx, hy, hz, c->m_box, c->d_bt,
			c->m_flag);
	k_flag_counts<<<nb, 256, 0, c->st>>>(c->m_flag, c->d_ml + 1, c->m_bcnt);
	k_scan_counts<<<1, 1024, 0, c->st>>>(c->m_bcnt, nb, c->m_bcnt + nb, c->d_ml + 2);
	k_compact_dev<<<nb, 256, 0, c->st>>>(c->m_list, c->m_flag, c->d_ml + 1, c->m_bcnt + nb, c->m_list2);
	k_copy_list<<<nb, 256, 0, c->st>>>(c->m_list2, c->d_ml, c->m_list);
	k_ml_copy<<<1, 1, 0, c->st>>>(c->d_ml);
	int rc;
	if ((rc = count_active(c, cap))) return rc;
	return refit_entities(c);
}

static int refit_entities(skg_ctx *c) {
	DevTree *t = &c->te;
	// with the sorted cut every cut entity's leaf is flagged; before it, all
	if (c->cutReady)
		k_leaf_ball_dirty<<<nblk(t->nLeaf, 256), 256, 0, c->st>>>(c->e_aux, t->n, t->B, t->nLeaf, t->nLeaf2,
				t->lstart, t->lo, c->leafDirty);
	else
		k_leaf_ball<<<nblk(t->nLeaf2, 256), 256, 0, c->st>>>(c->e_aux, t->n, t->B, t->nLeaf, t->nLeaf2, t->lstart,
				t->lo);
	int m = t->nLeaf2 >> 1;
	for (; m > 2048; m >>= 1) k_level_ball<<<nblk(m, 256), 256, 0, c->st>>>(t->lo, m, m);
	if (m >= 1) k_level_ball_top<<<1, 1024, 0, c->st>>>(t->lo, m);
	SKG_HIP(hipGetLastError());
	return 0;
}

// Host copies of the live-entity total and the active count.
static int read_counts(skg_ctx *c, int *nLive) {
	SKG_HIP(hipMemcpyAsync(c->h_pin, c->d_ml, 12, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipMemcpyAsync(c->h_pin + 4, c->d_cnt, 4, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	c->nActive = c->h_pin[1];
	c->nLive = c->h_pin[4];
	if (nLive) *nLive = c->nLive;
	return 0;
}

static void drop_graph(skg_ctx *c) {
	if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
	if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
	c->gcap = 0;
}
void skg_drop_graph(skg_ctx *c) { drop_graph(c); }

extern "C" int skg_init_move(skg_ctx *c, int *nMove) {
	if (!c) return SKG_E_PARAM;
	if (c->stage < 2) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	int rc, n = c->n;
	int *flag = nullptr, *off = nullptr;
	if (sget(c, 4, &flag, n) || sget(c, 5, &off, n)) return SKG_E_NOMEM;
	k_mover_flags<<<nblk(n, 256), 256, 0, c->st>>>(n, c->d_type, c->inType, c->d_rho, c->d_mass,
			c->d_temp, c->P.fDensMin, c->P.fTempMax, c->P.fMassMax, c->P.bGasAndDark, c->P.bGasOnly, flag);
	if ((rc = skg_exclusive_scan(c, flag, off, n))) return rc;
	int last[2];
	SKG_HIP(hipMemcpyAsync(&last[0], off + n - 1, 4, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipMemcpyAsync(&last[1], flag + n - 1, 4, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	const int nm = last[0] + last[1];
	c->nMove = nm;
	c->cutReady = false;
	drop_graph(c);
	const size_t nbm = (size_t)nblk(std::max(nm, 1), 256);
	if (dalloc(&c->m_ord, nm) || dalloc(&c->m_pos, nm) || dalloc(&c->m_old, nm) || dalloc(&c->m_list, nm) ||
	    dalloc(&c->m_wrap, nm) ||
	    dalloc(&c->m_list2, nm) || dalloc(&c->m_box, nm) || dalloc(&c->m_flag, nm) || dalloc(&c->m_bcnt, 2 * nbm))
		return SKG_E_NOMEM;
	if (!c->d_ml && dalloc(&c->d_ml, 4)) return SKG_E_NOMEM;
	if (!c->h_pin && hipHostMalloc((void **)&c->h_pin, 32, hipHostMallocDefault) != hipSuccess) return SKG_E_NOMEM;
	if (c->useLists && nm > c->subMax) {
		const size_t nt = (size_t)nblk(nm, 64);
		// Entity lists: a tile whose list outgrows tlCap walks the tree at
		// every step of its block, at about three times the cost of scanning a
		// list.  On 256 boxes of up to 2^19 points, 11% of the tiles' lists
		// pass 1024 entries and almost none 8192, so the capacity is 8192
		// within a budget of 16 GiB (and an eighth of the free device memory),
		// and at least 1024.
		if (c->useLists == 1 && !c->tlCapFixed) {
			size_t fr = 0, tot = 0;
			if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0;
			const size_t budget = std::min<size_t>((size_t)16 << 30, fr / 8);
			c->tlCap = (int)std::max<size_t>(1024, std::min<size_t>(8192, budget / (4 * nt)) & ~(size_t)63);
		}
		if (dalloc(&c->tl_list, nt * c->tlCap) || dalloc(&c->tl_cnt, nt)) return SKG_E_NOMEM;
	}
	if (skg_ensure_part(c, (size_t)nblk(std::max(c->nE, 1), 256))) return SKG_E_NOMEM;
	k_compact<<<nblk(n, 256), 256, 0, c->st>>>(nullptr, flag, off, n, c->m_ord);
	if (c->nE > 0) k_aux_rho<<<nblk(c->nE, 256), 256, 0, c->st>>>(c->e_aux, c->e_rho, c->nE);
	k_mover_init<<<nblk(nm, 256), 256, 0, c->st>>>(c->m_ord, nm, c->d_pos, c->d_boxOf, c->m_pos, c->m_old, c->m_box);
	c->nActive = nm;
	if (nm > 0 && (rc = mover_sort(c, nm)))
		return rc;
	ml_set(c, 0, 0);
	ml_set(c, 1, nm);
	if ((rc = count_active(c, nm))) return rc;
	SKG_HIP(hipStreamSynchronize(c->st));
	if (nMove) *nMove = nm;
	c->stage = 3;
	return SKG_OK;
}

// The density-sorted entity order for the remaining cuts (SKG_SORTED_CUT=0: a
// pass over every entity at every step instead).
static int build_sorted_cut(skg_ctx *c) {
	c->cutReady = false;
	const char *e = getenv("SKG_SORTED_CUT");
	if (e && atoi(e) == 0) return 0;
	const int ne = c->nE, nBox = c->nBox;
	unsigned long long *k0 = nullptr, *k1 = nullptr;
	int *v0 = nullptr, *v1 = nullptr, rc;
	const DevTree *t = &c->te;
	auto done = [&](int status) { return status; };
	if (sget(c, 6, &k0, ne) || sget(c, 7, &k1, ne) || sget(c, 8, &v0, ne) || sget(c, 9, &v1, ne) || dalloc(&c->cs_perm, ne) ||
	    dalloc(&c->cs_rho, ne) || dalloc(&c->cs_off, nBox + 1) || dalloc(&c->cs_pos, nBox) ||
	    dalloc(&c->e_leaf, ne) || dalloc(&c->leafDirty, t->nLeaf2))
		return done(SKG_E_NOMEM);
	k_entity_leaf<<<nblk(t->nLeaf2, 256), 256, 0, c->st>>>(t->n, t->B, t->nLeaf, t->nLeaf2, t->lstart, c->e_leaf,
			c->leafDirty);
	k_cut_keys<<<nblk(ne, 256), 256, 0, c->st>>>(c->e_live, c->e_box, c->e_rho, ne, k0, v0);
	if ((rc = skg_sort_pairs(c, k0, k1, v0, v1, ne, 64))) return done(rc);
	k_cut_bounds<<<nblk(std::max(ne, nBox + 1), 256), 256, 0, c->st>>>(k1, ne, nBox, c->cs_off, c->cs_pos,
			c->cs_rho, v1, c->cs_perm);
	hipError_t he = hipGetLastError();
	if (he == hipSuccess) he = hipStreamSynchronize(c->st);
	if (he != hipSuccess) { skg_note_hip(he, "build_sorted_cut", __FILE__, __LINE__); return done(SKG_E_HIP); }
	c->cutReady = true;
	return done(0);
}

extern "C" int skg_move_first(skg_ctx *c, int *nActive, int *nScatter) {
	if (!c) return SKG_E_PARAM;
	if (c->stage != 3) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	int rc, ns = c->nE;
	skg_timer_start(c);
	const int bInit = (c->inType == SKG_DARK) || c->P.bForceInitialCut;   // kd.c:696-699
	if (c->nActive > 0 && c->nE > 0) {
		if ((rc = enqueue_steps(c, c->nActive, c->nActive <= c->subMax, false, 1, 0, bInit, false))) return rc;
		if ((rc = refit_entities(c))) return rc;
		if ((rc = read_counts(c, &ns))) return rc;
		if ((rc = build_sorted_cut(c))) return rc;
	}
	skg_timer_stop(c, SKG_T_MOVE);
	if (nActive) *nActive = c->nActive;
	if (nScatter) *nScatter = ns;
	c->stage = 4;
	return SKG_OK;
}

extern "C" int skg_move_block(skg_ctx *c, int *nActive, int *nScatter) {
	if (!c) return SKG_E_PARAM;
	if (c->stage != 4) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	int rc, ns = c->nE;
	skg_timer_start(c);
	if (c->nActive > 0 && c->nE > 0) {
		const long na = c->nActive;
		const bool sub = na <= c->subMax;
		const bool lists = c->useLists && !sub && c->tl_list;
		if (c->useGraph && !c->prof) {
			if (!c->gexec || na * c->gShrink < c->gcap || sub != (c->gsub != 0)) {
				const auto t0 = std::chrono::steady_clock::now();
				drop_graph(c);
				hipGraph_t g = nullptr;
				SKG_HIP(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
				int e1 = enqueue_steps(c, na, sub, lists, 5, 0, 0, true);
				int e2 = e1 ? e1 : enqueue_prune(c, na);
				hipError_t ee = hipStreamEndCapture(c->st, &g);
				if (e1 || e2) { if (g) (void)hipGraphDestroy(g); return e1 ? e1 : e2; }
				SKG_HIP(ee);
				c->graph = g;
				SKG_HIP(hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
				c->gcap = na; c->gsub = sub;
				c->tms[SKG_T_N - 1] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
				c->nCaptures++;
			}
			SKG_HIP(hipGraphLaunch(c->gexec, c->st));
		} else {
			if ((rc = enqueue_steps(c, na, sub, lists, 5, 0, 0, false))) return rc;
			if ((rc = enqueue_prune(c, na))) return rc;
		}
		// nScatter of the Ittr line: the live count after the last step (the
		// prune does not change it)
		if ((rc = read_counts(c, &ns))) return rc;
		// Pruning thins the active list in its old order, so runs of 64 active
		// movers spread out; once a tenth are gone, re-sort the rest by where
		// they are now (per-mover sums do not depend on the order).  At most
		// s64Max movers run a wavefront each, where the order does not matter.
		if (c->nActive > c->s64Max && (double)c->nActive < mover_resort_frac() * (double)c->lastSort) {
			if ((rc = mover_resort(c))) return rc;
		}
	} else {
		c->nActive = 0;
	}
	if (c->nActive == 0) drop_graph(c);
	skg_timer_stop(c, SKG_T_MOVE);
	if (nActive) *nActive = c->nActive;
	if (nScatter) *nScatter = ns;
	return SKG_OK;
}

extern "C" int skg_microsteps(skg_ctx *c, int *nScatter) {
	if (!c) return SKG_E_PARAM;
	if (c->stage != 5) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	int rc, ns[5] = {c->nE, c->nE, c->nE, c->nE, c->nE};
	int *log = c->d_cnt + 32;                               // 5 ints: the count after each micro-step
	drop_graph(c);
	skg_timer_start(c);
	c->nActive = c->nMove;                                  // kdReactivateMove (kd.c:796-799)
	if (c->nMove > 0 && (rc = mover_sort(c, c->nMove))) return rc;
	ml_set(c, 1, c->nMove);
	if ((rc = count_active(c, c->nMove))) return rc;
	if (c->nMove > 0 && c->nE > 0) {
		const bool sub = c->nMove <= c->subMax;
		// MICRO_STEP*fStep (main.c:435)
		if ((rc = enqueue_steps(c, c->nMove, sub, c->useLists && !sub && c->tl_list, 5, 1, 0, false, log)))
			return rc;
		SKG_HIP(hipMemcpyAsync(ns, log, sizeof(ns), hipMemcpyDeviceToHost, c->st));
		if ((rc = read_counts(c, nullptr))) return rc;
	}
	SKG_HIP(hipStreamSynchronize(c->st));
	skg_timer_stop(c, SKG_T_MOVE);
	if (nScatter) memcpy(nScatter, ns, sizeof(ns));
	c->stage = 6;
	return SKG_OK;
}

extern "C" int skg_get_moved(skg_ctx *c, float *r) {
	if (!c || !r) return SKG_E_PARAM;
	if (c->stage < 3) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	float *d = nullptr;
	if (sget(c, 0, &d, 3 * (size_t)c->n)) return SKG_E_NOMEM;
	k_fill_f<<<nblk(3 * c->n, 256), 256, 0, c->st>>>(d, 3 * c->n, NAN);
	if (c->nMove > 0) k_scatter_moved<<<nblk(c->nMove, 256), 256, 0, c->st>>>(c->m_ord, c->nMove, c->m_pos, d);
	SKG_HIP(hipMemcpyAsync(r, d, 12 * (size_t)c->n, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	return SKG_OK;
}

extern "C" int skg_field_gradient(skg_ctx *c, int nq, const float *qpos, float *acc) {
	if (!c || nq < 0 || (nq && (!qpos || !acc))) return SKG_E_PARAM;
	if (c->stage < 2) return SKG_E_STATE;
	if (c->nBox > 1) return SKG_E_PARAM;   // a point does not name its box
	if (nq == 0) return SKG_OK;
	(void)hipSetDevice(c->dev);
	float4 *q4 = nullptr;
	float *out = nullptr;
	if (sget(c, 0, &q4, nq) || sget(c, 1, &out, 3 * (size_t)nq)) return SKG_E_NOMEM;
	std::vector<float4> h(nq);
	for (int i = 0; i < nq; ++i) h[i] = make_float4(qpos[3 * i], qpos[3 * i + 1], qpos[3 * i + 2], 0.f);
	SKG_HIP(hipMemcpyAsync(q4, h.data(), 16 * (size_t)nq, hipMemcpyHostToDevice, c->st));
	int rc = skg_gradient_points(c, q4, nq, out);
	if (rc) return rc;
	SKG_HIP(hipMemcpyAsync(acc, out, 12 * (size_t)nq, hipMemcpyDeviceToHost, c->st));
	SKG_HIP(hipStreamSynchronize(c->st));
	return SKG_OK;
}

// Calibration copy: a plain 16-byte-per-lane copy, grid-stride.
__global__ __launch_bounds__(256) void k_copy_calib(const float4 *__restrict__ a, float4 *__restrict__ b, long n) {
	for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

extern "C" int skg_device_copy(skg_ctx *c, long long bytes, int reps) {
	if (!c || bytes < 16 || reps < 1) return SKG_E_PARAM;
	(void)hipSetDevice(c->dev);
	const long n = (long)(bytes / 16);
	float4 *a = nullptr, *b = nullptr;
	auto done = [&](int status) { dfree(&a); dfree(&b); return status; };
	if (dalloc(&a, n) || dalloc(&b, n)) return done(SKG_E_NOMEM);
	k_fill_f<<<nblk(4 * n, 256), 256, 0, c->st>>>((float *)a, (int)std::min<long>(4 * n, INT_MAX), 1.0f);
	for (int r = 0; r < reps; ++r) k_copy_calib<<<4096, 256, 0, c->st>>>(a, b, n);
	hipError_t e = hipGetLastError();
	if (e == hipSuccess) e = hipStreamSynchronize(c->st);
	if (e != hipSuccess) { skg_note_hip(e, "k_copy_calib", __FILE__, __LINE__); return done(SKG_E_HIP); }
	return done(SKG_OK);
}

extern "C" int skg_stage_times(skg_ctx *c, double *ms) {
	if (!c || !ms) return SKG_E_PARAM;
	for (int i = 0; i < SKG_T_N; ++i) ms[i] = c->tms[i];
	return SKG_OK;
}

// =====================================================================
// Autotuner: time launch shapes on the loaded particles and keep the fastest.
// =====================================================================

static float elapsed_ms(skg_ctx *c) {
	float ms = 0.f;
	(void)hipEventRecord(c->e1, c->st);
	(void)hipEventSynchronize(c->e1);
	(void)hipEventElapsedTime(&ms, c->e0, c->e1);
	return ms;
}

extern "C" int skg_autotune(skg_ctx *c, int what, int *nBucket, long long *subMax, int *potTile) {
	if (!c) return SKG_E_PARAM;
	if (c->stage < 1) return SKG_E_STATE;
	(void)hipSetDevice(c->dev);
	int rc, na, ns;
	const int prof = c->prof;
	c->prof = 0;
	// 1. leaf size: density + the whole move loop (its dense early steps and
	// its sparse tail favour different leaves; a prefix of the loop has
	// picked a leaf 2% slower over the whole of it)
	if (what & SKG_TUNE_BUCKET) {
		float best = FLT_MAX;
		int bestB = c->P.nBucket;
		// host clock after a stream sync: the entry points below restart the
		// context's own stage events, so those cannot bracket them
		auto now_ms = [&]() {
			(void)hipStreamSynchronize(c->st);
			return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
					.count();
		};
		for (int B : {8, 16, 32, 64}) {
			c->P.nBucket = B;
			const double t0 = now_ms();
			if ((rc = skg_density(c, nullptr, nullptr, nullptr))) return rc;
			if ((rc = skg_init_move(c, nullptr))) return rc;
			if ((rc = skg_move_first(c, &na, &ns))) return rc;
			for (int blk = 0; na > 0 && blk < 1000; ++blk)
				if ((rc = skg_move_block(c, &na, &ns))) return rc;
			const float cost = (float)(now_ms() - t0);
			if (cost < best) { best = cost; bestB = B; }
		}
		c->P.nBucket = bestB;
	}
	// 2. mover mapping: the active count at and below which a block runs 16
	// lanes per mover instead of one lane per mover (with its tiles' lists),
	// timed over the whole move loop for each candidate (a block's cost with
	// lists includes its list build, which a single step on a subset misses)
	if (what & SKG_TUNE_MAPPING) {
		auto now_ms = [&]() {
			(void)hipStreamSynchronize(c->st);
			return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
					.count();
		};
		double best = 1e300;
		long bestM = c->subMax;
		for (long m : {1L << 14, 1L << 16, 1L << 18, 1L << 20}) {
			c->subMax = m;
			if ((rc = skg_density(c, nullptr, nullptr, nullptr))) return rc;   // the cuts restart from here
			if ((rc = skg_init_move(c, nullptr))) return rc;
			const double t0 = now_ms();
			if ((rc = skg_move_first(c, &na, &ns))) return rc;
			for (int blk = 0; na > 0 && blk < 1000; ++blk)
				if ((rc = skg_move_block(c, &na, &ns))) return rc;
			const double cost = now_ms() - t0;
			if (cost < best) { best = cost; bestM = m; }
		}
		c->subMax = bestM;
	}
	// 3. unbinding potential tile: timed on the groups of the next skg_unbind
	if (what & SKG_TUNE_POT_TILE) c->tunePot = 1;
	// 4. kNN shape: the kNN tree's leaf size and 1 or 2 wavefronts per kNN
	// workgroup, timed by the tree and kNN stages of skg_density (best of two
	// runs each, after one untimed run)
	if (what & SKG_TUNE_KNN) {
		double best = 1e300;
		int bestB = c->knnBucket, bestW = c->knnWaves;
		if ((rc = skg_density(c, nullptr, nullptr, nullptr))) return rc;
		for (int B : {8, 16, 32, 64})
			for (int W = 1; W <= (c->P.nSmooth <= 64 ? 2 : 1); ++W) {
				c->knnBucket = B;
				c->knnWaves = W;
				double t = 1e300;
				for (int r = 0; r < 2; ++r) {
					const double k0 = c->tms[SKG_T_TREE] + c->tms[SKG_T_KNN];
					if ((rc = skg_density(c, nullptr, nullptr, nullptr))) return rc;
					t = std::min(t, c->tms[SKG_T_TREE] + c->tms[SKG_T_KNN] - k0);
				}
				if (t < best) { best = t; bestB = B; bestW = W; }
			}
		c->knnBucket = bestB;
		c->knnWaves = bestW;
	}
	c->prof = prof;
	// the caller reruns the path from skg_density with the chosen shapes
	for (int i = 0; i < SKG_T_N; ++i) c->tms[i] = 0;
	c->stage = 1;
	if (nBucket) *nBucket = c->P.nBucket;
	if (subMax) *subMax = c->subMax;
	if (potTile) *potTile = c->potTile;
	return SKG_OK;
}

extern "C" int skg_knn_shape(skg_ctx *c, int *knnBucket, int *knnWaves) {
	if (!c) return SKG_E_PARAM;
	if (knnBucket) *knnBucket = c->knnBucket > 0 ? c->knnBucket : c->P.nBucket;
	if (knnWaves) *knnWaves = c->knnWaves;
	return SKG_OK;
}
