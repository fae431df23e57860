// skg_common.h -- shared state and device helpers of libskg (MI355X / gfx950).
//
// Layout in HBM (structure of arrays, 16-byte records where a kernel reads
// several fields of one item):
//   particles by iOrder : pos f32x3, vel f32x3, mass, soft, temp, type
//   scatterers (S)       : Morton-sorted float4 (x,y,z,m), iOrder, fBall2, rho parts
//   entities (S')        : S plus periodic replicas, Morton-sorted;
//                          float4 (x,y,z,m) + float4 (fBall2 live, ih2, grad norm, density norm)
//   movers               : persistent float4 pos / old pos by mover index, plus the
//                          active list (mover indices) re-sorted by Morton key each prune
//   trees                : complete binary tree over leaves of nBucket consecutive
//                          sorted items, heap-indexed, float4 lo (w: max fBall2) / hi.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include "../../include/skg.h"

#define SKG_WAVE 64
// Wave-per-tile kernels run four independent wavefronts per workgroup: a
// one-wave workgroup would cap residency at the per-CU workgroup limit.
#define SKG_TILE_T 256
// records allocated past the last entity of e_pos / e_aux (reads in blocks of 4)
#define SKG_EPAD 4
// scratch slots per context (skg_slot)
#define SKG_NSLOT 24

// Per-box constants of a batched context (box b = particles [off, off+n) by
// iOrder).  A single-box context has one entry holding the parameter block's
// values.
struct DevBox {
	int off, n;
	float fStep, fMicro;     // Denmax step and the micro-step 0.1*fStep (main.c:345, 435)
	float cvg2, tau2;        // fCvg^2 (kd.c:762), tau^2 (kd.c:875)
	float scoop2, pad;       // fScoop^2 (grav.c:63-135)
};
struct HostBox {
	int off, n, nGas, nDark, nStar;
	float tau, cvg, step, scoop;
};

struct DevTree {
	int n = 0;         // items
	int B = 16;        // bucket (leaf) size
	int nLeaf = 0;     // ceil(n/B)
	int nLeaf2 = 0;    // next power of two >= nLeaf (>= 1)
	float4 *lo = nullptr;   // [2*nLeaf2], index 0 unused; lo.w = max live fBall2 below
	float4 *hi = nullptr;   // hi.w: bits of the box range [b0, b1] below (b0 | b1 << 16)
	int *lstart = nullptr;  // [nLeaf2+1] item offset of each leaf (none: leaf l = [lB, lB+B))
};

struct skg_ctx {
	skg_params P;
	int dev = 0;
	hipStream_t st = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	// a second stream (default priority) for work that overlaps the main
	// stream's inside one entry point, forked and joined with evFork / evJoin
	hipStream_t st2 = nullptr;
	hipEvent_t evFork = nullptr, evJoin = nullptr;
	double tms[SKG_T_N] = {0};
	int stage = 0;            // 0 created, 1 loaded, 2 density, 3 movers, 4 moving done, 5 fof, 6 micro, 7 centers, 8 unbound, 9 final
	int n = 0, nGas = 0, nDark = 0, nStar = 0, inType = 0;

	// boxes (one unless loaded with skg_load_batch)
	int nBox = 1;
	HostBox *hbox = nullptr;
	DevBox *d_bt = nullptr;
	int *d_boxOf = nullptr;      // box by iOrder
	unsigned char *d_type = nullptr;   // SKG_GAS / SKG_DARK / SKG_STAR by iOrder
	int *h_eblk = nullptr;       // [nBox+1] first 256-entity block of each box (host)
	int *d_eblk = nullptr;       // same, on the device
	int *d_bact = nullptr, *d_bscat = nullptr;   // per-box active movers / live entities
	float *d_bthr = nullptr;     // per-box ScatterCut threshold
	float *d_bslot = nullptr;    // per-box threshold slots the move step's atomics spread over

	// by iOrder
	float *d_pos = nullptr, *d_vel = nullptr, *d_mass = nullptr, *d_soft = nullptr, *d_temp = nullptr;
	float *d_rho = nullptr, *d_ball2 = nullptr;
	int *d_grp = nullptr;

	// scatterers
	int nS = 0;
	float4 *s_pos = nullptr;
	int *s_ord = nullptr;
	int *s_inv = nullptr;        // iOrder -> scatterer index (scatterers only)
	float *s_ball2 = nullptr, *s_rhoG = nullptr, *s_rhoS = nullptr;
	// d rho / d ln nSmooth halves formed by the density kernels themselves once
	// skg_density_dlnk has been asked for (wantDlnk); dlnkReady: they belong
	// to the last skg_density
	float *s_dlnkG = nullptr, *s_dlnkS = nullptr;
	int wantDlnk = 0, dlnkReady = 0;
	int *s_box = nullptr;
	unsigned long long *s_heap = nullptr;   // k x n_S keys of the last kNN, slot-major (slot s of query i at s n_S + i)
	DevTree ts;

	// entities S' = S + replicas
	int nE = 0, nRep = 0, nLive = 0;
	float4 *e_pos = nullptr, *e_aux = nullptr;
	float4 *e_org = nullptr;     // source scatterer position; w = its image code (as e_code)
	int *e_code = nullptr;       // image code 9(ix+1) + 3(iy+1) + (iz+1) (13: the source itself)
	float *e_b2 = nullptr;       // fBall2 as built (all entities live)
	float *e_rho = nullptr;      // effective density (0 once zeroed by the initial cut)
	int *e_src = nullptr;        // scatterer (sorted S) index
	int *e_live = nullptr;       // 1 active, 0 cut
	int *e_hit = nullptr;        // step stamp of the last hit
	int *e_box = nullptr;        // box; each box's range starts on a 256-entity boundary,
	                             // filled up with dead entities (fBall2 -1, position NaN)
	DevTree te;
	DevTree tfull;               // entity tree with every radius as built (adjoint, field gradient)
	DevTree tfof;                // FOF tree (skg_fof, skg_fof_points)

	// movers
	int nMove = 0, nActive = 0, step = 0;
	float4 *m_pos = nullptr, *m_old = nullptr;
	int *m_ord = nullptr;
	int *m_box = nullptr;
	int *m_list = nullptr, *m_list2 = nullptr;
	float *d_thr = nullptr;      // [2]: threshold scratch
	int *d_cnt = nullptr;        // small counters

	// groups
	int nGroup = 0, nBigGroups = 0, nCoopFallbacks = 0;
	long long nFullPasses = 0;
	struct GroupState *gs = nullptr;

	// kernel timing: (class, start, stop, work) per instrumented launch
	int prof = 0;
	struct ProfRec { int k; hipEvent_t a, b; double work; };
	ProfRec *prec = nullptr; int nrec = 0, caprec = 0;
	double kms[SKG_K_N] = {0}, kwork[SKG_K_N] = {0};
	long long klaunch[SKG_K_N] = {0};

	// scratch
	void *tmp = nullptr; size_t tmpBytes = 0;
	unsigned long long *k64a = nullptr, *k64b = nullptr; size_t kcap = 0;
	int *v32a = nullptr, *v32b = nullptr;
	size_t vcap = 0;
	float *part = nullptr; size_t partCap = 0;   // per-block reduction partials
	void *slot[SKG_NSLOT] = {nullptr};           // stage temporaries kept between calls (skg_slot)
	size_t slotBytes[SKG_NSLOT] = {0};

	// launch shapes (set at create; SKG_COVER_SUB_MAX overrides): point sets of
	// at most subMax points use the sub-wave cover kernel
	long subMax = 0;
	int knnBucket = 0, knnWaves = 1;             // kNN tree leaf size (0: nBucket), wavefronts per kNN workgroup
	int fusedBuild = 1;                          // entity lists built by each block's first step
	long s64Max = 0;            // at most this many active movers: 64 lanes (a wavefront) per mover
	// per-tile lists of the current 5-step block: entities (SKG_TILE_LISTS=1,
	// default), leaves (2) or none (0: every step walks the tree); tlCap
	// entries per tile (a multiple of 4; for entity lists set from the mover
	// count when the move loop starts, unless tlCapFixed)
	int useLists = 1, tlCap = 1024, tlCapFixed = 0;
	int *tl_list = nullptr, *tl_cnt = nullptr;
	// device-resident move loop: ml[0] step stamp, ml[1] active, ml[2] next
	int *d_ml = nullptr;
	int *m_flag = nullptr, *m_bcnt = nullptr;   // prune flags; block counts then offsets
	int *m_wrap = nullptr;                       // per mover: wrapped across a periodic face since its tile's list was built
	long lastSort = 0;                          // active count at the last mover sort
	// ScatterCut over density-sorted entities (after the initial step): each
	// box's live entities in ascending density, its first still-live position
	int *cs_perm = nullptr, *cs_off = nullptr, *cs_pos = nullptr;
	float *cs_rho = nullptr;
	bool cutReady = false;
	int *e_leaf = nullptr;      // entity -> leaf of the entity tree (with the sorted cut)
	int *leafDirty = nullptr;   // leaves holding an entity cut since the last refit
	int *h_pin = nullptr;                        // pinned read-back
	int useGraph = 0;                            // SKG_GRAPH=1: replay each block as a captured graph
	hipGraph_t graph = nullptr;
	hipGraphExec_t gexec = nullptr;
	long gcap = 0;
	int gsub = 0, gShrink = 2, nCaptures = 0;   // recapture when the count falls below gcap/gShrink
	// unbinding potential: workgroup / LDS tile size; tunePot: time 64/128/256
	// on the next skg_unbind's groups and keep the fastest
	int potTile = 256, tunePot = 0;
};

// ---------------------------------------------------------------- errors
// A failing HIP call returns SKG_E_HIP; the call, its error and the place are
// kept per host thread for skg_last_error().
void skg_note_hip(hipError_t e, const char *call, const char *file, int line);
#define SKG_HIP(x) do { hipError_t _e = (x); if (_e != hipSuccess) { skg_note_hip(_e, #x, __FILE__, __LINE__); \
		return SKG_E_HIP; } } while (0)

// ---------------------------------------------------------------- wave helpers
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
	return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
	return v;
}
__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
	return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
	return v;
}

// Workgroups are dealt to the 8 XCDs round-robin (block b to XCD b % 8), and
// each XCD has its own L2.  Tile-ordered kernels remap the block index so that
// XCD x walks one contiguous range of tiles: neighbouring tiles (Morton
// neighbours, sharing most of their tree nodes and entities) then hit the
// same L2.  A bijection of [0, gridDim.x).
#define SKG_NXCD 8
__device__ __forceinline__ int xcd_tile(int b, int G) {
	const int q = G / SKG_NXCD, r = G % SKG_NXCD, x = b % SKG_NXCD;
	return x * q + min(x, r) + b / SKG_NXCD;
}

// Item range of leaf l (uniform leaves of B items, or the leaf offsets).
__device__ __forceinline__ int2 leaf_range(const int *lstart, int l, int B, int n) {
	if (lstart) return make_int2(lstart[l], lstart[l + 1]);
	const int s = l * B;
	return make_int2(s, min(n, s + B));
}

// ---------------------------------------------------------------- box ranges
// A tree node's hi.w holds the range [b0, b1] of boxes below it as raw bits;
// an empty node has b0 > b1.  Single-box trees hold [0, 0].
#define SKG_MAX_BOX 32767
__device__ __forceinline__ float brange_pack(int b0, int b1) {
	return __uint_as_float((unsigned)b0 | ((unsigned)b1 << 16));
}
__device__ __forceinline__ bool brange_has(float w, int b) {
	const unsigned r = __float_as_uint(w);
	return b >= (int)(r & 0xFFFFu) && b <= (int)(r >> 16);
}
__device__ __forceinline__ bool brange_meets(float w, int lo, int hi) {
	const unsigned r = __float_as_uint(w);
	return max(lo, (int)(r & 0xFFFFu)) <= min(hi, (int)(r >> 16));
}
__device__ __forceinline__ bool brange_is(float w, int b) {
	return __float_as_uint(w) == ((unsigned)b | ((unsigned)b << 16));
}

// ---------------------------------------------------------------- exact fp32 ops
// The path's distance arithmetic must match the reference's (no contraction);
// all HIP sources are built with -ffp-contract=off, and these make it explicit.
__device__ __forceinline__ float fsub(float a, float b) { return __fsub_rn(a, b); }
__device__ __forceinline__ float fmul(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ float fadd(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ float sq3(float dx, float dy, float dz) {
	return fadd(fadd(fmul(dx, dx), fmul(dy, dy)), fmul(dz, dz));
}

// Compensated (two-sum) accumulator: the group sums (centres, potentials).
struct Acc {
	float s, c;
	__device__ __forceinline__ void init() { s = 0.f; c = 0.f; }
	__device__ __forceinline__ void add(float x) {
		float t = fadd(s, x);
		float bp = fsub(t, s);
		float e = fadd(fsub(s, fsub(t, bp)), fsub(x, bp));
		c = fadd(c, e);
		s = t;
	}
	__device__ __forceinline__ float val() const { return fadd(s, c); }
	__device__ __forceinline__ float cc() const { return c; }
	// fold in another partial (its sum and its carried error)
	__device__ __forceinline__ void merge(float os, float oc) { add(os); c = fadd(c, oc); }
};
// The per-point sums of the cover kernels.  Sum<true> (the compensated mode)
// carries the sum of the fp32 terms in FP64: a term converts exactly and each
// add rounds once at 2^-53 relative, so the fp32 result is the terms' exact sum
// rounded once, up to n 2^-53 of the sum of their magnitudes (a two-sum
// accumulator leaves (n u)^2 of it, u = 2^-24).  It costs two full-rate
// instructions per term (convert, FP64 add) against a two-sum's seven.
// Sum<false> is the plain fp32 sum.  xor_merge folds in the partial of the
// lane o positions away (butterfly reductions over sub-wave groups).
template <bool COMP> struct Sum;
template <> struct Sum<true> {
	double s;
	__device__ __forceinline__ void init() { s = 0.0; }
	__device__ __forceinline__ void add(float x) { s = __dadd_rn(s, (double)x); }
	__device__ __forceinline__ float val() const { return (float)s; }
	__device__ __forceinline__ void xor_merge(int o) { s = __dadd_rn(s, __shfl_xor(s, o, 64)); }
};
template <> struct Sum<false> {
	float s;
	__device__ __forceinline__ void init() { s = 0.f; }
	__device__ __forceinline__ void add(float x) { s = fadd(s, x); }
	__device__ __forceinline__ float val() const { return s; }
	__device__ __forceinline__ void xor_merge(int o) { s = fadd(s, __shfl_xor(s, o, 64)); }
};

// ---------------------------------------------------------------- box distances
// All lower bounds are computed from the same fp32 operands as the per-pair
// deltas, so by monotonicity of rounding they never exceed a member's d2.
__device__ __forceinline__ float gap1(float alo, float ahi, float blo, float bhi) {
	float g = fmaxf(fsub(alo, bhi), fsub(blo, ahi));
	return fmaxf(g, 0.f);
}
template <bool PER>
__device__ __forceinline__ float gap_axis(float nlo, float nhi, float blo, float bhi, float L) {
	float g = gap1(nlo, nhi, blo, bhi);
	if (PER) {
		g = fminf(g, gap1(nlo, nhi, fadd(blo, L), fadd(bhi, L)));
		g = fminf(g, gap1(nlo, nhi, fsub(blo, L), fsub(bhi, L)));
	}
	return g;
}
template <bool PER>
__device__ __forceinline__ float box_box_d2(float4 nlo, float4 nhi, float3 blo, float3 bhi, float L) {
	float gx = gap_axis<PER>(nlo.x, nhi.x, blo.x, bhi.x, L);
	float gy = gap_axis<PER>(nlo.y, nhi.y, blo.y, bhi.y, L);
	float gz = gap_axis<PER>(nlo.z, nhi.z, blo.z, bhi.z, L);
	return sq3(gx, gy, gz);
}

__device__ __forceinline__ float point_box_d2_np(float4 lo, float4 hi, float4 p) {
	return sq3(gap1(lo.x, hi.x, p.x, p.x), gap1(lo.y, hi.y, p.y, p.y), gap1(lo.z, hi.z, p.z, p.z));
}

// Query-side nearest-image delta: d = q - p, re-evaluated from q -/+ L rounded
// first when |d| exceeds half a period (the reference's INTERSECT shift).
template <bool PER>
__device__ __forceinline__ float img_delta(float q, float qm, float qp, float p, float hL) {
	const float d = fsub(q, p);
	if (!PER) return d;
	// all three images, then a select (no divergent branches per candidate)
	const float dm = fsub(qm, p), dp = fsub(qp, p);
	return d > hL ? dm : (d < -hL ? dp : d);
}

// Periodic delta (s + shift) - p accurate to one rounding of the result: the
// rounding error of s - p is carried by a two-sum, and adding the shift to
// s - p (|s - p| close to |shift|) is exact.  The reference instead rounds
// s + shift first (ulp(L/2) absolute), which near the box faces costs ~1e-5
// relative in pair separations of ~1e-3.
__device__ __forceinline__ float exact_delta(float s, float p, float shift) {
	float t = fsub(s, p);
	float bp = fsub(t, s);
	float e = fadd(fsub(s, fsub(t, bp)), fsub(-p, bp));
	return fadd(fadd(t, shift), e);
}

// M4 spline density term in the reference's float/double mix (smooth1.c:257-262).
__device__ __forceinline__ float m4_term(float d2, float ih2, float fNorm) {
	float r2 = fmul(d2, ih2);
	float rs = (float)(2.0 - sqrt((double)r2));
	if (r2 < 1.0f) rs = (float)(1.0 - 0.75 * (double)rs * (double)r2);
	else rs = (float)(0.25 * (double)rs * (double)rs * (double)rs);
	return fmul(rs, fNorm);
}
// The same M4 term in fp32 for the compensated mode (no double operations in
// the per-pair code): q from the hardware square root (about 1 ulp), then
// 1 - 0.75 (2 - q) q^2 inside the unit radius and 0.25 (2 - q)^3 outside,
// where 2 - q is exact near q = 2.  A few ulp per term.
__device__ __forceinline__ float m4_term_f32(float d2, float ih2, float fNorm) {
	const float r2 = fmul(d2, ih2);
	const float t = fsub(2.0f, __builtin_amdgcn_sqrtf(r2));
	const float w = r2 < 1.0f ? fmaf(fmul(-0.75f, t), r2, 1.0f) : fmul(fmul(0.25f, t), fmul(t, t));
	return fmul(w, fNorm);
}
template <bool COMP> __device__ __forceinline__ float m4_w(float d2, float ih2, float fNorm) {
	return COMP ? m4_term_f32(d2, ih2, fNorm) : m4_term(d2, ih2, fNorm);
}

// Gradient factor g(q)*norm in the reference's float/double mix
// (smooth1.c:450-454): sqrt of a float in double rounds like a correctly
// rounded fp32 sqrt, and 2.25*rs is exact in double, so the fma is the
// reference's multiply-then-add.
__device__ __forceinline__ float grad_factor(float d2, float ih2, float gnorm) {
	float r2 = fmul(d2, ih2);
	float rs = __fsqrt_rn(r2);
	if (r2 < 1.0f) rs = (float)fma(2.25, (double)rs, -3.0);
	else rs = (float)((-3.0 / (double)rs + 3.0) - 0.75 * (double)rs);
	return fmul(rs, gnorm);
}

// The same factor in fp32 for the compensated mode.  For q >= 1 the
// reference's -3/q + 3 - 0.75 q equals -0.75 (2 - q)^2 / q, which has no
// cancellation near q = 2 (2 - q is exact there).  q and 1/q come from the
// hardware square root and reciprocal (about 1 ulp each), so the factor is
// good to a few ulp; both branches are evaluated and one is selected, so the
// move step's hit path carries no divide and no branch here (q = 0, the
// point on its own entity, selects the inner branch; the outer's inf is
// discarded).
__device__ __forceinline__ float grad_factor_f32(float d2, float ih2, float gnorm) {
	const float r2 = fmul(d2, ih2);
	const float q = __builtin_amdgcn_sqrtf(r2);
	const float t = fsub(2.0f, q);
	const float gin = fmaf(2.25f, q, -3.0f);
	const float gout = fmul(fmul(-0.75f, fmul(t, t)), __builtin_amdgcn_rcpf(q));
	return fmul(r2 < 1.0f ? gin : gout, gnorm);
}

// M4 kernel w(q) and its derivative w'(q) (Monaghan & Lattanzio 1985), fp32.
__device__ __forceinline__ void m4_derivs(float q, float &w, float &wp) {
	if (q < 1.0f) {
		w = fmaf(fmaf(0.75f, q, -1.5f), q * q, 1.0f);
		wp = q * fmaf(2.25f, q, -3.0f);
	} else if (q < 2.0f) {
		const float t = 2.0f - q;
		w = 0.25f * t * t * t;
		wp = -0.75f * t * t;
	} else {
		w = 0.f; wp = 0.f;
	}
}
// One pair's term of d rho / d ln nSmooth before the factor -ih^3/(6 pi): the
// smoothing lengths scale as nSmooth^(1/3), so h dW/dh = -(3w + q w')/(pi h^3)
// and d/d ln k = (1/3) h d/dh; ih = 1/h of the ball the term is taken with.
__device__ __forceinline__ float dlnk_term(float r, float ih, float m) {
	const float q = r * ih;
	float w, wp;
	m4_derivs(q, w, wp);
	return m * (3.0f * w + q * wp);
}

// ---------------------------------------------------------------- host helpers (skg_core.hip)
int skg_ensure_tmp(skg_ctx *c, size_t bytes);
int skg_ensure_keys(skg_ctx *c, size_t n);
int skg_ensure_part(skg_ctx *c, size_t n);
int skg_sort_pairs(skg_ctx *c, unsigned long long *kin, unsigned long long *kout, int *vin, int *vout,
		int n, int endBit);
int skg_build_tree(skg_ctx *c, DevTree *t, const float4 *pos, int n, int B, const float *ball2live,
		const int *ibox);
int skg_build_tree_prefix(skg_ctx *c, DevTree *t, const float4 *pos, const unsigned long long *keys, int n,
		int B, int minDepth, const float *ball2live, const int *ibox);
int skg_refit_tree(skg_ctx *c, DevTree *t, const float4 *pos, const float *ball2live, const int *ibox);
int skg_copy_tree_leaves(skg_ctx *c, DevTree *dst, const DevTree *src);
int skg_morton_sort(skg_ctx *c, const float4 *pos, const int *idx, int n, int *perm_out, const int *box);
int skg_hilbert_sort(skg_ctx *c, const float4 *pos, const int *idx, int n, int *perm_out, const int *box);
int skg_box_bits(const skg_ctx *c);
int skg_exclusive_scan(skg_ctx *c, const int *in, int *out, int n);
// The same on caller storage (tmp null: report the bytes needed in *need only).
int skg_sort_pairs_at(hipStream_t st, void *tmp, size_t tmpBytes, size_t *need, unsigned long long *kin,
		unsigned long long *kout, int *vin, int *vout, int n, int endBit);
int skg_exclusive_scan_at(hipStream_t st, void *tmp, size_t tmpBytes, size_t *need, const int *in, int *out,
		int n);
int skg_curve_keys(skg_ctx *c, const float4 *pos, const int *idx, int n, const int *box, bool hilbert,
		unsigned long long *keys, int *vals, float *part, unsigned *bb);
int skg_tree_leaves2(int n, int B, int *nLeaf);
int skg_build_tree_at(skg_ctx *c, DevTree *t, const float4 *pos, int n, int B, float4 *lo, float4 *hi);
int skg_noalloc_setup(int dev);                         // skg_fof.hip: the scratch array's address
void skg_free_tree(DevTree *t);
int skg_gradient_points(skg_ctx *c, const float4 *q, int nq, float *out);   // skg_adjoint.hip
void skg_timer_start(skg_ctx *c);
int skg_prof_begin(skg_ctx *c);                        // returns a record slot or -1
void skg_prof_end(skg_ctx *c, int slot, int which, double work);
void skg_prof_flush(skg_ctx *c);
void skg_timer_stop(skg_ctx *c, int slot);
// Every device allocation of the library goes through these two, which keep the
// process-wide byte count and its high-water mark (skg_device_memory).
int skg_dmalloc(void **p, size_t bytes);
void skg_dfree(void **p);
template <typename T> static inline int dalloc(T **p, size_t count) {
	if (count == 0) count = 1;
	return skg_dmalloc((void **)p, count * sizeof(T));
}
template <typename T> static inline void dfree(T **p) { skg_dfree((void **)p); }
// An entry point's own device temporaries, freed on every return path.  Declare
// it after the pointers it holds, so that it is destroyed before them; dfree
// skips a pointer already freed (null).
struct TmpFree {
	void **p[8];
	int n = 0;
	template <typename T> void add(T **q) { if (n < 8) p[n++] = (void **)q; }
	~TmpFree() { for (int i = n - 1; i >= 0; --i) skg_dfree(p[i]); }
};
// Stage temporaries: slot s of the context, grown to count items when needed
// and kept for the next call.  Within one entry point every temporary takes
// its own slot; entry points never overlap on a context.
int skg_slot(skg_ctx *c, int s, void **p, size_t bytes);
template <typename T> static inline int sget(skg_ctx *c, int s, T **p, size_t count) {
	return skg_slot(c, s, (void **)p, count * sizeof(T));
}
