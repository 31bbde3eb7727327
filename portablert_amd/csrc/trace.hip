// Nearest-hit traversal for nearest_hits (replaces BVH2::nearest_tri, bvh.hpp:224-265, and
// the thread split of CPUBackend::nearest_hits, intersect_cpu.hpp:20-43).
//
// One lane per ray, persistent waves. A wave takes rays from a global counter in batches:
// lanes whose ray has finished are counted with a ballot, one lane reserves that many ray
// indices with a single atomic and the base is broadcast with readlane, so finished lanes are
// refilled while the others keep traversing. The traversal stack lives in an LDS ring per lane
// (spilling its oldest entries to global memory when deep). Entries carry the entry distance
// of their box so they are re-culled when popped.
//
// Node tests are conservative versions of the reference slab test on the child boxes (fp32 or
// 8-bit quantized boxes rounded outward), so a node is never rejected when the exact test of a
// triangle below it would pass. Every triangle reached gets the exact reference predicates
// (slab on its own box, then Moller-Trumbore, operation by operation). The answer is the
// minimum t over all candidates with equal t resolved to the lowest primitive id.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>

#include "common.h"
#include "internal.h"

namespace prt {
namespace {

struct KArgs {
	const void *nodes;
	const NodeQ *nodes_q;   // quantized nodes: their byte 15 is the never-cull mask of either format
	int any_ncull;          // some triangle is flagged (or the count is unknown)
	int leaf_test;          // 0 Moller-Trumbore (the reference's), 1 watertight
	const TriRec *tris;
	const float *rays;
	uint8_t *out;
	uint32_t *counter;      // [0] work counter, [1..3] stats
	uint2 *spill;           // spill_depth entries per launched thread
	uint64_t n_rays;
	uint32_t n_nodes;
	uint32_t n_tris;
	uint32_t max_entries;   // stack bound 7 * depth + 8
	uint32_t stride;
	uint32_t mask;
	uint32_t spill_depth;
	uint32_t chunk;         // rays a wave reserves per atomic
	uint32_t leaf_phase;    // triangle phase once parked lanes >= live lanes * leaf_phase / 64
	uint32_t refill;        // load new rays once at least this many lanes are idle (or all are)
	float margin_rel;
	float margin_abs;
	int cull;
	int fast_scene;         // every vertex coordinate below 2^60 in magnitude
	// learned seeding (null when off): 4 cluster ids per ray, cluster entries and fp32 boxes
	const uint32_t *seeds;
	const uint32_t *cl_entry;
	const float *cl_box;
	uint32_t n_clusters;
	int n_seeds;            // predicted clusters searched first (1..4, the best ones)
	// several scenes (null when off): ray tiles {first lo, first hi, count, scene} claimed one
	// per atomic, and per scene {margin_abs, fast_scene}; BVH8 node s is the root of scene s
	const uint4 *tiles;
	uint32_t n_tiles;
	const float2 *scene_info;
};

// kWatertight: the plain traversal with the watertight leaf test (wt_candidate), no culling.
// kSeeded searches the best predicted cluster first, kSeededMulti the best 2..4.
enum : int { kPlain = 0, kSeeded = 1, kBatch = 2, kWatertight = 3, kSeededMulti = 4 };

constexpr uint32_t kInvalid = 0xFFFFFFFFu;

// Per-lane ray state.
struct Ray {
	float o[3], d[3], inv[3];
	float tb, ub, vb;
	uint32_t idb;
	float abs_margin;  // margin_abs / |d|
	bool degen;        // some 1/d is infinite: axis-parallel handling in node tests
	bool patho;        // non-finite origin or NaN direction: every node passes
	bool fast;         // |1/d| and |o| below 2^60 in a scene within 2^60: one-fma node bounds
};

__device__ __forceinline__ void load_ray(const float *rays, uint64_t idx, float margin_abs, bool fast_scene,
                                         Ray &r) {
	const float *p = rays + 6 * idx;  // 4-byte aligned AoS records of 24 bytes
#pragma unroll
	for (int k = 0; k < 3; ++k) {
		r.o[k] = p[k];
		r.d[k] = p[3 + k];
	}
#pragma unroll
	for (int k = 0; k < 3; ++k) r.inv[k] = 1.0f / r.d[k];  // IEEE division, as the reference
	r.degen = !(__builtin_isfinite(r.inv[0]) && __builtin_isfinite(r.inv[1]) && __builtin_isfinite(r.inv[2]));
	r.patho = !(__builtin_isfinite(r.o[0]) && __builtin_isfinite(r.o[1]) && __builtin_isfinite(r.o[2])) ||
	          __builtin_isnan(r.d[0]) || __builtin_isnan(r.d[1]) || __builtin_isnan(r.d[2]);
	// |1/d| below 2^30 (every |d_k| above 2^-30), |o| below 2^60 in a scene within 2^60
	const float big = 0x1p60f, inv_big = 0x1p30f;
	r.fast = fast_scene && !r.degen && !r.patho && fabsf(r.inv[0]) < inv_big && fabsf(r.inv[1]) < inv_big &&
	         fabsf(r.inv[2]) < inv_big && fabsf(r.o[0]) < big && fabsf(r.o[1]) < big && fabsf(r.o[2]) < big;
	// margin_abs / |d|, rounded up (1 + 2^-20 covers the reciprocal square root's error). An
	// infinite margin turns distance culling off for this ray: culling is exact only for rays
	// with every direction component within [2^-30, 2^30] (|d|^2 <= 2^60 bounds them above) and
	// |o| < 2^60 (cull_unsafe in common.h), i.e. fast rays; other rays (zero direction
	// components, huge or tiny values) traverse without it.
	const float len2 = r.d[0] * r.d[0] + r.d[1] * r.d[1] + r.d[2] * r.d[2];
	r.abs_margin = (r.fast && len2 <= 0x1p60f) ? margin_abs * (__builtin_amdgcn_rsqf(len2) * (1.0f + 0x1p-20f))
	                                           : __builtin_inff();
	r.tb = __builtin_inff();
	r.ub = r.vb = 0.0f;
	r.idb = kInvalid;
}

// Distance beyond which no box can hold a better candidate (+inf: no culling).
__device__ __forceinline__ float cull_limit(const Ray &r, const KArgs &a) {
	if (!a.cull || !__builtin_isfinite(r.tb)) return __builtin_inff();
	return r.tb + fmaxf(fabsf(r.tb) * a.margin_rel, r.abs_margin);
}

// Conservative slab test of one child box [lo, hi]. Returns the entry distance (tmin), or NaN
// when the box is rejected. WIDE (the watertight mode): the interval is widened by 2^-20 of its
// end points, so that a ray through a box corner (a vertex shared by several triangles) is not
// rejected by the rounding of tmin and tmax.
template <bool WIDE = false>
__device__ __forceinline__ float child_slab(const Ray &r, const float lo[3], const float hi[3], float limit) {
#pragma clang fp contract(off)
	float amin[3], amax[3];
	bool poison = false;
#pragma unroll
	for (int k = 0; k < 3; ++k) {
		const float t1 = (lo[k] - r.o[k]) * r.inv[k];
		const float t2 = (hi[k] - r.o[k]) * r.inv[k];
		amin[k] = fminf(t1, t2);
		amax[k] = fmaxf(t1, t2);
	}
	if (r.degen) {
		// 1/d infinite on an axis: a box side through the origin gives NaN in the exact test.
		// Treat any box containing the origin's coordinate on that axis as unbounded there;
		// on the x axis such a NaN makes the exact test pass outright (bvh.hpp:210-217 with
		// std::max/std::min argument order), so the box must pass.
#pragma unroll
		for (int k = 0; k < 3; ++k) {
			if (!__builtin_isfinite(r.inv[k]) && lo[k] <= r.o[k] && r.o[k] <= hi[k]) {
				amin[k] = -__builtin_inff();
				amax[k] = __builtin_inff();
				if (k == 0) poison = true;
			}
		}
	}
	const float tmin = fmaxf(fmaxf(amin[0], amin[1]), amin[2]);
	const float tmax = fminf(fminf(amax[0], amax[1]), amax[2]);
	if (poison || r.patho) return -__builtin_inff();
	const float slack = WIDE ? 0x1p-20f * (fabsf(tmin) + fabsf(tmax)) : 0.0f;
	const bool hit = (tmax >= -slack) && (tmin <= tmax + slack) && (tmin <= limit);
	return hit ? tmin : __builtin_nanf("");
}

// LDS ring stack: the newest `S` entries of each lane stay in LDS, older ones are moved to
// global memory (the entry at depth k of launched thread g lives at spill[(k - S) * threads + g]
// once evicted, so lanes spilling at equal depths write adjacent words). An entry is
// {work entry, sort key of its box entry distance}.
template <int S, int B>
struct Stack {
	uint2 *lds;        // [S][B]
	uint2 *spill;      // shared spill array
	uint32_t thread;   // launched thread index g
	uint32_t threads;  // launched threads
	uint32_t sp;
	__device__ __forceinline__ uint2 *spilled(uint32_t depth) const {
		return spill + (static_cast<uint64_t>(depth - S) * threads + thread);
	}
	__device__ __forceinline__ void push(uint32_t ref, uint32_t key) {
		uint2 *slot = lds + (sp & (S - 1)) * B;
		if (sp >= static_cast<uint32_t>(S)) *spilled(sp) = *slot;
		*slot = make_uint2(ref, key);
		++sp;
	}
	__device__ __forceinline__ uint2 pop() {
		--sp;
		uint2 *slot = lds + (sp & (S - 1)) * B;
		const uint2 e = *slot;
		if (sp >= static_cast<uint32_t>(S)) *slot = *spilled(sp);
		return e;
	}
};

// Sort key of a child hit: the order-preserving integer image of its entry distance with the
// low three bits replaced by the child slot. Clearing bits only lowers the distance the key
// stands for, so re-culling by key at pop stays conservative. -inf (a box that must pass) is
// clamped to -FLT_MAX, which every culling limit admits. Misses sort last.
constexpr uint32_t kMiss = 0xFFFFFFFFu;

// Key of a finite (or +inf) entry distance.
__device__ __forceinline__ uint32_t sort_key_finite(float tmin, int slot) {
	const uint32_t u = __float_as_uint(tmin);
	const uint32_t o = u ^ (static_cast<uint32_t>(static_cast<int32_t>(u) >> 31) | 0x80000000u);
	return (o & ~7u) | static_cast<uint32_t>(slot);
}

__device__ __forceinline__ uint32_t sort_key(float tmin, int slot) {
	return sort_key_finite(fmaxf(tmin, -3.402823466e38f), slot);
}

// Largest key whose decoded distance is within `limit` (not NaN): a key k passes iff
// k <= key_limit(limit). The decoded distance of k is at most the box's entry distance.
__device__ __forceinline__ uint32_t key_limit(float limit) {
	const uint32_t u = __float_as_uint(limit);
	return (u ^ (static_cast<uint32_t>(static_cast<int32_t>(u) >> 31) | 0x80000000u)) | 7u;
}

#define PRT_CE(i, j)                                                                               \
	{                                                                                              \
		const uint32_t lo_ = min(k[i], k[j]);                                                      \
		k[j] = max(k[i], k[j]);                                                                    \
		k[i] = lo_;                                                                                \
	}

// Sorting network for 8 keys (19 compare-exchanges).
__device__ __forceinline__ void sort8(uint32_t k[8]) {
	PRT_CE(0, 2) PRT_CE(1, 3) PRT_CE(4, 6) PRT_CE(5, 7) PRT_CE(0, 4) PRT_CE(1, 5) PRT_CE(2, 6) PRT_CE(3, 7)
	PRT_CE(0, 1) PRT_CE(2, 3) PRT_CE(4, 5) PRT_CE(6, 7) PRT_CE(2, 4) PRT_CE(3, 5) PRT_CE(1, 4) PRT_CE(3, 6)
	PRT_CE(1, 2) PRT_CE(3, 4) PRT_CE(5, 6)
}
#undef PRT_CE

// Children of one node that pass the box test, sorted by entry distance, plus what is needed
// to turn a key back into a stack entry.
struct NodeHits {
	uint32_t key[8];
	uint32_t child_base, lbase;
	uint64_t meta;  // the 8 meta bytes, slot k in bits 8k..8k+7
	__device__ __forceinline__ uint32_t entry(uint32_t key) const {
		const uint32_t m = static_cast<uint32_t>(meta >> ((key << 3) & 63u));
		return child_entry(m & 0xFFu, child_base, lbase);
	}
};

// Key of a child slot whose subtree holds a triangle the build flagged as ill-conditioned
// (cull_unsafe in common.h): the order-preserving image of -FLT_MAX, which no culling limit
// rejects, so such a box is never culled (and is visited before its siblings).
constexpr uint32_t kKeyNoCull = 0x00800000u;

// Tests the children of one node (fast_ok selects the one-fma bound evaluation). Slots whose bit
// is set in the node's never-cull mask pass whenever the box is hit, at any distance, and take
// the key kKeyNoCull | slot; the other slots of such a node are tested without the limit and
// culled only when popped. `ncull_q` is the quantized node array, which holds the mask of
// both formats (byte 15 of node idx).
template <class Node, bool NC, bool WIDE>
__device__ __forceinline__ void test_children(const Node *nodes, const NodeQ *ncull_q, uint32_t idx, const Ray &r,
                                              float limit, bool fast_ok, NodeHits &h);

// Byte k of a little-endian dword pair, as a float (one v_cvt_f32_ubyteN when k is constant).
__device__ __forceinline__ float byte_f(const uint32_t w[2], int k) {
	return static_cast<float>((w[k >> 2] >> (8 * (k & 3))) & 0xFFu);
}

// Flagged slots that were hit take the never-cull key (their boxes were tested without the
// culling limit).
__device__ __forceinline__ void no_cull_keys(uint32_t key[8], uint32_t ncull) {
#pragma unroll
	for (int k = 0; k < 8; ++k)
		if (((ncull >> k) & 1u) && key[k] != kMiss) key[k] = kKeyNoCull | static_cast<uint32_t>(k);
}

// NC = false: the scene has no flagged triangle (every mask is 0) and the masks are not read.
// WIDE: widened reference-formula intervals (child_slab).
template <class Node, bool NC, bool WIDE>
struct ChildTest;

template <bool NC, bool WIDE>
struct ChildTest<NodeQ, NC, WIDE> {
static __device__ __forceinline__ void run(const NodeQ *nodes, const NodeQ * /*ncull_q*/, uint32_t idx, const Ray &r,
                                           float limit, bool fast_ok, NodeHits &h) {
	const uint4 *src = reinterpret_cast<const uint4 *>(nodes + idx);
	const uint4 w0 = src[0], w1 = src[1], w2 = src[2], w3 = src[3], w4 = src[4];
	const float p[3] = {__uint_as_float(w0.x), __uint_as_float(w0.y), __uint_as_float(w0.z)};
	const float s[3] = {exp2i(static_cast<int8_t>(w0.w & 0xFF)), exp2i(static_cast<int8_t>((w0.w >> 8) & 0xFF)),
	                    exp2i(static_cast<int8_t>((w0.w >> 16) & 0xFF))};
	const uint32_t ncull = NC ? (w0.w >> 24) : 0u;  // never-cull mask (the node's byte 15)
	h.child_base = w1.x;
	h.lbase = leaf_base(w1.y);
	const uint32_t meta_lo = w1.z, meta_hi = w1.w;
	h.meta = static_cast<uint64_t>(meta_lo) | (static_cast<uint64_t>(meta_hi) << 32);
	// qlo x: w2.x w2.y, qlo y: w2.z w2.w, qlo z: w3.x w3.y, qhi x: w3.z w3.w, qhi y: w4.x w4.y, qhi z: w4.z w4.w
	const uint32_t ql[3][2] = {{w2.x, w2.y}, {w2.z, w2.w}, {w3.x, w3.y}};
	const uint32_t qh[3][2] = {{w3.z, w3.w}, {w4.x, w4.y}, {w4.z, w4.w}};

	// Fast test: on axis a the child bound q lies at t = q * step_a + base_a in real arithmetic,
	// with step_a = 2^e_a / d_a and base_a = (p_a - o_a) / d_a. One fma per bound replaces the
	// dequantize / subtract / multiply of the reference formula. Both fp32 evaluations stay
	// within 8 * 2^-24 * |1/d_a| * (255 * 2^e_a + |p_a| + |o_a|) of the real value, so moving the
	// near plane in and the far plane out by 2^-18 times that amount keeps every fast interval
	// around the interval the reference formula gives for the same dequantized box (which the
	// builder already rounded outward). fast_ok guarantees |1/d|, |o| and the scene coordinates
	// are below 2^60, so none of these values overflows; other rays take the reference formula.
	if (fast_ok) {
		// a slot is used iff its meta byte is not 0xFF, i.e. its byte of the complement is not 0
		const uint32_t unused_lo = ~meta_lo, unused_hi = ~meta_hi;
		float step[3], bnear[3], bfar[3];
		uint32_t qn[3][2], qf[3][2];
#pragma unroll
		for (int ax = 0; ax < 3; ++ax) {
			step[ax] = s[ax] * r.inv[ax];
			const float base = (p[ax] - r.o[ax]) * r.inv[ax];
			const float margin = (0x1p-18f * fabsf(r.inv[ax])) * fmaf(255.0f, s[ax], fabsf(p[ax]) + fabsf(r.o[ax]));
			bnear[ax] = base - margin;
			bfar[ax] = base + margin;
			// the near plane is the low bound for 1/d >= 0, the high bound otherwise
			const bool pos = r.inv[ax] >= 0.0f;
			qn[ax][0] = pos ? ql[ax][0] : qh[ax][0];
			qn[ax][1] = pos ? ql[ax][1] : qh[ax][1];
			qf[ax][0] = pos ? qh[ax][0] : ql[ax][0];
			qf[ax][1] = pos ? qh[ax][1] : ql[ax][1];
		}
		// a node with a flagged slot tests every slot without the limit (its unflagged slots are
		// still culled when popped); the flagged slots' keys are replaced below
		const float lim = ncull != 0u ? __builtin_inff() : limit;
#pragma unroll
		for (int k = 0; k < 8; ++k) {
			const uint32_t used = ((k < 4 ? unused_lo : unused_hi) >> (8 * (k & 3))) & 0xFFu;
			float tn[3], tf[3];
#pragma unroll
			for (int ax = 0; ax < 3; ++ax) {
				tn[ax] = fmaf(byte_f(qn[ax], k), step[ax], bnear[ax]);
				tf[ax] = fmaf(byte_f(qf[ax], k), step[ax], bfar[ax]);
			}
			const float tmin = fmaxf(fmaxf(tn[0], tn[1]), tn[2]);
			const float tmax = fminf(fminf(tf[0], tf[1]), tf[2]);
			// the fast bounds are finite (no overflow below 2^60), so no clamp in the key
			const bool hit = used != 0u && tmax >= 0.0f && tmin <= tmax && tmin <= lim;
			h.key[k] = hit ? sort_key_finite(tmin, k) : kMiss;
		}
		if (ncull != 0u) no_cull_keys(h.key, ncull);
		sort8(h.key);
		return;
	}
	// Reference formula on the dequantized box (also the path of the exact mode).
#pragma unroll
	for (int k = 0; k < 8; ++k) {
		const uint32_t meta = ((k < 4 ? meta_lo : meta_hi) >> (8 * (k & 3))) & 0xFFu;
		float lo[3], hi[3];
#pragma unroll
		for (int ax = 0; ax < 3; ++ax) {
			const uint32_t bl = (ql[ax][k >> 2] >> (8 * (k & 3))) & 0xFFu;
			const uint32_t bh = (qh[ax][k >> 2] >> (8 * (k & 3))) & 0xFFu;
			lo[ax] = fmaf(static_cast<float>(bl), s[ax], p[ax]);
			hi[ax] = fmaf(static_cast<float>(bh), s[ax], p[ax]);
		}
		const float tmin = child_slab<WIDE>(r, lo, hi, ncull != 0u ? __builtin_inff() : limit);
		const bool hit = meta != kMetaEmpty && !__builtin_isnan(tmin);
		h.key[k] = hit ? sort_key(tmin, k) : kMiss;
	}
	if (ncull != 0u) no_cull_keys(h.key, ncull);
	sort8(h.key);
}
};

template <bool NC, bool WIDE>
struct ChildTest<NodeF, NC, WIDE> {
static __device__ __forceinline__ void run(const NodeF *nodes, const NodeQ *ncull_q, uint32_t idx, const Ray &r,
                                           float limit, bool /*fast_ok*/, NodeHits &hh) {
	const NodeF *n = nodes + idx;
	const uint4 h = *reinterpret_cast<const uint4 *>(n);
	const uint32_t ncull = NC ? reinterpret_cast<const uint32_t *>(ncull_q + idx)[3] >> 24 : 0u;
	hh.child_base = h.x;
	hh.lbase = leaf_base(h.y);
	hh.meta = static_cast<uint64_t>(h.z) | (static_cast<uint64_t>(h.w) << 32);
	float lo[3][8], hi[3][8];
#pragma unroll
	for (int ax = 0; ax < 3; ++ax) {
		const float4 *l4 = reinterpret_cast<const float4 *>(n->lo[ax]);
		const float4 *h4 = reinterpret_cast<const float4 *>(n->hi[ax]);
		const float4 a = l4[0], b = l4[1], c = h4[0], d = h4[1];
		lo[ax][0] = a.x; lo[ax][1] = a.y; lo[ax][2] = a.z; lo[ax][3] = a.w;
		lo[ax][4] = b.x; lo[ax][5] = b.y; lo[ax][6] = b.z; lo[ax][7] = b.w;
		hi[ax][0] = c.x; hi[ax][1] = c.y; hi[ax][2] = c.z; hi[ax][3] = c.w;
		hi[ax][4] = d.x; hi[ax][5] = d.y; hi[ax][6] = d.z; hi[ax][7] = d.w;
	}
#pragma unroll
	for (int k = 0; k < 8; ++k) {
		const uint32_t meta = ((k < 4 ? h.z : h.w) >> (8 * (k & 3))) & 0xFFu;
		const float l[3] = {lo[0][k], lo[1][k], lo[2][k]};
		const float u[3] = {hi[0][k], hi[1][k], hi[2][k]};
		const float tmin = child_slab<WIDE>(r, l, u, ncull != 0u ? __builtin_inff() : limit);
		const bool hit = meta != kMetaEmpty && !__builtin_isnan(tmin);
		hh.key[k] = hit ? sort_key(tmin, k) : kMiss;
	}
	if (ncull != 0u) no_cull_keys(hh.key, ncull);
	sort8(hh.key);
}
};

template <class Node, bool NC, bool WIDE>
__device__ __forceinline__ void test_children(const Node *nodes, const NodeQ *ncull_q, uint32_t idx, const Ray &r,
                                              float limit, bool fast_ok, NodeHits &h) {
	ChildTest<Node, NC, WIDE>::run(nodes, ncull_q, idx, r, limit, fast_ok, h);
}

// Lane 0's value in every lane, as a wave-uniform (scalar) value; every lane of the wave must be
// active.
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
	const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
	const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
	return static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32);
}

// Writes the finished ray's record in the caller's layout.
__device__ __forceinline__ void write_record(const KArgs &a, uint64_t ray_idx, const Ray &r) {
	Hit h;
	h.t = r.tb;
	h.u = r.ub;
	h.v = r.vb;
	h.id = r.idb;
	write_any(a.mask, a.out + ray_idx * a.stride, h, r.o, r.d);
}

template <bool WT>
__device__ __forceinline__ void test_leaf(const TriRec *tris, uint32_t entry, Ray &r) {
	const uint32_t first = leaf_first(entry);
	const uint32_t count = leaf_count(entry);
	for (uint32_t m = 0; m < count; ++m) {
		const uint4 *src = reinterpret_cast<const uint4 *>(tris + first + m);
		const uint4 a = src[0], b = src[1], c = src[2];
		const float v[9] = {__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z),
		                    __uint_as_float(a.w), __uint_as_float(b.x), __uint_as_float(b.y),
		                    __uint_as_float(b.z), __uint_as_float(b.w), __uint_as_float(c.x)};
		const uint32_t id = c.y;
		float t, u, vv;
		const bool cand = WT ? wt_candidate(v, r.o, r.d, r.inv, t, u, vv) : exact_candidate(v, r.o, r.d, r.inv, t, u, vv);
		if (cand && better(t, id, r.tb, r.idb)) {
			r.tb = t;
			r.ub = u;
			r.vb = vv;
			r.idb = id;
		}
	}
}

// Waves per SIMD the register allocation must allow. Quantized nodes with an 8-entry LDS stack
// (4 KB per wave): 8 for the plain mode, i.e. at most 64 VGPRs; the traversal is partly
// latency-bound (dependent node fetches) and the extra waves hide more of it than the few
// spilled registers cost. The seeded and batch modes carry more state per lane (seed clusters,
// scene tiles) and spill more at 64 VGPRs: they run at 7 waves (72 VGPRs). Spills inside the
// loop are costly here: the multi-seed state (three more entries, their flags) spills 14
// registers at 72 VGPRs, and the one-seed mode, which carries one entry, runs 10% faster with
// 2. A 16-entry stack allows 5 waves by LDS alone. The fp32 node test needs about 100 VGPRs
// and is left to the compiler (1 = no constraint).
template <class Node, int S, int MODE>
constexpr int kWavesPerEu = std::is_same<Node, NodeQ>::value ? (S == 8 ? (MODE == kPlain ? 8 : 7) : 5) : 1;

template <class Node, int S, int B, bool COUNT, int MODE, bool NC>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(kWavesPerEu<Node, S, MODE>))) void k_trace(KArgs a) {
	constexpr bool MULTI = MODE == kSeededMulti, SEEDED = MODE == kSeeded || MULTI, BATCH = MODE == kBatch;
	__shared__ uint2 lds[S * B];
	const Node *nodes = static_cast<const Node *>(a.nodes);
	const uint32_t lane = threadIdx.x & 63u;
	Stack<S, B> st;
	st.lds = lds + threadIdx.x;
	st.spill = a.spill;
	st.thread = blockIdx.x * B + threadIdx.x;
	st.threads = gridDim.x * B;
	st.sp = 0;

	Ray r;
	uint64_t ray_idx = ~0ull;
	bool active = false;       // lane holds a ray
	uint32_t cur = kInvalid;   // next entry to process (node or leaf)
	uint32_t pend = kInvalid;  // parked leaf, tested in the next triangle phase
	// learned seeding: the entries of the seed clusters (sd[0] the best) and, with several
	// seeds, which of sd[1..3] have been taken from the stack
	uint32_t sd[4] = {kInvalid, kInvalid, kInvalid, kInvalid};
	uint32_t seen = 0u;
	// wave-uniform ray chunk [chunk_next, chunk_end) reserved with one atomic
	uint64_t chunk_next = 0, chunk_end = 0;
	uint32_t chunk_scene = 0;
	bool exhausted = false;
	unsigned long long c_nodes = 0, c_tris = 0, c_leaves = 0, c_nphase = 0, c_tphase = 0, c_culled = 0;

	// Issue priority: 1 while the wave may still reserve rays, 0 once the work is used up. At the
	// end of a launch the waves that still hold reserved rays (the last to finish) then win the
	// issue slots over waves draining their last rays. The SIMD's arbiter otherwise favours the
	// older waves, and the younger ones, which trace about half as many rays, end the launch.
	__builtin_amdgcn_s_setprio(1);
	while (true) {
		// ---- refill finished lanes from the wave's chunk (one ballot when none is idle); the
		// ray loads are amortised over at least `refill` lanes unless the whole wave is idle
		const uint64_t need = __ballot(!active);
		const uint32_t n_need = __popcll(need);
		if (n_need != 0u && !exhausted && (n_need >= a.refill || n_need == __popcll(__ballot(true)))) {
			if (chunk_next >= chunk_end && BATCH) {
				uint64_t tix = 0;
				if (lane == 0) tix = atomicAdd(reinterpret_cast<unsigned long long *>(a.counter), 1ull);
				tix = uniform64(tix);
				if (tix < a.n_tiles) {
					const uint4 tl = a.tiles[tix];
					chunk_next = static_cast<uint64_t>(tl.x) | (static_cast<uint64_t>(tl.y) << 32);
					chunk_end = chunk_next + tl.z;
					chunk_scene = tl.w;
				} else {
					exhausted = true;
					chunk_next = chunk_end = 0;
					__builtin_amdgcn_s_setprio(0);
				}
			} else if (chunk_next >= chunk_end) {
				uint64_t base = 0;
				if (lane == 0) base = atomicAdd(reinterpret_cast<unsigned long long *>(a.counter),
				                                static_cast<unsigned long long>(a.chunk));
				base = uniform64(base);
				chunk_next = base;
				chunk_end = base + a.chunk < a.n_rays ? base + a.chunk : a.n_rays;
				if (base >= a.n_rays) {
					exhausted = true;
					__builtin_amdgcn_s_setprio(0);
				}
			}
			const uint64_t avail = chunk_end > chunk_next ? chunk_end - chunk_next : 0ull;
			// idle lanes below this one
			const uint32_t rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(need >> 32),
			                                                __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(need), 0u));
			if (!active && rank < avail) {
				ray_idx = chunk_next + rank;
				active = true;
				if (BATCH) {
					const float2 si = a.scene_info[chunk_scene];
					load_ray(a.rays, ray_idx, si.x, si.y != 0.0f, r);
					cur = chunk_scene;
				} else {
					load_ray(a.rays, ray_idx, a.margin_abs, a.fast_scene != 0, r);
					cur = a.n_nodes > 0 ? 0u : kInvalid;
				}
				pend = kInvalid;
				st.sp = 0;
				if (SEEDED && a.n_nodes > 0) {
					// The lane starts at the best predicted cluster's entry instead of the root,
					// which waits at the bottom of the stack; with several seeds the others sit
					// above the root (the better on top) with their box entry distances. The
					// clusters' subtrees are intersected exactly first and seed the culling
					// distance; the root traversal afterwards covers every triangle except those
					// of the clusters already searched. The best cluster's box is not tested:
					// the leaf test is exact, so a ray that misses it only visits its entry.
					const uint4 s4 = reinterpret_cast<const uint4 *>(a.seeds)[ray_idx];  // best first
					sd[0] = s4.x < a.n_clusters ? a.cl_entry[s4.x] : 0u;
					if (sd[0] == 0u) sd[0] = kInvalid;  // no seed, or the cluster is the whole tree
					uint32_t key[4] = {};
					if (MULTI) {
						seen = 0u;
						const uint32_t sc[4] = {s4.x, s4.y, s4.z, s4.w};
#pragma unroll
						for (int q = 1; q < 4; ++q) {
							sd[q] = kInvalid;
							if (q < a.n_seeds && sc[q] < a.n_clusters) {
								const float *bx = a.cl_box + 6ull * sc[q];
								const float lo[3] = {bx[0], bx[1], bx[2]}, hi[3] = {bx[3], bx[4], bx[5]};
								const float tmin = child_slab(r, lo, hi, __builtin_inff());
								const uint32_t ce = a.cl_entry[sc[q]];
								if (ce != 0u && !__builtin_isnan(tmin)) {
									sd[q] = ce;
									key[q] = sort_key(tmin, 0);
								}
							}
						}
					}
					if (sd[0] != kInvalid || (MULTI && (sd[1] & sd[2] & sd[3]) != kInvalid)) {
						st.push(0u, sort_key(-3.402823466e38f, 0));
						if (MULTI) {
#pragma unroll
							for (int q = 3; q >= 1; --q)
								if (sd[q] != kInvalid) st.push(sd[q], key[q]);
						}
						cur = sd[0];  // kInvalid: the next pop takes the best pushed cluster
					}
				}
			}
			const uint64_t took = n_need < avail ? n_need : avail;
			chunk_next += took;
		}
		if (__ballot(active) == 0ull) {
			if (exhausted || chunk_next >= chunk_end) {
				if (exhausted) break;
				continue;  // chunk used up with every lane idle: reserve the next one
			}
		}
		if (!active) continue;

		// clusters searched in the seed phase are not searched again: their entries are dropped
		// when the root traversal takes them (from the stack or as a node's nearest child). The
		// clusters are disjoint subtrees, so a cluster's entry is never met inside another
		// cluster's search; the best one is never on the stack, so meeting its entry always
		// means the root traversal reached it. A pushed cluster counts once it has been taken
		// from the stack (one culled there is traversed if the root traversal reaches it).
		auto searched = [&](uint32_t e) {
			if (!SEEDED) return false;
			bool s = e == sd[0];
			if (MULTI) {
#pragma unroll
				for (int q = 1; q < 4; ++q) s |= ((seen >> q) & 1u) && e == sd[q];
			}
			return s;
		};

		// ---- fetch work: pop one entry (a culled one leaves the lane without work for this
		// iteration: no divergent loop over consecutive culled entries); park a leaf if the
		// parking slot is free
		if (cur == kInvalid && st.sp > 0) {
			const uint32_t klim = key_limit(cull_limit(r, a));
			{
				const uint2 e = st.pop();
				if (e.y <= klim) {
					cur = e.x;
					if (searched(cur)) {
						cur = kInvalid;
					} else if (MULTI) {
						// a pushed cluster taken here is searched before anything below it
#pragma unroll
						for (int q = 1; q < 4; ++q) seen |= cur == sd[q] ? (1u << q) : 0u;
					}
				} else if (COUNT) {
					++c_culled;
				}
			}
		}
		if (cur != kInvalid && (cur & kLeafBit) && pend == kInvalid) {
			pend = cur;
			cur = kInvalid;
		}
		if (cur == kInvalid && pend == kInvalid && st.sp == 0) {  // ray finished
			write_record(a, ray_idx, r);
			active = false;
			continue;
		}

		// ---- one phase per iteration for the whole wave: nodes, or the triangles of every
		// parked leaf. Lanes with a parked leaf keep traversing nodes meanwhile; the triangle
		// phase runs when no lane has a node to visit or enough lanes have parked a leaf.
		const uint64_t want_node = __ballot(cur != kInvalid && !(cur & kLeafBit));
		const uint64_t parked = __ballot(pend != kInvalid);
		const uint64_t live = __ballot(true);
		const bool tri_phase = want_node == 0ull || __popcll(parked) * 64u >= __popcll(live) * a.leaf_phase;
		if (COUNT && lane == static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(live)) - 1)) {
			if (tri_phase) ++c_tphase;
			else ++c_nphase;
		}
		if (tri_phase) {
			if (pend != kInvalid) {
				const uint32_t tri = leaf_first(pend);
				const uint32_t left = leaf_count(pend);
				if (tri + left > a.n_tris) {  // corrupt entry: report, never read
					atomicOr(a.counter + 8, 1u);
					pend = kInvalid;
					continue;
				}
				test_leaf<MODE == kWatertight>(a.tris, pend, r);  // the whole leaf (1..4 triangles)
				if (COUNT) c_tris += left;
				pend = kInvalid;
			}
		} else if (cur != kInvalid && !(cur & kLeafBit)) {
			if (cur >= a.n_nodes || st.sp + 7u > a.max_entries) {  // corrupt node or stack bound
				atomicOr(a.counter + 8, 2u);
				cur = kInvalid;
				continue;
			}
			NodeHits h;
			// the one-fma test's margins also cover the watertight mode's box corners
			test_children<Node, NC, MODE == kWatertight>(nodes, a.nodes_q, cur, r, cull_limit(r, a),
			                                             (MODE == kWatertight || a.cull) && r.fast, h);
			if (COUNT) ++c_nodes;
			cur = h.key[0] != kMiss ? h.entry(h.key[0]) : kInvalid;
			if (searched(cur)) cur = kInvalid;
			// farthest first, so the nearest remaining child is on top
#pragma unroll
			for (int k = 7; k >= 1; --k) {
				if (h.key[k] != kMiss) {
					const uint32_t e = h.entry(h.key[k]);
					if (COUNT) c_leaves += (e & kLeafBit) ? 1u : 0u;
					st.push(e, h.key[k]);
				}
			}
		}
	}
	if (COUNT) {
		atomicAdd(reinterpret_cast<unsigned long long *>(a.counter) + 1, c_nodes);
		atomicAdd(reinterpret_cast<unsigned long long *>(a.counter) + 2, c_tris);
		atomicAdd(reinterpret_cast<unsigned long long *>(a.counter) + 3, c_leaves);
		atomicAdd(reinterpret_cast<unsigned long long *>(a.counter) + 5, c_nphase);
		atomicAdd(reinterpret_cast<unsigned long long *>(a.counter) + 6, c_tphase);
		atomicAdd(reinterpret_cast<unsigned long long *>(a.counter) + 7, c_culled);
	}
}

__global__ void k_zero(unsigned long long *p, int n) {
	if (static_cast<int>(threadIdx.x) < n) p[threadIdx.x] = 0ull;
}

template <class Node, int S, int B, int MODE, bool NC>
void launch_nc(const KArgs &a, uint32_t grid, bool count, hipStream_t st) {
	if (count)
		hipLaunchKernelGGL((k_trace<Node, S, B, true, MODE, NC>), dim3(grid), dim3(B), 0, st, a);
	else
		hipLaunchKernelGGL((k_trace<Node, S, B, false, MODE, NC>), dim3(grid), dim3(B), 0, st, a);
}

// The batch always reads the never-cull masks (its builds are not read back); the other modes
// only when the scene has a flagged triangle.
template <class Node, int S, int B, int MODE>
void launch_m(const KArgs &a, uint32_t grid, bool count, hipStream_t st) {
	if constexpr (MODE == kBatch) {
		launch_nc<Node, S, B, MODE, true>(a, grid, count, st);
	} else if constexpr (MODE == kWatertight) {  // no distance culling: the masks are not read
		launch_nc<Node, S, B, MODE, false>(a, grid, count, st);
	} else {
		if (a.any_ncull) launch_nc<Node, S, B, MODE, true>(a, grid, count, st);
		else launch_nc<Node, S, B, MODE, false>(a, grid, count, st);
	}
}

template <class Node, int S, int B>
hipError_t launch(const KArgs &a, uint32_t grid, bool count, hipStream_t st) {
	if (a.tiles) launch_m<Node, S, B, kBatch>(a, grid, count, st);
	else if (a.seeds && a.n_seeds > 1) launch_m<Node, S, B, kSeededMulti>(a, grid, count, st);
	else if (a.seeds) launch_m<Node, S, B, kSeeded>(a, grid, count, st);
	else if (a.leaf_test == 1) launch_m<Node, S, B, kWatertight>(a, grid, count, st);
	else launch_m<Node, S, B, kPlain>(a, grid, count, st);
	return hipGetLastError();
}

template <class Node, int B>
hipError_t launch_s(const KArgs &a, uint32_t grid, int S, bool count, hipStream_t st) {
	switch (S) {
	case 8: return launch<Node, 8, B>(a, grid, count, st);
	case 16: return launch<Node, 16, B>(a, grid, count, st);
	default: return hipErrorInvalidValue;
	}
}

template <class Node>
hipError_t launch_b(const KArgs &a, uint32_t grid, int B, int S, bool count, hipStream_t st) {
	switch (B) {
	case 64: return launch_s<Node, 64>(a, grid, S, count, st);
	case 128: return launch_s<Node, 128>(a, grid, S, count, st);
	case 256: return launch_s<Node, 256>(a, grid, S, count, st);
	default: return hipErrorInvalidValue;
	}
}

}  // namespace

#define PRT_TRY(x)                                                                                 \
	do {                                                                                           \
		hipError_t e_ = (x);                                                                       \
		if (e_ != hipSuccess) return e_;                                                           \
	} while (0)

namespace {

// Launch shape and stack bound of one traversal.
struct TracePlan {
	uint32_t grid = 0;         // workgroups
	uint32_t max_entries = 0;  // stack bound 7 * depth + 8
	uint32_t spill_depth = 0;  // stack entries per thread kept in global memory
};

hipError_t plan_trace(const Bvh8 &b, TraceScratch &s, const TraceParams &p, uint64_t n, TracePlan &pl) {
	if (s.n_cus == 0) {
		int dev = 0;
		PRT_TRY(hipGetDevice(&dev));
		PRT_TRY(hipDeviceGetAttribute(&s.n_cus, hipDeviceAttributeMultiprocessorCount, dev));
	}
	const int B = p.block, S = p.lds_stack;
	// Automatic: about 6 rays per lane, 16 to 32 waves per CU. A launch ends when each wave's
	// slowest last ray does; with few rays per lane that end weighs more than the latency the
	// extra waves hide (1 Mi rays of the random soup: 0.415 ms at 16 waves per CU against 0.459
	// at 32; 8 Mi room rays: 32 is fastest).
	int wpc = p.waves_per_cu;
	if (wpc <= 0) {
		const uint64_t rays_per_wave_per_cu = static_cast<uint64_t>(s.n_cus) * 64u * 6u;  // 6 rays a lane
		wpc = static_cast<int>(
		    std::min<uint64_t>(32, std::max<uint64_t>(16, (n + rays_per_wave_per_cu - 1) / rays_per_wave_per_cu)));
	}
	const uint64_t waves = static_cast<uint64_t>(s.n_cus) * static_cast<uint64_t>(wpc);
	uint64_t grid64 = (waves * 64 + B - 1) / B;
	const uint64_t need_threads = (n + 63) / 64 * 64;
	grid64 = std::max<uint64_t>(1, std::min<uint64_t>(grid64, (need_threads + B - 1) / B));
	pl.grid = static_cast<uint32_t>(grid64);
	// stack bound: each visited node pushes at most 7 entries, so 7 * depth + 1 suffice
	pl.max_entries = 7u * std::max(1u, b.depth) + 8u;
	pl.spill_depth = pl.max_entries > static_cast<uint32_t>(S) ? pl.max_entries - S : 1u;
	return hipSuccess;
}

// Launches the traversal with the given counter (8 x u64, cleared here by a kernel) and spill
// array (spill_depth entries per launched thread).
hipError_t run_trace(const Bvh8 &b, const TraceParams &p, const float *rays, uint64_t n, uint32_t mask, void *out,
                     hipStream_t st, uint32_t *counter, uint2 *spill, const TracePlan &pl, float *ms,
                     TraceStats *stats, const SeedArgs *seeds, const BatchArgs *batch) {
	const int B = p.block, S = p.lds_stack;
	KArgs a;
	a.nodes = p.node_format == 0 ? b.nodes_f.ptr : b.nodes_q.ptr;
	a.nodes_q = b.nodes_q.as<NodeQ>();
	a.any_ncull = b.n_ncull != 0u ? 1 : 0;
	a.tris = b.tris.as<TriRec>();
	a.rays = rays;
	a.out = static_cast<uint8_t *>(out);
	a.counter = counter;
	a.spill = spill;
	a.n_rays = n;
	a.n_nodes = b.n_nodes;
	a.n_tris = static_cast<uint32_t>(b.n_tris);
	a.max_entries = pl.max_entries;
	a.stride = static_cast<uint32_t>(layout_of(mask).size);
	a.mask = mask;
	a.spill_depth = pl.spill_depth;
	a.chunk = static_cast<uint32_t>(std::max(64, std::min(1 << 16, p.chunk)));
	a.leaf_phase = static_cast<uint32_t>(std::max(1, std::min(64, p.leaf_phase)));
	a.refill = static_cast<uint32_t>(std::max(1, std::min(64, p.refill)));
	a.margin_rel = p.margin_rel;
	a.margin_abs = p.margin_abs;
	a.leaf_test = p.leaf_test;
	// the watertight test's t is not covered by the culling guard (derived for Moller-Trumbore)
	a.cull = p.leaf_test == 1 ? 0 : p.cull;
	if (p.leaf_test == 1 && (seeds || batch)) return hipErrorInvalidValue;  // plain traversal only
	a.seeds = seeds ? seeds->seeds : nullptr;
	a.cl_entry = seeds ? seeds->entry : nullptr;
	a.cl_box = seeds ? seeds->box : nullptr;
	a.n_clusters = seeds ? seeds->count : 0u;
	a.n_seeds = std::max(1, std::min(4, p.n_seeds));
	a.tiles = batch ? batch->tiles : nullptr;
	a.n_tiles = batch ? batch->n_tiles : 0u;
	a.scene_info = batch ? batch->info : nullptr;
	if (batch && (seeds || b.n_nodes < batch->n_scenes || !batch->tiles || !batch->info)) return hipErrorInvalidValue;
	a.fast_scene = 1;
	for (int k = 0; k < 6; ++k) a.fast_scene &= std::fabs(b.scene_box[k]) < 0x1p60f ? 1 : 0;

	hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, st, reinterpret_cast<unsigned long long *>(a.counter), 8);
	hipEvent_t e0 = nullptr, e1 = nullptr;
	if (ms) {
		PRT_TRY(hipEventCreate(&e0));
		PRT_TRY(hipEventCreate(&e1));
		PRT_TRY(hipEventRecord(e0, st));
	}
	const bool count = p.count_stats != 0;
	hipError_t e = p.node_format == 0 ? launch_b<NodeF>(a, pl.grid, B, S, count, st)
	                                  : launch_b<NodeQ>(a, pl.grid, B, S, count, st);
	PRT_TRY(e);
	if (ms) {
		PRT_TRY(hipEventRecord(e1, st));
		PRT_TRY(hipEventSynchronize(e1));
		PRT_TRY(hipEventElapsedTime(ms, e0, e1));
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
	}
	if (stats) {
		unsigned long long h[8];
		PRT_TRY(hipMemcpyAsync(h, a.counter, sizeof(h), hipMemcpyDeviceToHost, st));
		PRT_TRY(hipStreamSynchronize(st));
		stats->node_visits = h[1];
		stats->tri_tests = h[2];
		stats->leaf_visits = h[3];
		stats->errors = static_cast<uint32_t>(h[4]);
		stats->node_phases = h[5];
		stats->tri_phases = h[6];
		stats->culled_pops = h[7];
	}
	return hipSuccess;
}

}  // namespace

hipError_t trace_bvh8(const Bvh8 &b, TraceScratch &s, const TraceParams &p, const float *rays, uint64_t n,
                      uint32_t mask, void *out, hipStream_t st, float *ms, TraceStats *stats, const SeedArgs *seeds,
                      const BatchArgs *batch) {
	if (n == 0) return hipSuccess;
	TracePlan pl;
	PRT_TRY(plan_trace(b, s, p, n, pl));
	PRT_TRY(s.counter.ensure(sizeof(unsigned long long) * 8));
	PRT_TRY(s.spill.ensure(sizeof(uint2) * static_cast<uint64_t>(pl.spill_depth) * pl.grid * p.block));
	return run_trace(b, p, rays, n, mask, out, st, s.counter.as<uint32_t>(), s.spill.as<uint2>(), pl, ms, stats, seeds,
	                 batch);
}

// ------------------------------------------------------------ traversal without allocation
//
// A device array of fixed size that the library never allocates or frees: the counter in its
// first 256 bytes, the stack spill after it. Every GPU holds its own copy for the life of the
// process; its address on each device is looked up once (hipGetSymbolAddress, with that device
// current) and only that address is passed to kernels or copies. All calls on one GPU share it:
// calls that overlap on one GPU (other streams, other contexts) race.
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch_setup(int device) {
	if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
	std::lock_guard<std::mutex> g(g_scratch_mu);
	if (g_scratch_ptr[device]) return hipSuccess;
	int cur = -1;
	PRT_TRY(hipGetDevice(&cur));
	if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current
	void *ptr = nullptr;
	PRT_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
	if (!ptr) return hipErrorInvalidSymbol;
	g_scratch_ptr[device] = ptr;
	return hipSuccess;
}

void *noalloc_scratch(int device) {
	if (device < 0 || device >= kMaxDevices) return nullptr;
	std::lock_guard<std::mutex> g(g_scratch_mu);
	return g_scratch_ptr[device];
}

hipError_t trace_bvh8_noalloc(const Bvh8 &b, TraceScratch &s, const TraceParams &p, const float *rays, uint64_t n,
                              uint32_t mask, void *out, hipStream_t st, void *scratch_ptr, uint64_t *need_bytes) {
	if (n == 0) return hipSuccess;
	if (!scratch_ptr) return hipErrorInvalidValue;
	TracePlan pl;
	PRT_TRY(plan_trace(b, s, p, n, pl));
	// as many workgroups as the spill part holds (the kernel is persistent: fewer waves trace
	// the same rays)
	const uint64_t per_group = sizeof(uint2) * static_cast<uint64_t>(pl.spill_depth) * p.block;
	const uint64_t fit = (kNoAllocScratchBytes - kNoAllocCounterBytes) / per_group;
	*need_bytes = kNoAllocCounterBytes + per_group;
	if (fit == 0) return hipErrorInvalidValue;
	pl.grid = static_cast<uint32_t>(std::min<uint64_t>(pl.grid, fit));
	uint8_t *base = static_cast<uint8_t *>(scratch_ptr);
	return run_trace(b, p, rays, n, mask, out, st, reinterpret_cast<uint32_t *>(base),
	                 reinterpret_cast<uint2 *>(base + kNoAllocCounterBytes), pl, nullptr, nullptr, nullptr, nullptr);
}

hipError_t trace_errors_at(TraceScratch &s, const uint32_t *counter, hipStream_t st, uint32_t *flags) {
	*flags = 0;
	if (!counter) return hipSuccess;
	if (!s.host_err) {
		hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&s.host_err), sizeof(uint32_t));
		if (e != hipSuccess) return e;
	}
	hipError_t e = hipMemcpyAsync(s.host_err, counter + 8, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
	if (e == hipSuccess) e = hipStreamSynchronize(st);
	if (e == hipSuccess) *flags = *s.host_err;
	return e;
}

#undef PRT_TRY

hipError_t trace_errors(TraceScratch &s, hipStream_t st, uint32_t *flags) {
	return trace_errors_at(s, s.counter.ptr ? s.counter.as<uint32_t>() : nullptr, st, flags);
}

}  // namespace prt
