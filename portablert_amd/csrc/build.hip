// GPU BVH8 build for set_tris (replaces the reference's host binned-SAH BVH2,
// include/portableRT/bvh.hpp:58-193, whose result does not matter for the nearest hit except
// for tie order).
//
// Pipeline, all on one stream:
//   1. k_tri_prep     triangle boxes, per-scene centroid bounds, finiteness check
//   2. k_morton       30-bit Morton code of the box centre (Karras 2012) below the scene index
//   3. radix sort     stable LSD, 8-bit digits, hand-written (hist / scan / scatter)
//   4. k_karras       binary radix tree over the sorted codes (Karras 2012), index tiebreak;
//                     the node spanning exactly one scene's keys is that scene's root
//   5. k_bottomup     boxes leaf to root, second arrival continues (counts come from the
//                     Karras ranges)
//   6. k_collapse     per BVH8 level: open the child of largest surface area until `width`
//                     children, subtrees of <= leaf_size triangles become leaves; writes fp32 and
//                     8-bit quantized nodes (bounds rounded outward and verified in fp32)
//   7. k_emit_tris    leaf triangle records in the order the nodes address them
//
// Several scenes build together: the scene index sits above the Morton bits of each key, so
// every scene is one subtree of the radix tree and BVH8 node s is the root of scene s.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"
#include "internal.h"

namespace prt {
namespace {

constexpr int kBlock = 256;
constexpr int kSortItems = 16;                       // keys per thread per radix tile
constexpr int kSortTile = kBlock * kSortItems;       // 4096 keys per tile
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kSceneShift = 30;                      // key = scene << 30 | 30-bit Morton code

// misc buffer layout (uint32 words)
enum : int {
	M_ERR = 0,         // bit 0 non-finite input, 1 triangle not placed, 2 levels exhausted
	M_NODES = 1,       // BVH8 node allocation counter
	M_TRIS = 2,        // leaf triangle slot counter
	M_LVL_BEGIN = 3,   // current level [begin, end)
	M_LVL_END = 4,
	M_DEPTH = 5,
	M_PM = 6,          // clustering: live clusters, and after a step the next count
	M_PMNEXT = 7,
	M_PNEXT = 8,       // clustering: internal node allocation counter
	M_PDONE = 9,       // clustering: finished (a step merged nothing)
	M_PFINAL = 10,     // clustering: which cluster buffer holds the roots
	M_NCULL = 11,      // triangles flagged by cull_unsafe (never culled by distance)
	M_SCENE = 12,      // per scene 12 words: ordered centroid min xyz, max xyz, box min, max
	M_PER_SCENE = 12
};
constexpr uint32_t misc_words(uint32_t scenes) { return M_SCENE + M_PER_SCENE * scenes; }

__device__ __forceinline__ uint32_t f2o(float f) {
	const uint32_t u = __float_as_uint(f);
	return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float o2f(uint32_t u) {
	return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

__global__ __launch_bounds__(kBlock) void k_init_misc(uint32_t *misc, uint32_t words) {
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if (i < words) {
		uint32_t v = 0;
		if (i >= M_SCENE) {
			const uint32_t k = (i - M_SCENE) % M_PER_SCENE;  // min words start +inf, max words -inf
			v = f2o(((k / 3) & 1u) ? -__builtin_inff() : __builtin_inff());
		}
		misc[i] = v;
	}
}

// Scene of input triangle i: last s with off[s] <= i (off has scenes + 1 entries).
__device__ __forceinline__ uint32_t scene_of(const uint32_t *off, uint32_t scenes, uint32_t i) {
	uint32_t lo = 0, hi = scenes;  // off[lo] <= i < off[hi]
	while (hi - lo > 1) {
		const uint32_t mid = (lo + hi) >> 1;
		if (off[mid] <= i) lo = mid;
		else hi = mid;
	}
	return lo;
}

__device__ __forceinline__ void scene_atomics(uint32_t *w, const float c[3], const float cmax[3], const float smn[3],
                                              const float smx[3]) {
#pragma unroll
	for (int k = 0; k < 3; ++k) {
		atomicMin(w + k, f2o(c[k]));
		atomicMax(w + 3 + k, f2o(cmax[k]));
		atomicMin(w + 6 + k, f2o(smn[k]));
		atomicMax(w + 9 + k, f2o(smx[k]));
	}
}

constexpr int kPrepBlock = 1024;

// Triangle boxes (the scene index in the w word of the lower corner, the never-cull flag in the
// w word of the upper corner) and per-scene centroid and
// vertex bounds. The bounds are reduced over the workgroup before the atomics when all of its
// triangles belong to one scene (one set of 12 atomics per 1024 triangles: same-address
// atomics serialise), else per lane.
__global__ __launch_bounds__(kPrepBlock) void k_tri_prep(const float *__restrict__ tris, uint32_t n,
                                                         const uint32_t *__restrict__ off, uint32_t scenes,
                                                         float4 *__restrict__ tbox, uint32_t *misc) {
	__shared__ float red[kPrepBlock / 64][12];
	__shared__ uint32_t mixed;
	const uint32_t i = blockIdx.x * kPrepBlock + threadIdx.x;
	const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
	if (threadIdx.x == 0) mixed = 0u;
	float c[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
	float cmax[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
	float smn[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
	float smx[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
	const uint32_t scene = (i < n && scenes > 1) ? scene_of(off, scenes, i) : 0u;
	if (i < n) {
		float v[9];
#pragma unroll
		for (int k = 0; k < 9; ++k) v[k] = tris[9ull * i + k];
		bool finite = true;
#pragma unroll
		for (int k = 0; k < 9; ++k) finite &= __builtin_isfinite(v[k]);
		if (!finite) atomicOr(misc + M_ERR, 1u);
		float lo[3], hi[3];
#pragma unroll
		for (int k = 0; k < 3; ++k) {
			lo[k] = fminf(v[k], fminf(v[3 + k], v[6 + k]));
			hi[k] = fmaxf(v[k], fmaxf(v[3 + k], v[6 + k]));
			c[k] = cmax[k] = 0.5f * lo[k] + 0.5f * hi[k];
			smn[k] = lo[k];
			smx[k] = hi[k];
		}
		tbox[2ull * i] = make_float4(lo[0], lo[1], lo[2], __uint_as_float(scene));
		// w of the upper corner: the distance-culling guard's flag (cull_unsafe, common.h)
		const bool unsafe = cull_unsafe(v);
		if (unsafe) atomicAdd(misc + M_NCULL, 1u);
		tbox[2ull * i + 1] = make_float4(hi[0], hi[1], hi[2], __uint_as_float(unsafe ? 1u : 0u));
	}
	// the block's first triangle decides the scene the block reduces for
	const uint32_t first = blockIdx.x * kPrepBlock;
	const uint32_t s0 = scenes > 1 ? scene_of(off, scenes, first) : 0u;
	__syncthreads();
	if (i < n && scene != s0) mixed = 1u;
	__syncthreads();
	const bool block_mixed = mixed != 0u;  // the block spans a scene boundary (rare)
	const uint32_t ws = __shfl(scene, 0);
	if (block_mixed && __ballot(i < n && scene != ws) != 0ull) {  // this wave spans it: per lane
		if (i < n) scene_atomics(misc + M_SCENE + M_PER_SCENE * scene, c, cmax, smn, smx);
		return;
	}
#pragma unroll
	for (int k = 0; k < 3; ++k) {
		for (int sh = 32; sh > 0; sh >>= 1) {
			c[k] = fminf(c[k], __shfl_xor(c[k], sh));
			cmax[k] = fmaxf(cmax[k], __shfl_xor(cmax[k], sh));
			smn[k] = fminf(smn[k], __shfl_xor(smn[k], sh));
			smx[k] = fmaxf(smx[k], __shfl_xor(smx[k], sh));
		}
	}
	if (block_mixed) {  // uniform wave in a mixed block: one set of atomics per wave
		if (lane == 0 && !__builtin_isnan(c[0])) scene_atomics(misc + M_SCENE + M_PER_SCENE * ws, c, cmax, smn, smx);
		return;
	}
	if (lane == 0) {
#pragma unroll
		for (int k = 0; k < 3; ++k) {
			red[wave][k] = c[k];
			red[wave][3 + k] = cmax[k];
			red[wave][6 + k] = smn[k];
			red[wave][9 + k] = smx[k];
		}
	}
	__syncthreads();
	if (threadIdx.x < 12) {
		const uint32_t k = threadIdx.x;
		const bool is_min = ((k / 3) & 1u) == 0u;  // centroid min, centroid max, box min, box max
		float r = red[0][k];
		for (int w = 1; w < kPrepBlock / 64; ++w) r = is_min ? fminf(r, red[w][k]) : fmaxf(r, red[w][k]);
		if (!__builtin_isnan(r)) {  // lanes past n contribute the identities +-inf
			uint32_t *dst = misc + M_SCENE + M_PER_SCENE * s0 + k;
			if (is_min) atomicMin(dst, f2o(r));
			else atomicMax(dst, f2o(r));
		}
	}
}

__device__ __forceinline__ uint32_t spread10(uint32_t x) {
	x &= 0x3FFu;
	x = (x | (x << 16)) & 0x030000FFu;
	x = (x | (x << 8)) & 0x0300F00Fu;
	x = (x | (x << 4)) & 0x030C30C3u;
	x = (x | (x << 2)) & 0x09249249u;
	return x;
}

__global__ __launch_bounds__(kBlock) void k_morton(const float4 *__restrict__ tbox, uint32_t n,
                                                   const uint32_t *misc, uint64_t *__restrict__ keys,
                                                   uint32_t *__restrict__ vals) {
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n) return;
	const float4 lo = tbox[2ull * i], hi = tbox[2ull * i + 1];
	const uint32_t scene = __float_as_uint(lo.w);
	const uint32_t *cb = misc + M_SCENE + M_PER_SCENE * scene;
	const float c[3] = {0.5f * lo.x + 0.5f * hi.x, 0.5f * lo.y + 0.5f * hi.y, 0.5f * lo.z + 0.5f * hi.z};
	uint32_t q[3];
#pragma unroll
	for (int k = 0; k < 3; ++k) {
		const float mn = o2f(cb[k]), mx = o2f(cb[3 + k]);
		const float ext = mx - mn;
		float x = ext > 0.0f ? (c[k] - mn) / ext : 0.0f;
		x = fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
		q[k] = static_cast<uint32_t>(x);
	}
	keys[i] = (static_cast<uint64_t>(scene) << kSceneShift) | (spread10(q[0]) << 2) | (spread10(q[1]) << 1) |
	          spread10(q[2]);
	vals[i] = i;
}

// ------------------------------------------------------------------------- radix sort

__global__ __launch_bounds__(kBlock) void k_hist(const uint64_t *__restrict__ keys, uint32_t n, int shift,
                                                 uint32_t *__restrict__ hist, uint32_t tiles) {
	__shared__ uint32_t h[256];
	h[threadIdx.x] = 0;
	__syncthreads();
	const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSortTile;
#pragma unroll 4
	for (int r = 0; r < kSortItems; ++r) {
		const uint64_t i = base + static_cast<uint64_t>(r) * kBlock + threadIdx.x;
		if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFFu], 1u);
	}
	__syncthreads();
	hist[static_cast<uint64_t>(threadIdx.x) * tiles + blockIdx.x] = h[threadIdx.x];
}

constexpr int kScanItems = 16;                 // histogram entries per thread
constexpr int kScanSeg = kBlock * kScanItems;  // entries per segment

// Exclusive scan of hist (digit-major: entry d * tiles + t) in two levels: each segment of
// 4096 entries is scanned in place by one workgroup and its total goes to seg[]; k_scan then
// scans seg[] and the scatter adds seg[entry / 4096].
__global__ __launch_bounds__(kBlock) void k_scan_seg(uint32_t *hist, uint32_t total, uint32_t *seg) {
	__shared__ uint32_t part[kBlock];
	const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanSeg + static_cast<uint64_t>(threadIdx.x) * kScanItems;
	uint32_t v[kScanItems];
	uint32_t s = 0;
#pragma unroll
	for (int k = 0; k < kScanItems; ++k) {
		v[k] = base + k < total ? hist[base + k] : 0u;
		s += v[k];
	}
	part[threadIdx.x] = s;
	__syncthreads();
	for (int sh = 1; sh < kBlock; sh <<= 1) {
		const uint32_t a = threadIdx.x >= static_cast<uint32_t>(sh) ? part[threadIdx.x - sh] : 0u;
		__syncthreads();
		part[threadIdx.x] += a;
		__syncthreads();
	}
	uint32_t run = part[threadIdx.x] - s;
#pragma unroll
	for (int k = 0; k < kScanItems; ++k) {
		if (base + k < total) hist[base + k] = run;
		run += v[k];
	}
	if (threadIdx.x == kBlock - 1) seg[blockIdx.x] = part[threadIdx.x];
}

// Exclusive scan of a short array in place, one workgroup.
__global__ __launch_bounds__(1024) void k_scan(uint32_t *hist, uint32_t total) {
	__shared__ uint32_t part[1024];
	const uint32_t per = (total + 1023) / 1024;
	const uint32_t b = threadIdx.x * per;
	const uint32_t e = min(total, b + per);
	uint32_t s = 0;
	for (uint32_t i = b; i < e; ++i) s += hist[i];
	part[threadIdx.x] = s;
	__syncthreads();
	for (int off = 1; off < 1024; off <<= 1) {
		const uint32_t v = threadIdx.x >= static_cast<uint32_t>(off) ? part[threadIdx.x - off] : 0u;
		__syncthreads();
		part[threadIdx.x] += v;
		__syncthreads();
	}
	uint32_t run = part[threadIdx.x] - s;
	for (uint32_t i = b; i < e; ++i) {
		const uint32_t x = hist[i];
		hist[i] = run;
		run += x;
	}
}

// Stable scatter through LDS. A tile's keys are loaded at once (kSortItems per thread), ranked
// in index order (round-major, then wave, then lane; equal digits matched within a wave with
// eight ballots) and placed in LDS in digit order; the tile is then written out so that
// consecutive threads write consecutive positions of a digit's run. The tile's count of a digit
// is the difference of two consecutive entries of the scanned histogram: it is digit-major, so
// the entry after a digit's last tile is the next digit's first.
__global__ __launch_bounds__(kBlock) void k_scatter(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                    uint64_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                    uint32_t n, int shift, const uint32_t *__restrict__ hist,
                                                    const uint32_t *__restrict__ seg, uint32_t tiles) {
	__shared__ uint64_t skey[kSortTile];
	__shared__ uint32_t sval[kSortTile];
	__shared__ uint32_t run[256];                // next tile position of each digit
	__shared__ uint32_t delta[256];              // output position minus tile position, per digit
	__shared__ uint8_t wcnt[kBlock / 64][256];  // keys of each digit in each wave of a round (<= 64)
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const uint32_t d = threadIdx.x;
	const uint64_t total = 256ull * tiles;
	const uint64_t e = static_cast<uint64_t>(d) * tiles + blockIdx.x;
	const uint32_t g0 = hist[e] + seg[e / kScanSeg];
	const uint32_t g1 = e + 1 < total ? hist[e + 1] + seg[(e + 1) / kScanSeg] : n;
	const uint32_t cnt = g1 - g0;
	const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSortTile;
	uint64_t kk[kSortItems];
	uint32_t vv[kSortItems];
#pragma unroll
	for (int r = 0; r < kSortItems; ++r) {
		const uint64_t i = base + static_cast<uint64_t>(r) * kBlock + threadIdx.x;
		kk[r] = i < n ? kin[i] : 0ull;
		vv[r] = i < n ? vin[i] : 0u;
	}
	// exclusive scan of the tile's digit counts: each digit's first tile position
	run[d] = cnt;
	__syncthreads();
	for (int sh = 1; sh < 256; sh <<= 1) {
		const uint32_t a = d >= static_cast<uint32_t>(sh) ? run[d - sh] : 0u;
		__syncthreads();
		run[d] += a;
		__syncthreads();
	}
	const uint32_t lofs = run[d] - cnt;
	run[d] = lofs;
	delta[d] = g0 - lofs;
	const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
	for (int r = 0; r < kSortItems; ++r) {
#pragma unroll
		for (int w = 0; w < kBlock / 64; ++w) wcnt[w][threadIdx.x] = 0;
		__syncthreads();
		const uint64_t i = base + static_cast<uint64_t>(r) * kBlock + threadIdx.x;
		const bool act = i < n;
		const uint64_t k = kk[r];
		const uint32_t digit = static_cast<uint32_t>(k >> shift) & 0xFFu;
		uint64_t peers = __ballot(act);
#pragma unroll
		for (int b = 0; b < 8; ++b) {
			const bool bit = (digit >> b) & 1u;
			const uint64_t m = __ballot(bit);
			peers &= bit ? m : ~m;
		}
		const uint32_t rank = __popcll(peers & lt);
		if (act && rank == 0) wcnt[wave][digit] = static_cast<uint8_t>(__popcll(peers));
		__syncthreads();
		if (act) {
			uint32_t pos = run[digit] + rank;
			for (int w = 0; w < wave; ++w) pos += wcnt[w][digit];
			skey[pos] = k;
			sval[pos] = vv[r];
		}
		__syncthreads();
		uint32_t add = 0;
#pragma unroll
		for (int w = 0; w < kBlock / 64; ++w) add += wcnt[w][threadIdx.x];
		run[threadIdx.x] += add;
	}
	__syncthreads();
	const uint32_t m = static_cast<uint32_t>(min(static_cast<uint64_t>(kSortTile), n - base));
	for (uint32_t j = threadIdx.x; j < m; j += kBlock) {
		const uint64_t k = skey[j];
		const uint32_t g = delta[static_cast<uint32_t>(k >> shift) & 0xFFu] + j;
		kout[g] = k;
		vout[g] = sval[j];
	}
}

// ------------------------------------------------------------------------ Karras 2012

struct BinTree {
	uint32_t *left, *right, *parent, *first, *count, *flag;
	float4 *box;  // 2 per node: lo, hi
	uint32_t n;   // leaves; internal nodes 0..n-2, leaf j is node (n-1)+j
};

__device__ __forceinline__ int delta(const uint64_t *keys, uint32_t n, int i, int j) {
	if (j < 0 || j >= static_cast<int>(n)) return -1;
	const uint64_t a = keys[i], b = keys[j];
	if (a != b) return __clzll(a ^ b);
	return 64 + __clz(static_cast<uint32_t>(i) ^ static_cast<uint32_t>(j));
}

__global__ __launch_bounds__(kBlock) void k_karras(const uint64_t *__restrict__ keys, BinTree t,
                                                   uint32_t *__restrict__ scene_root) {
	const int n = static_cast<int>(t.n);
	const int i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n - 1) return;
	const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
	const int dmin = delta(keys, n, i, i - d);
	int lmax = 2;
	while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
	int l = 0;
	for (int s = lmax >> 1; s >= 1; s >>= 1)
		if (delta(keys, n, i, i + (l + s) * d) > dmin) l += s;
	const int j = i + l * d;
	const int dnode = delta(keys, n, i, j);
	int s = 0, div = 2, step;
	do {
		step = (l + div - 1) / div;
		if (delta(keys, n, i, i + (s + step) * d) > dnode) s += step;
		div <<= 1;
	} while (step > 1);
	const int gamma = i + s * d + min(d, 0);
	const uint32_t lc = (min(i, j) == gamma) ? static_cast<uint32_t>(n - 1 + gamma) : static_cast<uint32_t>(gamma);
	const uint32_t rc = (max(i, j) == gamma + 1) ? static_cast<uint32_t>(n - 1 + gamma + 1) : static_cast<uint32_t>(gamma + 1);
	t.left[i] = lc;
	t.right[i] = rc;
	t.parent[lc] = static_cast<uint32_t>(i);
	t.parent[rc] = static_cast<uint32_t>(i);
	const int lo = min(i, j), hi = max(i, j);
	t.first[i] = static_cast<uint32_t>(lo);
	t.count[i] = static_cast<uint32_t>(hi - lo + 1);
	t.flag[i] = 0;
	if (i == 0) t.parent[0] = kNone;
	// the keys of one scene share the bits above the Morton code and no other key does, so
	// exactly one node spans them: the one whose range ends at both scene boundaries
	const uint64_t sc = keys[lo] >> kSceneShift;
	if ((keys[hi] >> kSceneShift) == sc && (lo == 0 || (keys[lo - 1] >> kSceneShift) != sc) &&
	    (hi == n - 1 || (keys[hi + 1] >> kSceneShift) != sc))
		scene_root[sc] = static_cast<uint32_t>(i);
}

// Scene roots before the radix tree: none for an empty scene, the leaf for a one-triangle
// scene; scenes of two or more triangles get their internal root from k_karras.
__global__ __launch_bounds__(kBlock) void k_scene_init(const uint32_t *__restrict__ off, uint32_t scenes, uint32_t n,
                                                       uint32_t *__restrict__ scene_root) {
	const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
	if (s >= scenes) return;
	const uint32_t b = scenes > 1 ? off[s] : 0u, e = scenes > 1 ? off[s + 1] : n;
	scene_root[s] = (e - b == 1u) ? (n - 1u + b) : kNone;
}

// Boxes that one thread writes and a thread of another workgroup reads in the same launch go
// through device-scope atomic 8-byte stores and loads, which are coherent across the XCDs'
// separate L2 caches without writing back a whole L2 as a device-scope fence does.
__device__ __forceinline__ uint64_t pack2(float a, float b) {
	return static_cast<uint64_t>(__float_as_uint(a)) | (static_cast<uint64_t>(__float_as_uint(b)) << 32);
}
__device__ __forceinline__ void st_box(float4 *box, uint32_t node, float4 lo, float4 hi) {
	uint64_t *p = reinterpret_cast<uint64_t *>(box + 2ull * node);
	__hip_atomic_store(p + 0, pack2(lo.x, lo.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	__hip_atomic_store(p + 1, pack2(lo.z, 0.0f), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	__hip_atomic_store(p + 2, pack2(hi.x, hi.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	__hip_atomic_store(p + 3, pack2(hi.z, hi.w), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ld_box(const float4 *box, uint32_t node, float4 &lo, float4 &hi) {
	const uint64_t *p = reinterpret_cast<const uint64_t *>(box + 2ull * node);
	uint64_t w[4];
#pragma unroll
	for (int k = 0; k < 4; ++k) w[k] = __hip_atomic_load(p + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	lo = make_float4(__uint_as_float(static_cast<uint32_t>(w[0])), __uint_as_float(static_cast<uint32_t>(w[0] >> 32)),
	                 __uint_as_float(static_cast<uint32_t>(w[1])), 0.0f);
	hi = make_float4(__uint_as_float(static_cast<uint32_t>(w[2])), __uint_as_float(static_cast<uint32_t>(w[2] >> 32)),
	                 __uint_as_float(static_cast<uint32_t>(w[3])), __uint_as_float(static_cast<uint32_t>(w[3] >> 32)));
}
// Waits until this lane's earlier stores are complete (device-scope atomic stores are then
// visible to every CU), so the flag increment that follows publishes them.
__device__ __forceinline__ void stores_done() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// ------------------------------------------- collapse costs (surface-area heuristic, SAH)
//
// Cost-optimal collapse of the binary tree into nodes of at most `width` children, by the
// dynamic program of Ylitie, Karras and Laine 2017 ("Efficient incoherent ray traversal on GPUs
// through compressed wide BVHs", section 3.1). For binary node x and i in 1..7 slots, C(x, i) is
// the least expected cost of x's subtree when it may fill i child slots of the node above it:
//   C(x, 1) = min(leaf(x), A(x) * kCostNode + split(x, width))
//   C(x, i) = min(C(x, i - 1), split(x, i))                               for i >= 2
//   split(x, i) = min over k in 1..i-1 of C(left(x), k) + C(right(x), i - k)
//   leaf(x) = A(x) * kCostTri * P(x) if x holds at most leaf_size triangles, else +inf
// with A the half surface area of x's box relative to its scene's box (computed in double, so
// coordinates near 1e19 or 1e-19 neither overflow nor vanish) and P the triangle count; a single
// triangle costs A * kCostTri in any number of slots. A row holds C(x, 1..7) and, last, 1 when x
// is best kept as one leaf. Rows of internal nodes are computed where the tree is built (at the
// merge for the clustered tree, in the bottom-up box pass for the radix tree); rows of binary
// leaves are recomputed from their boxes wherever they are needed.
constexpr float kCostNode = 1.0f;
constexpr float kCostTri = 0.6f;  // swept 0.2-1.5 on the room: 0.6 traces fastest
constexpr int kCostRow = 8;

struct CostCtx {
	float *row;             // kCostRow floats per binary internal node, or null: greedy collapse
	const double *scale;    // per scene: 1 / half-area of the scene box
	const uint64_t *keys;   // sorted keys: the scene of binary node b is keys[first[b]] >> 30
	const uint32_t *param;  // per scene leaf_size | width << 8, or null: leaf_size and width below
	int leaf_size;
	int width;
};

// Leaf size and node width of scene `scene` (per-scene parameters when given).
__device__ __forceinline__ void scene_shape(const CostCtx &c, uint32_t scene, int &ls, int &w) {
	int leaf_size = c.leaf_size, width = c.width;
	if (c.param) {
		const uint32_t pr = c.param[scene];
		leaf_size = static_cast<int>(pr & 0xFFu);
		width = static_cast<int>((pr >> 8) & 0xFFu);
	}
	ls = max(1, min(kMaxLeaf, leaf_size));
	w = max(2, min(8, width));
}

// Scene of binary node x (leaf j holds sorted key j; an internal node, its first key's).
__device__ __forceinline__ uint32_t scene_of(const BinTree &t, const CostCtx &c, uint32_t x) {
	const uint32_t k = x >= t.n - 1 ? x - (t.n - 1) : t.first[x];
	return static_cast<uint32_t>(c.keys[k] >> kSceneShift);
}

// Half-area of [lo, hi] relative to the scene's.
__device__ __forceinline__ float rel_area(float4 lo, float4 hi, double scale) {
	const double dx = static_cast<double>(hi.x) - lo.x, dy = static_cast<double>(hi.y) - lo.y,
	             dz = static_cast<double>(hi.z) - lo.z;
	return static_cast<float>((dx * dy + dx * dz + dy * dz) * scale);
}

__device__ __forceinline__ void leaf_row(float area, float r[kCostRow]) {
#pragma unroll
	for (int k = 0; k < kCostRow - 1; ++k) r[k] = area * kCostTri;
	r[kCostRow - 1] = 1.0f;
}

// split(x, i) from the rows of x's children, and the k that attains it (the smallest on ties).
__device__ __forceinline__ float best_split(const float *cl, const float *cr, int i, int &kbest) {
	float best = cl[0] + cr[i - 2];
	kbest = 1;
	for (int k = 2; k < i; ++k) {
		const float v = cl[k - 1] + cr[i - k - 1];
		if (v < best) {
			best = v;
			kbest = k;
		}
	}
	return best;
}

// Row of an internal node from its area, triangle count, shape and its children's rows.
__device__ __forceinline__ void node_row(float area, uint32_t count, int ls, int w, const float cl[kCostRow],
                                         const float cr[kCostRow], float c[kCostRow]) {
	int kb;
	const bool can_leaf = count <= static_cast<uint32_t>(ls);
	const float as_leaf = can_leaf ? area * kCostTri * static_cast<float>(count) : __builtin_inff();
	const float as_node = area * kCostNode + best_split(cl, cr, w, kb);
	const bool leaf_best = can_leaf && as_leaf <= as_node;
	c[0] = leaf_best ? as_leaf : as_node;
	for (int i = 2; i < kCostRow; ++i) c[i - 1] = fminf(c[i - 2], best_split(cl, cr, i, kb));
	c[kCostRow - 1] = leaf_best ? 1.0f : 0.0f;
}

// Rows written by one thread and read by a thread of another workgroup in the same launch go
// through device-scope atomic 8-byte stores and loads, as the boxes do.
__device__ __forceinline__ void st_row(float *rows, uint32_t x, const float c[kCostRow]) {
	uint64_t *p = reinterpret_cast<uint64_t *>(rows + static_cast<uint64_t>(kCostRow) * x);
#pragma unroll
	for (int k = 0; k < kCostRow / 2; ++k)
		__hip_atomic_store(p + k, pack2(c[2 * k], c[2 * k + 1]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ld_row(const float *rows, uint32_t x, float c[kCostRow]) {
	const uint64_t *p = reinterpret_cast<const uint64_t *>(rows + static_cast<uint64_t>(kCostRow) * x);
#pragma unroll
	for (int k = 0; k < kCostRow / 2; ++k) {
		const uint64_t v = __hip_atomic_load(p + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		c[2 * k] = __uint_as_float(static_cast<uint32_t>(v));
		c[2 * k + 1] = __uint_as_float(static_cast<uint32_t>(v >> 32));
	}
}

// Row of binary node x written in an earlier launch (plain loads), or computed for a leaf.
__device__ __forceinline__ void row_of(const BinTree &t, const CostCtx &c, uint32_t x, float r[kCostRow]) {
	if (x >= t.n - 1) {
		leaf_row(rel_area(t.box[2ull * x], t.box[2ull * x + 1], c.scale[scene_of(t, c, x)]), r);
		return;
	}
	const float4 *p = reinterpret_cast<const float4 *>(c.row + static_cast<uint64_t>(kCostRow) * x);
	const float4 a = p[0], b = p[1];
	r[0] = a.x; r[1] = a.y; r[2] = a.z; r[3] = a.w;
	r[4] = b.x; r[5] = b.y; r[6] = b.z; r[7] = b.w;
}

// 1 / half-area of every scene's box (1 for an empty or flat scene).
__global__ __launch_bounds__(kBlock) void k_scene_scale(const uint32_t *misc, uint32_t scenes, double *scale) {
	const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
	if (s >= scenes) return;
	const uint32_t *b = misc + M_SCENE + M_PER_SCENE * s + 6;
	const double dx = static_cast<double>(o2f(b[3])) - o2f(b[0]), dy = static_cast<double>(o2f(b[4])) - o2f(b[1]),
	             dz = static_cast<double>(o2f(b[5])) - o2f(b[2]);
	const double a = dx * dy + dx * dz + dy * dz;
	scale[s] = (a > 0.0 && a < 1e300) ? 1.0 / a : 1.0;
}

// Boxes (and collapse-cost rows) from the leaves to the roots: each leaf's thread climbs, and at
// every internal node the second of its two children to arrive continues. A node whose leaf range
// lies inside the workgroup's own leaves (most of them: a radix-tree node's range starts or ends
// at its own index) is met by two threads of this workgroup, so its arrival counter and its
// children's boxes and rows live in LDS; only nodes spanning two workgroups use the global
// counters, which every XCD reaches through device-coherent atomics. Every node's box and row is
// also written to global memory for the collapse. Compiled once per cost mode: without the rows
// a workgroup holds 17 KB of LDS instead of 25 KB, so eight fit on a CU instead of six.
template <bool COST>
__global__ __launch_bounds__(kBlock) void k_bottomup(const float4 *__restrict__ tbox,
                                                     const uint32_t *__restrict__ sorted, BinTree t, CostCtx cc) {
	// slots 0..kBlock-1: internal nodes b0 + s; kBlock..2 kBlock-1: leaves of sorted positions b0 + s
	__shared__ uint32_t lflag[kBlock];
	__shared__ float4 lbox[2 * kBlock][2];
	__shared__ float lrow[COST ? kBlock : 1][kCostRow];
	const uint32_t b0 = blockIdx.x * kBlock;
	const uint32_t b1 = min(b0 + kBlock, t.n) - 1;  // last leaf position of this workgroup
	lflag[threadIdx.x] = 0u;
	__syncthreads();
	const uint32_t j = b0 + threadIdx.x;
	if (j >= t.n) return;
	uint32_t node = t.n - 1 + j;
	const uint32_t tri = sorted[j];
	float4 lo = tbox[2ull * tri], hi = tbox[2ull * tri + 1];
	lo.w = 0.0f;  // hi.w: the never-cull flag, ORed up the tree with the boxes
	st_box(t.box, node, lo, hi);
	lbox[kBlock + threadIdx.x][0] = lo;
	lbox[kBlock + threadIdx.x][1] = hi;
	t.count[node] = 1;
	t.first[node] = j;
	float c[kCostRow];  // this node's collapse row (when costs are on)
	if (COST) leaf_row(rel_area(lo, hi, cc.scale[scene_of(t, cc, node)]), c);
	uint32_t p = (t.n > 1) ? t.parent[node] : kNone;
	// the parent's links (range, children, its own parent); no thread of this kernel writes them,
	// so the next level's are fetched before this level's arrival counter and the climb waits
	// on the counters and the sibling boxes only
	uint32_t pf = 0, pc = 0, pl = 0, pr = 0, pp = kNone;
	if (p != kNone) {
		pf = t.first[p];
		pc = t.count[p];
		pl = t.left[p];
		pr = t.right[p];
		pp = t.parent[p];
	}
	// a radix tree over at most 2^29 keys of at most 46 bits is at most 46 + 29 levels deep
	for (int guard = 0; p != kNone && guard < 128; ++guard) {
		const bool local = pf >= b0 && pf + pc - 1u <= b1;
		const bool left = pl == node;
		const uint32_t sib = left ? pr : pl;
		// this thread's box (and row) stores complete before the device-scope count (the wait
		// is for every load too, so it comes before the next level's links are fetched)
		if (!local) stores_done();
		uint32_t qf = 0, qc = 0, ql = 0, qr = 0, qp = kNone;
		if (pp != kNone) {
			qf = t.first[pp];
			qc = t.count[pp];
			ql = t.left[pp];
			qr = t.right[pp];
			qp = t.parent[pp];
		}
		float4 sl, sh;
		if (local) {
			// this thread's box (and row) are in LDS already; the release orders them before the
			// count, the acquire the sibling's before the reads below
			if (__hip_atomic_fetch_add(&lflag[p - b0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u) return;
			const uint32_t ss = sib >= t.n - 1 ? kBlock + (sib - (t.n - 1) - b0) : sib - b0;
			sl = lbox[ss][0];
			sh = lbox[ss][1];
		} else {
			if (atomicAdd(&t.flag[p], 1u) == 0u) return;  // sibling not finished: it continues
			ld_box(t.box, sib, sl, sh);
		}
		lo.x = fminf(lo.x, sl.x);
		lo.y = fminf(lo.y, sl.y);
		lo.z = fminf(lo.z, sl.z);
		hi.x = fmaxf(hi.x, sh.x);
		hi.y = fmaxf(hi.y, sh.y);
		hi.z = fmaxf(hi.z, sh.z);
		hi.w = __uint_as_float(__float_as_uint(hi.w) | __float_as_uint(sh.w));
		st_box(t.box, p, lo, hi);
		if (COST) {
			float cs[kCostRow], cp[kCostRow];
			if (sib >= t.n - 1) leaf_row(rel_area(sl, sh, cc.scale[scene_of(t, cc, sib)]), cs);
			else if (local) {
#pragma unroll
				for (int k = 0; k < kCostRow; ++k) cs[k] = lrow[sib - b0][k];
			} else {
				ld_row(cc.row, sib, cs);
			}
			const uint32_t scene = scene_of(t, cc, p);
			int ls, w;
			scene_shape(cc, scene, ls, w);
			node_row(rel_area(lo, hi, cc.scale[scene]), pc, ls, w, left ? c : cs, left ? cs : c, cp);
#pragma unroll
			for (int k = 0; k < kCostRow; ++k) c[k] = cp[k];
			st_row(cc.row, p, c);
			if (local) {
#pragma unroll
				for (int k = 0; k < kCostRow; ++k) lrow[p - b0][k] = c[k];
			}
		}
		if (local) {
			lbox[p - b0][0] = lo;
			lbox[p - b0][1] = hi;
		}
		node = p;
		p = pp;
		pf = qf;
		pc = qc;
		pl = ql;
		pr = qr;
		pp = qp;
	}
}

// ------------------------------------------------- clustered tree (PLOC), the other builder
//
// Bottom-up agglomerative clustering over the key order (Meister and Bittner 2018, "Parallel
// Locally-Ordered Clustering"): every cluster finds its nearest neighbour among the clusters
// within kPlocR positions (same scene only) by the surface area of their union; mutual nearest
// neighbours merge into a new internal node; an ordered compaction removes the merged partners;
// repeat until a step merges nothing (one cluster per scene). Boxes and counts are computed at
// the merge, so no bottom-up pass is needed. Node numbering as for the radix tree: internal
// nodes 0..n-2 (allocation order), leaf of sorted triangle j = n-1+j.

constexpr int kPlocR = 16;
constexpr int kPlocMaxSteps = 4096;

__global__ __launch_bounds__(kBlock) void k_ploc_init(const float4 *__restrict__ tbox, const uint32_t *__restrict__ sorted,
                                                      BinTree t, uint32_t *__restrict__ clu, uint32_t *misc) {
	const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
	if (j == 0) {
		misc[M_PM] = t.n;
		misc[M_PMNEXT] = t.n;
		misc[M_PNEXT] = 0u;
		misc[M_PDONE] = 0u;
		misc[M_PFINAL] = 0u;
	}
	if (j >= t.n) return;
	const uint32_t node = t.n - 1 + j, tri = sorted[j];
	t.box[2ull * node] = tbox[2ull * tri];  // w: the scene index
	t.box[2ull * node + 1] = tbox[2ull * tri + 1];
	t.count[node] = 1u;
	t.first[node] = j;
	clu[j] = node;
}

// Nearest neighbour within +-kPlocR by (area of the union, lower index, higher index): a
// symmetric order, so a scene's smallest pair is always mutual and every step merges.
__global__ __launch_bounds__(kBlock) void k_ploc_nn(const uint32_t *__restrict__ clu, BinTree t, const uint32_t *misc,
                                                    int32_t *__restrict__ nn) {
	__shared__ float4 slo[kBlock + 2 * kPlocR], shi[kBlock + 2 * kPlocR];
	if (misc[M_PDONE]) return;
	const int m = static_cast<int>(misc[M_PM]);
	const int base = static_cast<int>(blockIdx.x) * kBlock - kPlocR;
	for (int k = threadIdx.x; k < kBlock + 2 * kPlocR; k += kBlock) {
		const int idx = base + k;
		if (idx >= 0 && idx < m) {
			const uint32_t node = clu[idx];
			slo[k] = t.box[2ull * node];
			shi[k] = t.box[2ull * node + 1];
		}
	}
	__syncthreads();
	const int i = static_cast<int>(blockIdx.x) * kBlock + static_cast<int>(threadIdx.x);
	if (i >= m) return;
	const int me = static_cast<int>(threadIdx.x) + kPlocR;
	const float4 a = slo[me], b = shi[me];
	float best = __builtin_inff();
	int bj = -1;
	for (int d = -kPlocR; d <= kPlocR; ++d) {
		const int j = i + d;
		if (d == 0 || j < 0 || j >= m) continue;
		const float4 c = slo[me + d], e = shi[me + d];
		if (__float_as_uint(c.w) != __float_as_uint(a.w)) continue;  // other scene
		const float dx = fmaxf(b.x, e.x) - fminf(a.x, c.x);
		const float dy = fmaxf(b.y, e.y) - fminf(a.y, c.y);
		const float dz = fmaxf(b.z, e.z) - fminf(a.z, c.z);
		const float ar = dx * dy + dx * dz + dy * dz;  // +inf for huge coordinates: ties then decide
		bool take = bj < 0 || ar < best;
		if (ar == best && bj >= 0) {
			const int lo = min(i, j), hi = max(i, j), blo = min(i, bj), bhi = max(i, bj);
			take = lo < blo || (lo == blo && hi < bhi);
		}
		if (take) {
			best = ar;
			bj = j;
		}
	}
	nn[i] = bj;
}

// Mutual pairs (i < j) merge at position i; the partner's position is dropped by the compaction.
__global__ __launch_bounds__(kBlock) void k_ploc_merge(const uint32_t *__restrict__ clu, const int32_t *__restrict__ nn,
                                                       BinTree t, uint32_t *__restrict__ clu_next,
                                                       uint32_t *__restrict__ flags, uint32_t n_flags, uint32_t *misc,
                                                       CostCtx cc) {
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if (i >= n_flags || misc[M_PDONE]) return;
	const uint32_t m = misc[M_PM];
	if (i >= m) {
		flags[i] = 0u;
		return;
	}
	const int32_t j = nn[i];
	uint32_t keep = 1u, node = clu[i];
	if (j >= 0 && nn[j] == static_cast<int32_t>(i)) {
		if (static_cast<int32_t>(i) < j) {
			const uint32_t id = atomicAdd(misc + M_PNEXT, 1u);
			if (id < t.n - 1) {
				const uint32_t l = node, r = clu[j];
				const float4 la = t.box[2ull * l], lb = t.box[2ull * l + 1];
				const float4 ra = t.box[2ull * r], rb = t.box[2ull * r + 1];
				t.box[2ull * id] = make_float4(fminf(la.x, ra.x), fminf(la.y, ra.y), fminf(la.z, ra.z), la.w);
				t.box[2ull * id + 1] = make_float4(fmaxf(lb.x, rb.x), fmaxf(lb.y, rb.y), fmaxf(lb.z, rb.z),
				                                   __uint_as_float(__float_as_uint(lb.w) | __float_as_uint(rb.w)));
				t.left[id] = l;
				t.right[id] = r;
				t.count[id] = t.count[l] + t.count[r];
				t.first[id] = t.first[l];
				node = id;
				if (cc.row) {
					// the children are leaves or merged in earlier steps: their rows are final
					float cl[kCostRow], cr[kCostRow], cp[kCostRow];
					row_of(t, cc, l, cl);
					row_of(t, cc, r, cr);
					const uint32_t scene = scene_of(t, cc, l);
					int ls, w;
					scene_shape(cc, scene, ls, w);
					const float4 plo = make_float4(fminf(la.x, ra.x), fminf(la.y, ra.y), fminf(la.z, ra.z), 0.0f);
					const float4 phi = make_float4(fmaxf(lb.x, rb.x), fmaxf(lb.y, rb.y), fmaxf(lb.z, rb.z), 0.0f);
					node_row(rel_area(plo, phi, cc.scale[scene]), t.count[l] + t.count[r], ls, w, cl, cr, cp);
					float4 *dst = reinterpret_cast<float4 *>(cc.row + static_cast<uint64_t>(kCostRow) * id);
					dst[0] = make_float4(cp[0], cp[1], cp[2], cp[3]);
					dst[1] = make_float4(cp[4], cp[5], cp[6], cp[7]);
				}
			} else {
				atomicOr(misc + M_ERR, 2u);  // more merges than internal nodes: never happens
			}
		} else {
			keep = 0u;
		}
	}
	clu_next[i] = node;
	flags[i] = keep;
}

// Ordered compaction (flags scanned in place: exclusive prefix per 4096-entry segment + seg[]).
__global__ __launch_bounds__(kBlock) void k_ploc_compact(const uint32_t *__restrict__ clu_next,
                                                         const int32_t *__restrict__ nn, const uint32_t *__restrict__ pre,
                                                         const uint32_t *__restrict__ seg, uint32_t *__restrict__ clu_out,
                                                         uint32_t *misc) {
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if (misc[M_PDONE]) return;
	const uint32_t m = misc[M_PM];
	if (i >= m) return;
	const int32_t j = nn[i];
	const bool keep = !(j >= 0 && nn[j] == static_cast<int32_t>(i) && static_cast<int32_t>(i) > j);
	const uint32_t pos = pre[i] + seg[i / kScanSeg];
	if (keep) clu_out[pos] = clu_next[i];
	if (i == m - 1) misc[M_PMNEXT] = pos + (keep ? 1u : 0u);
}

// End of a step: the next cluster count; a step that merged nothing ends the clustering, and
// the buffer it compacted into (parity) holds the roots.
__global__ void k_ploc_step(uint32_t *misc, uint32_t parity_out) {
	if (threadIdx.x != 0 || misc[M_PDONE]) return;
	const uint32_t m = misc[M_PM], mn = misc[M_PMNEXT];
	misc[M_PFINAL] = parity_out;
	misc[M_PM] = mn;
	if (mn == m) misc[M_PDONE] = 1u;
}

// One cluster per non-empty scene remains: it is the scene's root.
__global__ __launch_bounds__(kBlock) void k_ploc_roots(const uint32_t *__restrict__ c0, const uint32_t *__restrict__ c1,
                                                       BinTree t, uint32_t *__restrict__ scene_root, uint32_t scenes,
                                                       uint32_t *misc) {
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if (i == 0 && !misc[M_PDONE]) atomicOr(misc + M_ERR, 2u);  // clustering did not finish
	if (i >= misc[M_PM]) return;
	const uint32_t node = (misc[M_PFINAL] ? c1 : c0)[i];
	const uint32_t sc = __float_as_uint(t.box[2ull * node].w);
	if (sc >= scenes || atomicExch(scene_root + sc, node) != kNone) atomicOr(misc + M_ERR, 2u);
}

// ------------------------------------------------------------------------- BVH8 collapse

__device__ __forceinline__ float half_area(float4 lo, float4 hi) {
	const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
	return dx * dy + dx * dz + dy * dz;
}

// Quantizes [clo, chi] against a grid at p with step s = 2^e; both bounds round outward,
// checked with the same fmaf the traversal evaluates.
__device__ __forceinline__ void quantize(float clo, float chi, float p, float s, uint8_t &qlo, uint8_t &qhi) {
	int lo = static_cast<int>(floorf((clo - p) / s));
	lo = max(0, min(255, lo));
	while (lo > 0 && fmaf(static_cast<float>(lo), s, p) > clo) --lo;
	int hi = static_cast<int>(ceilf((chi - p) / s));
	hi = max(0, min(255, hi));
	while (hi < 255 && fmaf(static_cast<float>(hi), s, p) < chi) ++hi;
	qlo = static_cast<uint8_t>(lo);
	qhi = static_cast<uint8_t>(hi);
}

// Smallest e in [-126, 127] with p + 255 * 2^e >= top in fp32.
__device__ __forceinline__ int grid_exp(float p, float top) {
	const float ext = top - p;
	int e = -126;
	if (ext > 0.0f) {
		int ex;
		(void)frexpf(ext / 255.0f, &ex);  // ext/255 in [2^(ex-1), 2^ex)
		e = max(-126, min(127, ex - 1));
	}
	while (e < 127 && fmaf(255.0f, exp2i(e), p) < top) ++e;
	return e;
}

struct CollapseArgs {
	BinTree t;
	uint32_t *misc;
	uint32_t *bin_of;
	uint32_t *tri_dest;
	NodeQ *nq;
	NodeF *nf;
	CostCtx cc;  // keys, per-scene shape; cc.row null: greedy collapse
};

__global__ __launch_bounds__(kBlock) void k_level_advance(uint32_t *misc) {
	if (threadIdx.x == 0) {
		const uint32_t end = misc[M_LVL_END];
		const uint32_t nodes = misc[M_NODES];
		misc[M_LVL_BEGIN] = end;
		misc[M_LVL_END] = nodes;
		if (nodes > end) misc[M_DEPTH] += 1;
	}
}

// Collapses BVH8 node i (binary node bin_of[i]): picks its children, allocates its internal
// children (their binary nodes go to bin_of) and its leaves' triangle slots, and writes both
// node formats. Returns {first internal child, number of internal children}.
template <bool COST>
__device__ __forceinline__ uint2 collapse_node(const CollapseArgs &a, uint32_t i) {
	const BinTree &t = a.t;
	const uint32_t n_int = t.n - 1;  // binary internal nodes are ids < n_int
	const uint32_t b = a.bin_of[i];  // kNone: root of an empty scene, written as an empty node
	uint32_t el[8];
	float ar[8];
	int ne;
	if (b == kNone) {
		ne = 0;
	} else if (b < n_int) {
		el[0] = t.left[b];
		el[1] = t.right[b];
		ne = 2;
	} else {
		el[0] = b;
		ne = 1;
	}
	int leaf_size, width;
	scene_shape(a.cc, b != kNone ? scene_of(t, a.cc, b) : 0u, leaf_size, width);
	// is_leaf bit k: child k becomes a leaf (by the cost rows, or by its size when greedy)
	uint32_t is_leaf = 0;
	if (COST && ne == 2) {
		// the width slots split between the two subtrees where that costs least, and each
		// part again wherever using fewer slots does not cost as little (at most 8 parts)
		float rl[kCostRow], rr[kCostRow], cx[kCostRow];
		uint32_t sn[8];
		int sj[8], sp = 0, k;
		row_of(t, a.cc, el[0], rl);
		row_of(t, a.cc, el[1], rr);
		best_split(rl, rr, width, k);
		sn[sp] = el[1];
		sj[sp++] = width - k;
		sn[sp] = el[0];
		sj[sp++] = k;
		ne = 0;
		while (sp > 0) {
			--sp;
			const uint32_t x = sn[sp];
			int slots = sj[sp];
			if (x >= n_int) {  // one triangle: a leaf
				is_leaf |= 1u << ne;
				el[ne++] = x;
				continue;
			}
			row_of(t, a.cc, x, cx);
			while (slots > 1 && cx[slots - 1] == cx[slots - 2]) --slots;
			if (slots == 1) {
				if (cx[kCostRow - 1] != 0.0f) is_leaf |= 1u << ne;
				el[ne++] = x;
				continue;
			}
			const uint32_t xl = t.left[x], xr = t.right[x];
			row_of(t, a.cc, xl, rl);
			row_of(t, a.cc, xr, rr);
			best_split(rl, rr, slots, k);
			sn[sp] = xr;
			sj[sp++] = slots - k;
			sn[sp] = xl;
			sj[sp++] = k;
		}
	} else if (ne == 1) {
		is_leaf = 1u;  // a one-triangle scene
	}
#pragma unroll
	for (int k = 0; k < 8; ++k) ar[k] = -1.0f;
	for (int k = 0; k < ne; ++k)
		if (el[k] < n_int) ar[k] = half_area(t.box[2ull * el[k]], t.box[2ull * el[k] + 1]);
	while (!COST && ne < width) {
		int best = -1;
		float ba = -1.0f;
#pragma unroll
		for (int k = 0; k < 8; ++k)
			if (k < ne && ar[k] > ba) {
				ba = ar[k];
				best = k;
			}
		if (best < 0) break;
		const uint32_t x = el[best];
		const uint32_t l = t.left[x], r = t.right[x];
		el[best] = l;
		ar[best] = (l < n_int) ? half_area(t.box[2ull * l], t.box[2ull * l + 1]) : -1.0f;
		el[ne] = r;
		ar[ne] = (r < n_int) ? half_area(t.box[2ull * r], t.box[2ull * r + 1]) : -1.0f;
		++ne;
	}
	// classify: greedy collapse makes subtrees of at most leaf_size triangles leaves
	uint32_t cnt[8];
	int n_inner = 0;
	uint32_t n_leaf_tris = 0;
	for (int k = 0; k < ne; ++k) {
		cnt[k] = t.count[el[k]];
		if (!COST && cnt[k] <= static_cast<uint32_t>(leaf_size)) is_leaf |= 1u << k;
		if (!((is_leaf >> k) & 1u)) ++n_inner;
		else n_leaf_tris += cnt[k];
	}
	const uint32_t child_base = n_inner ? atomicAdd(a.misc + M_NODES, static_cast<uint32_t>(n_inner)) : 0u;
	const uint32_t tri_base = n_leaf_tris ? atomicAdd(a.misc + M_TRIS, n_leaf_tris) : 0u;
	// node box = union of children
	float4 nlo = make_float4(__builtin_inff(), __builtin_inff(), __builtin_inff(), 0.0f);
	float4 nhi = make_float4(-__builtin_inff(), -__builtin_inff(), -__builtin_inff(), 0.0f);
	for (int k = 0; k < ne; ++k) {
		const float4 lo = t.box[2ull * el[k]], hi = t.box[2ull * el[k] + 1];
		nlo.x = fminf(nlo.x, lo.x);
		nlo.y = fminf(nlo.y, lo.y);
		nlo.z = fminf(nlo.z, lo.z);
		nhi.x = fmaxf(nhi.x, hi.x);
		nhi.y = fmaxf(nhi.y, hi.y);
		nhi.z = fmaxf(nhi.z, hi.z);
	}
	NodeQ q;
	NodeF f;
	if (ne == 0) nlo = nhi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	const float np[3] = {nlo.x, nlo.y, nlo.z}, nt[3] = {nhi.x, nhi.y, nhi.z};
	float sc[3];
#pragma unroll
	for (int ax = 0; ax < 3; ++ax) {
		const int e = grid_exp(np[ax], nt[ax]);
		q.p[ax] = np[ax];
		q.e[ax] = static_cast<int8_t>(e);
		sc[ax] = exp2i(e);
	}
	q.child_base = f.child_base = child_base;
	q.tri_base = f.tri_base = tri_base;
	uint8_t ncull = 0;  // slots whose subtree holds a triangle flagged by cull_unsafe
	uint32_t inner = 0, toff = 0;
#pragma unroll
	for (int k = 0; k < 8; ++k) {
		uint8_t meta = kMetaEmpty;
		float lo3[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
		float hi3[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
		if (k < ne) {
			const float4 lo = t.box[2ull * el[k]], hi = t.box[2ull * el[k] + 1];
			lo3[0] = lo.x; lo3[1] = lo.y; lo3[2] = lo.z;
			hi3[0] = hi.x; hi3[1] = hi.y; hi3[2] = hi.z;
			if (__float_as_uint(hi.w) != 0u) ncull |= static_cast<uint8_t>(1u << k);
			if (!((is_leaf >> k) & 1u)) {
				meta = static_cast<uint8_t>(kMetaInternal | inner);
				a.bin_of[child_base + inner] = el[k];
				++inner;
			} else {
				meta = meta_leaf(toff, cnt[k]);
				// the subtree's triangles in depth-first (left to right) order: a range of
				// sorted positions for the radix tree, any set of them for the clustered tree
				uint32_t stk[kMaxLeaf + 1];
				int sp = 0;
				stk[sp++] = el[k];
				uint32_t m = 0;
				while (sp > 0 && m < cnt[k]) {
					const uint32_t x = stk[--sp];
					if (x >= n_int) {
						a.tri_dest[x - n_int] = tri_base + toff + m;
						++m;
					} else if (sp + 2 <= kMaxLeaf + 1) {
						stk[sp++] = t.right[x];
						stk[sp++] = t.left[x];
					}
				}
				toff += cnt[k];
			}
		}
		q.meta[k] = f.meta[k] = meta;
#pragma unroll
		for (int ax = 0; ax < 3; ++ax) {
			f.lo[ax][k] = lo3[ax];
			f.hi[ax][k] = hi3[ax];
			if (k < ne) {
				quantize(lo3[ax], hi3[ax], np[ax], sc[ax], q.qlo[ax][k], q.qhi[ax][k]);
			} else {
				q.qlo[ax][k] = 255;
				q.qhi[ax][k] = 0;
			}
		}
	}
	q.ncull = ncull;
	a.nq[i] = q;
	a.nf[i] = f;
	return make_uint2(child_base, static_cast<uint32_t>(n_inner));
}

template <bool COST>
__global__ __launch_bounds__(kBlock) void k_collapse(CollapseArgs a) {
	const uint32_t begin = a.misc[M_LVL_BEGIN], end = a.misc[M_LVL_END];
	for (uint32_t i = begin + blockIdx.x * kBlock + threadIdx.x; i < end; i += gridDim.x * kBlock) collapse_node<COST>(a, i);
}

// The rest of the tree below the current level, depth first, one frontier node per thread: a
// fixed-length build (capturable) runs a few level launches for the wide top of the tree and
// then this one launch instead of a long tail of mostly empty level launches. The stack holds
// one {first node, nodes left} range per level; a deeper tree sets error bit 2.
constexpr int kDfsStack = kFixedDfsDepth;

template <bool COST>
__global__ __launch_bounds__(kBlock) void k_collapse_dfs(CollapseArgs a) {
	const uint32_t begin = a.misc[M_LVL_BEGIN], end = a.misc[M_LVL_END];
	for (uint32_t i = begin + blockIdx.x * kBlock + threadIdx.x; i < end; i += gridDim.x * kBlock) {
		uint2 st[kDfsStack];
		int sp = 0;
		st[sp++] = make_uint2(i, 1u);
		while (sp > 0) {
			const uint32_t node = st[sp - 1].x;
			if (st[sp - 1].y > 1u) st[sp - 1] = make_uint2(node + 1u, st[sp - 1].y - 1u);
			else --sp;
			const uint2 ch = collapse_node<COST>(a, node);
			if (ch.y == 0u) continue;
			if (sp == kDfsStack) {
				atomicOr(a.misc + M_ERR, 4u);
				break;
			}
			st[sp++] = ch;
		}
	}
}

// Leaf triangle records; ids are scene-local (input index minus the scene's first index).
__global__ __launch_bounds__(kBlock) void k_emit_tris(const float *__restrict__ tris, const uint32_t *__restrict__ sorted,
                                                      const uint64_t *__restrict__ keys, const uint32_t *__restrict__ off,
                                                      const uint32_t *__restrict__ dest, uint32_t n,
                                                      TriRec *__restrict__ out, uint32_t *misc) {
	const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
	if (j >= n) return;
	const uint32_t src = sorted[j];
	if (dest[j] >= n || src >= n) {  // a triangle not placed by the collapse: report, never write
		atomicOr(misc + M_ERR, 2u);
		return;
	}
	TriRec r;
#pragma unroll
	for (int k = 0; k < 9; ++k) r.v[k] = tris[9ull * src + k];
	r.id = off ? src - off[keys[j] >> kSceneShift] : src;
	r.pad[0] = r.pad[1] = 0;
	out[dest[j]] = r;
}

__global__ __launch_bounds__(kBlock) void k_fill(uint32_t *p, uint32_t n, uint32_t v) {
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if (i < n) p[i] = v;
}

// BVH8 level 0: node s is the root of scene s.
__global__ __launch_bounds__(kBlock) void k_roots(uint32_t *misc, uint32_t *bin_of, const uint32_t *scene_root,
                                                  uint32_t scenes) {
	const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
	if (s < scenes) bin_of[s] = scene_root[s];
	if (s == 0) {
		misc[M_NODES] = scenes;
		misc[M_LVL_BEGIN] = 0;
		misc[M_LVL_END] = 0;
		misc[M_DEPTH] = 0;
	}
}

// After a fixed number of levels: flag a build the levels did not finish.
__global__ void k_check_done(uint32_t *misc, uint32_t n) {
	if (threadIdx.x == 0) {
		if (misc[M_NODES] != misc[M_LVL_END]) atomicOr(misc + M_ERR, 4u);
		if (misc[M_TRIS] != n) atomicOr(misc + M_ERR, 2u);
	}
}

// Per-scene vertex bounds as floats: min xyz, max xyz.
__global__ __launch_bounds__(kBlock) void k_scene_boxes(const uint32_t *misc, uint32_t scenes, float *box) {
	const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
	if (i < 6 * scenes) box[i] = o2f(misc[M_SCENE + M_PER_SCENE * (i / 6) + 6 + i % 6]);
}

inline uint32_t blocks_for(uint64_t n, int per = kBlock) { return static_cast<uint32_t>((n + per - 1) / per); }

// ------------------------------------------------------------ refit: new vertices, same tree
//
// Every leaf record takes its primitive's new vertices; then every node's child boxes are
// recomputed from the deepest level up (children have larger indices than their parent, and
// a level's nodes are one contiguous index range), and the quantized boxes re-derived with the
// build's outward rounding. Topology, child order and leaf membership do not change.

__global__ __launch_bounds__(kBlock) void k_refit_tris(const float *__restrict__ tri_in, TriRec *__restrict__ tris,
                                                       uint32_t n, uint32_t *misc) {
	const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
	if (j >= n) return;
	TriRec r = tris[j];
	if (r.id >= n) {  // not a record of this tree: report, never read
		atomicOr(misc + M_ERR, 2u);
		return;
	}
	bool finite = true;
#pragma unroll
	for (int k = 0; k < 9; ++k) {
		r.v[k] = tri_in[9ull * r.id + k];
		finite &= __builtin_isfinite(r.v[k]);
	}
	tris[j] = r;
	if (!finite) atomicOr(misc + M_ERR, 1u);
	if (cull_unsafe(r.v)) atomicAdd(misc + M_NCULL, 1u);
}

struct RefitArgs {
	const TriRec *tris;
	NodeQ *nq;
	NodeF *nf;
	float4 *box;  // [2 * n_nodes] node boxes (lo, hi), written here for the parent level
	uint32_t n_nodes;
	uint32_t n_tris;
	uint32_t *misc;
};

__global__ __launch_bounds__(kBlock) void k_refit_level(RefitArgs a, uint32_t begin, uint32_t end) {
	for (uint32_t i = begin + blockIdx.x * kBlock + threadIdx.x; i < end; i += gridDim.x * kBlock) {
		NodeF f = a.nf[i];
		NodeQ q = a.nq[i];
		float nlo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
		float nhi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
		uint32_t ncull = 0;  // never-cull mask from the moved triangles (cull_unsafe)
#pragma unroll
		for (int k = 0; k < 8; ++k) {
			const uint32_t m = f.meta[k];
			float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
			float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
			if (meta_internal(m)) {
				const uint32_t c = f.child_base + meta_offset(m);
				if (c <= i || c >= a.n_nodes) {
					atomicOr(a.misc + M_ERR, 2u);
				} else {
					const float4 cl = a.box[2ull * c], ch = a.box[2ull * c + 1];
					lo[0] = cl.x; lo[1] = cl.y; lo[2] = cl.z;
					hi[0] = ch.x; hi[1] = ch.y; hi[2] = ch.z;
					if (__float_as_uint(ch.w) != 0u) ncull |= 1u << k;
				}
			} else if (meta_used(m)) {
				const uint32_t first = f.tri_base + meta_offset(m), cnt = meta_count(m);
				if (first + cnt > a.n_tris) {
					atomicOr(a.misc + M_ERR, 2u);
				} else {
					for (uint32_t t = first; t < first + cnt; ++t) {
						for (int v = 0; v < 3; ++v)
#pragma unroll
							for (int ax = 0; ax < 3; ++ax) {
								lo[ax] = fminf(lo[ax], a.tris[t].v[3 * v + ax]);
								hi[ax] = fmaxf(hi[ax], a.tris[t].v[3 * v + ax]);
							}
						if (cull_unsafe(a.tris[t].v)) ncull |= 1u << k;
					}
				}
			}
#pragma unroll
			for (int ax = 0; ax < 3; ++ax) {
				f.lo[ax][k] = lo[ax];
				f.hi[ax][k] = hi[ax];
				nlo[ax] = fminf(nlo[ax], lo[ax]);
				nhi[ax] = fmaxf(nhi[ax], hi[ax]);
			}
		}
		a.box[2ull * i] = make_float4(nlo[0], nlo[1], nlo[2], 0.0f);
		a.box[2ull * i + 1] = make_float4(nhi[0], nhi[1], nhi[2], __uint_as_float(ncull != 0u ? 1u : 0u));
		q.ncull = static_cast<uint8_t>(ncull);
		float sc[3];
#pragma unroll
		for (int ax = 0; ax < 3; ++ax) {
			const int e = grid_exp(nlo[ax], nhi[ax]);
			q.p[ax] = nlo[ax];
			q.e[ax] = static_cast<int8_t>(e);
			sc[ax] = exp2i(e);
		}
#pragma unroll
		for (int k = 0; k < 8; ++k) {
#pragma unroll
			for (int ax = 0; ax < 3; ++ax) {
				if (meta_used(f.meta[k])) {
					quantize(f.lo[ax][k], f.hi[ax][k], nlo[ax], sc[ax], q.qlo[ax][k], q.qhi[ax][k]);
				} else {
					q.qlo[ax][k] = 255;
					q.qhi[ax][k] = 0;
				}
			}
		}
		a.nf[i] = f;
		a.nq[i] = q;
	}
}

}  // namespace

hipError_t build_bvh8(BuildScratch &s, uint64_t n64, const BuildParams &p, Bvh8 &out, hipStream_t st,
                      bool *bad_input, double *build_ms, const SceneSet *scenes) {
	auto h2f = [](uint32_t u) {
		union {
			uint32_t u;
			float f;
		} c;
		c.u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
		return c.f;
	};
#define PRT_TRY(x)                                                                                 \
	do {                                                                                           \
		hipError_t e_ = (x);                                                                       \
		if (e_ != hipSuccess) return e_;                                                           \
	} while (0)
	SceneSet one;
	const SceneSet &ss = scenes ? *scenes : one;
	const uint32_t S = ss.count;
	const bool fixed = ss.levels > 0;  // no host synchronisation: capturable in a graph
	*bad_input = false;
	out.levels.clear();
	out.n_tris = n64;
	out.n_nodes = 0;
	out.n_ncull = 0;
	out.depth = 0;
	out.leaf_size = static_cast<uint32_t>(p.leaf_size);
	if (S == 0 || S > kMaxScenes || (S > 1 && !ss.tri_off) || (fixed && build_ms)) return hipErrorInvalidValue;
	if (n64 == 0) return hipSuccess;
	if (n64 > static_cast<uint64_t>(kTriMask)) return hipErrorInvalidValue;
	const uint32_t n = static_cast<uint32_t>(n64);
	const uint32_t nb = 2 * n - 1;
	const uint32_t cap = n + S;  // BVH8 nodes: at most one per binary internal node, one per scene root
	const uint32_t words = misc_words(S);
	const uint32_t *off = S > 1 ? ss.tri_off : nullptr;

	PRT_TRY(s.tbox.ensure(sizeof(float4) * 2 * n));
	for (int k = 0; k < 2; ++k) {
		PRT_TRY(s.keys[k].ensure(sizeof(uint64_t) * n));
		PRT_TRY(s.vals[k].ensure(sizeof(uint32_t) * n));
	}
	const uint32_t tiles = blocks_for(n, kSortTile);
	const uint32_t hist_n = 256 * tiles, n_seg = blocks_for(hist_n, kScanSeg);
	PRT_TRY(s.hist.ensure(sizeof(uint32_t) * (static_cast<uint64_t>(hist_n) + n_seg)));
	PRT_TRY(s.bin.ensure(sizeof(uint32_t) * (6ull * nb + 8)));
	PRT_TRY(s.bbox.ensure(sizeof(float4) * 2ull * nb));
	PRT_TRY(s.bin_of.ensure(sizeof(uint32_t) * cap));
	PRT_TRY(s.tri_dest.ensure(sizeof(uint32_t) * n));
	const uint64_t scale_words = (S + 1ull) & ~1ull;  // doubles, keeps the rows 16-byte aligned
	if (p.collapse == 1) PRT_TRY(s.cost.ensure(sizeof(double) * scale_words + sizeof(float) * kCostRow * n));
	PRT_TRY(s.misc.ensure(sizeof(uint32_t) * words));
	PRT_TRY(s.scene_root.ensure(sizeof(uint32_t) * S));
	if (!s.host_misc) PRT_TRY(hipHostMalloc(&s.host_misc, sizeof(uint32_t) * misc_words(1)));
	PRT_TRY(out.nodes_q.ensure(sizeof(NodeQ) * cap));
	PRT_TRY(out.nodes_f.ensure(sizeof(NodeF) * cap));
	PRT_TRY(out.tris.ensure(sizeof(TriRec) * n));

	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	if (build_ms) {  // device time of the build kernels (allocations above are not counted)
		PRT_TRY(hipEventCreate(&ev0));
		PRT_TRY(hipEventCreate(&ev1));
		PRT_TRY(hipEventRecord(ev0, st));
	}
	uint32_t *misc = s.misc.as<uint32_t>();
	const float *tri_in = ss.tris ? ss.tris : s.tri_in.as<float>();
	hipLaunchKernelGGL(k_init_misc, dim3(blocks_for(words)), dim3(kBlock), 0, st, misc, words);
	hipLaunchKernelGGL(k_tri_prep, dim3(blocks_for(n, kPrepBlock)), dim3(kPrepBlock), 0, st, tri_in, n, off, S, s.tbox.as<float4>(),
	                   misc);
	hipLaunchKernelGGL(k_morton, dim3(blocks_for(n)), dim3(kBlock), 0, st, s.tbox.as<float4>(), n, misc,
	                   s.keys[0].as<uint64_t>(), s.vals[0].as<uint32_t>());
	int key_bits = kSceneShift;
	while (S > 1 && (1ull << (key_bits - kSceneShift)) < S) ++key_bits;
	int cur = 0;
	for (int shift = 0; shift < key_bits; shift += 8) {
		hipLaunchKernelGGL(k_hist, dim3(tiles), dim3(kBlock), 0, st, s.keys[cur].as<uint64_t>(), n, shift,
		                   s.hist.as<uint32_t>(), tiles);
		uint32_t *seg = s.hist.as<uint32_t>() + hist_n;
		hipLaunchKernelGGL(k_scan_seg, dim3(n_seg), dim3(kBlock), 0, st, s.hist.as<uint32_t>(), hist_n, seg);
		hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, seg, n_seg);
		hipLaunchKernelGGL(k_scatter, dim3(tiles), dim3(kBlock), 0, st, s.keys[cur].as<uint64_t>(),
		                   s.vals[cur].as<uint32_t>(), s.keys[cur ^ 1].as<uint64_t>(), s.vals[cur ^ 1].as<uint32_t>(),
		                   n, shift, s.hist.as<uint32_t>(), seg, tiles);
		cur ^= 1;
	}
	const uint64_t *keys = s.keys[cur].as<uint64_t>();
	BinTree t;
	uint32_t *bw = s.bin.as<uint32_t>();
	t.left = bw;
	t.right = bw + nb;
	t.parent = bw + 2ull * nb;
	t.first = bw + 3ull * nb;
	t.count = bw + 4ull * nb;
	t.flag = bw + 5ull * nb;
	t.box = s.bbox.as<float4>();
	t.n = n;
	uint32_t *scene_root = s.scene_root.as<uint32_t>();
	uint32_t *hm = static_cast<uint32_t *>(s.host_misc);
	const int leaf_size = std::max(1, std::min(kMaxLeaf, p.leaf_size)), width = std::max(2, std::min(8, p.width));
	CostCtx cc;
	cc.row = nullptr;
	cc.scale = nullptr;
	cc.keys = keys;
	cc.param = ss.param;
	cc.leaf_size = leaf_size;
	cc.width = width;
	if (p.collapse == 1) {
		double *scale = s.cost.as<double>();
		cc.scale = scale;
		cc.row = reinterpret_cast<float *>(scale + scale_words);
		hipLaunchKernelGGL(k_scene_scale, dim3(blocks_for(S)), dim3(kBlock), 0, st, misc, S, scale);
	}
	if (p.builder == 1 && !fixed) {
		// clustering: steps until one merges nothing, the host checks every 8 steps
		const uint32_t n_seg_p = blocks_for(n, kScanSeg);
		PRT_TRY(s.ploc.ensure(sizeof(uint32_t) * (5ull * n + n_seg_p + 8)));
		uint32_t *pw = s.ploc.as<uint32_t>();
		uint32_t *clu0 = pw, *clu1 = pw + n, *flags = pw + 3ull * n, *clu_next = pw + 4ull * n, *segp = pw + 5ull * n;
		int32_t *nnb = reinterpret_cast<int32_t *>(pw + 2ull * n);
		hipLaunchKernelGGL(k_fill, dim3(blocks_for(S)), dim3(kBlock), 0, st, scene_root, S, kNone);
		hipLaunchKernelGGL(k_ploc_init, dim3(blocks_for(n)), dim3(kBlock), 0, st, s.tbox.as<float4>(),
		                   s.vals[cur].as<uint32_t>(), t, clu0, misc);
		bool done = false;
		for (int it = 0; it < kPlocMaxSteps && !done; ++it) {
			uint32_t *cin = (it & 1) ? clu1 : clu0, *cout = (it & 1) ? clu0 : clu1;
			hipLaunchKernelGGL(k_ploc_nn, dim3(blocks_for(n)), dim3(kBlock), 0, st, cin, t, misc, nnb);
			hipLaunchKernelGGL(k_ploc_merge, dim3(blocks_for(n)), dim3(kBlock), 0, st, cin, nnb, t, clu_next, flags, n,
			                   misc, cc);
			hipLaunchKernelGGL(k_scan_seg, dim3(n_seg_p), dim3(kBlock), 0, st, flags, n, segp);
			hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, segp, n_seg_p);
			hipLaunchKernelGGL(k_ploc_compact, dim3(blocks_for(n)), dim3(kBlock), 0, st, clu_next, nnb, flags, segp, cout,
			                   misc);
			hipLaunchKernelGGL(k_ploc_step, dim3(1), dim3(64), 0, st, misc, (it & 1) ? 0u : 1u);
			if ((it & 7) == 7) {
				PRT_TRY(hipMemcpyAsync(hm, misc, sizeof(uint32_t) * misc_words(1), hipMemcpyDeviceToHost, st));
				PRT_TRY(hipStreamSynchronize(st));
				done = hm[M_PDONE] != 0u;
			}
		}
		hipLaunchKernelGGL(k_ploc_roots, dim3(blocks_for(n)), dim3(kBlock), 0, st, clu0, clu1, t, scene_root, S, misc);
	} else {
		hipLaunchKernelGGL(k_scene_init, dim3(blocks_for(S)), dim3(kBlock), 0, st, off, S, n, scene_root);
		if (n > 1)
			hipLaunchKernelGGL(k_karras, dim3(blocks_for(n - 1)), dim3(kBlock), 0, st, keys, t, scene_root);
		if (cc.row)
			hipLaunchKernelGGL(k_bottomup<true>, dim3(blocks_for(n)), dim3(kBlock), 0, st, s.tbox.as<float4>(),
			                   s.vals[cur].as<uint32_t>(), t, cc);
		else
			hipLaunchKernelGGL(k_bottomup<false>, dim3(blocks_for(n)), dim3(kBlock), 0, st, s.tbox.as<float4>(),
			                   s.vals[cur].as<uint32_t>(), t, cc);
	}
	hipLaunchKernelGGL(k_fill, dim3(blocks_for(n)), dim3(kBlock), 0, st, s.tri_dest.as<uint32_t>(), n, kNone);
	hipLaunchKernelGGL(k_roots, dim3(blocks_for(S)), dim3(kBlock), 0, st, misc, s.bin_of.as<uint32_t>(), scene_root, S);
	CollapseArgs ca;
	ca.t = t;
	ca.misc = misc;
	ca.bin_of = s.bin_of.as<uint32_t>();
	ca.tri_dest = s.tri_dest.as<uint32_t>();
	ca.nq = out.nodes_q.as<NodeQ>();
	ca.nf = out.nodes_f.as<NodeF>();
	ca.cc = cc;
	// a fixed-length build: a few level launches, then the rest of the tree depth first
	const int max_levels = fixed ? std::min(ss.levels, kFixedTopLevels) : 128;
	const uint32_t grid = std::min<uint32_t>(blocks_for(cap), 2048);
	uint64_t level_nodes = S;  // most BVH8 nodes a level can hold: the roots, then 8 children each
	for (int level = 0; level < max_levels; ++level) {
		// a grid for that many nodes (the kernel's loop covers any count), so that the small top
		// levels do not dispatch thousands of idle waves
		const uint32_t lgrid = std::min<uint32_t>(grid, blocks_for(level_nodes));
		level_nodes = std::min<uint64_t>(level_nodes * 8u, cap);
		hipLaunchKernelGGL(k_level_advance, dim3(1), dim3(64), 0, st, misc);
		if (ca.cc.row) hipLaunchKernelGGL(k_collapse<true>, dim3(lgrid), dim3(kBlock), 0, st, ca);
		else hipLaunchKernelGGL(k_collapse<false>, dim3(lgrid), dim3(kBlock), 0, st, ca);
		if (!fixed && (level & 3) == 3) {
			PRT_TRY(hipMemcpyAsync(hm, misc, sizeof(uint32_t) * misc_words(1), hipMemcpyDeviceToHost, st));
			PRT_TRY(hipStreamSynchronize(st));
			// done when the level just processed created no new node
			if (hm[M_NODES] == hm[M_LVL_END]) break;
		}
	}
	if (fixed) {
		hipLaunchKernelGGL(k_level_advance, dim3(1), dim3(64), 0, st, misc);
		if (ca.cc.row) hipLaunchKernelGGL(k_collapse_dfs<true>, dim3(grid), dim3(kBlock), 0, st, ca);
		else hipLaunchKernelGGL(k_collapse_dfs<false>, dim3(grid), dim3(kBlock), 0, st, ca);
		hipLaunchKernelGGL(k_level_advance, dim3(1), dim3(64), 0, st, misc);  // every node is done
	}
	hipLaunchKernelGGL(k_emit_tris, dim3(blocks_for(n)), dim3(kBlock), 0, st, tri_in, s.vals[cur].as<uint32_t>(), keys,
	                   off, s.tri_dest.as<uint32_t>(), n, out.tris.as<TriRec>(), misc);
	if (ss.box) hipLaunchKernelGGL(k_scene_boxes, dim3(blocks_for(6ull * S)), dim3(kBlock), 0, st, misc, S, ss.box);
	if (fixed) {
		hipLaunchKernelGGL(k_check_done, dim3(1), dim3(64), 0, st, misc, n);
		PRT_TRY(hipGetLastError());
		out.n_nodes = cap;  // an upper bound: the device holds the count
		out.depth = static_cast<uint32_t>(ss.levels);
		out.n_ncull = kNcullUnknown;  // not read back: traversals handle never-cull slots
		return hipSuccess;
	}
	PRT_TRY(hipGetLastError());
	PRT_TRY(hipMemcpyAsync(hm, misc, sizeof(uint32_t) * misc_words(1), hipMemcpyDeviceToHost, st));
	if (build_ms) PRT_TRY(hipEventRecord(ev1, st));
	PRT_TRY(hipStreamSynchronize(st));
	if (build_ms) {
		float ms = 0.0f;
		PRT_TRY(hipEventElapsedTime(&ms, ev0, ev1));
		*build_ms = ms;
		(void)hipEventDestroy(ev0);
		(void)hipEventDestroy(ev1);
	}
	*bad_input = (hm[M_ERR] & 1u) != 0;
	if (*bad_input) return hipSuccess;  // the caller reports the input; the tree is not used
	if (hm[M_ERR] & 2u) return hipErrorUnknown;  // inconsistent build
	for (int k = 0; k < 6; ++k) out.scene_box[k] = h2f(hm[M_SCENE + 6 + k]);
	out.n_nodes = hm[M_NODES];
	out.depth = hm[M_DEPTH];
	out.n_ncull = hm[M_NCULL];
	if (hm[M_TRIS] != n || hm[M_NODES] != hm[M_LVL_END]) return hipErrorUnknown;  // incomplete build
	return hipSuccess;
#undef PRT_TRY
}

hipError_t refit_bvh8(BuildScratch &s, Bvh8 &b, hipStream_t st, bool *bad_input, double *refit_ms) {
#define PRT_TRY(x)                                                                                 \
	do {                                                                                           \
		hipError_t e_ = (x);                                                                       \
		if (e_ != hipSuccess) return e_;                                                           \
	} while (0)
	*bad_input = false;
	const uint64_t n = b.n_tris;
	if (n == 0 || b.n_nodes == 0) return hipSuccess;
	if (b.levels.empty()) {
		// level ranges from the node headers: level 0 is the root; the next level ends after the
		// last internal child of the current one
		std::vector<NodeF> h;
		try {
			h.resize(b.n_nodes);
		} catch (...) {
			return hipErrorOutOfMemory;
		}
		PRT_TRY(hipMemcpyAsync(h.data(), b.nodes_f.ptr, sizeof(NodeF) * b.n_nodes, hipMemcpyDeviceToHost, st));
		PRT_TRY(hipStreamSynchronize(st));
		uint32_t begin = 0, end = 1;
		while (begin < end) {
			b.levels.push_back(begin);
			uint32_t next = end;
			for (uint32_t i = begin; i < end; ++i)
				for (int k = 0; k < 8; ++k)
					if (meta_internal(h[i].meta[k]))
						next = std::max(next, h[i].child_base + meta_offset(h[i].meta[k]) + 1u);
			if (next > b.n_nodes) {
				b.levels.clear();
				return hipErrorUnknown;  // a child outside the tree
			}
			begin = end;
			end = next;
		}
		if (begin != b.n_nodes) {
			b.levels.clear();
			return hipErrorUnknown;  // nodes not reachable level by level
		}
	}
	PRT_TRY(s.refit_box.ensure(sizeof(float4) * 2ull * b.n_nodes));
	PRT_TRY(s.misc.ensure(sizeof(uint32_t) * misc_words(1)));
	if (!s.host_misc) PRT_TRY(hipHostMalloc(&s.host_misc, sizeof(uint32_t) * misc_words(1)));
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	if (refit_ms) {
		PRT_TRY(hipEventCreate(&ev0));
		PRT_TRY(hipEventCreate(&ev1));
		PRT_TRY(hipEventRecord(ev0, st));
	}
	uint32_t *misc = s.misc.as<uint32_t>();
	hipLaunchKernelGGL(k_init_misc, dim3(blocks_for(misc_words(1))), dim3(kBlock), 0, st, misc, misc_words(1));
	hipLaunchKernelGGL(k_refit_tris, dim3(blocks_for(n)), dim3(kBlock), 0, st, s.tri_in.as<float>(), b.tris.as<TriRec>(),
	                   static_cast<uint32_t>(n), misc);
	RefitArgs ra;
	ra.tris = b.tris.as<TriRec>();
	ra.nq = b.nodes_q.as<NodeQ>();
	ra.nf = b.nodes_f.as<NodeF>();
	ra.box = s.refit_box.as<float4>();
	ra.n_nodes = b.n_nodes;
	ra.n_tris = static_cast<uint32_t>(n);
	ra.misc = misc;
	for (size_t l = b.levels.size(); l-- > 0;) {
		const uint32_t begin = b.levels[l], end = l + 1 < b.levels.size() ? b.levels[l + 1] : b.n_nodes;
		hipLaunchKernelGGL(k_refit_level, dim3(std::min<uint32_t>(blocks_for(end - begin), 2048)), dim3(kBlock), 0, st, ra,
		                   begin, end);
	}
	PRT_TRY(hipGetLastError());
	if (refit_ms) PRT_TRY(hipEventRecord(ev1, st));
	float4 root[2];
	PRT_TRY(hipMemcpyAsync(root, ra.box, sizeof(root), hipMemcpyDeviceToHost, st));
	PRT_TRY(hipMemcpyAsync(s.host_misc, misc, sizeof(uint32_t) * misc_words(1), hipMemcpyDeviceToHost, st));
	PRT_TRY(hipStreamSynchronize(st));
	if (refit_ms) {
		float ms = 0.0f;
		PRT_TRY(hipEventElapsedTime(&ms, ev0, ev1));
		*refit_ms = ms;
		(void)hipEventDestroy(ev0);
		(void)hipEventDestroy(ev1);
	}
	const uint32_t err = static_cast<const uint32_t *>(s.host_misc)[M_ERR];
	*bad_input = (err & 1u) != 0;
	if (*bad_input) return hipSuccess;  // the caller reports the input; the tree is not used
	if (err & 2u) return hipErrorUnknown;
	b.n_ncull = static_cast<const uint32_t *>(s.host_misc)[M_NCULL];
	b.scene_box[0] = root[0].x;
	b.scene_box[1] = root[0].y;
	b.scene_box[2] = root[0].z;
	b.scene_box[3] = root[1].x;
	b.scene_box[4] = root[1].y;
	b.scene_box[5] = root[1].z;
	return hipSuccess;
#undef PRT_TRY
}

hipError_t build_errors(BuildScratch &s, hipStream_t st, uint32_t *flags) {
	*flags = 0;
	if (!s.misc.ptr || !s.host_misc) return hipSuccess;
	uint32_t *hm = static_cast<uint32_t *>(s.host_misc);
	hipError_t e = hipMemcpyAsync(hm, s.misc.ptr, sizeof(uint32_t) * misc_words(1), hipMemcpyDeviceToHost, st);
	if (e == hipSuccess) e = hipStreamSynchronize(st);
	if (e == hipSuccess) *flags = hm[M_ERR];
	return e;
}

}  // namespace prt
