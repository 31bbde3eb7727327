// Shared device-side definitions of the MI355X nearest-hit backend: BVH8 node formats, the
// leaf triangle record, the exact reference predicates and the hit-record writer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace prt {

// ----------------------------------------------------------------------------- formats

// Output field bits (filter tags of the reference, hitreg.hpp:48-54).
enum : uint32_t { F_UV = 1, F_T = 2, F_ID = 4, F_P = 8, F_VALID = 16 };

// Leaf triangle record, 48 bytes: the three vertices as given to set_tris and the
// primitive id. The exact leaf test reads only this record.
struct alignas(16) TriRec {
	float v[9];
	uint32_t id;
	uint32_t pad[2];
};
static_assert(sizeof(TriRec) == 48, "TriRec");

// meta byte of a BVH8 child slot: 0xFF empty; internal child: 0x80 | offset from child_base
// (offset < 8); leaf child: (offset from tri_base) << 2 | (count - 1), count in 1..kMaxLeaf,
// offset < 32. A slot's stack entry is then one select and one add (child_entry).
constexpr int kMaxLeaf = 4;
constexpr uint8_t kMetaEmpty = 0xFF;
constexpr uint8_t kMetaInternal = 0x80;
__host__ __device__ inline bool meta_used(uint32_t m) { return m != kMetaEmpty; }
__host__ __device__ inline bool meta_internal(uint32_t m) { return m != kMetaEmpty && (m & 0x80u) != 0u; }
__host__ __device__ inline uint32_t meta_count(uint32_t m) { return (m & 3u) + 1u; }
__host__ __device__ inline uint32_t meta_offset(uint32_t m) { return (m & 0x80u) ? (m & 0x7Fu) : (m >> 2); }
__host__ __device__ inline uint8_t meta_leaf(uint32_t offset, uint32_t count) {
	return static_cast<uint8_t>((offset << 2) | (count - 1u));
}

// 8-bit quantized BVH8 node, 80 bytes. Child k's box on axis a is
// [p_a + qlo_a[k] * 2^e_a, p_a + qhi_a[k] * 2^e_a], evaluated in fp32 as fmaf(q, 2^e, p);
// the builder verifies each bound rounds outward from the exact child box. Bit k of ncull:
// child k's subtree holds a triangle flagged by cull_unsafe (never culled by distance); it
// serves the fp32 node of the same index too.
struct alignas(16) NodeQ {
	float p[3];
	int8_t e[3];
	uint8_t ncull;
	uint32_t child_base;
	uint32_t tri_base;
	uint8_t meta[8];
	uint8_t qlo[3][8];
	uint8_t qhi[3][8];
};
static_assert(sizeof(NodeQ) == 80, "NodeQ");

// fp32 BVH8 node with the same topology, 208 bytes.
struct alignas(16) NodeF {
	uint32_t child_base;
	uint32_t tri_base;
	uint8_t meta[8];
	float lo[3][8];
	float hi[3][8];
};
static_assert(sizeof(NodeF) == 208, "NodeF");

// Stack / work entry: bit 31 set = leaf (bits 2..30 = first triangle record, bits 0..1 =
// count-1); clear = internal node index. Adding 3 to a leaf entry advances it by one triangle
// with one fewer left.
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kTriMask = 0x1FFFFFFFu;  // largest first-triangle index
__host__ __device__ inline uint32_t leaf_entry(uint32_t first, uint32_t count) {
	return kLeafBit | (first << 2) | (count - 1u);
}
__host__ __device__ inline uint32_t leaf_first(uint32_t e) { return (e & ~kLeafBit) >> 2; }
__host__ __device__ inline uint32_t leaf_count(uint32_t e) { return (e & 3u) + 1u; }
// Entry base for a node's leaf children: leaf_base(tri_base) + meta byte = the leaf's entry.
__host__ __device__ inline uint32_t leaf_base(uint32_t tri_base) { return kLeafBit | (tri_base << 2); }
// Stack entry of a used child slot with meta byte m.
__host__ __device__ inline uint32_t child_entry(uint32_t m, uint32_t child_base, uint32_t lbase) {
	return ((m & 0x80u) ? child_base : lbase) + (m & 0x7Fu);
}

// Scale 2^e as a normal float; e is kept in [-126, 127] by the builder.
__host__ __device__ inline float exp2i(int e) {
	union {
		uint32_t u;
		float f;
	} c;
	c.u = static_cast<uint32_t>(e + 127) << 23;
	return c.f;
}

// ----------------------------------------------------------------- reference predicates

// std::min / std::max semantics (the first argument wins when the comparison is false).
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }

// Slab test of the reference (bvh.hpp:195-222) on one box, with the reciprocal passed in
// (1.0f / dir is the same IEEE value in every call). NaN behaviour is reproduced exactly.
__device__ __forceinline__ bool ref_slab(const float o[3], const float inv[3], const float lo[3],
                                         const float hi[3]) {
#pragma clang fp contract(off)
	const float t1 = (lo[0] - o[0]) * inv[0];
	const float t2 = (hi[0] - o[0]) * inv[0];
	const float t3 = (lo[1] - o[1]) * inv[1];
	const float t4 = (hi[1] - o[1]) * inv[1];
	const float t5 = (lo[2] - o[2]) * inv[2];
	const float t6 = (hi[2] - o[2]) * inv[2];
	const float tmin = smax(smax(smin(t1, t2), smin(t3, t4)), smin(t5, t6));
	const float tmax = smin(smin(smax(t1, t2), smax(t3, t4)), smax(t5, t6));
	return !(tmax < 0.0f) && !(tmin > tmax);
}

// Moller-Trumbore of the reference (core.hpp:27-65), every operation rounded on its own:
// rejects det == 0 exactly, inclusive u/v bounds, one IEEE division, t without a range check.
__device__ __forceinline__ bool ref_mt(const float v[9], const float o[3], const float d[3], float &t,
                                       float &u, float &vv) {
#pragma clang fp contract(off)
	const float e1x = v[3] - v[0], e1y = v[4] - v[1], e1z = v[5] - v[2];
	const float e2x = v[6] - v[0], e2y = v[7] - v[1], e2z = v[8] - v[2];
	const float px = d[1] * e2z - d[2] * e2y;
	const float py = d[2] * e2x - d[0] * e2z;
	const float pz = d[0] * e2y - d[1] * e2x;
	const float det = e1x * px + e1y * py + e1z * pz;
	const float inv = 1.0f / det;
	const float tx = o[0] - v[0], ty = o[1] - v[1], tz = o[2] - v[2];
	u = (tx * px + ty * py + tz * pz) * inv;
	const float qx = ty * e1z - tz * e1y;
	const float qy = tz * e1x - tx * e1z;
	const float qz = tx * e1y - ty * e1x;
	vv = (d[0] * qx + d[1] * qy + d[2] * qz) * inv;
	t = (e2x * qx + e2y * qy + e2z * qz) * inv;
	// the reference returns early on each failed test; evaluating all terms and testing the
	// same conditions afterwards gives the same accept/reject and the same t, u, v
	return !(det == 0.0f) && !(u < 0.0f || u > 1.0f) && !(vv < 0.0f || u + vv > 1.0f);
}

// Exact candidate test of one triangle: reference slab on its exact box, then the reference
// leaf test (bvh.hpp:238-246).
__device__ __forceinline__ bool exact_candidate(const float v[9], const float o[3], const float d[3],
                                                const float inv[3], float &t, float &u, float &vv) {
	float lo[3], hi[3];
#pragma unroll
	for (int k = 0; k < 3; ++k) {
		lo[k] = smin(v[k], smin(v[3 + k], v[6 + k]));
		hi[k] = smax(v[k], smax(v[3 + k], v[6 + k]));
	}
	const bool box = ref_slab(o, inv, lo, hi);
	const bool tri = ref_mt(v, o, d, t, u, vv);
	return box && tri;
}

// Watertight leaf test (Woop, Benthin and Wald, "Watertight Ray/Triangle Intersection", JCGT
// 2(1), 2013), the optional mode beside the reference's Moller-Trumbore. Semantics kept from
// core.hpp:27-65: both windings, no t range check, (u, v) the weights of vertices 1 and 2. The
// ray is sheared so that it runs along +z of the axes (kx, ky, kz), kz the largest |d|
// component; the 2-D edge functions U, V, W of the sheared vertices decide the hit by sign, and
// one that rounds to exactly 0 is re-evaluated from the fma-exact product errors, which gives
// its exact sign (fl(ab) == fl(ce) there, so ab - ce = (ab - fl(ab)) - (ce - fl(ce))): a ray
// through a shared edge is accepted by at least one of the two triangles. fp32 only.
__device__ __forceinline__ float wt_edge(float a, float b, float c, float e) {
#pragma clang fp contract(off)
	const float p = a * b, q = c * e;
	const float f = p - q;
	return f != 0.0f ? f : fmaf(a, b, -p) - fmaf(c, e, -q);
}

__device__ __forceinline__ float sel3(int k, float a0, float a1, float a2) { return k == 0 ? a0 : (k == 1 ? a1 : a2); }

__device__ __forceinline__ bool wt_test(const float v[9], const float o[3], const float d[3], float &t, float &u,
                                        float &vv) {
#pragma clang fp contract(off)
	int kz = 0;
	if (fabsf(d[1]) > fabsf(d[kz])) kz = 1;
	if (fabsf(d[2]) > fabsf(sel3(kz, d[0], d[1], d[2]))) kz = 2;
	int kx = kz == 2 ? 0 : kz + 1;
	int ky = kx == 2 ? 0 : kx + 1;
	const float dz = sel3(kz, d[0], d[1], d[2]);
	if (dz < 0.0f) {
		const int tmp = kx;
		kx = ky;
		ky = tmp;
	}
	const float sx = sel3(kx, d[0], d[1], d[2]) / dz, sy = sel3(ky, d[0], d[1], d[2]) / dz, sz = 1.0f / dz;
	float rel[3][3];  // vertex - origin
#pragma unroll
	for (int i = 0; i < 3; ++i)
#pragma unroll
		for (int k = 0; k < 3; ++k) rel[i][k] = v[3 * i + k] - o[k];
	float x[3], y[3], z[3];
#pragma unroll
	for (int i = 0; i < 3; ++i) {
		const float rx = sel3(kx, rel[i][0], rel[i][1], rel[i][2]);
		const float ry = sel3(ky, rel[i][0], rel[i][1], rel[i][2]);
		z[i] = sel3(kz, rel[i][0], rel[i][1], rel[i][2]);
		x[i] = rx - sx * z[i];
		y[i] = ry - sy * z[i];
	}
	const float U = wt_edge(x[2], y[1], y[2], x[1]);
	const float V = wt_edge(x[0], y[2], y[0], x[2]);
	const float W = wt_edge(x[1], y[0], y[1], x[0]);
	if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
	const float det = U + V + W;
	if (det == 0.0f) return false;
	const float T = U * (sz * z[0]) + V * (sz * z[1]) + W * (sz * z[2]);
	const float rcp = 1.0f / det;
	t = T * rcp;
	u = V * rcp;
	vv = W * rcp;
	return true;
}

// Candidate test of the watertight mode: the watertight test, and the triangle's exact box not
// entirely behind the origin (the reference slab test's tmax >= 0, bvh.hpp:213, with its
// operations; its tmin <= tmax part is left out because it can fail by an ulp at a box corner
// that a ray through a shared vertex grazes).
__device__ __forceinline__ bool wt_candidate(const float v[9], const float o[3], const float d[3],
                                             const float inv[3], float &t, float &u, float &vv) {
#pragma clang fp contract(off)
	float tmax = __builtin_inff();
#pragma unroll
	for (int k = 0; k < 3; ++k) {
		const float lo = smin(v[k], smin(v[3 + k], v[6 + k]));
		const float hi = smax(v[k], smax(v[3 + k], v[6 + k]));
		const float t1 = (lo - o[k]) * inv[k], t2 = (hi - o[k]) * inv[k];
		tmax = k == 0 ? smax(t1, t2) : smin(tmax, smax(t1, t2));
	}
	const bool box = !(tmax < 0.0f);
	const bool tri = wt_test(v, o, d, t, u, vv);
	return box && tri;
}

// ------------------------------------------------------- distance-culling guard per triangle
//
// Distance culling drops a box whose entry distance exceeds t_best + max(m_rel |t_best|, m_abs),
// so it is exact only if no triangle's Moller-Trumbore t falls further below its own box entry
// than that margin. With e1, e2 the fp32 edges the leaf test computes, n = e1 x e2 and
// W_j = |e1_k e2_l| + |e1_l e2_k| ({j, k, l} = {x, y, z}), the standard rounding analysis of the
// reference arithmetic gives, for a ray (o, d) and s = o - v0,
//   |det_fp - det| <= 5u (W . |d|),   |num_fp - num| <= 5u (W . |s|)     (u = 2^-24)
// with det = -(n . d): the relative error of t is driven by rho(d) = (W . |d|) / |n . d|.
// For a triangle in an axis plane (normal along axis k) every product involving the other two
// axes is exactly zero, so rho = kappa = W_k / |n_k| for every ray, the numerator's error is
// relative to s_k (and so to t) with the same factor, and the ray crosses the triangle's flat box
// exactly at the plane: |t_fp - t| <= (10 kappa + 8) u |t| for any ray. With kappa <= 1024 that
// is below 2^-10.3 |t|, inside the culling margin m_rel = 2^-10. A tilted triangle's rho also
// grows as 1 / |cos(angle between d and n)|: kappa_T = rho(n) (its value for the ray along the
// normal) is its ray-independent part.
//
// cull_unsafe() flags a triangle whose kappa_T may exceed 1024 (slivers and degenerate
// triangles, kappa_T computed in fp32 with fma-exact products and a threshold of 1000, which
// covers the evaluation's own error), or whose edges or normal lie outside the ranges the bound
// assumes: edge components 0 or within [2^-30, 2^30] and |n|_inf >= 2^-60. Culling is applied
// only to rays whose direction components all lie within [2^-30, 2^30] (and |o| < 2^60); then
// no product of the leaf test overflows, the products of d and the edges stay normal, and a
// product that underflows (an origin within 2^-100 of a vertex plane) moves t by less than
// 2^-69 |n_k|^-1 2^-48, below the absolute margin. A flagged triangle's box is never culled:
// the flag is ORed into every ancestor's child slot.
__host__ __device__ inline bool cull_unsafe(const float v[9]) {
#pragma clang fp contract(off)
	float e1[3], e2[3];
	for (int k = 0; k < 3; ++k) {
		e1[k] = v[3 + k] - v[k];
		e2[k] = v[6 + k] - v[k];
	}
	for (int k = 0; k < 3; ++k) {
		const float a = fabsf(e1[k]), b = fabsf(e2[k]);
		if (!((a == 0.0f || (a >= 0x1p-30f && a <= 0x1p30f)) && (b == 0.0f || (b >= 0x1p-30f && b <= 0x1p30f))))
			return true;  // also NaN and infinite edges
	}
	float n[3], w[3];
	for (int j = 0; j < 3; ++j) {
		const int k = (j + 1) % 3, l = (j + 2) % 3;
		const float p = e1[l] * e2[k];
		n[j] = fmaf(e1[k], e2[l], -p);  // e1_k e2_l - e1_l e2_k with one rounding after p's
		w[j] = fabsf(e1[k] * e2[l]) + fabsf(p);
	}
	const float m = fmaxf(fabsf(n[0]), fmaxf(fabsf(n[1]), fabsf(n[2])));
	if (!(m >= 0x1p-60f)) return true;  // (near) zero area
	float num = 0.0f, den = 0.0f;
	for (int j = 0; j < 3; ++j) {
		const float a = fabsf(n[j]) / m;  // in [0, 1]
		num += w[j] * a;
		den += a * a;
	}
	// kappa_T = (W . |n|) / (n . n) = num / (m * den); den >= 1
	return !(num <= 1000.0f * m * den);
}

// Keep the better of (t, id): smaller t, equal finite t resolved to the lower id. NaN and +inf
// never win (the reference's strict t < t_near with t_near starting at +inf).
__device__ __forceinline__ bool better(float t, uint32_t id, float tb, uint32_t idb) {
	return (t < tb) || (t == tb && t < __builtin_inff() && id < idb);
}

// ---------------------------------------------------------------------- hit record writer

// Byte layout of HitReg<Tags...> for a field mask (hitreg.hpp:29-44): members u v t id valid
// px py pz in that order; absent members are 1-byte empties, present floats/uint32 align to 4,
// valid is one byte. off[] = u v t id valid px py pz (-1 absent).
struct Layout {
	int32_t size;
	int32_t off[8];
};

__host__ __device__ constexpr Layout layout_of(uint32_t mask) {
	Layout L{};
	int pos = 0, align = 1;
	auto place = [&](bool present, int sz, int al) {
		if (!present) {  // Empty member: one byte, alignment 1
			pos += 1;
			return -1;
		}
		pos = (pos + al - 1) / al * al;
		const int o = pos;
		pos += sz;
		if (al > align) align = al;
		return o;
	};
	const bool uv = mask & F_UV, t = mask & F_T, id = mask & F_ID, p = mask & F_P, va = mask & F_VALID;
	L.off[0] = place(uv, 4, 4);
	L.off[1] = place(uv, 4, 4);
	L.off[2] = place(t, 4, 4);
	L.off[3] = place(id, 4, 4);
	L.off[4] = place(va, 1, 1);
	L.off[5] = place(p, 4, 4);
	L.off[6] = place(p, 4, 4);
	L.off[7] = place(p, 4, 4);
	L.size = (pos + align - 1) / align * align;
	return L;
}

struct Hit {
	float t, u, v;
	uint32_t id;
};

// Writes one full record (padding bytes zero) with dword stores. `Mask` is the field mask.
template <uint32_t Mask>
__device__ __forceinline__ void write_record(uint8_t *rec, const Hit &h, const float o[3], const float d[3]) {
#pragma clang fp contract(off)
	constexpr Layout L = layout_of(Mask);
	static_assert(L.size % 4 == 0 && L.size <= 32, "record size");
	uint32_t w[L.size / 4];
#pragma unroll
	for (int i = 0; i < L.size / 4; ++i) w[i] = 0u;
	const bool valid = h.t < __builtin_inff();
	if constexpr ((Mask & F_UV) != 0) {
		w[L.off[0] / 4] = __float_as_uint(valid ? h.u : 0.0f);
		w[L.off[1] / 4] = __float_as_uint(valid ? h.v : 0.0f);
	}
	if constexpr ((Mask & F_T) != 0) w[L.off[2] / 4] = __float_as_uint(h.t);
	if constexpr ((Mask & F_ID) != 0) w[L.off[3] / 4] = valid ? h.id : 0xFFFFFFFFu;
	if constexpr ((Mask & F_VALID) != 0) w[L.off[4] / 4] |= (valid ? 1u : 0u) << (8 * (L.off[4] % 4));
	if constexpr ((Mask & F_P) != 0) {
		// bvh.hpp:261-263
		w[L.off[5] / 4] = __float_as_uint(o[0] + h.t * d[0]);
		w[L.off[6] / 4] = __float_as_uint(o[1] + h.t * d[1]);
		w[L.off[7] / 4] = __float_as_uint(o[2] + h.t * d[2]);
	}
	if constexpr (L.size % 16 == 0) {
#pragma unroll
		for (int i = 0; i < L.size / 16; ++i)
			reinterpret_cast<uint4 *>(rec)[i] = make_uint4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
	} else if constexpr (L.size % 8 == 0) {
#pragma unroll
		for (int i = 0; i < L.size / 8; ++i) reinterpret_cast<uint2 *>(rec)[i] = make_uint2(w[2 * i], w[2 * i + 1]);
	} else {
#pragma unroll
		for (int i = 0; i < L.size / 4; ++i) reinterpret_cast<uint32_t *>(rec)[i] = w[i];
	}
}

// Dispatch on the runtime mask (uniform across the launch).
__device__ __forceinline__ void write_any(uint32_t mask, uint8_t *rec, const Hit &h, const float o[3],
                                          const float d[3]) {
	switch (mask) {
#define PRT_CASE(m)                                                                                \
	case m: write_record<m>(rec, h, o, d); break;
		PRT_CASE(0) PRT_CASE(1) PRT_CASE(2) PRT_CASE(3) PRT_CASE(4) PRT_CASE(5) PRT_CASE(6) PRT_CASE(7)
		PRT_CASE(8) PRT_CASE(9) PRT_CASE(10) PRT_CASE(11) PRT_CASE(12) PRT_CASE(13) PRT_CASE(14)
		PRT_CASE(15) PRT_CASE(16) PRT_CASE(17) PRT_CASE(18) PRT_CASE(19) PRT_CASE(20) PRT_CASE(21)
		PRT_CASE(22) PRT_CASE(23) PRT_CASE(24) PRT_CASE(25) PRT_CASE(26) PRT_CASE(27) PRT_CASE(28)
		PRT_CASE(29) PRT_CASE(30) PRT_CASE(31)
#undef PRT_CASE
	default: break;
	}
}

}  // namespace prt
