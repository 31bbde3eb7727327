// C-ABI shim over the reference CPU backend, compiled from the reference's own sources
// (src/portableRT.cpp, src/intersect_cpu.cpp, src/backend.cpp) where they lie. TEST
// INFRASTRUCTURE ONLY: it produces golden vectors and validates the oracle restatement; the
// result lands in oracle/_ref/ and is never part of the product.
#include <portableRT/portableRT.hpp>

#include <chrono>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {

// Builds the reference BVH2 and traces n rays with all tags (FullHitReg, 32 bytes per ray,
// the reference's own layout). Miss records carry the reference's indeterminate u/v/id.
int ref_trace_full(const float *tris, uint64_t n_tris, const float *rays, uint64_t n_rays, void *out) {
	portableRT::Tris t(n_tris);
	for (uint64_t i = 0; i < n_tris; ++i) std::memcpy(t[i].data(), tris + 9 * i, 9 * sizeof(float));
	std::vector<portableRT::Ray> r(n_rays);
	for (uint64_t i = 0; i < n_rays; ++i) {
		std::memcpy(r[i].origin.data(), rays + 6 * i, 3 * sizeof(float));
		std::memcpy(r[i].direction.data(), rays + 6 * i + 3, 3 * sizeof(float));
	}
	portableRT::select_backend(&portableRT::cpu_backend);
	portableRT::cpu_backend.set_tris(t);
	auto hits = portableRT::nearest_hits<ALL_TAGS>(r);
	static_assert(sizeof(portableRT::FullHitReg) == 32, "FullHitReg layout");
	std::memcpy(out, hits.data(), hits.size() * sizeof(portableRT::FullHitReg));
	return 0;
}

// Same, timing only the nearest_hits call (steady clock, seconds) after set_tris.
double ref_time_trace(const float *tris, uint64_t n_tris, const float *rays, uint64_t n_rays, int reps) {
	portableRT::Tris t(n_tris);
	for (uint64_t i = 0; i < n_tris; ++i) std::memcpy(t[i].data(), tris + 9 * i, 9 * sizeof(float));
	std::vector<portableRT::Ray> r(n_rays);
	for (uint64_t i = 0; i < n_rays; ++i) {
		std::memcpy(r[i].origin.data(), rays + 6 * i, 3 * sizeof(float));
		std::memcpy(r[i].direction.data(), rays + 6 * i + 3, 3 * sizeof(float));
	}
	portableRT::select_backend(&portableRT::cpu_backend);
	portableRT::cpu_backend.set_tris(t);
	double best = 1e300;
	for (int k = 0; k < reps; ++k) {
		auto a = std::chrono::steady_clock::now();
		auto hits = portableRT::nearest_hits<portableRT::filter::t, portableRT::filter::primitive_id>(r);
		auto b = std::chrono::steady_clock::now();
		double s = std::chrono::duration<double>(b - a).count();
		if (s < best) best = s;
		if (hits.size() != n_rays) return -1.0;
	}
	return best;
}

// The sizes and field offsets of HitReg for every field mask (bit 0 uv, 1 t, 2 primitive_id,
// 3 p, 4 valid), as the reference headers lay them out: size, then offsets u v t id valid px
// py pz (-1 for absent fields).
#define PRT_L(...)                                                                                 \
	{                                                                                              \
		using H = portableRT::HitReg<__VA_ARGS__>;                                                 \
		out[0] = sizeof(H);                                                                        \
		out[1] = H::has_uv::value ? (int)offsetof(H, u) : -1;                                      \
		out[2] = H::has_uv::value ? (int)offsetof(H, v) : -1;                                      \
		out[3] = H::has_t::value ? (int)offsetof(H, t) : -1;                                       \
		out[4] = H::has_primitive_id::value ? (int)offsetof(H, primitive_id) : -1;                 \
		out[5] = H::has_valid::value ? (int)offsetof(H, valid) : -1;                               \
		out[6] = H::has_p::value ? (int)offsetof(H, px) : -1;                                      \
		out[7] = H::has_p::value ? (int)offsetof(H, py) : -1;                                      \
		out[8] = H::has_p::value ? (int)offsetof(H, pz) : -1;                                      \
	}
int ref_hitreg_layout(int mask, int *out) {
	using namespace portableRT::filter;
	switch (mask) {
	case 0x01: PRT_L(uv) break;
	case 0x02: PRT_L(t) break;
	case 0x04: PRT_L(primitive_id) break;
	case 0x08: PRT_L(p) break;
	case 0x10: PRT_L(valid) break;
	case 0x03: PRT_L(uv, t) break;
	case 0x05: PRT_L(uv, primitive_id) break;
	case 0x09: PRT_L(uv, p) break;
	case 0x11: PRT_L(uv, valid) break;
	case 0x06: PRT_L(t, primitive_id) break;
	case 0x0A: PRT_L(t, p) break;
	case 0x12: PRT_L(t, valid) break;
	case 0x0C: PRT_L(primitive_id, p) break;
	case 0x14: PRT_L(primitive_id, valid) break;
	case 0x18: PRT_L(p, valid) break;
	case 0x07: PRT_L(uv, t, primitive_id) break;
	case 0x0B: PRT_L(uv, t, p) break;
	case 0x13: PRT_L(uv, t, valid) break;
	case 0x0D: PRT_L(uv, primitive_id, p) break;
	case 0x15: PRT_L(uv, primitive_id, valid) break;
	case 0x19: PRT_L(uv, p, valid) break;
	case 0x0E: PRT_L(t, primitive_id, p) break;
	case 0x16: PRT_L(t, primitive_id, valid) break;
	case 0x1A: PRT_L(t, p, valid) break;
	case 0x1C: PRT_L(primitive_id, p, valid) break;
	case 0x0F: PRT_L(uv, t, primitive_id, p) break;
	case 0x17: PRT_L(uv, t, primitive_id, valid) break;
	case 0x1B: PRT_L(uv, t, p, valid) break;
	case 0x1D: PRT_L(uv, primitive_id, p, valid) break;
	case 0x1E: PRT_L(t, primitive_id, p, valid) break;
	case 0x1F: PRT_L(uv, t, primitive_id, p, valid) break;
	default: return 1;
	}
	return 0;
}
}
