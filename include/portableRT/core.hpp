// Value types and helpers of the portableRT API (interface of include/portableRT/core.hpp in the
// reference, v0.3.3): Ray (core.hpp:19-22), Tri / Tris (core.hpp:24-25), clear_affinity,
// intersect_tri, the tag names and all_tags_combinations. The GPU reads Ray and Tri arrays in
// place, so their layout (24-byte Ray, 36-byte Tri) is part of the contract.
#pragma once

#include <sched.h>

#include <array>
#include <string>
#include <string_view>
#include <thread>
#include <vector>

#include "hitreg.hpp"

namespace portableRT {

// Lets the calling thread run on every CPU again (core.hpp:11-17: some runtimes pin the
// process to a subset).
inline void clear_affinity() {
	cpu_set_t all;
	CPU_ZERO(&all);
	const unsigned n = std::thread::hardware_concurrency();
	for (unsigned c = 0; c < n && c < CPU_SETSIZE; ++c) CPU_SET(c, &all);
	sched_setaffinity(0, sizeof(all), &all);
}

struct Ray {
	std::array<float, 3> origin;
	std::array<float, 3> direction;
};

using Tri = std::array<float, 9>;
using Tris = std::vector<Tri>;

static_assert(sizeof(Ray) == 24, "Ray must be 6 packed floats");
static_assert(sizeof(Tri) == 36, "Tri must be 9 packed floats");

// The reference's ray/triangle test (Moller-Trumbore, core.hpp:27-65), on the host: edges from
// vertex 0, p = d x e2, det = e1 . p (rejected only when exactly 0), one division 1 / det,
// inclusive barycentric bounds, t returned without a range check. Each operation is rounded on
// its own when the caller's compiler does not contract to fma (x86-64 without -march=...
// FMA, or -ffp-contract=off); the GPU backend evaluates the same sequence that way.
inline bool intersect_tri(const Tri &tri, const Ray &ray, float &t, float &u, float &v) {
	const auto &o = ray.origin;
	const auto &d = ray.direction;
	const float a[3] = {tri[3] - tri[0], tri[4] - tri[1], tri[5] - tri[2]};  // e1
	const float b[3] = {tri[6] - tri[0], tri[7] - tri[1], tri[8] - tri[2]};  // e2
	const float p[3] = {d[1] * b[2] - d[2] * b[1], d[2] * b[0] - d[0] * b[2], d[0] * b[1] - d[1] * b[0]};
	const float det = a[0] * p[0] + a[1] * p[1] + a[2] * p[2];
	if (det == 0.0f) return false;
	const float inv = 1.0f / det;
	const float s[3] = {o[0] - tri[0], o[1] - tri[1], o[2] - tri[2]};
	u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv;
	if (u < 0.0f || u > 1.0f) return false;
	const float q[3] = {s[1] * a[2] - s[2] * a[1], s[2] * a[0] - s[0] * a[2], s[0] * a[1] - s[1] * a[0]};
	v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
	if (v < 0.0f || u + v > 1.0f) return false;
	t = (b[0] * q[0] + b[1] * q[1] + b[2] * q[2]) * inv;
	return true;
}

// Tag names (core.hpp:67-91).
template <class Tag> inline constexpr std::string_view tag_name_v = "unknown";
template <> inline constexpr std::string_view tag_name_v<filter::uv> = "uv";
template <> inline constexpr std::string_view tag_name_v<filter::t> = "t";
template <> inline constexpr std::string_view tag_name_v<filter::primitive_id> = "primitive_id";
template <> inline constexpr std::string_view tag_name_v<filter::p> = "p";
template <> inline constexpr std::string_view tag_name_v<filter::valid> = "valid";

template <class... Tags> std::string hitreg_name() {
	std::string s;
	((s += tag_name_v<Tags>, s += '_'), ...);
	if (!s.empty()) s.pop_back();
	return s;
}

// hitreg_name of each of the 31 TAG_COMBOS entries, in that order (core.hpp:83-91): the
// stringised tag list with ", " joined by '_'.
inline std::vector<std::string> all_tags_combinations() {
	std::vector<std::string> out;
	auto name = [](std::string list) {
		std::string r;
		for (char c : list)
			if (c == ',') r += '_';
			else if (c != ' ') r += c;
		return r;
	};
#define X(...) out.push_back(name(#__VA_ARGS__));
	TAG_COMBOS
#undef X
	return out;
}

}  // namespace portableRT
