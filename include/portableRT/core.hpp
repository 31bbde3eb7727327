// Value types of the portableRT API (interface of include/portableRT/core.hpp in the
// reference, v0.3.3): Ray (core.hpp:19-22) and Tri / Tris (core.hpp:24-25). The GPU reads
// these arrays in place, so their layout (24-byte Ray, 36-byte Tri) is part of the contract.
#pragma once

#include <array>
#include <string>
#include <string_view>
#include <vector>

#include "hitreg.hpp"

namespace portableRT {

struct Ray {
	std::array<float, 3> origin;
	std::array<float, 3> direction;
};

using Tri = std::array<float, 9>;
using Tris = std::vector<Tri>;

static_assert(sizeof(Ray) == 24, "Ray must be 6 packed floats");
static_assert(sizeof(Tri) == 36, "Tri must be 9 packed floats");

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

}  // namespace portableRT
