// Hit records of the portableRT API (interface of include/portableRT/hitreg.hpp in the
// reference, v0.3.3). The member order, member types and the one-byte placeholder for absent
// fields are part of the drop-in contract: HitReg<Tags...> must have the same size and field
// offsets as the reference's, because records are written by the GPU with that byte layout.
#pragma once

#include <cstdint>
#include <iostream>
#include <type_traits>

#if defined(__HIPCC__)
#define PRT_HD __host__ __device__
#else
#define PRT_HD
#endif

namespace portableRT {

// Placeholder for a field that the tag list leaves out: one byte, swallows assignments,
// cannot be read (hitreg.hpp:16-23).
struct Empty {
	template <class T> PRT_HD constexpr Empty &operator=(T &&) noexcept { return *this; }
	template <class T> PRT_HD constexpr operator T() const = delete;
};

namespace filter {
struct uv {};
struct t {};
struct primitive_id {};
struct p {};
struct valid {};
}  // namespace filter

#define ALL_TAGS                                                                                   \
	portableRT::filter::uv, portableRT::filter::t, portableRT::filter::primitive_id,               \
	    portableRT::filter::p, portableRT::filter::valid

template <class T, class... List> constexpr bool has_tag = (std::is_same_v<T, List> || ...);

template <bool UV, bool T, bool ID, bool P, bool V> struct HitRegImpl {
	using has_uv = std::bool_constant<UV>;
	using has_t = std::bool_constant<T>;
	using has_primitive_id = std::bool_constant<ID>;
	using has_p = std::bool_constant<P>;
	using has_valid = std::bool_constant<V>;

	std::conditional_t<UV, float, Empty> u;
	std::conditional_t<UV, float, Empty> v;
	std::conditional_t<T, float, Empty> t;
	std::conditional_t<ID, uint32_t, Empty> primitive_id;
	std::conditional_t<V, bool, Empty> valid;
	std::conditional_t<P, float, Empty> px;
	std::conditional_t<P, float, Empty> py;
	std::conditional_t<P, float, Empty> pz;
};

template <class... Tags>
using HitReg = HitRegImpl<has_tag<filter::uv, Tags...>, has_tag<filter::t, Tags...>,
                          has_tag<filter::primitive_id, Tags...>, has_tag<filter::p, Tags...>,
                          has_tag<filter::valid, Tags...>>;

using FullHitReg = HitReg<ALL_TAGS>;

static_assert(sizeof(FullHitReg) == 32, "FullHitReg must match the reference layout");
static_assert(sizeof(HitReg<filter::t, filter::primitive_id>) == 16, "HitReg<t, primitive_id>");
static_assert(sizeof(HitReg<filter::t>) == 16, "HitReg<t>");
static_assert(sizeof(HitReg<filter::valid>) == 8, "HitReg<valid>");

// Field mask of a tag list as the C ABI encodes it (bit 0 uv, 1 t, 2 primitive_id, 3 p,
// 4 valid).
template <class... Tags> constexpr uint32_t field_mask() {
	return (has_tag<filter::uv, Tags...> ? 1u : 0u) | (has_tag<filter::t, Tags...> ? 2u : 0u) |
	       (has_tag<filter::primitive_id, Tags...> ? 4u : 0u) | (has_tag<filter::p, Tags...> ? 8u : 0u) |
	       (has_tag<filter::valid, Tags...> ? 16u : 0u);
}

// Projection of a full record onto a tag list (hitreg.hpp:63-84).
template <class... Tags> HitReg<Tags...> slice(const FullHitReg &h) {
	HitReg<Tags...> r;
	if constexpr (has_tag<filter::uv, Tags...>) {
		r.u = h.u;
		r.v = h.v;
	}
	if constexpr (has_tag<filter::t, Tags...>) r.t = h.t;
	if constexpr (has_tag<filter::primitive_id, Tags...>) r.primitive_id = h.primitive_id;
	if constexpr (has_tag<filter::p, Tags...>) {
		r.px = h.px;
		r.py = h.py;
		r.pz = h.pz;
	}
	if constexpr (has_tag<filter::valid, Tags...>) r.valid = h.valid;
	return r;
}

// Space-separated dump of the present fields (hitreg.hpp:86-102).
template <class... Tags> void print(const HitReg<Tags...> &h) {
	if constexpr (has_tag<filter::uv, Tags...>) std::cout << h.u << " " << h.v << " ";
	if constexpr (has_tag<filter::t, Tags...>) std::cout << h.t << " ";
	if constexpr (has_tag<filter::primitive_id, Tags...>) std::cout << h.primitive_id << " ";
	if constexpr (has_tag<filter::p, Tags...>) std::cout << h.px << " " << h.py << " " << h.pz << " ";
	if constexpr (has_tag<filter::valid, Tags...>) std::cout << h.valid << " ";
}

// The 31 non-empty tag combinations in canonical order, as an X-macro list (hitreg.hpp:146-177).
#define TAG_COMBOS                                                                                 \
	X(uv) X(t) X(primitive_id) X(p) X(valid) X(uv, t) X(uv, primitive_id) X(uv, p) X(uv, valid)    \
	X(t, primitive_id) X(t, p) X(t, valid) X(primitive_id, p) X(primitive_id, valid) X(p, valid)   \
	X(uv, t, primitive_id) X(uv, t, p) X(uv, t, valid) X(uv, primitive_id, p)                      \
	X(uv, primitive_id, valid) X(uv, p, valid) X(t, primitive_id, p) X(t, primitive_id, valid)     \
	X(t, p, valid) X(primitive_id, p, valid) X(uv, t, primitive_id, p) X(uv, t, primitive_id, valid) \
	X(uv, t, p, valid) X(uv, primitive_id, p, valid) X(t, primitive_id, p, valid)                  \
	X(uv, t, primitive_id, p, valid)

}  // namespace portableRT
