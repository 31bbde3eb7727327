// The MI355X (gfx950) backend behind the portableRT Backend interface. Header-only: the
// member template nearest_hits<Tags...> derives the field mask and the HitReg layout on the
// host and calls the C ABI of include/prt_mi355x.h, so every tag list and tag order works
// without explicit instantiations (the reference instantiates 31 combos per compiled backend,
// nearesthit_inst.hpp:3-5). Link with libprt_mi355x.so.
//
// Errors throw std::runtime_error with the library's message; results are never silently
// replaced. Not thread-safe, like the reference backends.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "../prt_mi355x.h"
#include "backend.hpp"
#include "core.hpp"

namespace portableRT {

class MI355XBackend : public InvokableBackend<MI355XBackend> {
  public:
	MI355XBackend() : InvokableBackend("MI355X") { static RegisterBackend reg(*this); }

	// Device probe: true when a gfx950 device is visible. Never throws (no GPU -> false).
	bool is_available() const override {
		int n = 0;
		return prt_mi355x_probe(&n, nullptr, 0) == PRT_OK && n > 0;
	}

	void init() override {
		if (ctx_) return;
		const int rc = prt_mi355x_create(0, &ctx_);
		if (rc != PRT_OK) throw std::runtime_error("MI355X: no usable gfx950 device (status " + std::to_string(rc) + ")");
	}

	void shutdown() override {
		prt_mi355x_destroy(ctx_);
		ctx_ = nullptr;
	}

	void set_tris(const Tris &tris) override {
		init();
		check(prt_mi355x_set_tris(ctx_, tris.empty() ? nullptr : tris.data()->data(), tris.size()));
	}

	// Move version (the reference's backend.hpp:18 to-do): the triangles live on the device
	// after the build, so the caller's host copy is released here.
	void set_tris(Tris &&tris) {
		set_tris(static_cast<const Tris &>(tris));
		Tris().swap(tris);
	}

	// Same triangles at new positions: keeps the tree, recomputes its boxes (prt_mi355x_refit).
	void refit(const Tris &tris) {
		if (!ctx_) throw std::runtime_error("MI355X: refit before set_tris");
		check(prt_mi355x_refit(ctx_, tris.empty() ? nullptr : tris.data()->data(), tris.size()));
	}

	std::string device_name() const override {
		char buf[256] = {0};
		if (ctx_) {
			prt_mi355x_device_name(ctx_, buf, sizeof(buf));
		} else {
			int n = 0;
			prt_mi355x_probe(&n, buf, sizeof(buf));
			if (n == 0) return "unsupported";
		}
		return buf;
	}

	template <class... Tags> std::vector<HitReg<Tags...>> nearest_hits(const std::vector<Ray> &rays) {
		using H = HitReg<Tags...>;
		std::vector<H> hits(rays.size());
		if (!ctx_) throw std::runtime_error("MI355X: backend not initialised (select_backend first)");
		const int32_t off[8] = {
		    H::has_uv::value ? static_cast<int32_t>(offsetof(H, u)) : -1,
		    H::has_uv::value ? static_cast<int32_t>(offsetof(H, v)) : -1,
		    H::has_t::value ? static_cast<int32_t>(offsetof(H, t)) : -1,
		    H::has_primitive_id::value ? static_cast<int32_t>(offsetof(H, primitive_id)) : -1,
		    H::has_valid::value ? static_cast<int32_t>(offsetof(H, valid)) : -1,
		    H::has_p::value ? static_cast<int32_t>(offsetof(H, px)) : -1,
		    H::has_p::value ? static_cast<int32_t>(offsetof(H, py)) : -1,
		    H::has_p::value ? static_cast<int32_t>(offsetof(H, pz)) : -1};
		// the field mask from HitRegImpl's has_* members only (hitreg.hpp:29-34), so the header
		// also compiles against the reference's own hitreg.hpp
		const uint32_t mask = (H::has_uv::value ? PRT_UV : 0u) | (H::has_t::value ? PRT_T : 0u) |
		                      (H::has_primitive_id::value ? PRT_PRIMITIVE_ID : 0u) |
		                      (H::has_p::value ? PRT_P : 0u) | (H::has_valid::value ? PRT_VALID : 0u);
		const float *r = rays.empty() ? nullptr : rays.data()->origin.data();
		check(prt_mi355x_nearest_hits(ctx_, r, rays.size(), mask, hits.data(), static_cast<uint32_t>(sizeof(H)), off));
		return hits;
	}

	// The C ABI context (for options, statistics and the device-resident entry points).
	prt_ctx *context() const { return ctx_; }

  private:
	void check(int rc) const {
		if (rc != PRT_OK) throw std::runtime_error(std::string("MI355X: ") + prt_mi355x_last_error(ctx_));
	}
	prt_ctx *ctx_ = nullptr;
};

static MI355XBackend mi355x_backend;

}  // namespace portableRT
