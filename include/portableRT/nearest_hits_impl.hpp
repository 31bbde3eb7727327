// Free nearest_hits, to_variant and select_backend of the portableRT API
// (nearest_hits_impl.hpp:28-36 and src/backend.cpp:24-56 in the reference). Inline here so the
// API layer is header-only.
#pragma once

#include <stdexcept>
#include <variant>

#include "backend.hpp"
#include "intersect_mi355x.hpp"

namespace portableRT {

inline BackendVar to_variant(Backend *backend) {
	if (auto *q = dynamic_cast<MI355XBackend *>(backend)) return q;
	throw std::runtime_error("Unknown backend");
}

inline void select_backend(Backend *backend) {
	if (selected_backend) selected_backend->shutdown();
	selected_backend = backend;
	var_selected = to_variant(backend);
	backend->init();
}

template <class... Tags> std::vector<HitReg<Tags...>> nearest_hits(const std::vector<Ray> &rays) {
	return std::visit(
	    [&](auto *b) -> std::vector<HitReg<Tags...>> {
		    if (!b) throw std::runtime_error("Unknown backend");
		    return b->template nearest_hits<Tags...>(rays);
	    },
	    var_selected);
}

}  // namespace portableRT
