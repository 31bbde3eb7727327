// Backend interface, registry and selection of the portableRT API (interface of
// include/portableRT/backend.hpp and src/backend.cpp in the reference, v0.3.3).
//
// Same names and behaviour: Backend with virtual set_tris / is_available / init / shutdown /
// device_name, InvokableBackend<D>, RegisterBackend (the first available backend becomes the
// selected one), all_backends(), available_backends(), selected_backend, select_backend()
// (shuts the old backend down, initialises the new one), the closed BackendVar variant and
// the free nearest_hits<Tags...>() dispatched through it. The variant here holds the MI355X
// backend. The registry vectors are function-local statics so that registration from static
// constructors in any translation unit is well ordered.
#pragma once

#include <stdexcept>
#include <string>
#include <utility>
#include <variant>
#include <vector>

#include "core.hpp"

namespace portableRT {

class Backend {
  public:
	Backend(std::string name, void *self_ptr) : self_{self_ptr}, name_{std::move(name)} {}
	virtual ~Backend() = default;

	const std::string &name() const { return name_; }

	virtual void set_tris(const Tris &tris) = 0;
	virtual bool is_available() const = 0;
	virtual void init() = 0;
	virtual void shutdown() = 0;
	virtual std::string device_name() const = 0;

	void *self_;

  private:
	std::string name_;
};

template <class Derived> class InvokableBackend : public Backend {
  public:
	explicit InvokableBackend(std::string name) : Backend{std::move(name), static_cast<void *>(this)} {}

  protected:
	~InvokableBackend() = default;
};

inline Backend *selected_backend = nullptr;

namespace detail {
inline std::vector<Backend *> &all_list() {
	static std::vector<Backend *> v;
	return v;
}
inline std::vector<Backend *> &available_list() {
	static std::vector<Backend *> v;
	return v;
}
}  // namespace detail

inline const std::vector<Backend *> &all_backends() { return detail::all_list(); }
inline const std::vector<Backend *> &available_backends() { return detail::available_list(); }

class MI355XBackend;

using BackendVar = std::variant<MI355XBackend *>;

inline BackendVar var_selected{static_cast<MI355XBackend *>(nullptr)};

template <class... Tags> std::vector<HitReg<Tags...>> nearest_hits(const std::vector<Ray> &rays);

inline std::vector<FullHitReg> nearest_hits(const std::vector<Ray> &rays) { return nearest_hits<ALL_TAGS>(rays); }

BackendVar to_variant(Backend *backend);
void select_backend(Backend *backend);

struct RegisterBackend {
	explicit RegisterBackend(Backend &b) {
		if (b.is_available()) {
			if (selected_backend == nullptr) select_backend(&b);
			detail::available_list().push_back(&b);
		}
		detail::all_list().push_back(&b);
	}
};

}  // namespace portableRT
