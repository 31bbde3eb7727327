// Exercises the portableRT C++ API with the MI355X backend: registration, selection, set_tris,
// free and member nearest_hits with several tag lists and tag orders, the known answer of the
// reference validation example (examples/validation/main.cpp:66-111), slice() and error paths.
// Exit status 0 on success. Without a gfx950 device it checks the no-backend behaviour instead.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <portableRT/portableRT.hpp>
#include <portableRT/version.hpp>

using namespace portableRT;

static int failures = 0;
#define CHECK(c)                                                                                   \
	do {                                                                                           \
		if (!(c)) {                                                                                \
			std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);                                \
			++failures;                                                                            \
		}                                                                                          \
	} while (0)

static bool near(float a, float b) { return std::fabs(a - b) <= 1e-4f; }

int main() {
	std::printf("portableRT %s, MI355X backend %s\n", PORTABLERT_VERSION, PORTABLERT_MI355X_VERSION);
	CHECK(all_backends().size() == 1);
	CHECK(all_backends()[0]->name() == "MI355X");
	if (available_backends().empty()) {
		std::printf("no gfx950 device: checking the no-backend path\n");
		CHECK(selected_backend == nullptr);
		bool threw = false;
		try {
			nearest_hits(std::vector<Ray>{Ray{{0, 0, 0}, {0, 0, 1}}});
		} catch (const std::runtime_error &e) {
			threw = std::string(e.what()) == "Unknown backend";
		}
		CHECK(threw);
		return failures ? 1 : 0;
	}
	for (auto *b : available_backends()) {
		select_backend(b);
		std::printf("backend %s on %s\n", b->name().c_str(), b->device_name().c_str());
		const Tri tri = {-1, -1, 0, 1, -1, 0, 0, 1, 0};
		Ray hit_ray{{0.1f, 0, -1}, {0, 0, 1}};
		Ray miss_ray{{-2, 0, -1}, {0, 0, 1}};
		b->set_tris({tri});
		auto h = nearest_hits({hit_ray});
		auto m = nearest_hits({miss_ray});
		CHECK(h[0].valid && near(h[0].t, 1.0f) && near(h[0].u, 0.3f) && near(h[0].v, 0.5f));
		CHECK(h[0].primitive_id == 0u && near(h[0].px, 0.1f) && near(h[0].py, 0.0f) && near(h[0].pz, 0.0f));
		CHECK(!m[0].valid && std::isinf(m[0].t));
		auto hv = nearest_hits<filter::valid>({hit_ray, miss_ray});
		CHECK(sizeof(hv[0]) == 8 && hv[0].valid && !hv[1].valid);
		// tag order does not matter; records equal the projection of the full record
		auto a = nearest_hits<filter::primitive_id, filter::t>({hit_ray});
		auto c = nearest_hits<filter::t, filter::primitive_id>({hit_ray});
		auto s = slice<filter::t, filter::primitive_id>(h[0]);
		CHECK(a[0].t == s.t && a[0].primitive_id == s.primitive_id && c[0].t == s.t);
		auto uvp = static_cast<MI355XBackend *>(b)->nearest_hits<filter::uv, filter::p>({hit_ray});
		CHECK(uvp[0].u == h[0].u && uvp[0].v == h[0].v && uvp[0].px == h[0].px);
		// many rays through a small grid of triangles
		Tris grid;
		for (int i = 0; i < 16; ++i)
			for (int j = 0; j < 16; ++j) {
				const float x = i * 0.25f, y = j * 0.25f;
				grid.push_back({x, y, 0, x + 0.25f, y, 0, x, y + 0.25f, 0});
			}
		b->set_tris(grid);
		std::vector<Ray> rays;
		for (int k = 0; k < 1000; ++k) rays.push_back(Ray{{0.01f + 0.0039f * k, 0.02f, -1}, {0, 0, 1}});
		auto g = nearest_hits<filter::t, filter::valid>(rays);
		int hits = 0;
		for (auto &r : g) hits += r.valid;
		CHECK(hits > 0);
		for (auto &r : g)
			if (r.valid) CHECK(r.t == 1.0f);
		// refit: the grid lifted by 0.5 keeps the tree, the hits move with it
		auto *mi = static_cast<MI355XBackend *>(b);
		Tris lifted = grid;
		for (auto &t : lifted)
			for (int v = 2; v < 9; v += 3) t[v] += 0.5f;
		mi->refit(lifted);
		auto gl = nearest_hits<filter::t, filter::valid>(rays);
		for (size_t k = 0; k < rays.size(); ++k) CHECK(gl[k].valid == g[k].valid && (!gl[k].valid || gl[k].t == 1.5f));
		// move overload, then a call above one chunk of the host pipeline (1,048,576 rays)
		mi->set_tris(std::move(lifted));
		std::vector<Ray> many;
		many.reserve(1500000);
		for (int k = 0; k < 1500000; ++k) many.push_back(rays[k % rays.size()]);
		auto gm = nearest_hits<filter::t, filter::primitive_id, filter::valid>(many);
		CHECK(gm.size() == many.size());
		int bad = 0;
		for (size_t k = 0; k < many.size(); ++k) bad += gm[k].valid != gl[k % rays.size()].valid || (gm[k].valid && gm[k].t != 1.5f);
		CHECK(bad == 0);
		// error path: non-finite vertex -> std::runtime_error
		bool threw = false;
		try {
			b->set_tris({{0, 0, NAN, 1, 0, 0, 0, 1, 0}});
		} catch (const std::runtime_error &) {
			threw = true;
		}
		CHECK(threw);
		b->set_tris({tri});
		CHECK(nearest_hits({hit_ray})[0].valid);
	}
	std::printf(failures ? "FAILED (%d)\n" : "all API checks passed\n", failures);
	return failures ? 1 : 0;
}
