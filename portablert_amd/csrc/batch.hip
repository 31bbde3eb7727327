// Batched multi-scene tracing (no counterpart in the reference, whose backends hold one scene,
// include/portableRT/backend.hpp:19-46, and whose HIP backend builds on the host and launches
// one thread per ray, src/intersect_hip.cpp:295-341).
//
//   * One build over all scenes (build.hip with a SceneSet): the scene index sits above the
//     Morton code, so one radix sort and one radix tree hold every scene, BVH8 node s is the root
//     of scene s and leaf records carry scene-local primitive ids. The build runs a fixed sequence
//     of launches (a few BVH8 level launches, then the rest of the tree depth first) with no
//     host synchronisation, so it can be captured in a graph.
//   * One persistent traversal launch over all rays: waves claim work tiles (a run of rays of
//     one scene) with one atomic each, largest scenes first; each lane starts at its scene's root
//     with that scene's culling margin.
//   * build + traversal are captured once in a hipGraph and replayed.
//   * The autotuner picks the BVH branching factor, leaf size and tile size per size bucket of
//     scenes, and the workgroup size, persistent waves per CU and LDS stack depth of the launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/prt_mi355x.h"
#include "batch.h"
#include "common.h"
#include "ctx.h"
#include "internal.h"

namespace prt {

std::vector<uint32_t> batch_plan(uint32_t n_scenes, const uint64_t *tri_off, const uint64_t *ray_off,
                                 const int chunk[kBatchBuckets], int only_bucket) {
	std::vector<uint32_t> order(n_scenes);
	std::iota(order.begin(), order.end(), 0u);
	std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
		return tri_off[a + 1] - tri_off[a] > tri_off[b + 1] - tri_off[b];
	});
	std::vector<uint32_t> tiles;
	for (uint32_t s : order) {
		const int bk = batch_bucket(tri_off[s + 1] - tri_off[s]);
		if (only_bucket >= 0 && bk != only_bucket) continue;
		const uint64_t c = static_cast<uint64_t>(std::max(1, chunk[bk]));
		for (uint64_t r = ray_off[s]; r < ray_off[s + 1]; r += c) {
			const uint64_t cnt = std::min(c, ray_off[s + 1] - r);
			tiles.push_back(static_cast<uint32_t>(r & 0xFFFFFFFFull));
			tiles.push_back(static_cast<uint32_t>(r >> 32));
			tiles.push_back(static_cast<uint32_t>(cnt));
			tiles.push_back(s);
		}
	}
	return tiles;
}

namespace {

#define PRT_TRY(x)                                                                                 \
	do {                                                                                           \
		hipError_t e_ = (x);                                                                       \
		if (e_ != hipSuccess) return e_;                                                           \
	} while (0)

// Per-scene culling margin (a fixed fraction of the scene extent, as for one scene) and the
// one-fma node test flag (every vertex coordinate below 2^60 in magnitude).
__global__ __launch_bounds__(256) void k_batch_info(const float *__restrict__ box, uint32_t scenes, int margin_log2,
                                                    float2 *__restrict__ info) {
	const uint32_t s = blockIdx.x * 256 + threadIdx.x;
	if (s >= scenes) return;
	const float *b = box + 6ull * s;
	float ext = 0.0f;
	bool fast = true;
	for (int k = 0; k < 3; ++k) ext = fmaxf(ext, b[3 + k] - b[k]);
	for (int k = 0; k < 6; ++k) fast &= fabsf(b[k]) < 0x1p60f;
	if (!__builtin_isfinite(ext)) ext = 0.0f;
	info[s] = make_float2(ldexpf(ext, -margin_log2 - 6), fast ? 1.0f : 0.0f);
}

uint64_t scene_tris(const Batch &b, uint32_t s) { return b.tri_off[s + 1] - b.tri_off[s]; }

hipError_t upload_params(Batch &b, hipStream_t st) {
	std::vector<uint32_t> p(b.n_scenes);
	for (uint32_t s = 0; s < b.n_scenes; ++s) {
		const BucketParams &q = b.bucket[batch_bucket(scene_tris(b, s))];
		p[s] = static_cast<uint32_t>(q.leaf) | (static_cast<uint32_t>(q.width) << 8);
	}
	PRT_TRY(b.d_param.ensure(sizeof(uint32_t) * b.n_scenes));
	PRT_TRY(hipMemcpyAsync(b.d_param.ptr, p.data(), sizeof(uint32_t) * b.n_scenes, hipMemcpyHostToDevice, st));
	PRT_TRY(hipStreamSynchronize(st));
	return hipSuccess;
}

hipError_t upload_tiles(const std::vector<uint32_t> &t, DevBuf &dst, hipStream_t st) {
	PRT_TRY(dst.ensure(std::max<size_t>(16, sizeof(uint32_t) * t.size())));
	if (!t.empty())
		PRT_TRY(hipMemcpyAsync(dst.ptr, t.data(), sizeof(uint32_t) * t.size(), hipMemcpyHostToDevice, st));
	PRT_TRY(hipStreamSynchronize(st));
	return hipSuccess;
}

hipError_t upload_all_tiles(Batch &b, hipStream_t st) {
	int chunk[kBatchBuckets];
	for (int k = 0; k < kBatchBuckets; ++k) chunk[k] = b.bucket[k].chunk;
	const std::vector<uint32_t> t = batch_plan(b.n_scenes, b.tri_off.data(), b.ray_off.data(), chunk);
	b.n_tiles = static_cast<uint32_t>(t.size() / 4);
	return upload_tiles(t, b.d_tiles, st);
}

TraceParams batch_trace_params(const prt_ctx *c, const Batch &b) {
	TraceParams p = c->tp;
	p.block = b.block;
	p.waves_per_cu = b.waves_per_cu;
	p.lds_stack = b.lds_stack;
	return p;
}

hipError_t enqueue_build(prt_ctx *c, Batch &b, const float *tris, hipStream_t st) {
	SceneSet ss;
	ss.count = b.n_scenes;
	ss.tri_off = b.d_tri_off.as<uint32_t>();
	ss.param = b.d_param.as<uint32_t>();
	ss.box = b.d_box.as<float>();
	ss.tris = tris;
	ss.levels = kBatchLevels;
	bool bad = false;
	BuildParams bp;
	// greedy collapse: on the batch's scenes the least-cost collapse traces no faster (7.07 ms per
	// step either way) and its cost rows add 1.3 ms to the bottom-up pass of every step's build
	bp.collapse = 0;
	PRT_TRY(build_bvh8(b.bs, b.tri_off.back(), bp, b.bvh, st, &bad, nullptr, &ss));
	hipLaunchKernelGGL(k_batch_info, dim3((b.n_scenes + 255) / 256), dim3(256), 0, st, b.d_box.as<float>(), b.n_scenes,
	                   c->margin_log2, b.d_info.as<float2>());
	return hipGetLastError();
}

hipError_t enqueue_trace(prt_ctx *c, Batch &b, const float *rays, void *out, hipStream_t st, const DevBuf &tiles,
                         uint32_t n_tiles, uint64_t n_rays, TraceStats *stats) {
	BatchArgs ba;
	ba.tiles = tiles.as<uint4>();
	ba.n_tiles = n_tiles;
	ba.info = b.d_info.as<float2>();
	ba.n_scenes = b.n_scenes;
	const TraceParams p = batch_trace_params(c, b);
	return trace_bvh8(b.bvh, b.ts, p, rays, n_rays, b.mask, out, st, nullptr, stats, nullptr, &ba);
}

hipError_t enqueue(prt_ctx *c, Batch &b, const float *tris, const float *rays, void *out, hipStream_t st,
                   TraceStats *stats) {
	PRT_TRY(enqueue_build(c, b, tris, st));
	return enqueue_trace(c, b, rays, out, st, b.d_tiles, b.n_tiles, b.ray_off.back(), stats);
}

// Smallest of `reps` event-timed runs of f on st, in ms.
template <class F>
hipError_t time_min(hipStream_t st, int reps, F f, double *ms) {
	hipEvent_t e0, e1;
	PRT_TRY(hipEventCreate(&e0));
	PRT_TRY(hipEventCreate(&e1));
	double best = 1e30;
	hipError_t err = hipSuccess;
	for (int r = 0; r < reps && err == hipSuccess; ++r) {
		err = hipEventRecord(e0, st);
		if (err == hipSuccess) err = f();
		if (err == hipSuccess) err = hipEventRecord(e1, st);
		if (err == hipSuccess) err = hipEventSynchronize(e1);
		float t = 0.0f;
		if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
		if (err == hipSuccess) best = std::min(best, static_cast<double>(t));
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	*ms = best;
	return err;
}

Batch *batch_of(prt_ctx *c) { return c ? c->batch.get() : nullptr; }

int check_io(prt_ctx *c, const void *tris, const void *rays, void *out) {
	const Batch &b = *c->batch;
	if (!tris) return fail(c, PRT_ERR_ARG, "batch: null triangle pointer");
	if (b.ray_off.back() > 0 && (!rays || !out)) return fail(c, PRT_ERR_ARG, "batch: null ray or record pointer");
	if ((reinterpret_cast<uintptr_t>(tris) | reinterpret_cast<uintptr_t>(rays)) & 3u)
		return fail(c, PRT_ERR_ARG, "batch: triangles and rays must be 4-byte aligned");
	if (reinterpret_cast<uintptr_t>(out) & 15u) return fail(c, PRT_ERR_ARG, "batch: records must be 16-byte aligned");
	return PRT_OK;
}

int batch_errors(prt_ctx *c, Batch &b) {
	hipError_t e = hipDeviceSynchronize();
	if (e != hipSuccess) return hip_fail(c, e, "batch");
	uint32_t bf = 0, tf = 0;
	e = build_errors(b.bs, c->stream, &bf);
	if (e == hipSuccess) e = trace_errors(b.ts, c->stream, &tf);
	if (e != hipSuccess) return hip_fail(c, e, "batch error flags");
	if (bf & 1u) return fail(c, PRT_ERR_INPUT, "batch: a triangle vertex is not finite");
	if (bf & 4u) return fail(c, PRT_ERR_HIP, "batch: a BVH deeper than " + std::to_string(kBatchLevels) + " levels");
	if (bf) return fail(c, PRT_ERR_HIP, "batch: inconsistent BVH build (flags " + std::to_string(bf) + ")");
	if (tf) return fail(c, PRT_ERR_HIP, "batch: traversal found a corrupt node or stack bound (flags " +
	                                       std::to_string(tf) + ")");
	return PRT_OK;
}

bool valid_bucket(const int32_t v[3]) {
	return v[0] >= 2 && v[0] <= 8 && v[1] >= 1 && v[1] <= kMaxLeaf && v[2] >= 1 && v[2] <= (1 << 20);
}
bool valid_shape(const int32_t v[3]) {
	return (v[0] == 64 || v[0] == 128 || v[0] == 256) && v[1] >= 1 && v[1] <= 64 && (v[2] == 8 || v[2] == 16);
}

#undef PRT_TRY

}  // namespace
}  // namespace prt

using prt::fail;
using prt::hip_fail;

extern "C" {

int prt_mi355x_batch_tiles(uint32_t n_scenes, const uint64_t *tri_counts, const uint64_t *ray_counts,
                           const int32_t chunk[4], uint32_t *tiles4, uint64_t *n_tiles) {
	return prt::guarded(nullptr, [&]() -> int {
		if (!n_tiles || (n_scenes && (!tri_counts || !ray_counts)) || !chunk) return PRT_ERR_ARG;
		std::vector<uint64_t> to(n_scenes + 1, 0), ro(n_scenes + 1, 0);
		for (uint32_t s = 0; s < n_scenes; ++s) {
			to[s + 1] = to[s] + tri_counts[s];
			ro[s + 1] = ro[s] + ray_counts[s];
		}
		int c[prt::kBatchBuckets];
		for (int k = 0; k < prt::kBatchBuckets; ++k) {
			if (chunk[k] < 1) return PRT_ERR_ARG;
			c[k] = chunk[k];
		}
		const std::vector<uint32_t> t = prt::batch_plan(n_scenes, to.data(), ro.data(), c);
		const uint64_t have = *n_tiles;
		*n_tiles = t.size() / 4;
		if (tiles4) {
			if (have < t.size() / 4) return PRT_ERR_ARG;
			std::memcpy(tiles4, t.data(), sizeof(uint32_t) * t.size());
		}
		return PRT_OK;
	});
}

int prt_mi355x_batch_setup(prt_ctx *c, uint32_t n_scenes, const uint64_t *tri_counts, const uint64_t *ray_counts,
                           uint32_t mask) {
	return prt::guarded(c, [&]() -> int {
		if (!c) return PRT_ERR_ARG;
		if (n_scenes == 0 || n_scenes > prt::kMaxScenes || !tri_counts || !ray_counts)
			return fail(c, PRT_ERR_ARG, "batch_setup: 1.." + std::to_string(prt::kMaxScenes) + " scenes with counts");
		if (mask > 31u) return fail(c, PRT_ERR_ARG, "batch_setup: field mask out of range");
		std::vector<uint64_t> to(n_scenes + 1, 0), ro(n_scenes + 1, 0);
		for (uint32_t s = 0; s < n_scenes; ++s) {
			to[s + 1] = to[s] + tri_counts[s];
			ro[s + 1] = ro[s] + ray_counts[s];
			if (to[s + 1] > static_cast<uint64_t>(prt::kTriMask) || ro[s + 1] > (1ull << 40))
				return fail(c, PRT_ERR_ARG, "batch_setup: more than 2^29 triangles or 2^40 rays");
		}
		if (to.back() == 0) return fail(c, PRT_ERR_ARG, "batch_setup: the batch holds no triangle");
		if (hipSetDevice(c->device) != hipSuccess) return fail(c, PRT_ERR_NODEV, "batch_setup: device");
		std::unique_ptr<prt::Batch> nb(new prt::Batch());
		if (c->batch) {  // keep tuned parameters across setups
			for (int k = 0; k < prt::kBatchBuckets; ++k) nb->bucket[k] = c->batch->bucket[k];
			nb->block = c->batch->block;
			nb->waves_per_cu = c->batch->waves_per_cu;
			nb->lds_stack = c->batch->lds_stack;
		}
		c->batch.reset();
		prt::Batch &b = *nb;
		b.n_scenes = n_scenes;
		b.tri_off = to;
		b.ray_off = ro;
		b.mask = mask;
		std::vector<uint32_t> off32(to.begin(), to.end());
		hipError_t e = b.d_tri_off.ensure(sizeof(uint32_t) * (n_scenes + 1));
		if (e == hipSuccess)
			e = hipMemcpyAsync(b.d_tri_off.ptr, off32.data(), sizeof(uint32_t) * (n_scenes + 1), hipMemcpyHostToDevice,
			                   c->stream);
		if (e == hipSuccess) e = b.d_box.ensure(sizeof(float) * 6 * n_scenes);
		if (e == hipSuccess) e = b.d_info.ensure(sizeof(float) * 2 * n_scenes);
		if (e == hipSuccess) e = prt::upload_params(b, c->stream);
		if (e == hipSuccess) e = prt::upload_all_tiles(b, c->stream);
		if (e != hipSuccess) return hip_fail(c, e, "batch_setup");
		c->batch = std::move(nb);
		return PRT_OK;
	});
}

int prt_mi355x_batch_run(prt_ctx *c, const float *d_tris9, const float *d_rays6, void *d_out, void *stream) {
	return prt::guarded(c, [&]() -> int {
		prt::Batch *b = prt::batch_of(c);
		if (!c) return PRT_ERR_ARG;
		if (!b) return fail(c, PRT_ERR_STATE, "batch_run before batch_setup");
		if (int r = prt::check_io(c, d_tris9, d_rays6, d_out)) return r;
		if (hipSetDevice(c->device) != hipSuccess) return fail(c, PRT_ERR_NODEV, "batch_run: device");
		hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
		const bool count = c->tp.count_stats != 0;
		hipError_t e = prt::enqueue(c, *b, d_tris9, d_rays6, d_out, st, count ? &c->stats : nullptr);
		if (e != hipSuccess) return hip_fail(c, e, "batch_run");
		b->ran = true;
		return PRT_OK;
	});
}

int prt_mi355x_batch_capture(prt_ctx *c, const float *d_tris9, const float *d_rays6, void *d_out) {
	return prt::guarded(c, [&]() -> int {
		prt::Batch *b = prt::batch_of(c);
		if (!c) return PRT_ERR_ARG;
		if (!b) return fail(c, PRT_ERR_STATE, "batch_capture before batch_setup");
		if (int r = prt::check_io(c, d_tris9, d_rays6, d_out)) return r;
		if (c->tp.count_stats) return fail(c, PRT_ERR_ARG, "batch_capture: statistics counting synchronises; turn it off");
		if (hipSetDevice(c->device) != hipSuccess) return fail(c, PRT_ERR_NODEV, "batch_capture: device");
		b->drop_graph();
		if (!b->ran) {  // size every buffer with one eager run: nothing may allocate under capture
			hipError_t e = prt::enqueue(c, *b, d_tris9, d_rays6, d_out, c->stream, nullptr);
			if (e != hipSuccess) return hip_fail(c, e, "batch_capture: sizing run");
			if (int r = prt::batch_errors(c, *b)) return r;
			b->ran = true;
		}
		hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
		if (e != hipSuccess) return hip_fail(c, e, "batch_capture: begin");
		hipError_t ee = prt::enqueue(c, *b, d_tris9, d_rays6, d_out, c->stream, nullptr);
		hipGraph_t g = nullptr;
		e = hipStreamEndCapture(c->stream, &g);
		if (ee != hipSuccess || e != hipSuccess) {
			if (g) (void)hipGraphDestroy(g);
			(void)hipGetLastError();
			return hip_fail(c, ee != hipSuccess ? ee : e, "batch_capture");
		}
		b->graph = g;
		e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
		if (e != hipSuccess) {
			b->drop_graph();
			return hip_fail(c, e, "batch_capture: instantiate");
		}
		return PRT_OK;
	});
}

int prt_mi355x_batch_replay(prt_ctx *c, void *stream) {
	prt::Batch *b = prt::batch_of(c);
	if (!c) return PRT_ERR_ARG;
	if (!b || !b->exec) return fail(c, PRT_ERR_STATE, "batch_replay before batch_capture");
	if (hipSetDevice(c->device) != hipSuccess) return fail(c, PRT_ERR_NODEV, "batch_replay: device");
	hipError_t e = hipGraphLaunch(b->exec, stream ? static_cast<hipStream_t>(stream) : c->stream);
	if (e != hipSuccess) return hip_fail(c, e, "batch_replay");
	return PRT_OK;
}

int prt_mi355x_batch_check(prt_ctx *c) {
	prt::Batch *b = prt::batch_of(c);
	if (!c) return PRT_ERR_ARG;
	if (!b) return fail(c, PRT_ERR_STATE, "batch_check before batch_setup");
	if (hipSetDevice(c->device) != hipSuccess) return fail(c, PRT_ERR_NODEV, "batch_check: device");
	return prt::batch_errors(c, *b);
}

int prt_mi355x_batch_get_params(const prt_ctx *c, int bucket, int32_t v[3]) {
	if (!c || !v) return PRT_ERR_ARG;
	const prt::Batch *b = c->batch.get();
	if (!b) return PRT_ERR_STATE;
	if (bucket == -1) {
		v[0] = b->block;
		v[1] = b->waves_per_cu;
		v[2] = b->lds_stack;
	} else if (bucket >= 0 && bucket < prt::kBatchBuckets) {
		v[0] = b->bucket[bucket].width;
		v[1] = b->bucket[bucket].leaf;
		v[2] = b->bucket[bucket].chunk;
	} else {
		return PRT_ERR_ARG;
	}
	return PRT_OK;
}

int prt_mi355x_batch_set_params(prt_ctx *c, int bucket, const int32_t v[3]) {
	return prt::guarded(c, [&]() -> int {
		prt::Batch *b = prt::batch_of(c);
		if (!c || !v) return PRT_ERR_ARG;
		if (!b) return fail(c, PRT_ERR_STATE, "batch_set_params before batch_setup");
		if (bucket == -1) {
			if (!prt::valid_shape(v)) return fail(c, PRT_ERR_ARG, "batch_set_params: block 64/128/256, waves 1..64, stack 8/16");
			b->block = v[0];
			b->waves_per_cu = v[1];
			b->lds_stack = v[2];
		} else if (bucket >= 0 && bucket < prt::kBatchBuckets) {
			if (!prt::valid_bucket(v)) return fail(c, PRT_ERR_ARG, "batch_set_params: width 2..8, leaf 1..4, chunk 1..2^20");
			b->bucket[bucket].width = v[0];
			b->bucket[bucket].leaf = v[1];
			b->bucket[bucket].chunk = v[2];
		} else {
			return fail(c, PRT_ERR_ARG, "batch_set_params: bucket -1..3");
		}
		if (hipSetDevice(c->device) != hipSuccess) return fail(c, PRT_ERR_NODEV, "batch_set_params: device");
		hipError_t e = prt::upload_params(*b, c->stream);
		if (e == hipSuccess) e = prt::upload_all_tiles(*b, c->stream);
		if (e != hipSuccess) return hip_fail(c, e, "batch_set_params");
		b->drop_graph();
		b->ran = false;
		return PRT_OK;
	});
}

// Autotuner. Stage 1: every (width, leaf) candidate is applied to all buckets at once, the batch
// is rebuilt, and each bucket's rays are traced alone; each bucket keeps its fastest candidate.
// Stage 2: the same for the tile size. Stage 3: the launch shape over the whole batch. The tuned
// parameters are kept only if build + traversal of the whole batch got faster.
int prt_mi355x_batch_tune(prt_ctx *c, const float *d_tris9, const float *d_rays6, void *d_out, int reps,
                          double *before_ms, double *after_ms) {
	return prt::guarded(c, [&]() -> int {
		prt::Batch *b = prt::batch_of(c);
		if (!c) return PRT_ERR_ARG;
		if (!b) return fail(c, PRT_ERR_STATE, "batch_tune before batch_setup");
		if (int r = prt::check_io(c, d_tris9, d_rays6, d_out)) return r;
		if (hipSetDevice(c->device) != hipSuccess) return fail(c, PRT_ERR_NODEV, "batch_tune: device");
		reps = std::max(1, std::min(16, reps));
		hipStream_t st = c->stream;
		const int saved_count = c->tp.count_stats;
		c->tp.count_stats = 0;
		b->drop_graph();
		prt::Batch &B = *b;
		const prt::BucketParams orig[prt::kBatchBuckets] = {B.bucket[0], B.bucket[1], B.bucket[2], B.bucket[3]};
		const int orig_shape[3] = {B.block, B.waves_per_cu, B.lds_stack};
		hipError_t e = hipSuccess;
		auto whole = [&]() { return prt::enqueue(c, B, d_tris9, d_rays6, d_out, st, nullptr); };
		double t_before = 0.0, t_after = 0.0;
		// one untimed run sizes the buffers
		if (e == hipSuccess) e = whole();
		if (e == hipSuccess) e = hipStreamSynchronize(st);
		if (e == hipSuccess) e = prt::time_min(st, reps, whole, &t_before);
		// per-bucket subsets of tiles
		auto bucket_time = [&](int bk, const int chunk[prt::kBatchBuckets], double *ms) -> hipError_t {
			const std::vector<uint32_t> t = prt::batch_plan(B.n_scenes, B.tri_off.data(), B.ray_off.data(), chunk, bk);
			*ms = -1.0;
			if (t.empty()) return hipSuccess;
			uint64_t n = 0;
			for (size_t i = 2; i < t.size(); i += 4) n += t[i];
			hipError_t r = prt::upload_tiles(t, B.d_tiles_sub, st);
			if (r != hipSuccess) return r;
			const uint32_t nt = static_cast<uint32_t>(t.size() / 4);
			return prt::time_min(st, reps, [&]() { return prt::enqueue_trace(c, B, d_rays6, d_out, st, B.d_tiles_sub, nt, n, nullptr); }, ms);
		};
		int cur_chunk[prt::kBatchBuckets];
		for (int k = 0; k < prt::kBatchBuckets; ++k) cur_chunk[k] = B.bucket[k].chunk;
		// stage 1: branching factor and leaf size
		const int cand[][2] = {{8, 1}, {8, 2}, {8, 3}, {8, 4}, {4, 1}, {4, 2}, {4, 4}, {6, 2}};
		double best[prt::kBatchBuckets];
		prt::BucketParams pick[prt::kBatchBuckets];
		for (int k = 0; k < prt::kBatchBuckets; ++k) {
			best[k] = 1e30;
			pick[k] = B.bucket[k];
		}
		for (const auto &cd : cand) {
			if (e != hipSuccess) break;
			for (int k = 0; k < prt::kBatchBuckets; ++k) {
				B.bucket[k].width = cd[0];
				B.bucket[k].leaf = cd[1];
			}
			e = prt::upload_params(B, st);
			if (e == hipSuccess) e = prt::enqueue_build(c, B, d_tris9, st);
			for (int k = 0; k < prt::kBatchBuckets && e == hipSuccess; ++k) {
				double ms = 0.0;
				e = bucket_time(k, cur_chunk, &ms);
				if (e == hipSuccess && ms >= 0.0 && ms < best[k]) {
					best[k] = ms;
					pick[k].width = cd[0];
					pick[k].leaf = cd[1];
				}
			}
		}
		for (int k = 0; k < prt::kBatchBuckets; ++k) B.bucket[k] = pick[k];
		// stage 2: tile size
		if (e == hipSuccess) e = prt::upload_params(B, st);
		if (e == hipSuccess) e = prt::enqueue_build(c, B, d_tris9, st);
		const int chunks[] = {64, 128, 256, 512};
		for (int k = 0; k < prt::kBatchBuckets; ++k) best[k] = 1e30;
		for (int ch : chunks) {
			int cc[prt::kBatchBuckets] = {ch, ch, ch, ch};
			for (int k = 0; k < prt::kBatchBuckets && e == hipSuccess; ++k) {
				double ms = 0.0;
				e = bucket_time(k, cc, &ms);
				if (e == hipSuccess && ms >= 0.0 && ms < best[k]) {
					best[k] = ms;
					B.bucket[k].chunk = ch;
				}
			}
		}
		// stage 3: launch shape over the whole batch
		if (e == hipSuccess) e = prt::upload_all_tiles(B, st);
		double best_shape = 1e30;
		int shape[3] = {B.block, B.waves_per_cu, B.lds_stack};
		const int blocks[] = {64, 128, 256}, stacks[] = {8, 16}, waves[] = {20, 24, 32};
		for (int bl : blocks)
			for (int sk : stacks)
				for (int wv : waves) {
					if (e != hipSuccess) break;
					B.block = bl;
					B.lds_stack = sk;
					B.waves_per_cu = wv;
					double ms = 0.0;
					e = prt::time_min(st, reps, [&]() { return prt::enqueue_trace(c, B, d_rays6, d_out, st, B.d_tiles, B.n_tiles, B.ray_off.back(), nullptr); }, &ms);
					if (e == hipSuccess && ms < best_shape) {
						best_shape = ms;
						shape[0] = bl;
						shape[1] = wv;
						shape[2] = sk;
					}
				}
		B.block = shape[0];
		B.waves_per_cu = shape[1];
		B.lds_stack = shape[2];
		if (e == hipSuccess) e = prt::time_min(st, reps, whole, &t_after);
		if (e == hipSuccess && t_after > t_before) {  // keep what was there
			for (int k = 0; k < prt::kBatchBuckets; ++k) B.bucket[k] = orig[k];
			B.block = orig_shape[0];
			B.waves_per_cu = orig_shape[1];
			B.lds_stack = orig_shape[2];
			e = prt::upload_params(B, st);
			if (e == hipSuccess) e = prt::upload_all_tiles(B, st);
			if (e == hipSuccess) e = prt::time_min(st, reps, whole, &t_after);
		}
		c->tp.count_stats = saved_count;
		if (e != hipSuccess) return hip_fail(c, e, "batch_tune");
		B.ran = true;
		if (int r = prt::batch_errors(c, B)) return r;
		if (before_ms) *before_ms = t_before;
		if (after_ms) *after_ms = t_after;
		return PRT_OK;
	});
}

}  // extern "C"
