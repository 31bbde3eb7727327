// C ABI of the MI355X nearest-hit backend (declared in include/prt_mi355x.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/prt_mi355x.h"
#include "common.h"
#include "ctx.h"
#include "internal.h"

using prt::fail;
using prt::hip_fail;

namespace {

bool is_gfx950(int dev, std::string *name) {
	hipDeviceProp_t prop;
	if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return false;
	if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return false;
	// the marketing name comes from libdrm's table and is empty where that table is missing
	if (name) *name = std::string(prop.name[0] ? prop.name : "AMD GPU") + " (" + prop.gcnArchName + ")";
	return true;
}

// Checks the caller's record layout against HitReg for the mask.
// Bytes reserved for error messages at context creation.
constexpr size_t kErrCapacity = 512;

int check_layout(prt_ctx *c, uint32_t mask, uint32_t stride, const int32_t off[8]) {
	if (mask > 31u) return fail(c, PRT_ERR_ARG, "field mask has bits above 'valid'");
	const prt::Layout L = prt::layout_of(mask);
	if (stride != static_cast<uint32_t>(L.size))
		return fail(c, PRT_ERR_LAYOUT, "stride " + std::to_string(stride) + " != sizeof(HitReg) " + std::to_string(L.size));
	if (off)
		for (int k = 0; k < 8; ++k)
			if (off[k] != L.off[k])
				return fail(c, PRT_ERR_LAYOUT, "field offset " + std::to_string(k) + " is " + std::to_string(off[k]) +
				                                   ", HitReg has " + std::to_string(L.off[k]));
	return PRT_OK;
}

void apply_margin(prt_ctx *c) {
	const float *b = c->bvh.scene_box;
	float ext = 0.0f;
	for (int k = 0; k < 3; ++k) ext = std::max(ext, b[3 + k] - b[k]);
	if (!std::isfinite(ext)) ext = 0.0f;
	c->tp.margin_rel = std::ldexp(1.0f, -c->margin_log2);
	// absolute part for hits near t = 0: a fixed fraction of the scene extent
	c->tp.margin_abs = std::ldexp(ext, -c->margin_log2 - 6);
}

int do_build(prt_ctx *c, uint64_t n) {
	bool bad = false;
	double ms = 0.0;
	hipError_t e = prt::build_bvh8(c->bs, n, c->bp, c->bvh, c->stream, &bad, &ms);
	if (e != hipSuccess) {
		c->has_scene = false;
		return hip_fail(c, e, "BVH build");
	}
	if (bad) {
		c->has_scene = false;
		return fail(c, PRT_ERR_INPUT, "set_tris: a triangle vertex is not finite");
	}
	c->build_ms = ms;
	c->has_scene = true;
	++c->scene_gen;
	apply_margin(c);
	return PRT_OK;
}

int over_limit(prt_ctx *c, uint64_t bytes, const char *what) {
	if (c->alloc_limit > 0 && bytes > static_cast<uint64_t>(c->alloc_limit))
		return fail(c, PRT_ERR_OOM, std::string(what) + ": " + std::to_string(bytes) + " bytes over the allocation limit");
	return PRT_OK;
}

}  // namespace

extern "C" {

int prt_mi355x_probe(int *ndev, char *name, size_t cap) {
	if (!ndev) return PRT_ERR_ARG;
	*ndev = 0;
	if (name && cap) name[0] = '\0';
	int count = 0;
	if (hipGetDeviceCount(&count) != hipSuccess) {
		(void)hipGetLastError();
		return PRT_OK;
	}
	std::string first;
	for (int d = 0; d < count; ++d) {
		std::string nm;
		if (is_gfx950(d, &nm)) {
			if (*ndev == 0) first = nm;
			++*ndev;
		}
	}
	if (name && cap) std::snprintf(name, cap, "%s", first.c_str());
	return PRT_OK;
}

int prt_mi355x_create(int device, prt_ctx **out) {
	if (!out) return PRT_ERR_ARG;
	*out = nullptr;
	int count = 0;
	if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
		(void)hipGetLastError();
		return PRT_ERR_NODEV;
	}
	std::string nm;
	if (!is_gfx950(device, &nm)) return PRT_ERR_NODEV;
	if (hipSetDevice(device) != hipSuccess) return PRT_ERR_HIP;
	prt_ctx *c = new prt_ctx();
	c->device = device;
	c->name = nm;
	c->err.reserve(kErrCapacity);  // the no-allocation entry point sets messages without allocating
	// the fixed traversal scratch's address on this device (prt_mi355x_nearest_hits_device_noalloc);
	// a failure is reported by that entry point, the others do not use it
	c->noalloc_setup = prt::noalloc_scratch_setup(device);
	// default priority (no priority streams anywhere in this library)
	if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
		delete c;
		return PRT_ERR_HIP;
	}
	if (hipEventCreateWithFlags(&c->busy, hipEventDisableTiming) != hipSuccess) {
		(void)hipStreamDestroy(c->stream);
		delete c;
		return PRT_ERR_HIP;
	}
	*out = c;
	return PRT_OK;
}

void prt_mi355x_destroy(prt_ctx *c) {
	if (!c) return;
	(void)hipSetDevice(c->device);
	(void)hipStreamSynchronize(c->stream);
	c->bvh.nodes_q.release();
	c->bvh.nodes_f.release();
	c->bvh.tris.release();
	prt::DevBuf *bufs[] = {&c->bs.tri_in, &c->bs.tbox, &c->bs.keys[0], &c->bs.keys[1], &c->bs.vals[0],
	                       &c->bs.vals[1], &c->bs.hist, &c->bs.bin, &c->bs.bbox, &c->bs.bin_of,
	                       &c->bs.tri_dest, &c->bs.misc, &c->ts.counter, &c->ts.spill, &c->rays_stage,
	                       &c->out_stage};
	for (auto *b : bufs) b->release();
	if (c->bs.host_misc) (void)hipHostFree(c->bs.host_misc);
	c->pool.reset();
	for (int b = 0; b < 2; ++b) {
		if (c->pin_in[b]) (void)hipHostFree(c->pin_in[b]);
		if (c->pin_out[b]) (void)hipHostFree(c->pin_out[b]);
		hipEvent_t evs[3] = {c->ev_in[b], c->ev_tr[b], c->ev_out[b]};
		for (hipEvent_t ev : evs)
			if (ev) (void)hipEventDestroy(ev);
	}
	if (c->s_in) (void)hipStreamDestroy(c->s_in);
	if (c->s_out) (void)hipStreamDestroy(c->s_out);
	if (c->ts.host_err) (void)hipHostFree(c->ts.host_err);
	c->batch.reset();
	if (c->busy) (void)hipEventDestroy(c->busy);
	(void)hipStreamDestroy(c->stream);
	delete c;
}

const char *prt_mi355x_last_error(const prt_ctx *c) { return c ? c->err.c_str() : "null context"; }

int prt_mi355x_device_name(const prt_ctx *c, char *buf, size_t cap) {
	if (!c || !buf || !cap) return PRT_ERR_ARG;
	std::snprintf(buf, cap, "%s", c->name.c_str());
	return PRT_OK;
}

// Refit from triangles already in c->bs.tri_in.
static int do_refit(prt_ctx *c) {
	bool bad = false;
	double ms = 0.0;
	const hipError_t e = prt::refit_bvh8(c->bs, c->bvh, c->stream, &bad, &ms);
	if (e != hipSuccess) {
		c->has_scene = false;
		return hip_fail(c, e, "BVH refit");
	}
	if (bad) {
		c->has_scene = false;
		return fail(c, PRT_ERR_INPUT, "refit: a triangle vertex is not finite");
	}
	c->refit_ms = ms;
	++c->scene_gen;  // boxes changed: learned clusters must be set up again
	apply_margin(c);
	return PRT_OK;
}

static int refit_common(prt_ctx *c, const float *tris, uint64_t n, hipMemcpyKind kind) {
	if (!c) return PRT_ERR_ARG;
	c->err.clear();
	if (!c->has_scene) return fail(c, PRT_ERR_STATE, "refit called before set_tris");
	if (n != c->bvh.n_tris)
		return fail(c, PRT_ERR_ARG, "refit: " + std::to_string(n) + " triangles, the tree has " + std::to_string(c->bvh.n_tris));
	if (n == 0) return PRT_OK;
	if (!tris) return fail(c, PRT_ERR_ARG, "refit: null triangle pointer");
	if (hipSetDevice(c->device) != hipSuccess) return fail(c, PRT_ERR_HIP, "hipSetDevice");
	hipError_t e = prt::wait_use(c);  // traversals on caller streams still read the old boxes
	if (e == hipSuccess) e = c->bs.tri_in.ensure(n * 36ull);
	if (e == hipSuccess) e = hipMemcpyAsync(c->bs.tri_in.ptr, tris, n * 36ull, kind, c->stream);
	if (e != hipSuccess) return hip_fail(c, e, "refit: triangle copy");
	return do_refit(c);
}

int prt_mi355x_set_tris(prt_ctx *c, const float *tris, uint64_t n) {
	if (!c) return PRT_ERR_ARG;
	c->err.clear();
	if (!tris && n) return fail(c, PRT_ERR_ARG, "set_tris: null triangle pointer");
	if (hipSetDevice(c->device) != hipSuccess) return fail(c, PRT_ERR_HIP, "hipSetDevice");
	if (int r = over_limit(c, n * 36ull, "set_tris")) return r;
	hipError_t e = prt::wait_use(c);  // traversals on caller streams still read the old BVH
	if (e == hipSuccess) e = c->bs.tri_in.ensure(std::max<uint64_t>(1, n) * 36ull);
	if (e != hipSuccess) return hip_fail(c, e, "set_tris: triangle buffer");
	if (n) {
		e = hipMemcpyAsync(c->bs.tri_in.ptr, tris, n * 36ull, hipMemcpyHostToDevice, c->stream);
		if (e != hipSuccess) return hip_fail(c, e, "set_tris: copy");
	}
	return do_build(c, n);
}

int prt_mi355x_refit(prt_ctx *c, const float *tris, uint64_t n) {
	return refit_common(c, tris, n, hipMemcpyHostToDevice);
}

int prt_mi355x_refit_device(prt_ctx *c, const float *d_tris, uint64_t n) {
	return refit_common(c, d_tris, n, hipMemcpyDeviceToDevice);
}

int prt_mi355x_set_tris_device(prt_ctx *c, const float *d_tris, uint64_t n) {
	if (!c) return PRT_ERR_ARG;
	c->err.clear();
	if (!d_tris && n) return fail(c, PRT_ERR_ARG, "set_tris_device: null triangle pointer");
	if (hipSetDevice(c->device) != hipSuccess) return fail(c, PRT_ERR_HIP, "hipSetDevice");
	if (int r = over_limit(c, n * 36ull, "set_tris")) return r;
	hipError_t e = prt::wait_use(c);  // traversals on caller streams still read the old BVH
	if (e == hipSuccess) e = c->bs.tri_in.ensure(std::max<uint64_t>(1, n) * 36ull);
	if (e != hipSuccess) return hip_fail(c, e, "set_tris: triangle buffer");
	if (n) {
		e = hipMemcpyAsync(c->bs.tri_in.ptr, d_tris, n * 36ull, hipMemcpyDeviceToDevice, c->stream);
		if (e != hipSuccess) return hip_fail(c, e, "set_tris: copy");
	}
	return do_build(c, n);
}

static int trace_common(prt_ctx *c, const float *d_rays, uint64_t n, uint32_t mask, void *d_out, hipStream_t st) {
	float ms = 0.0f;
	prt::TraceStats stats;
	hipError_t e = prt::order_use(c, st);
	if (e == hipSuccess)
		e = prt::trace_bvh8(c->bvh, c->ts, c->tp, d_rays, n, mask, d_out, st, c->time_kernel ? &ms : nullptr,
		                    c->tp.count_stats ? &stats : nullptr);
	if (e == hipSuccess) e = prt::mark_use(c, st);
	if (e != hipSuccess) return hip_fail(c, e, "nearest_hits: traversal");
	if (c->time_kernel) c->trace_ms = ms;
	if (c->tp.count_stats) c->stats = stats;
	return PRT_OK;
}

// Rays per chunk of the host-buffer entry point.
constexpr uint64_t kHostChunk = uint64_t{1} << 20;

static hipError_t ensure_pinned(void *buf[2], size_t *have, size_t want) {
	if (want <= *have) return hipSuccess;
	for (int b = 0; b < 2; ++b) {
		if (buf[b]) (void)hipHostFree(buf[b]);
		buf[b] = nullptr;
	}
	*have = 0;
	for (int b = 0; b < 2; ++b) {
		const hipError_t e = hipHostMalloc(&buf[b], want);
		if (e != hipSuccess) return e;
	}
	*have = want;
	return hipSuccess;
}

// Host rays -> records in chunks: chunk i is copied into a pinned buffer by the copy threads
// and sent to the device on s_in while the context's stream traces chunk i - 1 and s_out
// brings its records back; the host copies chunk i - 1's records out of their pinned buffer
// meanwhile. Each chunk's error flags are kept (every traversal launch clears them).
static int chunked_pass(prt_ctx *c, const float *rays, uint64_t n, uint32_t mask, void *out, uint32_t stride);

static int nearest_hits_chunked(prt_ctx *c, const float *rays, uint64_t n, uint32_t mask, void *out, uint32_t stride) {
	if (!c->pool) {
		try {
			c->pool.reset(new prt::CopyPool(prt::CopyPool::default_threads()));
		} catch (...) {
			return fail(c, PRT_ERR_OOM, "nearest_hits: copy threads");
		}
	}
	const int r = chunked_pass(c, rays, n, mask, out, stride);
	if (r != PRT_OK) {
		// nothing may still read or write the pinned buffers when the next call reuses them
		if (c->s_in) (void)hipStreamSynchronize(c->s_in);
		if (c->s_out) (void)hipStreamSynchronize(c->s_out);
		(void)hipStreamSynchronize(c->stream);
	}
	return r;
}

static int chunked_pass(prt_ctx *c, const float *rays, uint64_t n, uint32_t mask, void *out, uint32_t stride) {
	const uint64_t K = (n + kHostChunk - 1) / kHostChunk;
	hipError_t e = hipSuccess;
	e = ensure_pinned(c->pin_in, &c->pin_in_bytes, kHostChunk * 24);
	if (e == hipSuccess) e = ensure_pinned(c->pin_out, &c->pin_out_bytes, kHostChunk * stride);
	if (e == hipSuccess && !c->s_in) e = hipStreamCreateWithFlags(&c->s_in, hipStreamNonBlocking);
	if (e == hipSuccess && !c->s_out) e = hipStreamCreateWithFlags(&c->s_out, hipStreamNonBlocking);
	for (int b = 0; b < 2 && e == hipSuccess; ++b) {
		hipEvent_t *evs[3] = {&c->ev_in[b], &c->ev_tr[b], &c->ev_out[b]};
		for (hipEvent_t *ev : evs)
			if (e == hipSuccess && !*ev) e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
	}
	if (e == hipSuccess) e = c->chunk_err.ensure(K * sizeof(uint32_t));
	if (e != hipSuccess) return hip_fail(c, e, "nearest_hits: staging");
	float *d_rays = c->rays_stage.as<float>();
	uint8_t *d_out = c->out_stage.as<uint8_t>(), *h_out = static_cast<uint8_t *>(out);
	for (uint64_t i = 0; i <= K; ++i) {
		const int b = static_cast<int>(i & 1);
		if (i < K) {
			const uint64_t first = i * kHostChunk, cnt = std::min(kHostChunk, n - first);
			if (i >= 2) e = hipEventSynchronize(c->ev_in[b]);  // the upload of chunk i - 2 has read pin_in[b]
			if (e != hipSuccess) return hip_fail(c, e, "nearest_hits: ray upload");
			c->pool->copy(c->pin_in[b], rays + 6 * first, cnt * 24);
			e = hipMemcpyAsync(d_rays + 6 * first, c->pin_in[b], cnt * 24, hipMemcpyHostToDevice, c->s_in);
			if (e == hipSuccess) e = hipEventRecord(c->ev_in[b], c->s_in);
			if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, c->ev_in[b], 0);
			if (e != hipSuccess) return hip_fail(c, e, "nearest_hits: ray upload");
			if (int r = trace_common(c, d_rays + 6 * first, cnt, mask, d_out + first * stride, c->stream)) return r;
			e = hipMemcpyAsync(c->chunk_err.as<uint32_t>() + i, c->ts.counter.as<uint32_t>() + 8, sizeof(uint32_t),
			                   hipMemcpyDeviceToDevice, c->stream);
			if (e == hipSuccess) e = hipEventRecord(c->ev_tr[b], c->stream);
			if (e == hipSuccess) e = hipStreamWaitEvent(c->s_out, c->ev_tr[b], 0);
			// pin_out[b] was emptied by the host when it copied out chunk i - 2 (previous iteration)
			if (e == hipSuccess)
				e = hipMemcpyAsync(c->pin_out[b], d_out + first * stride, cnt * stride, hipMemcpyDeviceToHost, c->s_out);
			if (e == hipSuccess) e = hipEventRecord(c->ev_out[b], c->s_out);
			if (e != hipSuccess) return hip_fail(c, e, "nearest_hits: record download");
		}
		if (i >= 1) {
			const uint64_t j = i - 1, first = j * kHostChunk, cnt = std::min(kHostChunk, n - first);
			e = hipEventSynchronize(c->ev_out[j & 1]);
			if (e != hipSuccess) return hip_fail(c, e, "nearest_hits: record download");
			c->pool->copy(h_out + first * stride, c->pin_out[j & 1], cnt * stride);
		}
	}
	std::vector<uint32_t> errs(K);
	e = hipMemcpyAsync(errs.data(), c->chunk_err.ptr, K * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
	if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
	if (e != hipSuccess) return hip_fail(c, e, "nearest_hits: error flags");
	uint32_t flags = 0;
	for (uint32_t f : errs) flags |= f;
	if (flags)
		return fail(c, PRT_ERR_HIP, "nearest_hits: traversal found a corrupt BVH entry (flags " + std::to_string(flags) + ")");
	return PRT_OK;
}

int prt_mi355x_nearest_hits(prt_ctx *c, const float *rays, uint64_t n, uint32_t mask, void *out, uint32_t stride,
                            const int32_t off[8]) {
	return prt::guarded(c, [&]() -> int {
		if (!c) return PRT_ERR_ARG;
		c->err.clear();
		if (int r = check_layout(c, mask, stride, off)) return r;
		if (!c->has_scene) return fail(c, PRT_ERR_STATE, "nearest_hits called before set_tris");
		if (n == 0) return PRT_OK;
		if (!rays || !out) return fail(c, PRT_ERR_ARG, "nearest_hits: null pointer");
		if (hipSetDevice(c->device) != hipSuccess) return fail(c, PRT_ERR_HIP, "hipSetDevice");
		const auto t0 = std::chrono::steady_clock::now();
		if (int r = over_limit(c, n * (24ull + stride), "nearest_hits")) return r;
		hipError_t e = c->rays_stage.ensure(n * 24ull);
		if (e == hipSuccess) e = c->out_stage.ensure(n * static_cast<uint64_t>(stride));
		if (e != hipSuccess) return hip_fail(c, e, "nearest_hits: staging buffers");
		if (n > kHostChunk && !c->tp.count_stats && !c->time_kernel) {
			// large calls: pipelined through pinned buffers (statistics and kernel timing describe
			// one launch, so those modes take the single-launch path below)
			const int r = nearest_hits_chunked(c, rays, n, mask, out, stride);
			if (r == PRT_OK) c->api_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
			return r;
		}
		e = hipMemcpyAsync(c->rays_stage.ptr, rays, n * 24ull, hipMemcpyHostToDevice, c->stream);
		if (e != hipSuccess) return hip_fail(c, e, "nearest_hits: ray copy");
		if (int r = trace_common(c, c->rays_stage.as<float>(), n, mask, c->out_stage.ptr, c->stream)) return r;
		e = hipMemcpyAsync(out, c->out_stage.ptr, n * static_cast<uint64_t>(stride), hipMemcpyDeviceToHost, c->stream);
		uint32_t flags = 0;
		if (e == hipSuccess) e = prt::trace_errors(c->ts, c->stream, &flags);
		if (e != hipSuccess) return hip_fail(c, e, "nearest_hits: result copy");
		if (flags)
			return fail(c, PRT_ERR_HIP, "nearest_hits: traversal found a corrupt BVH entry (flags " + std::to_string(flags) + ")");
		c->api_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
		return PRT_OK;
	});
}

int prt_mi355x_nearest_hits_device(prt_ctx *c, const float *d_rays, uint64_t n, uint32_t mask, void *d_out,
                                   uint32_t stride, const int32_t off[8], void *stream) {
	if (!c) return PRT_ERR_ARG;
	c->err.clear();
	if (int r = check_layout(c, mask, stride, off)) return r;
	if (!c->has_scene) return fail(c, PRT_ERR_STATE, "nearest_hits called before set_tris");
	if (n == 0) return PRT_OK;
	if (!d_rays || !d_out) return fail(c, PRT_ERR_ARG, "nearest_hits_device: null pointer");
	if (reinterpret_cast<uintptr_t>(d_rays) % 4 || reinterpret_cast<uintptr_t>(d_out) % 16)
		return fail(c, PRT_ERR_ARG, "nearest_hits_device: rays need 4-byte, records 16-byte alignment");
	if (hipSetDevice(c->device) != hipSuccess) return fail(c, PRT_ERR_HIP, "hipSetDevice");
	hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
	return trace_common(c, d_rays, n, mask, d_out, st);
}

// Sets a message without allocating (the context reserved kErrCapacity bytes).
static int fail_lit(prt_ctx *c, int code, const char *msg) {
	c->err.assign(msg);
	return code;
}

int prt_mi355x_nearest_hits_device_noalloc(prt_ctx *c, const float *d_rays, uint64_t n, uint32_t mask, void *d_out,
                                           uint32_t stride, const int32_t off[8], void *stream) {
	if (!c) return PRT_ERR_ARG;
	c->err.clear();
	if (mask > 31u) return fail_lit(c, PRT_ERR_ARG, "field mask has bits above 'valid'");
	const prt::Layout L = prt::layout_of(mask);
	if (stride != static_cast<uint32_t>(L.size)) return fail_lit(c, PRT_ERR_LAYOUT, "stride != sizeof(HitReg)");
	if (off)
		for (int k = 0; k < 8; ++k)
			if (off[k] != L.off[k]) return fail_lit(c, PRT_ERR_LAYOUT, "field offsets differ from HitReg");
	if (!c->has_scene) return fail_lit(c, PRT_ERR_STATE, "nearest_hits called before set_tris");
	if (n == 0) return PRT_OK;
	if (!d_rays || !d_out) return fail_lit(c, PRT_ERR_ARG, "nearest_hits_device_noalloc: null pointer");
	if (reinterpret_cast<uintptr_t>(d_rays) % 4 || reinterpret_cast<uintptr_t>(d_out) % 16)
		return fail_lit(c, PRT_ERR_ARG, "nearest_hits_device_noalloc: rays need 4-byte, records 16-byte alignment");
	if (hipSetDevice(c->device) != hipSuccess) return fail_lit(c, PRT_ERR_HIP, "hipSetDevice");
	void *scratch = prt::noalloc_scratch(c->device);
	if (!scratch) {  // set-up failed at create: once more, on this first call
		c->noalloc_setup = prt::noalloc_scratch_setup(c->device);
		scratch = prt::noalloc_scratch(c->device);
	}
	if (!scratch) return fail_lit(c, PRT_ERR_HIP, "nearest_hits_device_noalloc: scratch address lookup failed");
	hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
	hipError_t e = prt::order_use(c, st);
	uint64_t need = 0;
	if (e == hipSuccess) e = prt::trace_bvh8_noalloc(c->bvh, c->ts, c->tp, d_rays, n, mask, d_out, st, scratch, &need);
	if (e == hipErrorInvalidValue && need > prt::kNoAllocScratchBytes)
		return fail_lit(c, PRT_ERR_ARG,
		                "nearest_hits_device_noalloc: the BVH is too deep for the 128 MiB traversal scratch (one "
		                "workgroup's stack spill does not fit)");
	if (e == hipSuccess) e = prt::mark_use(c, st);
	if (e != hipSuccess) return fail_lit(c, PRT_ERR_HIP, "nearest_hits_device_noalloc: traversal launch failed");
	c->used_noalloc = true;
	return PRT_OK;
}

int prt_mi355x_synchronize(prt_ctx *c) {
	if (!c) return PRT_ERR_ARG;
	hipError_t e = hipDeviceSynchronize();  // device entry points may use caller streams
	uint32_t flags = 0, flags_na = 0;
	if (e == hipSuccess) e = prt::trace_errors(c->ts, c->stream, &flags);
	if (e == hipSuccess && c->used_noalloc)
		e = prt::trace_errors_at(c->ts, static_cast<const uint32_t *>(prt::noalloc_scratch(c->device)), c->stream,
		                         &flags_na);
	flags |= flags_na;
	if (e != hipSuccess) return hip_fail(c, e, "synchronize");
	if (flags) return fail(c, PRT_ERR_HIP, "traversal found a corrupt BVH entry (flags " + std::to_string(flags) + ")");
	return PRT_OK;
}

int prt_mi355x_set_option(prt_ctx *c, int opt, int64_t v) {
	if (!c) return PRT_ERR_ARG;
	auto bad = [&]() { return fail(c, PRT_ERR_ARG, "option " + std::to_string(opt) + " value " + std::to_string(v)); };
	switch (opt) {
	case PRT_OPT_NODE_FORMAT:
		if (v != 0 && v != 1) return bad();
		c->tp.node_format = static_cast<int>(v);
		break;
	case PRT_OPT_CULL:
		if (v != 0 && v != 1) return bad();
		c->tp.cull = static_cast<int>(v);
		break;
	case PRT_OPT_LEAF_SIZE:
		if (v < 1 || v > prt::kMaxLeaf) return bad();
		c->bp.leaf_size = static_cast<int>(v);
		break;
	case PRT_OPT_BLOCK:
		if (v != 64 && v != 128 && v != 256) return bad();
		c->tp.block = static_cast<int>(v);
		break;
	case PRT_OPT_WAVES_PER_CU:
		if (v < 0 || v > 64) return bad();
		c->tp.waves_per_cu = static_cast<int>(v);
		break;
	case PRT_OPT_LDS_STACK:
		if (v != 8 && v != 16) return bad();
		c->tp.lds_stack = static_cast<int>(v);
		break;
	case PRT_OPT_CHUNK:
		if (v < 64 || v > 65536) return bad();
		c->tp.chunk = static_cast<int>(v);
		break;
	case PRT_OPT_LEAF_PHASE:
		if (v < 1 || v > 64) return bad();
		c->tp.leaf_phase = static_cast<int>(v);
		break;
	case PRT_OPT_REFILL:
		if (v < 1 || v > 64) return bad();
		c->tp.refill = static_cast<int>(v);
		break;
	case PRT_OPT_COUNT_STATS: c->tp.count_stats = v ? 1 : 0; break;
	case PRT_OPT_TIME_KERNEL: c->time_kernel = v ? 1 : 0; break;
	case PRT_OPT_ALLOC_LIMIT:
		if (v < 0) return bad();
		c->alloc_limit = v;
		break;
	case PRT_OPT_MARGIN_REL_LOG2:
		// the build's never-cull flag covers leaf-test errors up to 2^-10 |t|: a smaller margin
		// would not be exact
		if (v < 4 || v > 10) return bad();
		c->margin_log2 = static_cast<int>(v);
		apply_margin(c);
		break;
	case PRT_OPT_LEAF_TEST:
		if (v != 0 && v != 1) return bad();
		c->tp.leaf_test = static_cast<int>(v);
		break;
	case PRT_OPT_SEEDS:
		if (v < 1 || v > 4) return bad();
		c->tp.n_seeds = static_cast<int>(v);
		break;
	case PRT_OPT_BUILDER:
		if (v != 0 && v != 1) return bad();
		c->bp.builder = static_cast<int>(v);
		break;
	case PRT_OPT_COLLAPSE:
		if (v != 0 && v != 1) return bad();
		c->bp.collapse = static_cast<int>(v);
		break;
	default: return fail(c, PRT_ERR_ARG, "unknown option " + std::to_string(opt));
	}
	return PRT_OK;
}

int prt_mi355x_get_option(const prt_ctx *c, int opt, int64_t *v) {
	if (!c || !v) return PRT_ERR_ARG;
	switch (opt) {
	case PRT_OPT_NODE_FORMAT: *v = c->tp.node_format; break;
	case PRT_OPT_CULL: *v = c->tp.cull; break;
	case PRT_OPT_LEAF_SIZE: *v = c->bp.leaf_size; break;
	case PRT_OPT_BLOCK: *v = c->tp.block; break;
	case PRT_OPT_WAVES_PER_CU: *v = c->tp.waves_per_cu; break;
	case PRT_OPT_LDS_STACK: *v = c->tp.lds_stack; break;
	case PRT_OPT_CHUNK: *v = c->tp.chunk; break;
	case PRT_OPT_LEAF_PHASE: *v = c->tp.leaf_phase; break;
	case PRT_OPT_REFILL: *v = c->tp.refill; break;
	case PRT_OPT_COUNT_STATS: *v = c->tp.count_stats; break;
	case PRT_OPT_TIME_KERNEL: *v = c->time_kernel; break;
	case PRT_OPT_ALLOC_LIMIT: *v = c->alloc_limit; break;
	case PRT_OPT_MARGIN_REL_LOG2: *v = c->margin_log2; break;
	case PRT_OPT_BUILDER: *v = c->bp.builder; break;
	case PRT_OPT_COLLAPSE: *v = c->bp.collapse; break;
	case PRT_OPT_LEAF_TEST: *v = c->tp.leaf_test; break;
	case PRT_OPT_SEEDS: *v = c->tp.n_seeds; break;
	default: return PRT_ERR_ARG;
	}
	return PRT_OK;
}

int prt_mi355x_get_stat(const prt_ctx *c, int stat, double *v) {
	if (!c || !v) return PRT_ERR_ARG;
	switch (stat) {
	case PRT_STAT_BUILD_MS: *v = c->build_ms; break;
	case PRT_STAT_REFIT_MS: *v = c->refit_ms; break;
	case PRT_STAT_NCULL_TRIS: *v = c->bvh.n_ncull == prt::kNcullUnknown ? -1.0 : static_cast<double>(c->bvh.n_ncull); break;
	case PRT_STAT_TRACE_MS: *v = c->trace_ms; break;
	case PRT_STAT_NODES: *v = c->bvh.n_nodes; break;
	case PRT_STAT_DEPTH: *v = c->bvh.depth; break;
	case PRT_STAT_NODE_VISITS: *v = static_cast<double>(c->stats.node_visits); break;
	case PRT_STAT_TRI_TESTS: *v = static_cast<double>(c->stats.tri_tests); break;
	case PRT_STAT_LEAF_VISITS: *v = static_cast<double>(c->stats.leaf_visits); break;
	case PRT_STAT_NODE_BYTES: *v = c->tp.node_format == 0 ? sizeof(prt::NodeF) : sizeof(prt::NodeQ); break;
	case PRT_STAT_API_MS: *v = c->api_ms; break;
	case PRT_STAT_N_TRIS: *v = static_cast<double>(c->bvh.n_tris); break;
	case PRT_STAT_NODE_PHASES: *v = static_cast<double>(c->stats.node_phases); break;
	case PRT_STAT_TRI_PHASES: *v = static_cast<double>(c->stats.tri_phases); break;
	case PRT_STAT_CULLED_POPS: *v = static_cast<double>(c->stats.culled_pops); break;
	case PRT_STAT_MLP_MS: *v = c->mlp_ms; break;
	case PRT_STAT_SEEDED_MS: *v = c->seeded_ms; break;
	case PRT_STAT_TRAIN_MS: *v = c->train_ms; break;
	case PRT_STAT_TRAIN_LOSS: *v = c->train_loss; break;
	case PRT_STAT_CLUSTERS: *v = c->clusters.count; break;
	case PRT_STAT_SEED_TOP1: *v = c->seed_top1; break;
	case PRT_STAT_SEED_TOP4: *v = c->seed_top4; break;
	default: return PRT_ERR_ARG;
	}
	return PRT_OK;
}

int prt_mi355x_layout(uint32_t mask, int32_t out[9]) {
	if (!out || mask > 31u) return PRT_ERR_ARG;
	const prt::Layout L = prt::layout_of(mask);
	out[0] = L.size;
	for (int k = 0; k < 8; ++k) out[1 + k] = L.off[k];
	return PRT_OK;
}

int prt_mi355x_export(prt_ctx *c, int which, void *out, uint64_t *bytes) {
	if (!c || !bytes) return PRT_ERR_ARG;
	const prt::DevBuf *b = nullptr;
	uint64_t sz = 0;
	switch (which) {
	case 0: b = &c->bvh.nodes_q; sz = sizeof(prt::NodeQ) * uint64_t(c->bvh.n_nodes); break;
	case 1: b = &c->bvh.nodes_f; sz = sizeof(prt::NodeF) * uint64_t(c->bvh.n_nodes); break;
	case 2: b = &c->bvh.tris; sz = sizeof(prt::TriRec) * c->bvh.n_tris; break;
	default: return fail(c, PRT_ERR_ARG, "export: unknown buffer");
	}
	if (!c->has_scene) sz = 0;
	*bytes = sz;
	if (!out || sz == 0) return PRT_OK;
	hipError_t e = hipMemcpyAsync(out, b->ptr, sz, hipMemcpyDeviceToHost, c->stream);
	if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
	return e == hipSuccess ? PRT_OK : hip_fail(c, e, "export");
}

}  // extern "C"
