// The C ABI context (prt_ctx) shared by api.hip, learned.hip and batch.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <new>
#include <string>

#include "../../include/prt_mi355x.h"
#include "batch.h"
#include "host_copy.h"
#include "internal.h"
#include "mlp.h"

struct prt_ctx {
	int device = 0;
	hipStream_t stream = nullptr;
	// recorded after every traversal, on the stream it was enqueued on (busy_stream): rebuilding
	// the BVH and traversals on other streams wait for it
	hipEvent_t busy = nullptr;
	bool busy_set = false;
	hipStream_t busy_stream = nullptr;
	std::string name;
	std::string err;
	prt::Bvh8 bvh;
	prt::BuildScratch bs;
	prt::BuildParams bp;
	prt::TraceParams tp;
	prt::TraceScratch ts;
	prt::DevBuf rays_stage, out_stage;
	// host-buffer entry point, in chunks: pinned double buffers, copy threads, and the streams
	// that move data while the context's stream traces (created at the first large call)
	std::unique_ptr<prt::CopyPool> pool;
	void *pin_in[2] = {nullptr, nullptr}, *pin_out[2] = {nullptr, nullptr};
	size_t pin_in_bytes = 0, pin_out_bytes = 0;
	hipStream_t s_in = nullptr, s_out = nullptr;
	hipEvent_t ev_in[2] = {nullptr, nullptr}, ev_tr[2] = {nullptr, nullptr}, ev_out[2] = {nullptr, nullptr};
	prt::DevBuf chunk_err;
	bool has_scene = false;
	// traversal without allocation: result of the scratch address look-up at creation, and
	// whether this context has launched with that scratch (its error flags live there)
	hipError_t noalloc_setup = hipSuccess;
	bool used_noalloc = false;
	int time_kernel = 0;
	int margin_log2 = 10;
	int64_t alloc_limit = 0;
	double build_ms = 0.0, trace_ms = 0.0, api_ms = 0.0, refit_ms = 0.0;
	prt::TraceStats stats;
	// learned seeding
	prt::Mlp mlp;
	prt::Clusters clusters;
	prt::DevBuf top4, labels, hits_tmp;
	double mlp_ms = 0.0, seeded_ms = 0.0, train_ms = 0.0, train_loss = 0.0;
	double seed_top1 = 0.0, seed_top4 = 0.0;
	uint64_t learned_scene = 0;  // set_tris generation the clusters were built for
	uint64_t scene_gen = 0;
	// batched multi-scene tracing
	std::unique_ptr<prt::Batch> batch;
};

namespace prt {

inline int fail(prt_ctx *c, int code, const std::string &msg) {
	if (c) c->err = msg;
	return code;
}

// Marks the BVH, the traversal scratch and the learned buffers as used by work enqueued on `st`.
inline hipError_t mark_use(prt_ctx *c, hipStream_t st) {
	if (!c->busy) return hipSuccess;
	c->busy_set = true;
	c->busy_stream = st;
	return hipEventRecord(c->busy, st);
}

// Makes `st` wait for the last traversal when it was enqueued on another stream: the launches
// of one context share its work counter, stack spill buffer and learned buffers, so calls on
// different streams run one after the other.
inline hipError_t order_use(prt_ctx *c, hipStream_t st) {
	if (!c->busy_set || c->busy_stream == st) return hipSuccess;
	return hipStreamWaitEvent(st, c->busy, 0);
}

// Makes the context's stream wait for the last traversal (before the BVH is rebuilt or moved).
inline hipError_t wait_use(prt_ctx *c) {
	const hipError_t e = order_use(c, c->stream);
	c->busy_set = false;
	return e;
}

// Runs an entry point's body; a failed host allocation (or any other C++ exception) becomes an
// error code instead of crossing the C ABI.
template <class F>
inline int guarded(prt_ctx *c, F &&body) {
	try {
		return body();
	} catch (const std::bad_alloc &) {
		return fail(c, PRT_ERR_OOM, "host memory allocation failed");
	} catch (...) {
		return fail(c, PRT_ERR_HIP, "unexpected host error");
	}
}

inline int hip_fail(prt_ctx *c, hipError_t e, const char *what) {
	const int code = (e == hipErrorOutOfMemory) ? PRT_ERR_OOM : PRT_ERR_HIP;
	return fail(c, code, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace prt
