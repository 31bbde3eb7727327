// Learned t_max seeding around the plain traversal (no counterpart in the reference):
//   * the BVH8 is cut into at most 4096 subtree clusters (largest subtree split first);
//   * training rays are traced with the plain traversal, the hit triangle's cluster (or "miss")
//     is the label, and the MLP of mlp.hip is trained with softmax cross-entropy and Adam;
//   * at inference the 4 best clusters per ray are pushed above the root on the traversal stack,
//     so their subtrees are intersected exactly first and seed the culling distance, and the
//     root traversal that follows keeps the result identical to the plain traversal.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <queue>
#include <vector>

#include "../../include/prt_mi355x.h"
#include "common.h"
#include "ctx.h"
#include "internal.h"
#include "mlp.h"

namespace prt {
namespace {

#define PRT_TRY(x)                                                                                 \
	do {                                                                                           \
		hipError_t e_ = (x);                                                                       \
		if (e_ != hipSuccess) return e_;                                                           \
	} while (0)

// label = cluster of the hit triangle, kClusters for a miss (records are HitReg<t, primitive_id>)
__global__ __launch_bounds__(256) void k_labels(const uint4 *__restrict__ hits, uint64_t n,
                                                const uint32_t *__restrict__ tri_cluster, uint32_t n_tris,
                                                uint32_t *__restrict__ labels) {
	const uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
	if (i >= n) return;
	const uint4 h = hits[i];
	const float t = __uint_as_float(h.y);
	labels[i] = (t < __builtin_inff() && h.z < n_tris) ? tri_cluster[h.z] : static_cast<uint32_t>(kClusters);
}

// counts rays whose true label is the top-1 / among the top-4 predictions
__global__ __launch_bounds__(256) void k_seed_quality(const uint32_t *__restrict__ labels, const uint4 *__restrict__ top4,
                                                      uint64_t n, unsigned long long *cnt) {
	const uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
	uint32_t c1 = 0, c4 = 0;
	if (i < n) {
		const uint32_t y = labels[i];
		const uint4 t = top4[i];
		c1 = t.x == y;
		c4 = (t.x == y) | (t.y == y) | (t.z == y) | (t.w == y);
	}
	for (int off = 32; off > 0; off >>= 1) {
		c1 += __shfl_xor(c1, off);
		c4 += __shfl_xor(c4, off);
	}
	if ((threadIdx.x & 63) == 0) {
		atomicAdd(cnt, static_cast<unsigned long long>(c1));
		atomicAdd(cnt + 1, static_cast<unsigned long long>(c4));
	}
}

inline unsigned blocks(uint64_t n) { return static_cast<unsigned>((n + 255) / 256); }

// Host: cut the BVH8 into clusters and map every primitive to its cluster.
hipError_t build_clusters(prt_ctx *c) {
	const Bvh8 &b = c->bvh;
	const uint32_t nn = b.n_nodes;
	std::vector<NodeF> nodes(nn);
	std::vector<TriRec> tris(b.n_tris);
	PRT_TRY(hipMemcpyAsync(nodes.data(), b.nodes_f.ptr, sizeof(NodeF) * nn, hipMemcpyDeviceToHost, c->stream));
	PRT_TRY(hipMemcpyAsync(tris.data(), b.tris.ptr, sizeof(TriRec) * b.n_tris, hipMemcpyDeviceToHost, c->stream));
	PRT_TRY(hipStreamSynchronize(c->stream));
	// triangles per subtree; children always have larger indices than their parent
	std::vector<uint64_t> cnt(nn, 0);
	for (int64_t i = static_cast<int64_t>(nn) - 1; i >= 0; --i) {
		uint64_t s = 0;
		for (int k = 0; k < 8; ++k) {
			const uint32_t m = nodes[i].meta[k];
			if (!meta_used(m)) continue;
			s += meta_internal(m) ? cnt[nodes[i].child_base + meta_offset(m)] : meta_count(m);
		}
		cnt[i] = s;
	}
	struct Item {
		uint64_t count;
		uint32_t entry;
		float box[6];
		bool operator<(const Item &o) const { return count < o.count; }
	};
	std::priority_queue<Item> open;  // internal subtrees that may still be split
	std::vector<Item> done;          // leaves and subtrees kept whole
	{
		Item root{cnt[0], 0u, {}};
		for (int a = 0; a < 3; ++a) {
			root.box[a] = __builtin_inff();
			root.box[3 + a] = -__builtin_inff();
		}
		for (int k = 0; k < 8; ++k)
			if (meta_used(nodes[0].meta[k]))
				for (int a = 0; a < 3; ++a) {
					root.box[a] = std::min(root.box[a], nodes[0].lo[a][k]);
					root.box[3 + a] = std::max(root.box[3 + a], nodes[0].hi[a][k]);
				}
		open.push(root);
	}
	while (!open.empty()) {
		const Item it = open.top();
		const NodeF &n = nodes[it.entry];
		int kids = 0;
		for (int k = 0; k < 8; ++k) kids += meta_used(n.meta[k]);
		if (open.size() + done.size() - 1 + kids > static_cast<size_t>(kClusters)) break;
		open.pop();
		for (int k = 0; k < 8; ++k) {
			const uint32_t m = n.meta[k];
			if (!meta_used(m)) continue;
			Item ch{};
			for (int a = 0; a < 3; ++a) {
				ch.box[a] = n.lo[a][k];
				ch.box[3 + a] = n.hi[a][k];
			}
			if (meta_internal(m)) {
				ch.entry = n.child_base + meta_offset(m);
				ch.count = cnt[ch.entry];
				open.push(ch);
			} else {
				ch.entry = leaf_entry(n.tri_base + meta_offset(m), meta_count(m));
				ch.count = meta_count(m);
				done.push_back(ch);
			}
		}
	}
	while (!open.empty()) {
		done.push_back(open.top());
		open.pop();
	}
	// primitive -> cluster by walking each cluster's subtree
	std::vector<uint32_t> tri_cluster(b.n_tris, 0xFFFFFFFFu);
	std::vector<uint32_t> entries(done.size());
	std::vector<float> boxes(done.size() * 6);
	std::vector<uint32_t> stack;
	for (size_t ci = 0; ci < done.size(); ++ci) {
		entries[ci] = done[ci].entry;
		std::memcpy(&boxes[ci * 6], done[ci].box, sizeof(float) * 6);
		stack.assign(1, done[ci].entry);
		while (!stack.empty()) {
			const uint32_t e = stack.back();
			stack.pop_back();
			if (e & kLeafBit) {
				const uint32_t first = leaf_first(e), count = leaf_count(e);
				for (uint32_t q = 0; q < count; ++q) tri_cluster[tris[first + q].id] = static_cast<uint32_t>(ci);
			} else {
				const NodeF &n = nodes[e];
				for (int k = 0; k < 8; ++k) {
					const uint32_t m = n.meta[k];
					if (!meta_used(m)) continue;
					stack.push_back(meta_internal(m) ? n.child_base + meta_offset(m)
					                                 : leaf_entry(n.tri_base + meta_offset(m), meta_count(m)));
				}
			}
		}
	}
	Clusters &cl = c->clusters;
	cl.count = static_cast<uint32_t>(done.size());
	PRT_TRY(cl.entry.ensure(sizeof(uint32_t) * std::max<size_t>(1, done.size())));
	PRT_TRY(cl.box.ensure(sizeof(float) * 6 * std::max<size_t>(1, done.size())));
	PRT_TRY(cl.tri_cluster.ensure(sizeof(uint32_t) * std::max<uint64_t>(1, b.n_tris)));
	PRT_TRY(hipMemcpyAsync(cl.entry.ptr, entries.data(), sizeof(uint32_t) * done.size(), hipMemcpyHostToDevice, c->stream));
	PRT_TRY(hipMemcpyAsync(cl.box.ptr, boxes.data(), sizeof(float) * 6 * done.size(), hipMemcpyHostToDevice, c->stream));
	PRT_TRY(hipMemcpyAsync(cl.tri_cluster.ptr, tri_cluster.data(), sizeof(uint32_t) * b.n_tris, hipMemcpyHostToDevice,
	                       c->stream));
	PRT_TRY(hipStreamSynchronize(c->stream));
	cl.host_tri_cluster.swap(tri_cluster);
	return hipSuccess;
}

// Labels for n device rays from the plain traversal.
hipError_t make_labels(prt_ctx *c, const float *d_rays, uint64_t n, uint32_t *labels) {
	PRT_TRY(c->hits_tmp.ensure(16 * n));
	PRT_TRY(order_use(c, c->stream));
	PRT_TRY(trace_bvh8(c->bvh, c->ts, c->tp, d_rays, n, F_T | F_ID, c->hits_tmp.ptr, c->stream, nullptr, nullptr));
	PRT_TRY(mark_use(c, c->stream));
	hipLaunchKernelGGL(k_labels, dim3(blocks(n)), dim3(256), 0, c->stream, c->hits_tmp.as<uint4>(), n,
	                   c->clusters.tri_cluster.as<uint32_t>(), static_cast<uint32_t>(c->bvh.n_tris), labels);
	return hipGetLastError();
}

constexpr uint64_t kInferChunk = 1u << 24;  // rays per inference launch: 8 Mi rays in one (1% faster than 1 Mi launches)

hipError_t infer_all(prt_ctx *c, const float *d_rays, uint64_t n, hipStream_t st) {
	PRT_TRY(c->top4.ensure(16 * std::max<uint64_t>(1, n)));
	for (uint64_t f = 0; f < n; f += kInferChunk) {
		const int rows = static_cast<int>(std::min<uint64_t>(kInferChunk, n - f));
		PRT_TRY(mlp_infer_top4(c->mlp, d_rays, f, rows, c->top4.as<uint32_t>() + 4 * f, st));
	}
	return hipSuccess;
}

int check_learned(prt_ctx *c) {
	if (!c->has_scene) return fail(c, PRT_ERR_STATE, "learned: set_tris first");
	if (!c->mlp.ready || c->learned_scene != c->scene_gen)
		return fail(c, PRT_ERR_STATE, "learned: call prt_mi355x_learned_setup after set_tris");
	return PRT_OK;
}

}  // namespace
}  // namespace prt

using namespace prt;

extern "C" {

int prt_mi355x_learned_setup(prt_ctx *c, uint64_t seed) {
	return prt::guarded(c, [&]() -> int {
		if (!c) return PRT_ERR_ARG;
		c->err.clear();
		if (!c->has_scene) return fail(c, PRT_ERR_STATE, "learned_setup: set_tris first");
		if (c->bvh.n_tris == 0) return fail(c, PRT_ERR_STATE, "learned_setup: empty scene");
		if (hipSetDevice(c->device) != hipSuccess) return fail(c, PRT_ERR_HIP, "hipSetDevice");
		hipError_t e = build_clusters(c);
		if (e != hipSuccess) return hip_fail(c, e, "learned_setup: clusters");
		for (int a = 0; a < 3; ++a) {
			const float lo = c->bvh.scene_box[a], hi = c->bvh.scene_box[3 + a];
			c->mlp.scene_lo[a] = lo;
			c->mlp.scene_inv_ext[a] = hi > lo ? 1.0f / (hi - lo) : 1.0f;
		}
		e = mlp_init(c->mlp, seed, c->stream);
		if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
		if (e != hipSuccess) return hip_fail(c, e, "learned_setup: network");
		c->learned_scene = c->scene_gen;
		return PRT_OK;
	});
}

int prt_mi355x_learned_train(prt_ctx *c, const float *d_rays, uint64_t n, int epochs, int batch, float lr) {
	if (!c) return PRT_ERR_ARG;
	c->err.clear();
	if (int r = check_learned(c)) return r;
	if (!d_rays || n < static_cast<uint64_t>(batch) || batch < 32 || batch % 32 || epochs < 1 || !(lr > 0.0f))
		return fail(c, PRT_ERR_ARG, "learned_train: need rays, batch a positive multiple of 32 <= n, epochs >= 1, lr > 0");
	if (hipSetDevice(c->device) != hipSuccess) return fail(c, PRT_ERR_HIP, "hipSetDevice");
	const auto t0 = std::chrono::steady_clock::now();
	hipError_t e = c->labels.ensure(4 * n);
	if (e == hipSuccess) e = make_labels(c, d_rays, n, c->labels.as<uint32_t>());
	const uint64_t steps = n / static_cast<uint64_t>(batch);
	for (int ep = 0; ep < epochs && e == hipSuccess; ++ep) {
		for (uint64_t s = 0; s < steps && e == hipSuccess; ++s) {
			const uint64_t first = s * batch;
			e = mlp_encode(c->mlp, d_rays, first, batch, c->stream);
			if (e == hipSuccess) e = mlp_forward_logits(c->mlp, batch, c->stream);
			if (e == hipSuccess) e = mlp_backward(c->mlp, c->labels.as<uint32_t>() + first, batch, c->stream);
			if (e == hipSuccess) e = mlp_adam(c->mlp, lr, c->stream);
		}
	}
	float loss = 0.0f;
	if (e == hipSuccess) e = hipMemcpyAsync(&loss, c->mlp.loss.ptr, 4, hipMemcpyDeviceToHost, c->stream);
	if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
	if (e != hipSuccess) return hip_fail(c, e, "learned_train");
	c->train_loss = loss;
	c->train_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
	return PRT_OK;
}

int prt_mi355x_nearest_hits_learned_device(prt_ctx *c, const float *d_rays, uint64_t n, uint32_t mask, void *d_out,
                                           uint32_t stride, const int32_t off[8], void *stream) {
	if (!c) return PRT_ERR_ARG;
	c->err.clear();
	if (int r = check_learned(c)) return r;
	if (mask > 31u) return fail(c, PRT_ERR_ARG, "field mask has bits above 'valid'");
	const Layout L = layout_of(mask);
	if (stride != static_cast<uint32_t>(L.size)) return fail(c, PRT_ERR_LAYOUT, "stride != sizeof(HitReg)");
	if (off)
		for (int k = 0; k < 8; ++k)
			if (off[k] != L.off[k]) return fail(c, PRT_ERR_LAYOUT, "field offsets differ from HitReg");
	if (n == 0) return PRT_OK;
	if (!d_rays || !d_out) return fail(c, PRT_ERR_ARG, "null pointer");
	if (reinterpret_cast<uintptr_t>(d_rays) % 4 || reinterpret_cast<uintptr_t>(d_out) % 16)
		return fail(c, PRT_ERR_ARG, "rays need 4-byte, records 16-byte alignment");
	if (c->tp.leaf_test != 0)
		return fail(c, PRT_ERR_STATE, "learned seeding uses the Moller-Trumbore leaf test (set leaf_test to 0)");
	if (hipSetDevice(c->device) != hipSuccess) return fail(c, PRT_ERR_HIP, "hipSetDevice");
	hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
	hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
	hipError_t e = order_use(c, st);  // the inference buffers and the traversal scratch are shared
	if (e == hipSuccess && c->time_kernel) {
		e = hipEventCreate(&e0);
		if (e == hipSuccess) e = hipEventCreate(&e1);
		if (e == hipSuccess) e = hipEventCreate(&e2);
		if (e == hipSuccess) e = hipEventRecord(e0, st);
	}
	if (e == hipSuccess) e = infer_all(c, d_rays, n, st);
	if (e == hipSuccess && c->time_kernel) e = hipEventRecord(e1, st);
	SeedArgs sa{c->top4.as<uint32_t>(), c->clusters.entry.as<uint32_t>(), c->clusters.box.as<float>(), c->clusters.count};
	prt::TraceStats stats;
	if (e == hipSuccess)
		e = trace_bvh8(c->bvh, c->ts, c->tp, d_rays, n, mask, d_out, st, nullptr, c->tp.count_stats ? &stats : nullptr, &sa);
	if (e == hipSuccess) e = mark_use(c, st);
	if (e == hipSuccess && c->time_kernel) {
		e = hipEventRecord(e2, st);
		if (e == hipSuccess) e = hipEventSynchronize(e2);
		float a = 0.f, b = 0.f;
		if (e == hipSuccess) e = hipEventElapsedTime(&a, e0, e1);
		if (e == hipSuccess) e = hipEventElapsedTime(&b, e1, e2);
		c->mlp_ms = a;
		c->seeded_ms = b;
	}
	if (e0) (void)hipEventDestroy(e0);
	if (e1) (void)hipEventDestroy(e1);
	if (e2) (void)hipEventDestroy(e2);
	if (e != hipSuccess) return hip_fail(c, e, "nearest_hits_learned");
	if (c->tp.count_stats) c->stats = stats;
	return PRT_OK;
}

int prt_mi355x_learned_eval(prt_ctx *c, const float *d_rays, uint64_t n, double *top1, double *top4) {
	if (!c || !top1 || !top4) return PRT_ERR_ARG;
	c->err.clear();
	if (int r = check_learned(c)) return r;
	if (!d_rays || n == 0) return fail(c, PRT_ERR_ARG, "learned_eval: no rays");
	hipError_t e = c->labels.ensure(4 * n);
	if (e == hipSuccess) e = make_labels(c, d_rays, n, c->labels.as<uint32_t>());
	if (e == hipSuccess) e = infer_all(c, d_rays, n, c->stream);
	DevBuf cnt;
	if (e == hipSuccess) e = cnt.ensure(16);
	if (e == hipSuccess) e = hipMemsetAsync(cnt.ptr, 0, 16, c->stream);
	if (e == hipSuccess) {
		hipLaunchKernelGGL(k_seed_quality, dim3(blocks(n)), dim3(256), 0, c->stream, c->labels.as<uint32_t>(),
		                   c->top4.as<uint4>(), n, cnt.as<unsigned long long>());
		e = hipGetLastError();
	}
	unsigned long long h[2] = {0, 0};
	if (e == hipSuccess) e = hipMemcpyAsync(h, cnt.ptr, 16, hipMemcpyDeviceToHost, c->stream);
	if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
	cnt.release();
	if (e != hipSuccess) return hip_fail(c, e, "learned_eval");
	*top1 = static_cast<double>(h[0]) / static_cast<double>(n);
	*top4 = static_cast<double>(h[1]) / static_cast<double>(n);
	c->seed_top1 = *top1;
	c->seed_top4 = *top4;
	return PRT_OK;
}

int prt_mi355x_learned_clusters(prt_ctx *c, uint32_t *count, uint32_t *tri_cluster) {
	if (!c || !count) return PRT_ERR_ARG;
	if (int r = check_learned(c)) return r;
	*count = c->clusters.count;
	if (tri_cluster)
		std::memcpy(tri_cluster, c->clusters.host_tri_cluster.data(), sizeof(uint32_t) * c->clusters.host_tri_cluster.size());
	return PRT_OK;
}

// Parameters of layer l (W^T [out][in] fp32 and bias [out]; the last layer has 4097 valid
// rows of 4224). get != 0 copies them out, get == 0 replaces them (and the bf16 copies).
int prt_mi355x_mlp_params(prt_ctx *c, int layer, int get, float *w, float *b) {
	if (!c || layer < 0 || layer > 3 || !w || !b) return PRT_ERR_ARG;
	if (int r = check_learned(c)) return r;
	static const int in[4] = {kMlpIn, kMlpHidden, kMlpHidden, kMlpHidden};
	const int out = layer == 3 ? kMlpOut : kMlpHidden;
	const size_t nw = static_cast<size_t>(out) * in[layer];
	hipError_t e;
	if (get) {
		e = hipMemcpyAsync(w, c->mlp.w[layer].ptr, nw * 4, hipMemcpyDeviceToHost, c->stream);
		if (e == hipSuccess) e = hipMemcpyAsync(b, c->mlp.b[layer].ptr, out * 4, hipMemcpyDeviceToHost, c->stream);
	} else {
		e = hipMemcpyAsync(c->mlp.w[layer].ptr, w, nw * 4, hipMemcpyHostToDevice, c->stream);
		if (e == hipSuccess) e = hipMemcpyAsync(c->mlp.b[layer].ptr, b, out * 4, hipMemcpyHostToDevice, c->stream);
		if (e == hipSuccess) e = mlp_refresh_bf16(c->mlp, c->stream);
	}
	if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
	return e == hipSuccess ? PRT_OK : hip_fail(c, e, "mlp_params");
}

// One forward + backward on n host rays (n multiple of 32) with host labels, without an
// update: returns the loss, the encoded inputs (bf16 [n][192]) and leaves the gradients for
// prt_mi355x_mlp_grads. For checking the kernels against a CPU implementation.
int prt_mi355x_mlp_grad(prt_ctx *c, const float *rays, uint64_t n, const uint32_t *labels, float *loss, uint16_t *x0) {
	if (!c || !rays || !labels || !loss || n == 0 || n % 32 || n > 65536) return PRT_ERR_ARG;
	if (int r = check_learned(c)) return r;
	DevBuf dr, dl;
	hipError_t e = dr.ensure(n * 24);
	if (e == hipSuccess) e = dl.ensure(n * 4);
	if (e == hipSuccess) e = hipMemcpyAsync(dr.ptr, rays, n * 24, hipMemcpyHostToDevice, c->stream);
	if (e == hipSuccess) e = hipMemcpyAsync(dl.ptr, labels, n * 4, hipMemcpyHostToDevice, c->stream);
	const int rows = static_cast<int>(n);
	if (e == hipSuccess) e = mlp_encode(c->mlp, dr.as<float>(), 0, rows, c->stream);
	if (e == hipSuccess) e = mlp_forward_logits(c->mlp, rows, c->stream);
	if (e == hipSuccess) e = mlp_backward(c->mlp, dl.as<uint32_t>(), rows, c->stream);
	if (e == hipSuccess) e = hipMemcpyAsync(loss, c->mlp.loss.ptr, 4, hipMemcpyDeviceToHost, c->stream);
	if (e == hipSuccess && x0) e = hipMemcpyAsync(x0, c->mlp.x0.ptr, n * kMlpIn * 2, hipMemcpyDeviceToHost, c->stream);
	if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
	dr.release();
	dl.release();
	return e == hipSuccess ? PRT_OK : hip_fail(c, e, "mlp_grad");
}

int prt_mi355x_mlp_grads(prt_ctx *c, int layer, float *gw, float *gb) {
	if (!c || layer < 0 || layer > 3 || !gw || !gb) return PRT_ERR_ARG;
	if (int r = check_learned(c)) return r;
	static const int in[4] = {kMlpIn, kMlpHidden, kMlpHidden, kMlpHidden};
	const int out = layer == 3 ? kMlpOut : kMlpHidden;
	hipError_t e = hipMemcpyAsync(gw, c->mlp.gw[layer].ptr, static_cast<size_t>(out) * in[layer] * 4, hipMemcpyDeviceToHost,
	                              c->stream);
	if (e == hipSuccess) e = hipMemcpyAsync(gb, c->mlp.gb[layer].ptr, out * 4, hipMemcpyDeviceToHost, c->stream);
	if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
	return e == hipSuccess ? PRT_OK : hip_fail(c, e, "mlp_grads");
}

// Inference on n host rays: the top-4 clusters per ray, and optionally the full logits
// (fp32 [n][4097]) from the training forward path.
int prt_mi355x_mlp_top4(prt_ctx *c, const float *rays, uint64_t n, uint32_t *top4, float *logits) {
	if (!c || !rays || !top4 || n == 0 || n > 65536) return PRT_ERR_ARG;
	if (int r = check_learned(c)) return r;
	DevBuf dr, dt;
	hipError_t e = dr.ensure(n * 24);
	if (e == hipSuccess) e = dt.ensure(n * 16);
	if (e == hipSuccess) e = hipMemcpyAsync(dr.ptr, rays, n * 24, hipMemcpyHostToDevice, c->stream);
	const int rows = static_cast<int>(n);
	if (e == hipSuccess) e = mlp_infer_top4(c->mlp, dr.as<float>(), 0, rows, dt.as<uint32_t>(), c->stream);
	if (e == hipSuccess) e = hipMemcpyAsync(top4, dt.ptr, n * 16, hipMemcpyDeviceToHost, c->stream);
	if (e == hipSuccess && logits) {
		e = mlp_encode(c->mlp, dr.as<float>(), 0, rows, c->stream);
		if (e == hipSuccess) e = mlp_forward_logits(c->mlp, rows, c->stream);
		if (e == hipSuccess)
			e = hipMemcpy2DAsync(logits, kMlpOut * 4, c->mlp.logits.ptr, kMlpOutPad * 4, kMlpOut * 4, n,
			                     hipMemcpyDeviceToHost, c->stream);
	}
	if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
	dr.release();
	dt.release();
	return e == hipSuccess ? PRT_OK : hip_fail(c, e, "mlp_top4");
}

}  // extern "C"
