// State and host entry points of the learned t_max seeding network (mlp.hip, learned.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "internal.h"

namespace prt {

constexpr int kMlpIn = 192;       // 6 inputs x 16 frequencies x (sin, cos)
constexpr int kMlpHidden = 256;
constexpr int kClusters = 4096;   // cluster logits; logit kClusters is "miss"
constexpr int kMlpOut = kClusters + 1;
constexpr int kMlpOutPad = 4224;  // 33 tiles of 128 columns

struct Mlp {
	bool ready = false;
	int step = 0;
	float scene_lo[3] = {0, 0, 0}, scene_inv_ext[3] = {1, 1, 1};
	// parameters: fp32 master W^T [out][in] and bias, bf16 W^T and W, Adam moments, gradients
	DevBuf w[4], b[4], wt_bf[4], w_bf[4], mw[4], vw[4], mb[4], vb[4], gw[4], gb[4];
	// activations and work buffers (grown to the largest batch)
	DevBuf x0, h[3], logits, dlog, dlog_t, dh[3], act_t, dact_t, split, loss;
};

hipError_t mlp_init(Mlp &m, uint64_t seed, hipStream_t st);
hipError_t mlp_refresh_bf16(Mlp &m, hipStream_t st);
// Encodes rays [first, first + rows) into m.x0.
hipError_t mlp_encode(Mlp &m, const float *rays, uint64_t first, int rows, hipStream_t st);
// x0 -> logits (fp32, m.logits [rows][kMlpOutPad]).
hipError_t mlp_forward_logits(Mlp &m, int rows, hipStream_t st);
// Softmax cross-entropy against labels, backward into m.gw / m.gb (rows multiple of 32).
hipError_t mlp_backward(Mlp &m, const uint32_t *labels, int rows, hipStream_t st);
hipError_t mlp_adam(Mlp &m, float lr, hipStream_t st);
// Encodes and runs rays [first, first + rows); writes the 4 best cluster logits' indices per
// ray (descending, lower index first at equal logits) to top4 [rows][4].
hipError_t mlp_infer_top4(Mlp &m, const float *rays, uint64_t first, int rows, uint32_t *top4, hipStream_t st);

// Clusters of a BVH8: up to kClusters subtrees that partition the triangles.
struct Clusters {
	uint32_t count = 0;
	DevBuf entry;       // uint32 [count]: stack entry (node index or leaf entry) of each cluster
	DevBuf box;         // float [count][6]: fp32 box (lo xyz, hi xyz)
	DevBuf tri_cluster; // uint32 [n_tris]: cluster of each primitive id
	std::vector<uint32_t> host_tri_cluster;
};

}  // namespace prt
