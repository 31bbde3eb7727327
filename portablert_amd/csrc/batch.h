// Batched multi-scene tracing: many scenes of uneven sizes built and traced by one set of
// launches (one persistent traversal launch), capturable in a hipGraph, with per-size-bucket
// parameters chosen by the autotuner in batch.hip. No counterpart in the reference, whose
// backends hold one scene and trace one ray vector per call (include/portableRT/backend.hpp:19-46).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "internal.h"

namespace prt {

constexpr int kBatchBuckets = 4;   // scenes by triangle count: < 4096, < 16384, < 65536, more
constexpr int kBatchLevels = kFixedTopLevels + kFixedDfsDepth;  // deepest BVH8 of the capturable build

inline int batch_bucket(uint64_t tris) { return tris < 4096 ? 0 : tris < 16384 ? 1 : tris < 65536 ? 2 : 3; }

// Per-bucket parameters: BVH branching factor, leaf size, rays per work tile.
struct BucketParams {
	int width = 8;
	int leaf = 2;
	int chunk = 64;
};

struct Batch {
	uint32_t n_scenes = 0;
	std::vector<uint64_t> tri_off, ray_off;  // host prefix offsets, n_scenes + 1 each
	uint32_t mask = 0;
	BucketParams bucket[kBatchBuckets];
	int block = 128, waves_per_cu = 32, lds_stack = 8;  // launch shape of the one traversal launch
	DevBuf d_tri_off;  // uint32 prefix offsets
	DevBuf d_param;    // uint32 per scene: leaf | width << 8
	DevBuf d_box;      // float[6] per scene, written by the build
	DevBuf d_info;     // float2 per scene: culling margin, fast-path flag
	DevBuf d_tiles;    // uint4 work tiles, largest scenes first
	DevBuf d_tiles_sub;  // tiles of one bucket (autotuner)
	uint32_t n_tiles = 0;
	Bvh8 bvh;
	BuildScratch bs;
	TraceScratch ts;
	bool ran = false;  // an eager run has sized every buffer for the current parameters
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	Batch() = default;
	Batch(const Batch &) = delete;
	Batch &operator=(const Batch &) = delete;
	void drop_graph() {
		if (exec) (void)hipGraphExecDestroy(exec);
		if (graph) (void)hipGraphDestroy(graph);
		exec = nullptr;
		graph = nullptr;
	}
	~Batch() {
		drop_graph();
		if (bs.host_misc) (void)hipHostFree(bs.host_misc);
		if (ts.host_err) (void)hipHostFree(ts.host_err);
	}
};

// Work tiles {first lo, first hi, count, scene}: every ray of every scene exactly once, tiles of
// chunk(bucket of the scene) rays, scenes in decreasing triangle count (then index).
std::vector<uint32_t> batch_plan(uint32_t n_scenes, const uint64_t *tri_off, const uint64_t *ray_off,
                                 const int chunk[kBatchBuckets], int only_bucket = -1);

}  // namespace prt
