// Host-side interfaces shared by the translation units of libprt_mi355x.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "common.h"

namespace prt {

// Device buffer owned by a context, grown on demand and reused across calls.
struct DevBuf {
	void *ptr = nullptr;
	size_t bytes = 0;
	DevBuf() = default;
	DevBuf(const DevBuf &) = delete;
	DevBuf &operator=(const DevBuf &) = delete;
	~DevBuf() { release(); }
	hipError_t ensure(size_t want) {
		if (want <= bytes) return hipSuccess;
		if (ptr) {
			hipError_t e = hipFree(ptr);
			ptr = nullptr;
			bytes = 0;
			if (e != hipSuccess) return e;
		}
		if (want == 0) return hipSuccess;
		hipError_t e = hipMalloc(&ptr, want);
		if (e == hipSuccess) bytes = want;
		return e;
	}
	void release() {
		if (ptr) (void)hipFree(ptr);
		ptr = nullptr;
		bytes = 0;
	}
	template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// A BVH8 over one triangle set, resident in HBM.
constexpr uint32_t kNcullUnknown = 0xFFFFFFFFu;

struct Bvh8 {
	uint64_t n_tris = 0;
	uint32_t n_nodes = 0;    // BVH8 nodes (both formats share topology)
	uint32_t depth = 0;      // BVH8 levels
	uint32_t leaf_size = 0;  // maximum triangles per leaf child
	float scene_box[6] = {0, 0, 0, 0, 0, 0};  // min xyz, max xyz of all vertices
	uint32_t n_ncull = 0;    // triangles flagged by cull_unsafe (kNcullUnknown: not counted)
	DevBuf nodes_q;          // NodeQ[n_nodes]
	DevBuf nodes_f;          // NodeF[n_nodes]
	DevBuf tris;             // TriRec[n_tris], in leaf order
	std::vector<uint32_t> levels;  // first node of each BVH8 level (found at the first refit)
};

// Scratch of the LBVH build, kept across set_tris calls.
struct BuildScratch {
	DevBuf tri_in;      // input triangles (9 floats each)
	DevBuf tbox;        // float4[2*n] box per input triangle
	DevBuf keys[2];     // uint64 Morton keys (ping-pong)
	DevBuf vals[2];     // uint32 triangle indices (ping-pong)
	DevBuf hist;        // radix histograms
	DevBuf bin;         // binary tree: left, right, parent, first, count, flags
	DevBuf bbox;        // float4[2*(2n-1)] binary node boxes
	DevBuf bin_of;      // BVH8 node -> binary node
	DevBuf tri_dest;    // sorted position -> TriRec slot
	DevBuf misc;        // counters and per-scene bounds
	DevBuf scene_root;  // binary root node of every scene
	DevBuf ploc;        // clustering builder: cluster lists, neighbours, flags
	DevBuf refit_box;   // refit: float4[2 * n_nodes] node boxes
	DevBuf cost;        // cost-optimal collapse: per-scene area scales, 8 floats per internal node
	void *host_misc = nullptr;  // pinned mirror of the counters and scene 0's bounds
};

struct BuildParams {
	int leaf_size = 3;  // 1..kMaxLeaf
	int width = 8;      // children per BVH8 node, 2..8 (the node format has 8 slots)
	int builder = 1;    // binary tree: 0 radix tree over the Morton order (Karras), 1 clustering
	                    // (PLOC) over the same order; graph-capturable builds use the radix tree
	int collapse = 1;   // binary tree to BVH8: 0 greedy largest-area opening, 1 least SAH cost
};

constexpr uint32_t kMaxScenes = 65536;

// Fixed-length (graph-capturable) builds: kFixedTopLevels BVH8 level launches, then one launch
// that collapses the rest depth first with a stack of kFixedDfsDepth levels.
constexpr int kFixedTopLevels = 6;
constexpr int kFixedDfsDepth = 64;

// Several scenes in one build (device arrays). BVH8 node s is the root of scene s (an empty node
// for an empty scene) and leaf triangle records carry scene-local ids.
struct SceneSet {
	uint32_t count = 1;
	const uint32_t *tri_off = nullptr;  // count + 1 prefix offsets into the triangle array
	const uint32_t *param = nullptr;    // per scene leaf_size | width << 8 (null: BuildParams)
	float *box = nullptr;               // out: per scene vertex bounds, min xyz max xyz, or null
	const float *tris = nullptr;        // triangles (null: the scratch's tri_in)
	int levels = 0;                     // > 0: a fixed-length build of at most this many BVH8
	                                    // levels (min(levels, kFixedTopLevels) level launches,
	                                    // then the depth-first rest) and no host synchronisation,
	                                    // so it can be captured in a graph once its buffers
	                                    // exist; completion is checked on the device
};

// Builds `out` from n triangles (device, 9 floats each; scratch.tri_in unless the scene set
// names them). Returns hipSuccess, or an error; *bad_input is set when a vertex is not finite.
// With scenes->levels > 0, out.n_nodes is the node capacity, out.depth the level count, and
// errors are read afterwards with build_errors.
hipError_t build_bvh8(BuildScratch &s, uint64_t n, const BuildParams &p, Bvh8 &out, hipStream_t st,
                      bool *bad_input, double *build_ms, const SceneSet *scenes = nullptr);

// Error bits of the last build on `s`: 1 non-finite vertex, 2 triangle not placed, 4 the BVH8
// deeper than a fixed-length build allows. Synchronises st.
hipError_t build_errors(BuildScratch &s, hipStream_t st, uint32_t *flags);

// New vertex positions (s.tri_in, the same n triangles in input order) for the tree in `b`:
// leaf records and every node box are recomputed bottom-up, the topology is kept. Sets
// *bad_input for a non-finite vertex (the tree is then not usable). Synchronises st.
hipError_t refit_bvh8(BuildScratch &s, Bvh8 &b, hipStream_t st, bool *bad_input, double *refit_ms);

// Traversal tunables.
struct TraceParams {
	int node_format = 1;      // 0 fp32 nodes, 1 8-bit quantized nodes
	int cull = 1;             // 0 exact (no distance culling), 1 distance culling with margin
	int block = 128;          // threads per workgroup
	int waves_per_cu = 0;     // persistent waves per CU; 0: from the ray count (16..32)
	int lds_stack = 8;        // LDS stack entries per thread (power of two)
	int chunk = 128;          // rays a wave reserves per atomic (refilled lane by lane)
	int leaf_phase = 20;      // triangle phase when parked leaves >= live lanes * leaf_phase / 64
	int refill = 8;           // load new rays once at least this many lanes of a wave are idle
	int count_stats = 0;      // count node visits and triangle tests
	int leaf_test = 0;        // 0 Moller-Trumbore (reference), 1 watertight (no distance culling)
	int n_seeds = 1;          // learned traversal: predicted clusters searched first (1..4)
	float margin_rel = 0x1p-10f;
	float margin_abs = 0.0f;  // absolute margin in scene units (set from the scene size)
};

struct TraceStats {
	unsigned long long node_visits = 0, tri_tests = 0, leaf_visits = 0;
	uint32_t errors = 0;  // bit 0 corrupt leaf entry, bit 1 corrupt node / stack bound
	unsigned long long node_phases = 0, tri_phases = 0, culled_pops = 0;  // wave iterations
};

struct TraceScratch {
	DevBuf counter;  // work counter + stats
	DevBuf spill;    // per-thread stack spill
	int n_cus = 0;
	uint32_t *host_err = nullptr;  // pinned copy of the last launch's error flags
};

// Reads the error flags of the last traversal launched with `s` (synchronises `st`).
hipError_t trace_errors(TraceScratch &s, hipStream_t st, uint32_t *flags);

// Traces n device rays against `b`, writing HitReg records at `out` (device) with the layout
// of `mask`. Launches on `st`; does not synchronise. `ms` receives the duration of the
// traversal kernel measured with HIP events when non-null (synchronises then).
// Per-ray seed clusters for the learned traversal: 4 cluster ids per ray (>= count: none),
// and each cluster's stack entry and fp32 box.
struct SeedArgs {
	const uint32_t *seeds;
	const uint32_t *entry;
	const float *box;
	uint32_t count;
};

// Several scenes in one launch over a BVH built from a SceneSet: rays are claimed in tiles
// {first lo, first hi, count, scene} (each inside one scene), and info holds per scene
// {absolute culling margin, 1 if every vertex is below 2^60 in magnitude else 0}. n is the total
// ray count and records go to out + ray index * stride.
struct BatchArgs {
	const uint4 *tiles;
	uint32_t n_tiles;
	const float2 *info;
	uint32_t n_scenes;
};

hipError_t trace_bvh8(const Bvh8 &b, TraceScratch &s, const TraceParams &p, const float *rays, uint64_t n,
                      uint32_t mask, void *out, hipStream_t st, float *ms, TraceStats *stats,
                      const SeedArgs *seeds = nullptr, const BatchArgs *batch = nullptr);

// Traversal without allocation: the counter and stack spill live in a fixed device array of
// kNoAllocScratchBytes (trace.hip). noalloc_scratch_setup(device) looks up that array's address
// on `device` (which must be current) once; noalloc_scratch(device) returns it (null before the
// set-up). trace_bvh8_noalloc launches as many workgroups as the array's spill part holds and
// returns hipErrorInvalidValue, with *need_bytes the size one workgroup needs, when not even one
// fits (a tree deeper than the array allows). No timing and no statistics.
constexpr uint64_t kNoAllocScratchBytes = uint64_t{128} << 20;
constexpr uint64_t kNoAllocCounterBytes = 256;
hipError_t noalloc_scratch_setup(int device);
void *noalloc_scratch(int device);
hipError_t trace_bvh8_noalloc(const Bvh8 &b, TraceScratch &s, const TraceParams &p, const float *rays, uint64_t n,
                              uint32_t mask, void *out, hipStream_t st, void *scratch_ptr, uint64_t *need_bytes);
// Error flags of the last launch that used `counter` (synchronises st).
hipError_t trace_errors_at(TraceScratch &s, const uint32_t *counter, hipStream_t st, uint32_t *flags);

}  // namespace prt
