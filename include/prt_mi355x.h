/* C ABI of the MI355X (gfx950) nearest-hit backend for portableRT.
 *
 * Plain pointers and sizes only. Every entry point returns a prt_status; on failure the
 * context's last error holds a message (prt_mi355x_last_error). Nothing aborts the process:
 * with no GPU, probe reports zero devices. A context is not thread-safe (as the reference
 * backends, include/portableRT/backend.hpp:39-46); calls on one context block unless named
 * *_device.
 *
 * Triangle layout: 9 floats v0 v1 v2 per triangle (reference Tri, core.hpp:24-25).
 * Ray layout:      6 floats origin, direction (reference Ray, core.hpp:19-22), 24 bytes.
 * Hit layout:      the reference HitReg<Tags...> for a field mask (hitreg.hpp:29-61).
 */
#ifndef PRT_MI355X_H
#define PRT_MI355X_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct prt_ctx prt_ctx;

typedef enum {
	PRT_OK = 0,
	PRT_ERR_ARG = 1,      /* bad argument (null pointer, size, option) */
	PRT_ERR_NODEV = 2,    /* no gfx950 device / device index out of range */
	PRT_ERR_HIP = 3,      /* HIP runtime error (message in last_error) */
	PRT_ERR_OOM = 4,      /* device allocation failed or over the configured limit */
	PRT_ERR_STATE = 5,    /* nearest_hits before set_tris */
	PRT_ERR_LAYOUT = 6,   /* stride/offsets do not match HitReg for the field mask */
	PRT_ERR_INPUT = 7     /* non-finite triangle vertex */
} prt_status;

/* Field mask bits: the reference filter tags (hitreg.hpp:48-54). */
enum { PRT_UV = 1, PRT_T = 2, PRT_PRIMITIVE_ID = 4, PRT_P = 8, PRT_VALID = 16 };

/* Tunables (prt_mi355x_set_option). */
enum {
	PRT_OPT_NODE_FORMAT = 1,  /* 0 fp32 BVH8 nodes (208 B), 1 8-bit quantized (80 B, default) */
	PRT_OPT_CULL = 2,         /* 0 no distance culling; 1 culling with margin (default), boxes holding
	                             an ill-conditioned triangle never culled */
	PRT_OPT_LEAF_SIZE = 3,    /* max triangles per leaf, 1..4 (default 3); takes effect at set_tris */
	PRT_OPT_BLOCK = 4,        /* traversal workgroup size: 64, 128 (default), 256 */
	PRT_OPT_WAVES_PER_CU = 5, /* persistent waves per CU, 1..64; 0 (default): from the ray count */
	PRT_OPT_LDS_STACK = 6,    /* LDS stack entries per lane: 8 (default) or 16 */
	PRT_OPT_CHUNK = 7,        /* rays a wave reserves per atomic, 64..65536 (default 128) */
	PRT_OPT_LEAF_PHASE = 8,   /* triangle phase when parked leaves >= live lanes * v / 64, 1..64 (default 20) */
	PRT_OPT_COUNT_STATS = 9,  /* 1: count node visits / triangle tests (slower) */
	PRT_OPT_TIME_KERNEL = 10, /* 1: time the traversal kernel with HIP events (synchronises) */
	PRT_OPT_ALLOC_LIMIT = 11, /* refuse device allocations above this many bytes (0 = none) */
	PRT_OPT_MARGIN_REL_LOG2 = 12, /* culling margin relative part 2^-k, k in 4..10 (default 10) */
	PRT_OPT_BUILDER = 13,     /* binary tree under the BVH8: 0 radix tree over the Morton order, 1
	                             clustering of Morton-order neighbours (default); at set_tris */
	PRT_OPT_REFILL = 14,      /* a wave loads new rays once at least v of its lanes are idle, 1..64 (default 8) */
	PRT_OPT_COLLAPSE = 15,    /* binary tree to BVH8: 0 greedy (open the largest child until full), 1
	                             least surface-area-heuristic cost (default); at set_tris */
	PRT_OPT_LEAF_TEST = 16,   /* 0 Moller-Trumbore, the reference's test (default); 1 watertight
	                             (Woop, Benthin, Wald 2013; plain traversal only, no distance
	                             culling) */
	PRT_OPT_SEEDS = 17        /* learned traversal: the best 1..4 predicted clusters searched
	                             first (default 1, the fastest measured) */
};

/* Statistics (prt_mi355x_get_stat). */
enum {
	PRT_STAT_BUILD_MS = 1,    /* last set_tris, device time */
	PRT_STAT_TRACE_MS = 2,    /* last traversal kernel (needs PRT_OPT_TIME_KERNEL) */
	PRT_STAT_NODES = 3,       /* BVH8 nodes */
	PRT_STAT_DEPTH = 4,       /* BVH8 levels */
	PRT_STAT_NODE_VISITS = 5, /* last trace, with PRT_OPT_COUNT_STATS */
	PRT_STAT_TRI_TESTS = 6,
	PRT_STAT_LEAF_VISITS = 7,
	PRT_STAT_NODE_BYTES = 8,  /* bytes per node of the selected format */
	PRT_STAT_API_MS = 9,      /* last host-buffer nearest_hits, wall time incl. copies */
	PRT_STAT_N_TRIS = 10,
	PRT_STAT_NODE_PHASES = 11, /* wave iterations spent in node phases (with COUNT_STATS) */
	PRT_STAT_TRI_PHASES = 12,  /* wave iterations spent in triangle phases */
	PRT_STAT_CULLED_POPS = 13, /* stack entries discarded by distance culling */
	PRT_STAT_MLP_MS = 14,      /* last learned call: encoding + MLP + top-4 (PRT_OPT_TIME_KERNEL) */
	PRT_STAT_SEEDED_MS = 15,   /* last learned call: seeded traversal kernel */
	PRT_STAT_TRAIN_MS = 16,    /* last learned_train, wall time */
	PRT_STAT_TRAIN_LOSS = 17,  /* cross-entropy of the last training batch */
	PRT_STAT_CLUSTERS = 18,    /* BVH8 subtree clusters */
	PRT_STAT_SEED_TOP1 = 19,   /* last learned_eval: fraction of rays whose top-1 cluster is true */
	PRT_STAT_SEED_TOP4 = 20,   /* ... whose true cluster is among the top 4 */
	PRT_STAT_REFIT_MS = 21,    /* last refit, device time */
	PRT_STAT_NCULL_TRIS = 22   /* triangles whose boxes are never culled by distance (ill-conditioned
	                              for the leaf test; counted by set_tris and refit) */
};

/* Number of usable gfx950 devices and the name of device 0 (name may be NULL). Returns PRT_OK
 * with *ndev = 0 when there is no GPU. Replaces the device probe in the reference HIP
 * backend's is_available (src/intersect_hip.cpp:353-372). */
int prt_mi355x_probe(int *ndev, char *name, size_t cap);

/* Context on device `device`: one HIP stream, the BVH and reusable staging buffers. */
int prt_mi355x_create(int device, prt_ctx **out);
void prt_mi355x_destroy(prt_ctx *ctx);

/* Message of the last failed call on ctx ("" if none). Valid until the next call. */
const char *prt_mi355x_last_error(const prt_ctx *ctx);

/* Device name ("AMD Instinct MI355X (gfx950)"). Backend::device_name (backend.hpp:23). */
int prt_mi355x_device_name(const prt_ctx *ctx, char *buf, size_t cap);

/* Builds the BVH over n triangles in host memory (copied; the previous BVH is replaced).
 * Backend::set_tris (backend.hpp:19, intersect_cpu.cpp:12). */
int prt_mi355x_set_tris(prt_ctx *ctx, const float *tris9, uint64_t n);

/* Same with the triangles already in device memory (read during the call only). The copy runs
 * on the context's own stream, which does not wait for other streams: the caller must have
 * completed its writes to d_tris9 (synchronise the stream that wrote them) before the call. */
int prt_mi355x_set_tris_device(prt_ctx *ctx, const float *d_tris9, uint64_t n);

/* Refit for dynamic geometry: the same n triangles (same order and ids) at new positions. Keeps
 * the tree's topology and recomputes every leaf record and node box (fp32 and quantized), so a
 * frame of moving geometry costs a few launches instead of a build; records equal those of
 * set_tris on the new positions (tree quality degrades as triangles move far). Errors:
 * PRT_ERR_STATE before set_tris, PRT_ERR_ARG when n differs from the tree's count,
 * PRT_ERR_INPUT for a non-finite vertex (the scene is then cleared). Learned seeding must be
 * set up again afterwards. No counterpart in the reference, which rebuilds on every set_tris
 * and notes a move version as a to-do (backend.hpp:18-19, intersect_cpu.cpp:12). */
int prt_mi355x_refit(prt_ctx *ctx, const float *tris9, uint64_t n);

/* Same with the triangles in device memory (their writes completed before the call, as for
 * prt_mi355x_set_tris_device). */
int prt_mi355x_refit_device(prt_ctx *ctx, const float *d_tris9, uint64_t n);

/* Nearest hit of n rays in host memory, records written to host memory `out` with the
 * HitReg layout of `mask`: `stride` = sizeof(HitReg), off[8] = offsets of u v t primitive_id
 * valid px py pz (-1 for absent fields). Blocking. The member template
 * nearest_hits<Tags...>(const std::vector<Ray>&) (intersect_cpu.hpp:20-43,
 * nearest_hits_impl.hpp:28-36). Miss records: t = +inf, valid = 0, u = v = 0,
 * primitive_id = 0xFFFFFFFF, p = o + inf * d. */
int prt_mi355x_nearest_hits(prt_ctx *ctx, const float *rays6, uint64_t n, uint32_t mask, void *out,
                            uint32_t stride, const int32_t off[8]);

/* Same with rays and records in device memory; enqueued on `stream` (hipStream_t, NULL = the
 * context's stream) without synchronising unless PRT_OPT_TIME_KERNEL is set. Calls of one
 * context run in order: a call on another stream than the previous call's waits for it on the
 * device (they share the context's scratch); use one context per stream to run traversals
 * concurrently. No counterpart in the reference, which copies on every call
 * (src/intersect_hip.cpp:322-339). */
int prt_mi355x_nearest_hits_device(prt_ctx *ctx, const float *d_rays6, uint64_t n, uint32_t mask,
                                   void *d_out, uint32_t stride, const int32_t off[8], void *stream);

/* prt_mi355x_nearest_hits_device for callers that cannot allocate device memory (code under
 * hipGraph stream capture, allocation-free serving loops): no counterpart in the reference. Same
 * arguments and records; the traversal's work counter and stack spill live in a fixed 128 MiB
 * device array that each GPU holds from the library's load to the end of the process (its
 * address is looked up once per device, at prt_mi355x_create), so the call makes no device (or
 * host) allocation. It launches as many persistent workgroups as that array holds (fewer than
 * prt_mi355x_nearest_hits_device for deep trees) and fails with PRT_ERR_ARG when not even one
 * fits (a BVH deeper than about 18,000 levels at the default workgroup of 128 threads). All
 * calls on one GPU share the array and each launch clears its counter with a kernel: calls that
 * overlap on one GPU (other streams, other contexts, other threads) race. No timing or statistics options apply; error flags are read
 * by prt_mi355x_synchronize. */
int prt_mi355x_nearest_hits_device_noalloc(prt_ctx *ctx, const float *d_rays6, uint64_t n, uint32_t mask,
                                           void *d_out, uint32_t stride, const int32_t off[8], void *stream);

/* Waits for all work on the context's stream. */
int prt_mi355x_synchronize(prt_ctx *ctx);

int prt_mi355x_set_option(prt_ctx *ctx, int option, int64_t value);
int prt_mi355x_get_option(const prt_ctx *ctx, int option, int64_t *value);
int prt_mi355x_get_stat(const prt_ctx *ctx, int stat, double *value);

/* HitReg layout for a field mask: out[0] = size, out[1..8] = offsets (u v t id valid px py pz),
 * -1 for absent fields. Needs no device. */
int prt_mi355x_layout(uint32_t mask, int32_t out[9]);

/* Copies the device BVH to host memory for inspection: which = 0 quantized nodes, 1 fp32
 * nodes, 2 leaf triangle records. *bytes is the size; with out == NULL only the size is
 * returned. */
int prt_mi355x_export(prt_ctx *ctx, int which, void *out, uint64_t *bytes);

/* ---- Learned t_max seeding. No counterpart in the reference: a BF16 MLP (192 -> 256 -> 256 ->
 * 256 -> 4097, MFMA kernels) predicts the BVH8 subtree cluster holding each ray's hit; the 4
 * best clusters are intersected exactly before the full traversal, so results are identical to
 * prt_mi355x_nearest_hits_device. */

/* Cuts the current BVH into at most 4096 subtree clusters and initialises the network
 * (deterministic from `seed`). Call after every set_tris. */
int prt_mi355x_learned_setup(prt_ctx *ctx, uint64_t seed);

/* Traces n device rays with the plain traversal, labels each with its hit cluster (or miss)
 * and trains the network for `epochs` passes in batches of `batch` rays (multiple of 32), Adam
 * with learning rate lr. */
int prt_mi355x_learned_train(prt_ctx *ctx, const float *d_rays6, uint64_t n, int epochs, int batch, float lr);

/* prt_mi355x_nearest_hits_device with learned seeding (same records, same arguments). */
int prt_mi355x_nearest_hits_learned_device(prt_ctx *ctx, const float *d_rays6, uint64_t n, uint32_t mask,
                                           void *d_out, uint32_t stride, const int32_t off[8], void *stream);

/* Fraction of n device rays whose true cluster is the top-1 / among the top-4 predictions. */
int prt_mi355x_learned_eval(prt_ctx *ctx, const float *d_rays6, uint64_t n, double *top1, double *top4);

/* Number of clusters and (tri_cluster != NULL) the cluster of every primitive id. */
int prt_mi355x_learned_clusters(prt_ctx *ctx, uint32_t *count, uint32_t *tri_cluster);

/* Network access for checking the kernels on the host: parameters of layer 0..3 (W^T [out][in]
 * and bias, fp32; out = 4097 for the last layer), gradients of one batch, inference. */
int prt_mi355x_mlp_params(prt_ctx *ctx, int layer, int get, float *w, float *b);
int prt_mi355x_mlp_grad(prt_ctx *ctx, const float *rays6, uint64_t n, const uint32_t *labels, float *loss,
                        uint16_t *x0_bf16);
int prt_mi355x_mlp_grads(prt_ctx *ctx, int layer, float *gw, float *gb);
int prt_mi355x_mlp_top4(prt_ctx *ctx, const float *rays6, uint64_t n, uint32_t *top4, float *logits);

/* ---- Batched multi-scene tracing. No counterpart in the reference, whose backends hold one
 * scene per set_tris and trace one ray vector per nearest_hits (backend.hpp:19-46): here many
 * scenes of uneven sizes are built by one set of launches and traced by ONE persistent launch,
 * capturable in a hipGraph. Triangles of all scenes are concatenated (scene s owns
 * tri_counts[s] consecutive triangles), and so are rays; each record equals what set_tris(scene
 * s) + nearest_hits(rays of s) returns (primitive ids are scene-local). Scenes are grouped in 4
 * size buckets by triangle count (< 4096, < 16384, < 65536, more), each with its own BVH
 * branching factor, leaf size and work-tile size. */

/* Declares the batch: scene count (1..65536), per-scene triangle and ray counts (at least one
 * triangle in total), record layout (field mask). Replaces any previous batch; tuned parameters
 * are kept. */
int prt_mi355x_batch_setup(prt_ctx *ctx, uint32_t n_scenes, const uint64_t *tri_counts, const uint64_t *ray_counts,
                           uint32_t mask);

/* Builds every scene and traces every ray, enqueued on `stream` (NULL = the context's stream)
 * without synchronising. Triangles (9 floats), rays (6 floats) and records in device memory. */
int prt_mi355x_batch_run(prt_ctx *ctx, const float *d_tris9, const float *d_rays6, void *d_out, void *stream);

/* Captures the same work with these buffers into a hipGraph (after one eager sizing run when
 * needed); batch_replay launches it. The graph reads the buffers at replay time. */
int prt_mi355x_batch_capture(prt_ctx *ctx, const float *d_tris9, const float *d_rays6, void *d_out);
int prt_mi355x_batch_replay(prt_ctx *ctx, void *stream);

/* Waits for the device and reports errors of the last batch run or replay (non-finite vertex:
 * PRT_ERR_INPUT; BVH deeper than the fixed 48 levels or corrupt: PRT_ERR_HIP). */
int prt_mi355x_batch_check(prt_ctx *ctx);

/* Parameters: bucket 0..3 -> {width 2..8, leaf 1..4, tile rays 1..2^20}; bucket -1 -> launch
 * shape {workgroup 64/128/256, waves per CU 1..64, LDS stack 8/16}. Setting drops the graph. */
int prt_mi355x_batch_get_params(const prt_ctx *ctx, int bucket, int32_t v[3]);
int prt_mi355x_batch_set_params(prt_ctx *ctx, int bucket, const int32_t v[3]);

/* Autotuner: times candidate parameters on the batch (rebuilding and tracing it, `reps` runs
 * each, smallest taken) and keeps the fastest; returns whole-batch build + trace time before and
 * after. Leaves the records of the final run in d_out. Synchronous. */
int prt_mi355x_batch_tune(prt_ctx *ctx, const float *d_tris9, const float *d_rays6, void *d_out, int reps,
                          double *before_ms, double *after_ms);

/* The work tiles {first ray lo, hi, count, scene} the batch traversal claims, for given tile
 * sizes per bucket, largest scenes first. Needs no device: *n_tiles receives the count; with
 * tiles4 != NULL, *n_tiles must hold its capacity in tiles on entry. */
int prt_mi355x_batch_tiles(uint32_t n_scenes, const uint64_t *tri_counts, const uint64_t *ray_counts,
                           const int32_t chunk[4], uint32_t *tiles4, uint64_t *n_tiles);

#ifdef __cplusplus
}
#endif
#endif
