/* CPU oracle for the nearest-hit path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library,
 * and only as the checker / the timed CPU baseline. The product (libprt_mi355x) never links it.
 *
 * It restates the reference CPU backend (portableRT v0.3.3) in plain C, built with
 * -ffp-contract=off so every float operation rounds on its own as in the reference's default
 * x86-64 build:
 *   orc_build        BVH2 binned-SAH build          include/portableRT/bvh.hpp:13-193
 *   orc_trace        traversal + exact leaf test     bvh.hpp:195-265, core.hpp:27-65
 *                    threading as CPUBackend         intersect_cpu.hpp:12-43
 *   orc_brute        O(N*M) predicate oracle (slab(aabb(T)) and Moller-Trumbore, lowest id)
 *   orc_wt           watertight leaf test (Woop, Benthin, Wald 2013), the optional GPU mode
 *   orc_fp64         double-precision re-intersection of a given (ray, triangle) pair
 *
 * Pinning (what fixes each function independently of this restatement) is listed in
 * tests/test_oracle.py; every function here has at least one such check.
 */
#ifndef PRT_ORACLE_H
#define PRT_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Tie rule at equal t, optionally ORed with ORC_WATERTIGHT (leaf test). */
enum {
	ORC_DFS = 0,       /* reference: first found in depth-first order (strict <, bvh.hpp:247) */
	ORC_LOWEST_ID = 1, /* same traversal, equal t resolved to the lowest primitive id */
	ORC_WATERTIGHT = 0x100, /* leaf test: orc_wt instead of the reference's Moller-Trumbore */
};

/* Full hit record in the reference FullHitReg layout (hitreg.hpp:29-44): 32 bytes. For misses
 * the reference leaves u, v and primitive_id indeterminate (bvh.hpp:232); the oracle writes
 * u = v = 0 and primitive_id = 0xFFFFFFFF. */
typedef struct {
	float u, v, t;
	uint32_t primitive_id;
	uint8_t valid;
	uint8_t pad[3];
	float px, py, pz;
} orc_hit;

typedef struct orc_bvh orc_bvh;

/* Builds the reference BVH2 over n triangles (9 floats each). NULL on allocation failure. */
orc_bvh *orc_build(const float *tris, size_t n);
void orc_free(orc_bvh *b);
size_t orc_node_count(const orc_bvh *b);
/* Node i of the preorder array: is_leaf, tri (0xFFFFFFFF if none), left, right, bounds[6]. */
void orc_node(const orc_bvh *b, size_t i, uint32_t *is_leaf, uint32_t *tri, uint32_t *left,
              uint32_t *right, float *bounds6);

/* Traces n rays (6 floats each) on `threads` threads (0 = hardware concurrency), statically
 * partitioned as the reference does. Returns 0. */
int orc_trace(const orc_bvh *b, const float *rays, size_t n, int tie_mode, int threads,
              orc_hit *out);

/* Brute force over all triangles in id order with strict <, i.e. minimum t, lowest id. */
int orc_brute(const float *tris, size_t n_tris, const float *rays, size_t n, int threads,
              orc_hit *out);

/* The two predicates on one (ray, triangle) pair, exactly as the reference evaluates them. */
int orc_slab(const float *ray6, const float *bmin3, const float *bmax3);
int orc_mt(const float *tri9, const float *ray6, float *t, float *u, float *v);
/* Watertight test (Woop, Benthin and Wald 2013, fp32 with an fma-exact zero fallback); no
 * reference counterpart: the optional leaf test of the GPU backend. */
int orc_wt(const float *tri9, const float *ray6, float *t, float *u, float *v);
/* orc_brute with a leaf-test choice (mode & ORC_WATERTIGHT); minimum t, lowest id. */
int orc_brute_mode(const float *tris, size_t n_tris, const float *rays, size_t n, int mode,
                   int threads, orc_hit *out);

/* Moller-Trumbore in double on the fp32 inputs for pair (ray i, triangle ids[i]); ids equal to
 * 0xFFFFFFFF give NaN. Writes t, u, v in double. */
int orc_fp64(const float *tris, size_t n_tris, const float *rays, const uint32_t *ids, size_t n,
             double *t, double *u, double *v);

#ifdef __cplusplus
}
#endif
#endif
