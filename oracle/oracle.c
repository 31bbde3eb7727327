/* CPU oracle for the nearest-hit path -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restatement of the reference CPU backend. Each function cites the reference lines it
 * follows. Compiled with -ffp-contract=off: no fused multiply-add, each operation rounded to
 * float as in the reference's default x86-64 build (no -march, CMakeLists.txt:17-19).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* std::min / std::max as libstdc++ defines them: the first argument wins when the comparison
 * is false, so NaN propagates asymmetrically (bvh.hpp:30-46, :210-211). */
static inline float smin(float a, float b) { return (b < a) ? b : a; }
static inline float smax(float a, float b) { return (a < b) ? b : a; }

typedef struct {
	float lo[3], hi[3];
} aabb_t;

typedef struct {
	uint32_t is_leaf, tri, left, right;
	aabb_t box;
} node_t;

struct orc_bvh {
	node_t *nodes;
	size_t n_nodes, cap;
	float *tris;
	size_t n_tris;
};

/* BVH2::empty_aabb (bvh.hpp:13-18) */
static aabb_t empty_aabb(void) {
	aabb_t b;
	for (int k = 0; k < 3; ++k) {
		b.lo[k] = FLT_MAX;
		b.hi[k] = -FLT_MAX;
	}
	return b;
}

/* BVH2::make_aabb (bvh.hpp:28-37) */
static aabb_t make_aabb(const float *t) {
	aabb_t b;
	for (int k = 0; k < 3; ++k) {
		b.lo[k] = smin(t[k], smin(t[3 + k], t[6 + k]));
		b.hi[k] = smax(t[k], smax(t[3 + k], t[6 + k]));
	}
	return b;
}

/* BVH2::extend_aabb (bvh.hpp:39-48) */
static aabb_t extend_aabb(aabb_t a, aabb_t b) {
	aabb_t r;
	for (int k = 0; k < 3; ++k) {
		r.lo[k] = smin(a.lo[k], b.lo[k]);
		r.hi[k] = smax(a.hi[k], b.hi[k]);
	}
	return r;
}

/* BVH2::aabb_harea (bvh.hpp:51-56) */
static float harea(aabb_t b) {
	const float dx = b.hi[0] - b.lo[0];
	const float dy = b.hi[1] - b.lo[1];
	const float dz = b.hi[2] - b.lo[2];
	return (dx * dy + dx * dz + dy * dz);
}

/* ------------------------------------------------------------------------------ build */

typedef struct {
	const float *tris;
	aabb_t *tbox;  /* make_aabb of every triangle, computed once (same values as recomputing) */
	float *cent;   /* 3 centroids per triangle, (a+b+c)/3.0f as bvh.hpp:83 */
	node_t *nodes;
	size_t n_nodes, cap;
	int oom;
} builder_t;

static uint32_t push_node(builder_t *B) {
	if (B->n_nodes == B->cap) {
		size_t nc = B->cap ? 2 * B->cap : 1024;
		node_t *nn = (node_t *)realloc(B->nodes, nc * sizeof(node_t));
		if (!nn) {
			B->oom = 1;
			return 0;
		}
		B->nodes = nn;
		B->cap = nc;
	}
	node_t *n = &B->nodes[B->n_nodes];
	n->is_leaf = 0;
	n->tri = 0xFFFFFFFFu;
	n->left = 0xFFFFFFFFu;
	n->right = 0xFFFFFFFFu;
	n->box = empty_aabb();
	return (uint32_t)B->n_nodes++;
}

/* BVH2::divide_sah (bvh.hpp:58-129). `idx` is the node's triangle list in input order; the
 * left and right lists keep that order. The scan for every (bin, axis) pair folds the boxes in
 * input order, as the reference does, so NaN-carrying inputs fold identically. */
static void divide_sah(builder_t *B, const uint32_t *idx, size_t n, uint32_t *left, size_t *nl,
                       uint32_t *right, size_t *nr) {
	const size_t bin_count = 14;
	float min_cost = FLT_MAX;
	int best_axis = 0;
	int best_bin = (int)(bin_count / 2);

	aabb_t node_box = empty_aabb();
	for (size_t i = 0; i < n; ++i) node_box = extend_aabb(node_box, B->tbox[idx[i]]);

	for (int bin = 0; bin < (int)bin_count; ++bin) {
		for (int axis = 0; axis < 3; ++axis) {
			/* (max - min) * float(bin) / bin_count, then added to min (bvh.hpp:73-75) */
			const float bin_pos = node_box.lo[axis] + (node_box.hi[axis] - node_box.lo[axis]) *
			                                              (float)bin / (float)bin_count;
			aabb_t lb = empty_aabb(), rb = empty_aabb();
			int lc = 0, rc = 0;
			for (size_t i = 0; i < n; ++i) {
				const uint32_t t = idx[i];
				if (B->cent[3 * t + axis] < bin_pos) {
					lb = extend_aabb(lb, B->tbox[t]);
					++lc;
				} else {
					rb = extend_aabb(rb, B->tbox[t]);
					++rc;
				}
			}
			const float cost = harea(lb) * (float)lc + harea(rb) * (float)rc;
			if (cost < min_cost) {
				min_cost = cost;
				best_axis = axis;
				best_bin = bin;
			}
		}
	}

	const float bin_pos = node_box.lo[best_axis] +
	                      (node_box.hi[best_axis] - node_box.lo[best_axis]) * (float)best_bin /
	                          (float)bin_count;
	*nl = *nr = 0;
	for (size_t i = 0; i < n; ++i) {
		if (B->cent[3 * idx[i] + best_axis] < bin_pos)
			left[(*nl)++] = idx[i];
		else
			right[(*nr)++] = idx[i];
	}
	if (*nl == 0 || *nr == 0) { /* median split by input order (bvh.hpp:118-128) */
		*nl = *nr = 0;
		for (size_t i = 0; i < n; ++i) {
			if (i < n / 2)
				left[(*nl)++] = idx[i];
			else
				right[(*nr)++] = idx[i];
		}
	}
}

/* BVH2::build_aux (bvh.hpp:131-156): preorder node numbering, one triangle per leaf. */
static uint32_t build_aux(builder_t *B, const uint32_t *idx, size_t n) {
	const uint32_t me = push_node(B);
	if (B->oom) return 0;
	if (n == 0) {
		B->nodes[me].is_leaf = 1;
	} else if (n == 1) {
		B->nodes[me].is_leaf = 1;
		B->nodes[me].tri = idx[0];
		B->nodes[me].box = B->tbox[idx[0]];
	} else {
		aabb_t box = empty_aabb();
		for (size_t i = 0; i < n; ++i) box = extend_aabb(box, B->tbox[idx[i]]);
		B->nodes[me].box = box;
		uint32_t *l = (uint32_t *)malloc(n * sizeof(uint32_t));
		uint32_t *r = (uint32_t *)malloc(n * sizeof(uint32_t));
		if (!l || !r) {
			free(l);
			free(r);
			B->oom = 1;
			return 0;
		}
		size_t nl, nr;
		divide_sah(B, idx, n, l, &nl, r, &nr);
		const uint32_t li = build_aux(B, l, nl);
		free(l);
		if (B->oom) {
			free(r);
			return 0;
		}
		B->nodes[me].left = li;
		const uint32_t ri = build_aux(B, r, nr);
		free(r);
		B->nodes[me].right = ri;
	}
	return me;
}

/* BVH2::build (bvh.hpp:158-193), including the single-triangle special case (:165-181). */
orc_bvh *orc_build(const float *tris, size_t n) {
	builder_t B;
	memset(&B, 0, sizeof(B));
	orc_bvh *res = (orc_bvh *)calloc(1, sizeof(orc_bvh));
	if (!res) return NULL;
	res->tris = (float *)malloc((n ? n : 1) * 9 * sizeof(float));
	B.tbox = (aabb_t *)malloc((n ? n : 1) * sizeof(aabb_t));
	B.cent = (float *)malloc((n ? n : 1) * 3 * sizeof(float));
	uint32_t *idx = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
	if (!res->tris || !B.tbox || !B.cent || !idx) goto fail;
	memcpy(res->tris, tris, n * 9 * sizeof(float));
	res->n_tris = n;
	B.tris = res->tris;
	for (size_t i = 0; i < n; ++i) {
		const float *t = tris + 9 * i;
		B.tbox[i] = make_aabb(t);
		for (int a = 0; a < 3; ++a) B.cent[3 * i + a] = (t[a] + t[a + 3] + t[a + 6]) / 3.0f;
		idx[i] = (uint32_t)i;
	}
	if (n == 1) {
		uint32_t r = push_node(&B), l = push_node(&B), e = push_node(&B);
		if (B.oom) goto fail;
		B.nodes[r].is_leaf = 0;
		B.nodes[r].box = B.tbox[0];
		B.nodes[r].left = l;
		B.nodes[r].right = e;
		B.nodes[l].is_leaf = 1;
		B.nodes[l].tri = 0;
		B.nodes[l].box = B.tbox[0];
		B.nodes[e].is_leaf = 1;
	} else {
		build_aux(&B, idx, n);
		if (B.oom) goto fail;
	}
	free(B.tbox);
	free(B.cent);
	free(idx);
	res->nodes = B.nodes;
	res->n_nodes = B.n_nodes;
	res->cap = B.cap;
	return res;
fail:
	free(B.tbox);
	free(B.cent);
	free(idx);
	free(B.nodes);
	free(res->tris);
	free(res);
	return NULL;
}

void orc_free(orc_bvh *b) {
	if (!b) return;
	free(b->nodes);
	free(b->tris);
	free(b);
}

size_t orc_node_count(const orc_bvh *b) { return b ? b->n_nodes : 0; }

void orc_node(const orc_bvh *b, size_t i, uint32_t *is_leaf, uint32_t *tri, uint32_t *left,
              uint32_t *right, float *bounds6) {
	const node_t *n = &b->nodes[i];
	*is_leaf = n->is_leaf;
	*tri = n->tri;
	*left = n->left;
	*right = n->right;
	for (int k = 0; k < 3; ++k) {
		bounds6[k] = n->box.lo[k];
		bounds6[3 + k] = n->box.hi[k];
	}
}

/* ---------------------------------------------------------------------------- queries */

/* BVH2::ray_box_intersect (bvh.hpp:195-222): reciprocal recomputed per call, std::min/max
 * chains in the reference's order, rejects only tmax < 0 or tmin > tmax. */
int orc_slab(const float *r, const float *bmin, const float *bmax) {
	float df[3];
	df[0] = 1.0f / r[3];
	df[1] = 1.0f / r[4];
	df[2] = 1.0f / r[5];
	const float t1 = (bmin[0] - r[0]) * df[0];
	const float t2 = (bmax[0] - r[0]) * df[0];
	const float t3 = (bmin[1] - r[1]) * df[1];
	const float t4 = (bmax[1] - r[1]) * df[1];
	const float t5 = (bmin[2] - r[2]) * df[2];
	const float t6 = (bmax[2] - r[2]) * df[2];
	const float tmin = smax(smax(smin(t1, t2), smin(t3, t4)), smin(t5, t6));
	const float tmax = smin(smin(smax(t1, t2), smax(t3, t4)), smax(t5, t6));
	if (tmax < 0) return 0;
	if (tmin > tmax) return 0;
	return 1;
}

/* intersect_tri (core.hpp:27-65): Moller-Trumbore, rejects det == 0 exactly, inclusive u/v
 * bounds, one IEEE division, t returned without a range check. */
int orc_mt(const float *tri, const float *r, float *t, float *u, float *v) {
	const float e1[3] = {tri[3] - tri[0], tri[4] - tri[1], tri[5] - tri[2]};
	const float e2[3] = {tri[6] - tri[0], tri[7] - tri[1], tri[8] - tri[2]};
	const float *d = r + 3;
	const float p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2],
	                    d[0] * e2[1] - d[1] * e2[0]};
	const float det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
	if (det == 0.0f) return 0;
	const float inv = 1.0f / det;
	const float tv[3] = {r[0] - tri[0], r[1] - tri[1], r[2] - tri[2]};
	*u = (tv[0] * p[0] + tv[1] * p[1] + tv[2] * p[2]) * inv;
	if (*u < 0.0f || *u > 1.0f) return 0;
	const float q[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2],
	                    tv[0] * e1[1] - tv[1] * e1[0]};
	*v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
	if (*v < 0.0f || *u + *v > 1.0f) return 0;
	*t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
	return 1;
}

/* Watertight ray/triangle test (Woop, Benthin and Wald, "Watertight Ray/Triangle Intersection",
 * JCGT 2(1), 2013), an optional leaf test beside the reference's Moller-Trumbore.
 * The reference has no such test; the semantics kept from core.hpp:27-65 are: both windings
 * accepted, t returned without a range check, (u, v) the weights of vertices 1 and 2. Steps:
 *   kz = axis of the largest |d| (first on ties), kx = kz + 1, ky = kx + 1 (mod 3), kx and ky
 *   swapped when d[kz] < 0; sx = d[kx] / d[kz], sy = d[ky] / d[kz], sz = 1 / d[kz];
 *   A, B, C = v0 - o, v1 - o, v2 - o; sheared ax = A[kx] - sx A[kz], ay = A[ky] - sy A[kz]
 *   (same for B, C); edge functions U = cx by - cy bx, V = ax cy - ay cx, W = bx ay - by ax;
 *   reject when the signs differ (some < 0 and some > 0); det = U + V + W, reject det == 0;
 *   T = U sz A[kz] + V sz B[kz] + W sz C[kz]; t = T / det, u = V / det, v = W / det (one
 *   reciprocal, then products).
 * An edge function that rounds to exactly 0 is re-evaluated from the exact product errors
 * (fmaf): then fl(a b) == fl(c e), so a b - c e = (a b - fl(a b)) - (c e - fl(c e)) exactly and
 * the rounded difference has the exact sign. This replaces the paper's double-precision
 * fallback; no double is used. The shared edge of two triangles gives exactly opposite values
 * (the same two products, subtracted in the other order), so a ray through it is accepted by at
 * least one of them: no ray leaks through a closed mesh. */
static inline float edge_fn(float a, float b, float c, float e) {
	const float p = a * b, q = c * e;
	const float f = p - q;
	if (f != 0.0f) return f;
	return fmaf(a, b, -p) - fmaf(c, e, -q);
}

int orc_wt(const float *tri, const float *r, float *t, float *u, float *v) {
	const float *o = r, *d = r + 3;
	int kz = 0;
	if (fabsf(d[1]) > fabsf(d[kz])) kz = 1;
	if (fabsf(d[2]) > fabsf(d[kz])) kz = 2;
	int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
	if (d[kz] < 0.0f) {
		const int tmp = kx;
		kx = ky;
		ky = tmp;
	}
	const float sx = d[kx] / d[kz], sy = d[ky] / d[kz], sz = 1.0f / d[kz];
	const float A[3] = {tri[0] - o[0], tri[1] - o[1], tri[2] - o[2]};
	const float B[3] = {tri[3] - o[0], tri[4] - o[1], tri[5] - o[2]};
	const float C[3] = {tri[6] - o[0], tri[7] - o[1], tri[8] - o[2]};
	const float ax = A[kx] - sx * A[kz], ay = A[ky] - sy * A[kz];
	const float bx = B[kx] - sx * B[kz], by = B[ky] - sy * B[kz];
	const float cx = C[kx] - sx * C[kz], cy = C[ky] - sy * C[kz];
	const float U = edge_fn(cx, by, cy, bx);
	const float V = edge_fn(ax, cy, ay, cx);
	const float W = edge_fn(bx, ay, by, ax);
	if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
	const float det = U + V + W;
	if (det == 0.0f) return 0;
	const float az = sz * A[kz], bz = sz * B[kz], cz = sz * C[kz];
	const float T = U * az + V * bz + W * cz;
	const float rcp = 1.0f / det;
	*t = T * rcp;
	*u = V * rcp;
	*v = W * rcp;
	return 1;
}

/* Entry and exit distances of the reference slab test (bvh.hpp:195-222, the same operations). */
static void slab_interval(const float *r, const float *bmin, const float *bmax, float *tmin, float *tmax) {
	const float df[3] = {1.0f / r[3], 1.0f / r[4], 1.0f / r[5]};
	const float t1 = (bmin[0] - r[0]) * df[0];
	const float t2 = (bmax[0] - r[0]) * df[0];
	const float t3 = (bmin[1] - r[1]) * df[1];
	const float t4 = (bmax[1] - r[1]) * df[1];
	const float t5 = (bmin[2] - r[2]) * df[2];
	const float t6 = (bmax[2] - r[2]) * df[2];
	*tmin = smax(smax(smin(t1, t2), smin(t3, t4)), smin(t5, t6));
	*tmax = smin(smin(smax(t1, t2), smax(t3, t4)), smax(t5, t6));
}

static float slab_tmax(const float *r, const float *bmin, const float *bmax) {
	float lo, hi;
	slab_interval(r, bmin, bmax, &lo, &hi);
	return hi;
}

/* Candidate test of a triangle in a trace mode (the leaf of bvh.hpp:238-246). The reference:
 * its own box passes the slab test, then Moller-Trumbore. ORC_WATERTIGHT: the watertight test,
 * and the own box not entirely behind the origin (the slab test's tmax >= 0 alone; its
 * tmin <= tmax part can fail by an ulp at a box corner a ray grazes, which would let rays through
 * shared vertices of a closed mesh). */
static inline int candidate(int mode, const float *tri, const float *r, float *t, float *u, float *v) {
	const aabb_t box = make_aabb(tri);
	if (mode & ORC_WATERTIGHT) return !(slab_tmax(r, box.lo, box.hi) < 0.0f) && orc_wt(tri, r, t, u, v);
	return orc_slab(r, box.lo, box.hi) && orc_mt(tri, r, t, u, v);
}

/* Result write-out of nearest_tri (bvh.hpp:259-263). */
static void finish(const float *r, float t_near, float u, float v, uint32_t id, orc_hit *h) {
	const int valid = t_near < INFINITY;
	h->u = valid ? u : 0.0f;
	h->v = valid ? v : 0.0f;
	h->t = t_near;
	h->primitive_id = valid ? id : 0xFFFFFFFFu;
	h->valid = (uint8_t)valid;
	h->pad[0] = h->pad[1] = h->pad[2] = 0;
	h->px = r[0] + t_near * r[3];
	h->py = r[1] + t_near * r[4];
	h->pz = r[2] + t_near * r[5];
}

/* BVH2::nearest_tri (bvh.hpp:224-265): depth-first stack walk, box test on every popped node,
 * push left then right so the right child is visited first. */
static void nearest_tri(const orc_bvh *b, const float *r, int mode, orc_hit *h) {
	uint32_t stack[1024];
	uint32_t sp = 0;
	stack[sp++] = 0;
	float t_near = INFINITY, bu = 0.0f, bv = 0.0f;
	uint32_t bid = 0xFFFFFFFFu;
	while (sp > 0) {
		const node_t *n = &b->nodes[stack[--sp]];
		if (mode & ORC_WATERTIGHT) {
			/* the candidate test decides; node boxes only need to be conservative: the slab
			 * interval widened by 2^-20 of its end points, so that a ray through a box corner
			 * (a vertex shared by several triangles) is never rejected by rounding */
			float lo_t, hi_t;
			slab_interval(r, n->box.lo, n->box.hi, &lo_t, &hi_t);
			const float slack = 0x1p-20f * (fabsf(lo_t) + fabsf(hi_t));
			if (hi_t < -slack || lo_t > hi_t + slack) continue;
		} else if (!orc_slab(r, n->box.lo, n->box.hi)) {
			continue;
		}
		if (n->is_leaf) {
			if (n->tri == 0xFFFFFFFFu) continue;
			float t, u, v;
			if ((mode & ORC_WATERTIGHT) ? candidate(mode, b->tris + 9 * (size_t)n->tri, r, &t, &u, &v)
			                            : orc_mt(b->tris + 9 * (size_t)n->tri, r, &t, &u, &v)) {
				const int take = (t < t_near) || ((mode & 0xFF) == ORC_LOWEST_ID && t == t_near &&
				                                   t_near < INFINITY && n->tri < bid);
				if (take) {
					t_near = t;
					bu = u;
					bv = v;
					bid = n->tri;
				}
			}
		} else {
			if (sp + 2 > 1024) break; /* cannot happen for trees of depth < 1023 */
			stack[sp++] = n->left;
			stack[sp++] = n->right;
		}
	}
	finish(r, t_near, bu, bv, bid, h);
}

static void brute_one(const float *tris, size_t n_tris, const float *r, int mode, orc_hit *h) {
	float t_near = INFINITY, bu = 0.0f, bv = 0.0f;
	uint32_t bid = 0xFFFFFFFFu;
	for (size_t i = 0; i < n_tris; ++i) {
		const float *tri = tris + 9 * i;
		float t, u, v;
		if (candidate(mode, tri, r, &t, &u, &v) && t < t_near) {
			t_near = t;
			bu = u;
			bv = v;
			bid = (uint32_t)i;
		}
	}
	finish(r, t_near, bu, bv, bid, h);
}

/* --------------------------------------------------------------------------- threading */

typedef struct {
	const orc_bvh *b;
	const float *tris;
	size_t n_tris;
	const float *rays;
	orc_hit *out;
	size_t first, count;
	int mode;
	int brute;
} job_t;

static void *run_job(void *arg) {
	job_t *j = (job_t *)arg;
	for (size_t i = 0; i < j->count; ++i) {
		const size_t r = j->first + i;
		if (j->brute)
			brute_one(j->tris, j->n_tris, j->rays + 6 * r, j->mode, &j->out[r]);
		else
			nearest_tri(j->b, j->rays + 6 * r, j->mode, &j->out[r]);
	}
	return NULL;
}

/* CPUBackend::nearest_hits (intersect_cpu.hpp:20-43): n threads over contiguous blocks of
 * size/n rays, the caller handles the remainder. */
static int run(const orc_bvh *b, const float *tris, size_t n_tris, const float *rays, size_t n,
               int mode, int brute, int threads, orc_hit *out) {
	if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
	if (threads <= 0) threads = 1;
	pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
	job_t *jobs = (job_t *)malloc(sizeof(job_t) * (size_t)(threads + 1));
	if (!th || !jobs) {
		free(th);
		free(jobs);
		return 1;
	}
	const size_t per = n / (size_t)threads;
	int *ok = (int *)calloc((size_t)threads, sizeof(int));
	if (!ok) {
		free(th);
		free(jobs);
		return 1;
	}
	for (int i = 0; i <= threads; ++i) {
		job_t *j = &jobs[i];
		j->b = b;
		j->tris = tris;
		j->n_tris = n_tris;
		j->rays = rays;
		j->out = out;
		j->mode = mode;
		j->brute = brute;
		j->first = (size_t)i * per;
		j->count = (i < threads) ? per : n % (size_t)threads;
	}
	for (int i = 0; i < threads; ++i) {
		ok[i] = pthread_create(&th[i], NULL, run_job, &jobs[i]) == 0;
		if (!ok[i]) run_job(&jobs[i]); /* no thread available: do the block here */
	}
	for (int i = 0; i < threads; ++i)
		if (ok[i]) pthread_join(th[i], NULL);
	run_job(&jobs[threads]);
	free(ok);
	free(th);
	free(jobs);
	return 0;
}

int orc_trace(const orc_bvh *b, const float *rays, size_t n, int tie_mode, int threads,
              orc_hit *out) {
	if (!b) return 1;
	return run(b, b->tris, b->n_tris, rays, n, tie_mode, 0, threads, out);
}

int orc_brute(const float *tris, size_t n_tris, const float *rays, size_t n, int threads,
              orc_hit *out) {
	return run(NULL, tris, n_tris, rays, n, ORC_LOWEST_ID, 1, threads, out);
}

int orc_brute_mode(const float *tris, size_t n_tris, const float *rays, size_t n, int mode,
                   int threads, orc_hit *out) {
	return run(NULL, tris, n_tris, rays, n, ORC_LOWEST_ID | (mode & ORC_WATERTIGHT), 1, threads, out);
}

int orc_fp64(const float *tris, size_t n_tris, const float *rays, const uint32_t *ids, size_t n,
             double *t, double *u, double *v) {
	for (size_t i = 0; i < n; ++i) {
		if (ids[i] >= n_tris) {
			t[i] = u[i] = v[i] = NAN;
			continue;
		}
		const float *T = tris + 9 * (size_t)ids[i];
		const float *r = rays + 6 * i;
		double e1[3], e2[3], p[3], tv[3], q[3];
		for (int k = 0; k < 3; ++k) {
			e1[k] = (double)T[3 + k] - (double)T[k];
			e2[k] = (double)T[6 + k] - (double)T[k];
			tv[k] = (double)r[k] - (double)T[k];
		}
		const double d[3] = {r[3], r[4], r[5]};
		p[0] = d[1] * e2[2] - d[2] * e2[1];
		p[1] = d[2] * e2[0] - d[0] * e2[2];
		p[2] = d[0] * e2[1] - d[1] * e2[0];
		const double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
		q[0] = tv[1] * e1[2] - tv[2] * e1[1];
		q[1] = tv[2] * e1[0] - tv[0] * e1[2];
		q[2] = tv[0] * e1[1] - tv[1] * e1[0];
		u[i] = (tv[0] * p[0] + tv[1] * p[1] + tv[2] * p[2]) / det;
		v[i] = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) / det;
		t[i] = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) / det;
	}
	return 0;
}
