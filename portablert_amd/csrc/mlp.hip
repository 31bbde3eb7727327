// The learned t_max seeding network: a 192 -> 256 -> 256 -> 256 -> 4097 MLP (ReLU hidden
// layers, BF16 weights and activations, FP32 accumulation) trained and run with hand-written
// MFMA kernels on gfx950. It predicts, per ray, the BVH8 cluster (subtree) that holds the hit
// triangle, or "miss" (logit 4096). No counterpart in the reference.
//
// Every matrix product is C[M,N] = A[M,K] * B[N,K]^T with both operands K-contiguous (bf16),
// computed by k_gemm with v_mfma_f32_16x16x32_bf16 (lane layout checked on the device:
// A[l%16][(l/16)*8+i], B[(l/16)*8+i][l%16], C[(l/16)*4+i][l%16]). Weights are kept as
// W^T [out, in] (forward) and W [in, out] (backward data gradient); weight gradients use
// transposed activations so that the batch is the contiguous reduction dimension.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "internal.h"
#include "mlp.h"

namespace prt {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 128, kBN = 128, kBK = 32, kThreads = 256;
constexpr int kCPad = 132;  // fp32 row stride of the epilogue tile in LDS
constexpr size_t kSplitFloats = static_cast<size_t>(8) * kMlpOutPad * kMlpHidden;  // split-K / column-sum scratch

__device__ __forceinline__ uint16_t f2bf(float f) {
	// round to nearest even (inputs are finite here)
	const uint32_t u = __float_as_uint(f);
	const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
	return static_cast<uint16_t>(r >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

enum Epi : int {
	E_BIAS_RELU_BF16 = 0,  // out bf16 = relu(acc + bias[n])
	E_BIAS_F32 = 1,        // out f32 = acc + bias[n]
	E_MASK_BF16 = 2,       // out bf16 = acc * (aux_bf16[m][n] > 0)
	E_F32 = 3              // out f32 = acc (split-K partial: out += split * M * N)
};

struct GemmArgs {
	const uint16_t *A;   // [M][lda] bf16
	const uint16_t *B;   // [N][ldb] bf16
	void *out;
	const float *bias;   // [N] (E_BIAS_*)
	const uint16_t *aux; // [M][ldo] bf16 (E_MASK_BF16)
	int M, N, K, lda, ldb, ldo;
	int k_split;         // K range per z-slice (split-K), multiple of kBK
};

// C tile (128 x 128) of one workgroup: 4 waves, each 64 x 64 as 4 x 4 MFMA 16x16 tiles.
template <int EPI>
__global__ __launch_bounds__(kThreads) void k_gemm(GemmArgs g) {
	__shared__ __align__(16) unsigned char smem[kBM * kCPad * 4];
	uint16_t *As = reinterpret_cast<uint16_t *>(smem);            // [128][32]
	uint16_t *Bs = As + kBM * kBK;                                // [128][32]
	float *Cs = reinterpret_cast<float *>(smem);                  // [128][kCPad] after the K loop
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
	const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
	const int kbeg = blockIdx.z * g.k_split;
	const int kend = min(g.K, kbeg + g.k_split);
	f32x4 acc[4][4];
#pragma unroll
	for (int i = 0; i < 4; ++i)
#pragma unroll
		for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

	for (int k0 = kbeg; k0 < kend; k0 += kBK) {
		// global -> LDS: 512 chunks of 16 bytes per operand, two per thread
#pragma unroll
		for (int c2 = 0; c2 < 2; ++c2) {
			const int c = tid + c2 * kThreads;
			const int row = c >> 2, kq = (c & 3) * 8;
			uint4 va = make_uint4(0, 0, 0, 0), vb = make_uint4(0, 0, 0, 0);
			if (m0 + row < g.M) va = *reinterpret_cast<const uint4 *>(g.A + static_cast<size_t>(m0 + row) * g.lda + k0 + kq);
			if (n0 + row < g.N) vb = *reinterpret_cast<const uint4 *>(g.B + static_cast<size_t>(n0 + row) * g.ldb + k0 + kq);
			*reinterpret_cast<uint4 *>(As + row * kBK + kq) = va;
			*reinterpret_cast<uint4 *>(Bs + row * kBK + kq) = vb;
		}
		__syncthreads();
		bf16x8 af[4], bfr[4];
#pragma unroll
		for (int s = 0; s < 4; ++s) {
			af[s] = *reinterpret_cast<const bf16x8 *>(As + (wr + s * 16 + (lane & 15)) * kBK + (lane >> 4) * 8);
			bfr[s] = *reinterpret_cast<const bf16x8 *>(Bs + (wc + s * 16 + (lane & 15)) * kBK + (lane >> 4) * 8);
		}
#pragma unroll
		for (int i = 0; i < 4; ++i)
#pragma unroll
			for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
		__syncthreads();
	}
	// accumulators -> LDS tile (row-major fp32)
#pragma unroll
	for (int i = 0; i < 4; ++i)
#pragma unroll
		for (int j = 0; j < 4; ++j)
#pragma unroll
			for (int e = 0; e < 4; ++e)
				Cs[(wr + i * 16 + (lane >> 4) * 4 + e) * kCPad + wc + j * 16 + (lane & 15)] = acc[i][j][e];
	__syncthreads();

	// element-wise epilogues: each thread handles 64 consecutive columns... of half a row
	const int row = tid >> 1, cbase = (tid & 1) * 64;
	const int m = m0 + row;
	if (m >= g.M) return;
	for (int c = cbase; c < cbase + 64; ++c) {
		const int n = n0 + c;
		if (n >= g.N) break;
		const float x = Cs[row * kCPad + c];
		const size_t o = static_cast<size_t>(m) * g.ldo + n;
		if constexpr (EPI == E_BIAS_RELU_BF16) {
			reinterpret_cast<uint16_t *>(g.out)[o] = f2bf(fmaxf(x + g.bias[n], 0.0f));
		} else if constexpr (EPI == E_BIAS_F32) {
			reinterpret_cast<float *>(g.out)[o] = x + g.bias[n];
		} else if constexpr (EPI == E_MASK_BF16) {
			reinterpret_cast<uint16_t *>(g.out)[o] = bf2f(g.aux[o]) > 0.0f ? f2bf(x) : static_cast<uint16_t>(0);
		} else if constexpr (EPI == E_F32) {
			reinterpret_cast<float *>(g.out)[static_cast<size_t>(blockIdx.z) * g.M * g.ldo + o] = x;
		}
	}
}

// Positional encoding: 6 inputs (origin scaled to [-1,1] by the scene box, direction) times
// 16 frequencies 2^(j/2) * pi, sin and cos: 192 features, written as bf16 in the order
// (input, frequency, sin/cos). encode_inputs writes the features of inputs [a0, a0 + 3).
__device__ __forceinline__ void encode_inputs(const float *r, float4 lo, float4 inv_ext, int a0, uint16_t *o) {
	float in[6];
	in[0] = (r[0] - lo.x) * inv_ext.x * 2.0f - 1.0f;
	in[1] = (r[1] - lo.y) * inv_ext.y * 2.0f - 1.0f;
	in[2] = (r[2] - lo.z) * inv_ext.z * 2.0f - 1.0f;
	in[3] = r[3];
	in[4] = r[4];
	in[5] = r[5];
#pragma unroll
	for (int q = 0; q < 3; ++q) {
		const int a = a0 + q;
		const float x = a0 == 0 ? in[q] : in[3 + q];
#pragma unroll
		for (int j = 0; j < 16; ++j) {
			// v_sin/v_cos take revolutions: sin(2 pi y); y = in * 2^(j/2) / 2
			const float y = x * (0.5f * exp2f(0.5f * static_cast<float>(j)));
			o[a * 32 + j * 2] = f2bf(__builtin_amdgcn_sinf(y));
			o[a * 32 + j * 2 + 1] = f2bf(__builtin_amdgcn_cosf(y));
		}
	}
}

// Inputs a0..a0+2, frequencies j0..j0+7 of one ray, each (sin, cos) pair as one 32-bit store:
// the fused kernel's four threads per ray (the same values as encode_inputs).
__device__ __forceinline__ void encode_quarter(const float *r, float4 lo, float4 inv_ext, int a0, int j0, uint16_t *o) {
	float in[3];
	if (a0 == 0) {
		in[0] = (r[0] - lo.x) * inv_ext.x * 2.0f - 1.0f;
		in[1] = (r[1] - lo.y) * inv_ext.y * 2.0f - 1.0f;
		in[2] = (r[2] - lo.z) * inv_ext.z * 2.0f - 1.0f;
	} else {
		in[0] = r[3];
		in[1] = r[4];
		in[2] = r[5];
	}
	uint32_t *o32 = reinterpret_cast<uint32_t *>(o);
#pragma unroll
	for (int q = 0; q < 3; ++q) {
#pragma unroll
		for (int jj = 0; jj < 8; ++jj) {
			const int j = j0 + jj;
			const float y = in[q] * (0.5f * exp2f(0.5f * static_cast<float>(j)));
			o32[(a0 + q) * 16 + j] = static_cast<uint32_t>(f2bf(__builtin_amdgcn_sinf(y))) |
			                         (static_cast<uint32_t>(f2bf(__builtin_amdgcn_cosf(y))) << 16);
		}
	}
}

__global__ __launch_bounds__(256) void k_encode(const float *__restrict__ rays, uint64_t first, uint32_t n, float4 lo,
                                                float4 inv_ext, uint16_t *__restrict__ x) {
	const uint32_t i = blockIdx.x * 256 + threadIdx.x;
	if (i >= n) return;
	const float *r = rays + 6 * (first + i);
	uint16_t *o = x + static_cast<size_t>(i) * kMlpIn;
	encode_inputs(r, lo, inv_ext, 0, o);
	encode_inputs(r, lo, inv_ext, 3, o);
}

// ------------------------------------------------------------- fused inference with top-4
//
// One workgroup (4 waves) takes 64 rays through the whole network: encoding, the three hidden
// layers and the 4097 output logits, keeping the activations in LDS (X, 64 x 256 bf16) and
// streaming W^T tiles of 64 output neurons (prefetched into registers one tile ahead) from L2,
// with v_mfma_f32_16x16x32_bf16 over K in steps of 32. In every layer each wave loads all 64 rows
// of the layer's input into registers and owns 16 of each tile's 64 neurons, so each W fragment
// read from LDS serves 64 rows; a hidden layer writes its outputs to X as its tiles finish
// (two barriers per layer). In the output layer each wave fetches and stages only its own 16
// rows of W^T (in its slice of two tile buffers, one of them in X's space), and passes its
// 64 x 16 logits through the same slice to its 64 lanes, one ray
// each, which keep a running top-4 (value descending, index ascending). No wave reads another's
// slice, so the 65 output tiles run without a barrier; the four waves' lists of a ray are merged
// at the end, and no logit leaves the chip. The tiles are software-pipelined: a wave issues tile
// t + 1's MFMAs interleaved with the compares of tile t's logits against its running bound.
// The workgroup uses 66 KB of LDS, so two share a CU.
constexpr int kFRows = 64;     // rays per workgroup
constexpr int kFCols = 64;     // output neurons per W tile
constexpr int kFStride = 264;  // bf16 row stride of X and W (16-byte reads without bank conflicts)

struct FusedArgs {
	const float *rays;
	uint64_t first;
	int rows;
	float4 lo, inv_ext;
	const uint16_t *w[4];  // W^T bf16 [out][in]
	const float *b[4];
	uint32_t *top4;        // [rows][4]
};

// One hidden layer computed per wave, without barriers between its tiles: every wave loads all 64
// rows of X[:, :K] into registers, then owns 16 of each 64-neuron tile's columns and stages only
// its own 16 rows of W^T in its slice of the W region (the next tile's rows prefetched into
// registers during the MFMAs), so each W fragment read from LDS serves 64 rows. Once every wave
// holds X, X's region takes the outputs as each tile finishes. Two barriers per layer; each
// output element is the same MFMA chain over K in the same order as before.
template <int K>
__device__ __forceinline__ void hidden_layer_pw(uint16_t *X, uint16_t *W, const uint16_t *wt, const float *bias,
                                                int lane, int wave) {
	constexpr int KS = K / kBK;       // MFMA K steps
	constexpr int RC = K / 8;         // 16-byte chunks per W^T row
	constexpr int CH = 16 * RC / 64;  // chunks per lane of a 16-row slice
	bf16x8 xr[4][KS];
#pragma unroll
	for (int i = 0; i < 4; ++i)
#pragma unroll
		for (int kk = 0; kk < KS; ++kk)
			xr[i][kk] = *reinterpret_cast<const bf16x8 *>(X + (i * 16 + (lane & 15)) * kFStride + kk * kBK + (lane >> 4) * 8);
	uint16_t *ws = W + wave * 16 * kFStride;  // this wave's slice
	auto fetch = [&](int t, u32x4 q[CH]) {
#pragma unroll
		for (int j = 0; j < CH; ++j) {
			const int c = lane + 64 * j;
			q[j] = *reinterpret_cast<const u32x4 *>(wt + static_cast<size_t>(t * kFCols + wave * 16 + c / RC) * K + (c % RC) * 8);
		}
	};
	auto store = [&](const u32x4 q[CH]) {
#pragma unroll
		for (int j = 0; j < CH; ++j) {
			const int c = lane + 64 * j;
			*reinterpret_cast<u32x4 *>(ws + (c / RC) * kFStride + (c % RC) * 8) = q[j];
		}
	};
	u32x4 pre[CH];
	fetch(0, pre);
	__syncthreads();  // every wave holds X in registers: X's region takes the outputs
	store(pre);
#pragma unroll
	for (int t = 0; t < 4; ++t) {
		if (t + 1 < 4) fetch(t + 1, pre);
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		f32x4 acc[4];
#pragma unroll
		for (int i = 0; i < 4; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
		for (int kk = 0; kk < KS; ++kk) {
			const bf16x8 bfr = *reinterpret_cast<const bf16x8 *>(ws + (lane & 15) * kFStride + kk * kBK + (lane >> 4) * 8);
#pragma unroll
			for (int i = 0; i < 4; ++i) acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xr[i][kk], bfr, acc[i], 0, 0, 0);
		}
		if (t + 1 < 4) store(pre);  // this wave's MFMAs have read the slice (LDS operations of one wave complete in order)
		const int col = t * kFCols + wave * 16 + (lane & 15);
		const float bj = bias[col];
#pragma unroll
		for (int i = 0; i < 4; ++i)
#pragma unroll
			for (int e = 0; e < 4; ++e)
				X[(i * 16 + (lane >> 4) * 4 + e) * kFStride + col] = f2bf(fmaxf(acc[i][e] + bj, 0.0f));
	}
	__syncthreads();
}

__device__ __forceinline__ bool top_better(float va, uint32_t ia, float vb, uint32_t ib) {
	return va > vb || (va == vb && ia < ib);
}

// Inserts (x, n) into a list of 4 kept in (value descending, index ascending) order.
__device__ __forceinline__ void top_insert(float v[4], uint32_t ix[4], float x, uint32_t n) {
	if (!top_better(x, n, v[3], ix[3])) return;
	v[3] = x;
	ix[3] = n;
#pragma unroll
	for (int k = 3; k > 0; --k) {
		if (top_better(v[k], ix[k], v[k - 1], ix[k - 1])) {
			const float tv = v[k]; v[k] = v[k - 1]; v[k - 1] = tv;
			const uint32_t ti = ix[k]; ix[k] = ix[k - 1]; ix[k - 1] = ti;
		}
	}
}

// Merges the list of the lane `mask` away into this lane's list.
__device__ __forceinline__ void top_merge(float v[4], uint32_t ix[4], int mask) {
	float ov[4];
	uint32_t oi[4];
#pragma unroll
	for (int q = 0; q < 4; ++q) {
		ov[q] = __shfl_xor(v[q], mask);
		oi[q] = __shfl_xor(ix[q], mask);
	}
#pragma unroll
	for (int q = 0; q < 4; ++q) top_insert(v, ix, ov[q], oi[q]);
}

__global__ __launch_bounds__(256, 2) void k_mlp_top4(FusedArgs a) {
	__shared__ __align__(16) uint16_t X[kFRows * kFStride];
	__shared__ __align__(16) uint16_t W[kFCols * kFStride];
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int r0 = blockIdx.x * kFRows;
	// encoding: two threads per ray, three inputs each; rows past the end are zero
	{
		// four threads per ray: inputs 0-2 or 3-5, frequencies 0-7 or 8-15
		const int row = tid >> 2, half = (tid >> 1) & 1, fq = tid & 1;
		uint16_t *o = X + row * kFStride;
		if (r0 + row < a.rows) {
			encode_quarter(a.rays + 6 * (a.first + r0 + row), a.lo, a.inv_ext, half * 3, fq * 8, o);
		} else {
			uint32_t *o32 = reinterpret_cast<uint32_t *>(o);
#pragma unroll
			for (int q = 0; q < 3; ++q)
#pragma unroll
				for (int jj = 0; jj < 8; ++jj) o32[(half * 3 + q) * 16 + fq * 8 + jj] = 0u;
		}
	}
	__syncthreads();
	hidden_layer_pw<kMlpIn>(X, W, a.w[0], a.b[0], lane, wave);
	hidden_layer_pw<kMlpHidden>(X, W, a.w[1], a.b[1], lane, wave);
	hidden_layer_pw<kMlpHidden>(X, W, a.w[2], a.b[2], lane, wave);
	// output layer: all 64 rows of the last hidden layer in every wave's registers; each wave
	// computes 16 of a tile's 64 columns, so each W fragment read from LDS serves 64 rows
	const int oc = wave * 16;
	bf16x8 ar[4][kMlpHidden / kBK];
#pragma unroll
	for (int i = 0; i < 4; ++i)
#pragma unroll
		for (int kk = 0; kk < kMlpHidden / kBK; ++kk)
			ar[i][kk] = *reinterpret_cast<const bf16x8 *>(X + (i * 16 + (lane & 15)) * kFStride + kk * kBK + (lane >> 4) * 8);
	// Output tiles without barriers: every wave fetches, stages and reads only its own 16 rows of
	// each W^T tile (columns oc .. oc + 15 of the 64 output neurons), in its own slice of the two
	// tile buffers, and passes its 64 x 16 block of logits through that slice (free once its
	// MFMAs have read it) to its 64 lanes, one row each: lane r keeps the running top-4 of row r
	// over this wave's columns. The four waves' lists of a row are merged at the end.
	constexpr int kSliceStride = 20;  // fp32 row stride of a wave's 64 x 16 logit block (16-byte rows, no bank conflicts)
	auto slice = [&](auto *buf) { return buf + oc * kFStride; };
	// this wave's 16 rows of tile t's W^T, 16 x 512 bytes: chunk c = lane + 64 j is row c / 32,
	// 16-byte column group c % 32
	auto w_chunk = [&](int t, int j) {
		const int c = lane + 64 * j;
		return *reinterpret_cast<const u32x4 *>(a.w[3] + static_cast<size_t>(t * kFCols + oc + c / 32) * kMlpHidden +
		                                        (c % 32) * 8);
	};
	auto fetch_own = [&](int t, u32x4 q[8]) {
#pragma unroll
		for (int j = 0; j < 8; ++j) q[j] = w_chunk(t, j);
	};
	auto store_own = [&](uint16_t *buf, const u32x4 q[8]) {
#pragma unroll
		for (int j = 0; j < 8; ++j) {
			const int c = lane + 64 * j;
			*reinterpret_cast<u32x4 *>(slice(buf) + (c / 32) * kFStride + (c % 32) * 8) = q[j];
		}
	};
	float v[4] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
	uint32_t ix[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
	constexpr int kOutTiles = (kMlpOut + kFCols - 1) / kFCols;  // the last one holds the miss logit
	// bias of this lane's column (padding columns: -inf, never taken)
	auto bias_of = [&](int t) {
		const int col = t * kFCols + oc + (lane & 15);
		const float b = a.b[3][col < kMlpOut ? col : kMlpOut - 1];  // no branch around the load
		return col < kMlpOut ? b : -__builtin_inff();
	};
	auto mma = [&](const uint16_t *buf, f32x4 acc[4]) {
#pragma unroll
		for (int i = 0; i < 4; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
		for (int kk = 0; kk < kMlpHidden / kBK; ++kk) {
			const bf16x8 bfr = *reinterpret_cast<const bf16x8 *>(slice(buf) + (lane & 15) * kFStride + kk * kBK + (lane >> 4) * 8);
#pragma unroll
			for (int i = 0; i < 4; ++i) acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar[i][kk], bfr, acc[i], 0, 0, 0);
		}
	};
	// this wave's logits into its slice of buf, whose W the MFMAs that made them have read (LDS
	// operations of one wave complete in order)
	auto put_logits = [&](uint16_t *buf, const f32x4 acc[4], float bias) {
		float *C = reinterpret_cast<float *>(slice(buf));
#pragma unroll
		for (int i = 0; i < 4; ++i)
#pragma unroll
			for (int e = 0; e < 4; ++e) C[(i * 16 + (lane >> 4) * 4 + e) * kSliceStride + (lane & 15)] = acc[i][e] + bias;
	};
	auto sync_wave = [&]() {
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
	};
	// lane r compares row r's 16 logits of tile t with its 4th best (a strict comparison keeps
	// the lower index first; held entries come from earlier tiles)
	auto scan_mask = [&](const uint16_t *buf) {
		const float *crow = reinterpret_cast<const float *>(slice(buf)) + lane * kSliceStride;
		uint32_t cand = 0;
#pragma unroll
		for (int c4 = 0; c4 < 4; ++c4) {
			const float4 q = *reinterpret_cast<const float4 *>(crow + c4 * 4);
			const float xs[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
			for (int e = 0; e < 4; ++e) cand |= xs[e] > v[3] ? (1u << (c4 * 4 + e)) : 0u;
		}
		return cand;
	};
	auto insert = [&](const uint16_t *buf, uint32_t cand, int t) {
		const float *crow = reinterpret_cast<const float *>(slice(buf)) + lane * kSliceStride;
		const int nbase = t * kFCols + oc;
		while (cand) {
			const int c = __builtin_ctz(cand);
			cand &= cand - 1u;
			const float x = crow[c];
			if (x > v[3]) {
				v[3] = x;
				ix[3] = static_cast<uint32_t>(nbase + c);
#pragma unroll
				for (int k = 3; k > 0; --k) {
					if (v[k] > v[k - 1]) {
						const float tv = v[k]; v[k] = v[k - 1]; v[k - 1] = tv;
						const uint32_t ti = ix[k]; ix[k] = ix[k - 1]; ix[k - 1] = ti;
					}
				}
			}
		}
	};
	// Software pipeline over the output tiles: while the matrix core computes tile t + 1, the
	// wave scans tile t's logits. Buffer roles alternate: at the top of iteration t, `cur` is
	// free (its W was read by tile t's MFMAs, whose results are in acc) and `nxt` holds tile
	// t + 1's W.
	f32x4 acc[4];
	uint16_t *cur = W, *nxt = X;
	{
		u32x4 p0[8];
		fetch_own(0, p0);
		__syncthreads();  // every wave has its rows of X in registers: X's space becomes a tile buffer
		store_own(cur, p0);
		fetch_own(1, p0);
		store_own(nxt, p0);
		sync_wave();
		mma(cur, acc);
	}
	float bias = bias_of(0);
	for (int t = 0; t < kOutTiles - 1; ++t) {
		put_logits(cur, acc, bias);
		bias = bias_of(t + 1);
		sync_wave();
		// Tile t + 1's MFMAs with tile t's scan in hand-ordered steps (a scheduling barrier after
		// each keeps the order): the four logit reads and the first two B fragment reads, then per
		// K step the B fragment read two steps ahead, four MFMAs, one W prefetch load and two of
		// the sixteen compares, which issue while the matrix core works.
		const float *crow = reinterpret_cast<const float *>(slice(cur)) + lane * kSliceStride;
		float lf[16];
#pragma unroll
		for (int c4 = 0; c4 < 4; ++c4) {
			const float4 q = *reinterpret_cast<const float4 *>(crow + c4 * 4);
			lf[c4 * 4] = q.x;
			lf[c4 * 4 + 1] = q.y;
			lf[c4 * 4 + 2] = q.z;
			lf[c4 * 4 + 3] = q.w;
		}
		const uint16_t *bsrc = slice(nxt) + (lane & 15) * kFStride + (lane >> 4) * 8;
		bf16x8 bq[kMlpHidden / kBK];
		bq[0] = *reinterpret_cast<const bf16x8 *>(bsrc);
		bq[1] = *reinterpret_cast<const bf16x8 *>(bsrc + kBK);
#pragma unroll
		for (int i = 0; i < 4; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
		__builtin_amdgcn_sched_barrier(0);
		u32x4 pre[8];
		// tile t + 2's W rows of this wave: chunk kk of w_chunk is 2 kk rows further
		const uint16_t *wsrc = a.w[3] + static_cast<size_t>(((t + 2) % kOutTiles) * kFCols + oc + (lane >> 5)) * kMlpHidden +
		                       (lane & 31) * 8;
		uint32_t cand = 0;
		const float thr = v[3];
#pragma unroll
		for (int kk = 0; kk < kMlpHidden / kBK; ++kk) {
			if (kk + 2 < kMlpHidden / kBK) bq[kk + 2] = *reinterpret_cast<const bf16x8 *>(bsrc + (kk + 2) * kBK);
			pre[kk] = *reinterpret_cast<const u32x4 *>(wsrc + 2 * kk * kMlpHidden);
#pragma unroll
			for (int i = 0; i < 4; ++i) acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar[i][kk], bq[kk], acc[i], 0, 0, 0);
			cand |= (lf[2 * kk] > thr ? (1u << (2 * kk)) : 0u) | (lf[2 * kk + 1] > thr ? (1u << (2 * kk + 1)) : 0u);
			__builtin_amdgcn_sched_barrier(0);
		}
		insert(cur, cand, t);
		sync_wave();
		store_own(cur, pre);  // tile t + 2's W into the buffer tile t's logits have left
		uint16_t *tmp = cur; cur = nxt; nxt = tmp;
	}
	put_logits(cur, acc, bias);
	sync_wave();
	insert(cur, scan_mask(cur), kOutTiles - 1);
	// merge the four waves' lists of each row (value descending, index ascending)
	__syncthreads();
	float *mv = reinterpret_cast<float *>(W);
	uint32_t *mi = reinterpret_cast<uint32_t *>(W) + 4 * 64 * 4;
#pragma unroll
	for (int k = 0; k < 4; ++k) {
		mv[(wave * 64 + lane) * 4 + k] = v[k];
		mi[(wave * 64 + lane) * 4 + k] = ix[k];
	}
	__syncthreads();
	const int orow = tid >> 2, opart = tid & 3;
#pragma unroll
	for (int k = 0; k < 4; ++k) {
		v[k] = mv[(opart * 64 + orow) * 4 + k];
		ix[k] = mi[(opart * 64 + orow) * 4 + k];
	}
	top_merge(v, ix, 1);
	top_merge(v, ix, 2);
	if (opart == 0 && r0 + orow < a.rows) {
		const uint32_t res[4] = {ix[0], ix[1], ix[2], ix[3]};
		reinterpret_cast<uint4 *>(a.top4)[r0 + orow] = make_uint4(res[0], res[1], res[2], res[3]);
	}
}

// Softmax cross-entropy over the first n_valid logits of each row. Writes dlogits =
// (softmax - onehot) / batch as bf16 (0 in padding columns) and accumulates the mean loss.
__global__ __launch_bounds__(256) void k_softmax_ce(const float *__restrict__ logits, const uint32_t *__restrict__ label,
                                                    int rows, int ld, int n_valid, float inv_batch,
                                                    uint16_t *__restrict__ dlog, float *loss) {
	const int row = blockIdx.x;
	if (row >= rows) return;
	__shared__ float red[256];
	const float *L = logits + static_cast<size_t>(row) * ld;
	float mx = -__builtin_inff();
	for (int c = threadIdx.x; c < n_valid; c += 256) mx = fmaxf(mx, L[c]);
	red[threadIdx.x] = mx;
	__syncthreads();
	for (int s = 128; s > 0; s >>= 1) {
		if (static_cast<int>(threadIdx.x) < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
		__syncthreads();
	}
	mx = red[0];
	__syncthreads();
	float sum = 0.0f;
	for (int c = threadIdx.x; c < n_valid; c += 256) sum += expf(L[c] - mx);
	red[threadIdx.x] = sum;
	__syncthreads();
	for (int s = 128; s > 0; s >>= 1) {
		if (static_cast<int>(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
		__syncthreads();
	}
	sum = red[0];
	const uint32_t y = label[row];
	const float inv = 1.0f / sum;
	uint16_t *D = dlog + static_cast<size_t>(row) * ld;
	for (int c = threadIdx.x; c < ld; c += 256) {
		float gval = 0.0f;
		if (c < n_valid) gval = (expf(L[c] - mx) * inv - (static_cast<uint32_t>(c) == y ? 1.0f : 0.0f)) * inv_batch;
		D[c] = f2bf(gval);
	}
	if (threadIdx.x == 0) atomicAdd(loss, (logf(sum) + mx - L[y]) * inv_batch);
}

// out[c][r] = in[r][c] for a bf16 [rows][cols] matrix (32 x 32 tiles through LDS).
__global__ __launch_bounds__(256) void k_transpose(const uint16_t *__restrict__ in, int rows, int cols, int ld_in,
                                                   uint16_t *__restrict__ out, int ld_out) {
	__shared__ uint16_t t[32][33];
	const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
	const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
	for (int j = ty; j < 32; j += 8) {
		const int r = by + j, c = bx + tx;
		t[j][tx] = (r < rows && c < cols) ? in[static_cast<size_t>(r) * ld_in + c] : static_cast<uint16_t>(0);
	}
	__syncthreads();
	for (int j = ty; j < 32; j += 8) {
		const int c = bx + j, r = by + tx;
		if (c < cols && r < rows) out[static_cast<size_t>(c) * ld_out + r] = t[tx][j];
	}
}

// Sums split-K partials: g[i] = sum_s part[s][i].
__global__ __launch_bounds__(256) void k_reduce_split(const float *__restrict__ part, int splits, size_t n,
                                                      float *__restrict__ g) {
	const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
	if (i >= n) return;
	float s = 0.0f;
	for (int k = 0; k < splits; ++k) s += part[k * n + i];
	g[i] = s;
}

// Column sums of a bf16 [rows][ld] matrix over its first cols columns (bias gradient), first
// stage: part[chunk][c] = sum of rows [chunk * per, (chunk + 1) * per) of column c. The second
// stage (k_reduce_split) adds the chunks in a fixed order.
__global__ __launch_bounds__(256) void k_colsum_part(const uint16_t *__restrict__ d, int rows, int cols, int ld, int per,
                                                     float *__restrict__ part) {
	const int c = blockIdx.x * 256 + threadIdx.x;
	if (c >= cols) return;
	const int r0 = blockIdx.y * per, r1 = min(rows, r0 + per);
	float s = 0.0f;
	for (int r = r0; r < r1; ++r) s += bf2f(d[static_cast<size_t>(r) * ld + c]);
	part[static_cast<size_t>(blockIdx.y) * cols + c] = s;
}

// Adam on fp32 master parameters; refreshes the bf16 copy (W^T [rows][cols]) and, when wt is
// non-null, the transposed bf16 copy W [cols][rows].
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                                              float *__restrict__ v, size_t n, float lr, float b1, float b2, float eps,
                                              float c1, float c2, uint16_t *__restrict__ bf, uint16_t *__restrict__ bft,
                                              int rows, int cols) {
	const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
	if (i >= n) return;
	const float gi = g[i];
	const float mi = b1 * m[i] + (1.0f - b1) * gi;
	const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
	m[i] = mi;
	v[i] = vi;
	const float pi = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
	p[i] = pi;
	if (bf) {
		const uint16_t h = f2bf(pi);
		bf[i] = h;
		if (bft) {
			const int r = static_cast<int>(i / cols), c = static_cast<int>(i % cols);
			bft[static_cast<size_t>(c) * rows + r] = h;
		}
	}
}

// Deterministic initialisation (uniform +-sqrt(6/(fan_in+fan_out)) from a counter hash).
__global__ __launch_bounds__(256) void k_init(float *__restrict__ w, size_t n, int fan_in, int fan_out, uint64_t seed,
                                              int rows_valid, int cols) {
	const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
	if (i >= n) return;
	uint64_t z = seed + i * 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	z ^= z >> 31;
	const float u = static_cast<float>(z >> 40) * 0x1p-24f * 2.0f - 1.0f;
	const float a = sqrtf(6.0f / static_cast<float>(fan_in + fan_out));
	w[i] = (static_cast<int>(i / cols) < rows_valid) ? u * a : 0.0f;
}

__global__ __launch_bounds__(256) void k_fill_f32(float *p, size_t n, float v) {
	const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
	if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_to_bf16(const float *__restrict__ p, size_t n, uint16_t *__restrict__ bf,
                                                 uint16_t *__restrict__ bft, int rows, int cols) {
	const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
	if (i >= n) return;
	const uint16_t h = f2bf(p[i]);
	bf[i] = h;
	if (bft) bft[static_cast<size_t>(i % cols) * rows + i / cols] = h;
}

inline unsigned blocks(size_t n, int per = 256) { return static_cast<unsigned>((n + per - 1) / per); }

template <int EPI>
hipError_t gemm(const GemmArgs &g, int splits, hipStream_t st) {
	dim3 grid((g.N + kBN - 1) / kBN, (g.M + kBM - 1) / kBM, splits);
	hipLaunchKernelGGL((k_gemm<EPI>), grid, dim3(kThreads), 0, st, g);
	return hipGetLastError();
}

}  // namespace

// ------------------------------------------------------------------------------ host side

const int kLayerIn[4] = {kMlpIn, kMlpHidden, kMlpHidden, kMlpHidden};
const int kLayerOut[4] = {kMlpHidden, kMlpHidden, kMlpHidden, kMlpOutPad};

hipError_t mlp_init(Mlp &m, uint64_t seed, hipStream_t st) {
#define PRT_TRY(x)                                                                                 \
	do {                                                                                           \
		hipError_t e_ = (x);                                                                       \
		if (e_ != hipSuccess) return e_;                                                           \
	} while (0)
	for (int l = 0; l < 4; ++l) {
		const size_t n = static_cast<size_t>(kLayerOut[l]) * kLayerIn[l];
		PRT_TRY(m.w[l].ensure(n * 4));
		PRT_TRY(m.b[l].ensure(kLayerOut[l] * 4));
		PRT_TRY(m.wt_bf[l].ensure(n * 2));
		PRT_TRY(m.w_bf[l].ensure(n * 2));
		PRT_TRY(m.mw[l].ensure(n * 4));
		PRT_TRY(m.vw[l].ensure(n * 4));
		PRT_TRY(m.mb[l].ensure(kLayerOut[l] * 4));
		PRT_TRY(m.vb[l].ensure(kLayerOut[l] * 4));
		PRT_TRY(m.gw[l].ensure(n * 4));
		PRT_TRY(m.gb[l].ensure(kLayerOut[l] * 4));
		const int valid = l == 3 ? kMlpOut : kLayerOut[l];
		hipLaunchKernelGGL(k_init, dim3(blocks(n)), dim3(256), 0, st, m.w[l].as<float>(), n, kLayerIn[l], valid,
		                   seed * 31 + l, valid, kLayerIn[l]);
		hipLaunchKernelGGL(k_fill_f32, dim3(blocks(kLayerOut[l])), dim3(256), 0, st, m.b[l].as<float>(),
		                   static_cast<size_t>(kLayerOut[l]), 0.0f);
		if (l == 3) {  // padding logits never win and get no probability mass
			hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(256), 0, st, m.b[l].as<float>() + kMlpOut,
			                   static_cast<size_t>(kMlpOutPad - kMlpOut), -1e30f);
		}
		hipLaunchKernelGGL(k_fill_f32, dim3(blocks(n)), dim3(256), 0, st, m.mw[l].as<float>(), n, 0.0f);
		hipLaunchKernelGGL(k_fill_f32, dim3(blocks(n)), dim3(256), 0, st, m.vw[l].as<float>(), n, 0.0f);
		hipLaunchKernelGGL(k_fill_f32, dim3(blocks(kLayerOut[l])), dim3(256), 0, st, m.mb[l].as<float>(),
		                   static_cast<size_t>(kLayerOut[l]), 0.0f);
		hipLaunchKernelGGL(k_fill_f32, dim3(blocks(kLayerOut[l])), dim3(256), 0, st, m.vb[l].as<float>(),
		                   static_cast<size_t>(kLayerOut[l]), 0.0f);
		hipLaunchKernelGGL(k_to_bf16, dim3(blocks(n)), dim3(256), 0, st, m.w[l].as<float>(), n, m.wt_bf[l].as<uint16_t>(),
		                   m.w_bf[l].as<uint16_t>(), kLayerOut[l], kLayerIn[l]);
	}
	m.step = 0;
	m.ready = true;
	return hipGetLastError();
}

hipError_t mlp_refresh_bf16(Mlp &m, hipStream_t st) {
	for (int l = 0; l < 4; ++l) {
		const size_t n = static_cast<size_t>(kLayerOut[l]) * kLayerIn[l];
		hipLaunchKernelGGL(k_to_bf16, dim3(blocks(n)), dim3(256), 0, st, m.w[l].as<float>(), n, m.wt_bf[l].as<uint16_t>(),
		                   m.w_bf[l].as<uint16_t>(), kLayerOut[l], kLayerIn[l]);
	}
	return hipGetLastError();
}

static hipError_t ensure_act(Mlp &m, int rows, bool train) {
	PRT_TRY(m.x0.ensure(static_cast<size_t>(rows) * kMlpIn * 2));
	for (int l = 0; l < 3; ++l) PRT_TRY(m.h[l].ensure(static_cast<size_t>(rows) * kMlpHidden * 2));
	if (train) {
		PRT_TRY(m.logits.ensure(static_cast<size_t>(rows) * kMlpOutPad * 4));
		PRT_TRY(m.dlog.ensure(static_cast<size_t>(rows) * kMlpOutPad * 2));
		PRT_TRY(m.dlog_t.ensure(static_cast<size_t>(rows) * kMlpOutPad * 2));
		for (int l = 0; l < 3; ++l) PRT_TRY(m.dh[l].ensure(static_cast<size_t>(rows) * kMlpHidden * 2));
		PRT_TRY(m.act_t.ensure(static_cast<size_t>(rows) * kMlpHidden * 2));
		PRT_TRY(m.dact_t.ensure(static_cast<size_t>(rows) * kMlpHidden * 2));
		PRT_TRY(m.split.ensure(kSplitFloats * 4));
		PRT_TRY(m.loss.ensure(4));
	}
	return hipSuccess;
}

hipError_t mlp_encode(Mlp &m, const float *rays, uint64_t first, int rows, hipStream_t st) {
	// the encoding may be the first kernel of a batch: size its output here
	PRT_TRY(m.x0.ensure(static_cast<size_t>(rows) * kMlpIn * 2));
	const float4 lo = make_float4(m.scene_lo[0], m.scene_lo[1], m.scene_lo[2], 0.f);
	const float4 ie = make_float4(m.scene_inv_ext[0], m.scene_inv_ext[1], m.scene_inv_ext[2], 0.f);
	hipLaunchKernelGGL(k_encode, dim3(blocks(rows)), dim3(256), 0, st, rays, first, static_cast<uint32_t>(rows), lo, ie,
	                   m.x0.as<uint16_t>());
	return hipGetLastError();
}

// Hidden layers: x0 -> h0 -> h1 -> h2.
static hipError_t forward_hidden(Mlp &m, int rows, hipStream_t st) {
	const uint16_t *in = m.x0.as<uint16_t>();
	for (int l = 0; l < 3; ++l) {
		GemmArgs g{};
		g.A = in;
		g.B = m.wt_bf[l].as<uint16_t>();
		g.out = m.h[l].ptr;
		g.bias = m.b[l].as<float>();
		g.M = rows;
		g.N = kMlpHidden;
		g.K = kLayerIn[l];
		g.lda = kLayerIn[l];
		g.ldb = kLayerIn[l];
		g.ldo = kMlpHidden;
		g.k_split = g.K;
		PRT_TRY(gemm<E_BIAS_RELU_BF16>(g, 1, st));
		in = m.h[l].as<uint16_t>();
	}
	return hipSuccess;
}

hipError_t mlp_forward_logits(Mlp &m, int rows, hipStream_t st) {
	PRT_TRY(ensure_act(m, rows, true));
	PRT_TRY(forward_hidden(m, rows, st));
	GemmArgs g{};
	g.A = m.h[2].as<uint16_t>();
	g.B = m.wt_bf[3].as<uint16_t>();
	g.out = m.logits.ptr;
	g.bias = m.b[3].as<float>();
	g.M = rows;
	g.N = kMlpOutPad;
	g.K = kMlpHidden;
	g.lda = g.ldb = kMlpHidden;
	g.ldo = kMlpOutPad;
	g.k_split = g.K;
	return gemm<E_BIAS_F32>(g, 1, st);
}

hipError_t mlp_infer_top4(Mlp &m, const float *rays, uint64_t first, int rows, uint32_t *top4, hipStream_t st) {
	if (rows <= 0) return hipSuccess;
	FusedArgs f;
	f.rays = rays;
	f.first = first;
	f.rows = rows;
	f.lo = make_float4(m.scene_lo[0], m.scene_lo[1], m.scene_lo[2], 0.f);
	f.inv_ext = make_float4(m.scene_inv_ext[0], m.scene_inv_ext[1], m.scene_inv_ext[2], 0.f);
	for (int l = 0; l < 4; ++l) {
		f.w[l] = m.wt_bf[l].as<uint16_t>();
		f.b[l] = m.b[l].as<float>();
	}
	f.top4 = top4;
	hipLaunchKernelGGL(k_mlp_top4, dim3((rows + kFRows - 1) / kFRows), dim3(256), 0, st, f);
	return hipGetLastError();
}

// Backward of one batch whose logits are in m.logits; gradients land in m.gw / m.gb.
hipError_t mlp_backward(Mlp &m, const uint32_t *labels, int rows, hipStream_t st) {
	PRT_TRY(ensure_act(m, rows, true));
	hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(256), 0, st, m.loss.as<float>(), static_cast<size_t>(1), 0.0f);
	hipLaunchKernelGGL(k_softmax_ce, dim3(rows), dim3(256), 0, st, m.logits.as<float>(), labels, rows, kMlpOutPad, kMlpOut,
	                   1.0f / static_cast<float>(rows), m.dlog.as<uint16_t>(), m.loss.as<float>());
	const uint16_t *dy = m.dlog.as<uint16_t>();
	int dy_cols = kMlpOutPad;
	for (int l = 3; l >= 0; --l) {
		const int in = kLayerIn[l], out = kLayerOut[l];
		const uint16_t *x = l == 0 ? m.x0.as<uint16_t>() : m.h[l - 1].as<uint16_t>();
		// bias gradient: column sums of dy (chunks of >= 128 rows whose partial sums fit the
		// split-K scratch, then the chunks in order)
		const int max_chunks = static_cast<int>(std::min<size_t>(65535, kSplitFloats / static_cast<size_t>(out)));
		const int per = std::max(128, (rows + max_chunks - 1) / max_chunks);
		const int chunks = (rows + per - 1) / per;
		hipLaunchKernelGGL(k_colsum_part, dim3(blocks(out), chunks), dim3(256), 0, st, dy, rows, out, dy_cols, per,
		                   m.split.as<float>());
		hipLaunchKernelGGL(k_reduce_split, dim3(blocks(out)), dim3(256), 0, st, m.split.as<float>(), chunks,
		                   static_cast<size_t>(out), m.gb[l].as<float>());
		// weight gradient gW^T[out][in] = dy^T[out][rows] * x^T[in][rows]^T (split over the batch)
		uint16_t *dyt = (l == 3) ? m.dlog_t.as<uint16_t>() : m.dact_t.as<uint16_t>();
		hipLaunchKernelGGL(k_transpose, dim3((out + 31) / 32, (rows + 31) / 32), dim3(256), 0, st, dy, rows, out, dy_cols,
		                   dyt, rows);
		hipLaunchKernelGGL(k_transpose, dim3((in + 31) / 32, (rows + 31) / 32), dim3(256), 0, st, x, rows, in, in,
		                   m.act_t.as<uint16_t>(), rows);
		const int splits = 8;
		GemmArgs g{};
		g.A = dyt;
		g.B = m.act_t.as<uint16_t>();
		g.out = m.split.ptr;
		g.M = out;
		g.N = in;
		g.K = rows;
		g.lda = g.ldb = rows;
		g.ldo = in;
		g.k_split = ((rows / splits + kBK - 1) / kBK) * kBK;
		PRT_TRY(gemm<E_F32>(g, splits, st));
		const size_t nw = static_cast<size_t>(out) * in;
		hipLaunchKernelGGL(k_reduce_split, dim3(blocks(nw)), dim3(256), 0, st, m.split.as<float>(), splits, nw,
		                   m.gw[l].as<float>());
		if (l == 0) break;
		// data gradient dh[l-1] = (dy * W) masked by relu'(h[l-1])
		GemmArgs d{};
		d.A = dy;
		d.B = m.w_bf[l].as<uint16_t>();
		d.out = m.dh[l - 1].ptr;
		d.aux = m.h[l - 1].as<uint16_t>();
		d.M = rows;
		d.N = in;
		d.K = out;
		d.lda = dy_cols;
		d.ldb = out;
		d.ldo = in;
		d.k_split = d.K;
		PRT_TRY(gemm<E_MASK_BF16>(d, 1, st));
		dy = m.dh[l - 1].as<uint16_t>();
		dy_cols = in;
	}
	return hipGetLastError();
}

hipError_t mlp_adam(Mlp &m, float lr, hipStream_t st) {
	++m.step;
	const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
	const float c1 = 1.0f / (1.0f - std::pow(b1, static_cast<float>(m.step)));
	const float c2 = 1.0f / (1.0f - std::pow(b2, static_cast<float>(m.step)));
	for (int l = 0; l < 4; ++l) {
		const size_t n = static_cast<size_t>(kLayerOut[l]) * kLayerIn[l];
		hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(256), 0, st, m.w[l].as<float>(), m.gw[l].as<float>(),
		                   m.mw[l].as<float>(), m.vw[l].as<float>(), n, lr, b1, b2, eps, c1, c2, m.wt_bf[l].as<uint16_t>(),
		                   m.w_bf[l].as<uint16_t>(), kLayerOut[l], kLayerIn[l]);
		const size_t nb = l == 3 ? static_cast<size_t>(kMlpOut) : static_cast<size_t>(kLayerOut[l]);
		hipLaunchKernelGGL(k_adam, dim3(blocks(nb)), dim3(256), 0, st, m.b[l].as<float>(), m.gb[l].as<float>(),
		                   m.mb[l].as<float>(), m.vb[l].as<float>(), nb, lr, b1, b2, eps, c1, c2,
		                   static_cast<uint16_t *>(nullptr), static_cast<uint16_t *>(nullptr), 1, 1);
	}
	return hipGetLastError();
#undef PRT_TRY
}

}  // namespace prt
