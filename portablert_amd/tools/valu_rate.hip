// VALU issue-rate probe for gfx950: wave-instructions per second for independent chains of
// fp32 fma, 32-bit integer add/xor, and byte-to-float conversions with max3 (the instruction
// kinds of the traversal's node test). Used as the VALU ceiling of the traversal kernel.
//
//   portablert_amd/lib/valu_rate        (one GPU, about a second)
//
// Each workgroup is 4 waves; waves per CU = 8, 16, 32. Instructions per inner iteration are
// counted from the generated gfx950 code: 128 (fma), 256 (int), 146 (convert/max).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                         \
	do {                                                                                 \
		hipError_t e_ = (x);                                                             \
		if (e_ != hipSuccess) {                                                          \
			std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                 \
			std::exit(1);                                                                \
		}                                                                                \
	} while (0)

template <int MODE>
__global__ __launch_bounds__(256) void k_rate(float *out, int iters, float s) {
	float a0 = threadIdx.x * 1e-3f, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, a4 = a0 + 4, a5 = a0 + 5, a6 = a0 + 6,
	      a7 = a0 + 7;
	uint32_t u0 = threadIdx.x, u1 = u0 * 3, u2 = u0 * 5, u3 = u0 * 7, u4 = u0 + 11, u5 = u0 + 13, u6 = u0 + 17,
	         u7 = u0 + 19;
	for (int i = 0; i < iters; ++i) {
#pragma unroll
		for (int r = 0; r < 16; ++r) {
			if (MODE == 0) {
				a0 = __builtin_fmaf(a0, s, 1.0f); a1 = __builtin_fmaf(a1, s, 1.0f);
				a2 = __builtin_fmaf(a2, s, 1.0f); a3 = __builtin_fmaf(a3, s, 1.0f);
				a4 = __builtin_fmaf(a4, s, 1.0f); a5 = __builtin_fmaf(a5, s, 1.0f);
				a6 = __builtin_fmaf(a6, s, 1.0f); a7 = __builtin_fmaf(a7, s, 1.0f);
			} else if (MODE == 1) {
				u0 = (u0 + 0x9E3779B9u) ^ u1; u1 = (u1 + 0x7F4A7C15u) ^ u2;
				u2 = (u2 + 0x94D049BBu) ^ u3; u3 = (u3 + 0x2545F491u) ^ u0;
				u4 = (u4 + 0x9E3779B9u) ^ u5; u5 = (u5 + 0x7F4A7C15u) ^ u6;
				u6 = (u6 + 0x94D049BBu) ^ u7; u7 = (u7 + 0x2545F491u) ^ u4;
			} else {
				a0 = fmaxf(fmaxf(a0, static_cast<float>(u0 & 255u)), a1);
				a1 = fmaxf(fmaxf(a1, static_cast<float>((u0 >> 8) & 255u)), a2);
				a2 = fmaxf(fmaxf(a2, static_cast<float>((u0 >> 16) & 255u)), a3);
				a3 = fmaxf(fmaxf(a3, static_cast<float>(u0 >> 24)), a0);
				u0 += 0x01010101u;
			}
		}
	}
	// keep every chain live; the store never happens for these inputs
	if (a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7 == 12345.0f || (u0 ^ u1 ^ u2 ^ u3 ^ u4 ^ u5 ^ u6 ^ u7) == 7u) out[0] = a0;
}

template <int MODE>
static double wave_inst_per_s(float *out, int blocks, int iters, int per_iter) {
	hipEvent_t e0, e1;
	CHECK(hipEventCreate(&e0));
	CHECK(hipEventCreate(&e1));
	hipLaunchKernelGGL(k_rate<MODE>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f);  // warm-up
	CHECK(hipGetLastError());
	CHECK(hipEventRecord(e0));
	hipLaunchKernelGGL(k_rate<MODE>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f);
	CHECK(hipEventRecord(e1));
	CHECK(hipEventSynchronize(e1));
	float ms = 0.0f;
	CHECK(hipEventElapsedTime(&ms, e0, e1));
	CHECK(hipEventDestroy(e0));
	CHECK(hipEventDestroy(e1));
	return 4.0 * blocks * static_cast<double>(iters) * per_iter / (ms * 1e-3);
}

int main() {
	float *out = nullptr;
	CHECK(hipMalloc(&out, sizeof(float)));
	int cus = 0;
	CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
	const int iters = 4096;
	std::printf("CUs %d\n", cus);
	for (int wpc : {8, 16, 32}) {
		const int blocks = cus * wpc / 4;
		std::printf("waves/CU %2d: fp32 fma %.1f G/s, int32 add/xor %.1f G/s, byte convert + max3 %.1f G/s\n", wpc,
		            wave_inst_per_s<0>(out, blocks, iters, 128) * 1e-9, wave_inst_per_s<1>(out, blocks, iters, 256) * 1e-9,
		            wave_inst_per_s<2>(out, blocks, iters, 146) * 1e-9);
	}
	CHECK(hipFree(out));
	return 0;
}
