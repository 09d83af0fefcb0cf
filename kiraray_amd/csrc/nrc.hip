// nrc.hip - neural radiance cache on gfx950 matrix cores (krr_nrc.h).
//
// Network: x0 = enc(vertex) (64), z_l = W_l^T x_l, x_{l+1} = relu(z_l) for
// l = 0..4, y = W_5^T x_5 (3 outputs, no activation, no biases; the encoding
// carries two constant 1 inputs).  W_l is stored [in][out].
//
// MFMA formulation.  Each wave owns a tile of 32 samples and computes the
// transposed products Z^T = W^T X^T in 8-deep k-steps: the A operand is a
// 32x8 slice of W^T (LDS), the B operand an 8x32 slice of X^T held by the
// lanes (lane m + 32g holds features 8s + 4g + j, j < 4, of sample m for
// k-step s), and the 32x32 accumulator tile leaves feature 8(r/4) + 4g + r%4
// of sample m in register r of lane m + 32g - exactly the B-operand layout of
// the next layer, so activations never leave the registers between layers.
// Pairs of k-steps run as one v_mfma_f32_32x32x16_bf16.  Backward uses the
// same layout with A = W (dX^T = W dZ^T).  Weight gradients
// dW_l = X_l dZ_l^T are a split-K GEMM over the batch on X^T and dZ^T stored
// feature-major in HBM, reduced in a fixed order (deterministic), then Adam
// updates the FP32 master weights and rewrites both BF16 copies.
//
// Encoding (64 features, lane half g = 0/1 computes slots q = 4s + j of
// feature 8s + 4g + j):
//   half 0: q = 10c + 2f + t: sin (t = 0) / cos (t = 1) of 2^f pi pos_c,
//           c < 3, f < 5 (hardware sine/cosine of pi pos_c, doubled f
//           times); q = 30, 31: 1.
//   half 1: q 0..2 pos; 3..6 / 7..10 one-blob(4) of dir theta / phi;
//           11..14 / 15..18 of normal theta / phi; 19..22 of roughness;
//           23..25 diffuse; 26..28 specular; 29..31: 0.
//   one-blob(4) of x in [0,1]: exp(-0.5 (4 (x - (i + 0.5) / 4))^2), i < 4.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "detmath.h"
#include "errors.h"
#include "json.h"
#include "krr_nrc.h"
#include "nrc_dev.h"

namespace krr {
namespace nrc {

// ---------------------------------------------------------------- device

// Round to nearest even (inf stays inf, a nan becomes a quiet nan): the
// hardware conversion, v_cvt_pk_bf16_f32 on gfx950, two values per
// instruction where the compiler pairs them.
__device__ __forceinline__ uint16_t toBf16(float f) { return __builtin_bit_cast(uint16_t, (__bf16) f); }
__device__ __forceinline__ float fromBf16(uint16_t h) { return __uint_as_float((uint32_t) h << 16); }

__device__ __forceinline__ short4v pack4(float a, float b, float c, float d) {
	short4v v;
	v[0] = (short) toBf16(a);
	v[1] = (short) toBf16(b);
	v[2] = (short) toBf16(c);
	v[3] = (short) toBf16(d);
	return v;
}

// The encoding uses the hardware exp2 / sine / cosine (a learned input needs
// no bit-exact agreement with the CPU; the checker compares to 2e-5).
__device__ __forceinline__ float oneBlob(float x, int i) {
	const float d = 4.0f * (x - ((float) i + 0.5f) * 0.25f);
	return __builtin_amdgcn_exp2f(-0.72134752044448170368f * d * d); // exp(-d^2 / 2)
}

// Encoding slot q (0..31) of lane half g for vertex record v (16 floats).
__device__ float encodeSlot(const float *v, int g, int q) {
	if (g == 0) {
		if (q >= 30) return 1.0f;
		const int c = q / 10, f = (q % 10) / 2, t = q & 1;
		float sn = __builtin_amdgcn_sinf(0.5f * v[c]), cs = __builtin_amdgcn_cosf(0.5f * v[c]);
		for (int k = 0; k < f; k++) {
			const float s2 = 2.0f * sn * cs, c2 = (cs - sn) * (cs + sn);
			sn = s2;
			cs = c2;
		}
		return t ? cs : sn;
	}
	if (q < 3) return v[q];						  // pos
	if (q < 7) return oneBlob(v[3], q - 3);		  // dir theta
	if (q < 11) return oneBlob(v[4], q - 7);	  // dir phi
	if (q < 15) return oneBlob(v[5], q - 11);	  // normal theta
	if (q < 19) return oneBlob(v[6], q - 15);	  // normal phi
	if (q < 23) return oneBlob(v[7], q - 19);	  // roughness
	if (q < 29) return v[8 + (q - 23)];			  // diffuse, specular
	return 0.0f;
}

// B operands of layer 0 for sample record v: b[s] = features 8s + 4g + j
// (slots q = 4s + j of half g).
__device__ void encodeB(const float *v, int g, short4v b[8]) {
	float e[32];
	if (g == 0) {
		// sin / cos of pi x from the hardware (argument in revolutions: x / 2),
		// higher frequencies by the double-angle recurrence
#pragma unroll
		for (int c = 0; c < 3; c++) {
			float sn = __builtin_amdgcn_sinf(0.5f * v[c]), cs = __builtin_amdgcn_cosf(0.5f * v[c]);
#pragma unroll
			for (int f = 0; f < 5; f++) {
				e[10 * c + 2 * f]	  = sn;
				e[10 * c + 2 * f + 1] = cs;
				const float s2 = 2.0f * sn * cs, c2 = (cs - sn) * (cs + sn);
				sn = s2;
				cs = c2;
			}
		}
		e[30] = e[31] = 1.0f;
	} else {
#pragma unroll
		for (int q = 0; q < 32; q++) e[q] = encodeSlot(v, 1, q);
	}
#pragma unroll
	for (int s = 0; s < 8; s++) b[s] = pack4(e[4 * s], e[4 * s + 1], e[4 * s + 2], e[4 * s + 3]);
}

__device__ __forceinline__ float16v zero16() {
	float16v z;
#pragma unroll
	for (int i = 0; i < 16; i++) z[i] = 0.f;
	return z;
}

// Two consecutive 8-deep K-steps (4 bf16 per lane half each) in one gfx950
// v_mfma_f32_32x32x16_bf16: lane half g passes the A row elements (a0, a1)
// and the B column elements (b0, b1) in the same order, so the 16 products
// summed are those of the two 32x32x8 steps (a0, b0) and (a1, b1); only the
// FP32 order of the additions inside the instruction differs.  Half the
// instructions of the 32x32x8 form, each at the same rate.
typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));
typedef short short8v __attribute__((ext_vector_type(8)));
__device__ __forceinline__ float16v mfma2(short4v a0, short4v a1, short4v b0, short4v b1, float16v c) {
	const short8v a = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
	const short8v b = __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7);
	return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8v, a), __builtin_bit_cast(bf16x8v, b), c, 0, 0, 0);
}

// A operand: 4 consecutive bf16 at row `row` (rows `stride` bf16 apart), column 8s + 4g.
__device__ __forceinline__ short4v ldsA(const uint16_t *m, int row, int s, int g, int stride = kRowH) {
	return *reinterpret_cast<const short4v *>(m + row * stride + 8 * s + 4 * g);
}

// One hidden layer, forward: out (B layout, relu, bf16) = W^T x.
__device__ __forceinline__ void hiddenForward(const uint16_t *wt, int lane, int g, const short4v x[8], short4v out[8], float16v &acc0,
											  float16v &acc1) {
	acc0 = zero16();
	acc1 = zero16();
	const int r0 = lane & 31;
#pragma unroll
	for (int s = 0; s < 8; s += 2) {
		acc0 = mfma2(ldsA(wt, r0, s, g), ldsA(wt, r0, s + 1, g), x[s], x[s + 1], acc0);
		acc1 = mfma2(ldsA(wt, 32 + r0, s, g), ldsA(wt, 32 + r0, s + 1, g), x[s], x[s + 1], acc1);
	}
#pragma unroll
	for (int s = 0; s < 4; s++) {
		out[s]	   = pack4(fmaxf(acc0[4 * s], 0.f), fmaxf(acc0[4 * s + 1], 0.f), fmaxf(acc0[4 * s + 2], 0.f), fmaxf(acc0[4 * s + 3], 0.f));
		out[4 + s] = pack4(fmaxf(acc1[4 * s], 0.f), fmaxf(acc1[4 * s + 1], 0.f), fmaxf(acc1[4 * s + 2], 0.f), fmaxf(acc1[4 * s + 3], 0.f));
	}
}

// Output layer forward: y_c (c < 3) lands in acc[c] of lanes with g == 0.
__device__ __forceinline__ float16v outputForward(const uint16_t *wt5, int lane, int g, const short4v x[8]) {
	float16v acc = zero16();
	const int r0 = lane & 31;
#pragma unroll
	for (int s = 0; s < 8; s += 2) acc = mfma2(ldsA(wt5, r0, s, g), ldsA(wt5, r0, s + 1, g), x[s], x[s + 1], acc);
	return acc;
}

// Copies `rows` dense rows of rowLen bf16 into LDS rows of rowLen + kPad.
// Each thread issues up to twelve 16-byte loads before it stores them (two
// 8-byte LDS writes each: padded rows are 8-byte aligned), so a block of 256
// threads fetches one copy of all layers' weights (~48 KB) in one round trip.  One
// load per store would put a memory round trip between consecutive loads,
// and the training kernel runs one wave per SIMD with nothing to hide it.
__device__ void loadRows(uint16_t *lds, const uint16_t *src, int rows, int rowLen) {
	constexpr int kBatch = 12;
	const int per = rowLen / 8; // 16-byte pieces per row
	const int total = rows * per, stride = (int) blockDim.x;
	for (int base = (int) threadIdx.x; base < total; base += kBatch * stride) {
		uint4 v[kBatch];
#pragma unroll
		for (int k = 0; k < kBatch; k++) {
			const int i = base + k * stride;
			if (i < total) {
				const int r = i / per, c = i - r * per;
				v[k]		= *reinterpret_cast<const uint4 *>(src + r * rowLen + 8 * c);
			}
		}
#pragma unroll
		for (int k = 0; k < kBatch; k++) {
			const int i = base + k * stride;
			if (i < total) {
				const int r = i / per, c = i - r * per;
				uint2 *d = reinterpret_cast<uint2 *>(lds + r * (rowLen + kPad) + 8 * c);
				d[0]	 = make_uint2(v[k].x, v[k].y);
				d[1]	 = make_uint2(v[k].z, v[k].w);
			}
		}
	}
}

// Forward weights (W^T of every layer, output rows padded to 32) into LDS.
__device__ void loadWeights(uint16_t *lds, const uint16_t *wtAll) { loadRows(lds, wtAll, kHidden * 64 + 32, 64); }

// Backward weights (W of every layer, output columns padded to 32) into LDS.
__device__ void loadWeightsBwd(uint16_t *lds, const uint16_t *wbAll) {
	loadRows(lds, wbAll, kHidden * 64, 64);
	loadRows(lds + kHidden * 64 * kRowH, wbAll + kHidden * 4096, 64, 32);
}

// ----------------------------------------------------------------- kernels

__global__ void __launch_bounds__(256) k_encode(const float *vertices, int64_t n, float *out) {
	const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n * 64) return;
	const int64_t row = i / 64;
	const int k		  = (int) (i % 64);
	const int s = k / 8, g = (k / 4) & 1, j = k & 3;
	out[i] = encodeSlot(vertices + row * 16, g, 4 * s + j);
}

// Inference: each wave evaluates tiles of 32 samples.
__global__ void __launch_bounds__(256) k_infer(const uint16_t *wtAll, const float *vertices, int64_t n, float *radiance) {
	extern __shared__ uint16_t ldsW[];
	loadWeights(ldsW, wtAll);
	__syncthreads();
	const int lane = threadIdx.x & 63, g = lane >> 5, m = lane & 31;
	const int wavesPerBlock = blockDim.x >> 6;
	const int64_t tiles		= (n + 31) / 32;
	for (int64_t t = (int64_t) blockIdx.x * wavesPerBlock + (threadIdx.x >> 6); t < tiles; t += (int64_t) gridDim.x * wavesPerBlock) {
		const int64_t i = t * 32 + m;
		float v[16];
		if (i < n) {
			const float4 *p = reinterpret_cast<const float4 *>(vertices + i * 16);
#pragma unroll
			for (int q = 0; q < 4; q++) {
				const float4 f = p[q];
				v[4 * q]	   = f.x;
				v[4 * q + 1]   = f.y;
				v[4 * q + 2]   = f.z;
				v[4 * q + 3]   = f.w;
			}
		} else {
#pragma unroll
			for (int q = 0; q < 16; q++) v[q] = 0.f;
		}
		short4v x[8], y[8];
		encodeB(v, g, x);
		float16v a0, a1;
		// not unrolled: an unrolled loop hoists every layer's weight loads
#pragma unroll 1
		for (int l = 0; l < kHidden; l++) {
			hiddenForward(ldsW + l * 64 * kRowH, lane, g, x, y, a0, a1);
#pragma unroll
			for (int s = 0; s < 8; s++) x[s] = y[s];
		}
		const float16v o = outputForward(ldsW + kHidden * 64 * kRowH, lane, g, x);
		if (g == 0 && i < n) {
			radiance[3 * i]		= o[0];
			radiance[3 * i + 1] = o[1];
			radiance[3 * i + 2] = o[2];
		}
	}
}

// Threads per k_infer_queue block: the block's waves share one LDS copy of
// the weights (52 KB), which limits a CU to 3 copies; blocks of 8 waves,
// compiled for 5 waves per SIMD (96 VGPRs, 20 B of spills), run 5 waves per SIMD
// where blocks of 4 ran 3.  On the 1080p field (0.36 M queries per launch):
// 27.8 us per launch against 32.4 us; 6 waves per SIMD (80 VGPRs, 84 B of
// spills) 33.6 us, 512 threads without a wave target (97 VGPRs, 4 waves)
// 29.4 us.
constexpr int kInferBlock = 512;

// Inference on a query queue whose length lives on the device; adds
// thp * max(y, 0) to L[pixel].  A wave's next tile of queries is loaded while
// it evaluates the current one, and the pixel's radiance is read as soon as
// the query names the pixel: the network in between hides both loads.  A
// launch answers one sample's queries, at most one per pixel, so no other
// lane writes that pixel's radiance meanwhile.
__device__ __forceinline__ void loadQuery(const float4 *queries, int stride, int i, int n, float v[16], float4 &pay) {
	pay = make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
	if (i < n) {
		const float4 *p = queries + (int64_t) i * stride;
#pragma unroll
		for (int q = 0; q < 4; q++) {
			const float4 f = p[q];
			v[4 * q]	   = f.x;
			v[4 * q + 1]   = f.y;
			v[4 * q + 2]   = f.z;
			v[4 * q + 3]   = f.w;
		}
		pay = p[4];
	} else {
#pragma unroll
		for (int q = 0; q < 16; q++) v[q] = 0.f;
	}
}

__global__ void __launch_bounds__(kInferBlock) __attribute__((amdgpu_waves_per_eu(5, 8))) k_infer_queue(const uint16_t *wtAll, const float4 *queries, int stride, const int *countPtr, int capacity,
													 float4 *L) {
	extern __shared__ uint16_t ldsW[];
	loadWeights(ldsW, wtAll);
	__syncthreads();
	const int n	   = min(max(*countPtr, 0), capacity);
	const int lane = threadIdx.x & 63, g = lane >> 5, m = lane & 31;
	const int wavesPerBlock = blockDim.x >> 6;
	const int tiles			= (n + 31) / 32;
	const int tStride		= gridDim.x * wavesPerBlock;
	int t					= blockIdx.x * wavesPerBlock + (threadIdx.x >> 6);
	float v[16];
	float4 pay;
	loadQuery(queries, stride, t * 32 + m, n, v, pay);
	for (; t < tiles; t += tStride) {
		const int i		= t * 32 + m;
		const int pixel = __float_as_int(pay.w);
		const bool mine = g == 0 && i < n;
		const float4 thp = pay;
		float4 acc		 = make_float4(0.f, 0.f, 0.f, 0.f);
		if (mine) acc = L[pixel];
		short4v x[8], y[8];
		encodeB(v, g, x);
		loadQuery(queries, stride, (t + tStride) * 32 + m, n, v, pay); // the next tile's queries
		float16v a0, a1;
		// not unrolled: an unrolled loop hoists every layer's weight loads
#pragma unroll 1
		for (int l = 0; l < kHidden; l++) {
			hiddenForward(ldsW + l * 64 * kRowH, lane, g, x, y, a0, a1);
#pragma unroll
			for (int s = 0; s < 8; s++) x[s] = y[s];
		}
		const float16v o = outputForward(ldsW + kHidden * 64 * kRowH, lane, g, x);
		if (mine) {
			acc.x += thp.x * fmaxf(o[0], 0.f);
			acc.y += thp.y * fmaxf(o[1], 0.f);
			acc.z += thp.z * fmaxf(o[2], 0.f);
			L[pixel] = acc;
		}
	}
}

// Writes the B-layout bf16 operand (features 8s + 4g + j of sample m) to a
// feature-major [feature][ld] array at column `col` (`dst` is a layer's
// 64 x ld block).  Buffer stores: the lane's column offset is computed once
// and the feature row's offset, (8s + j) ld, is wave-uniform (soffset), so a
// store takes no per-lane address arithmetic; stores past the block are
// dropped by the buffer's range check.  The rows of features >= `features`
// (s >= features / 8, for every g) are not written.
__device__ __forceinline__ void storeFeatureMajor(uint16_t *dst, int64_t ld, int64_t col, int g, const short4v x[8], int features = 64) {
	// ld <= 2^24 (max_batch): the largest row offset, 63 ld bf16, stays below
	// 2^31 bytes; the range, 64 ld bf16 <= 2^31 bytes, is read as unsigned
	const int ld32 = (int) ld;
	const uint32_t range = 64u * (uint32_t) ld32 * (uint32_t) sizeof(uint16_t);
	const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(dst, (short) 0, (int) range, 0x00020000);
	const int voff = (4 * g * ld32 + (int) col) * (int) sizeof(uint16_t);
#pragma unroll
	for (int s = 0; s < 8; s++)
#pragma unroll
		for (int j = 0; j < 4; j++)
			if (8 * s < features) __builtin_amdgcn_raw_buffer_store_b16((unsigned short) x[s][j], rsrc, voff, (8 * s + j) * ld32 * (int) sizeof(uint16_t), 0);
}

// Forward + backward of one tile of 32 samples per wave.  Writes X_l^T and
// dZ_l^T (bf16, feature-major, leading dimension ld = padded batch) and the
// tile's loss sum.  Samples i >= n contribute nothing.
template <bool BWD_LDS>
__global__ void __launch_bounds__(256) k_train_tile(const uint16_t *wtAll, const uint16_t *wAll, const float *vertices, const float *targets,
													int64_t nHost, const int *nDev, int64_t capTotal, int step, int steps, int64_t nTiles,
													int64_t ld, float lossEps, uint16_t *xT, uint16_t *dzT, float *tileLoss, int *adamStep) {
	// a training step (not a bare gradient) counts Adam's step here: the
	// kernels after this one in the stream read the new count
	if (adamStep && blockIdx.x == 0 && threadIdx.x == 0) *adamStep += 1;
	// samples: nHost, or with a device count, the step-th of `steps` equal
	// slices of the first min(*nDev, capTotal) samples
	int64_t n = nHost;
	if (nDev) {
		const int64_t cnt = min((int64_t) max(*nDev, 0), capTotal);
		const int64_t off = (int64_t) step * cnt / steps;
		n				  = min((int64_t) (step + 1) * cnt / steps - off, nHost);
		vertices += off * 16;
		targets += off * 3;
	}
	// A operands from LDS; with BWD_LDS false (a device whose blocks cannot
	// hold both copies) the backward ones come straight from global memory
	// (45 KB, cache resident), unpadded
	extern __shared__ uint16_t lds[];
	uint16_t *ldsWt = lds;
	loadWeights(ldsWt, wtAll);
	const uint16_t *ldsW = wAll;
	constexpr int sH = BWD_LDS ? kRowH : 64, sO = BWD_LDS ? kRowO : 32;
	if (BWD_LDS) {
		loadWeightsBwd(lds + kFwdLdsElems, wAll);
		ldsW = lds + kFwdLdsElems;
	}
	__syncthreads();
	const int lane = threadIdx.x & 63, g = lane >> 5, m = lane & 31;
	const int wavesPerBlock = blockDim.x >> 6;
	const float invCount	= n > 0 ? 1.0f / (3.0f * (float) n) : 0.f;
	for (int64_t t = (int64_t) blockIdx.x * wavesPerBlock + (threadIdx.x >> 6); t < nTiles; t += (int64_t) gridDim.x * wavesPerBlock) {
		const int64_t i	  = t * 32 + m;
		const bool valid = i < n;
		float v[16];
		float tg[3] = {0.f, 0.f, 0.f}; // loaded with the inputs: the loss needs them after the forward pass
		if (valid && g == 0) {
#pragma unroll
			for (int c = 0; c < 3; c++) tg[c] = targets[3 * i + c];
		}
		if (valid) {
			const float4 *p = reinterpret_cast<const float4 *>(vertices + i * 16);
#pragma unroll
			for (int q = 0; q < 4; q++) {
				const float4 f = p[q];
				v[4 * q]	   = f.x;
				v[4 * q + 1]   = f.y;
				v[4 * q + 2]   = f.z;
				v[4 * q + 3]   = f.w;
			}
		} else {
#pragma unroll
			for (int q = 0; q < 16; q++) v[q] = 0.f;
		}
		// forward: each layer input goes to X^T as it is produced; the
		// backward pass needs only its ReLU mask, one bit per held feature
		// (bit 4s + j of mask[l] <=> feature 8s + 4g + j of x_l is > 0)
		short4v x[8], y[8];
		uint32_t mask[kHidden + 1];
		encodeB(v, g, x);
		float16v a0, a1;
#pragma unroll 1
		for (int l = 0; l < kHidden; l++) {
			storeFeatureMajor(xT + (int64_t) l * 64 * ld, ld, i, g, x);
			hiddenForward(ldsWt + l * 64 * kRowH, lane, g, x, y, a0, a1);
			uint32_t mk = 0;
#pragma unroll
			for (int s = 0; s < 8; s++) {
				x[s] = y[s];
#pragma unroll
				for (int j = 0; j < 4; j++) mk |= (fromBf16((uint16_t) y[s][j]) > 0.f ? 1u : 0u) << (4 * s + j);
			}
			mask[l + 1] = mk;
		}
		storeFeatureMajor(xT + (int64_t) kHidden * 64 * ld, ld, i, g, x);
		const float16v o = outputForward(ldsWt + kHidden * 64 * kRowH, lane, g, x);
		// relative L2 loss and its gradient (lanes g == 0 hold the 3 outputs)
		float d[3]	 = {0.f, 0.f, 0.f};
		float lsum	 = 0.f;
		if (g == 0 && valid) {
#pragma unroll
			for (int c = 0; c < 3; c++) {
				const float yv = o[c], r = yv - tg[c];
				const float den = yv * yv + lossEps;
				lsum += r * r / den;
				d[c] = 2.0f * r / den * invCount;
			}
		}
		// wave sum of the loss
		for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
		if (lane == 0) tileLoss[t] = lsum * invCount;
		// dZ_5 as a B operand: outputs n = 8s + 4g + j, only n < 3 nonzero
		short4v dz[8];
#pragma unroll
		for (int s = 0; s < 8; s++) dz[s] = pack4(0.f, 0.f, 0.f, 0.f);
		if (g == 0) dz[0] = pack4(d[0], d[1], d[2], 0.f);
		storeFeatureMajor(dzT + (int64_t) kHidden * 64 * ld, ld, i, g, dz, 32); // 32 padded output rows
		// backward through the output layer: dX_5^T = W_5 dZ_5^T (K = 32 padded outputs)
		const int r0 = lane & 31;
		{
			const uint16_t *w5 = ldsW + kHidden * 64 * sH;
			a0 = zero16();
			a1 = zero16();
#pragma unroll
			for (int s = 0; s < 4; s += 2) {
				a0 = mfma2(ldsA(w5, r0, s, g, sO), ldsA(w5, r0, s + 1, g, sO), dz[s], dz[s + 1], a0);
				a1 = mfma2(ldsA(w5, 32 + r0, s, g, sO), ldsA(w5, 32 + r0, s + 1, g, sO), dz[s], dz[s + 1], a1);
			}
		}
#pragma unroll 1
		for (int l = kHidden - 1; l >= 0; l--) {
			// dZ_l = dX_{l+1} * (x_{l+1} > 0), as the next B operand
			const uint32_t mk = mask[l + 1];
#pragma unroll
			for (int s = 0; s < 4; s++)
#pragma unroll
				for (int j = 0; j < 4; j++) {
					dz[s][j]	 = (short) toBf16((mk >> (4 * s + j)) & 1u ? a0[4 * s + j] : 0.f);
					dz[4 + s][j] = (short) toBf16((mk >> (16 + 4 * s + j)) & 1u ? a1[4 * s + j] : 0.f);
				}
			storeFeatureMajor(dzT + (int64_t) l * 64 * ld, ld, i, g, dz);
			if (l > 0) {
				const uint16_t *w = ldsW + l * 64 * sH;
				a0				  = zero16();
				a1				  = zero16();
#pragma unroll
				for (int s = 0; s < 8; s += 2) {
					a0 = mfma2(ldsA(w, r0, s, g, sH), ldsA(w, r0, s + 1, g, sH), dz[s], dz[s + 1], a0);
					a1 = mfma2(ldsA(w, 32 + r0, s, g, sH), ldsA(w, 32 + r0, s + 1, g, sH), dz[s], dz[s + 1], a1);
				}
			}
		}
	}
}

// Split-K weight gradients of all layers in one launch: blockIdx.z = layer,
// blockIdx.y = chunk of `chunk` samples; partial[layer][chunk][k][nOut] =
// sum over the chunk of X_l^T[k][m] dZ_l^T[n][m].  One wave per 32x32
// output tile (4 for hidden layers, 2 for the 64x3 output layer).
__global__ void __launch_bounds__(256) k_weight_grad(const uint16_t *xTAll, const uint16_t *dzTAll, int64_t ld, int chunk, int chunks,
													 float *partialAll) {
	const int layer = blockIdx.z;
	const int nOut = layer < kHidden ? 64 : 3, nOutTiles = layer < kHidden ? 2 : 1;
	const int lane = threadIdx.x & 63, g = lane >> 5, r0 = lane & 31;
	const int tile = threadIdx.x >> 6;
	if (tile >= 2 * nOutTiles) return;
	const int kt = tile / nOutTiles, nt = tile % nOutTiles;
	const int64_t m0	 = (int64_t) blockIdx.y * chunk;
	const uint16_t *xT	 = xTAll + (int64_t) layer * 64 * ld;
	const uint16_t *dzT	 = dzTAll + (int64_t) layer * 64 * ld;
	float16v acc		 = zero16();
	const uint16_t *arow = xT + (int64_t) (32 * kt + r0) * ld + m0;
	const uint16_t *brow = dzT + (int64_t) (32 * nt + r0) * ld + m0;
	// eight k-steps of loads in flight before their MFMAs (same order of
	// accumulation); chunk is a multiple of 64 (checked by the caller)
	constexpr int kU = 8;
	for (int s0 = 0; s0 < chunk / 8; s0 += kU) {
		short4v a[kU], b[kU];
#pragma unroll
		for (int u = 0; u < kU; u++) {
			a[u] = *reinterpret_cast<const short4v *>(arow + 8 * (s0 + u) + 4 * g);
			b[u] = *reinterpret_cast<const short4v *>(brow + 8 * (s0 + u) + 4 * g);
		}
#pragma unroll
		for (int u = 0; u < kU; u += 2) acc = mfma2(a[u], a[u + 1], b[u], b[u + 1], acc);
	}
	// D[i = k][j = n]: lane holds column n = r0, rows k = 8(r/4) + 4g + r%4
	float *dst = partialAll + (int64_t) layer * chunks * 4096 + (int64_t) blockIdx.y * 64 * nOut;
#pragma unroll
	for (int r = 0; r < 16; r++) {
		const int k = 32 * kt + 8 * (r / 4) + 4 * g + (r % 4);
		const int n = 32 * nt + r0;
		if (n < nOut) dst[k * nOut + n] = acc[r];
	}
}

// Sum over chunks (fixed order) of the layer's partials of weight p
// (layer-major like the master weights).
__device__ __forceinline__ float chunkSum(const float *partialAll, int chunks, int p) {
	const int layer = min(p / 4096, kHidden);
	const int q		= p - layer * 4096;
	const int count = layer < kHidden ? 4096 : 64 * 3;
	const float *src = partialAll + (int64_t) layer * chunks * 4096 + q;
	float s			 = 0.f;
	for (int c = 0; c < chunks; c++) s += src[(int64_t) c * count];
	return s;
}

__global__ void k_reduce_chunks(const float *partialAll, int chunks, float *grad) {
	const int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= KRR_NRC_PARAMS) return;
	grad[p] = chunkSum(partialAll, chunks, p);
}

// Deterministic loss total over tiles (one block of 256 threads).
__device__ void reduceLoss(const float *tileLoss, int64_t nTiles, float *out) {
	__shared__ float part[256];
	float s = 0.f;
	for (int64_t t = threadIdx.x; t < nTiles; t += blockDim.x) s += tileLoss[t];
	part[threadIdx.x] = s;
	__syncthreads();
	for (int w = blockDim.x / 2; w > 0; w >>= 1) {
		if ((int) threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0) *out = part[0];
}

__global__ void k_reduce_loss(const float *tileLoss, int64_t nTiles, float *out) { reduceLoss(tileLoss, nTiles, out); }

// Rewrites both bf16 copies from the fp32 master weights.
__device__ void writeCopies(const float *w, uint16_t *wt, uint16_t *wb, int p) {
	// p indexes the master array: layer l < 5 at 4096 l + 64 k + n; layer 5 at 20480 + 3 k + n.
	if (p < kHidden * 4096) {
		const int l = p / 4096, k = (p / 64) % 64, n = p % 64;
		wt[l * 4096 + n * 64 + k] = toBf16(w[p]);
		wb[l * 4096 + k * 64 + n] = toBf16(w[p]);
	} else {
		const int q = p - kHidden * 4096, k = q / 3, n = q % 3;
		wt[kHidden * 4096 + n * 64 + k] = toBf16(w[p]);
		wb[kHidden * 4096 + k * 32 + n] = toBf16(w[p]);
	}
}

__global__ void k_write_copies(const float *w, uint16_t *wt, uint16_t *wb) {
	const int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p < KRR_NRC_PARAMS) writeCopies(w, wt, wb, p);
}

// Adam (Kingma & Ba 2015) with bias corrections c1 = 1 - beta1^t, c2 = 1 - beta2^t.
// Adam's step count lives in device memory, so that training steps recorded
// in a graph count up on every replay.
__device__ __forceinline__ void adamUpdate(float *w, float *mo, float *ve, float gr, float lr, float b1, float b2, float eps, float t,
										   uint16_t *wt, uint16_t *wb, int p) {
	const float c1 = 1.f - powf(b1, t), c2 = 1.f - powf(b2, t); // bias corrections
	const float m  = b1 * mo[p] + (1.f - b1) * gr;
	const float v  = b2 * ve[p] + (1.f - b2) * gr * gr;
	mo[p]		   = m;
	ve[p]		   = v;
	w[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
	writeCopies(w, wt, wb, p);
}

// One training step's tail in one launch: per weight the fixed-order chunk
// sum (the gradient, also stored), the Adam update and the bf16 copies; the
// last block also totals the loss.  The step count was raised by
// k_train_tile.
//
// A block of 256 threads takes 32 weights: thread (group q, weight j) sums
// the chunks of its eighth of the chunk range in order, and the eight group
// sums are added in group order (a fixed order: deterministic).  One thread
// per weight would leave most of the chip idle on a few hundred
// latency-bound sequential loads.
constexpr int kReduceWeights = 32, kReduceGroups = 8;
__global__ void __launch_bounds__(256) k_reduce_adam(const float *partialAll, int chunks, float *grad, float *w, float *mo, float *ve,
													 float lr, float b1, float b2, float eps, const int *tDev, uint16_t *wt, uint16_t *wb,
													 const float *tileLoss, int64_t nTiles, float *loss) {
	__shared__ float part[kReduceGroups][kReduceWeights];
	const int j = threadIdx.x % kReduceWeights, q = threadIdx.x / kReduceWeights;
	const int p = blockIdx.x * kReduceWeights + j;
	float s		= 0.f;
	if (p < KRR_NRC_PARAMS) {
		const int layer	 = min(p / 4096, kHidden);
		const int count	 = layer < kHidden ? 4096 : 64 * 3;
		const float *src = partialAll + (int64_t) layer * chunks * 4096 + (p - layer * 4096);
		const int c0 = q * chunks / kReduceGroups, c1 = (q + 1) * chunks / kReduceGroups;
		for (int c = c0; c < c1; c++) s += src[(int64_t) c * count];
	}
	part[q][j] = s;
	__syncthreads();
	if (q == 0 && p < KRR_NRC_PARAMS) {
		float gr = part[0][j];
		for (int k = 1; k < kReduceGroups; k++) gr += part[k][j];
		grad[p] = gr;
		adamUpdate(w, mo, ve, gr, lr, b1, b2, eps, (float) *tDev, wt, wb, p);
	}
	if (blockIdx.x == gridDim.x - 1) reduceLoss(tileLoss, nTiles, loss);
}

} // namespace nrc
} // namespace krr

// ------------------------------------------------------------------- host

using namespace krr::nrc;

namespace {

struct Buf {
	void *p	  = nullptr;
	size_t sz = 0;
	~Buf() {
		if (p) (void) hipFree(p);
	}
	hipError_t alloc(size_t bytes) {
		if (p) (void) hipFree(p);
		p  = nullptr;
		sz = 0;
		hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
		if (e == hipSuccess) sz = bytes;
		return e;
	}
	template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

#define NRC_HIP(expr)                                                                                                  \
	do {                                                                                                               \
		hipError_t e_ = (expr);                                                                                        \
		if (e_ != hipSuccess) return krr::fail(KRR_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));                  \
	} while (0)

constexpr int kChunk	= 1024; // batch granularity of the training buffers
constexpr int kWChunk = 256;	// samples per split-K chunk of the weight gradient
static_assert(kWChunk % 64 == 0, "k_weight_grad loads eight 8-sample k-steps at a time");

} // namespace

struct krr_nrc {
	float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.99f, eps = 1e-8f, lossEps = 0.01f;
	uint64_t seed	= 1;
	int64_t maxBatch = 65536, ld = 0;
	bool bwdLds		 = false; // the device's blocks hold both padded weight copies in LDS
	int inferBlocks	 = 1024;  // k_infer_queue blocks resident at once on the device (each loads the weights once)
	Buf w, m, v, grad, wt, wb, xT, dzT, partial, tileLoss, loss, step;
};

namespace {

int uploadWeights(krr_nrc *h, const std::vector<float> &host) {
	NRC_HIP(hipMemcpy(h->w.p, host.data(), sizeof(float) * KRR_NRC_PARAMS, hipMemcpyHostToDevice));
	NRC_HIP(hipMemset(h->m.p, 0, sizeof(float) * KRR_NRC_PARAMS));
	NRC_HIP(hipMemset(h->v.p, 0, sizeof(float) * KRR_NRC_PARAMS));
	NRC_HIP(hipMemset(h->wt.p, 0, h->wt.sz));
	NRC_HIP(hipMemset(h->wb.p, 0, h->wb.sz));
	hipLaunchKernelGGL(k_write_copies, dim3((KRR_NRC_PARAMS + 255) / 256), dim3(256), 0, 0, h->w.as<float>(), h->wt.as<uint16_t>(),
					   h->wb.as<uint16_t>());
	NRC_HIP(hipGetLastError());
	NRC_HIP(hipDeviceSynchronize());
	NRC_HIP(hipMemset(h->step.p, 0, sizeof(int)));
	NRC_HIP(hipDeviceSynchronize());
	return KRR_OK;
}

// Seeded Xavier-uniform initial weights (PCG32 stream of the seed).
std::vector<float> initialWeights(uint64_t seed) {
	std::vector<float> w(KRR_NRC_PARAMS);
	uint64_t state = 0, inc = (seed << 1u) | 1u;
	auto next = [&]() {
		uint64_t old = state;
		state		 = old * 6364136223846793005ULL + inc;
		uint32_t xs	 = (uint32_t) (((old >> 18u) ^ old) >> 27u);
		uint32_t rot = (uint32_t) (old >> 59u);
		return (xs >> rot) | (xs << ((-rot) & 31));
	};
	next();
	state += 0x853c49e6748fea9bULL;
	next();
	for (int p = 0; p < KRR_NRC_PARAMS; p++) {
		const int fanOut = p < kHidden * 4096 ? 64 : 3;
		const float a	 = std::sqrt(6.0f / (64.0f + (float) fanOut));
		const float u	 = (float) (next() >> 8) * (1.0f / 16777216.0f);
		w[p]			 = (2.0f * u - 1.0f) * a;
	}
	return w;
}

bool validN(int64_t n) { return n >= 0 && n <= ((int64_t) 1 << 31); }

// Gradient of one batch; with adam, also the Adam step (3 launches: tiles,
// split-K weight gradients, reduction + update).
int gradientImpl(krr_nrc *h, const krr_nrc_vertex *vert, const float *targets, int64_t n, hipStream_t st, bool adam, const int *nDev = nullptr,
				 int64_t capTotal = 0, int step = 0, int steps = 1) {
	const int64_t chunks = (n + kChunk - 1) / kChunk;
	const int64_t nTiles = chunks * (kChunk / 32);
	const bool bwdLds	 = h->bwdLds;
	const int lds		 = (kFwdLdsElems + (bwdLds ? kBwdLdsElems : 0)) * (int) sizeof(uint16_t);
	const int blocks	 = (int) std::min<int64_t>(1024, (nTiles + 3) / 4);
	hipLaunchKernelGGL(bwdLds ? k_train_tile<true> : k_train_tile<false>, dim3(blocks), dim3(256), lds, st, h->wt.as<uint16_t>(),
					   h->wb.as<uint16_t>(), (const float *) vert,
					   targets, n, nDev, capTotal, step, steps, nTiles, h->ld, h->lossEps, h->xT.as<uint16_t>(), h->dzT.as<uint16_t>(),
					   h->tileLoss.as<float>(), adam ? h->step.as<int>() : nullptr);
	NRC_HIP(hipGetLastError());
	const int wchunks = (int) ((nTiles * 32 + kWChunk - 1) / kWChunk);
	hipLaunchKernelGGL(k_weight_grad, dim3(1, (unsigned) wchunks, kHidden + 1), dim3(256), 0, st, h->xT.as<uint16_t>(), h->dzT.as<uint16_t>(),
					   h->ld, kWChunk, wchunks, h->partial.as<float>());
	NRC_HIP(hipGetLastError());
	if (adam) {
		hipLaunchKernelGGL(k_reduce_adam, dim3((KRR_NRC_PARAMS + kReduceWeights - 1) / kReduceWeights), dim3(kReduceWeights * kReduceGroups), 0,
						   st, h->partial.as<float>(), wchunks,
						   h->grad.as<float>(), h->w.as<float>(), h->m.as<float>(), h->v.as<float>(), h->lr, h->beta1, h->beta2, h->eps,
						   (const int *) h->step.as<int>(), h->wt.as<uint16_t>(), h->wb.as<uint16_t>(), h->tileLoss.as<float>(), nTiles,
						   h->loss.as<float>());
		NRC_HIP(hipGetLastError());
		return KRR_OK;
	}
	hipLaunchKernelGGL(k_reduce_chunks, dim3((KRR_NRC_PARAMS + 255) / 256), dim3(256), 0, st, h->partial.as<float>(), wchunks,
					   h->grad.as<float>());
	NRC_HIP(hipGetLastError());
	hipLaunchKernelGGL(k_reduce_loss, dim3(1), dim3(256), 0, st, h->tileLoss.as<float>(), nTiles, h->loss.as<float>());
	NRC_HIP(hipGetLastError());
	return KRR_OK;
}

} // namespace

extern "C" {

int krr_nrc_create(const char *json, krr_nrc **out) {
	krr::clearError();
	if (!out) return krr::fail(KRR_ERR_INVALID_ARG, "krr_nrc_create: null out");
	*out   = nullptr;
	auto h = std::make_unique<krr_nrc>();
	if (json && *json) {
		try {
			krr::Json j = krr::Json::parse(json);
			if (!j.isObject()) return krr::fail(KRR_ERR_BAD_CONFIG, "nrc config must be a JSON object");
			h->lr		= (float) j.value("lr", 1e-3);
			h->beta1	= (float) j.value("beta1", 0.9);
			h->beta2	= (float) j.value("beta2", 0.99);
			h->eps		= (float) j.value("eps", 1e-8);
			h->lossEps	= (float) j.value("loss_eps", 0.01);
			h->seed		= (uint64_t) j.value("seed", 1.0);
			h->maxBatch = (int64_t) j.value("max_batch", 65536.0);
		} catch (const std::exception &e) {
			return krr::fail(KRR_ERR_BAD_CONFIG, "nrc config: %s", e.what());
		}
	}
	if (!(h->lr > 0) || !(h->beta1 >= 0 && h->beta1 < 1) || !(h->beta2 >= 0 && h->beta2 < 1) || !(h->eps > 0) || !(h->lossEps > 0))
		return krr::fail(KRR_ERR_BAD_CONFIG, "nrc config: lr, eps and loss_eps must be > 0 and betas in [0, 1)");
	if (h->maxBatch < 32 || h->maxBatch > (1 << 24)) return krr::fail(KRR_ERR_BAD_CONFIG, "nrc config: max_batch must be in [32, 2^24]");
	int nDev = 0;
	if (hipGetDeviceCount(&nDev) != hipSuccess || nDev < 1) return krr::fail(KRR_ERR_NO_DEVICE, "krr_nrc_create: no HIP device");
	h->ld = ((h->maxBatch + kChunk - 1) / kChunk) * kChunk;
	const int64_t chunks = h->ld / kChunk;
	NRC_HIP(h->w.alloc(sizeof(float) * KRR_NRC_PARAMS));
	NRC_HIP(h->m.alloc(sizeof(float) * KRR_NRC_PARAMS));
	NRC_HIP(h->v.alloc(sizeof(float) * KRR_NRC_PARAMS));
	NRC_HIP(h->grad.alloc(sizeof(float) * KRR_NRC_PARAMS));
	{
		int dev = 0, maxLds = 0;
		NRC_HIP(hipGetDevice(&dev));
		NRC_HIP(hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
		h->bwdLds = maxLds >= (kFwdLdsElems + kBwdLdsElems) * (int) sizeof(uint16_t);
		int cus = 0, perCu = 0;
		NRC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
		NRC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, k_infer_queue, kInferBlock, kFwdLdsElems * sizeof(uint16_t)));
		h->inferBlocks = std::max(1, cus * std::max(perCu, 1));
	}
	NRC_HIP(h->wt.alloc(sizeof(uint16_t) * kFwdElems));
	NRC_HIP(h->wb.alloc(sizeof(uint16_t) * kBwdElems));
	NRC_HIP(h->xT.alloc(sizeof(uint16_t) * 64 * (kHidden + 1) * h->ld));
	NRC_HIP(h->dzT.alloc(sizeof(uint16_t) * 64 * (kHidden + 1) * h->ld));
	NRC_HIP(h->partial.alloc(sizeof(float) * 4096 * (kHidden + 1) * (h->ld / kWChunk)));
	NRC_HIP(h->tileLoss.alloc(sizeof(float) * (h->ld / 32)));
	NRC_HIP(h->loss.alloc(sizeof(float)));
	NRC_HIP(h->step.alloc(sizeof(int)));
	// padded rows of dZ_5^T (outputs 3..31) and of X^T are written per tile;
	// clear once so that unused feature rows are finite
	NRC_HIP(hipMemset(h->xT.p, 0, h->xT.sz));
	NRC_HIP(hipMemset(h->dzT.p, 0, h->dzT.sz));
	if (int rc = uploadWeights(h.get(), initialWeights(h->seed))) return rc;
	*out = h.release();
	return KRR_OK;
}

int krr_nrc_destroy(krr_nrc *nrc) {
	delete nrc;
	return KRR_OK;
}

int krr_nrc_get_weights(const krr_nrc *nrc, float *host_out) {
	krr::clearError();
	if (!nrc || !host_out) return krr::fail(KRR_ERR_INVALID_ARG, "krr_nrc_get_weights: null argument");
	NRC_HIP(hipDeviceSynchronize());
	NRC_HIP(hipMemcpy(host_out, nrc->w.p, sizeof(float) * KRR_NRC_PARAMS, hipMemcpyDeviceToHost));
	return KRR_OK;
}

int krr_nrc_set_weights(krr_nrc *nrc, const float *host_in) {
	krr::clearError();
	if (!nrc || !host_in) return krr::fail(KRR_ERR_INVALID_ARG, "krr_nrc_set_weights: null argument");
	std::vector<float> w(host_in, host_in + KRR_NRC_PARAMS);
	for (float x : w)
		if (!std::isfinite(x)) return krr::fail(KRR_ERR_INVALID_ARG, "krr_nrc_set_weights: non-finite weight");
	return uploadWeights(nrc, w);
}

int krr_nrc_encode(const krr_nrc *nrc, const krr_nrc_vertex *vertices_dev, int64_t n, float *encoded_dev, void *hip_stream) {
	krr::clearError();
	if (!nrc || !validN(n) || (n > 0 && (!vertices_dev || !encoded_dev))) return krr::fail(KRR_ERR_INVALID_ARG, "krr_nrc_encode: bad argument");
	if (n == 0) return KRR_OK;
	const int64_t total = n * 64;
	hipLaunchKernelGGL(k_encode, dim3((unsigned) ((total + 255) / 256)), dim3(256), 0, (hipStream_t) hip_stream, (const float *) vertices_dev, n,
					   encoded_dev);
	NRC_HIP(hipGetLastError());
	return KRR_OK;
}

int krr_nrc_infer(krr_nrc *nrc, const krr_nrc_vertex *vertices_dev, int64_t n, float *radiance_dev, void *hip_stream) {
	krr::clearError();
	if (!nrc || !validN(n) || (n > 0 && (!vertices_dev || !radiance_dev))) return krr::fail(KRR_ERR_INVALID_ARG, "krr_nrc_infer: bad argument");
	if (n == 0) return KRR_OK;
	const int64_t tiles = (n + 31) / 32;
	const int blocks	= (int) std::min<int64_t>(2048, (tiles + 3) / 4);
	hipLaunchKernelGGL(k_infer, dim3(blocks), dim3(256), kFwdLdsElems * (int) sizeof(uint16_t), (hipStream_t) hip_stream, nrc->wt.as<uint16_t>(),
					   (const float *) vertices_dev, n, radiance_dev);
	NRC_HIP(hipGetLastError());
	return KRR_OK;
}

int krr_nrc_gradient(krr_nrc *nrc, const krr_nrc_vertex *vertices_dev, const float *targets_dev, int64_t n, float *grad_dev,
					 float *loss_out, void *hip_stream) {
	krr::clearError();
	if (!nrc || n < 1 || !vertices_dev || !targets_dev || !grad_dev) return krr::fail(KRR_ERR_INVALID_ARG, "krr_nrc_gradient: bad argument");
	if (n > nrc->maxBatch) return krr::fail(KRR_ERR_CAPACITY, "krr_nrc_gradient: batch %lld exceeds max_batch %lld", (long long) n, (long long) nrc->maxBatch);
	hipStream_t st = (hipStream_t) hip_stream;
	if (int rc = gradientImpl(nrc, vertices_dev, targets_dev, n, st, false)) return rc;
	NRC_HIP(hipMemcpyAsync(grad_dev, nrc->grad.p, sizeof(float) * KRR_NRC_PARAMS, hipMemcpyDeviceToDevice, st));
	if (loss_out) {
		NRC_HIP(hipMemcpyAsync(loss_out, nrc->loss.p, sizeof(float), hipMemcpyDeviceToHost, st));
		NRC_HIP(hipStreamSynchronize(st));
	}
	return KRR_OK;
}

} // extern "C"

int nrcInferAccumulate(krr_nrc *nrc, const float4 *queries, int strideF4, const int *count, int capacity, float4 *L, hipStream_t st) {
	if (capacity <= 0) return KRR_OK;
	// one resident wave of blocks: each block loads the weights into LDS once
	// and then takes query tiles until the queue is done
	const int tilesPerPass = kInferBlock / 64; // one 32-query tile per wave
	const int blocks	   = std::max(1, std::min(nrc->inferBlocks, (capacity + 32 * tilesPerPass - 1) / (32 * tilesPerPass)));
	hipLaunchKernelGGL(k_infer_queue, dim3(blocks), dim3(kInferBlock), kFwdLdsElems * (int) sizeof(uint16_t), st, nrc->wt.as<uint16_t>(), queries, strideF4,
					   count, capacity, L);
	NRC_HIP(hipGetLastError());
	return KRR_OK;
}

int nrcTrainDevice(krr_nrc *nrc, const float *vertices, const float *targets, const int *count, int capacity, int steps, hipStream_t st) {
	const int64_t cap = std::min<int64_t>(capacity, nrc->maxBatch);
	if (cap <= 0 || steps <= 0) return KRR_OK;
	const int64_t m = (cap + steps - 1) / steps; // largest mini-batch
	for (int b = 0; b < steps; b++)
		if (int rc = gradientImpl(nrc, (const krr_nrc_vertex *) vertices, targets, m, st, true, count, cap, b, steps)) return rc;
	return KRR_OK;
}

float nrcLastLoss(krr_nrc *nrc, hipStream_t st) {
	float l = 0.f;
	if (hipMemcpyAsync(&l, nrc->loss.p, sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess) return -1.f;
	if (hipStreamSynchronize(st) != hipSuccess) return -1.f;
	return l;
}

extern "C" {

int krr_nrc_train_step(krr_nrc *nrc, const krr_nrc_vertex *vertices_dev, const float *targets_dev, int64_t n, float *loss_out,
					   void *hip_stream) {
	krr::clearError();
	if (!nrc || n < 1 || !vertices_dev || !targets_dev) return krr::fail(KRR_ERR_INVALID_ARG, "krr_nrc_train_step: bad argument");
	if (n > nrc->maxBatch) return krr::fail(KRR_ERR_CAPACITY, "krr_nrc_train_step: batch %lld exceeds max_batch %lld", (long long) n, (long long) nrc->maxBatch);
	hipStream_t st = (hipStream_t) hip_stream;
	if (int rc = gradientImpl(nrc, vertices_dev, targets_dev, n, st, true)) return rc;
	if (loss_out) {
		NRC_HIP(hipMemcpyAsync(loss_out, nrc->loss.p, sizeof(float), hipMemcpyDeviceToHost, st));
		NRC_HIP(hipStreamSynchronize(st));
	}
	return KRR_OK;
}

int64_t krr_nrc_steps(const krr_nrc *nrc) {
	if (!nrc) return -1;
	int t = 0;
	if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(&t, nrc->step.p, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return -1;
	return t;
}

} // extern "C"
