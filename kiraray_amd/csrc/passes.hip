// passes.hip - AccumulatePass counterpart (krr_accumulate).
//
// Per pixel, as AccumulatePass::render (src/render/passes/accumulate/
// accumulate.cu:30-79): with w = 1 / (count + 1), the first frame is copied;
// later frames are added (sum mode, until max_count frames when max_count >
// 0) or blended with weight w (moving-average mode); the frame buffer is
// rewritten with the accumulated mean (sum * w in sum mode).  The
// accumulation buffer holds float4 or double4 per pixel.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "errors.h"
#include "krr_wavefront.h"

namespace krr {
namespace {

template <class T> struct Vec4 {
	T x, y, z, w;
};

template <class T>
__global__ void k_accumulate_frames(Vec4<T> *accum, float4 *frame, int64_t n, int64_t count, int64_t maxCount, int mode) {
	const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const T wgt		 = T(1) / (T) (count + 1);
	const float4 cur = frame[i];
	Vec4<T> c{(T) cur.x, (T) cur.y, (T) cur.z, (T) cur.w};
	Vec4<T> a = accum[i];
	if (count > 0) {
		if (mode == KRR_ACCUM_MOVING_AVERAGE) {
			a.x = a.x + (c.x - a.x) * wgt;
			a.y = a.y + (c.y - a.y) * wgt;
			a.z = a.z + (c.z - a.z) * wgt;
			a.w = a.w + (c.w - a.w) * wgt;
		} else if (!maxCount || count < maxCount) {
			a.x += c.x;
			a.y += c.y;
			a.z += c.z;
			a.w += c.w;
		}
	} else {
		a = c;
	}
	accum[i] = a;
	if (mode == KRR_ACCUM_MOVING_AVERAGE) frame[i] = make_float4((float) a.x, (float) a.y, (float) a.z, (float) a.w);
	else frame[i] = make_float4((float) (a.x * wgt), (float) (a.y * wgt), (float) (a.z * wgt), (float) (a.w * wgt));
}

} // namespace
} // namespace krr

extern "C" int krr_accumulate(void *accum_dev, float *frame_dev, int64_t n_pixels, int64_t accum_count, int64_t max_count, int32_t mode,
							  int32_t precision, void *hip_stream) {
	krr::clearError();
	if (n_pixels < 0 || accum_count < 0 || max_count < 0 || (n_pixels > 0 && (!accum_dev || !frame_dev)))
		return krr::fail(KRR_ERR_INVALID_ARG, "krr_accumulate: bad argument");
	if (mode != KRR_ACCUM_SUM && mode != KRR_ACCUM_MOVING_AVERAGE) return krr::fail(KRR_ERR_INVALID_ARG, "krr_accumulate: unknown mode %d", mode);
	if (precision != 32 && precision != 64) return krr::fail(KRR_ERR_INVALID_ARG, "krr_accumulate: precision must be 32 or 64");
	if (n_pixels == 0) return KRR_OK;
	const unsigned blocks = (unsigned) ((n_pixels + 255) / 256);
	hipStream_t st		  = (hipStream_t) hip_stream;
	if (precision == 32)
		hipLaunchKernelGGL(krr::k_accumulate_frames<float>, dim3(blocks), dim3(256), 0, st, (krr::Vec4<float> *) accum_dev, (float4 *) frame_dev,
						   n_pixels, accum_count, max_count, mode);
	else
		hipLaunchKernelGGL(krr::k_accumulate_frames<double>, dim3(blocks), dim3(256), 0, st, (krr::Vec4<double> *) accum_dev,
						   (float4 *) frame_dev, n_pixels, accum_count, max_count, mode);
	const hipError_t e = hipGetLastError();
	if (e != hipSuccess) return krr::fail(KRR_ERR_HIP, "krr_accumulate: %s", hipGetErrorString(e));
	return KRR_OK;
}
