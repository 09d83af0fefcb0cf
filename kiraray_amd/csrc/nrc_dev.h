// nrc_dev.h - network geometry and matrix-core operand types of the neural
// radiance cache (nrc.hip).
#pragma once

#include <stdint.h>

#include "krr_nrc.h"

namespace krr {
namespace nrc {

constexpr int kHidden = KRR_NRC_HIDDEN_LAYERS; // hidden layers of width 64 (= matrices before the output one)
// bf16 weight copies: forward A operands W^T [out][in] (output layer padded
// to 32 rows) and backward A operands W [in][out] (output layer padded to 32
// columns)
constexpr int kFwdElems = kHidden * 64 * 64 + 32 * 64;
constexpr int kBwdElems = kHidden * 64 * 64 + 64 * 32;
// The copies in LDS pad every row by 4 bf16 (8 bytes): the 32 lanes that read
// an A operand read 32 consecutive rows, and unpadded 128-byte rows would put
// all of them on one bank.  With the padding consecutive rows start 2 banks
// apart, so a wave's 8-byte reads spread over every bank.
constexpr int kPad			= 4;
constexpr int kRowH			= 64 + kPad; // LDS row of a 64-wide matrix
constexpr int kRowO			= 32 + kPad; // LDS row of the backward output matrix (32 padded columns)
constexpr int kFwdLdsElems = (kHidden * 64 + 32) * kRowH;
constexpr int kBwdLdsElems = kHidden * 64 * kRowH + 64 * kRowO;

typedef short short4v __attribute__((ext_vector_type(4)));	// 4 x bf16 (bit patterns)
typedef float float16v __attribute__((ext_vector_type(16))); // 32x32 accumulator slice of one lane

} // namespace nrc
} // namespace krr

#ifdef __HIPCC__
#include <hip/hip_runtime.h>

// Entry points for the path tracer, with counts read on the device (no host
// synchronisation inside a frame).
//
// Queries: `count` records of `strideF4` float4 each; float4 0..3 are the
// krr_nrc_vertex, float4 4 = {thp.xyz, pixel bits}.  Adds thp * max(cache, 0)
// to L[pixel] (one query per pixel per sample, so no atomics).
int nrcInferAccumulate(krr_nrc *nrc, const float4 *queries, int strideF4, const int *count, int capacity, float4 *L, hipStream_t st);
// `steps` Adam steps, one per equal consecutive slice of the first
// min(*count, capacity) vertices (16 floats each) and targets (3 floats
// each), the slices computed on the device; kernels sized for `capacity`.
int nrcTrainDevice(krr_nrc *nrc, const float *vertices, const float *targets, const int *count, int capacity, int steps, hipStream_t st);
// Loss of the last training step (synchronises the stream).
float nrcLastLoss(krr_nrc *nrc, hipStream_t st);
#endif
