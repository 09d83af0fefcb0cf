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

typedef short short4v __attribute__((ext_vector_type(4)));	// 4 x bf16 (bit patterns)
typedef float float16v __attribute__((ext_vector_type(16))); // 32x32 accumulator slice of one lane

} // namespace nrc
} // namespace krr
