// vecmath.h - 3-vector / RGB arithmetic for device shading code.
//
// Component-wise operations evaluate in the same order as the Eigen
// expressions of the reference shading code (x op y per component, sums left
// to right), and every kernel is compiled with -ffp-contract=off, so each
// result is one IEEE binary32 operation per source-level operator.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "detmath.h"

#define KRR_DEV __device__ __forceinline__
#define KRR_HD __host__ __device__ __forceinline__

namespace krr {

struct f3 {
	float x, y, z;
	KRR_HD f3() : x(0), y(0), z(0) {}
	KRR_HD f3(float a, float b, float c) : x(a), y(b), z(c) {}
	KRR_HD explicit f3(float a) : x(a), y(a), z(a) {}
	// over values: a conditional over the members is a load through a
	// selected address, which keeps the vector in scratch memory
	KRR_HD float operator[](int i) const {
		const float a = x, b = y, c = z;
		return i == 0 ? a : (i == 1 ? b : c);
	}
	KRR_HD f3 operator+(const f3 &b) const { return {x + b.x, y + b.y, z + b.z}; }
	KRR_HD f3 operator-(const f3 &b) const { return {x - b.x, y - b.y, z - b.z}; }
	KRR_HD f3 operator*(const f3 &b) const { return {x * b.x, y * b.y, z * b.z}; }
	KRR_HD f3 operator/(const f3 &b) const { return {x / b.x, y / b.y, z / b.z}; }
	KRR_HD f3 operator*(float s) const { return {x * s, y * s, z * s}; }
	KRR_HD f3 operator/(float s) const { return {x / s, y / s, z / s}; }
	KRR_HD f3 operator-() const { return {-x, -y, -z}; }
	KRR_HD bool any() const { return x != 0.f || y != 0.f || z != 0.f; }
	KRR_HD bool isZero() const { return x == 0.f && y == 0.f && z == 0.f; }
	KRR_HD float mean() const { return (x + y + z) / 3.f; }
	KRR_HD float maxCoeff() const {
		float m = x > y ? x : y;
		return m > z ? m : z;
	}
	KRR_HD bool hasNaN() const { return x != x || y != y || z != z; }
};

KRR_HD f3 operator*(float s, const f3 &v) { return {s * v.x, s * v.y, s * v.z}; }
KRR_HD float dot(const f3 &a, const f3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
KRR_HD f3 cross(const f3 &a, const f3 &b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
KRR_HD float sqnorm(const f3 &v) { return dot(v, v); }
KRR_HD float length(const f3 &v) { return sqrtf(sqnorm(v)); }
// Eigen's normalized(): v / sqrt(|v|^2) if |v|^2 > 0, else v unchanged.
KRR_HD f3 normalize(const f3 &v) {
	float z = sqnorm(v);
	return z > 0.f ? v / sqrtf(z) : v;
}
KRR_HD f3 vabs(const f3 &v) { return {fabsf(v.x), fabsf(v.y), fabsf(v.z)}; }
KRR_HD f3 vmaxs(const f3 &v, float s) { return {v.x > s ? v.x : s, v.y > s ? v.y : s, v.z > s ? v.z : s}; }
KRR_HD f3 vmins(const f3 &v, float s) { return {v.x < s ? v.x : s, v.y < s ? v.y : s, v.z < s ? v.z : s}; }
KRR_HD f3 vsqrt(const f3 &v) { return {sqrtf(v.x), sqrtf(v.y), sqrtf(v.z)}; }
KRR_HD f3 vexp(const f3 &v) { return {detmath::expf(v.x), detmath::expf(v.y), detmath::expf(v.z)}; }
KRR_HD float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
KRR_HD float pow2(float x) { return x * x; }
KRR_HD float maxf(float a, float b) { return a > b ? a : b; }
KRR_HD float minf(float a, float b) { return a < b ? a : b; }

// Constants of src/core/math/include/krrmath/constants.h (float literals).
constexpr float kPi		 = 3.14159265358979323846f;
constexpr float k2Pi	 = 6.28318530717958647693f;
constexpr float kInvPi	 = 0.318309886183790671538f;
constexpr float kInv2Pi	 = 0.15915494309189533577f;
constexpr float kInv4Pi	 = 0.07957747154594766788f;
constexpr float kRayEps	 = 1e-4f; // KRR_RAY_EPS
constexpr float kFloatMax = 3.402823466e+38f;

} // namespace krr
