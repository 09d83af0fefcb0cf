// core.h - shading, sampling and intersection core of the CPU checker.
//
// TEST INFRASTRUCTURE: the checker for kiraray_amd's HIP tracer.  Only the
// tests, the smoke check and the CPU-baseline leg of the benchmark load it;
// the product never does.
//
// Written from the reference's algorithm (file:line citations per function,
// paths relative to the KiRaRay repository root), not from the HIP code.  The
// template parameter R is float (parity mode: the same IEEE binary32
// operations in the same order as the GPU, elementary functions from
// detmath.h) or double (reference-image mode, libm).
#pragma once
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "detmath.h"

// Planted mistakes for the pin tests: built with -DORACLE_PLANT=k, the
// checker carries one deliberate error (a dropped term, a wrong index or a
// wrong constant) in the medium code, and the pin that covers that code must
// then fail.  0, the default, is the restatement itself.
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif

namespace oracle {

// MurmurHash64A (Austin Appleby's 64-bit MurmurHash2 for x64), the hash of
// src/util/hash.h:37-76 behind HashFloat (hash.h:130-134): 8-byte blocks read
// little-endian, then the 1..7 tail bytes, then the final avalanche.
inline uint64_t murmurHash64A(const unsigned char *key, size_t len, uint64_t seed) {
	const uint64_t m = 0xc6a4a7935bd1e995ull;
	uint64_t h		 = seed ^ (len * m);
	const size_t nb	 = len / 8;
	for (size_t i = 0; i < nb; i++) {
		uint64_t k;
		std::memcpy(&k, key + 8 * i, 8);
		k *= m;
		k ^= k >> 47;
		k *= m;
		h ^= k;
		h *= m;
	}
	const size_t rest = len & 7;
	if (rest) {
		uint64_t t = 0;
		for (size_t i = 0; i < rest; i++) t |= (uint64_t) key[8 * nb + i] << (8 * i);
		h ^= t;
		h *= m;
	}
	h ^= h >> 47;
	h *= m;
	h ^= h >> 47;
	return h;
}

// ----------------------------------------------------------------------------
// Scalar math: binary32 routes through detmath, binary64 through libm.
// ----------------------------------------------------------------------------
inline float m_sqrt(float x) { return std::sqrt(x); }
inline double m_sqrt(double x) { return std::sqrt(x); }
inline float m_sin(float x) { return detmath::sinf(x); }
inline double m_sin(double x) { return std::sin(x); }
inline float m_cos(float x) { return detmath::cosf(x); }
inline double m_cos(double x) { return std::cos(x); }
inline float m_log(float x) { return detmath::logf(x); }
inline double m_log(double x) { return std::log(x); }
inline float m_exp(float x) { return detmath::expf(x); }
inline double m_exp(double x) { return std::exp(x); }
inline float m_tan(float x) { return detmath::tanf(x); }
inline double m_tan(double x) { return std::tan(x); }
inline float m_atan2(float y, float x) { return detmath::atan2f(y, x); }
inline double m_atan2(double y, double x) { return std::atan2(y, x); }
inline float m_acos(float x) { return detmath::acosf(x); }
inline double m_acos(double x) { return std::acos(x < -1.0 ? -1.0 : (x > 1.0 ? 1.0 : x)); }
inline float m_floor(float x) { return std::floor(x); }
inline double m_floor(double x) { return std::floor(x); }
template <class R> inline R m_abs(R x) { return std::fabs(x); }
template <class R> inline R m_min(R a, R b) { return a < b ? a : b; }
template <class R> inline R m_max(R a, R b) { return a > b ? a : b; }
template <class R> inline R clampv(R v, R lo, R hi) { return v < lo ? lo : (v > hi ? hi : v); }
template <class R> inline R pow2(R x) { return x * x; }
template <class R> inline bool is_inf(R x) { return std::isinf(x); }
template <class R> inline R safe_sqrt(R x) { return m_sqrt(m_max(R(0), x)); }

// Constants (src/core/math/include/krrmath/constants.h:17-27; the Windows
// build redefines M_PI as a float literal).
template <class R> constexpr R kPi() { if constexpr (sizeof(R) == 4) return 3.14159265358979323846f; else return 3.14159265358979323846; }
template <class R> constexpr R k2Pi() { if constexpr (sizeof(R) == 4) return 6.28318530717958647693f; else return 6.28318530717958647693; }
template <class R> constexpr R kInvPi() { if constexpr (sizeof(R) == 4) return 0.318309886183790671538f; else return 0.318309886183790671538; }
template <class R> constexpr R kInv2Pi() { if constexpr (sizeof(R) == 4) return 0.15915494309189533577f; else return 0.15915494309189533577; }
template <class R> constexpr R kInv4Pi() { if constexpr (sizeof(R) == 4) return 0.07957747154594766788f; else return 0.07957747154594766788; }
constexpr float kRayEps = 1e-4f;   // KRR_RAY_EPS, src/core/raytracing.h:10

template <class R> struct V3 {
	R x, y, z;
	V3() : x(0), y(0), z(0) {}
	V3(R a, R b, R c) : x(a), y(b), z(c) {}
	explicit V3(R a) : x(a), y(a), z(a) {}
	R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
	R &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
	V3 operator+(const V3 &b) const { return {x + b.x, y + b.y, z + b.z}; }
	V3 operator-(const V3 &b) const { return {x - b.x, y - b.y, z - b.z}; }
	V3 operator*(const V3 &b) const { return {x * b.x, y * b.y, z * b.z}; }
	V3 operator/(const V3 &b) const { return {x / b.x, y / b.y, z / b.z}; }
	V3 operator*(R s) const { return {x * s, y * s, z * s}; }
	V3 operator/(R s) const { return {x / s, y / s, z / s}; }
	V3 operator-() const { return {-x, -y, -z}; }
	V3 &operator+=(const V3 &b) { x = x + b.x; y = y + b.y; z = z + b.z; return *this; }
	V3 &operator*=(const V3 &b) { x = x * b.x; y = y * b.y; z = z * b.z; return *this; }
	V3 &operator*=(R s) { x = x * s; y = y * s; z = z * s; return *this; }
	V3 &operator/=(R s) { x = x / s; y = y / s; z = z / s; return *this; }
	bool any() const { return x != 0 || y != 0 || z != 0; }
	bool isZero() const { return x == 0 && y == 0 && z == 0; }
	R mean() const { return (x + y + z) / R(3); }
	R maxCoeff() const { return m_max(m_max(x, y), z); }
	bool hasNaN() const { return x != x || y != y || z != z; }
};
template <class R> inline V3<R> operator*(R s, const V3<R> &v) { return {s * v.x, s * v.y, s * v.z}; }
template <class R> inline R dot(const V3<R> &a, const V3<R> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(const V3<R> &a, const V3<R> &b) {
	return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R squaredNorm(const V3<R> &v) { return dot(v, v); }
template <class R> inline R norm(const V3<R> &v) { return m_sqrt(squaredNorm(v)); }
// Eigen normalized(): v / sqrt(|v|^2) when |v|^2 > 0.
template <class R> inline V3<R> normalize(const V3<R> &v) {
	R z = squaredNorm(v);
	return z > R(0) ? v / m_sqrt(z) : v;
}
template <class R> inline V3<R> vabs(const V3<R> &v) { return {m_abs(v.x), m_abs(v.y), m_abs(v.z)}; }
template <class R> inline V3<R> vexp(const V3<R> &v) { return {m_exp(v.x), m_exp(v.y), m_exp(v.z)}; }
template <class R> inline V3<R> vmaxs(const V3<R> &v, R s) { return {m_max(v.x, s), m_max(v.y, s), m_max(v.z, s)}; }
template <class R> inline V3<R> vmins(const V3<R> &v, R s) { return {m_min(v.x, s), m_min(v.y, s), m_min(v.z, s)}; }
template <class R> inline V3<R> vsqrt(const V3<R> &v) { return {m_sqrt(v.x), m_sqrt(v.y), m_sqrt(v.z)}; }
template <class R, class S> inline V3<R> cvt(const V3<S> &v) { return {R(v.x), R(v.y), R(v.z)}; }

using V3f = V3<float>;

// ----------------------------------------------------------------------------
// PCG32 sampler (src/core/sampler.h:13-91) and Morton interleave
// (src/util/hash.h:12-27).
// ----------------------------------------------------------------------------
constexpr uint64_t kPcgMult = 0x5851f42d4c957f2dULL;

struct Pcg {
	uint64_t state = 0, inc = 1;
	uint32_t nextUint() {
		uint64_t old = state;
		state		 = old * kPcgMult + inc;
		uint32_t xs	 = (uint32_t) (((old >> 18u) ^ old) >> 27u);
		uint32_t rot = (uint32_t) (old >> 59u);
		return (xs >> rot) | (xs << ((~rot + 1u) & 31));
	}
	void setSeed(uint64_t initstate, uint64_t initseq) {
		state = 0u;
		inc	  = (initseq << 1u) | 1u;
		nextUint();
		state += initstate;
		nextUint();
	}
	void advance(int64_t delta) {
		uint64_t cur_mult = kPcgMult, cur_plus = inc, acc_mult = 1u, acc_plus = 0u;
		while (delta > 0) {
			if (delta & 1) {
				acc_mult *= cur_mult;
				acc_plus = acc_plus * cur_mult + cur_plus;
			}
			cur_plus = (cur_mult + 1) * cur_plus;
			cur_mult *= cur_mult;
			delta /= 2;
		}
		state = acc_mult * state + acc_plus;
	}
	float get1D() {
		uint32_t u = (nextUint() >> 9) | 0x3f800000u;
		float f;
		std::memcpy(&f, &u, 4);
		return f - 1.0f;
	}
};

inline uint32_t interleave32(uint32_t xv, uint32_t yv) {
	uint32_t x = xv & 0x0000ffff, y = yv & 0x0000ffff;
	x = (x | (x << 8)) & 0x00FF00FF;
	x = (x | (x << 4)) & 0x0F0F0F0F;
	x = (x | (x << 2)) & 0x33333333;
	x = (x | (x << 1)) & 0x55555555;
	y = (y | (y << 8)) & 0x00FF00FF;
	y = (y | (y << 4)) & 0x0F0F0F0F;
	y = (y | (y << 2)) & 0x33333333;
	y = (y | (y << 1)) & 0x55555555;
	return x | (y << 1);
}

// ----------------------------------------------------------------------------
// BSDF type bits (src/render/materials/bxdf.h:42-62).
// ----------------------------------------------------------------------------
enum : int {
	BSDF_UNSET = 0, BSDF_NULL = 1, BSDF_REFLECTION = 2, BSDF_TRANSMISSION = 4, BSDF_DIFFUSE = 8,
	BSDF_GLOSSY = 16, BSDF_SPECULAR = 32,
	BSDF_SMOOTH = BSDF_DIFFUSE | BSDF_GLOSSY, BSDF_DELTA = BSDF_SPECULAR | BSDF_NULL,
};

// ----------------------------------------------------------------------------
// Sampling helpers (src/render/sampling.h:27-105).
// ----------------------------------------------------------------------------
template <class R> V3<R> uniformSampleSphere(R u0, R u1) {
	R z	  = R(1) - R(2) * u0;
	R r	  = m_sqrt(m_max(R(0), R(1) - z * z));
	R phi = k2Pi<R>() * u1;
	return {r * m_cos(phi), r * m_sin(phi), z};
}

template <class R> void uniformSampleDisk(R u0, R u1, R &dx, R &dy) {
	R ox = R(2) * u0 - R(1), oy = R(2) * u1 - R(1);
	if (ox == 0 && oy == 0) { dx = 0; dy = 0; return; }
	R theta, r;
	if (m_abs(ox) > m_abs(oy)) {
		r	  = ox;
		theta = kPi<R>() / R(4) * (oy / ox);
	} else {
		r	  = oy;
		theta = kPi<R>() / R(2) - kPi<R>() / R(4) * (ox / oy);
	}
	dx = r * m_cos(theta);
	dy = r * m_sin(theta);
}

template <class R> V3<R> cosineSampleHemisphere(R u0, R u1) {
	R dx, dy;
	uniformSampleDisk(u0, u1, dx, dy);
	R z = m_sqrt(m_max(R(0), R(1) - dx * dx - dy * dy));
	return {dx, dy, z};
}

template <class R> void uniformSampleTriangle(R u0, R u1, R &b0, R &b1, R &b2) {
	if (u0 < u1) {
		b0 = u0 / R(2);
		b1 = u1 - b0;
	} else {
		b1 = u1 / R(2);
		b0 = u0 - b1;
	}
	b2 = R(1) - b0 - b1;
}

template <class R> R sampleExponential(R u, R a) { return -m_log(R(1) - u) / a; }

inline float nextFloatDown(float v) {
	if (std::isinf(v) && v < 0.f) return v;
	if (v == 0.f) v = -0.f;
	uint32_t ui;
	std::memcpy(&ui, &v, 4);
	if (v > 0) --ui;
	else ++ui;
	std::memcpy(&v, &ui, 4);
	return v;
}
inline double nextFloatDown(double v) { return std::nextafter(v, -std::numeric_limits<double>::infinity()); }

template <class R> int sampleDiscrete3(const R w[3], R u, R *pmf) {
	R sumWeights = R(0);
	for (int i = 0; i < 3; i++) sumWeights += w[i];
	R up = u * sumWeights;
	if (up == sumWeights) up = nextFloatDown(up);
	int offset = 0;
	R sum	   = R(0);
	while (offset < 2 && sum + w[offset] <= up) sum += w[offset++];
	if (pmf) *pmf = w[offset] / sumWeights;
	return offset;
}

// getPerpendicular (src/util/math_utils.h:120-130).
template <class R> V3<R> getPerpendicular(const V3<R> &u) {
	V3<R> a		 = vabs(u);
	uint32_t uyx = (a.x - a.y) < 0 ? 1 : 0;
	uint32_t uzx = (a.x - a.z) < 0 ? 1 : 0;
	uint32_t uzy = (a.y - a.z) < 0 ? 1 : 0;
	uint32_t xm	 = uyx & uzx;
	uint32_t ym	 = (1 ^ xm) & uzy;
	uint32_t zm	 = 1 ^ (xm | ym);
	return normalize(cross(u, V3<R>(R(xm), R(ym), R(zm))));
}

// ----------------------------------------------------------------------------
// Fresnel and microfacet model (src/render/materials/{matutils,fresnel,
// microfacet}.h).
// ----------------------------------------------------------------------------
template <class R> inline bool SameHemisphere(const V3<R> &w, const V3<R> &wp) { return w.z * wp.z > 0; }
template <class R> inline R CosTheta(const V3<R> &w) { return w.z; }
template <class R> inline R Cos2Theta(const V3<R> &w) { return w.z * w.z; }
template <class R> inline R AbsCosTheta(const V3<R> &w) { return m_abs(w.z); }
template <class R> inline R Sin2Theta(const V3<R> &w) { return m_max(R(0), R(1) - Cos2Theta(w)); }
template <class R> inline R SinTheta(const V3<R> &w) { return m_sqrt(Sin2Theta(w)); }
template <class R> inline R TanTheta(const V3<R> &w) { return SinTheta(w) / CosTheta(w); }
template <class R> inline R Tan2Theta(const V3<R> &w) { return Sin2Theta(w) / Cos2Theta(w); }
template <class R> inline R CosPhi(const V3<R> &w) {
	R s = SinTheta(w);
	return (s == 0) ? R(1) : clampv(w.x / s, R(-1), R(1));
}
template <class R> inline R SinPhi(const V3<R> &w) {
	R s = SinTheta(w);
	return (s == 0) ? R(0) : clampv(w.y / s, R(-1), R(1));
}
template <class R> inline R Cos2Phi(const V3<R> &w) { return CosPhi(w) * CosPhi(w); }
template <class R> inline R Sin2Phi(const V3<R> &w) { return SinPhi(w) * SinPhi(w); }
template <class R> inline R AbsDot(const V3<R> &a, const V3<R> &b) { return m_abs(dot(a, b)); }
template <class R> inline V3<R> Reflect(const V3<R> &wo, const V3<R> &n) { return -wo + R(2) * dot(wo, n) * n; }
template <class R> inline V3<R> FaceForward(const V3<R> &w, const V3<R> &wp) { return dot(w, wp) > 0 ? w : -w; }
template <class R> inline V3<R> ToSameHemisphere(const V3<R> &w, const V3<R> &wp) {
	return {w.x, w.y, w.z * wp.z > 0 ? w.z : -w.z};
}

// Refract with absolute eta (matutils.h:85-99).
template <class R> bool Refract(const V3<R> &wi, const V3<R> &n, R eta, R *etap, V3<R> *wt) {
	R cosThetaI = dot(n, wi);
	if (wi.z < 0) eta = R(1) / eta;
	R sin2ThetaI = m_max(R(0), R(1) - pow2(cosThetaI));
	R sin2ThetaT = sin2ThetaI / pow2(eta);
	if (sin2ThetaT >= 1) return false;
	R cosThetaT = m_sqrt(R(1) - sin2ThetaT);
	*wt			= -wi / eta + (cosThetaI / eta - cosThetaT) * n;
	if (etap) *etap = eta;
	return true;
}

// FrDielectric (fresnel.h:24-44).
template <class R> R FrDielectric(R cosTheta_i, R eta) {
	cosTheta_i = clampv(cosTheta_i, R(-1), R(1));
	if (cosTheta_i < 0) {
		eta		   = R(1) / eta;
		cosTheta_i = -cosTheta_i;
	}
	R sin2Theta_i = R(1) - pow2(cosTheta_i);
	R sin2Theta_t = sin2Theta_i / pow2(eta);
	if (sin2Theta_t >= 1) return R(1);
	R cosTheta_t = m_sqrt(m_max(R(1) - sin2Theta_t, R(0)));
	R r_parl	 = (eta * cosTheta_i - cosTheta_t) / (eta * cosTheta_i + cosTheta_t);
	R r_perp	 = (cosTheta_i - eta * cosTheta_t) / (cosTheta_i + eta * cosTheta_t);
	return (pow2(r_parl) + pow2(r_perp)) / R(2);
}

// Complex arithmetic of src/core/math/include/krrmath/complex.h:11-60 (real
// operands are promoted to Complex(re, 0) as the reference's implicit
// conversions do).
template <class R> struct Cx {
	R re, im;
	Cx(R r) : re(r), im(0) {}
	Cx(R r, R i) : re(r), im(i) {}
	Cx operator+(Cx z) const { return {re + z.re, im + z.im}; }
	Cx operator-(Cx z) const { return {re - z.re, im - z.im}; }
	Cx operator*(Cx z) const { return {re * z.re - im * z.im, re * z.im + im * z.re}; }
	Cx operator/(Cx z) const {
		R scale = R(1) / (z.re * z.re + z.im * z.im);
		return {scale * (re * z.re + im * z.im), scale * (im * z.re - re * z.im)};
	}
	R norm() const { return re * re + im * im; }
	R abs() const { return m_sqrt(norm()); }
	Cx sqrt() const {
		R n = this->abs(), t1 = m_sqrt(R(.5) * (n + m_abs(re))), t2 = R(.5) * im / t1;
		if (n == 0) return Cx(R(0));
		if (re >= 0) return {t1, t2};
		return {m_abs(t2), std::copysign(t1, im)};
	}
};

// FrComplex (fresnel.h:46-58) for one channel.
template <class R> R FrComplex1(R cosTheta_i, Cx<R> eta) {
	cosTheta_i		  = clampv(cosTheta_i, R(0), R(1));
	R sin2Theta_i	  = R(1) - pow2(cosTheta_i);
	Cx<R> sin2Theta_t = Cx<R>(sin2Theta_i) / (eta * eta);
	Cx<R> cosTheta_t  = (Cx<R>(R(1)) - sin2Theta_t).sqrt();
	Cx<R> ci(cosTheta_i);
	Cx<R> r_parl = (eta * ci - cosTheta_t) / (eta * ci + cosTheta_t);
	Cx<R> r_perp = (ci - eta * cosTheta_t) / (ci + eta * cosTheta_t);
	return (r_parl.norm() + r_perp.norm()) / R(2);
}

template <class R> V3<R> FrComplex(R cosTheta_i, const V3<R> &eta, const V3<R> &k) {
	return {FrComplex1(cosTheta_i, Cx<R>(eta.x, k.x)), FrComplex1(cosTheta_i, Cx<R>(eta.y, k.y)),
			FrComplex1(cosTheta_i, Cx<R>(eta.z, k.z))};
}

// GGXMicrofacetDistribution (microfacet.h:17-70) with the legacy slope
// sampler GGXSample11/GGXSample (microfacet.h:86-154, selected by
// KRR_GGX_SAMPLE_LEGACY at microfacet.h:159).
template <class R> struct GGX {
	R ax, ay;
	GGX() : ax(R(1e-3f)), ay(R(1e-3f)) {}
	GGX(R a, R b) : ax(m_max(R(1e-3f), a)), ay(m_max(R(1e-3f), b)) {}
	bool isDelta() const { return m_max(ax, ay) <= R(1e-3f); }
	R Lambda(const V3<R> &w) const {
		R absTanTheta = m_abs(TanTheta(w));
		if (is_inf(absTanTheta)) return R(0);
		R alpha			 = m_sqrt(Cos2Phi(w) * ax * ax + Sin2Phi(w) * ay * ay);
		R alpha2Tan2Theta = pow2(alpha * absTanTheta);
		return (R(-1) + m_sqrt(R(1) + alpha2Tan2Theta)) / R(2);
	}
	R G1(const V3<R> &w) const { return R(1) / (R(1) + Lambda(w)); }
	R G(const V3<R> &wo, const V3<R> &wi) const { return R(1) / (R(1) + Lambda(wo) + Lambda(wi)); }
	R D(const V3<R> &wh) const {
		R tan2Theta = Tan2Theta(wh);
		if (is_inf(tan2Theta)) return R(0);
		const R cos4Theta = pow2(Cos2Theta(wh));
		R e				  = (pow2(CosPhi(wh) / ax) + pow2(SinPhi(wh) / ay)) * tan2Theta;
		return R(1) / (kPi<R>() * ax * ay * cos4Theta * pow2(R(1) + e));
	}
	R Pdf(const V3<R> &wo, const V3<R> &wh) const { return D(wh) * G1(wo) * m_abs(dot(wo, wh)) / AbsCosTheta(wo); }
	V3<R> Sample(const V3<R> &wo, R u0, R u1) const {
		bool flip = wo.z < 0;
		V3<R> wh  = GGXSample(flip ? -wo : wo, u0, u1);
		return flip ? -wh : wh;
	}
	static void GGXSample11(R cosTheta, R U1, R U2, R *slope_x, R *slope_y) {
		if (cosTheta > R(.9999f)) {
			R r	  = m_sqrt(U1 / (R(1) - U1));
			R phi = k2Pi<R>() * U2;
			*slope_x = r * m_cos(phi);
			*slope_y = r * m_sin(phi);
			return;
		}
		R sinTheta = m_sqrt(m_max(R(0), R(1) - cosTheta * cosTheta));
		R tanTheta = sinTheta / cosTheta;
		R a		   = R(1) / tanTheta;
		R G1v	   = R(2) / (R(1) + m_sqrt(R(1) + R(1) / (a * a)));
		R A		   = R(2) * U1 / G1v - R(1);
		R tmp	   = R(1) / (A * A - R(1));
		if ((double) tmp > 1e10) tmp = R(1e10);
		R B			= tanTheta;
		R Dv		= m_sqrt(m_max(R(B * B * tmp * tmp - (A * A - B * B) * tmp), R(0)));
		R slope_x_1 = B * tmp - Dv;
		R slope_x_2 = B * tmp + Dv;
		*slope_x	= (A < 0 || slope_x_2 > R(1) / tanTheta) ? slope_x_1 : slope_x_2;
		R S;
		if (U2 > R(0.5f)) {
			S  = R(1);
			U2 = R(2) * (U2 - R(.5f));
		} else {
			S  = R(-1);
			U2 = R(2) * (R(.5f) - U2);
		}
		R z = (U2 * (U2 * (U2 * R(0.27385f) - R(0.73369f)) + R(0.46341f))) /
			  (U2 * (U2 * (U2 * R(0.093073f) + R(0.309420f)) - R(1.000000f)) + R(0.597999f));
		*slope_y = S * z * m_sqrt(R(1) + *slope_x * *slope_x);
	}
	V3<R> GGXSample(const V3<R> &wi, R U1, R U2) const {
		V3<R> wiStretched = normalize(V3<R>(ax * wi.x, ay * wi.y, wi.z));
		R slope_x, slope_y;
		GGXSample11(CosTheta(wiStretched), U1, U2, &slope_x, &slope_y);
		R tmp	= CosPhi(wiStretched) * slope_x - SinPhi(wiStretched) * slope_y;
		slope_y = SinPhi(wiStretched) * slope_x + CosPhi(wiStretched) * slope_y;
		slope_x = tmp;
		slope_x = ax * slope_x;
		slope_y = ay * slope_y;
		return normalize(V3<R>(-slope_x, -slope_y, R(1)));
	}
};

// ----------------------------------------------------------------------------
// BSDFs (src/render/bsdf.h:19-54; materials/{diffuse,conductor,dielectric,
// null}.h) on the shading data of prepareSurfaceInteraction.
// ----------------------------------------------------------------------------
template <class R> struct BSDFSample {
	V3<R> f;
	V3<R> wi;
	R pdf	  = 0;
	int flags = BSDF_UNSET;
};

template <class R> struct ShadingData {
	int type = KRR_MAT_NULL;
	R IoR	 = R(1.5f);
	V3<R> diffuse, specular;
	R specularTransmission = 0, roughness = 1, anisotropic = 0;
	// BSDFData::getBsdfType (src/render/shared.h:46-73)
	int bsdfType() const {
		int t = BSDF_UNSET;
		switch (type) {
			case KRR_MAT_NULL: t = BSDF_NULL; break;
			case KRR_MAT_DIFFUSE: t = BSDF_DIFFUSE | BSDF_REFLECTION; break;
			case KRR_MAT_DIELECTRIC:
				t = roughness <= R(1e-3f) ? BSDF_SPECULAR : BSDF_GLOSSY;
				t = t | BSDF_REFLECTION | BSDF_TRANSMISSION;
				break;
			case KRR_MAT_CONDUCTOR: t = BSDF_REFLECTION | (roughness <= R(1e-3f) ? BSDF_SPECULAR : BSDF_GLOSSY); break;
			default: break;
		}
		return t;
	}
};

template <class R> struct BSDF {
	int type;
	// diffuse
	V3<R> diffuse;
	// conductor
	V3<R> eta, k, reflectance;
	// dielectric
	R etaS = 1;
	V3<R> baseColor;
	GGX<R> dist;

	explicit BSDF(const ShadingData<R> &sd) : type(sd.type) {
		if (type == KRR_MAT_DIFFUSE) {
			diffuse = sd.diffuse;
		} else if (type == KRR_MAT_CONDUCTOR) { // conductor.h:12-32
			reflectance = vmins(sd.diffuse, R(0.9999f));
			R aspect	= m_sqrt(R(1) - sd.anisotropic * R(.9f));
			R axv		= m_max(R(.001f), pow2(sd.roughness) / aspect);
			R ayv		= m_max(R(.001f), pow2(sd.roughness) * aspect);
			dist		= GGX<R>(axv, ayv);
			eta			= V3<R>(sd.IoR);
			k			= (R(2) * vsqrt(reflectance)) / vsqrt(vmaxs(V3<R>(R(1)) - reflectance, R(0)));
		} else if (type == KRR_MAT_DIELECTRIC) { // dielectric.h:88-93
			baseColor = sd.diffuse * sd.specularTransmission;
			etaS	  = sd.IoR;
			R alpha	  = pow2(sd.roughness);
			dist	  = GGX<R>(alpha, alpha);
		}
	}

	V3<R> f(const V3<R> &wo, const V3<R> &wi) const {
		switch (type) {
			case KRR_MAT_DIFFUSE:
				if (!SameHemisphere(wo, wi)) return V3<R>();
				return diffuse * kInvPi<R>();
			case KRR_MAT_CONDUCTOR: { // conductor.h:34-50
				if (!SameHemisphere(wo, wi)) return V3<R>();
				if (dist.isDelta()) return V3<R>();
				R cosTheta_o = AbsCosTheta(wo), cosTheta_i = AbsCosTheta(wi);
				if (cosTheta_i == 0 || cosTheta_o == 0) return V3<R>();
				V3<R> wm = wi + wo;
				if (squaredNorm(wm) == 0) return V3<R>();
				wm		= normalize(wm);
				V3<R> F = FrComplex(AbsDot(wo, wm), eta, k);
				return dist.D(wm) * F * dist.G(wo, wi) / (R(4) * cosTheta_i * cosTheta_o);
			}
			case KRR_MAT_DIELECTRIC: { // dielectric.h:95-129
				if (etaS == 1 || dist.isDelta()) return V3<R>();
				R cosTheta_o = CosTheta(wo), cosTheta_i = CosTheta(wi);
				bool reflect = cosTheta_i * cosTheta_o > 0;
				R etap		 = 1;
				if (!reflect) etap = cosTheta_o > 0 ? etaS : (R(1) / etaS);
				V3<R> wm = wi * etap + wo;
				if (cosTheta_i == 0 || cosTheta_o == 0 || !wm.any()) return V3<R>();
				wm = FaceForward(normalize(wm), V3<R>(0, 0, 1));
				if (dot(wm, wi) * cosTheta_i < 0 || dot(wm, wo) * cosTheta_o < 0) return V3<R>();
				R F		= FrDielectric(std::copysign(dot(wo, wm), wo.z), etaS);
				V3<R> Rs = baseColor * F, Ts = baseColor * (R(1) - F);
				if (reflect) {
					return dist.D(wm) * dist.G(wo, wi) * Rs / m_abs(R(4) * cosTheta_i * cosTheta_o);
				} else {
					R denom	 = pow2(dot(wi, wm) + dot(wo, wm) / etap) * cosTheta_i * cosTheta_o;
					V3<R> ft = Ts * dist.D(wm) * dist.G(wo, wi) * m_abs(dot(wi, wm) * dot(wo, wm) / denom);
					ft /= pow2(etap);
					return ft;
				}
			}
			default: return V3<R>();
		}
	}

	R pdf(const V3<R> &wo, const V3<R> &wi) const {
		switch (type) {
			case KRR_MAT_DIFFUSE:
				if (!SameHemisphere(wo, wi)) return 0;
				return m_abs(wi.z) * kInvPi<R>();
			case KRR_MAT_CONDUCTOR: { // conductor.h:80-89
				if (!SameHemisphere(wo, wi)) return 0;
				if (dist.isDelta()) return 0;
				V3<R> wm = wo + wi;
				if (wm.isZero()) return 0;
				wm = FaceForward(normalize(wm), V3<R>(0, 0, 1));
				return dist.Pdf(wo, wm) / (R(4) * AbsDot(wo, wm));
			}
			case KRR_MAT_DIELECTRIC: { // dielectric.h:190-232
				if (etaS == 1 || dist.isDelta()) return 0;
				R cosTheta_o = CosTheta(wo), cosTheta_i = CosTheta(wi);
				bool reflect = cosTheta_i * cosTheta_o > 0;
				R etap		 = 1;
				if (!reflect) etap = cosTheta_o > 0 ? etaS : (R(1) / etaS);
				V3<R> wm = wi * etap + wo;
				if (cosTheta_i == 0 || cosTheta_o == 0 || squaredNorm(wm) == 0) return 0;
				wm = FaceForward(normalize(wm), V3<R>(0, 0, 1));
				if (dot(wm, wi) * cosTheta_i < 0 || dot(wm, wo) * cosTheta_o < 0) return 0;
				R F		 = FrDielectric(CosTheta(wo), etaS);
				V3<R> Rs = baseColor * F, Ts = baseColor * (R(1) - F);
				R pr = Rs.mean(), pt = Ts.mean();
				if (pr == 0 && pt == 0) return 0;
				if (reflect) return dist.Pdf(wo, wm) / (R(4) * m_abs(dot(wo, wm))) * pr / (pr + pt);
				R denom	  = pow2(dot(wi, wm) + dot(wo, wm) / etap);
				R dwm_dwi = m_abs(dot(wi, wm)) / denom;
				return dist.Pdf(wo, wm) * dwm_dwi * pt / (pr + pt);
			}
			default: return 0;
		}
	}

	template <class S> BSDFSample<R> sample(const V3<R> &wo, S &sg) const {
		BSDFSample<R> s;
		switch (type) {
			case KRR_MAT_NULL: // null.h:23-30
				s.wi	= -wo;
				s.pdf	= 1;
				s.f		= V3<R>(R(1)) / AbsCosTheta(wo);
				s.flags = BSDF_NULL;
				return s;
			case KRR_MAT_DIFFUSE: { // diffuse.h:27-37
				R u0 = sg.get1D(), u1 = sg.get1D();
				V3<R> wi = cosineSampleHemisphere(u0, u1);
				s.wi	 = ToSameHemisphere(wi, wo);
				s.f		 = f(wo, s.wi);
				s.pdf	 = pdf(wo, s.wi);
				s.flags	 = BSDF_DIFFUSE | BSDF_REFLECTION;
				return s;
			}
			case KRR_MAT_CONDUCTOR: { // conductor.h:52-78
				if (dist.isDelta()) {
					V3<R> wi(-wo.x, -wo.y, wo.z);
					s.f		= FrComplex(AbsCosTheta(wi), eta, k) / AbsCosTheta(wi);
					s.wi	= wi;
					s.pdf	= 1;
					s.flags = BSDF_SPECULAR | BSDF_REFLECTION;
					return s;
				}
				if (wo.z == 0) return s;
				R u0 = sg.get1D(), u1 = sg.get1D();
				V3<R> wm = dist.Sample(wo, u0, u1);
				V3<R> wi = Reflect(wo, wm);
				if (!SameHemisphere(wo, wi)) return BSDFSample<R>();
				R pdfv		 = dist.Pdf(wo, wm) / (R(4) * AbsDot(wo, wm));
				R cosTheta_o = AbsCosTheta(wo), cosTheta_i = AbsCosTheta(wi);
				if (cosTheta_i == 0 || cosTheta_o == 0) return BSDFSample<R>();
				V3<R> F = FrComplex(AbsDot(wo, wm), eta, k);
				s.f		= dist.D(wm) * F * dist.G(wo, wi) / (R(4) * cosTheta_i * cosTheta_o);
				s.wi	= wi;
				s.pdf	= pdfv;
				s.flags = BSDF_GLOSSY | BSDF_REFLECTION;
				return s;
			}
			case KRR_MAT_DIELECTRIC: { // dielectric.h:131-188
				if (etaS == 1 || dist.isDelta()) {
					R F		 = FrDielectric(CosTheta(wo), etaS);
					V3<R> Rs = baseColor * F, Ts = baseColor * (R(1) - F);
					R pr = Rs.mean(), pt = Ts.mean();
					if (pr == 0 && pt == 0) return s;
					if (R(sg.get1D()) < pr / (pr + pt)) {
						V3<R> wi(-wo.x, -wo.y, wo.z);
						s.f		= Rs / AbsCosTheta(wi);
						s.wi	= wi;
						s.pdf	= pr / (pr + pt);
						s.flags = BSDF_SPECULAR | BSDF_REFLECTION;
						return s;
					}
					V3<R> wi;
					R etap;
					bool valid = Refract(wo, V3<R>(0, 0, std::copysign(R(1), wo.z)), etaS, &etap, &wi);
					if (!valid) return BSDFSample<R>();
					V3<R> ft = Ts / AbsCosTheta(wi);
					ft /= pow2(etap);
					s.f		= ft;
					s.wi	= wi;
					s.pdf	= pt / (pr + pt);
					s.flags = BSDF_SPECULAR | BSDF_TRANSMISSION;
					return s;
				}
				R u0 = sg.get1D(), u1 = sg.get1D();
				V3<R> wm = dist.Sample(wo, u0, u1);
				R F		 = FrDielectric(std::copysign(dot(wo, wm), wo.z), etaS);
				V3<R> Rs = baseColor * F, Ts = baseColor * (R(1) - F);
				R pr = Rs.mean(), pt = Ts.mean();
				if (pr == 0 && pt == 0) return s;
				if (R(sg.get1D()) < pr / (pr + pt)) {
					V3<R> wi = Reflect(wo, wm);
					if (!SameHemisphere(wo, wi)) return BSDFSample<R>();
					s.pdf	= dist.Pdf(wo, wm) / (R(4) * m_abs(dot(wo, wm))) * pr / (pr + pt);
					s.f		= dist.D(wm) * dist.G(wo, wi) * Rs / (R(4) * CosTheta(wi) * CosTheta(wo));
					s.wi	= wi;
					s.flags = BSDF_GLOSSY | BSDF_REFLECTION;
					return s;
				}
				R etap;
				V3<R> wi;
				bool tir = !Refract(wo, wm, etaS, &etap, &wi);
				if (SameHemisphere(wo, wi) || wi.z == 0 || tir) return BSDFSample<R>();
				R denom	  = pow2(dot(wi, wm) + dot(wo, wm) / etap);
				R dwm_dwi = m_abs(dot(wi, wm)) / denom;
				s.pdf	  = dist.Pdf(wo, wm) * dwm_dwi * pt / (pr + pt);
				V3<R> ft  = Ts * dist.D(wm) * dist.G(wo, wi) *
						   m_abs(dot(wi, wm) * dot(wo, wm) / (CosTheta(wi) * CosTheta(wo) * denom));
				ft /= pow2(etap);
				s.f		= ft;
				s.wi	= wi;
				s.flags = BSDF_GLOSSY | BSDF_TRANSMISSION;
				return s;
			}
			default: return s;
		}
	}
};

// ----------------------------------------------------------------------------
// Scene view: triangles, meshes, lights, media.
// ----------------------------------------------------------------------------
struct MajorantGrid {
	int res[3] = {64, 64, 64};
	float bmin[3], bmax[3];
	std::vector<float> voxels;
	float lookup(int x, int y, int z) const { return voxels[x + res[0] * (y + res[1] * z)]; }
};

struct Medium {
	const krr_medium *m = nullptr;
	float inv_voxel[3]	= {0, 0, 0};
	MajorantGrid majorant;
};

// Light list entry: DiffuseAreaLight of one emissive triangle
// (src/core/mesh.cpp:39-59) or an InfiniteLight.
struct LightRec {
	int type; // 0 area, 1 infinite
	int prim;
	float Le[3];
	float scale;
	const float *img = nullptr; // infinite light image: RGBA rows from v = 0 (null: constant)
	int iw = 0, ih = 0;
	float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}; // row-major, world from light
	bool hasUV = false;	// area light with an emissive texture (img): its mesh has texture coordinates
};

struct SceneView {
	const krr_scene_desc *d = nullptr;
	std::vector<int> triMesh;
	std::vector<int> triLight; // light index of an emissive triangle or -1
	std::vector<LightRec> lights;
	std::vector<Medium> media;
	float sceneRadius = 1e5f;
};

// Majorant grid build (src/render/media.cpp:18-75) on the dense grid.
static void buildMajorant(Medium &md) {
	const krr_medium &m = *md.m;
	MajorantGrid &g		= md.majorant;
	for (int a = 0; a < 3; a++) {
		g.bmin[a]		= m.bounds_min[a];
		g.bmax[a]		= m.bounds_max[a];
		md.inv_voxel[a] = (float) m.res[a] / (m.bounds_max[a] - m.bounds_min[a]);
	}
	g.voxels.assign((size_t) 64 * 64 * 64, 0.f);
	for (int z = 0; z < 64; z++)
		for (int y = 0; y < 64; y++)
			for (int x = 0; x < 64; x++) {
				int c[3] = {x, y, z};
				int n0[3], n1[3];
				for (int a = 0; a < 3; a++) {
					float diag = g.bmax[a] - g.bmin[a];
					float lo   = g.bmin[a] + (float(c[a]) / 64.f) * diag;
					float hi   = g.bmin[a] + (float(c[a] + 1) / 64.f) * diag;
					float ilo  = (lo - g.bmin[a]) * md.inv_voxel[a];
					float ihi  = (hi - g.bmin[a]) * md.inv_voxel[a];
					n0[a]	   = std::max((int) (ilo - 1.f), 0);
					n1[a]	   = std::min((int) (ihi + 1.f), m.res[a] - 1);
				}
				float maxValue = 0;
				for (int nz = n0[2]; nz <= n1[2]; ++nz)
					for (int ny = n0[1]; ny <= n1[1]; ++ny)
						for (int nx = n0[0]; nx <= n1[0]; ++nx)
							maxValue = std::max(maxValue, m.density[(size_t) nx + (size_t) m.res[0] * ((size_t) ny + (size_t) m.res[1] * nz)]);
				g.voxels[x + 64 * (y + 64 * z)] = maxValue;
			}
}

static void buildSceneView(const krr_scene_desc *d, SceneView &sv) {
	sv.d = d;
	sv.triMesh.assign((size_t) d->n_triangles, -1);
	sv.triLight.assign((size_t) d->n_triangles, -1);
	for (int mi = 0; mi < d->n_meshes; mi++) {
		const krr_mesh &m = d->meshes[mi];
		bool emissive	  = m.emission[0] > 0 || m.emission[1] > 0 || m.emission[2] > 0;
		float scale		  = std::max(std::max(m.emission[0], m.emission[1]), m.emission[2]);
		for (int t = 0; t < m.n_triangles; t++) {
			int prim		  = m.first_triangle + t;
			sv.triMesh[prim] = mi;
			if (emissive) {
				sv.triLight[prim] = (int) sv.lights.size();
				LightRec L{0, prim, {m.emission[0] / scale, m.emission[1] / scale, m.emission[2] / scale}, scale};
				const int et = m.material >= 0 ? d->materials[m.material].emissive_texture : 0;
				if (et > 0) { // the material's emissive texture (mesh.cpp:39-59)
					L.img	= d->textures[et - 1].rgba;
					L.iw	= d->textures[et - 1].width;
					L.ih	= d->textures[et - 1].height;
					L.hasUV = m.has_texcoords && d->texcoords;
				}
				sv.lights.push_back(L);
			}
		}
	}
	for (int i = 0; i < d->n_infinite_lights; i++) {
		const krr_infinite_light &il = d->infinite_lights[i];
		LightRec L{1, -1, {il.tint[0], il.tint[1], il.tint[2]}, il.scale};
		bool any = false;
		for (int k = 0; k < 9; k++) any |= il.rotation[k] != 0.f;
		if (any)
			for (int k = 0; k < 9; k++) L.R[k] = il.rotation[k];
		if (il.image_rgba && il.image_width > 0 && il.image_height > 0) {
			L.img = il.image_rgba;
			L.iw  = il.image_width;
			L.ih  = il.image_height;
		}
		sv.lights.push_back(L);
	}
	sv.media.resize(d->n_media);
	for (int i = 0; i < d->n_media; i++) {
		sv.media[i].m = &d->media[i];
		if (d->media[i].type == KRR_MEDIUM_GRID) buildMajorant(sv.media[i]);
	}
}

inline V3f vtx(const krr_scene_desc *d, int i) {
	return {d->positions[3 * (size_t) i], d->positions[3 * (size_t) i + 1], d->positions[3 * (size_t) i + 2]};
}
inline V3f nrm(const krr_scene_desc *d, int i) {
	return {d->normals[3 * (size_t) i], d->normals[3 * (size_t) i + 1], d->normals[3 * (size_t) i + 2]};
}

// ----------------------------------------------------------------------------
// Ray-triangle intersection: watertight test of Woop, Benthin and Wald
// (JCGT 2(1), 2013), binary32 with a binary64 recomputation of edge functions
// that come out exactly zero.  Hit iff t > 0 and (t < tbest or t == tbest with
// a lower primitive index) - the tie rule that makes the result independent of
// traversal order.
// ----------------------------------------------------------------------------
struct RayPre {
	float o[3], d[3];
	int kx, ky, kz;
	float Sx, Sy, Sz;
};

inline RayPre preprocess(const float o[3], const float d[3]) {
	RayPre r;
	for (int i = 0; i < 3; i++) { r.o[i] = o[i]; r.d[i] = d[i]; }
	int kz = 0;
	if (std::fabs(d[1]) > std::fabs(d[kz])) kz = 1;
	if (std::fabs(d[2]) > std::fabs(d[kz])) kz = 2;
	int kx = kz + 1 == 3 ? 0 : kz + 1;
	int ky = kx + 1 == 3 ? 0 : kx + 1;
	if (d[kz] < 0.0f) std::swap(kx, ky);
	r.kx = kx; r.ky = ky; r.kz = kz;
	r.Sx = d[kx] / d[kz];
	r.Sy = d[ky] / d[kz];
	r.Sz = 1.0f / d[kz];
	return r;
}

// Returns true and t, b1, b2 on a hit with t > 0.
inline bool intersectTri(const RayPre &r, const float *v0, const float *v1, const float *v2, float &t, float &b1, float &b2) {
	const float A[3] = {v0[0] - r.o[0], v0[1] - r.o[1], v0[2] - r.o[2]};
	const float B[3] = {v1[0] - r.o[0], v1[1] - r.o[1], v1[2] - r.o[2]};
	const float C[3] = {v2[0] - r.o[0], v2[1] - r.o[1], v2[2] - r.o[2]};
	const float Ax = A[r.kx] - r.Sx * A[r.kz], Ay = A[r.ky] - r.Sy * A[r.kz];
	const float Bx = B[r.kx] - r.Sx * B[r.kz], By = B[r.ky] - r.Sy * B[r.kz];
	const float Cx = C[r.kx] - r.Sx * C[r.kz], Cy = C[r.ky] - r.Sy * C[r.kz];
	float U = Cx * By - Cy * Bx;
	float V = Ax * Cy - Ay * Cx;
	float W = Bx * Ay - By * Ax;
	if (U == 0.0f || V == 0.0f || W == 0.0f) {
		U = (float) ((double) Cx * (double) By - (double) Cy * (double) Bx);
		V = (float) ((double) Ax * (double) Cy - (double) Ay * (double) Cx);
		W = (float) ((double) Bx * (double) Ay - (double) By * (double) Ax);
	}
	if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
	const float det = U + V + W;
	if (det == 0.0f) return false;
	const float Az = r.Sz * A[r.kz], Bz = r.Sz * B[r.kz], Cz = r.Sz * C[r.kz];
	const float T	= U * Az + V * Bz + W * Cz;
	const float inv = 1.0f / det;
	t				= T * inv;
	if (!(t > 0.0f)) return false;
	b1 = V * inv;
	b2 = W * inv;
	return true;
}

// Simple binned-SAH BVH2 for the CPU checker (independent of the GPU
// structure: hits only depend on the tie rule).
struct BvhNode {
	float lo[3], hi[3];
	int left = -1, right = -1; // children, or -1
	int first = 0, count = 0;  // leaf range in order[]
};

struct Bvh {
	std::vector<BvhNode> nodes;
	std::vector<int> order;
};

static void buildBvh(const krr_scene_desc *d, Bvh &bvh) {
	size_t n = (size_t) d->n_triangles;
	std::vector<float> blo(3 * n), bhi(3 * n), cen(3 * n);
	for (size_t i = 0; i < n; i++) {
		for (int a = 0; a < 3; a++) {
			float p0 = d->positions[3 * (size_t) d->indices[3 * i] + a];
			float p1 = d->positions[3 * (size_t) d->indices[3 * i + 1] + a];
			float p2 = d->positions[3 * (size_t) d->indices[3 * i + 2] + a];
			blo[3 * i + a] = std::min(p0, std::min(p1, p2));
			bhi[3 * i + a] = std::max(p0, std::max(p1, p2));
			cen[3 * i + a] = 0.5f * (blo[3 * i + a] + bhi[3 * i + a]);
		}
	}
	bvh.order.resize(n);
	for (size_t i = 0; i < n; i++) bvh.order[i] = (int) i;
	bvh.nodes.clear();
	bvh.nodes.reserve(2 * n + 1);
	struct Task { int node, first, count; };
	std::vector<Task> stack;
	bvh.nodes.push_back(BvhNode());
	stack.push_back({0, 0, (int) n});
	while (!stack.empty()) {
		Task tk = stack.back();
		stack.pop_back();
		BvhNode nd;
		float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
		for (int a = 0; a < 3; a++) { nd.lo[a] = INFINITY; nd.hi[a] = -INFINITY; }
		for (int i = tk.first; i < tk.first + tk.count; i++) {
			int p = bvh.order[i];
			for (int a = 0; a < 3; a++) {
				nd.lo[a] = std::min(nd.lo[a], blo[3 * p + a]);
				nd.hi[a] = std::max(nd.hi[a], bhi[3 * p + a]);
				clo[a]	 = std::min(clo[a], cen[3 * p + a]);
				chi[a]	 = std::max(chi[a], cen[3 * p + a]);
			}
		}
		nd.first = tk.first;
		nd.count = tk.count;
		int bestAxis = -1;
		float bestSplit = 0, bestCost = INFINITY;
		if (tk.count > 4) {
			const int NB = 16;
			for (int a = 0; a < 3; a++) {
				float ext = chi[a] - clo[a];
				if (!(ext > 0)) continue;
				int cnt[NB] = {0};
				float bl[NB][3], bh[NB][3];
				for (int b = 0; b < NB; b++)
					for (int k = 0; k < 3; k++) { bl[b][k] = INFINITY; bh[b][k] = -INFINITY; }
				for (int i = tk.first; i < tk.first + tk.count; i++) {
					int p = bvh.order[i];
					int b = std::min(NB - 1, (int) ((cen[3 * p + a] - clo[a]) / ext * NB));
					cnt[b]++;
					for (int k = 0; k < 3; k++) {
						bl[b][k] = std::min(bl[b][k], blo[3 * p + k]);
						bh[b][k] = std::max(bh[b][k], bhi[3 * p + k]);
					}
				}
				auto area = [](const float *l, const float *h) {
					float e0 = h[0] - l[0], e1 = h[1] - l[1], e2 = h[2] - l[2];
					return e0 < 0 ? 0.f : (e0 * e1 + e1 * e2 + e2 * e0);
				};
				float ra[NB];
				int rc[NB];
				float al[3] = {INFINITY, INFINITY, INFINITY}, ah[3] = {-INFINITY, -INFINITY, -INFINITY};
				int ac = 0;
				for (int b = NB - 1; b >= 1; b--) {
					for (int k = 0; k < 3; k++) { al[k] = std::min(al[k], bl[b][k]); ah[k] = std::max(ah[k], bh[b][k]); }
					ac += cnt[b];
					rc[b] = ac;
					ra[b] = area(al, ah);
				}
				for (int k = 0; k < 3; k++) { al[k] = INFINITY; ah[k] = -INFINITY; }
				ac = 0;
				for (int b = 0; b < NB - 1; b++) {
					for (int k = 0; k < 3; k++) { al[k] = std::min(al[k], bl[b][k]); ah[k] = std::max(ah[k], bh[b][k]); }
					ac += cnt[b];
					if (ac == 0 || rc[b + 1] == 0) continue;
					float cost = area(al, ah) * ac + ra[b + 1] * rc[b + 1];
					if (cost < bestCost) {
						bestCost  = cost;
						bestAxis  = a;
						bestSplit = clo[a] + ext * (float) (b + 1) / NB;
					}
				}
			}
		}
		if (bestAxis >= 0) {
			int *beg = bvh.order.data() + tk.first;
			int *mid = std::partition(beg, beg + tk.count,
									  [&](int p) { return cen[3 * p + bestAxis] < bestSplit; });
			int nl = (int) (mid - beg);
			if (nl == 0 || nl == tk.count) { // degenerate: median split
				std::nth_element(beg, beg + tk.count / 2, beg + tk.count,
								 [&](int a, int b) { return cen[3 * a + bestAxis] < cen[3 * b + bestAxis]; });
				nl = tk.count / 2;
			}
			int li = (int) bvh.nodes.size();
			bvh.nodes.push_back(BvhNode());
			bvh.nodes.push_back(BvhNode());
			nd.left	 = li;
			nd.right = li + 1;
			nd.count = 0;
			stack.push_back({li, tk.first, nl});
			stack.push_back({li + 1, tk.first + nl, tk.count - nl});
		} else if (tk.count > 8) { // no usable split axis (all centroids equal)
			int nl	 = tk.count / 2;
			int li	 = (int) bvh.nodes.size();
			bvh.nodes.push_back(BvhNode());
			bvh.nodes.push_back(BvhNode());
			nd.left	 = li;
			nd.right = li + 1;
			nd.count = 0;
			stack.push_back({li, tk.first, nl});
			stack.push_back({li + 1, tk.first + nl, tk.count - nl});
		}
		// Widen the box by a relative 2^-20 so that no hit the watertight test
		// reports (its edge decisions round) falls outside it.
		for (int a = 0; a < 3; a++) {
			float pad = (std::fabs(nd.lo[a]) + std::fabs(nd.hi[a]) + (nd.hi[a] - nd.lo[a])) * 9.5367431640625e-07f;
			nd.lo[a] -= pad;
			nd.hi[a] += pad;
		}
		bvh.nodes[tk.node] = nd;
	}
}

// Conservative slab test in binary64 (near/far plane chosen by the direction
// sign so that 0*inf only ever produces a NaN that the comparisons ignore).
inline bool boxHit(const BvhNode &n, const double o[3], const double inv[3], const int sgn[3], double tmax, double &tnear) {
	double t0 = 0.0, t1 = tmax * (1.0 + 1e-5);
	for (int a = 0; a < 3; a++) {
		double lo = sgn[a] ? n.hi[a] : n.lo[a];
		double hi = sgn[a] ? n.lo[a] : n.hi[a];
		double tn = (lo - o[a]) * inv[a];
		double tf = (hi - o[a]) * inv[a];
		if (tn > t0) t0 = tn;
		if (tf < t1) t1 = tf;
	}
	tnear = t0;
	return t0 <= t1;
}

} // namespace oracle
