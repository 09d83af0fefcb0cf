// shading.h - device shading library of the wavefront tracer.
//
// Counterparts of the reference's shading code (paths relative to the
// KiRaRay repository root):
//   PCG32 sampler            src/core/sampler.h:13-91, src/util/hash.h:12-27
//   sampling helpers         src/render/sampling.h:27-105
//   BSDF type / shading data src/render/shared.h:46-97
//   BSDFs                    src/render/bsdf.h:19-54, src/render/materials/*.h
//   lights + light sampler   src/core/light.h:152-259, src/core/shape.h:23-123,
//                            src/render/lightsampler.h:18-40
//   interactions             src/core/raytracing.h:106-180
//   surface interaction      src/render/shading.h:115-226
//   media, majorant DDA, HG  src/render/media.h:41-209, src/render/media.cpp:18-111
// All material, light and medium data lives in flat device arrays uploaded by
// the pass (the reference's TaggedPointer/VariantClass dispatch becomes a
// switch on a type field).
#pragma once

#include "texture.h"
#include "vecmath.h"

namespace krr {

// ---------------------------------------------------------------------------
// BSDF type bits (src/render/materials/bxdf.h:42-62).
// ---------------------------------------------------------------------------
enum : int {
	BSDF_UNSET = 0, BSDF_NULL = 1, BSDF_REFLECTION = 2, BSDF_TRANSMISSION = 4, BSDF_DIFFUSE = 8,
	BSDF_GLOSSY = 16, BSDF_SPECULAR = 32,
	BSDF_SMOOTH = BSDF_DIFFUSE | BSDF_GLOSSY, BSDF_DELTA = BSDF_SPECULAR | BSDF_NULL,
};

enum : int { MAT_NULL = 0, MAT_DIFFUSE = 1, MAT_DIELECTRIC = 2, MAT_CONDUCTOR = 3 };

// ---------------------------------------------------------------------------
// Device scene arrays.
// ---------------------------------------------------------------------------
struct DevMaterial {
	int type;
	float roughness;   // 1 - specular.a (SpecularGlossiness)
	float ior;
	float specTrans;
	float anisotropic;
	float diffuse[3];
	int texDiffuse, texSpecular; // textures replacing diffuse / specular.a (-1 none)
	int texNormal;				 // tangent-space normal map (-1 none)
};

struct DevTexture {
	const float4 *img; // rows from v = 0
	int w, h;
};

struct DevMesh {
	int material;	  // -1: null material
	int lightBase;	  // light index of the first triangle, -1 if not emissive
	int firstTri;
	int mediumIn, mediumOut; // interface media (-1 none); transition iff different
	int hasNormals;
	int hasUV;
	int hasTangents;
};

struct DevLight {
	int type; // 0: diffuse area light on triangle `prim`, 1: infinite
	int prim;
	float Le[3];
	float scale;
	// infinite light image (null: constant), rows from v = 0, and the
	// light's rotation (row-major, world from light)
	const float4 *img;
	int iw, ih;
	float R[9];
	int hasUV; // area light with an emissive texture: its mesh has texture coordinates
};

struct DevMedium {
	int type; // 0 homogeneous, 1 grid
	float sigma_t[3], albedo[3], Le[3];
	float g, scale;
	int res[3];
	float bmin[3], bmax[3], invVoxel[3];
	const float *density;  // res^3, x fastest
	const float *majorant; // 64^3
};

struct DevScene {
	const float *pos;	  // 3 per vertex
	const float *nrm;	  // 3 per vertex (may be null)
	const int *idx;		  // 3 per triangle
	const int *triMesh;	  // mesh of each triangle
	const DevMesh *meshes;
	const DevMaterial *materials;
	const DevLight *lights;
	int nLights;
	const DevMedium *media;
	int nMedia;
	const float *uv; // 2 per vertex (may be null)
	const float *tan; // 3 per vertex (may be null)
	const DevTexture *textures;
	float sceneRadius;
	float bmin[3], binv[3]; // scene box: p -> (p - bmin) * binv maps it to [0,1]^3
};

KRR_DEV f3 loadV(const float *a, int i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }

// ---------------------------------------------------------------------------
// PCG32 (src/core/sampler.h:13-91).  The stream increment is the same for
// every pixel of a frame ((frameIndex*spp) << 1 | 1), so only the 64-bit
// state is stored per pixel.
// ---------------------------------------------------------------------------
constexpr uint64_t kPcgMult = 0x5851f42d4c957f2dULL;

struct Pcg {
	uint64_t state, inc;
	KRR_HD uint32_t nextUint() {
		uint64_t old = state;
		state		 = old * kPcgMult + inc;
		uint32_t xs	 = (uint32_t) (((old >> 18u) ^ old) >> 27u);
		uint32_t rot = (uint32_t) (old >> 59u);
		return (xs >> rot) | (xs << ((~rot + 1u) & 31));
	}
	KRR_HD void setSeed(uint64_t initstate, uint64_t initseq) {
		state = 0u;
		inc	  = (initseq << 1u) | 1u;
		nextUint();
		state += initstate;
		nextUint();
	}
	KRR_HD void advance(int64_t delta) {
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
	KRR_HD float get1D() { return __builtin_bit_cast(float, (nextUint() >> 9) | 0x3f800000u) - 1.0f; }
};

KRR_HD uint32_t interleave32(uint32_t xv, uint32_t yv) {
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

// ---------------------------------------------------------------------------
// Sampling (src/render/sampling.h).
// ---------------------------------------------------------------------------
KRR_DEV f3 uniformSampleSphere(float u0, float u1) {
	float z = 1.f - 2.f * u0;
	float r = sqrtf(maxf(0.f, 1.f - z * z));
	float s, c;
	detmath::sincosf(k2Pi * u1, &s, &c);
	return {r * c, r * s, z};
}

KRR_DEV void uniformSampleDisk(float u0, float u1, float &dx, float &dy) {
	float ox = 2.f * u0 - 1.f, oy = 2.f * u1 - 1.f;
	if (ox == 0.f && oy == 0.f) {
		dx = 0.f;
		dy = 0.f;
		return;
	}
	float theta, r;
	if (fabsf(ox) > fabsf(oy)) {
		r	  = ox;
		theta = kPi / 4.f * (oy / ox);
	} else {
		r	  = oy;
		theta = kPi / 2.f - kPi / 4.f * (ox / oy);
	}
	float s, c;
	detmath::sincosf(theta, &s, &c);
	dx = r * c;
	dy = r * s;
}

KRR_DEV f3 cosineSampleHemisphere(float u0, float u1) {
	float dx, dy;
	uniformSampleDisk(u0, u1, dx, dy);
	float z = sqrtf(maxf(0.f, 1.f - dx * dx - dy * dy));
	return {dx, dy, z};
}

KRR_DEV float sampleExponential(float u, float a) { return -detmath::logf(1.f - u) / a; }

KRR_DEV float nextFloatDown(float v) {
	if (__builtin_isinf(v) && v < 0.f) return v;
	if (v == 0.f) v = -0.f;
	uint32_t ui = __builtin_bit_cast(uint32_t, v);
	if (v > 0) --ui;
	else ++ui;
	return __builtin_bit_cast(float, ui);
}

// (x, y, z) by an axis given as a0 = (axis == 0), a1 = (axis == 1), over
// values: a conditional over array elements or members would be a load
// through a selected address.
KRR_DEV float sel3(float x, float y, float z, bool a0, bool a1) { return a0 ? x : (a1 ? y : z); }
KRR_DEV int sel3(int x, int y, int z, bool a0, bool a1) { return a0 ? x : (a1 ? y : z); }

// The scan over the three weights is written out with constant indices: a
// loop indexing the array by its running offset keeps it in scratch memory.
KRR_DEV int sampleDiscrete3(const float w[3], float u) {
	const float w0 = w[0], w1 = w[1], w2 = w[2];
	float sumWeights = 0.f;
	sumWeights += w0;
	sumWeights += w1;
	sumWeights += w2;
	float up = u * sumWeights;
	if (up == sumWeights) up = nextFloatDown(up);
	float sum = 0.f;
	if (!(sum + w0 <= up)) return 0;
	sum += w0;
	if (!(sum + w1 <= up)) return 1;
	return 2;
}

// getPerpendicular (src/util/math_utils.h:120-130).
KRR_DEV f3 getPerpendicular(const f3 &u) {
	f3 a		 = vabs(u);
	uint32_t uyx = (a.x - a.y) < 0 ? 1 : 0;
	uint32_t uzx = (a.x - a.z) < 0 ? 1 : 0;
	uint32_t uzy = (a.y - a.z) < 0 ? 1 : 0;
	uint32_t xm	 = uyx & uzx;
	uint32_t ym	 = (1 ^ xm) & uzy;
	uint32_t zm	 = 1 ^ (xm | ym);
	return normalize(cross(u, f3((float) xm, (float) ym, (float) zm)));
}

// ---------------------------------------------------------------------------
// Microfacet / Fresnel helpers (src/render/materials/{matutils,fresnel,
// microfacet}.h).
// ---------------------------------------------------------------------------
KRR_DEV bool SameHemisphere(const f3 &w, const f3 &wp) { return w.z * wp.z > 0; }
KRR_DEV float Cos2Theta(const f3 &w) { return w.z * w.z; }
KRR_DEV float AbsCosTheta(const f3 &w) { return fabsf(w.z); }
KRR_DEV float Sin2Theta(const f3 &w) { return maxf(0.f, 1.f - Cos2Theta(w)); }
KRR_DEV float SinTheta(const f3 &w) { return sqrtf(Sin2Theta(w)); }
KRR_DEV float TanTheta(const f3 &w) { return SinTheta(w) / w.z; }
KRR_DEV float Tan2Theta(const f3 &w) { return Sin2Theta(w) / Cos2Theta(w); }
KRR_DEV float CosPhi(const f3 &w) {
	float s = SinTheta(w);
	return (s == 0) ? 1.f : clampf(w.x / s, -1.f, 1.f);
}
KRR_DEV float SinPhi(const f3 &w) {
	float s = SinTheta(w);
	return (s == 0) ? 0.f : clampf(w.y / s, -1.f, 1.f);
}
KRR_DEV float AbsDot(const f3 &a, const f3 &b) { return fabsf(dot(a, b)); }
KRR_DEV f3 Reflect(const f3 &wo, const f3 &n) { return -wo + 2.f * dot(wo, n) * n; }
KRR_DEV f3 FaceForwardZ(const f3 &w) { return w.z > 0 ? w : -w; } // FaceForward(w, (0,0,1))
KRR_DEV f3 ToSameHemisphere(const f3 &w, const f3 &wp) { return {w.x, w.y, w.z * wp.z > 0 ? w.z : -w.z}; }

KRR_DEV bool Refract(const f3 &wi, const f3 &n, float eta, float *etap, f3 *wt) {
	float cosThetaI = dot(n, wi);
	if (wi.z < 0) eta = 1.f / eta;
	float sin2ThetaI = maxf(0.f, 1.f - pow2(cosThetaI));
	float sin2ThetaT = sin2ThetaI / pow2(eta);
	if (sin2ThetaT >= 1) return false;
	float cosThetaT = sqrtf(1.f - sin2ThetaT);
	*wt				= -wi / eta + (cosThetaI / eta - cosThetaT) * n;
	*etap			= eta;
	return true;
}

KRR_DEV float FrDielectric(float cosTheta_i, float eta) {
	cosTheta_i = clampf(cosTheta_i, -1.f, 1.f);
	if (cosTheta_i < 0) {
		eta		   = 1.f / eta;
		cosTheta_i = -cosTheta_i;
	}
	float sin2Theta_i = 1.f - pow2(cosTheta_i);
	float sin2Theta_t = sin2Theta_i / pow2(eta);
	if (sin2Theta_t >= 1) return 1.f;
	float cosTheta_t = sqrtf(maxf(1.f - sin2Theta_t, 0.f));
	float r_parl	 = (eta * cosTheta_i - cosTheta_t) / (eta * cosTheta_i + cosTheta_t);
	float r_perp	 = (cosTheta_i - eta * cosTheta_t) / (cosTheta_i + eta * cosTheta_t);
	return (pow2(r_parl) + pow2(r_perp)) / 2.f;
}

// Complex<float> of src/core/math/include/krrmath/complex.h.
struct cpx {
	float re, im;
	KRR_DEV cpx operator+(cpx z) const { return {re + z.re, im + z.im}; }
	KRR_DEV cpx operator-(cpx z) const { return {re - z.re, im - z.im}; }
	KRR_DEV cpx operator*(cpx z) const { return {re * z.re - im * z.im, re * z.im + im * z.re}; }
	KRR_DEV cpx operator/(cpx z) const {
		float scale = 1.f / (z.re * z.re + z.im * z.im);
		return {scale * (re * z.re + im * z.im), scale * (im * z.re - re * z.im)};
	}
	KRR_DEV float norm() const { return re * re + im * im; }
	KRR_DEV cpx csqrt() const {
		float n = sqrtf(norm()), t1 = sqrtf(.5f * (n + fabsf(re))), t2 = .5f * im / t1;
		if (n == 0) return {0.f, 0.f};
		if (re >= 0) return {t1, t2};
		return {fabsf(t2), copysignf(t1, im)};
	}
};

KRR_DEV float FrComplex1(float cosTheta_i, cpx eta) {
	cosTheta_i		  = clampf(cosTheta_i, 0.f, 1.f);
	float sin2Theta_i = 1.f - pow2(cosTheta_i);
	cpx sin2Theta_t	  = cpx{sin2Theta_i, 0.f} / (eta * eta);
	cpx cosTheta_t	  = (cpx{1.f, 0.f} - sin2Theta_t).csqrt();
	cpx ci{cosTheta_i, 0.f};
	cpx r_parl = (eta * ci - cosTheta_t) / (eta * ci + cosTheta_t);
	cpx r_perp = (ci - eta * cosTheta_t) / (ci + eta * cosTheta_t);
	return (r_parl.norm() + r_perp.norm()) / 2.f;
}

KRR_DEV f3 FrComplex(float c, const f3 &eta, const f3 &k) {
	return {FrComplex1(c, {eta.x, k.x}), FrComplex1(c, {eta.y, k.y}), FrComplex1(c, {eta.z, k.z})};
}

struct GGX {
	float ax, ay;
	KRR_DEV void init(float a, float b) {
		ax = maxf(1e-3f, a);
		ay = maxf(1e-3f, b);
	}
	KRR_DEV bool isDelta() const { return maxf(ax, ay) <= 1e-3f; }
	KRR_DEV float Lambda(const f3 &w) const {
		float absTanTheta = fabsf(TanTheta(w));
		if (__builtin_isinf(absTanTheta)) return 0.f;
		float cp = CosPhi(w), sp = SinPhi(w);
		float alpha			  = sqrtf((cp * cp) * ax * ax + (sp * sp) * ay * ay);
		float alpha2Tan2Theta = pow2(alpha * absTanTheta);
		return (-1.f + sqrtf(1.f + alpha2Tan2Theta)) / 2.f;
	}
	KRR_DEV float G1(const f3 &w) const { return 1.f / (1.f + Lambda(w)); }
	KRR_DEV float G(const f3 &wo, const f3 &wi) const { return 1.f / (1.f + Lambda(wo) + Lambda(wi)); }
	KRR_DEV float D(const f3 &wh) const {
		float tan2Theta = Tan2Theta(wh);
		if (__builtin_isinf(tan2Theta)) return 0.f;
		const float cos4Theta = pow2(Cos2Theta(wh));
		float e				  = (pow2(CosPhi(wh) / ax) + pow2(SinPhi(wh) / ay)) * tan2Theta;
		return 1.f / (kPi * ax * ay * cos4Theta * pow2(1.f + e));
	}
	KRR_DEV float Pdf(const f3 &wo, const f3 &wh) const { return D(wh) * G1(wo) * fabsf(dot(wo, wh)) / AbsCosTheta(wo); }
	// GGXSample11 + GGXSample (microfacet.h:86-154).
	KRR_DEV static void sample11(float cosTheta, float U1, float U2, float *slope_x, float *slope_y) {
		if (cosTheta > .9999f) {
			float r = sqrtf(U1 / (1.f - U1));
			float s, c;
			detmath::sincosf(k2Pi * U2, &s, &c);
			*slope_x = r * c;
			*slope_y = r * s;
			return;
		}
		float sinTheta = sqrtf(maxf(0.f, 1.f - cosTheta * cosTheta));
		float tanTheta = sinTheta / cosTheta;
		float a		   = 1.f / tanTheta;
		float G1v	   = 2.f / (1.f + sqrtf(1.f + 1.f / (a * a)));
		float A		   = 2.f * U1 / G1v - 1.f;
		float tmp	   = 1.f / (A * A - 1.f);
		if ((double) tmp > 1e10) tmp = 1e10f;
		float B			= tanTheta;
		float Dv		= sqrtf(maxf(B * B * tmp * tmp - (A * A - B * B) * tmp, 0.f));
		float slope_x_1 = B * tmp - Dv;
		float slope_x_2 = B * tmp + Dv;
		*slope_x		= (A < 0 || slope_x_2 > 1.f / tanTheta) ? slope_x_1 : slope_x_2;
		float S;
		if (U2 > 0.5f) {
			S  = 1.f;
			U2 = 2.f * (U2 - .5f);
		} else {
			S  = -1.f;
			U2 = 2.f * (.5f - U2);
		}
		float z = (U2 * (U2 * (U2 * 0.27385f - 0.73369f) + 0.46341f)) /
				  (U2 * (U2 * (U2 * 0.093073f + 0.309420f) - 1.000000f) + 0.597999f);
		*slope_y = S * z * sqrtf(1.f + *slope_x * *slope_x);
	}
	KRR_DEV f3 sampleStretched(const f3 &wi, float U1, float U2) const {
		f3 ws = normalize(f3(ax * wi.x, ay * wi.y, wi.z));
		float slope_x, slope_y;
		sample11(ws.z, U1, U2, &slope_x, &slope_y);
		float cp = CosPhi(ws), sp = SinPhi(ws);
		float tmp = cp * slope_x - sp * slope_y;
		slope_y	  = sp * slope_x + cp * slope_y;
		slope_x	  = tmp;
		slope_x	  = ax * slope_x;
		slope_y	  = ay * slope_y;
		return normalize(f3(-slope_x, -slope_y, 1.f));
	}
	KRR_DEV f3 Sample(const f3 &wo, float u0, float u1) const {
		bool flip = wo.z < 0;
		f3 wh	  = sampleStretched(flip ? -wo : wo, u0, u1);
		return flip ? -wh : wh;
	}
};

// ---------------------------------------------------------------------------
// BSDF over the shading data of one surface interaction.
// ---------------------------------------------------------------------------
struct BSDFSample {
	f3 f, wi;
	float pdf = 0.f;
	int flags = BSDF_UNSET;
};

// BSDFData::getBsdfType (src/render/shared.h:46-73).
KRR_DEV int bsdfTypeOf(int matType, float roughness) {
	switch (matType) {
		case MAT_NULL: return BSDF_NULL;
		case MAT_DIFFUSE: return BSDF_DIFFUSE | BSDF_REFLECTION;
		case MAT_DIELECTRIC: return (roughness <= 1e-3f ? BSDF_SPECULAR : BSDF_GLOSSY) | BSDF_REFLECTION | BSDF_TRANSMISSION;
		case MAT_CONDUCTOR: return BSDF_REFLECTION | (roughness <= 1e-3f ? BSDF_SPECULAR : BSDF_GLOSSY);
		default: return BSDF_UNSET;
	}
}

struct BSDF {
	int type;
	f3 color;	// diffuse albedo, conductor reflectance or dielectric base colour
	f3 eta, k;	// conductor
	float etaS; // dielectric eta
	GGX dist;

	// setup() of DiffuseBrdf / ConductorBsdf / DielectricBsdf.  Every field
	// is written once, after the branches: field stores that differ per
	// branch get merged into one store through a selected address, which
	// keeps the whole object in scratch memory instead of registers.
	KRR_DEV void setup(const DevMaterial &m) {
		const f3 d(m.diffuse[0], m.diffuse[1], m.diffuse[2]);
		f3 c = d, e, kk;
		float eS = 1.f, ax = 1.f, ay = 1.f;
		if (m.type == MAT_CONDUCTOR) {
			c			 = vmins(d, 0.9999f);
			float aspect = sqrtf(1.f - m.anisotropic * .9f);
			ax			 = maxf(.001f, pow2(m.roughness) / aspect);
			ay			 = maxf(.001f, pow2(m.roughness) * aspect);
			e			 = f3(m.ior);
			kk			 = (2.f * vsqrt(c)) / vsqrt(vmaxs(f3(1.f) - c, 0.f));
		} else if (m.type == MAT_DIELECTRIC) {
			c  = d * m.specTrans;
			eS = m.ior;
			ax = ay = pow2(m.roughness);
		}
		type  = m.type;
		color = c;
		eta	  = e;
		k	  = kk;
		etaS  = eS;
		dist.init(ax, ay);
	}

	KRR_DEV f3 f(const f3 &wo, const f3 &wi) const {
		if (type == MAT_DIFFUSE) {
			if (!SameHemisphere(wo, wi)) return f3();
			return color * kInvPi;
		}
		if (type == MAT_CONDUCTOR) {
			if (!SameHemisphere(wo, wi)) return f3();
			if (dist.isDelta()) return f3();
			float cosTheta_o = AbsCosTheta(wo), cosTheta_i = AbsCosTheta(wi);
			if (cosTheta_i == 0 || cosTheta_o == 0) return f3();
			f3 wm = wi + wo;
			if (sqnorm(wm) == 0) return f3();
			wm	 = normalize(wm);
			f3 F = FrComplex(AbsDot(wo, wm), eta, k);
			return dist.D(wm) * F * dist.G(wo, wi) / (4.f * cosTheta_i * cosTheta_o);
		}
		if (type == MAT_DIELECTRIC) {
			if (etaS == 1 || dist.isDelta()) return f3();
			float cosTheta_o = wo.z, cosTheta_i = wi.z;
			bool reflect = cosTheta_i * cosTheta_o > 0;
			float etap	 = 1.f;
			if (!reflect) etap = cosTheta_o > 0 ? etaS : (1.f / etaS);
			f3 wm = wi * etap + wo;
			if (cosTheta_i == 0 || cosTheta_o == 0 || !wm.any()) return f3();
			wm = FaceForwardZ(normalize(wm));
			if (dot(wm, wi) * cosTheta_i < 0 || dot(wm, wo) * cosTheta_o < 0) return f3();
			float F = FrDielectric(copysignf(dot(wo, wm), wo.z), etaS);
			f3 Rs = color * F, Ts = color * (1.f - F);
			if (reflect) return dist.D(wm) * dist.G(wo, wi) * Rs / fabsf(4.f * cosTheta_i * cosTheta_o);
			float denom = pow2(dot(wi, wm) + dot(wo, wm) / etap) * cosTheta_i * cosTheta_o;
			f3 ft		= Ts * dist.D(wm) * dist.G(wo, wi) * fabsf(dot(wi, wm) * dot(wo, wm) / denom);
			ft			= ft / pow2(etap);
			return ft;
		}
		return f3();
	}

	KRR_DEV float pdf(const f3 &wo, const f3 &wi) const {
		if (type == MAT_DIFFUSE) {
			if (!SameHemisphere(wo, wi)) return 0.f;
			return fabsf(wi.z) * kInvPi;
		}
		if (type == MAT_CONDUCTOR) {
			if (!SameHemisphere(wo, wi)) return 0.f;
			if (dist.isDelta()) return 0.f;
			f3 wm = wo + wi;
			if (wm.isZero()) return 0.f;
			wm = FaceForwardZ(normalize(wm));
			return dist.Pdf(wo, wm) / (4.f * AbsDot(wo, wm));
		}
		if (type == MAT_DIELECTRIC) {
			if (etaS == 1 || dist.isDelta()) return 0.f;
			float cosTheta_o = wo.z, cosTheta_i = wi.z;
			bool reflect = cosTheta_i * cosTheta_o > 0;
			float etap	 = 1.f;
			if (!reflect) etap = cosTheta_o > 0 ? etaS : (1.f / etaS);
			f3 wm = wi * etap + wo;
			if (cosTheta_i == 0 || cosTheta_o == 0 || sqnorm(wm) == 0) return 0.f;
			wm = FaceForwardZ(normalize(wm));
			if (dot(wm, wi) * cosTheta_i < 0 || dot(wm, wo) * cosTheta_o < 0) return 0.f;
			float F = FrDielectric(wo.z, etaS);
			f3 Rs = color * F, Ts = color * (1.f - F);
			float pr = Rs.mean(), pt = Ts.mean();
			if (pr == 0 && pt == 0) return 0.f;
			if (reflect) return dist.Pdf(wo, wm) / (4.f * fabsf(dot(wo, wm))) * pr / (pr + pt);
			float denom	  = pow2(dot(wi, wm) + dot(wo, wm) / etap);
			float dwm_dwi = fabsf(dot(wi, wm)) / denom;
			return dist.Pdf(wo, wm) * dwm_dwi * pt / (pr + pt);
		}
		return 0.f;
	}

	KRR_DEV BSDFSample sample(const f3 &wo, Pcg &sg) const {
		BSDFSample s;
		if (type == MAT_DIFFUSE) {
			float u0 = sg.get1D(), u1 = sg.get1D();
			f3 wi	= cosineSampleHemisphere(u0, u1);
			s.wi	= ToSameHemisphere(wi, wo);
			s.f		= f(wo, s.wi);
			s.pdf	= pdf(wo, s.wi);
			s.flags = BSDF_DIFFUSE | BSDF_REFLECTION;
			return s;
		}
		if (type == MAT_CONDUCTOR) {
			if (dist.isDelta()) {
				f3 wi(-wo.x, -wo.y, wo.z);
				s.f		= FrComplex(AbsCosTheta(wi), eta, k) / AbsCosTheta(wi);
				s.wi	= wi;
				s.pdf	= 1.f;
				s.flags = BSDF_SPECULAR | BSDF_REFLECTION;
				return s;
			}
			if (wo.z == 0) return s;
			float u0 = sg.get1D(), u1 = sg.get1D();
			f3 wm = dist.Sample(wo, u0, u1);
			f3 wi = Reflect(wo, wm);
			if (!SameHemisphere(wo, wi)) return BSDFSample();
			float pdfv		 = dist.Pdf(wo, wm) / (4.f * AbsDot(wo, wm));
			float cosTheta_o = AbsCosTheta(wo), cosTheta_i = AbsCosTheta(wi);
			if (cosTheta_i == 0 || cosTheta_o == 0) return BSDFSample();
			f3 F	= FrComplex(AbsDot(wo, wm), eta, k);
			s.f		= dist.D(wm) * F * dist.G(wo, wi) / (4.f * cosTheta_i * cosTheta_o);
			s.wi	= wi;
			s.pdf	= pdfv;
			s.flags = BSDF_GLOSSY | BSDF_REFLECTION;
			return s;
		}
		if (type == MAT_DIELECTRIC) {
			if (etaS == 1 || dist.isDelta()) {
				float F = FrDielectric(wo.z, etaS);
				f3 Rs = color * F, Ts = color * (1.f - F);
				float pr = Rs.mean(), pt = Ts.mean();
				if (pr == 0 && pt == 0) return s;
				if (sg.get1D() < pr / (pr + pt)) {
					f3 wi(-wo.x, -wo.y, wo.z);
					s.f		= Rs / AbsCosTheta(wi);
					s.wi	= wi;
					s.pdf	= pr / (pr + pt);
					s.flags = BSDF_SPECULAR | BSDF_REFLECTION;
					return s;
				}
				f3 wi;
				float etap;
				if (!Refract(wo, f3(0.f, 0.f, copysignf(1.f, wo.z)), etaS, &etap, &wi)) return BSDFSample();
				f3 ft	= Ts / AbsCosTheta(wi);
				ft		= ft / pow2(etap);
				s.f		= ft;
				s.wi	= wi;
				s.pdf	= pt / (pr + pt);
				s.flags = BSDF_SPECULAR | BSDF_TRANSMISSION;
				return s;
			}
			float u0 = sg.get1D(), u1 = sg.get1D();
			f3 wm	= dist.Sample(wo, u0, u1);
			float F = FrDielectric(copysignf(dot(wo, wm), wo.z), etaS);
			f3 Rs = color * F, Ts = color * (1.f - F);
			float pr = Rs.mean(), pt = Ts.mean();
			if (pr == 0 && pt == 0) return s;
			if (sg.get1D() < pr / (pr + pt)) {
				f3 wi = Reflect(wo, wm);
				if (!SameHemisphere(wo, wi)) return BSDFSample();
				s.pdf	= dist.Pdf(wo, wm) / (4.f * fabsf(dot(wo, wm))) * pr / (pr + pt);
				s.f		= dist.D(wm) * dist.G(wo, wi) * Rs / (4.f * wi.z * wo.z);
				s.wi	= wi;
				s.flags = BSDF_GLOSSY | BSDF_REFLECTION;
				return s;
			}
			float etap;
			f3 wi;
			bool tir = !Refract(wo, wm, etaS, &etap, &wi);
			if (SameHemisphere(wo, wi) || wi.z == 0 || tir) return BSDFSample();
			float denom	  = pow2(dot(wi, wm) + dot(wo, wm) / etap);
			float dwm_dwi = fabsf(dot(wi, wm)) / denom;
			s.pdf		  = dist.Pdf(wo, wm) * dwm_dwi * pt / (pr + pt);
			f3 ft		  = Ts * dist.D(wm) * dist.G(wo, wi) * fabsf(dot(wi, wm) * dot(wo, wm) / (wi.z * wo.z * denom));
			ft			  = ft / pow2(etap);
			s.f			  = ft;
			s.wi		  = wi;
			s.flags		  = BSDF_GLOSSY | BSDF_TRANSMISSION;
			return s;
		}
		// null material (null.h:23-30)
		s.wi	= -wo;
		s.pdf	= 1.f;
		s.f		= f3(1.f) / AbsCosTheta(wo);
		s.flags = BSDF_NULL;
		return s;
	}
};

// ---------------------------------------------------------------------------
// Interactions (src/core/raytracing.h:106-180).
// ---------------------------------------------------------------------------
KRR_DEV f3 offsetRayOrigin(const f3 &p, const f3 &n, const f3 &w) {
	f3 offset = n * kRayEps;
	if (dot(n, w) < 0.f) offset = -offset;
	return p + offset;
}

struct SurfaceInteraction {
	f3 p, wo, n, tangent, bitangent;
	int material; // -1: null
	int light;	  // -1: none
	int medium;	  // medium of the incoming ray
	int mIn, mOut;
	bool interface_;
	float u, v; // texture coordinates (0 without them)

	KRR_DEV f3 toLocal(const f3 &v) const { return {dot(tangent, v), dot(bitangent, v), dot(n, v)}; }
	KRR_DEV f3 toWorld(const f3 &v) const { return tangent * v.x + bitangent * v.y + n * v.z; }
	// Values, not a conditional over the members: a conditional over
	// lvalues is a load through a selected address, which keeps the whole
	// interaction in scratch memory.
	KRR_DEV int getMedium(const f3 &w) const {
		const int in = mIn, out = mOut, own = medium;
		if (interface_) return dot(w, n) > 0.f ? out : in;
		return own;
	}
};

// prepareSurfaceInteraction (src/render/shading.h:115-226) for an identity
// instance transform; b1, b2 are the barycentrics of vertices 1 and 2.  The
// tangent is the mesh's interpolated tangent made orthogonal to n when the
// mesh has tangents (shading.h:140-145), else getPerpendicular(n); a normal
// map (shading.h:187-203) then tilts the frame on meshes with texture
// coordinates.  TEX = false: no mesh has tangents and no material a normal
// map (the lookups are compiled out).
template <bool TEX = true>
KRR_DEV void prepareSurfaceInteraction(const DevScene &sc, int prim, float b1, float b2, const f3 &rayDir, int rayMedium,
									   SurfaceInteraction &it) {
	const DevMesh m = sc.meshes[sc.triMesh[prim]];
	float b0		= 1.f - b1 - b2;
	int v0 = sc.idx[3 * prim], v1 = sc.idx[3 * prim + 1], v2 = sc.idx[3 * prim + 2];
	it.medium = rayMedium;
	it.wo	  = normalize(-normalize(rayDir));
	f3 p0 = loadV(sc.pos, v0), p1 = loadV(sc.pos, v1), p2 = loadV(sc.pos, v2);
	it.p = b0 * p0 + b1 * p1 + b2 * p2;
	if (m.hasNormals) it.n = normalize(b0 * loadV(sc.nrm, v0) + b1 * loadV(sc.nrm, v1) + b2 * loadV(sc.nrm, v2));
	else it.n = normalize(cross(p1 - p0, p2 - p0));
	if (TEX && m.hasTangents) {
		it.tangent = normalize(b0 * loadV(sc.tan, v0) + b1 * loadV(sc.tan, v1) + b2 * loadV(sc.tan, v2));
		it.tangent = normalize(it.tangent - it.n * dot(it.n, it.tangent));
	} else it.tangent = getPerpendicular(it.n);
	it.bitangent  = normalize(cross(it.n, it.tangent));
	it.light	  = m.lightBase >= 0 ? m.lightBase + (prim - m.firstTri) : -1;
	it.material	  = m.material;
	it.interface_ = m.mediumIn != m.mediumOut;
	it.mIn		  = m.mediumIn;
	it.mOut		  = m.mediumOut;
	it.n		  = normalize(it.n);
	it.tangent	  = normalize(it.tangent);
	it.bitangent  = normalize(it.bitangent);
	it.u = it.v = 0.f;
	if (m.hasUV) { // shading.h:147-149: b0 t0 + b1 t1 + b2 t2
		it.u = b0 * sc.uv[2 * v0] + b1 * sc.uv[2 * v1] + b2 * sc.uv[2 * v2];
		it.v = b0 * sc.uv[2 * v0 + 1] + b1 * sc.uv[2 * v1 + 1] + b2 * sc.uv[2 * v2 + 1];
		if (TEX && it.material >= 0) {
			const int tn = sc.materials[it.material].texNormal;
			if (tn >= 0) { // rgbToNormal (shading.h:32): 2 rgb - 1, in the (T, B, N) frame
				const DevTexture t = sc.textures[tn];
				const float4 c	   = texBilinear4(t.img, t.w, t.h, it.u, it.v);
				const float nx = 2.f * c.x - 1.f, ny = 2.f * c.y - 1.f, nz = 2.f * c.z - 1.f;
				it.n		 = normalize(it.tangent * nx + it.bitangent * ny + it.n * nz);
				it.tangent	 = normalize(it.tangent - it.n * dot(it.tangent, it.n));
				it.bitangent = normalize(cross(it.n, it.tangent));
			}
		}
	}
}

// ---------------------------------------------------------------------------
// Lights.
// ---------------------------------------------------------------------------
struct LightSample {
	f3 p, n; // n = 0 for infinite lights
	f3 L;
	float pdf;
};

KRR_DEV float triArea(const DevScene &sc, int prim) {
	f3 p0 = loadV(sc.pos, sc.idx[3 * prim]), p1 = loadV(sc.pos, sc.idx[3 * prim + 1]), p2 = loadV(sc.pos, sc.idx[3 * prim + 2]);
	return 0.5f * length(cross(p1 - p0, p2 - p0));
}

// DiffuseAreaLight::L (light.h:188-193): one-sided; scale * the emissive
// texture's rgb at (u, v) when the light has one, else scale * Le.
// TEX = false: no light of the scene has a texture.
template <bool TEX = true> KRR_DEV f3 areaLightL(const DevLight &L, const f3 &n, const f3 &w, float u, float v) {
	if (dot(n, w) < 0.f) return f3(); // one-sided
	if (TEX && L.img) return L.scale * texBilinear(L.img, L.iw, L.ih, u, v);
	return L.scale * f3(L.Le[0], L.Le[1], L.Le[2]);
}

// The material at a hit with its textures applied (shading.h:187-226,
// SpecularGlossiness): the diffuse texture's rgb replaces the diffuse
// colour, the specular texture's alpha the glossiness (roughness = 1 - a).
// TEX = false: the scene has no textures (the plain material).
template <bool TEX = true> KRR_DEV DevMaterial materialAt(const DevScene &sc, const SurfaceInteraction &it) {
	DevMaterial m = sc.materials[it.material];
	if (!TEX) return m;
	if (m.texDiffuse >= 0) {
		const DevTexture t = sc.textures[m.texDiffuse];
		const float4 c	   = texBilinear4(t.img, t.w, t.h, it.u, it.v);
		m.diffuse[0]	   = c.x;
		m.diffuse[1]	   = c.y;
		m.diffuse[2]	   = c.z;
	}
	if (m.texSpecular >= 0) {
		const DevTexture t = sc.textures[m.texSpecular];
		m.roughness		   = 1.f - texBilinear4(t.img, t.w, t.h, it.u, it.v).w;
	}
	return m;
}

// InfiniteLight::Li (light.h:241-246): scale * tint * image(worldToLatLong(
// R^T w)), worldToLatLong of math_utils.h:133-139.
// TEX = false: no infinite light of the scene has an image.
template <bool TEX = true> KRR_DEV f3 infiniteLi(const DevLight &L, const f3 &w) {
	f3 c(L.Le[0], L.Le[1], L.Le[2]);
	if (TEX && L.img) {
		const float *R = L.R;
		const f3 p	   = normalize(f3(R[0] * w.x + R[3] * w.y + R[6] * w.z, R[1] * w.x + R[4] * w.y + R[7] * w.z, R[2] * w.x + R[5] * w.y + R[8] * w.z));
		const float u  = detmath::atan2f(p.x, -p.z) * kInv2Pi + 0.5f;
		const float v  = detmath::acosf(p.y) * kInvPi;
		c			   = c * texBilinear(L.img, L.iw, L.ih, u, v);
	}
	return L.scale * c;
}

// Light::sampleLi for a DiffuseAreaLight (light.h:167-180 + shape.h:54-108)
// or an InfiniteLight (light.h:226-235).
template <bool TEX = true> KRR_DEV LightSample sampleLi(const DevScene &sc, const DevLight &L, float u0, float u1, const f3 &ctxP) {
	LightSample ls;
	if (L.type == 1) {
		f3 wi  = uniformSampleSphere(u0, u1);
		ls.p   = ctxP + wi * 2.f * sc.sceneRadius;
		ls.n   = f3();
		ls.L   = infiniteLi<TEX>(L, wi);
		ls.pdf = kInv4Pi;
		return ls;
	}
	const int prim = L.prim;
	const int i0 = sc.idx[3 * prim], i1 = sc.idx[3 * prim + 1], i2 = sc.idx[3 * prim + 2];
	f3 p0 = loadV(sc.pos, i0), p1 = loadV(sc.pos, i1), p2 = loadV(sc.pos, i2);
	float b0, b1;
	if (u0 < u1) {
		b0 = u0 / 2.f;
		b1 = u1 - b0;
	} else {
		b1 = u1 / 2.f;
		b0 = u0 - b1;
	}
	float b2 = 1.f - b0 - b1;
	f3 p	 = b0 * p0 + b1 * p1 + b2 * p2;
	f3 n	 = normalize(cross(p1 - p0, p2 - p0));
	f3 ns	 = normalize(b0 * loadV(sc.nrm, i0) + b1 * loadV(sc.nrm, i1) + b2 * loadV(sc.nrm, i2));
	if (dot(n, ns) < 0.f) n = n * -1.f;
	n		  = normalize(n);
	float pdf = 1.f / triArea(sc, prim);
	f3 wi	  = normalize(p - ctxP);
	pdf		  = pdf / (fabsf(dot(n, wi)) / sqnorm(ctxP - p));
	if (length(wi) == 0.f || __builtin_isinf(pdf)) pdf = 0.f;
	ls.p   = p;
	ls.n   = n;
	ls.pdf = pdf;
	float su = 0.f, sv = 0.f; // Triangle::sample's uv (shape.h:73-78)
	if (TEX && L.img && L.hasUV) {
		su = b0 * sc.uv[2 * i0] + b1 * sc.uv[2 * i1] + b2 * sc.uv[2 * i2];
		sv = b0 * sc.uv[2 * i0 + 1] + b1 * sc.uv[2 * i1 + 1] + b2 * sc.uv[2 * i2 + 1];
	}
	ls.L = areaLightL<TEX>(L, n, normalize(ctxP - p), su, sv);
	return ls;
}

// Triangle::pdf(sample, ctx) (shape.h:114-123) / InfiniteLight::pdfLi.
KRR_DEV float lightPdfLi(const DevScene &sc, const DevLight &L, const f3 &p, const f3 &n, const f3 &ctxP) {
	if (L.type == 1) return kInv4Pi;
	f3 wi	  = normalize(p - ctxP);
	float pdf = (1.f / triArea(sc, L.prim)) / (fabsf(dot(n, -wi)) / sqnorm(ctxP - p));
	if (length(wi) == 0.f || __builtin_isinf(pdf)) pdf = 0.f;
	return pdf;
}

// ---------------------------------------------------------------------------
// Media.
// ---------------------------------------------------------------------------
struct MediumProps {
	f3 sigma_a, sigma_s, Le;
	float g;
};

KRR_DEV float gridDensity(const DevMedium &m, const f3 &p) {
	float idx[3] = {(p.x - m.bmin[0]) * m.invVoxel[0], (p.y - m.bmin[1]) * m.invVoxel[1], (p.z - m.bmin[2]) * m.invVoxel[2]};
	int i0[3];
	float w[3];
	for (int a = 0; a < 3; a++) {
		float f = floorf(idx[a]);
		i0[a]	= (int) f;
		w[a]	= idx[a] - f;
	}
	auto at = [&](int x, int y, int z) -> float {
		if (x < 0 || y < 0 || z < 0 || x >= m.res[0] || y >= m.res[1] || z >= m.res[2]) return 0.f;
		return m.density[(size_t) x + (size_t) m.res[0] * ((size_t) y + (size_t) m.res[1] * z)];
	};
	auto lerp = [](float a, float b, float t) { return a + t * (b - a); };
	int x = i0[0], y = i0[1], z = i0[2];
	float c00 = lerp(at(x, y, z), at(x, y, z + 1), w[2]);
	float c01 = lerp(at(x, y + 1, z), at(x, y + 1, z + 1), w[2]);
	float c10 = lerp(at(x + 1, y, z), at(x + 1, y, z + 1), w[2]);
	float c11 = lerp(at(x + 1, y + 1, z), at(x + 1, y + 1, z + 1), w[2]);
	return lerp(lerp(c00, c01, w[1]), lerp(c10, c11, w[1]), w[0]);
}

// Medium::samplePoint (media.h:121-126 homogeneous, 161-172 grid).
KRR_DEV MediumProps samplePoint(const DevMedium &m, const f3 &p) {
	MediumProps mp;
	f3 st(m.sigma_t[0], m.sigma_t[1], m.sigma_t[2]);
	f3 al(m.albedo[0], m.albedo[1], m.albedo[2]);
	mp.g = m.g;
	if (m.type == 0) {
		f3 ss	   = st * al;
		mp.sigma_a = st - ss;
		mp.sigma_s = ss;
		mp.Le	   = f3(m.Le[0], m.Le[1], m.Le[2]);
		return mp;
	}
	float dens = gridDensity(m, p);
	f3 sts	   = dens * m.scale * st;
	f3 ss	   = sts * al;
	mp.sigma_a = sts - ss;
	mp.sigma_s = ss;
	mp.Le	   = f3();
	return mp;
}

struct MajorantSeg {
	float tMin, tMax;
	f3 sigma_maj;
};

// MajorantIterator (media.h:41-106).
struct MajorantIter {
	float tMin, tMax;
	f3 sigma_t;
	const float *grid; // 64^3 majorants or null (homogeneous)
	float nextCrossingT[3], deltaT[3];
	int step[3], voxelLimit[3], voxel[3];

	KRR_DEV void initEmpty() {
		tMin = kFloatMax;
		tMax = 1.17549435e-38f; // numeric_limits<float>::min()
		grid = nullptr;
	}
	KRR_DEV void init(const f3 &o, const f3 &d, float tmin, float tmax, const f3 &st, const DevMedium *m) {
		tMin	= tmin;
		tMax	= tmax;
		sigma_t = st;
		grid	= m ? m->majorant : nullptr;
		if (!grid) return;
		float oo[3], dd[3];
		for (int a = 0; a < 3; a++) {
			float diag = m->bmax[a] - m->bmin[a];
			oo[a]	   = (o[a] - m->bmin[a]) / (diag == 0.f ? 1.f : diag);
			dd[a]	   = d[a] / diag;
		}
		for (int axis = 0; axis < 3; ++axis) {
			float gi			= oo[axis] + dd[axis] * tMin;
			voxel[axis]			= (int) clampf(gi * 64.f, 0.f, 64.f - 1.f);
			deltaT[axis]		= 1.f / (fabsf(dd[axis]) * 64.f);
			if (dd[axis] == -0.f) dd[axis] = 0.f;
			if (dd[axis] >= 0) {
				float nextVoxelPos	= (float) (voxel[axis] + 1) / 64.f;
				nextCrossingT[axis] = tMin + (nextVoxelPos - gi) / dd[axis];
				step[axis]			= 1;
				voxelLimit[axis]	= 64;
			} else {
				float nextVoxelPos	= (float) voxel[axis] / 64.f;
				nextCrossingT[axis] = tMin + (nextVoxelPos - gi) / dd[axis];
				step[axis]			= -1;
				voxelLimit[axis]	= -1;
			}
		}
	}
	KRR_DEV bool next(MajorantSeg &seg) {
		if (tMin >= tMax) return false;
		if (!grid) {
			seg	 = {tMin, tMax, sigma_t};
			tMin = tMax;
			return true;
		}
		int bits = ((nextCrossingT[0] < nextCrossingT[1]) << 2) + ((nextCrossingT[0] < nextCrossingT[2]) << 1) +
				   ((nextCrossingT[1] < nextCrossingT[2]));
		const int stepAxis = (0x00221212 >> (4 * bits)) & 3; // cmpToAxis = {2,1,2,1,2,2,0,0}
		// Per-axis state is read and written through selects on the axis with
		// constant indices: indexing the arrays by stepAxis would keep the
		// iterator in scratch memory.
		const bool a0 = stepAxis == 0, a1 = stepAxis == 1, a2 = stepAxis == 2;
		const float nct	 = sel3(nextCrossingT[0], nextCrossingT[1], nextCrossingT[2], a0, a1);
		float tVoxelExit = fminf(tMax, nct);
		seg				 = {tMin, tVoxelExit, sigma_t * grid[voxel[0] + 64 * (voxel[1] + 64 * voxel[2])]};
		tMin			 = tVoxelExit;
		if (nct > tMax) tMin = tMax;
		const int v = sel3(voxel[0], voxel[1], voxel[2], a0, a1) + sel3(step[0], step[1], step[2], a0, a1);
		if (a0) voxel[0] = v;
		if (a1) voxel[1] = v;
		if (a2) voxel[2] = v;
		if (v == sel3(voxelLimit[0], voxelLimit[1], voxelLimit[2], a0, a1)) tMin = tMax;
		const float nctNext = nct + sel3(deltaT[0], deltaT[1], deltaT[2], a0, a1);
		if (a0) nextCrossingT[0] = nctNext;
		if (a1) nextCrossingT[1] = nctNext;
		if (a2) nextCrossingT[2] = nctNext;
		return true;
	}
};

// AABB3f::intersect (src/core/math/include/krrmath/aabb.h:58-76).
KRR_DEV bool aabbIntersect(const float bmin[3], const float bmax[3], const f3 &o, const f3 &d, float tMax, float *t0o, float *t1o) {
	float t0 = 0.f, t1 = tMax;
	for (int i = 0; i < 3; ++i) {
		float invRayDir = 1.f / d[i];
		float tNear		= (bmin[i] - o[i]) * invRayDir;
		float tFar		= (bmax[i] - o[i]) * invRayDir;
		if (tNear > tFar) {
			float t = tNear;
			tNear	= tFar;
			tFar	= t;
		}
		t0 = tNear > t0 ? tNear : t0;
		t1 = tFar < t1 ? tFar : t1;
		if (t0 > t1) return false;
	}
	*t0o = t0;
	*t1o = t1;
	return true;
}

// Medium::sampleRay (media.h:128-132, 174-185).
KRR_DEV void sampleRay(const DevMedium &m, const f3 &o, const f3 &d, float raytMax, MajorantIter &it) {
	f3 st(m.sigma_t[0], m.sigma_t[1], m.sigma_t[2]);
	if (m.type == 0) {
		it.init(o, d, 0.f, raytMax, st, nullptr);
		return;
	}
	float tMin = 0.f, tMax = raytMax;
	if (!aabbIntersect(m.bmin, m.bmax, o, d, raytMax, &tMin, &tMax)) {
		it.initEmpty();
		return;
	}
	it.init(o, d, tMin, tMax, m.scale * st, &m);
}

// HGPhaseFunction (media.cpp:83-106).
KRR_DEV float hgP(float gv, const f3 &wo, const f3 &wi) {
	float g		= clampf(gv, -.99f, .99f);
	float denom = 1.f + pow2(g) + 2.f * g * dot(wo, wi);
	return kInv4Pi * (1.f - pow2(g)) / (denom * sqrtf(maxf(0.f, denom)));
}

KRR_DEV f3 hgSample(float gv, const f3 &wo, float u0, float u1, float *pdf) {
	float g = clampf(gv, -.99f, .99f);
	float cosTheta;
	if (fabsf(g) < 1e-3f) cosTheta = 1.f - 2.f * u0;
	else cosTheta = -1.f / (2.f * g) * (1.f + pow2(g) - pow2((1.f - pow2(g)) / (1.f + g - 2.f * g * u0)));
	float sinTheta = sqrtf(maxf(0.f, 1.f - pow2(cosTheta)));
	float phi	   = k2Pi * u1;
	f3 won		   = normalize(wo);
	f3 T		   = getPerpendicular(won);
	f3 B		   = normalize(cross(won, T));
	float sp, cp;
	detmath::sincosf(phi, &sp, &cp);
	f3 local(sinTheta * cp, sinTheta * sp, cosTheta);
	f3 wi = T * local.x + B * local.y + won * local.z;
	*pdf  = hgP(gv, normalize(wo), wi);
	return wi;
}

} // namespace krr
