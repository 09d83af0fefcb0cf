// texture.h - texture lookups shared by shading and the traversal's alpha
// test (src/core/texture.h:186-205, texture.cpp:236-243).
#pragma once

#include "vecmath.h"

namespace krr {

// Bilinear lookup with wrap-around of a normalised-coordinate RGBA float
// texture (texture.cpp:236-243: cudaAddressModeWrap, cudaFilterModeLinear;
// weights in binary32).
KRR_DEV float4 texBilinear4(const float4 *img, int w, int h, float u, float v) {
	float x = u * (float) w - 0.5f, y = v * (float) h - 0.5f;
	if (!(x > -1e6f && x < 1e6f)) x = 0.f; // NaN / far out of range: any texel keeps the index in bounds
	if (!(y > -1e6f && y < 1e6f)) y = 0.f;
	const float x0 = floorf(x), y0 = floorf(y);
	const float fx = x - x0, fy = y - y0;
	int i0 = (int) x0 % w, j0 = (int) y0 % h;
	if (i0 < 0) i0 += w;
	if (j0 < 0) j0 += h;
	const int i1 = i0 + 1 == w ? 0 : i0 + 1, j1 = j0 + 1 == h ? 0 : j0 + 1;
	const float4 a = img[(size_t) j0 * w + i0], b = img[(size_t) j0 * w + i1];
	const float4 c = img[(size_t) j1 * w + i0], d = img[(size_t) j1 * w + i1];
	const float gx = 1.f - fx, gy = 1.f - fy;
	return make_float4(gy * (gx * a.x + fx * b.x) + fy * (gx * c.x + fx * d.x), gy * (gx * a.y + fx * b.y) + fy * (gx * c.y + fx * d.y),
					   gy * (gx * a.z + fx * b.z) + fy * (gx * c.z + fx * d.z), gy * (gx * a.w + fx * b.w) + fy * (gx * c.w + fx * d.w));
}

KRR_DEV f3 texBilinear(const float4 *img, int w, int h, float u, float v) {
	const float4 c = texBilinear4(img, w, h, u, v);
	return f3(c.x, c.y, c.z);
}

} // namespace krr
