// traverse.h - software BVH4 traversal and watertight ray/triangle test for
// gfx950 (replaces optixTrace, src/render/wavefront/device.cu:12-29).
//
// Each lane traverses its own ray through the 4-wide BVH (bvh.h): one node
// fetch = seven 16-byte loads of one 128-byte line, four slab tests, hit
// children (inner nodes and leaves) sorted front to back with the far ones
// pushed on a per-lane stack in LDS (stride = block size, so the 64 lanes of
// a wave hit 64 different banks); a leaf is one more step of the loop.
//
// Hit rule (shared with the CPU checker): the Woop-Benthin-Wald watertight
// test (JCGT 2(1), 2013), hit iff t > 0 and t < tbest, or t == tbest with a
// lower primitive index; the closest hit is then independent of the tree.
// The slab test picks near/far planes by direction sign and uses NaN-ignoring
// min/max, so a zero direction component (0*inf) only ever drops that
// axis; the far distance is widened by 1 + 2*gamma(3) (Ize, "Robust BVH Ray
// Traversal", JCGT 2(2), 2013) and boxes are padded by 2^-20 relative in
// the builder.
#pragma once

#include "bvh.h"
#include "texture.h"
#include "vecmath.h"

namespace krr {

// Alpha test data (alphaKilled, src/render/shading.h:89-113), present only
// for scenes with a transmission texture or constant: per triangle the
// texture coordinates of its vertices (zero for meshes without them) and its
// material; per material the transmission image, or the constant when img is
// null.
struct AlphaTri {
	float u0, v0, u1, v1, u2, v2;
	int material, pad;
};
struct AlphaSrc {
	const float4 *img;
	int w, h;
	float T[3];
	int pad;
};

struct TraceBvh {
	const Bvh4Node *nodes;
	const LeafTri *tris;
	int stackLds; // traversal stack entries held in LDS per lane
	int steps;	  // node count + leaf triangle count: a traversal takes each node and leaf at most once
	const AlphaTri *alphaTri = nullptr; // leaves flagged kLeafAlpha index these
	const AlphaSrc *alphaSrc = nullptr;
	const Bvh4QNode *qnodes	 = nullptr; // the same tree quantised; when set ("node_format": "quantized"), the traversal reads these
};

// HashFloat(o, d) of src/util/hash.h:130-134: MurmurHash64A (Austin Appleby,
// 64-bit MurmurHash2 for x64; hash.h:37-76) with seed 0 over the 24 bytes of
// the ray origin and direction, low 32 bits scaled by 2^-32.
KRR_DEV float hashRayFloat(const f3 &o, const f3 &d) {
	const uint64_t m = 0xc6a4a7935bd1e995ull;
	const uint64_t w[3] = {(uint64_t) __float_as_uint(o.x) | (uint64_t) __float_as_uint(o.y) << 32,
						   (uint64_t) __float_as_uint(o.z) | (uint64_t) __float_as_uint(d.x) << 32,
						   (uint64_t) __float_as_uint(d.y) | (uint64_t) __float_as_uint(d.z) << 32};
	uint64_t h = 24ull * m;
#pragma unroll
	for (int i = 0; i < 3; i++) {
		uint64_t k = w[i] * m;
		k ^= k >> 47;
		k *= m;
		h ^= k;
		h *= m;
	}
	h ^= h >> 47;
	h *= m;
	h ^= h >> 47;
	return (float) (uint32_t) h * 0x1p-32f;
}

// alphaKilled (shading.h:89-113): alpha = 1 - luminance(transmission at the
// hit's texture coordinates), luminance weights (0.299, 0.587, 0.114)
// (spectrum.h:545-547); a hit is ignored when alpha <= 0, or when alpha < 1
// and HashFloat(o, d) > alpha.  Null-material triangles are never flagged.
KRR_DEV bool alphaKilled(const TraceBvh &bvh, int prim, float b1, float b2, const f3 &o, const f3 &d) {
	const AlphaTri a = bvh.alphaTri[prim];
	const AlphaSrc s = bvh.alphaSrc[a.material];
	const float b0	 = 1.f - b1 - b2;
	float r = s.T[0], g = s.T[1], bl = s.T[2];
	if (s.img) {
		const float u  = b0 * a.u0 + b1 * a.u1 + b2 * a.u2;
		const float v  = b0 * a.v0 + b1 * a.v1 + b2 * a.v2;
		const float4 c = texBilinear4(s.img, s.w, s.h, u, v);
		r			   = c.x;
		g			   = c.y;
		bl			   = c.z;
	}
	const float alpha = 1.f - (r * 0.299f + g * 0.587f + bl * 0.114f);
	if (alpha >= 1.f) return false;
	if (alpha <= 0.f) return true;
	return hashRayFloat(o, d) > alpha;
}

// Per-ray constants of the watertight test: origin, the axis permutation
// (kx, ky, kz) as the two bits of each index, and the shear.
struct RayPre {
	f3 o;
	bool x1, x2, y1, y2, z1, z2;
	float Sx, Sy, Sz;
};

// (x, y, z)[k] for k = b1 + 2*b2 in {0, 1, 2}: two selects on per-ray
// booleans, over values (a select chain on an integer index, or a select of
// two fields of a referenced struct, compiles to a divergent branch or to a
// dynamically indexed private array in scratch memory on gfx950).
KRR_DEV float pick3(float x, float y, float z, bool b1, bool b2) {
	const float xy = b1 ? y : x;
	return b2 ? z : xy;
}

KRR_DEV RayPre rayPre(const f3 &o, const f3 &d) {
	RayPre r;
	r.o		 = o;
	const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
	// kz = largest |d| component, first index on ties
	int kz = 0;
	float m = ax;
	if (ay > m) {
		kz = 1;
		m  = ay;
	}
	if (az > m) kz = 2;
	int kx = kz + 1 == 3 ? 0 : kz + 1;
	int ky = kx + 1 == 3 ? 0 : kx + 1;
	const float dz = kz == 0 ? d.x : (kz == 1 ? d.y : d.z);
	if (dz < 0.0f) {
		int t = kx;
		kx	  = ky;
		ky	  = t;
	}
	r.x1 = kx & 1;
	r.x2 = kx & 2;
	r.y1 = ky & 1;
	r.y2 = ky & 2;
	r.z1 = kz & 1;
	r.z2 = kz & 2;
	r.Sx = pick3(d.x, d.y, d.z, r.x1, r.x2) / dz;
	r.Sy = pick3(d.x, d.y, d.z, r.y1, r.y2) / dz;
	r.Sz = 1.0f / dz;
	return r;
}

// Watertight test; true with t > 0 and the barycentrics of v1, v2.
KRR_DEV bool intersectTri(const RayPre &r, const f3 &v0, const f3 &v1, const f3 &v2, float &t, float &b1, float &b2) {
	const float a0 = v0.x - r.o.x, a1 = v0.y - r.o.y, a2 = v0.z - r.o.z;
	const float b0 = v1.x - r.o.x, bb1 = v1.y - r.o.y, bb2 = v1.z - r.o.z;
	const float c0 = v2.x - r.o.x, c1 = v2.y - r.o.y, c2 = v2.z - r.o.z;
	const float Akz = pick3(a0, a1, a2, r.z1, r.z2), Bkz = pick3(b0, bb1, bb2, r.z1, r.z2), Ckz = pick3(c0, c1, c2, r.z1, r.z2);
	const float Ax = pick3(a0, a1, a2, r.x1, r.x2) - r.Sx * Akz, Ay = pick3(a0, a1, a2, r.y1, r.y2) - r.Sy * Akz;
	const float Bx = pick3(b0, bb1, bb2, r.x1, r.x2) - r.Sx * Bkz, By = pick3(b0, bb1, bb2, r.y1, r.y2) - r.Sy * Bkz;
	const float Cx = pick3(c0, c1, c2, r.x1, r.x2) - r.Sx * Ckz, Cy = pick3(c0, c1, c2, r.y1, r.y2) - r.Sy * Ckz;
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
	const float T	= U * (r.Sz * Akz) + V * (r.Sz * Bkz) + W * (r.Sz * Ckz);
	const float inv = 1.0f / det;
	t				= T * inv;
	if (!(t > 0.0f)) return false;
	b1 = V * inv;
	b2 = W * inv;
	return true;
}

struct TraceResult {
	float t;
	int prim;
	float b1, b2;
};

// Per-lane traversal stack: the first nLds entries in LDS (entries `stride`
// ints apart, so a wave's 64 lanes hit 64 banks), deeper entries in a global
// overflow area laid out lane-fastest (entries `ovStride` ints apart:
// coalesced when the lanes of a wave spill together).  A small LDS footprint
// keeps the traversal kernels at high occupancy; deep stacks are rare but
// stay exact.  nLds is uniform over the launch (kShortStack by default).
//
// Each lane owns nLds + 1 LDS slots.  Every push writes LDS slot
// min(sp, nLds) - the extra slot absorbs the pushes that belong to the
// overflow area - and every pop reads that slot, so the common case is one
// ds_write / ds_read; only entries at sp >= nLds also take a (rarely
// entered) branch to global memory.  A pointer chosen between LDS and
// global memory would compile to flat loads and stores on every access.
constexpr int kShortStack = 16;
struct LaneStack {
	int *lds;
	int stride;
	int *ovf;
	int ovStride;
	int nLds;
	// LDS slot min(sp, nLds) * stride: both factors are below 2^23, so the
	// full-rate 24-bit multiply gives the product
	KRR_DEV int slot(int sp) const { return __mul24(min(sp, nLds), stride); }
	KRR_DEV void put(int sp, int v) const {
		lds[slot(sp)] = v;
		if (sp >= nLds) ovf[(size_t) (sp - nLds) * ovStride] = v;
	}
	KRR_DEV int get(int sp) const {
		int v = lds[slot(sp)];
		// keeps the LDS read a ds_read: without the barrier the compiler folds
		// both reads into one flat load of a selected address
		asm volatile("" : "+v"(v));
		if (sp >= nLds) v = ovf[(size_t) (sp - nLds) * ovStride];
		return v;
	}
};

// Per-ray constants of the slab test.  The distance to plane p along axis a,
// (p - o_a) / d_a, is one fma p * i_a + b_a with i_a = 1/d_a (rounded) and
// b_a = -(o_a i_a) - w_a for near planes, -(o_a i_a) + w_a for far planes,
// w_a = 2^-22 |o_a i_a| + 2^-40 |i_a|.  The roundings of o_a i_a and of b_a
// are each at most 2^-24 |o_a i_a|, so near distances stay at or below
// (p - o_a) i_a and far distances at or above it before the fma's own
// rounding; what remains is relative error (the rounding of i_a and of the
// fma, gamma(2)), which the far distance's 1 + 2 gamma(3) widening covers as
// it covers the subtract-then-multiply form (Ize, "Robust BVH Ray
// Traversal", JCGT 2(2), 2013).  In world units w_a is a margin of
// 2^-22 |o_a| + 2^-40 on each side of a slab.
//
// |i_a| is capped at 2^100: a zero (or tiny) direction component is taken
// as +-2^-100, a ray that drifts from the true one by at most 2^-100 t along
// that axis, less than the 2^-40 margin for every t below 2^60, so the test
// stays conservative and the axis still culls.  (An infinite i_a would make
// o_a i_a infinite and the axis' distances NaN: no culling on that axis,
// and a ray along a coordinate plane would visit every box its projection
// crosses.)  Only an o_a beyond 2^27 overflows o_a i_a; that axis then stops
// culling, which is still conservative.
constexpr float kSlabMaxInv = 0x1p100f;
struct SlabRay {
	float ix, iy, iz;
	float bnx, bny, bnz; // near planes
	float bfx, bfy, bfz; // far planes
	// byte offsets of the near planes inside a node (lo of axis a at 32a, hi
	// at 32a + 16, by direction sign); the far plane is at offset ^ 16
	uint32_t nX, nY, nZ;
};

KRR_DEV SlabRay slabRay(const f3 &o, const f3 &d) {
	SlabRay s;
	const float rx = 1.0f / d.x, ry = 1.0f / d.y, rz = 1.0f / d.z;
	s.ix		   = fabsf(rx) <= kSlabMaxInv ? rx : copysignf(kSlabMaxInv, d.x);
	s.iy		   = fabsf(ry) <= kSlabMaxInv ? ry : copysignf(kSlabMaxInv, d.y);
	s.iz		   = fabsf(rz) <= kSlabMaxInv ? rz : copysignf(kSlabMaxInv, d.z);
	const float qx = o.x * s.ix, qy = o.y * s.iy, qz = o.z * s.iz;
	const float wx = fmaf(fabsf(s.ix), 0x1p-40f, fabsf(qx) * 0x1p-22f);
	const float wy = fmaf(fabsf(s.iy), 0x1p-40f, fabsf(qy) * 0x1p-22f);
	const float wz = fmaf(fabsf(s.iz), 0x1p-40f, fabsf(qz) * 0x1p-22f);
	s.bnx		   = -qx - wx;
	s.bny		   = -qy - wy;
	s.bnz		   = -qz - wz;
	s.bfx		   = -qx + wx;
	s.bfy		   = -qy + wy;
	s.bfz		   = -qz + wz;
	s.nX		   = __builtin_signbit(d.x) ? 16u : 0u;
	s.nY		   = __builtin_signbit(d.y) ? 48u : 32u;
	s.nZ		   = __builtin_signbit(d.z) ? 80u : 64u;
	return s;
}

constexpr float kWiden = 1.00000036f;			 // 1 + 2*gamma(3)
constexpr float kTie   = 1.00000095367431640625f; // 1 + 2^-20

// Slab test of the four children of inner node `entry` against [0, tlim]:
// k[i] = entry distance of child i, +inf when missed; c[i] = its reference.
// Float nodes are read from LDS for entries below nLdsNodes (the block's copy
// of the top of the tree), from global memory otherwise: seven 16-byte loads
// of the node's 128-byte line (the six planes the ray needs and the children).
KRR_DEV void nodeTest(const TraceBvh &bvh, const SlabRay &sr, int entry, float tlim, const float4 *nodeLds, int nLdsNodes, float &k0,
					  float &k1, float &k2, float &k3, int &c0, int &c1, int &c2, int &c3) {
#define KRR_SLAB(kc, px, qx, py, qy, pz, qz)                                                                                                   \
	{                                                                                                                                          \
		const float t0 = fmaxf(fmaxf(fmaf(px, sr.ix, sr.bnx), fmaf(py, sr.iy, sr.bny)), fmaxf(fmaf(pz, sr.iz, sr.bnz), 0.0f));                 \
		const float t1 = fminf(fminf(fminf(fmaf(qx, sr.ix, sr.bfx), fmaf(qy, sr.iy, sr.bfy)), fmaf(qz, sr.iz, sr.bfz)) * kWiden, tlim);         \
		kc			   = t0 <= t1 ? t0 : __builtin_inff();                                                                                      \
	}
	if (bvh.qnodes) {
		// One 64-byte quantised node: {org, e, valid}, planes lox hix loy
		// hiy, planes loz hiz + children 0 1, children 2 3.  Plane c of a
		// dword is byte c; org + q 2^e is exact (one fma, no rounding).
		const uint4 *np = reinterpret_cast<const uint4 *>(bvh.qnodes + entry);
		const uint4 A = np[0], B = np[1], C = np[2], D = np[3];
		const float sx = __int_as_float(((int) (int8_t) (A.w & 0xffu) + 127) << 23);
		const float sy = __int_as_float(((int) (int8_t) ((A.w >> 8) & 0xffu) + 127) << 23);
		const float sz = __int_as_float(((int) (int8_t) ((A.w >> 16) & 0xffu) + 127) << 23);
		const float gx = __uint_as_float(A.x), gy = __uint_as_float(A.y), gz = __uint_as_float(A.z);
		const uint32_t valid = A.w >> 24;
		const uint32_t qnx = sr.nX ? B.y : B.x, qfx = sr.nX ? B.x : B.y;
		const uint32_t qny = sr.nY == 48u ? B.w : B.z, qfy = sr.nY == 48u ? B.z : B.w;
		const uint32_t qnz = sr.nZ == 80u ? C.y : C.x, qfz = sr.nZ == 80u ? C.x : C.y;
#define KRR_DEQ(w, c, sc, g) fmaf((float) (((w) >> (8 * (c))) & 0xffu), sc, g)
#define KRR_QSLAB(kc, c)                                                                                                                       \
	KRR_SLAB(kc, KRR_DEQ(qnx, c, sx, gx), KRR_DEQ(qfx, c, sx, gx), KRR_DEQ(qny, c, sy, gy), KRR_DEQ(qfy, c, sy, gy), KRR_DEQ(qnz, c, sz, gz),  \
			 KRR_DEQ(qfz, c, sz, gz))                                                                                                          \
	if (!((valid >> (c)) & 1u)) kc = __builtin_inff();
		KRR_QSLAB(k0, 0)
		KRR_QSLAB(k1, 1)
		KRR_QSLAB(k2, 2)
		KRR_QSLAB(k3, 3)
#undef KRR_QSLAB
#undef KRR_DEQ
		c0 = (int) C.z;
		c1 = (int) C.w;
		c2 = (int) D.x;
		c3 = (int) D.y;
	} else {
		float4 nx, fx, ny, fy, nz, fz;
		int4 ch;
		if (entry < nLdsNodes) {
			const float4 *np = nodeLds + (uint32_t) entry * 8u;
			nx				 = np[sr.nX >> 4];
			fx				 = np[(sr.nX ^ 16u) >> 4];
			ny				 = np[sr.nY >> 4];
			fy				 = np[(sr.nY ^ 16u) >> 4];
			nz				 = np[sr.nZ >> 4];
			fz				 = np[(sr.nZ ^ 16u) >> 4];
			const float4 c	 = np[6];
			// keeps these LDS reads ds_reads: without the barrier the compiler
			// may merge both branches into flat loads of a selected address
			asm volatile("" : "+v"(nx.x));
			ch = make_int4(__float_as_int(c.x), __float_as_int(c.y), __float_as_int(c.z), __float_as_int(c.w));
		} else {
			const char *nodeBase = reinterpret_cast<const char *>(bvh.nodes);
			const uint32_t nb	 = (uint32_t) entry * 128u; // node array < 4 GiB (checked at build)
#define KRR_PLANE(off) (*reinterpret_cast<const float4 *>(nodeBase + (nb + (off))))
			nx = KRR_PLANE(sr.nX);
			fx = KRR_PLANE(sr.nX ^ 16u);
			ny = KRR_PLANE(sr.nY);
			fy = KRR_PLANE(sr.nY ^ 16u);
			nz = KRR_PLANE(sr.nZ);
			fz = KRR_PLANE(sr.nZ ^ 16u);
#undef KRR_PLANE
			ch = *reinterpret_cast<const int4 *>(nodeBase + (nb + 96u));
		}
		KRR_SLAB(k0, nx.x, fx.x, ny.x, fy.x, nz.x, fz.x)
		KRR_SLAB(k1, nx.y, fx.y, ny.y, fy.y, nz.y, fz.y)
		KRR_SLAB(k2, nx.z, fx.z, ny.z, fy.z, nz.z, fz.z)
		KRR_SLAB(k3, nx.w, fx.w, ny.w, fy.w, nz.w, fz.w)
		c0 = ch.x;
		c1 = ch.y;
		c2 = ch.z;
		c3 = ch.w;
	}
#undef KRR_SLAB
	// Hit children nearest first (a five-exchange sorting network on
	// selects; equal distances keep their slot order).  Missed children
	// (+inf) end up last.
#define KRR_CSWAP(ka, ca, kb, cb)                                                                                                              \
	{                                                                                                                                          \
		const bool sw = kb < ka;                                                                                                               \
		const float lo = sw ? kb : ka, hi = sw ? ka : kb;                                                                                      \
		const int clo = sw ? cb : ca, chi = sw ? ca : cb;                                                                                      \
		ka			  = lo;                                                                                                                     \
		kb			  = hi;                                                                                                                     \
		ca			  = clo;                                                                                                                    \
		cb			  = chi;                                                                                                                    \
	}
	KRR_CSWAP(k0, c0, k1, c1)
	KRR_CSWAP(k2, c2, k3, c3)
	KRR_CSWAP(k0, c0, k2, c2)
	KRR_CSWAP(k1, c1, k3, c3)
	KRR_CSWAP(k1, c1, k2, c2)
#undef KRR_CSWAP
}

// Traverses one ray.  ANY: stop at the first hit of a non-null-material
// triangle (shadow rays, device.cu:83-100).  ALPHA: hits on triangles flagged
// kLeafAlpha pass the any-hit alpha test first (device.cu:65-67,86-87,111-113;
// megakernel/device.cu:134-142); the result does not depend on the order of
// the candidates, since the test is a function of the ray and the hit.
//
// Each loop iteration takes one entry: an inner node (slab test of its four
// children, hit children - inner and leaf alike - sorted near to far, the
// nearest taken next, the others pushed) or a leaf (its triangles tested).
// A lane therefore runs at most one leaf's tests per iteration, and the wave
// runs the leaf code once per iteration for all lanes that hold a leaf:
// testing the leaves of a node on the spot, slot by slot, made the wave run
// up to four leaf blocks per node visit, each for the few lanes whose child in
// that slot was a hit leaf.  The closest hit does not depend on the order.
template <bool ANY, bool COUNT, bool ALPHA = false>
KRR_DEV TraceResult traverse(const TraceBvh &bvh, const f3 &o, const f3 &d, float tmax, const LaneStack &stack,
							 uint32_t &nodeVisits, uint32_t &triTests) {
	TraceResult res{tmax, -1, 0.f, 0.f};
	const RayPre rp	 = rayPre(o, d);
	const SlabRay sr = slabRay(o, d);
	int sp			 = 0;
	int entry		 = 0; // >= 0: inner node; < 0: leaf code ~(first * 8 + count - 1)
	// Each node and leaf is taken at most once; more steps can only come from
	// a corrupt stack or tree, and the budget turns that into a finished
	// (wrong) traversal instead of a wave that never ends.
	int budget = bvh.steps;
	while (--budget >= 0) {
		if (entry < 0) {
			const int code	= ~entry;
			const int first = code >> 3, cnt = (code & 7) + 1;
			for (int k = 0; k < cnt; k++) {
				const float4 *tp = reinterpret_cast<const float4 *>(bvh.tris + first + k);
				const float4 a = tp[0], b = tp[1], e = tp[2];
				if (ANY && (__float_as_int(b.w) & kLeafNull)) continue;
				if (COUNT) triTests++;
				float t, b1, b2;
				if (!intersectTri(rp, f3(a.x, a.y, a.z), f3(b.x, b.y, b.z), f3(e.x, e.y, e.z), t, b1, b2)) continue;
				const int prim = __float_as_int(a.w);
				if (t < res.t || (t == res.t && res.prim >= 0 && prim < res.prim)) {
					if constexpr (ALPHA) {
						if ((__float_as_int(b.w) & kLeafAlpha) && alphaKilled(bvh, prim, b1, b2, o, d)) continue;
					}
					res.t	 = t;
					res.prim = prim;
					res.b1	 = b1;
					res.b2	 = b2;
					if (ANY) return res;
				}
			}
			if (sp == 0) break;
			entry = stack.get(--sp);
			continue;
		}
		if (COUNT) nodeVisits++;
		float k0, k1, k2, k3;
		int c0, c1, c2, c3;
		nodeTest(bvh, sr, entry, res.t * kTie, nullptr, 0, k0, k1, k2, k3, c0, c1, c2, c3);
		if (k0 == __builtin_inff()) {
			if (sp == 0) break;
			entry = stack.get(--sp);
			continue;
		}
		if (k3 != __builtin_inff()) stack.put(sp++, c3);
		if (k2 != __builtin_inff()) stack.put(sp++, c2);
		if (k1 != __builtin_inff()) stack.put(sp++, c1);
		entry = c0;
	}
	return res;
}

// Default refill threshold of the persistent launch (pass param
// refill_lanes): on the 1.31 M-triangle 1080p field, rounds that end below
// 8 traversing lanes of 64 took 4.12 ms per frame against 4.57 ms for rays
// run to their end, 4.23 ms at 16 and 4.49 ms at 32.
constexpr int kDefaultRefillLanes = 8;
// Smallest tree (nodes) for which refill_lanes -1 turns the refill launch on.
constexpr int64_t kRefillMinNodes = 4096;
// Default leaf share of the refill launch's traversal (pass param leaf_share):
// on the same field, two sweeps on two GPUs: iterations that take only leaves
// once 30-50 % of the traversing lanes hold one ran 3.86-3.99 ms per frame
// (35 %: 3.90 and 3.93) against 4.08 and 4.12 ms for iterations that take both
// kinds of entry (20 %: 4.02, 65 %: 4.09, 80 %: 4.40).
constexpr int kDefaultLeafShare = 35;
// Largest per-block LDS copy of the top tree nodes (128 KB of 128-byte nodes).
constexpr int kMaxLdsNodes = 1024;

// A traversal that stops and resumes between steps: the entry to take next,
// the stack depth (the entries stay in the lane's stack), the step budget and
// the best hit so far.  A ray advanced by traverseSteps until it is done takes
// exactly the steps of traverse() and ends with the same result.
struct TravState {
	int entry, sp, budget;
	float t;
	int prim;
	float b1, b2;
};

KRR_DEV void travInit(TravState &s, const TraceBvh &bvh, float tmax) {
	s.entry	 = 0;
	s.sp	 = 0;
	s.budget = bvh.steps;
	s.t		 = tmax;
	s.prim	 = -1;
	s.b1	 = 0.f;
	s.b2	 = 0.f;
}

// Advances the rays of the lanes with `going` set, one entry per lane per
// iteration, until fewer than minLanes lanes of the wave are still going
// (minLanes 1: until every ray is done).  A lane clears `going` when its ray
// is done.  any: the lane's ray is a shadow ray (stop at the first hit of a
// non-null-material triangle); otherwise the closest hit is kept.  All lanes
// of the wave call this together; the loop exit is wave-uniform.
//
// nodeLds: the first nLdsNodes nodes of the tree (float nodes; the top levels
// of a breadth-first numbered tree) copied into LDS by the block; those are
// read from there.
//
// leafShare (percent, 0 = off): an iteration takes either the leaves or the
// inner nodes of the traversing lanes, not both: leaves once at least
// leafShare % of the traversing lanes hold one, inner nodes otherwise; the
// other lanes wait.  A mixed iteration runs both code paths, each for part of
// the wave.  Each ray still takes its own entries in its own order.
template <bool COUNT, bool ALPHA>
KRR_DEV void traverseSteps(const TraceBvh &bvh, const f3 &o, const f3 &d, const bool any, bool &going, TravState &s,
						   const LaneStack &stack, const int minLanes, uint32_t &nodeVisits, uint32_t &triTests,
						   const float4 *nodeLds = nullptr, int nLdsNodes = 0, int leafShare = 0) {
	const RayPre rp	 = rayPre(o, d);
	const SlabRay sr = slabRay(o, d);
	while (true) {
		const uint64_t lanes = __ballot(going);
		const int nGoing	 = __popcll(lanes);
		if (lanes == 0 || nGoing < minLanes) break;
		// wave-uniform: with leafShare 0 both kinds of entry are taken
		bool takeLeaf = true, takeNode = true;
		if (leafShare > 0) {
			const int nLeaf = __popcll(__ballot(going && s.entry < 0));
			takeLeaf		= nLeaf * 100 >= leafShare * nGoing; // all lanes at leaves: true; none: false
			takeNode		= !takeLeaf;
		}
		if (!going || !(s.entry < 0 ? takeLeaf : takeNode)) continue;
		if (--s.budget < 0) {
			going = false;
			continue;
		}
		if (s.entry < 0) {
			const int code	= ~s.entry;
			const int first = code >> 3, cnt = (code & 7) + 1;
			bool stop = false;
			for (int k = 0; k < cnt; k++) {
				const float4 *tp = reinterpret_cast<const float4 *>(bvh.tris + first + k);
				const float4 a = tp[0], b = tp[1], e = tp[2];
				if (any && (__float_as_int(b.w) & kLeafNull)) continue;
				if (COUNT) triTests++;
				float t, b1, b2;
				if (!intersectTri(rp, f3(a.x, a.y, a.z), f3(b.x, b.y, b.z), f3(e.x, e.y, e.z), t, b1, b2)) continue;
				const int prim = __float_as_int(a.w);
				if (t < s.t || (t == s.t && s.prim >= 0 && prim < s.prim)) {
					if constexpr (ALPHA) {
						if ((__float_as_int(b.w) & kLeafAlpha) && alphaKilled(bvh, prim, b1, b2, o, d)) continue;
					}
					s.t	   = t;
					s.prim = prim;
					s.b1   = b1;
					s.b2   = b2;
					if (any) {
						stop = true;
						break;
					}
				}
			}
			if (stop || s.sp == 0) {
				going = false;
				continue;
			}
			s.entry = stack.get(--s.sp);
			continue;
		}
		if (COUNT) nodeVisits++;
		float k0, k1, k2, k3;
		int c0, c1, c2, c3;
		nodeTest(bvh, sr, s.entry, s.t * kTie, nodeLds, nLdsNodes, k0, k1, k2, k3, c0, c1, c2, c3);
		if (k0 == __builtin_inff()) {
			if (s.sp == 0) {
				going = false;
				continue;
			}
			s.entry = stack.get(--s.sp);
			continue;
		}
		if (k3 != __builtin_inff()) stack.put(s.sp++, c3);
		if (k2 != __builtin_inff()) stack.put(s.sp++, c2);
		if (k1 != __builtin_inff()) stack.put(s.sp++, c1);
		s.entry = c0;
	}
}

} // namespace krr
