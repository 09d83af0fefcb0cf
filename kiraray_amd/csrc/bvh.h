// bvh.h - 4-wide BVH layout shared by the host builder and the traversal
// kernels.  Replaces the OptiX GAS/IAS (src/core/device/optix.cpp:143-250,
// 357-398) with a structure the gfx950 kernels traverse in software.
#pragma once

#include <stdint.h>
#include <vector>

namespace krr {

// One 128-byte node = one gfx950 L2 line.  Child boxes are stored as
// structure-of-arrays so that a lane reads them with eight 16-byte loads.
// child[c] >= 0: inner node index.  child[c] < 0: leaf, ~child[c] =
// first * 8 + (count - 1) over the leaf triangle array (count 1..8).
// Unused slots carry an inverted box (lo = +inf, hi = -inf) that no ray hits.
struct alignas(128) Bvh4Node {
	float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
	int32_t child[4];
	int32_t pad[4];
};
static_assert(sizeof(Bvh4Node) == 128, "Bvh4Node must be 128 bytes");

// Quantised node: the same tree in 64 bytes (two nodes per 128-byte line).
// Per axis a, the child boxes are 8-bit steps on a grid of spacing 2^e[a]
// from org[a], a multiple of 2^e[a] small enough that every plane
// org[a] + q * 2^e[a] (q < 256) is a float exactly: the traversal decodes a
// plane with one exact fma and runs the same robust slab test as on the
// float node.  Each decoded box contains the float box it stands for (the
// planes are rounded outward), so a traversal over quantised nodes visits a
// superset of the boxes and finds the same closest hit.
struct alignas(64) Bvh4QNode {
	float org[3];
	int8_t e[3];
	uint8_t valid;	  // bit c set: child slot c is used
	uint32_t q[6];	  // planes lox, hix, loy, hiy, loz, hiz; byte c = child c
	int32_t child[4]; // as Bvh4Node::child
	int32_t pad[2];
};
static_assert(sizeof(Bvh4QNode) == 64, "Bvh4QNode must be 64 bytes");

// Leaf triangle: three 16-byte records {v0, prim}, {v1, flags}, {v2, 0};
// flags: kLeafNull = null material (skipped by shadow rays), kLeafAlpha =
// the material has an alpha test (transmission texture or constant).
constexpr int32_t kLeafNull = 1, kLeafAlpha = 2;
struct LeafTri {
	float v0[3];
	int32_t prim;
	float v1[3];
	int32_t flags;
	float v2[3];
	int32_t pad;
};
static_assert(sizeof(LeafTri) == 48, "LeafTri must be 48 bytes");

struct Bvh4Host {
	std::vector<Bvh4Node> nodes;
	std::vector<LeafTri> tris;
	int maxDepth = 0;	// depth of the deepest inner node (root = 1)
	int64_t nLeaves = 0;
	double buildSeconds = 0;
};

// Default SAH cost of a binary-tree node, in triangle tests: the fastest of
// 0.15 .. 3 for the gfx950 traversal on the 1.31 M-triangle icosphere field
// (low values give small leaves; half the binary nodes vanish in the 4-wide
// collapse, and a watertight triangle test costs more than a 4-wide box test).
constexpr float kSahNodeCost = 0.25f;

struct BvhBuildOptions {
	int maxLeafSize = 4;			// triangles per leaf (<= 8)
	int bins		= 16;			// SAH bins per axis
	float nodeCost	= kSahNodeCost; // SAH cost of a node visit, in triangle tests
	int threads		= 0;			// 0 = hardware concurrency (capped at 16)
};

// pos: 3 floats per vertex; idx: 3 ints per triangle; nullFlag: the
// LeafTri flags of each triangle (kLeafNull | kLeafAlpha), may be null.
void buildBvh4(const float *pos, const int32_t *idx, int64_t nTris, const uint8_t *nullFlag, const BvhBuildOptions &opt,
			   Bvh4Host &out);

// Stack entries a traversal needs for this BVH.
inline int bvhStackSize(const Bvh4Host &b) { return 3 * b.maxDepth + 4; }

} // namespace krr
