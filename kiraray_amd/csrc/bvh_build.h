// bvh_build.h - binned-SAH BVH builder on the GPU (bvh_build.hip), output in
// the 4-wide layout of bvh.h.
#pragma once

#include <stdint.h>
#include <string>

#include "bvh.h"

namespace krr {

// Device arrays (hipMalloc'd; the caller frees them) and build statistics.
struct DeviceBvh {
	void *nodes = nullptr; // Bvh4Node[nNodes]
	void *tris	= nullptr; // LeafTri[nTris], in reference order
	int64_t nNodes = 0, nTris = 0, nBinary = 0, smallTrees = 0;
	int maxDepth = 0, largeLevels = 0;
	double seconds = 0;		 // whole call, including the upload of the triangle arrays
	double buildSeconds = 0; // after the upload
	double largeSeconds = 0, smallSeconds = 0, collapseSeconds = 0; // phases of buildSeconds
};

// pos: 3 floats per vertex, idx: 3 ints per triangle (validated by the
// caller), nullFlag: one byte per triangle or null; all in host memory.
// Returns 0, or -1 with err set.
int buildBvh4Device(const float *pos, int64_t nVerts, const int32_t *idx, int64_t nTris, const uint8_t *nullFlag, const BvhBuildOptions &opt,
					DeviceBvh &out, std::string &err);

} // namespace krr
