// pathtracer.h - launch interface of the megakernel path tracer
// (MegakernelPathTracer, src/render/megakernel/device.cu:50-195).
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"

namespace krr {

struct PtArgs {
	DevScene sc;
	TraceBvh bvh;
	CameraDev cam;
	int W, H;
	int spp, maxDepth, nee;
	float rr;
	uint32_t sampleIndex; // setPixelSample(pixel, frame * 512), device.cu:157
	float4 *film;		  // W x H, row H-1-y
	int *ovf;			  // traversal stack overflow: (stack size - LDS entries) * grid * tb ints
	unsigned long long *rays; // [0] closest rays, [1] shadow rays (added to)
	int tilesX, nTiles;		  // 8 x 8-pixel tiles, one per wave at a time
	bool textured;			  // textures, image-lit infinite lights or alpha tests: the kernel with those lookups
};

// One persistent launch: grid blocks of tb threads (a multiple of 64), lds
// bytes of traversal stack per block.  The ray counters are cleared first.
hipError_t launchPathTracer(const PtArgs &a, int grid, int tb, size_t lds, hipStream_t st);

// Resident blocks per CU of the kernel for this block size and LDS (>= 1).
int pathTracerBlocksPerCU(int tb, size_t lds, bool textured);

} // namespace krr
