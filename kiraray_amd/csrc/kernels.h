// kernels.h - device-side state and wave-level helpers of the wavefront tracer.
//
// Queue layout (array-of-structures in 16-byte records, so that one lane reads
// one item with 16-byte loads and a wave reads consecutive cache lines):
//   ray queue item (96 B):  ray[2i]   = {o.xyz, tmax}
//                           ray[2i+1] = {d.xyz, pixelId}
//                           st[4i]    = {thp.xyz, pu.x}
//                           st[4i+1]  = {pu.y, pu.z, pl.x, pl.y}
//                           st[4i+2]  = {pl.z, ctx.p.xyz}
//                           st[4i+3]  = {ctx.n.xyz, meta}
//   meta = depth | bsdfType << 8 | (medium + 1) << 16
//   hit record (16 B):      {t, prim, b1, b2}
//   shadow item (48/80 B):  ray[2i], ray[2i+1] = {o, tmax}, {d, pixelId}
//                           c[i] = {Ld / mean(pu + pl), medium + 1}   (no media)
//                           ext[3i..3i+2] = {Ld, pu, pl}               (media)
// The reference keeps one SoA array per field (src/render/wavefront/
// workitem.soa:79-152) and pushes with one atomicAdd per item
// (workqueue.h:54-70); here a wave reserves its slots with one atomic.
#pragma once

#include "shading.h"
#include "traverse.h"

namespace krr {

struct RayQueueDev {
	float4 *ray; // 2 per item
	float4 *st;	 // 4 per item
};

struct ShadowQueueDev {
	float4 *ray; // 2 per item
	float4 *c;	 // 1 per item
	float4 *ext; // 3 per item (media only)
};

struct PixelStateDev {
	float4 *L;		// radiance of the current sample
	float4 *pixel;	// sum over samples
	uint64_t *rng;	// PCG32 state
	int *channel;	// hero channel of the medium code (SampledWavelengths::mainIndex)
	const int *order; // fetch order of the camera paths (null: pixel order)
	uint8_t *plen;	  // length (vertices - 1) of the pixel's last path, capped at 255 (null: not recorded)
};

struct Counters {
	int rayCount[2];
	int shadowCount;
	int pad0;
	unsigned long long closestTotal;
	unsigned long long shadowTotal;
	unsigned long long nodeVisits;	   // closest-hit traversal of the wavefront kernels
	unsigned long long triTests;
	unsigned long long shadowNodeVisits; // shadow traversal of the wavefront kernels
	unsigned long long shadowTriTests;
	unsigned long long closestLaunches;
	unsigned long long megaClosest;	   // closest-hit rays traced in the tail megakernel
	unsigned long long megaShadow;	   // shadow rays traced in the tail megakernel
	unsigned long long megaNodeVisits; // closest + shadow traversal in the megakernel
	unsigned long long megaTriTests;
	unsigned long long nrcQueriesTotal; // cache queries answered this frame
	int nrcTrainDone;					// training vertices whose targets are written
	// Counters that take an atomic per wave during a launch, one 128-byte
	// line each, so that their atomics do not serialise against each other.
	alignas(128) int work[4]; // chunk counters of the persistent kernels: 0 closest, 1 shadow, 2 megakernel
	// radiance-cache queries of the current sample and training vertices
	// pushed this frame (may exceed the batch): adjacent, so that one 64-bit
	// atomic reserves slots in both
	alignas(128) int nrcQueries;
	int nrcTrain;
	alignas(128) int nrcCandidates; // training-path vertices seen this frame (kept or not)
	int nrcPrevCandidates;			// ... in the previous frame
	float nrcKeep;					// probability of keeping a candidate this frame
};

// Neural radiance cache queues of the path tracer.  A query is 5 float4:
// the 64-byte krr_nrc_vertex and {thp.xyz, pixel}; a training vertex is 6
// float4: the vertex, {thp.xyz, pixel}, {L.xyz so far, 0}.
struct NrcDev {
	float4 *query;
	float4 *train;
	int *queryCount, *trainCount; // Counters::nrcQueries, Counters::nrcTrain (adjacent, 8-byte aligned)
	int *candCount;				  // Counters::nrcCandidates
	const float *keep;			  // Counters::nrcKeep
	float keepProb;				  // *keep, loaded once per launch
	int capQuery, capTrain;
	int trainMod;	// pixel trains when trainHash(pixel) % trainMod == phase
	int phase;		// frame index % trainMod
	float spread;	// path-spread threshold c of the termination rule (0: first vertex after a bounce)
};

KRR_DEV uint32_t trainHash(uint32_t pixel) {
	uint32_t h = pixel * 0x9E3779B1u;
	h ^= h >> 15;
	h *= 0x85EBCA77u;
	h ^= h >> 13;
	return h;
}

struct CameraDev {
	float filmH, focalLength, focalDistance, lensRadius, aspect;
	float M[12];
	int medium;
};

struct WfParams {
	int W, H, N;
	int spp;
	int maxDepth;
	float rr;
	int rrMinDepth;
	int nee;
	int enableMedium;
	int enableClamp;
	float clampMax;
	int legacySpp;
	uint64_t inc;		  // PCG32 increment of this frame
	uint32_t sampleIndex; // frameIndex * spp (as uint)
	int refillLanes;	  // persistent launch without media: end a traversal round below this many traversing lanes (1..64)
	int ldsNodes;		  // refill launch: tree nodes 0..ldsNodes-1 are read from a per-block copy in LDS
	int leafShare;		  // refill launch: leaf-only iterations once this percent of traversing lanes hold a leaf (0: off)
};

// Values that change from frame to frame, read from device memory by the
// multi-view kernels so that a captured graph can be replayed.
struct FrameDev {
	uint64_t inc;		  // PCG32 increment: (frameIndex * spp) << 1 | 1
	uint32_t sampleIndex; // frameIndex * spp (as uint)
	int phase;			  // radiance-cache training phase: frameIndex % trainMod
};

// One camera view of a multi-view frame; its film is W x H float4, row H-1-y.
struct ViewDev {
	CameraDev cam;
	int W, H;
	float4 *film;
};

struct ViewsDev {
	const ViewDev *view;
	const int *start; // n + 1 entries: first global pixel of each view, total last
	int n;
	const FrameDev *frame;
	const int *tileStart; // n + 1 entries: first global 8x8 tile of each view, total last
};

KRR_DEV int laneId() { return __lane_id(); }

// Wave-aggregated queue push: one atomic per wave per queue, each active lane
// that wants a slot gets base + (number of wanting lanes below it).
KRR_DEV int wavePush(int *counter, bool want) {
	const uint64_t mask = __ballot(want);
	if (mask == 0) return -1;
	const int leader = __ffsll((unsigned long long) mask) - 1;
	int base		 = 0;
	if (laneId() == leader) base = atomicAdd(counter, __popcll(mask));
	base = __shfl(base, leader);
	if (!want) return -1;
	const uint64_t below = laneId() == 0 ? 0ull : (mask & (~0ull >> (64 - laneId())));
	return base + __popcll(below);
}

// Two wave-aggregated pushes with one 64-bit atomic on the pair of int
// counters {a, b} at `pair` (a in the low word): ia / ib are the slots of the
// lanes that want them, -1 elsewhere.
KRR_DEV void wavePush2(int *pair, bool wantA, bool wantB, int &ia, int &ib) {
	const uint64_t ma = __ballot(wantA), mb = __ballot(wantB);
	ia = ib = -1;
	if ((ma | mb) == 0) return;
	const int leader			 = __ffsll((unsigned long long) (ma | mb)) - 1;
	unsigned long long base = 0;
	if (laneId() == leader)
		base = atomicAdd(reinterpret_cast<unsigned long long *>(pair),
						 (unsigned long long) __popcll(ma) | ((unsigned long long) __popcll(mb) << 32));
	base				 = __shfl(base, leader);
	const uint64_t below = laneId() == 0 ? 0ull : (~0ull >> (64 - laneId()));
	if (wantA) ia = (int) (uint32_t) base + __popcll(ma & below);
	if (wantB) ib = (int) (uint32_t) (base >> 32) + __popcll(mb & below);
}

// Wave-aggregated count (no slot returned, so the atomic is not waited for).
KRR_DEV void waveCount(int *counter, bool want) {
	const uint64_t mask = __ballot(want);
	if (mask != 0 && laneId() == __ffsll((unsigned long long) mask) - 1) atomicAdd(counter, __popcll(mask));
}

// Persistent-kernel work distribution: each wave takes the next 64 items of
// the queue with one atomic (wave-uniform result).
KRR_DEV int waveFetch(int *work) {
	int base = 0;
	if (laneId() == 0) base = atomicAdd(work, 64);
	return __shfl(base, 0);
}

// Sum of a per-lane counter over the wave, added to a 64-bit global counter
// by one lane.  Must be called by every lane of the wave.
KRR_DEV void waveAddU64(unsigned long long *dst, uint32_t v) {
	unsigned long long s = v;
	for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
	if (laneId() == 0 && s) atomicAdd(dst, s);
}

KRR_DEV uint32_t packMeta(int depth, int bsdfType, int medium) {
	return (uint32_t) depth | ((uint32_t) bsdfType << 8) | ((uint32_t) (medium + 1) << 16);
}
KRR_DEV int metaDepth(uint32_t m) { return (int) (m & 0xff); }
KRR_DEV int metaBsdf(uint32_t m) { return (int) ((m >> 8) & 0xff); }
KRR_DEV int metaMedium(uint32_t m) { return (int) (m >> 16) - 1; }

struct RayItem {
	f3 o, d;
	f3 thp, pu, pl, cp, cn;
	int depth, bsdfType, medium;
	int pixel;
	// path spread of the radiance-cache termination rule (registers only,
	// not queued): primary footprint a0, running sum of the secondary
	// footprints' square roots, solid-angle pdf of the last bounce (0 = delta)
	float sp0, spSum, pdfPrev;
};

KRR_DEV RayItem loadRayItem(const RayQueueDev &q, int i) {
	RayItem r;
	const float4 a = q.ray[2 * i], b = q.ray[2 * i + 1];
	const float4 s0 = q.st[4 * i], s1 = q.st[4 * i + 1], s2 = q.st[4 * i + 2], s3 = q.st[4 * i + 3];
	r.o		 = f3(a.x, a.y, a.z);
	r.d		 = f3(b.x, b.y, b.z);
	r.pixel	 = __float_as_int(b.w);
	r.thp	 = f3(s0.x, s0.y, s0.z);
	r.pu	 = f3(s0.w, s1.x, s1.y);
	r.pl	 = f3(s1.z, s1.w, s2.x);
	r.cp	 = f3(s2.y, s2.z, s2.w);
	r.cn	 = f3(s3.x, s3.y, s3.z);
	uint32_t m = (uint32_t) __float_as_int(s3.w);
	r.depth	   = metaDepth(m);
	r.bsdfType = metaBsdf(m);
	r.medium   = metaMedium(m);
	r.sp0	   = 0.f;
	r.spSum	   = 0.f;
	r.pdfPrev  = 0.f;
	return r;
}

KRR_DEV void storeRayItem(const RayQueueDev &q, int i, const RayItem &r) {
	q.ray[2 * i]	 = make_float4(r.o.x, r.o.y, r.o.z, __builtin_inff());
	q.ray[2 * i + 1] = make_float4(r.d.x, r.d.y, r.d.z, __int_as_float(r.pixel));
	q.st[4 * i]		 = make_float4(r.thp.x, r.thp.y, r.thp.z, r.pu.x);
	q.st[4 * i + 1]	 = make_float4(r.pu.y, r.pu.z, r.pl.x, r.pl.y);
	q.st[4 * i + 2]	 = make_float4(r.pl.z, r.cp.x, r.cp.y, r.cp.z);
	q.st[4 * i + 3]	 = make_float4(r.cn.x, r.cn.y, r.cn.z, __int_as_float((int) packMeta(r.depth, r.bsdfType, r.medium)));
}

// CameraData::generateSample + getRay (camera.h:32-58): the five camera draws
// (pFilm 2, pLens 2, time 1) in order, pixel (x, y) of a W x H film.
KRR_DEV RayItem cameraRay(const CameraDev &cam, int W, int H, int x, int y, Pcg &s) {
	float cs[5];
	for (int k = 0; k < 5; k++) cs[k] = s.get1D();
	const float p0	 = (float) x + 0.5f + cs[0];
	const float p1	 = (float) y + 0.5f + cs[1];
	const float ndc0 = (2.f * p0) / (float) W + (-1.f);
	const float ndc1 = (2.f * p1) / (float) H + (-1.f);
	const float fov	 = detmath::atan2f(cam.filmH * 0.5f, cam.focalLength);
	const f3 fd		 = normalize(f3(detmath::tanf(fov) * cam.aspect * ndc0, detmath::tanf(fov) * ndc1, -1.f));
	f3 o, d;
	if (cam.lensRadius > 1e-5f) {
		float ax, ay;
		uniformSampleDisk(cs[2], cs[3], ax, ay);
		o = f3(cam.lensRadius * ax, cam.lensRadius * ay, 0.f);
		d = normalize(fd * cam.focalDistance - o);
	} else {
		o = f3();
		d = fd;
	}
	const float *M = cam.M;
	RayItem r;
	r.o		   = f3(M[0] * o.x + M[1] * o.y + M[2] * o.z + M[3], M[4] * o.x + M[5] * o.y + M[6] * o.z + M[7],
					M[8] * o.x + M[9] * o.y + M[10] * o.z + M[11]);
	r.d		   = f3(M[0] * d.x + M[1] * d.y + M[2] * d.z, M[4] * d.x + M[5] * d.y + M[6] * d.z, M[8] * d.x + M[9] * d.y + M[10] * d.z);
	r.thp	   = f3(1.f);
	r.pu	   = f3(1.f);
	r.pl	   = f3(1.f);
	r.cp	   = f3();
	r.cn	   = f3();
	r.depth	   = 0;
	r.bsdfType = BSDF_UNSET;
	r.medium   = cam.medium;
	r.pixel	   = 0;
	r.sp0	   = 0.f;
	r.spSum	   = 0.f;
	r.pdfPrev  = 0.f;
	return r;
}

// This lane's traversal stack: LDS slots of the block, overflow slots of the
// grid (ovf holds (stack size - kShortStack) * grid threads ints).
KRR_DEV LaneStack laneStack(const TraceBvh &bvh, int *lds, int *ovf) {
	return LaneStack{lds + threadIdx.x, (int) blockDim.x, ovf + (size_t) blockIdx.x * blockDim.x + threadIdx.x, (int) (gridDim.x * blockDim.x),
					 bvh.stackLds};
}

} // namespace krr
