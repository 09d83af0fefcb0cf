// pathtracer.hip - the megakernel path tracer for gfx950
// (MegakernelPathTracer, src/render/megakernel/device.cu:50-195).
//
// One lane runs one pixel's samples from the camera ray to the end of each
// path: closest hit, emission with its MIS weight, Russian roulette from the
// first vertex on, a light sample on smooth surfaces (power heuristic,
// src/render/sampling.h:8-28), then the BSDF sample.  A persistent grid of
// waves takes 8 x 8-pixel tiles in turn, so a wave's primary rays are
// coherent; lanes whose paths end early idle until the longest path of the
// wave ends, which is what this integrator costs against the wavefront pass
// on a SIMD machine.  No media: rays ignore medium interfaces, as in the
// reference, and meshes without material are rejected by the host.
#include "pathtracer.h"

namespace krr {

// evalMIS(1, p0, 1, p1) with MIS_POWER_HEURISTIC (sampling.h:20-28).
KRR_DEV float misPower(float p0, float p1) {
	const float q0 = p0 * p0, q1 = p1 * p1;
	return q0 / (q0 + q1);
}

KRR_DEV bool finiteWeight(float w) { return !(__builtin_isnan(w) || __builtin_isinf(w)); }

// The samples of pixel (x, y); returns their sum.  TEX = false: the scene
// has no textures, image-lit infinite lights or alpha tests.
template <bool TEX>
KRR_DEV f3 pathTracePixel(const PtArgs &A, const LaneStack &stack, int x, int y, uint32_t &nClosest, uint32_t &nShadow) {
	Pcg rng;
	rng.setSeed((uint64_t) interleave32((uint32_t) x, (uint32_t) y), (uint64_t) A.sampleIndex);
	const DevScene &sc = A.sc;
	const int nL	   = sc.nLights;
	const float selPdf = 1.f / (float) (nL > 0 ? nL : 1);
	uint32_t nv = 0, nt = 0;
	f3 color;
	for (int s = 0; s < A.spp; s++) {
		const RayItem cr = cameraRay(A.cam, A.W, A.H, x, y, rng);
		(void) rng.get1D(); // SampledWavelengths::sampleUniform
		f3 o = cr.o, d = cr.d, thp(1.f), L;
		float pdf	 = 0.f;
		int bsdfType = BSDF_UNSET;
		f3 ctxP;
		for (int depth = 0;; depth++) {
			nClosest++;
			const TraceResult h = traverse<false, false, TEX>(A.bvh, o, d, __builtin_inff(), stack, nv, nt);
			if (h.prim < 0) { // handleMiss (device.cu:67-79)
				for (int li = 0; li < nL; li++) {
					const DevLight Lr = sc.lights[li];
					if (Lr.type != 1) continue;
					float w = 1.f;
					if (A.nee && depth > 0 && !(bsdfType & BSDF_SPECULAR)) {
						w = misPower(pdf, kInv4Pi * selPdf);
						if (!finiteWeight(w)) w = 1.f;
					}
					L = L + thp * w * infiniteLi<TEX>(Lr, d);
				}
				break;
			}
			SurfaceInteraction intr;
			prepareSurfaceInteraction<TEX>(sc, h.prim, h.b1, h.b2, d, -1, intr);
			if (intr.light >= 0) { // handleHit (device.cu:52-65)
				const DevLight Lr = sc.lights[intr.light];
				const f3 Le		  = areaLightL<TEX>(Lr, intr.n, intr.wo, intr.u, intr.v);
				float w			  = 1.f;
				if (A.nee && depth > 0) {
					const float lightPdf = lightPdfLi(sc, Lr, intr.p, intr.n, ctxP) * selPdf;
					if (!(bsdfType & BSDF_SPECULAR)) w = misPower(pdf, lightPdf);
					if (!finiteWeight(w)) w = 1.f;
				}
				L = L + Le * w * thp;
			}
			if (depth == A.maxDepth || (A.rr < 1.f && rng.get1D() > A.rr)) break;
			thp = thp / A.rr;
			const DevMaterial mat = materialAt<TEX>(sc, intr);
			BSDF bsdf;
			bsdf.setup(mat);
			const f3 woL = intr.toLocal(intr.wo);
			if (A.nee && (bsdfTypeOf(mat.type, mat.roughness) & BSDF_SMOOTH) && nL > 0) { // evalDirect (device.cu:101-106)
				const float ul		 = rng.get1D();
				const int li		 = min((int) (uint32_t) (ul * (float) nL), nL - 1);
				const float u0		 = rng.get1D(), u1 = rng.get1D();
				const LightSample ls = sampleLi<TEX>(sc, sc.lights[li], u0, u1, intr.p);
				const f3 wiL		 = intr.toLocal(normalize(ls.p - intr.p));
				const float lightPdf = selPdf * ls.pdf;
				if (lightPdf != 0.f) {
					const float bp = bsdf.pdf(woL, wiL);
					const f3 bv	   = bsdf.f(woL, wiL) * fabsf(wiL.z);
					const float w  = misPower(lightPdf, bp);
					if (finiteWeight(w) && bv.any()) {
						const f3 to	  = offsetRayOrigin(ls.p, ls.n, intr.p - ls.p);
						const f3 p_o  = offsetRayOrigin(intr.p, intr.n, to - intr.p);
						const TraceResult sh = traverse<true, false, TEX>(A.bvh, p_o, to - p_o, 1.f, stack, nv, nt);
						nShadow++;
						if (sh.prim < 0) L = L + thp * bv * w / lightPdf * ls.L;
					}
				}
			}
			// generateScatterRay (device.cu:108-122)
			const BSDFSample bs = bsdf.sample(woL, rng);
			if (bs.pdf == 0.f || !bs.f.any()) break;
			const f3 wiW = intr.toWorld(bs.wi);
			bsdfType	 = bs.flags;
			o			 = offsetRayOrigin(intr.p, intr.n, wiW);
			d			 = wiW;
			ctxP		 = intr.p;
			pdf			 = bs.pdf;
			thp			 = thp * (bs.f * fabsf(bs.wi.z) / bs.pdf);
			if (!thp.any()) break;
		}
		color = color + L;
	}
	return color;
}

template <bool TEX> __global__ void __launch_bounds__(256) k_path_tracer(PtArgs A) {
	extern __shared__ int ldsStack[];
	const LaneStack stack = laneStack(A.bvh, ldsStack, A.ovf);
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
	uint32_t nClosest = 0, nShadow = 0;
	// wave-uniform loop over tiles: lanes of a wave stay together
	for (int tile = blockIdx.x * waves + wave; tile < A.nTiles; tile += gridDim.x * waves) {
		const int x = (tile % A.tilesX) * 8 + (lane & 7);
		const int y = (tile / A.tilesX) * 8 + (lane >> 3);
		if (x < A.W && y < A.H) {
			const f3 c		 = pathTracePixel<TEX>(A, stack, x, y, nClosest, nShadow);
			const float inv	 = (float) A.spp;
			A.film[(size_t) (A.H - 1 - y) * A.W + x] = make_float4(c.x / inv, c.y / inv, c.z / inv, 1.f);
		}
	}
	if (nClosest) atomicAdd(&A.rays[0], (unsigned long long) nClosest);
	if (nShadow) atomicAdd(&A.rays[1], (unsigned long long) nShadow);
}

__global__ void k_path_tracer_clear(unsigned long long *rays) {
	if (threadIdx.x < 2) rays[threadIdx.x] = 0ull;
}

hipError_t launchPathTracer(const PtArgs &a, int grid, int tb, size_t lds, hipStream_t st) {
	hipLaunchKernelGGL(k_path_tracer_clear, dim3(1), dim3(64), 0, st, a.rays);
	if (a.textured) hipLaunchKernelGGL(k_path_tracer<true>, dim3(grid), dim3(tb), lds, st, a);
	else hipLaunchKernelGGL(k_path_tracer<false>, dim3(grid), dim3(tb), lds, st, a);
	return hipGetLastError();
}

int pathTracerBlocksPerCU(int tb, size_t lds, bool textured) {
	int nb = 0;
	const hipError_t e = textured ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_path_tracer<true>, tb, lds)
								  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_path_tracer<false>, tb, lds);
	if (e != hipSuccess || nb < 1) nb = 1;
	return nb;
}

} // namespace krr
