// wavefront.hip - gfx950 wavefront path tracer: kernels and the C ABI.
//
// Per sample and loop depth the reference launches camera generation,
// traceClosest (OptiX raygen + closest-hit/miss programs), handleHit,
// handleMiss, generateScatterRays and traceShadow as separate kernels that
// hand work over through SoA queues (src/render/wavefront/integrator.cpp:
// 328-372).  Here the closest-hit routing, medium sampling, handleHit,
// handleMiss, medium scattering and generateScatterRays of one ray run in a
// single "shade" kernel: a path owns at most one item per queue per stage, so
// running those stages back to back in one thread keeps the reference's
// per-pixel order of radiance additions and RNG draws while the hit-light,
// miss, scatter and medium queues (464 B per path per bounce in the
// reference layout) never touch HBM.  What stays in memory is what crosses a
// traversal: the ray queues (double-buffered), the hit records and the shadow
// queue.  Queue slots are reserved with one atomic per wave.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "bvh.h"
#include "bvh_build.h"
#include "errors.h"
#include "json.h"
#include "kernels.h"
#include "nrc_dev.h"
#include "pathtracer.h"
#include "krr_wavefront.h"

namespace krr {

std::string &lastErrorStorage() {
	static thread_local std::string s;
	return s;
}

// ===========================================================================
// Kernels
// ===========================================================================

__global__ void k_reset_totals(Counters *cnt) {
	cnt->closestTotal	 = 0;
	cnt->shadowTotal	 = 0;
	cnt->nodeVisits		  = 0;
	cnt->triTests		  = 0;
	cnt->shadowNodeVisits = 0;
	cnt->shadowTriTests	  = 0;
	cnt->closestLaunches  = 0;
	cnt->megaClosest	  = 0;
	cnt->megaShadow		  = 0;
	cnt->megaNodeVisits	  = 0;
	cnt->megaTriTests	  = 0;
	cnt->nrcQueries		  = 0;
	cnt->nrcTrain		  = 0;
	cnt->nrcTrainDone	  = 0;
	cnt->nrcQueriesTotal  = 0;
	cnt->rayCount[0] = cnt->rayCount[1] = cnt->shadowCount = 0;
}

// beginFrame (integrator.cpp:205-221): reset L and pixel, seed the sampler,
// draw the wavelength sample whose main index picks the hero channel.
__global__ void k_begin_frame(WfParams P, PixelStateDev px) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= P.N) return;
	const int x = i % P.W, y = i / P.W;
	px.L[i]		= make_float4(0.f, 0.f, 0.f, 0.f);
	px.pixel[i] = make_float4(0.f, 0.f, 0.f, 0.f);
	Pcg s;
	s.setSeed(interleave32((uint32_t) x, (uint32_t) y), (uint64_t) P.sampleIndex);
	s.advance((int64_t) (256 * i));
	const float u		= s.get1D();
	const float lambda0 = (1.f - u) * 360.f + u * 830.f;
	int ch				= (int) ((lambda0 - 360.f) / 470.f * 3);
	px.channel[i]		= ch < 0 ? 0 : (ch > 2 ? 2 : ch);
	px.rng[i]			= s.state;
}

// generateCameraRays (integrator.cpp:166-179).  Camera rays go to slot =
// pixelId, so the first trace sees image-coherent rays.
__global__ void k_camera(WfParams P, CameraDev cam, PixelStateDev px, RayQueueDev q, Counters *cnt) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i == 0) cnt->rayCount[0] = P.N;
	if (i >= P.N) return;
	if (!P.legacySpp) px.L[i] = make_float4(0.f, 0.f, 0.f, 0.f);
	Pcg s{px.rng[i], P.inc};
	RayItem r = cameraRay(cam, P.W, P.H, i % P.W, i / P.W, s);
	r.pixel	  = i;
	storeRayItem(q, i, r);
	px.rng[i] = s.state;
}

// ---------------------------------------------------------------------------
// Many views in one launch.  The pixels of all views are numbered one after
// the other (view v owns [start[v], start[v+1])); pixel state, queues and the
// radiance cache work on that global number.  Frame-varying values come from
// a FrameDev in device memory, so one captured graph serves every frame.
// ---------------------------------------------------------------------------

// View of the global pixel i (largest v with start[v] <= i).
KRR_DEV int findView(const ViewsDev &V, int i) {
	int lo = 0, hi = V.n - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (V.start[mid] <= i) lo = mid;
		else hi = mid - 1;
	}
	return lo;
}

__global__ void k_set_frame(FrameDev *f, FrameDev v) { *f = v; }

// beginFrame for every view.  View v seeds its pixel (x, y) with initstate
// interleave(x, y) | v << 32, so view 0 gets exactly the single-view seed.
__global__ void k_begin_views(WfParams P, ViewsDev V, PixelStateDev px) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= P.N) return;
	const int v		= findView(V, i);
	const int local = i - V.start[v];
	const int W		= V.view[v].W;
	const int x = local % W, y = local / W;
	px.L[i]		= make_float4(0.f, 0.f, 0.f, 0.f);
	px.pixel[i] = make_float4(0.f, 0.f, 0.f, 0.f);
	Pcg s;
	s.setSeed((uint64_t) interleave32((uint32_t) x, (uint32_t) y) | ((uint64_t) (uint32_t) v << 32), (uint64_t) V.frame->sampleIndex);
	s.advance((int64_t) (256 * local));
	const float u		= s.get1D();
	const float lambda0 = (1.f - u) * 360.f + u * 830.f;
	int ch				= (int) ((lambda0 - 360.f) / 470.f * 3);
	px.channel[i]		= ch < 0 ? 0 : (ch > 2 ? 2 : ch);
	px.rng[i]			= s.state;
}

// The pixel state a path updates (PixelStateBuffer's L, sampler and hero
// channel, workqueue.h:15-26): the queue kernels read it from memory and
// write it back per stage; the persistent kernel keeps it in registers from
// the path's first vertex to its last.  The radiance additions are the same
// in number and order either way.
struct PathRegs {
	f3 L;
	uint64_t rng;
	int channel;
};

KRR_DEV PathRegs loadPathRegs(const PixelStateDev &px, int pixel, bool media) {
	const float4 L = px.L[pixel];
	return PathRegs{f3(L.x, L.y, L.z), px.rng[pixel], media ? px.channel[pixel] : 0};
}

KRR_DEV void storePathRegs(const PixelStateDev &px, int pixel, const PathRegs &pr) {
	px.L[pixel]	  = make_float4(pr.L.x, pr.L.y, pr.L.z, 0.f);
	px.rng[pixel] = pr.rng;
}

// Camera ray of global pixel i for the current sample (k_camera per view);
// the pixel's radiance (zero unless the reference's spp loop keeps it) and
// sampler state after the camera draws go to pr.
KRR_DEV RayItem viewCameraRay(const WfParams &P, const ViewsDev &V, const PixelStateDev &px, int i, PathRegs &pr, bool media) {
	const int v		= findView(V, i);
	const int local = i - V.start[v];
	const ViewDev &vw = V.view[v];
	pr				  = loadPathRegs(px, i, media);
	if (!P.legacySpp) pr.L = f3();
	Pcg s{pr.rng, P.inc};
	RayItem r = cameraRay(vw.cam, vw.W, vw.H, local % vw.W, local / vw.W, s);
	r.pixel	  = i;
	pr.rng	  = s.state;
	return r;
}

// Film write of every view (k_film per view).
__global__ void k_film_views(WfParams P, ViewsDev V, PixelStateDev px) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= P.N) return;
	const int v		  = findView(V, i);
	const int local	  = i - V.start[v];
	const ViewDev &vw = V.view[v];
	const int x = local % vw.W, y = local / vw.W;
	const float4 p = px.pixel[i];
	float r = p.x / (float) P.spp, g = p.y / (float) P.spp, b = p.z / (float) P.spp;
	if (P.enableClamp) {
		r = clampf(r, 0.f, P.clampMax);
		g = clampf(g, 0.f, P.clampMax);
		b = clampf(b, 0.f, P.clampMax);
	}
	vw.film[(size_t) (vw.H - 1 - y) * vw.W + x] = make_float4(r, g, b, 1.f);
}

__global__ void k_depth_begin(Counters *cnt, int cur) {
	const int nxt = cur ^ 1;
	cnt->closestTotal += (unsigned long long) cnt->rayCount[cur];
	cnt->closestLaunches += 1;
	cnt->rayCount[nxt] = 0;
	cnt->shadowCount   = 0;
	cnt->work[0] = cnt->work[1] = 0;
}

// Closest-hit trace of one ray queue entry into its hit record.
template <bool COUNT>
KRR_DEV void traceClosestItem(const TraceBvh &bvh, const float4 *__restrict__ rays, float4 *__restrict__ hits, int i,
							  const LaneStack &stack, uint32_t &nv, uint32_t &nt) {
	const float4 a = rays[2 * i], b = rays[2 * i + 1];
	TraceResult r  = traverse<false, COUNT, true>(bvh, f3(a.x, a.y, a.z), f3(b.x, b.y, b.z), a.w, stack, nv, nt);
	hits[i]		   = make_float4(r.prim >= 0 ? r.t : __builtin_inff(), __int_as_float(r.prim), r.b1, r.b2);
}

// Persistent closest-hit trace of the current ray queue: a grid sized to the
// resident capacity of the GPU, waves taking 64 rays at a time until the
// count in device memory is drained (no host round trip, no launch of empty
// workgroups when the queue is short, per-wave load balance).
template <bool COUNT>
__global__ void k_trace_closest_queue(TraceBvh bvh, const float4 *__restrict__ rays, float4 *__restrict__ hits, const int *countPtr,
									  int *ovf, Counters *cnt) {
	extern __shared__ int ldsStack[];
	const LaneStack stack = laneStack(bvh, ldsStack, ovf);
	const int count		  = *countPtr;
	uint32_t nv = 0, nt = 0;
	while (true) {
		const int base = waveFetch(&cnt->work[0]);
		if (base >= count) break;
		const int i = base + laneId();
		if (i < count) traceClosestItem<COUNT>(bvh, rays, hits, i, stack, nv, nt);
	}
	if (COUNT) {
		waveAddU64(&cnt->nodeVisits, nv);
		waveAddU64(&cnt->triTests, nt);
	}
}

__global__ void k_accel_begin(Counters *cnt) { cnt->work[0] = 0; }

// Closest-hit trace of a caller's ray array (krr_accel_trace_closest),
// persistent like the queue version.
template <bool COUNT>
__global__ void k_trace_closest(TraceBvh bvh, const float4 *__restrict__ rays, float4 *__restrict__ hits, int count, int *ovf,
								Counters *cnt) {
	extern __shared__ int ldsStack[];
	const LaneStack stack = laneStack(bvh, ldsStack, ovf);
	uint32_t nv = 0, nt = 0;
	while (true) {
		const int base = waveFetch(&cnt->work[0]);
		if (base >= count) break;
		const int i = base + laneId();
		if (i < count) traceClosestItem<COUNT>(bvh, rays, hits, i, stack, nv, nt);
	}
	if (COUNT) {
		waveAddU64(&cnt->nodeVisits, nv);
		waveAddU64(&cnt->triTests, nt);
	}
}

// Any-hit trace with a visibility result per ray (krr_accel_trace_any).
template <bool COUNT>
__global__ void k_trace_any(TraceBvh bvh, const float4 *__restrict__ rays, int *__restrict__ visible, int count, int *ovf,
							Counters *cnt) {
	extern __shared__ int ldsStack[];
	const LaneStack stack = laneStack(bvh, ldsStack, ovf);
	uint32_t nv = 0, nt = 0;
	while (true) {
		const int base = waveFetch(&cnt->work[0]);
		if (base >= count) break;
		const int i = base + laneId();
		if (i < count) {
			const float4 a = rays[2 * i], b = rays[2 * i + 1];
			TraceResult r  = traverse<true, COUNT, true>(bvh, f3(a.x, a.y, a.z), f3(b.x, b.y, b.z), a.w, stack, nv, nt);
			visible[i]	   = r.prim < 0 ? 1 : 0;
		}
	}
	if (COUNT) {
		waveAddU64(&cnt->nodeVisits, nv);
		waveAddU64(&cnt->triTests, nt);
	}
}

// traceShadow without media (device.cu:83-100): any hit on a non-null
// triangle in (0, 1) occludes; visible shadow rays add Ld / mean(pu + pl).
template <bool COUNT>
__global__ void k_trace_shadow(TraceBvh bvh, ShadowQueueDev sq, PixelStateDev px, int *ovf, Counters *cnt);

// sampleT_maj (wavefront.h:34-78) on the medium of the ray.  callback(p,
// mp, sigma_maj, T_maj) returns false to stop; the result is the remaining
// majorant transmittance (or 1 when stopped by the callback).
template <class F>
KRR_DEV f3 sampleTmaj(const DevScene &sc, int medium, f3 o, f3 d, float tMax, Pcg &rng, int channel, F &&callback) {
	tMax			   = tMax * length(d);
	d				   = normalize(d);
	f3 T_maj		   = f3(1.f);
	const DevMedium &m = sc.media[medium];
	MajorantIter it;
	sampleRay(m, o, d, tMax, it);
	MajorantSeg seg;
	while (true) {
		if (!it.next(seg)) return T_maj;
		if (seg.sigma_maj[channel] == 0.f) {
			float dt = seg.tMax - seg.tMin;
			if (__builtin_isinf(dt)) dt = kFloatMax;
			T_maj = T_maj * vexp(-dt * seg.sigma_maj);
			continue;
		}
		float tMin = seg.tMin;
		while (true) {
			const float t = tMin + sampleExponential(rng.get1D(), seg.sigma_maj[channel]);
			if (t < seg.tMax) {
				T_maj			  = T_maj * vexp(-(t - tMin) * seg.sigma_maj);
				const f3 p		  = o + d * t;
				MediumProps mp	  = samplePoint(m, p);
				if (!callback(p, mp, seg.sigma_maj, T_maj)) return f3(1.f);
				T_maj = f3(1.f);
				tMin  = t;
			} else {
				float dt = seg.tMax - tMin;
				if (__builtin_isinf(dt)) dt = kFloatMax;
				T_maj = T_maj * vexp(-dt * seg.sigma_maj);
				break;
			}
		}
	}
}

// The fused shading stage for the rays of the current queue.  Per ray, in the
// reference's launch order: closest-hit / miss routing (device.cu:43-74),
// sampleMediumInteraction (medium.cpp:13-103), handleHit
// (integrator.cpp:78-90), handleMiss (:197-213), the depth == maxDepth
// break, sampleMediumScattering (medium.cpp:105-153) and
// generateScatterRays (integrator.cpp:110-164).
struct ShadowOut {
	f3 o, d, Ld, pu, pl;
	int medium;
};


// One ray of the queue through the stages above; new work is returned in
// (wantNext, nx) and (wantShadow, so), radiance and RNG state in pr.
// Spherical angles of a unit vector, mapped to [0,1]: theta / pi, phi / (2 pi) + 1/2.
KRR_DEV void unitAngles(const f3 &w, float &th, float &ph) {
	const float z = fminf(1.f, fmaxf(-1.f, w.z));
	th			  = detmath::atan2f(sqrtf(w.x * w.x + w.y * w.y), z) * kInvPi;
	ph			  = detmath::atan2f(w.y, w.x) * (0.5f * kInvPi) + 0.5f;
}

// The krr_nrc_vertex of a surface vertex: position in the scene box,
// outgoing direction, shading normal, roughness, diffuse albedo (diffuse
// materials) and specular albedo (the base colour of conductors and
// dielectrics).
template <bool TEX> KRR_DEV void nrcVertex(const DevScene &sc, const SurfaceInteraction &intr, float4 *out) {
	const DevMaterial mat = materialAt<TEX>(sc, intr);
	const float px		  = fminf(1.f, fmaxf(0.f, (intr.p.x - sc.bmin[0]) * sc.binv[0]));
	const float py		  = fminf(1.f, fmaxf(0.f, (intr.p.y - sc.bmin[1]) * sc.binv[1]));
	const float pz		  = fminf(1.f, fmaxf(0.f, (intr.p.z - sc.bmin[2]) * sc.binv[2]));
	float dt, dp, nt, np;
	unitAngles(intr.wo, dt, dp);
	unitAngles(intr.n, nt, np);
	const bool diffuse = mat.type == KRR_MAT_DIFFUSE;
	const float d0 = diffuse ? mat.diffuse[0] : 0.f, d1 = diffuse ? mat.diffuse[1] : 0.f, d2 = diffuse ? mat.diffuse[2] : 0.f;
	const float s0 = diffuse ? 0.f : mat.diffuse[0], s1 = diffuse ? 0.f : mat.diffuse[1], s2 = diffuse ? 0.f : mat.diffuse[2];
	out[0] = make_float4(px, py, pz, dt);
	out[1] = make_float4(dp, nt, np, mat.roughness);
	out[2] = make_float4(d0, d1, d2, s0);
	out[3] = make_float4(s1, s2, 0.f, 0.f);
}

// NRC (template flag) in the shading stage, before generateScatterRays.
// A render path ends at the first surface vertex (depth >= 1) where the
// path's spread has grown past c * a0 (Mueller et al. 2021, sec. 5: a0 =
// |x0 - x1|^2 / (4 pi |cos1|), spread = (sum_i>=2 sqrt(|x_i-1 - x_i|^2 /
// (pdf_i-1 |cos_i|)))^2, delta bounces add nothing): the vertex's emission is
// added, its light sample (NEE) is traced as usual, and a cache query stands
// for the rest - the radiance the BSDF-sampled continuation brings.  A
// training path (pixel chosen by trainHash) offers every scattering vertex;
// a uniform hash keeps each with probability Counters::nrcKeep, so the kept
// vertices spread over the whole frame.  A kept vertex records its throughput
// now and the radiance gathered up to and including its own NEE afterwards
// (*trainSlot, written by the caller after the shadow ray), and the path
// continues unbiased.
// PLAIN: no textures and no image-lit infinite lights in the scene; the
// texture lookups are compiled out.
template <bool MEDIA, bool NRC = false, bool PLAIN = false>
KRR_DEV void shadeCore(const WfParams &P, const DevScene &sc, const RayItem &r, const float4 h, PathRegs &pr, int loopDepth,
					   bool &wantNext, RayItem &nx, bool &wantShadow, ShadowOut &so, const NrcDev *nrc = nullptr, bool trainPath = false,
					   int *trainSlot = nullptr) {
	const int pixel = r.pixel;
	f3 &sO = so.o, &sD = so.d, &sLd = so.Ld, &sPu = so.pu, &sPl = so.pl;
	int &sMedium = so.medium;
	const int prim	= __float_as_int(h.y);
	Pcg rng{pr.rng, P.inc};
	const int channel = MEDIA ? pr.channel : 0;
	f3 L			  = pr.L;
	f3 thp = r.thp, pu = r.pu, pl = r.pl;
	bool hasHL = false, hasScatter = false, hasMiss = false, hasMS = false, hasMSc = false;
	SurfaceInteraction intr;
	float msTmax   = __builtin_inff();
	bool msHasIntr = false;
	f3 mscP, mscThp, mscPu, mscWo;
	float mscG = 0.f;
	if (prim >= 0) {
		prepareSurfaceInteraction<!PLAIN>(sc, prim, h.z, h.w, r.d, r.medium, intr);
		if (MEDIA && P.enableMedium && r.medium >= 0) {
			hasMS	  = true;
			msTmax	  = h.x;
			msHasIntr = true;
		} else if (intr.material < 0) {
			wantNext  = true;
			nx		  = r;
			nx.o	  = offsetRayOrigin(intr.p, intr.n, r.d);
			nx.medium = intr.getMedium(r.d);
		} else {
			if (intr.light >= 0) hasHL = true;
			if (thp.any()) hasScatter = true;
		}
	} else {
		if (MEDIA && P.enableMedium && r.medium >= 0) hasMS = true;
		else hasMiss = true;
	}

	if (MEDIA && hasMS) {
		f3 Lm;
		bool scattered = false;
		const f3 T_maj = sampleTmaj(sc, r.medium, r.o, r.d, msTmax, rng, channel,
			[&](const f3 &p, const MediumProps &mp, const f3 &sigma_maj, const f3 &Tm) -> bool {
				if (r.depth < P.maxDepth && mp.Le.any()) {
					const float pr = sigma_maj[channel] * Tm[channel];
					const f3 pe	   = pu * sigma_maj * Tm / pr;
					if (pe.any()) Lm = Lm + thp * mp.sigma_a * Tm * mp.Le / (pr * pe.mean());
				}
				const float pAbsorb	 = mp.sigma_a[channel] / sigma_maj[channel];
				const float pScatter = mp.sigma_s[channel] / sigma_maj[channel];
				const float pNull	 = maxf(0.f, 1.f - pAbsorb - pScatter);
				const float w[3]	 = {pAbsorb, pScatter, pNull};
				const int mode		 = sampleDiscrete3(w, rng.get1D());
				if (mode == 0) {
					thp = f3();
					return false;
				} else if (mode == 1) {
					const float pr = Tm[channel] * mp.sigma_s[channel];
					thp			   = thp * (Tm * mp.sigma_s / pr);
					pu			   = pu * (Tm * mp.sigma_s / pr);
					if (thp.any() && pu.any()) {
						hasMSc = true;
						mscP   = p;
						mscThp = thp;
						mscPu  = pu;
						mscWo  = -r.d;
						mscG   = mp.g;
					}
					scattered = true;
					return false;
				} else {
					const f3 sigma_n = vmaxs(sigma_maj - mp.sigma_a - mp.sigma_s, 0.f);
					const float pr	 = Tm[channel] * sigma_n[channel];
					thp				 = thp * (Tm * sigma_n / pr);
					if (pr == 0.f) thp = f3();
					pu = pu * (Tm * sigma_n / pr);
					pl = pl * (Tm * sigma_maj / pr);
					return thp.any() && pu.any();
				}
			});
		if (Lm.any()) L = L + Lm;
		if (!scattered && thp.any()) {
			thp = thp * (T_maj / T_maj[channel]);
			pu	= pu * (T_maj / T_maj[channel]);
			pl	= pl * (T_maj / T_maj[channel]);
		}
		if (!(scattered || !thp.any() || !pu.any() || r.depth == P.maxDepth)) {
			if (!msHasIntr) hasMiss = true;
			else if (intr.material < 0) {
				wantNext  = true;
				nx		  = r;
				nx.o	  = offsetRayOrigin(intr.p, intr.n, r.d);
				nx.medium = intr.getMedium(r.d);
				nx.thp	  = thp;
				nx.pu	  = pu;
				nx.pl	  = pl;
			} else {
				if (intr.light >= 0) hasHL = true;
				hasScatter = true;
			}
		}
	}

	const float lightSelectPdf = 1.f / (float) sc.nLights;
	if (hasHL) { // handleHit
		const DevLight Lr = sc.lights[intr.light];
		f3 Le			  = areaLightL<!PLAIN>(Lr, intr.n, intr.wo, intr.u, intr.v) * thp;
		if (P.nee && r.depth && !(r.bsdfType & BSDF_DELTA)) {
			const float lightPdf = lightPdfLi(sc, Lr, intr.p, intr.n, r.cp) * lightSelectPdf;
			Le					 = Le / (pl * lightPdf + pu).mean();
		} else Le = Le / pu.mean();
		L = L + Le;
	}
	if (hasMiss) { // handleMiss
		f3 Lm;
		for (int li = 0; li < sc.nLights; li++) {
			const DevLight Lr = sc.lights[li];
			if (Lr.type != 1) continue;
			const f3 Li = infiniteLi<!PLAIN>(Lr, r.d);
			if (P.nee && r.depth && !(r.bsdfType & BSDF_DELTA)) {
				const float lp = kInv4Pi * lightSelectPdf;
				Lm			   = Lm + Li / (pu + pl * lp).mean();
			} else Lm = Lm + Li / pu.mean();
		}
		L = L + thp * Lm;
	}

	bool nrcEnd = false; // this vertex ends the path in the cache
	float sp0 = r.sp0, spSum = r.spSum;
	if (NRC) {
		const bool live = hasScatter && loopDepth != P.maxDepth;
		if (live) { // path spread at this vertex
			const float t  = h.x;
			const float cs = fmaxf(fabsf(dot(intr.n, r.d)), 1e-4f);
			if (r.depth == 0) sp0 = t * t / (4.f * kPi * cs);
			else if (r.pdfPrev > 0.f) spSum += sqrtf(t * t / (r.pdfPrev * cs));
		}
		const bool wantQ = live && !trainPath && r.depth >= 1 && spSum * spSum > nrc->spread * sp0;
		const bool cand	 = live && trainPath;
		uint32_t hk		 = trainHash((uint32_t) pixel ^ ((uint32_t) r.depth * 0x9E3779B9u) ^ (uint32_t) (P.inc >> 1) * 0x85EBCA6Bu);
		const bool wantT = cand && (float) (hk >> 8) * (1.f / 16777216.f) < nrc->keepProb;
		waveCount(nrc->candCount, cand);
		int qi, ti;
		wavePush2(nrc->queryCount, wantQ, wantT, qi, ti);
		if (wantQ) {
			nrcEnd = true;
			if (qi < nrc->capQuery) {
				float4 *q = nrc->query + (size_t) qi * 5;
				nrcVertex<!PLAIN>(sc, intr, q);
				q[4] = make_float4(thp.x, thp.y, thp.z, __int_as_float(pixel));
			}
		}
		if (wantT && ti < nrc->capTrain) {
			float4 *t = nrc->train + (size_t) ti * 6;
			nrcVertex<!PLAIN>(sc, intr, t);
			t[4] = make_float4(thp.x, thp.y, thp.z, __int_as_float(pixel));
			*trainSlot = ti;
		}
	}

	if (loopDepth != P.maxDepth) {
		if (MEDIA && hasMSc) { // sampleMediumScattering
			if (P.nee && sc.nLights > 0) {
				const float ul	 = rng.get1D();
				const int li	 = min((int) (uint32_t) (ul * (float) sc.nLights), sc.nLights - 1);
				const float u0	 = rng.get1D(), u1 = rng.get1D();
				const LightSample ls = sampleLi<!PLAIN>(sc, sc.lights[li], u0, u1, mscP);
				const f3 to		 = offsetRayOrigin(ls.p, ls.n, mscP - ls.p);
				const f3 p_o	 = offsetRayOrigin(mscP, f3(), to - mscP);
				const f3 sdir	 = to - p_o;
				const f3 wi		 = normalize(sdir);
				const f3 thpS	 = mscThp * hgP(mscG, mscWo, wi);
				const float lightPdf = lightSelectPdf * ls.pdf;
				const float phasePdf = hgP(mscG, mscWo, wi);
				const f3 Ld			 = thpS * ls.L;
				if (Ld.any() && lightPdf > 0.f) {
					wantShadow = true;
					sO		   = p_o;
					sD		   = sdir;
					sLd		   = Ld;
					sPu		   = mscPu * phasePdf;
					sPl		   = mscPu * lightPdf;
					sMedium	   = r.medium;
				}
			}
			const float u0 = rng.get1D(), u1 = rng.get1D();
			float ppdf;
			const f3 wi	   = hgSample(mscG, mscWo, u0, u1, &ppdf);
			f3 thpN		   = mscThp * ppdf / ppdf;
			const float rrProb = (thpN / mscPu.mean()).maxCoeff();
			bool alive		   = true;
			if (r.depth >= 1 && rrProb < 1.f) {
				if (rng.get1D() >= rrProb) alive = false;
				else thpN = thpN / rrProb;
			}
			if (alive && !thpN.isZero() && !thpN.hasNaN()) {
				wantNext	= true;
				nx.o		= mscP;
				nx.d		= wi;
				nx.medium	= r.medium;
				nx.cp		= mscP;
				nx.cn		= f3();
				nx.thp		= thpN;
				nx.pu		= mscPu;
				nx.pl		= mscPu / ppdf;
				nx.depth	= r.depth + 1;
				nx.bsdfType = BSDF_SMOOTH;
				nx.pixel	= pixel;
			}
		}
		if (hasScatter) { // generateScatterRays
			bool alive = true;
			if (r.depth >= P.rrMinDepth && !nrcEnd) {
				if (rng.get1D() >= P.rr) alive = false;
				else thp = thp / P.rr;
			}
			if (alive) {
				const f3 woL		  = intr.toLocal(intr.wo);
				const DevMaterial mat = materialAt<!PLAIN>(sc, intr);
				const int bt		  = bsdfTypeOf(mat.type, mat.roughness);
				BSDF bsdf;
				bsdf.setup(mat);
				if (P.nee && (bt & BSDF_SMOOTH) && sc.nLights > 0) {
					const float ul		 = rng.get1D();
					const int li		 = min((int) (uint32_t) (ul * (float) sc.nLights), sc.nLights - 1);
					const float u0		 = rng.get1D(), u1 = rng.get1D();
					const LightSample ls = sampleLi<!PLAIN>(sc, sc.lights[li], u0, u1, intr.p);
					const f3 to			 = offsetRayOrigin(ls.p, ls.n, intr.p - ls.p);
					const f3 p_o		 = offsetRayOrigin(intr.p, intr.n, to - intr.p);
					const f3 sdir		 = to - p_o;
					const f3 wiL		 = intr.toLocal(normalize(sdir));
					const float lightPdf = lightSelectPdf * ls.pdf;
					const f3 bv			 = bsdf.f(woL, wiL);
					const float bp		 = bsdf.pdf(woL, wiL);
					if (lightPdf > 0.f && bv.any()) {
						const f3 Ld = ls.L * thp * bv * fabsf(wiL.z);
						if (Ld.any()) {
							wantShadow = true;
							sO		   = p_o;
							sD		   = sdir;
							sLd		   = Ld;
							sPu		   = pu * bp;
							sPl		   = pu * lightPdf;
							sMedium	   = intr.getMedium(sdir);
						}
					}
				}
				const BSDFSample s = nrcEnd ? BSDFSample{} : bsdf.sample(woL, rng);
				if (!nrcEnd && s.pdf != 0.f && s.f.any()) {
					const f3 wiW  = intr.toWorld(s.wi);
					const f3 nthp = thp * s.f * fabsf(s.wi.z) / s.pdf;
					if (nthp.any()) {
						wantNext	= true;
						nx.bsdfType = s.flags;
						nx.pu		= pu;
						nx.pl		= pu / s.pdf;
						nx.o		= offsetRayOrigin(intr.p, intr.n, wiW);
						nx.d		= wiW;
						nx.medium	= intr.getMedium(wiW);
						nx.cp		= intr.p;
						nx.cn		= intr.n;
						nx.depth	= r.depth + 1;
						nx.thp		= nthp;
						nx.pixel	= pixel;
						nx.sp0		= sp0;
						nx.spSum	= spSum;
						nx.pdfPrev	= (s.flags & BSDF_DELTA) ? 0.f : s.pdf;
					}
				}
			}
		}
	}
	pr.L   = L;
	pr.rng = rng.state;
}

template <bool MEDIA>
KRR_DEV void shadeItem(const WfParams &P, const DevScene &sc, const RayQueueDev &cur, const float4 *__restrict__ hits,
					   const RayQueueDev &nxt, const ShadowQueueDev &sq, const PixelStateDev &px, Counters *cnt, int curIdx,
					   int loopDepth, const int i, const int count) {
	bool wantNext = false, wantShadow = false;
	RayItem nx;
	ShadowOut so;
	so.medium = -1;
	int pixel = 0;
	if (i < count) {
		const RayItem r = loadRayItem(cur, i);
		pixel			= r.pixel;
		PathRegs pr		= loadPathRegs(px, pixel, MEDIA);
		shadeCore<MEDIA>(P, sc, r, hits[i], pr, loopDepth, wantNext, nx, wantShadow, so);
		storePathRegs(px, pixel, pr);
	}
	const f3 &sO = so.o, &sD = so.d, &sLd = so.Ld, &sPu = so.pu, &sPl = so.pl;
	const int sMedium = so.medium;
	const int slotN = wavePush(&cnt->rayCount[curIdx ^ 1], wantNext);
	const int slotS = wavePush(&cnt->shadowCount, wantShadow);
	if (wantNext) storeRayItem(nxt, slotN, nx);
	if (wantShadow) {
		sq.ray[2 * slotS]	  = make_float4(sO.x, sO.y, sO.z, 1.f);
		sq.ray[2 * slotS + 1] = make_float4(sD.x, sD.y, sD.z, __int_as_float(pixel));
		if (MEDIA && P.enableMedium) {
			sq.c[slotS]			  = make_float4(0.f, 0.f, 0.f, __int_as_float(sMedium + 1));
			sq.ext[3 * slotS]	  = make_float4(sLd.x, sLd.y, sLd.z, 0.f);
			sq.ext[3 * slotS + 1] = make_float4(sPu.x, sPu.y, sPu.z, 0.f);
			sq.ext[3 * slotS + 2] = make_float4(sPl.x, sPl.y, sPl.z, 0.f);
		} else {
			const f3 c	= sLd / (sPl + sPu).mean();
			sq.c[slotS] = make_float4(c.x, c.y, c.z, __int_as_float(sMedium + 1));
		}
	}
}

// Grid-stride loop over the current queue (count read on the device); the
// loop bound is block-uniform, so every lane of a wave reaches the
// wave-aggregated pushes together.
template <bool MEDIA>
__global__ void __launch_bounds__(256) k_shade(WfParams P, DevScene sc, RayQueueDev cur, const float4 *__restrict__ hits,
											  RayQueueDev nxt, ShadowQueueDev sq, PixelStateDev px, Counters *cnt, int curIdx,
											  int loopDepth) {
	const int count = cnt->rayCount[curIdx];
	for (int base = blockIdx.x * blockDim.x; base < count; base += gridDim.x * blockDim.x)
		shadeItem<MEDIA>(P, sc, cur, hits, nxt, sq, px, cnt, curIdx, loopDepth, base + threadIdx.x, count);
}

// traceShadow with media: ShadowTr raygen + traceTransmittance
// (device.cu:102-128, wavefront.h:80-139).  Ratio tracking through media,
// stepping across null-material surfaces.
constexpr int kMaxNullCrossings = 1024;

// traceTransmittance of one shadow ray (tMax = 1 to the offset light point):
// ratio tracking through media, stepping across null-material surfaces.
template <bool COUNT, bool ALPHA = true>
KRR_DEV void transmittanceItem(const WfParams &P, const DevScene &sc, const TraceBvh &bvh, f3 o, f3 d, const float tMax,
							   const int origMedium, const f3 &Ld, const f3 &spu, const f3 &spl, PathRegs &pr, const LaneStack &stack,
							   uint32_t &nv, uint32_t &nt) {
	Pcg rng{pr.rng, P.inc};
	const int channel = pr.channel;
	int medium		  = origMedium;
	const f3 pLight	  = o + d * tMax;
	f3 T_ray(1.f), pu(1.f), pl(1.f);
	SurfaceInteraction intr;
	const float zeroEps = kRayEps * 2.f;
	// The reference loops until the segment is shorter than 2 KRR_RAY_EPS;
	// the cap only bounds pathological stacks of coplanar null surfaces.
	for (int crossing = 0; crossing < kMaxNullCrossings && !(fabsf(d.x) <= zeroEps && fabsf(d.y) <= zeroEps && fabsf(d.z) <= zeroEps);
		 crossing++) {
		const TraceResult h = traverse<false, COUNT, ALPHA>(bvh, o, d, tMax, stack, nv, nt);
		const bool visible	= h.prim < 0;
		// only a null-material surface lets the ray on, and it has no normal map
		if (!visible) prepareSurfaceInteraction<false>(sc, h.prim, h.b1, h.b2, d, origMedium, intr);
		if (!visible && intr.material >= 0) {
			T_ray = f3();
			break;
		}
		if (medium >= 0) {
			const float tEnd = visible ? tMax : length(intr.p - o) / length(d);
			const f3 T_maj	 = sampleTmaj(sc, medium, o, d, tEnd, rng, channel,
				  [&](const f3 &, const MediumProps &mp, const f3 &sigma_maj, const f3 &Tm) -> bool {
					  const f3 sigma_n = vmaxs(sigma_maj - mp.sigma_a - mp.sigma_s, 0.f);
					  const float pr   = Tm[channel] * sigma_maj[channel];
					  T_ray			   = T_ray * (Tm * sigma_n / pr);
					  pl			   = pl * (Tm * sigma_maj / pr);
					  pu			   = pu * (Tm * sigma_n / pr);
					  const f3 Tr	   = T_ray / (pu + pl).mean();
					  if (Tr.maxCoeff() < 0.05f) {
						  if (rng.get1D() < 0.75f) T_ray = f3();
						  else T_ray = T_ray / 0.25f;
					  }
					  return T_ray.any();
				  });
			T_ray = T_ray * (T_maj / T_maj[channel]);
			pu	  = pu * (T_maj / T_maj[channel]);
			pl	  = pl * (T_maj / T_maj[channel]);
		}
		if (visible || T_ray.isZero()) break;
		const f3 p_o = offsetRayOrigin(intr.p, intr.n, pLight - intr.p);
		d			 = pLight - p_o;
		o			 = p_o;
		medium		 = intr.getMedium(d);
	}
	if (T_ray.any()) {
		const f3 c = Ld * T_ray / (spu * pu + spl * pl).mean();
		pr.L	   = pr.L + c;
	}
	pr.rng = rng.state;
}

// Non-media shadow test of one shadow ray: true when nothing occludes it
// (the caller then adds Ld / mean(pu + pl)).
template <bool COUNT, bool ALPHA = true>
KRR_DEV bool occluded(const TraceBvh &bvh, const f3 &o, const f3 &d, float tMax, const LaneStack &stack, uint32_t &nv, uint32_t &nt) {
	return traverse<true, COUNT, ALPHA>(bvh, o, d, tMax, stack, nv, nt).prim >= 0;
}

// traceShadow with media: ShadowTr raygen + traceTransmittance
// (device.cu:102-128, wavefront.h:80-139).
template <bool COUNT>
__global__ void k_shadow_tr(WfParams P, DevScene sc, TraceBvh bvh, ShadowQueueDev sq, PixelStateDev px, int *ovf, Counters *cnt) {
	extern __shared__ int ldsStack[];
	const LaneStack stack = laneStack(bvh, ldsStack, ovf);
	const int count = cnt->shadowCount;
	if (blockIdx.x == 0 && threadIdx.x == 0) cnt->shadowTotal += (unsigned long long) count;
	uint32_t nv = 0, nt = 0;
	while (true) {
		const int base = waveFetch(&cnt->work[1]);
		if (base >= count) break;
		const int i = base + laneId();
		if (i >= count) continue;
		const float4 a = sq.ray[2 * i], b = sq.ray[2 * i + 1];
		const float4 e0 = sq.ext[3 * i], e1 = sq.ext[3 * i + 1], e2 = sq.ext[3 * i + 2];
		const int pixel = __float_as_int(b.w);
		PathRegs pr		= loadPathRegs(px, pixel, true);
		transmittanceItem<COUNT>(P, sc, bvh, f3(a.x, a.y, a.z), f3(b.x, b.y, b.z), a.w, __float_as_int(sq.c[i].w) - 1, f3(e0.x, e0.y, e0.z),
								 f3(e1.x, e1.y, e1.z), f3(e2.x, e2.y, e2.z), pr, stack, nv, nt);
		storePathRegs(px, pixel, pr);
	}
	if (COUNT) {
		waveAddU64(&cnt->shadowNodeVisits, nv);
		waveAddU64(&cnt->shadowTriTests, nt);
	}
}

template <bool COUNT>
__global__ void k_trace_shadow(TraceBvh bvh, ShadowQueueDev sq, PixelStateDev px, int *ovf, Counters *cnt) {
	extern __shared__ int ldsStack[];
	const LaneStack stack = laneStack(bvh, ldsStack, ovf);
	const int count		  = cnt->shadowCount;
	if (blockIdx.x == 0 && threadIdx.x == 0) cnt->shadowTotal += (unsigned long long) count;
	uint32_t nv = 0, nt = 0;
	while (true) {
		const int base = waveFetch(&cnt->work[1]);
		if (base >= count) break;
		const int i = base + laneId();
		if (i >= count) continue;
		const float4 a = sq.ray[2 * i], b = sq.ray[2 * i + 1], c = sq.c[i];
		if (!occluded<COUNT>(bvh, f3(a.x, a.y, a.z), f3(b.x, b.y, b.z), a.w, stack, nv, nt)) {
			const int pixel = __float_as_int(b.w);
			float4 L		= px.L[pixel];
			L.x				= L.x + c.x;
			L.y				= L.y + c.y;
			L.z				= L.z + c.z;
			px.L[pixel]		= L;
		}
	}
	if (COUNT) {
		waveAddU64(&cnt->shadowNodeVisits, nv);
		waveAddU64(&cnt->shadowTriTests, nt);
	}
}

constexpr int kTileKeys = 64, kTileBlock = 1024; // sort keys (4 x mean path length), tiles per histogram block

// Camera paths in the order the persistent launch takes them up: 8 x 8-pixel
// tiles by the mean length of their pixels' paths in the previous sample,
// longest first, pixels row by row inside a tile.  Long paths then start
// early and the last paths a launch takes up are short (many leave the
// scene at once), so the frame's tail - waves kept alive by their longest
// path once no new path is left - is shorter, while the 64 camera rays of a
// wave stay coherent.  The images do not depend on the order.  A stable
// counting sort of the tiles on 64 keys (4 x the mean length), then the
// pixel positions by a scan of the tiles' pixel counts.
__global__ void __launch_bounds__(64) k_tile_keys(const uint8_t *plen, int W, int H, int tilesX, uint8_t *key, int *cnt) {
	const int t = blockIdx.x, x = (t % tilesX) * 8 + (threadIdx.x & 7), y = (t / tilesX) * 8 + (threadIdx.x >> 3);
	const bool v = x < W && y < H;
	int s = v ? plen[(size_t) y * W + x] : 0;
	for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
	const int c = __popcll(__ballot(v)); // >= 1: every tile holds a pixel
	if (threadIdx.x == 0) {
		key[t] = (uint8_t) min(kTileKeys - 1, (4 * s) / c);
		cnt[t] = c;
	}
}

// Per block of kTileBlock tiles, the count of each key, laid out key-major
// from the largest key down so that one exclusive scan gives every (key,
// block) its first slot.
__global__ void __launch_bounds__(kTileBlock) k_tile_hist(const uint8_t *key, int nTiles, int nBlk, int *hist) {
	__shared__ int h[kTileKeys];
	if (threadIdx.x < kTileKeys) h[threadIdx.x] = 0;
	__syncthreads();
	const int t = blockIdx.x * kTileBlock + threadIdx.x;
	if (t < nTiles) atomicAdd(&h[key[t]], 1);
	__syncthreads();
	if (threadIdx.x < kTileKeys) hist[(kTileKeys - 1 - threadIdx.x) * nBlk + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of n counts in place, the total at a[n] (one block of 1024
// threads).  The array is taken in chunks of 4096: each thread takes four
// consecutive counts, wave sums by shuffles, a carry between chunks.
__global__ void __launch_bounds__(1024) k_order_scan(int *a, int n) {
	__shared__ int wsum[16];
	__shared__ int carry;
	const int t = threadIdx.x, lane = t & 63, w = t >> 6;
	if (t == 0) carry = 0;
	__syncthreads();
	for (int base = 0; base < n; base += 4096) {
		const int i0 = base + 4 * t;
		int v[4];
		for (int k = 0; k < 4; k++) v[k] = i0 + k < n ? a[i0 + k] : 0;
		const int tsum = v[0] + v[1] + v[2] + v[3];
		int incl = tsum; // inclusive scan of the thread sums inside the wave
		for (int o = 1; o < 64; o <<= 1) {
			const int u = __shfl_up(incl, o);
			if (lane >= o) incl += u;
		}
		if (lane == 63) wsum[w] = incl;
		__syncthreads();
		int before = carry;
		for (int k = 0; k < w; k++) before += wsum[k];
		int run = before + incl - tsum;
		for (int k = 0; k < 4; k++) {
			if (i0 + k < n) a[i0 + k] = run;
			run += v[k];
		}
		__syncthreads();
		if (t == 1023) carry = run; // run of the last thread = carry + chunk total
		__syncthreads();
	}
	if (t == 0) a[n] = carry;
}

// Sorted slot of each tile (stable: equal keys keep tile order), and its
// pixel count at that slot.  A tile's rank among the block's tiles of its key:
// the lanes of its wave below it with the same key (six ballots, one per key
// bit), plus the tiles of that key in the lower waves (per-wave key counts in
// LDS).
static_assert(kTileKeys == 64, "six key bits");
__global__ void __launch_bounds__(kTileBlock) k_tile_rank(const uint8_t *key, const int *cnt, int nTiles, int nBlk, const int *hist, int *sorted,
														  int *sortedCnt) {
	constexpr int kWaves = kTileBlock / 64;
	__shared__ int wcnt[kWaves][kTileKeys];
	const int t = blockIdx.x * kTileBlock + threadIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	const bool valid = t < nTiles;
	const int my	 = valid ? key[t] : 0;
	uint64_t eq		 = __ballot(valid);
#pragma unroll
	for (int bit = 0; bit < 6; bit++) {
		const uint64_t m = __ballot(valid && ((my >> bit) & 1));
		eq &= ((my >> bit) & 1) ? m : ~m;
	}
	const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	int r				 = __popcll(eq & below);
	for (int i = (int) threadIdx.x; i < kWaves * kTileKeys; i += (int) blockDim.x) (&wcnt[0][0])[i] = 0;
	__syncthreads();
	if (valid && r == 0) wcnt[w][my] = __popcll(eq); // the key's first lane in the wave
	__syncthreads();
	if (!valid) return;
	for (int v = 0; v < w; v++) r += wcnt[v][my];
	const int slot	= hist[(kTileKeys - 1 - my) * nBlk + blockIdx.x] + r;
	sorted[slot]	= t;
	sortedCnt[slot] = cnt[t];
}

__global__ void __launch_bounds__(64) k_tile_pixels(const int *sorted, const int *offset, int W, int H, int tilesX, int *order) {
	const int s = blockIdx.x, t = sorted[s];
	const int x = (t % tilesX) * 8 + (threadIdx.x & 7), y = (t / tilesX) * 8 + (threadIdx.x >> 3);
	const bool v	  = x < W && y < H;
	const uint64_t b  = __ballot(v);
	const int lane	  = threadIdx.x;
	const uint64_t lo = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	if (v) order[offset[s] + __popcll(b & lo)] = y * W + x;
}

// The same for many views: tiles of each view numbered after those of the
// previous views (ViewsDev::tileStart), pixels by global number.
KRR_DEV int findTileView(const ViewsDev &V, int t) {
	int lo = 0, hi = V.n - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (V.tileStart[mid] <= t) lo = mid;
		else hi = mid - 1;
	}
	return lo;
}

__global__ void __launch_bounds__(64) k_tile_keys_views(const uint8_t *plen, ViewsDev V, uint8_t *key, int *cnt) {
	const int t = blockIdx.x, v = findTileView(V, t), lt = t - V.tileStart[v];
	const int W = V.view[v].W, H = V.view[v].H, tilesX = (W + 7) / 8;
	const int x = (lt % tilesX) * 8 + (threadIdx.x & 7), y = (lt / tilesX) * 8 + (threadIdx.x >> 3);
	const bool ok = x < W && y < H;
	int s		  = ok ? plen[V.start[v] + y * W + x] : 0;
	for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
	const int c = __popcll(__ballot(ok));
	if (threadIdx.x == 0) {
		key[t] = (uint8_t) min(kTileKeys - 1, (4 * s) / c);
		cnt[t] = c;
	}
}

__global__ void __launch_bounds__(64) k_tile_pixels_views(const int *sorted, const int *offset, ViewsDev V, int *order) {
	const int s = blockIdx.x, t = sorted[s], v = findTileView(V, t), lt = t - V.tileStart[v];
	const int W = V.view[v].W, H = V.view[v].H, tilesX = (W + 7) / 8;
	const int x = (lt % tilesX) * 8 + (threadIdx.x & 7), y = (lt / tilesX) * 8 + (threadIdx.x >> 3);
	const bool ok	  = x < W && y < H;
	const uint64_t b  = __ballot(ok);
	const int lane	  = threadIdx.x;
	const uint64_t lo = lane == 0 ? 0ull : (~0ull >> (64 - lane));
	if (ok) order[offset[s] + __popcll(b & lo)] = V.start[v] + y * W + x;
}

__global__ void k_order_identity(int N, int *order) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < N) order[i] = i;
}

__global__ void k_mega_begin(Counters *cnt) {
	cnt->work[2]	= 0;
	cnt->nrcQueries = 0;
	cnt->closestLaunches += 1;
}

// Tail of the sample: every path left in the queue after the wavefront depths
// runs to completion in one thread - closest trace, shade, shadow trace,
// next depth - in the same per-path stage order as the wavefront, so the
// radiance additions and RNG draws of each pixel are unchanged.  Short queues
// at deep bounces would otherwise cost one latency-bound launch per depth,
// each as long as its slowest ray.
//
// Each loop iteration advances every lane's path by one depth; a lane whose
// path has ended takes the next queued path at once (one atomic per wave for
// all such lanes), so no lane idles while the queue still holds work.
//
// VIEWS: the paths are the camera samples of all views (global pixels
// 0..P.N-1), generated at fetch instead of read from the ray queue; the
// frame's PCG increment and training phase come from V.frame.
template <bool MEDIA, bool COUNT, bool NRC = false, bool VIEWS = false, bool PLAIN = false>
__global__ void k_megakernel(WfParams P0, DevScene sc, TraceBvh bvh, RayQueueDev q, const int *countPtr, int startDepth, PixelStateDev px,
							 int *ovf, Counters *cnt, NrcDev nrc = NrcDev{}, ViewsDev V = ViewsDev{}, CameraDev cam = CameraDev{},
							 int camFetch = 0) {
	extern __shared__ int ldsStack[];
	const LaneStack stack = laneStack(bvh, ldsStack, ovf);
	WfParams P			  = P0;
	if (VIEWS) {
		P.inc = V.frame->inc;
		if (NRC) nrc.phase = V.frame->phase;
	}
	if (NRC) nrc.keepProb = *nrc.keep;
	// camFetch: the camera rays of this sample are generated here, at fetch
	// (generateCameraRays, integrator.cpp:166-179), instead of read from q
	const int count = (VIEWS || camFetch) ? P.N : *countPtr;
	uint32_t nv = 0, nt = 0, snv = 0, snt = 0, nClosest = 0, nShadow = 0;
	bool active = false, exhausted = false, trainPath = false;
	int depth	= startDepth;
	RayItem r;
	PathRegs pr; // the path's pixel state, in registers until the path ends
	while (true) {
		if (!exhausted) {
			const int i = wavePush(&cnt->work[2], !active);
			if (__ballot(!active && i >= count)) exhausted = true;
			if (!active && i < count) {
				if (VIEWS) r = viewCameraRay(P, V, px, px.order ? px.order[i] : i, pr, MEDIA);
				else if (camFetch) {
					const int pix = px.order ? px.order[i] : i;
					pr			  = loadPathRegs(px, pix, MEDIA);
					if (!P.legacySpp) pr.L = f3();
					Pcg cs{pr.rng, P.inc};
					r		= cameraRay(cam, P.W, P.H, pix % P.W, pix / P.W, cs);
					r.pixel = pix;
					pr.rng	= cs.state;
				} else {
					r  = loadRayItem(q, (px.order && startDepth == 0) ? px.order[i] : i);
					pr = loadPathRegs(px, r.pixel, MEDIA);
				}
				depth	  = startDepth;
				active	  = true;
				trainPath = NRC && trainHash((uint32_t) r.pixel) % (uint32_t) nrc.trainMod == (uint32_t) nrc.phase;
			}
		}
		if (__ballot(active) == 0) break;
		if (!active) continue;
		nClosest++;
		const TraceResult h = traverse<false, COUNT, !PLAIN>(bvh, r.o, r.d, __builtin_inff(), stack, nv, nt);
		const float4 hv		= make_float4(h.prim >= 0 ? h.t : __builtin_inff(), __int_as_float(h.prim), h.b1, h.b2);
		bool wantNext = false, wantShadow = false;
		RayItem nx;
		ShadowOut so;
		so.medium = -1;
		int trainSlot = -1;
		shadeCore<MEDIA, NRC, PLAIN>(P, sc, r, hv, pr, depth, wantNext, nx, wantShadow, so, &nrc, trainPath, &trainSlot);
		if (wantShadow) {
			nShadow++;
			if (MEDIA && P.enableMedium)
				transmittanceItem<COUNT, !PLAIN>(P, sc, bvh, so.o, so.d, 1.f, so.medium, so.Ld, so.pu, so.pl, pr, stack, snv, snt);
			else if (!occluded<COUNT, !PLAIN>(bvh, so.o, so.d, 1.f, stack, snv, snt))
				pr.L = pr.L + so.Ld / (so.pl + so.pu).mean();
		}
		if (NRC && trainSlot >= 0) // radiance gathered up to and including this vertex's NEE
			nrc.train[(size_t) trainSlot * 6 + 5] = make_float4(pr.L.x, pr.L.y, pr.L.z, 0.f);
		if (wantNext && depth < P.maxDepth) {
			r = nx;
			depth++;
		} else {
			storePathRegs(px, r.pixel, pr);
			if (px.plen) px.plen[r.pixel] = (uint8_t) min(depth, 255);
			active = false;
		}
	}
	waveAddU64(&cnt->closestTotal, nClosest);
	waveAddU64(&cnt->shadowTotal, nShadow);
	waveAddU64(&cnt->megaClosest, nClosest);
	waveAddU64(&cnt->megaShadow, nShadow);
	if (COUNT) {
		waveAddU64(&cnt->megaNodeVisits, nv + snv);
		waveAddU64(&cnt->megaTriTests, nt + snt);
	}
}

// The persistent per-path launch of k_megakernel (scenes without media) with
// dynamic ray refill: each lane has one ray in flight, its path's closest ray
// or its shadow ray, and all lanes of a wave traverse together until fewer
// than P.refillLanes of them are still traversing.  The lanes whose ray is
// done then shade (closest) or add the light sample (shadow) and set up their
// path's next ray, or take a new path; the others keep their traversal state
// (next entry, stack, best hit) and go on in the next round.  A traversal
// loop otherwise runs as long as its slowest lane, with the other lanes of the
// wave idle.  Each path still runs closest trace, shading, shadow trace and
// next depth in this order, and each ray takes the same traversal steps, so
// the radiance additions, RNG draws and hits are those of k_megakernel.  Once
// no path is left to take, the rounds run every traversal to its end.
// Waves per SIMD the refill launch is compiled for: 3 (168 VGPRs, no
// scratch) for the single-view launch, where it runs as fast as 4 waves with
// 120 bytes per lane of spills (3.80 against 3.79-3.83 ms per frame on the
// 1.31 M-triangle 1080p field); 4 for the multi-view launch, which 3 waves
// slow by 7 % (259 against 243 ms for 256 views).
constexpr int refillWaves(bool views) { return views ? 4 : 3; }

template <bool COUNT, bool NRC = false, bool VIEWS = false, bool PLAIN = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(refillWaves(VIEWS), refillWaves(VIEWS))))
k_megakernel_refill(WfParams P0, DevScene sc, TraceBvh bvh, RayQueueDev q, const int *countPtr, int startDepth,
									PixelStateDev px, int *ovf, Counters *cnt, NrcDev nrc = NrcDev{}, ViewsDev V = ViewsDev{},
									CameraDev cam = CameraDev{}, int camFetch = 0) {
	extern __shared__ int ldsStack[];
	const LaneStack stack = laneStack(bvh, ldsStack, ovf);
	WfParams P			  = P0;
	// the top P.ldsNodes nodes of the tree, copied into LDS behind the stack
	float4 *nodeLds = reinterpret_cast<float4 *>(ldsStack + (size_t) (bvh.stackLds + 1) * blockDim.x);
	if (P.ldsNodes > 0) {
		const float4 *src = reinterpret_cast<const float4 *>(bvh.nodes);
		for (int i = (int) threadIdx.x; i < P.ldsNodes * 8; i += (int) blockDim.x) nodeLds[i] = src[i];
		__syncthreads();
	}
	if (VIEWS) {
		P.inc = V.frame->inc;
		if (NRC) nrc.phase = V.frame->phase;
	}
	if (NRC) nrc.keepProb = *nrc.keep;
	const int count = (VIEWS || camFetch) ? P.N : *countPtr;
	uint32_t nv = 0, nt = 0, nClosest = 0, nShadow = 0;
	bool active = false, exhausted = false, trainPath = false;
	bool going = false, shadowRay = false, pendingNext = false;
	int depth = startDepth, trainSlot = -1;
	RayItem r; // the path's current vertex ray; while its shadow ray is in flight, the next one
	PathRegs pr;
	f3 to, td;	 // the ray in flight
	f3 contrib;	 // the shadow ray's radiance when unoccluded
	TravState ts;
	while (true) {
		if (!exhausted) {
			const int i = wavePush(&cnt->work[2], !active);
			if (__ballot(!active && i >= count)) exhausted = true;
			if (!active && i < count) {
				if (VIEWS) r = viewCameraRay(P, V, px, px.order ? px.order[i] : i, pr, false);
				else if (camFetch) {
					const int pix = px.order ? px.order[i] : i;
					pr			  = loadPathRegs(px, pix, false);
					if (!P.legacySpp) pr.L = f3();
					Pcg cs{pr.rng, P.inc};
					r		= cameraRay(cam, P.W, P.H, pix % P.W, pix / P.W, cs);
					r.pixel = pix;
					pr.rng	= cs.state;
				} else {
					r  = loadRayItem(q, (px.order && startDepth == 0) ? px.order[i] : i);
					pr = loadPathRegs(px, r.pixel, false);
				}
				depth	  = startDepth;
				active	  = true;
				trainPath = NRC && trainHash((uint32_t) r.pixel) % (uint32_t) nrc.trainMod == (uint32_t) nrc.phase;
				to		  = r.o;
				td		  = r.d;
				shadowRay = false;
				going	  = true;
				travInit(ts, bvh, __builtin_inff());
				nClosest++;
			}
		}
		if (__ballot(active) == 0) break;
		traverseSteps<COUNT, !PLAIN>(bvh, to, td, shadowRay, going, ts, stack, exhausted ? 1 : P.refillLanes, nv, nt, nodeLds, P.ldsNodes,
									 P.leafShare);
		if (!active || going) continue;
		bool next = false; // go on with the path's next closest ray
		if (shadowRay) {
			if (ts.prim < 0) pr.L = pr.L + contrib;
			if (NRC && trainSlot >= 0) nrc.train[(size_t) trainSlot * 6 + 5] = make_float4(pr.L.x, pr.L.y, pr.L.z, 0.f);
			next = pendingNext;
		} else {
			const float4 hv = make_float4(ts.prim >= 0 ? ts.t : __builtin_inff(), __int_as_float(ts.prim), ts.b1, ts.b2);
			bool wantNext = false, wantShadow = false;
			RayItem nx;
			ShadowOut so;
			so.medium = -1;
			trainSlot = -1;
			shadeCore<false, NRC, PLAIN>(P, sc, r, hv, pr, depth, wantNext, nx, wantShadow, so, &nrc, trainPath, &trainSlot);
			const bool more = wantNext && depth < P.maxDepth;
			if (more) r = nx;
			if (wantShadow) {
				nShadow++;
				contrib		= so.Ld / (so.pl + so.pu).mean();
				to			= so.o;
				td			= so.d;
				shadowRay	= true;
				pendingNext = more;
				going		= true;
				travInit(ts, bvh, 1.f);
				continue;
			}
			if (NRC && trainSlot >= 0) nrc.train[(size_t) trainSlot * 6 + 5] = make_float4(pr.L.x, pr.L.y, pr.L.z, 0.f);
			next = more;
		}
		if (next) {
			depth++;
			to		  = r.o;
			td		  = r.d;
			shadowRay = false;
			going	  = true;
			travInit(ts, bvh, __builtin_inff());
			nClosest++;
		} else {
			storePathRegs(px, r.pixel, pr);
			if (px.plen) px.plen[r.pixel] = (uint8_t) min(depth, 255);
			active = false;
		}
	}
	waveAddU64(&cnt->closestTotal, nClosest);
	waveAddU64(&cnt->shadowTotal, nShadow);
	waveAddU64(&cnt->megaClosest, nClosest);
	waveAddU64(&cnt->megaShadow, nShadow);
	if (COUNT) {
		waveAddU64(&cnt->megaNodeVisits, nv);
		waveAddU64(&cnt->megaTriTests, nt);
	}
}

// Training targets of the vertices recorded in this sample: the radiance the
// path gathered after the vertex, divided by the throughput at the vertex.
__global__ void k_nrc_targets(NrcDev nrc, const int *donePtr, const float4 *L, float4 *batchV, float *batchT) {
	const int count = min(*nrc.trainCount, nrc.capTrain), done = *donePtr;
	for (int e = done + (int) (blockIdx.x * blockDim.x + threadIdx.x); e < count; e += (int) (gridDim.x * blockDim.x)) {
		const float4 *t	 = nrc.train + (size_t) e * 6;
		const float4 pay = t[4], li = t[5];
		const float4 Lf	 = L[__float_as_int(pay.w)];
		batchV[4 * e]		= t[0];
		batchV[4 * e + 1]	= t[1];
		batchV[4 * e + 2]	= t[2];
		batchV[4 * e + 3]	= t[3];
		batchT[3 * e]		= pay.x > 0.f ? fmaxf(0.f, (Lf.x - li.x) / pay.x) : 0.f;
		batchT[3 * e + 1]	= pay.y > 0.f ? fmaxf(0.f, (Lf.y - li.y) / pay.y) : 0.f;
		batchT[3 * e + 2]	= pay.z > 0.f ? fmaxf(0.f, (Lf.z - li.z) / pay.z) : 0.f;
	}
}

// Keep probability of this frame's training candidates from the previous
// frame's count, so that about capTrain vertices are kept, spread uniformly.
__global__ void k_nrc_begin(Counters *cnt, int capTrain) {
	const int prev		   = cnt->nrcCandidates;
	cnt->nrcPrevCandidates = prev;
	cnt->nrcKeep		   = prev > 0 ? fminf(1.f, 0.95f * (float) capTrain / (float) prev) : 1.f;
	cnt->nrcCandidates	   = 0;
}

__global__ void k_nrc_mark_done(Counters *cnt, int capTrain, int capQuery) {
	cnt->nrcTrainDone = min(cnt->nrcTrain, capTrain);
	cnt->nrcQueriesTotal += (unsigned long long) min(cnt->nrcQueries, capQuery);
}

// pixel += L (integrator.cpp:257-260).
__global__ void k_accumulate(int N, PixelStateDev px) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= N) return;
	const float4 L = px.L[i];
	float4 p	   = px.pixel[i];
	p.x			   = p.x + L.x;
	p.y			   = p.y + L.y;
	p.z			   = p.z + L.z;
	px.pixel[i]	   = p;
}

// Film write (integrator.cpp:262-266, CudaRenderTarget::write,
// src/core/device/cuda.h:33-45): RGBA(pixel / spp, 1) at row H-1-y.  The
// frame's last sample is added to the pixel sum here (k_accumulate's
// pixel += L, the same sum), saving a pass over the pixels.
__global__ void k_film(WfParams P, PixelStateDev px, float4 *film) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= P.N) return;
	const int x = i % P.W, y = i / P.W;
	float4 p	   = px.pixel[i];
	const float4 L = px.L[i];
	p.x			   = p.x + L.x;
	p.y			   = p.y + L.y;
	p.z			   = p.z + L.z;
	px.pixel[i]	   = p;
	float r		   = p.x / (float) P.spp, g = p.y / (float) P.spp, b = p.z / (float) P.spp;
	if (P.enableClamp) {
		r = clampf(r, 0.f, P.clampMax);
		g = clampf(g, 0.f, P.clampMax);
		b = clampf(b, 0.f, P.clampMax);
	}
	film[(size_t) (P.H - 1 - y) * P.W + x] = make_float4(r, g, b, 1.f);
}

// Error metrics (metrics.cu:63-128): per-pixel value clamped at 100, FP64
// sum per block, then one atomic per block.
__global__ void k_metric(int metric, const float4 *img, const float4 *ref, int64_t n, double *out) {
	__shared__ double part[256];
	double acc = 0.0;
	for (int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
		const float4 y = img[i], r = ref[i];
		float e		   = 0.f;
		const bool bad = __builtin_isinf(r.x) || __builtin_isinf(r.y) || __builtin_isinf(r.z) || r.x != r.x || r.y != r.y || r.z != r.z;
		if (!bad) {
			float yy[3] = {y.x, y.y, y.z}, rr[3] = {r.x, r.y, r.z}, a[3];
			for (int c = 0; c < 3; c++) {
				const float dd = fabsf(yy[c] - rr[c]);
				switch (metric) {
					case KRR_METRIC_MSE: a[c] = dd * dd; break;
					case KRR_METRIC_MAPE: a[c] = dd / (rr[c] + 0.f); break;
					case KRR_METRIC_SMAPE: a[c] = dd / (rr[c] + yy[c] + 0.f); break;
					default: a[c] = rr[c] == 0.f ? 0.f : (dd / rr[c]) * (dd / rr[c]); break;
				}
			}
			e = (a[0] + a[1] + a[2]) / 3.f;
		}
		acc += (double) fminf(e, 100.f);
	}
	part[threadIdx.x] = acc;
	__syncthreads();
	for (int s = blockDim.x / 2; s > 0; s >>= 1) {
		if ((int) threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
		__syncthreads();
	}
	if (threadIdx.x == 0) atomicAdd(out, part[0]);
}

__global__ void k_zero_double(double *p) { *p = 0.0; }

} // namespace krr

// ===========================================================================
// Host side
// ===========================================================================

using namespace krr;

#define KRR_HIP(call)                                                                                               \
	do {                                                                                                            \
		hipError_t e_ = (call);                                                                                     \
		if (e_ != hipSuccess) return krr::fail(e_ == hipErrorOutOfMemory ? KRR_ERR_OUT_OF_MEMORY : KRR_ERR_HIP,      \
											   "%s failed: %s", #call, hipGetErrorString(e_));                      \
	} while (0)

namespace {

bool deviceAvailable() {
	int n			 = 0;
	hipError_t e	 = hipGetDeviceCount(&n);
	return e == hipSuccess && n > 0;
}

// Device allocation owned by the library.
struct DevBuf {
	void *p	  = nullptr;
	size_t sz = 0;
	DevBuf() = default;
	DevBuf(const DevBuf &) = delete;
	DevBuf &operator=(const DevBuf &) = delete;
	~DevBuf() { release(); }
	void release() {
		if (p) (void) hipFree(p);
		p  = nullptr;
		sz = 0;
	}
	hipError_t alloc(size_t bytes) {
		release();
		if (bytes == 0) bytes = 16;
		hipError_t e = hipMalloc(&p, bytes);
		if (e == hipSuccess) sz = bytes;
		else p = nullptr;
		return e;
	}
	hipError_t upload(const void *src, size_t bytes) {
		hipError_t e = alloc(bytes);
		if (e != hipSuccess) return e;
		return src && bytes ? hipMemcpy(p, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
	}
	template <class T> T *as() const { return reinterpret_cast<T *>(p); }
	// takes ownership of a hipMalloc'd block
	void adopt(void *q, size_t bytes) {
		release();
		p  = q;
		sz = bytes;
	}
};

struct AccelData {
	DevBuf nodes, tris;
	DevBuf qnodes;			 // the same tree as Bvh4QNode (64 B per node)
	bool quantized = false; // traversal reads qnodes ("node_format": "quantized") or nodes ("float", the default)
	DevBuf alphaTri, alphaSrc, alphaImg; // alpha test data; empty when no material has one
	int stackSize = 0;
	int stackLds  = kShortStack; // stack entries per lane kept in LDS (the rest spill to global memory)
	Bvh4Host host; // kept for statistics (nodes/tris arrays cleared after upload)
	int64_t nNodes = 0, nLeafTris = 0;
	int maxDepth	= 0;
	double buildSeconds = 0;
	bool onDevice		= false; // built by the GPU builder
	int ldsEntries() const { return std::min(stackSize, stackLds); }
	TraceBvh dev() const {
		return TraceBvh{nodes.as<Bvh4Node>(), tris.as<LeafTri>(), ldsEntries(), (int) std::min<int64_t>(INT32_MAX, nNodes + nLeafTris),
						alphaTri.as<AlphaTri>(), alphaSrc.as<AlphaSrc>(), quantized ? qnodes.as<Bvh4QNode>() : nullptr};
	}
};

// A material has the any-hit alpha test when it has a transmission texture
// or a non-zero transmission constant (rt::TextureData::isValid of the
// Transmission slot, shading.h:92-95; pbrt.cpp:145-156 sets it only then).
bool materialAlphaTested(const krr_material &m) {
	return m.transmission_texture > 0 || m.transmission[0] != 0.f || m.transmission[1] != 0.f || m.transmission[2] != 0.f;
}

// Per-triangle texture coordinates and material, per-material transmission
// source (images copied into one buffer), for the triangles flagged
// kLeafAlpha.  Releases the buffers when no triangle is flagged.
int uploadAlpha(const krr_scene_desc *s, const std::vector<uint8_t> &flags, AccelData &a) {
	a.alphaTri.release();
	a.alphaSrc.release();
	a.alphaImg.release();
	bool any = false;
	for (uint8_t f : flags) any |= (f & kLeafAlpha) != 0;
	if (!any) return KRR_OK;
	std::vector<AlphaTri> tri((size_t) s->n_triangles);
	for (int m = 0; m < s->n_meshes; m++) {
		const krr_mesh &mm = s->meshes[m];
		for (int t = 0; t < mm.n_triangles; t++) {
			const int64_t p = (int64_t) mm.first_triangle + t;
			AlphaTri &at	= tri[(size_t) p];
			at.material		= std::max(mm.material, 0);
			if (mm.has_texcoords && s->texcoords) { // shading.h:100-102: zero without texture coordinates
				const int32_t *v = s->indices + 3 * p;
				at.u0 = s->texcoords[2 * v[0]], at.v0 = s->texcoords[2 * v[0] + 1];
				at.u1 = s->texcoords[2 * v[1]], at.v1 = s->texcoords[2 * v[1] + 1];
				at.u2 = s->texcoords[2 * v[2]], at.v2 = s->texcoords[2 * v[2] + 1];
			}
		}
	}
	std::vector<AlphaSrc> src((size_t) std::max(s->n_materials, 1));
	std::vector<size_t> off((size_t) std::max(s->n_materials, 1), 0);
	size_t texels = 0;
	for (int i = 0; i < s->n_materials; i++) {
		const krr_material &m = s->materials[i];
		for (int c = 0; c < 3; c++) src[i].T[c] = m.transmission[c];
		if (m.transmission_texture > 0) {
			const krr_texture &t = s->textures[m.transmission_texture - 1];
			src[i].w = t.width;
			src[i].h = t.height;
			off[i]	 = texels;
			texels	+= (size_t) t.width * t.height;
		}
	}
	KRR_HIP(a.alphaImg.alloc(texels * sizeof(float4)));
	for (int i = 0; i < s->n_materials; i++) {
		const krr_material &m = s->materials[i];
		if (m.transmission_texture <= 0) continue;
		const krr_texture &t = s->textures[m.transmission_texture - 1];
		src[i].img			 = a.alphaImg.as<float4>() + off[i];
		KRR_HIP(hipMemcpy((void *) src[i].img, t.rgba, (size_t) t.width * t.height * sizeof(float4), hipMemcpyHostToDevice));
	}
	KRR_HIP(a.alphaTri.upload(tri.data(), tri.size() * sizeof(AlphaTri)));
	KRR_HIP(a.alphaSrc.upload(src.data(), src.size() * sizeof(AlphaSrc)));
	return KRR_OK;
}

} // namespace

// Quantised copy of a float node (Bvh4QNode): per axis, the grid spacing
// 2^e is the smallest that spans the children's extent in 254 steps and keeps
// every plane org + q 2^e (q < 256, org a multiple of 2^e) exact in binary32;
// planes are rounded outward and then checked, so each decoded child box
// contains its float box.  The arithmetic is on doubles, where these values
// and the differences taken are exact.
__device__ void quantizeNode(const Bvh4Node &nd, Bvh4QNode &q) {
	const float *lo[3] = {nd.lox, nd.loy, nd.loz}, *hi[3] = {nd.hix, nd.hiy, nd.hiz};
	uint32_t valid = 0;
	for (int c = 0; c < 4; c++)
		if (nd.lox[c] <= nd.hix[c] && nd.loy[c] <= nd.hiy[c] && nd.loz[c] <= nd.hiz[c]) valid |= 1u << c;
	for (int a = 0; a < 3; a++) {
		double L = 0.0, H = 0.0;
		bool any = false;
		for (int c = 0; c < 4; c++)
			if ((valid >> c) & 1u) {
				L	= any ? fmin(L, (double) lo[a][c]) : (double) lo[a][c];
				H	= any ? fmax(H, (double) hi[a][c]) : (double) hi[a][c];
				any = true;
			}
		const double m = fmax(fabs(L), fabs(H));
		int e		   = -126;
		if (m > 0.0) e = max(e, ilogb(m) - 22);
		if (H > L) e = max(e, (int) ceil(log2((H - L) / 254.0)));
		double sc = ldexp(1.0, e), org = floor(L / sc) * sc;
		while ((H - org) / sc > 255.0 || (double) (float) org != org) { // spacing too fine for the extent or the origin
			e++;
			sc	= ldexp(1.0, e);
			org = floor(L / sc) * sc;
		}
		q.org[a] = (float) org;
		q.e[a]	 = (int8_t) e;
		uint32_t wl = 0, wh = 0;
		for (int c = 0; c < 4; c++) {
			int ql = 255, qh = 0; // unused slot: an inverted box, also masked by `valid`
			if ((valid >> c) & 1u) {
				ql = (int) fmin(255.0, fmax(0.0, floor(((double) lo[a][c] - org) / sc)));
				qh = (int) fmin(255.0, fmax(0.0, ceil(((double) hi[a][c] - org) / sc)));
				while (ql > 0 && org + ql * sc > (double) lo[a][c]) ql--;
				while (qh < 255 && org + qh * sc < (double) hi[a][c]) qh++;
			}
			wl |= (uint32_t) ql << (8 * c);
			wh |= (uint32_t) qh << (8 * c);
		}
		q.q[2 * a]	   = wl;
		q.q[2 * a + 1] = wh;
	}
	q.valid = (uint8_t) valid;
	for (int c = 0; c < 4; c++) q.child[c] = nd.child[c];
	q.pad[0] = q.pad[1] = 0;
}

__global__ void k_quantize_nodes(const Bvh4Node *in, int64_t n, Bvh4QNode *out) {
	const int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x;
	if (i < n) quantizeNode(in[i], out[i]);
}

namespace {

int quantizeAccel(AccelData &a) {
	KRR_HIP(a.qnodes.alloc((size_t) std::max<int64_t>(a.nNodes, 1) * sizeof(Bvh4QNode)));
	if (a.nNodes > 0) {
		hipLaunchKernelGGL(k_quantize_nodes, dim3((unsigned) ((a.nNodes + 255) / 256)), dim3(256), 0, 0, a.nodes.as<Bvh4Node>(), a.nNodes,
						   a.qnodes.as<Bvh4QNode>());
		KRR_HIP(hipGetLastError());
		KRR_HIP(hipDeviceSynchronize());
	}
	return KRR_OK;
}

int buildAccel(const krr_scene_desc *s, AccelData &a, int leafSize, float nodeCost, bool onDevice) {
	if (!s || s->n_triangles < 0 || (s->n_triangles > 0 && (!s->positions || !s->indices)))
		return krr::fail(KRR_ERR_INVALID_ARG, "scene: missing triangle arrays");
	if (s->n_triangles > (int64_t) 0x0fffffff) return krr::fail(KRR_ERR_CAPACITY, "scene: at most 2^28 triangles");
	for (int64_t i = 0; i < 3 * s->n_triangles; i++)
		if (s->indices[i] < 0 || s->indices[i] >= s->n_vertices) return krr::fail(KRR_ERR_INVALID_ARG, "scene: vertex index %lld out of range", (long long) i);
	std::vector<uint8_t> nullFlag((size_t) s->n_triangles, 0);
	for (int m = 0; m < s->n_meshes; m++) {
		const krr_mesh &mm = s->meshes[m];
		if (mm.first_triangle < 0 || mm.n_triangles < 0 || (int64_t) mm.first_triangle + mm.n_triangles > s->n_triangles)
			return krr::fail(KRR_ERR_INVALID_ARG, "scene: mesh %d triangle range out of bounds", m);
		if (mm.material < 0)
			for (int t = 0; t < mm.n_triangles; t++) nullFlag[mm.first_triangle + t] = kLeafNull;
		else if (mm.material < s->n_materials && s->materials && materialAlphaTested(s->materials[mm.material]))
			for (int t = 0; t < mm.n_triangles; t++) nullFlag[mm.first_triangle + t] = kLeafAlpha;
	}
	for (int i = 0; i < s->n_materials && s->materials; i++) {
		const int tt = s->materials[i].transmission_texture;
		const bool ok = tt == 0 || (tt > 0 && tt <= s->n_textures && s->textures && s->textures[tt - 1].rgba && s->textures[tt - 1].width > 0 &&
									s->textures[tt - 1].height > 0);
		if (!ok) return krr::fail(KRR_ERR_INVALID_ARG, "scene: material %d transmission_texture %d is not a texture", i, tt);
	}
	if (int rc = uploadAlpha(s, nullFlag, a)) return rc;
	BvhBuildOptions opt;
	opt.maxLeafSize = leafSize;
	opt.nodeCost	= nodeCost;
	a.onDevice		= onDevice && s->n_triangles > 0;
	if (a.onDevice) {
		DeviceBvh d;
		std::string err;
		if (buildBvh4Device(s->positions, s->n_vertices, s->indices, s->n_triangles, nullFlag.data(), opt, d, err))
			return krr::fail(KRR_ERR_HIP, "device BVH build: %s", err.c_str());
		if ((size_t) d.nNodes * sizeof(Bvh4Node) >= ((size_t) 1 << 32)) {
			(void) hipFree(d.nodes);
			(void) hipFree(d.tris);
			return krr::fail(KRR_ERR_CAPACITY, "BVH of %lld nodes exceeds the 4 GiB node array limit", (long long) d.nNodes);
		}
		a.nodes.adopt(d.nodes, (size_t) d.nNodes * sizeof(Bvh4Node));
		a.tris.adopt(d.tris, (size_t) d.nTris * sizeof(LeafTri));
		a.maxDepth	   = d.maxDepth;
		a.stackSize	   = 3 * d.maxDepth + 4; // bvhStackSize
		a.nNodes	   = d.nNodes;
		a.nLeafTris	   = d.nTris;
		a.buildSeconds = d.buildSeconds;
		if (a.stackSize > 96) return krr::fail(KRR_ERR_CAPACITY, "BVH depth %d needs a %d-entry stack (max 96)", a.maxDepth, a.stackSize);
		return quantizeAccel(a);
	}
	buildBvh4(s->positions, s->indices, s->n_triangles, nullFlag.data(), opt, a.host);
	// traversal addresses nodes with 32-bit byte offsets
	if (a.host.nodes.size() * sizeof(Bvh4Node) >= ((size_t) 1 << 32))
		return krr::fail(KRR_ERR_CAPACITY, "BVH of %zu nodes exceeds the 4 GiB node array limit", a.host.nodes.size());
	KRR_HIP(a.nodes.upload(a.host.nodes.data(), a.host.nodes.size() * sizeof(Bvh4Node)));
	KRR_HIP(a.tris.upload(a.host.tris.data(), a.host.tris.size() * sizeof(LeafTri)));
	a.stackSize	   = bvhStackSize(a.host);
	a.nNodes	   = (int64_t) a.host.nodes.size();
	a.nLeafTris	   = (int64_t) a.host.tris.size();
	a.maxDepth	   = a.host.maxDepth;
	a.buildSeconds = a.host.buildSeconds;
	a.host.nodes.clear();
	a.host.nodes.shrink_to_fit();
	a.host.tris.clear();
	a.host.tris.shrink_to_fit();
	if (a.stackSize > 96) return krr::fail(KRR_ERR_CAPACITY, "BVH depth %d needs a %d-entry stack (max 96)", a.maxDepth, a.stackSize);
	return quantizeAccel(a);
}

// Blocks that fit on the GPU at once for this kernel and launch shape
// (persistent kernels launch exactly that many), capped at maxBlocks.
template <class K> int persistentGrid(K kernel, int block, size_t lds, int maxBlocks) {
	int dev = 0, cus = 0, nb = 0;
	if (hipGetDevice(&dev) != hipSuccess) return maxBlocks;
	if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) return maxBlocks;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, lds) != hipSuccess || nb < 1) nb = 1;
	return std::max(1, std::min(maxBlocks, nb * cus));
}

// LDS bytes per block of the short traversal stack, and ints per grid thread
// of its global overflow.
// nLds entries + one slot that absorbs the pushes bound for the overflow area
size_t stackLds(const AccelData &a, int block) { return (size_t) (a.ldsEntries() + 1) * block * sizeof(int); }
size_t stackOverflowInts(const AccelData &a) { return (size_t) (a.stackSize - a.ldsEntries()); }

// A fraction of a persistent grid (at least one block).
int scaledGrid(int g, float frac) { return std::max(1, (int) std::lround((double) g * (double) frac)); }

// Grows buf to hold the traversal-stack overflow of `threads` grid threads.
int reserveOverflow(DevBuf &buf, const AccelData &a, size_t threads) {
	const size_t bytes = stackOverflowInts(a) * threads * sizeof(int);
	if (bytes > buf.sz) KRR_HIP(buf.alloc(bytes));
	return KRR_OK;
}

// The persistent kernel variants of every frame kind.
// REFILL: k_megakernel_refill (scenes without media).
// nodeBytes: the refill launch's LDS copy of the top tree nodes, behind the stack.
template <bool MEDIA, bool COUNT, bool NRC, bool VIEWS, bool PLAIN, bool REFILL = false> struct Mega {
	static int grid(int tb, size_t lds, int maxBlocks, size_t nodeBytes = 0) {
		if constexpr (REFILL) return persistentGrid(k_megakernel_refill<COUNT, NRC, VIEWS, PLAIN>, tb, lds + nodeBytes, maxBlocks);
		else return persistentGrid(k_megakernel<MEDIA, COUNT, NRC, VIEWS, PLAIN>, tb, lds, maxBlocks);
	}
	static void launch(unsigned g, int tb, size_t lds, hipStream_t st, const WfParams &P, const DevScene &ds, const TraceBvh &bvh,
					   const RayQueueDev &q, const int *countPtr, const PixelStateDev &px, int *ovf, Counters *cnt, const NrcDev &nrc,
					   const ViewsDev &V, int startDepth = 0, const CameraDev &cam = CameraDev{}, int camFetch = 0) {
		if constexpr (REFILL)
			hipLaunchKernelGGL((k_megakernel_refill<COUNT, NRC, VIEWS, PLAIN>), dim3(g), dim3(tb), lds + (size_t) P.ldsNodes * sizeof(Bvh4Node), st, P,
							   ds, bvh, q, countPtr, startDepth, px, ovf, cnt, nrc, V, cam, camFetch);
		else
			hipLaunchKernelGGL((k_megakernel<MEDIA, COUNT, NRC, VIEWS, PLAIN>), dim3(g), dim3(tb), lds, st, P, ds, bvh, q, countPtr, startDepth, px, ovf,
							   cnt, nrc, V, cam, camFetch);
	}
};

template <bool VIEWS, bool REFILL, class F> auto withMegaR(bool media, bool count, bool nrc, bool plain, F &&f) {
	if (nrc && plain) return count ? f(Mega<false, true, true, VIEWS, true, REFILL>{}) : f(Mega<false, false, true, VIEWS, true, REFILL>{});
	if (nrc) return count ? f(Mega<false, true, true, VIEWS, false, REFILL>{}) : f(Mega<false, false, true, VIEWS, false, REFILL>{});
	if (media) return count ? f(Mega<true, true, false, VIEWS, false>{}) : f(Mega<true, false, false, VIEWS, false>{});
	if (plain) return count ? f(Mega<false, true, false, VIEWS, true, REFILL>{}) : f(Mega<false, false, false, VIEWS, true, REFILL>{});
	return count ? f(Mega<false, true, false, VIEWS, false, REFILL>{}) : f(Mega<false, false, false, VIEWS, false, REFILL>{});
}

// The persistent kernel for these flags; the refill launch for scenes
// without media when refill_lanes > 0.
template <bool VIEWS, class F> auto withMega(bool media, bool count, bool nrc, bool plain, bool refill, F &&f) {
	if (refill && !media) return withMegaR<VIEWS, true>(media, count, nrc, plain, f);
	return withMegaR<VIEWS, false>(media, count, nrc, plain, f);
}

} // namespace

struct krr_accel {
	AccelData a;
	DevBuf counters, ovf;
	int gClosest = 0, gAny = 0;
};

struct krr_wf {
	// params (integrator.h:96-103 + additions)
	bool nee = true, enableMedium = true, enableClamp = false, legacySpp = false, countTraversal = false, timeKernels = false;
	int maxDepth = 10, spp = 1, rrMinDepth = 0, blockSize = 256, traceBlockSize = 128, leafSize = 4, wavefrontDepths = 0;
	int stackLdsEntries = kShortStack;
	// persistent launch without media: traversal rounds end when fewer than
	// this many lanes of a wave are still traversing (0: each ray to its end)
	int refillLanes = -1; // -1: by the tree's size (refillLanesOf)
	// refill launch with float nodes: the first ldsNodes nodes of the tree
	// (its top levels: the device builder numbers nodes level by level) are
	// read from a per-block copy in LDS
	int ldsNodes = 0;
	// refill launch: iterations take leaves once this share (percent) of the
	// traversing lanes hold one, inner nodes otherwise (0: both every iteration)
	int leafShare = kDefaultLeafShare;
	float rr = 0.8f, clampMax = 1e3f, sahNodeCost = kSahNodeCost;
	bool deviceBuilder = true; // "bvh_builder": "device" | "host"
	bool quantizedNodes = false; // "node_format": "float" (default) | "quantized"
	// neural radiance cache ("nrc" params)
	bool nrcOn = false;
	int nrcTrainMod = 16, nrcBatch = 65536, nrcSteps = 4;
	float nrcSpread = 0.01f;
	std::string nrcConfig; // JSON for krr_nrc_create
	krr_nrc *nrc = nullptr;
	DevBuf nrcQuery, nrcTrainBuf, nrcBatchV, nrcBatchT;
	int gMegaNrc = 1;
	// persistent launches: fraction of the resident blocks launched
	float occupancy = 1.f;
	// multi-view frame (krr_wf_set_views / krr_wf_render_views): pixel state
	// of all views' pixels, numbered view after view
	std::vector<krr_view> views;
	int NV = 0;
	uint64_t viewsVersion = 0, sceneVersion = 0;
	DevBuf viewBuf, viewStart, frameBuf, vL, vPixel, vRng, vChannel, vNrcQuery;
	uint64_t viewsGridKey = ~0ull;
	int gMegaViews = 1;
	// the multi-view frame is recorded once into a graph and replayed
	bool useGraph = true;
	hipStream_t capStream = nullptr;
	hipGraph_t graph	  = nullptr;
	hipGraphExec_t graphExec = nullptr;
	uint64_t graphKey		 = ~0ull;
	bool haveCamera			 = false;
	void dropGraph() {
		if (graphExec) (void) hipGraphExecDestroy(graphExec);
		if (graph) (void) hipGraphDestroy(graph);
		graphExec = nullptr;
		graph	  = nullptr;
		graphKey  = ~0ull;
	}
	// scene
	bool haveScene = false;
	AccelData accel;
	DevBuf pos, nrm, idx, triMesh, meshes, materials, lights, media;
	std::vector<std::unique_ptr<DevBuf>> mediumBufs, imageBufs, textureBufs;
	DevBuf uv, tan, textures;
	DevScene ds{};
	bool sceneHasMedia = false;
	bool sceneTextured = false; // textures or image-lit infinite lights: the kernels with texture lookups
	// frame
	int W = 0, H = 0;
	DevBuf rayQ[2][2], hits, shadowRay, shadowC, shadowExt, L, pixel, rng, channel, counters, ovf;
	DevBuf plen, order, tileKey, tileCnt, tileHist, tileSorted, tileOff; // camera-path order of the persistent launch (k_tile_*)
	DevBuf vPlen, vOrder, vTileStart, vTileKey, vTileCnt, vTileHist, vTileSorted, vTileOff; // the same over all views
	int vTiles = 0;
	krr_camera cam{};
	uint64_t frameIndex = 0;
	uint64_t gridKey	= ~0ull;
	int gClosest = 1, gShadow = 1, gShade = 1, gMega = 1;
	bool beginPending	= false;
	hipStream_t lastStream = nullptr;
	krr_stats stats{};
	// HIP-event timing of kernel launches ("time_kernels"): kind 0 closest
	// trace, 1 shade, 2 shadow trace, 3 whole render, 4 tail megakernel,
	// 5 radiance cache (inference, targets, training).
	struct Span {
		int kind, begin, end;
	};
	std::vector<hipEvent_t> events;
	std::vector<Span> spans;
	int eventsUsed = 0;

	hipEvent_t event(int k) {
		while ((int) events.size() <= k) {
			hipEvent_t e = nullptr;
			if (hipEventCreate(&e) != hipSuccess) return nullptr;
			events.push_back(e);
		}
		return events[k];
	}
	int mark(hipStream_t st) {
		int k = eventsUsed++;
		hipEvent_t e = event(k);
		if (e) (void) hipEventRecord(e, st);
		return k;
	}
	~krr_wf() {
		dropGraph();
		if (capStream) (void) hipStreamDestroy(capStream);
		for (hipEvent_t e : events) (void) hipEventDestroy(e);
		if (nrc) krr_nrc_destroy(nrc);
	}
};

namespace {

int parseParams(krr_wf *wf, const char *json) {
	if (!json || !*json) return KRR_OK;
	try {
		Json j = Json::parse(json);
		if (!j.isObject()) return krr::fail(KRR_ERR_BAD_CONFIG, "pass params must be a JSON object");
		wf->nee			   = j.value("nee", true);
		wf->enableMedium   = j.value("enable_medium", true);
		wf->maxDepth	   = (int) j.value("max_depth", 10.0);
		wf->rr			   = (float) j.value("rr", 0.8);
		wf->enableClamp	   = j.value("enable_clamp", false);
		wf->clampMax	   = (float) j.value("clamp_max", 1e3);
		wf->spp			   = (int) j.value("spp", 1.0);
		wf->rrMinDepth	   = (int) j.value("rr_min_depth", 0.0);
		wf->legacySpp	   = j.value("legacy_spp_accumulate", false);
		wf->countTraversal = j.value("count_traversal", false);
		wf->timeKernels	   = j.value("time_kernels", false);
		wf->blockSize	   = (int) j.value("block_size", 256.0);
		wf->traceBlockSize = (int) j.value("trace_block_size", 128.0);
		wf->leafSize	   = (int) j.value("leaf_size", 4.0);
		wf->wavefrontDepths = (int) j.value("wavefront_depths", 0.0);
		wf->stackLdsEntries = (int) j.value("stack_lds_entries", (double) kShortStack);
		wf->refillLanes		= (int) j.value("refill_lanes", -1.0);
		wf->ldsNodes		= (int) j.value("lds_nodes", 0.0);
		wf->leafShare		= (int) j.value("leaf_share", (double) kDefaultLeafShare);
		if (j.contains("nrc")) {
			const Json &n = j.at("nrc");
			if (n.isBool()) wf->nrcOn = n.asBool();
			else if (n.isObject()) {
				wf->nrcOn		= n.value("enable", true);
				wf->nrcTrainMod = (int) n.value("train_mod", 16.0);
				wf->nrcBatch	= (int) n.value("batch", 65536.0);
				wf->nrcSteps	= (int) n.value("steps_per_frame", 4.0);
				wf->nrcSpread	= (float) n.value("spread", 0.01);
				Json c			= Json::object();
				for (const auto &kv : n.members())
					if (kv.first != "enable" && kv.first != "train_mod" && kv.first != "batch" && kv.first != "steps_per_frame" &&
						kv.first != "spread")
						c.set(kv.first, kv.second);
				c.set("max_batch", Json((double) wf->nrcBatch));
				wf->nrcConfig = c.dump();
			} else throw std::runtime_error("\"nrc\" must be a boolean or an object");
			if (wf->nrcConfig.empty()) wf->nrcConfig = "{\"max_batch\": " + std::to_string(wf->nrcBatch) + "}";
		}
		wf->sahNodeCost		= (float) j.value("sah_node_cost", (double) kSahNodeCost);
		wf->occupancy		= (float) j.value("occupancy", 1.0);
		const std::string bb = j.value("bvh_builder", "device");
		if (bb != "device" && bb != "host") throw std::runtime_error("bvh_builder must be \"device\" or \"host\"");
		wf->deviceBuilder = bb == "device";
		const std::string nf = j.value("node_format", "float");
		if (nf != "quantized" && nf != "float") throw std::runtime_error("node_format must be \"float\" or \"quantized\"");
		wf->quantizedNodes = nf == "quantized";
		wf->useGraph		= j.value("graph", true);
	} catch (const std::exception &e) {
		return krr::fail(KRR_ERR_BAD_CONFIG, "pass params: %s", e.what());
	}
	if (wf->maxDepth < 0 || wf->maxDepth > 250) return krr::fail(KRR_ERR_BAD_CONFIG, "max_depth must be in [0, 250]");
	if (wf->spp < 1 || wf->spp > 1 << 20) return krr::fail(KRR_ERR_BAD_CONFIG, "spp must be >= 1");
	if (wf->blockSize < 64 || wf->blockSize > 256 || wf->blockSize % 64) return krr::fail(KRR_ERR_BAD_CONFIG, "block_size must be 64, 128, 192 or 256");
	if (wf->traceBlockSize < 64 || wf->traceBlockSize > 256 || wf->traceBlockSize % 64)
		return krr::fail(KRR_ERR_BAD_CONFIG, "trace_block_size must be 64, 128, 192 or 256");
	if (wf->leafSize < 1 || wf->leafSize > 8) return krr::fail(KRR_ERR_BAD_CONFIG, "leaf_size must be in [1, 8]");
	if (wf->wavefrontDepths < 0) return krr::fail(KRR_ERR_BAD_CONFIG, "wavefront_depths must be >= 0");
	if (wf->stackLdsEntries < 0 || wf->stackLdsEntries > 96) return krr::fail(KRR_ERR_BAD_CONFIG, "stack_lds_entries must be in [0, 96]");
	if (wf->refillLanes < -1 || wf->refillLanes > 64) return krr::fail(KRR_ERR_BAD_CONFIG, "refill_lanes must be -1 (by scene) or in [0, 64]");
	if (wf->ldsNodes < 0 || wf->ldsNodes > kMaxLdsNodes) return krr::fail(KRR_ERR_BAD_CONFIG, "lds_nodes must be in [0, %d]", kMaxLdsNodes);
	if (wf->leafShare < 0 || wf->leafShare > 100) return krr::fail(KRR_ERR_BAD_CONFIG, "leaf_share must be in [0, 100]");
	if (wf->nrcTrainMod < 1 || wf->nrcTrainMod > 1 << 16) return krr::fail(KRR_ERR_BAD_CONFIG, "nrc.train_mod must be in [1, 65536]");
	if (wf->nrcBatch < 1024 || wf->nrcBatch > 1 << 22) return krr::fail(KRR_ERR_BAD_CONFIG, "nrc.batch must be in [1024, 2^22]");
	if (wf->nrcSteps < 1 || wf->nrcSteps > 64) return krr::fail(KRR_ERR_BAD_CONFIG, "nrc.steps_per_frame must be in [1, 64]");
	if (!(wf->nrcSpread >= 0.f && wf->nrcSpread <= 1e6f)) return krr::fail(KRR_ERR_BAD_CONFIG, "nrc.spread must be >= 0");
	if (!(wf->sahNodeCost >= 0.f && wf->sahNodeCost <= 64.f)) return krr::fail(KRR_ERR_BAD_CONFIG, "sah_node_cost must be in [0, 64]");
	if (!(wf->occupancy > 0.f && wf->occupancy <= 1.f)) return krr::fail(KRR_ERR_BAD_CONFIG, "occupancy must be in (0, 1]");
	return KRR_OK;
}

// Majorant grid build of src/render/media.cpp:18-75 on the dense grid (max
// density over the voxels each 1/64 cell touches, one voxel of slop).
void buildMajorant(const krr_medium &m, const float invVoxel[3], std::vector<float> &g) {
	g.assign((size_t) 64 * 64 * 64, 0.f);
	for (int z = 0; z < 64; z++)
		for (int y = 0; y < 64; y++)
			for (int x = 0; x < 64; x++) {
				const int c[3] = {x, y, z};
				int n0[3], n1[3];
				for (int a = 0; a < 3; a++) {
					const float diag = m.bounds_max[a] - m.bounds_min[a];
					const float lo	 = m.bounds_min[a] + (float(c[a]) / 64.f) * diag;
					const float hi	 = m.bounds_min[a] + (float(c[a] + 1) / 64.f) * diag;
					const float ilo	 = (lo - m.bounds_min[a]) * invVoxel[a];
					const float ihi	 = (hi - m.bounds_min[a]) * invVoxel[a];
					n0[a]			 = std::max((int) (ilo - 1.f), 0);
					n1[a]			 = std::min((int) (ihi + 1.f), m.res[a] - 1);
				}
				float mx = 0.f;
				for (int nz = n0[2]; nz <= n1[2]; ++nz)
					for (int ny = n0[1]; ny <= n1[1]; ++ny)
						for (int nx = n0[0]; nx <= n1[0]; ++nx)
							mx = std::max(mx, m.density[(size_t) nx + (size_t) m.res[0] * ((size_t) ny + (size_t) m.res[1] * nz)]);
				g[x + 64 * (y + 64 * z)] = mx;
			}
}

// Host-side checks of a scene description that need no device: array
// pointers against their counts and every mesh's triangle range, before any
// per-triangle table is written.
int validateSceneDesc(const krr_scene_desc *s) {
	if (s->n_triangles < 0 || (s->n_triangles > 0 && (!s->positions || !s->indices)))
		return krr::fail(KRR_ERR_INVALID_ARG, "scene: missing triangle arrays");
	if (s->n_vertices < 0 || (s->n_vertices > 0 && !s->positions)) return krr::fail(KRR_ERR_INVALID_ARG, "scene: missing positions");
	if (s->n_meshes < 0 || (s->n_meshes > 0 && !s->meshes)) return krr::fail(KRR_ERR_INVALID_ARG, "scene: missing meshes");
	for (int mi = 0; mi < s->n_meshes; mi++) {
		const krr_mesh &m = s->meshes[mi];
		if (m.first_triangle < 0 || m.n_triangles < 0 || (int64_t) m.first_triangle + m.n_triangles > s->n_triangles)
			return krr::fail(KRR_ERR_INVALID_ARG, "scene: mesh %d triangle range out of bounds", mi);
	}
	return KRR_OK;
}

int uploadScene(krr_wf *wf, const krr_scene_desc *s) {
	if (s->n_vertices < 0 || (s->n_vertices > 0 && !s->positions)) return krr::fail(KRR_ERR_INVALID_ARG, "scene: missing positions");
	if (s->n_meshes < 0 || (s->n_meshes > 0 && !s->meshes)) return krr::fail(KRR_ERR_INVALID_ARG, "scene: missing meshes");
	if (s->n_materials < 0 || (s->n_materials > 0 && !s->materials)) return krr::fail(KRR_ERR_INVALID_ARG, "scene: missing materials");
	if (s->n_media < 0 || (s->n_media > 0 && !s->media)) return krr::fail(KRR_ERR_INVALID_ARG, "scene: missing media");
	if (s->n_infinite_lights < 0 || (s->n_infinite_lights > 0 && !s->infinite_lights))
		return krr::fail(KRR_ERR_INVALID_ARG, "scene: missing infinite lights");
	if (int rc = validateSceneDesc(s)) return rc;
	// triangle -> mesh, mesh table with light bases (mesh lights in mesh order,
	// then infinite lights: RTScene::uploadSceneLightData, device/scene.cpp:80-135)
	std::vector<int> triMesh((size_t) s->n_triangles, -1);
	std::vector<DevMesh> meshes((size_t) s->n_meshes);
	std::vector<DevLight> lights;
	for (int mi = 0; mi < s->n_meshes; mi++) {
		const krr_mesh &m = s->meshes[mi];
		if (m.material >= s->n_materials) return krr::fail(KRR_ERR_INVALID_ARG, "scene: mesh %d material out of range", mi);
		if (m.medium_inside >= s->n_media || m.medium_outside >= s->n_media)
			return krr::fail(KRR_ERR_INVALID_ARG, "scene: mesh %d medium out of range", mi);
		DevMesh &d	 = meshes[mi];
		d.material	 = m.material < 0 ? -1 : m.material;
		d.firstTri	 = m.first_triangle;
		d.mediumIn	 = m.medium_inside < 0 ? -1 : m.medium_inside;
		d.mediumOut	 = m.medium_outside < 0 ? -1 : m.medium_outside;
		d.hasNormals = m.has_normals ? 1 : 0;
		if (m.has_normals && !s->normals) return krr::fail(KRR_ERR_INVALID_ARG, "scene: mesh %d has_normals but normals is NULL", mi);
		d.hasUV = m.has_texcoords ? 1 : 0;
		if (m.has_tangents && !s->tangents) return krr::fail(KRR_ERR_INVALID_ARG, "scene: mesh %d has_tangents but tangents is NULL", mi);
		d.hasTangents = m.has_tangents ? 1 : 0;
		if (m.has_texcoords && !s->texcoords) return krr::fail(KRR_ERR_INVALID_ARG, "scene: mesh %d has_texcoords but texcoords is NULL", mi);
		const bool emissive = m.emission[0] > 0 || m.emission[1] > 0 || m.emission[2] > 0;
		d.lightBase			= -1;
		if (emissive) {
			if (!m.has_normals) return krr::fail(KRR_ERR_INVALID_ARG, "scene: emissive mesh %d needs vertex normals", mi);
			d.lightBase		  = (int) lights.size();
			const float scale = std::max(std::max(m.emission[0], m.emission[1]), m.emission[2]);
			for (int t = 0; t < m.n_triangles; t++) {
				DevLight L{};
				L.type	= 0;
				L.prim	= m.first_triangle + t;
				L.Le[0] = m.emission[0] / scale;
				L.Le[1] = m.emission[1] / scale;
				L.Le[2] = m.emission[2] / scale;
				L.scale = scale;
				lights.push_back(L);
			}
		}
		for (int t = 0; t < m.n_triangles; t++) triMesh[m.first_triangle + t] = mi;
	}
	for (int64_t t = 0; t < s->n_triangles; t++)
		if (triMesh[t] < 0) return krr::fail(KRR_ERR_INVALID_ARG, "scene: triangle %lld belongs to no mesh", (long long) t);
	wf->imageBufs.clear();
	for (int i = 0; i < s->n_infinite_lights; i++) {
		const krr_infinite_light &il = s->infinite_lights[i];
		DevLight L{};
		L.type = 1;
		L.prim = -1;
		for (int a = 0; a < 3; a++) L.Le[a] = il.tint[a];
		L.scale	 = il.scale;
		bool any = false;
		for (int k = 0; k < 9; k++) any |= il.rotation[k] != 0.f;
		for (int k = 0; k < 9; k++) L.R[k] = any ? il.rotation[k] : (k % 4 == 0 ? 1.f : 0.f);
		if (il.image_rgba) {
			if (il.image_width < 1 || il.image_height < 1 || il.image_width > 65536 || il.image_height > 65536)
				return krr::fail(KRR_ERR_INVALID_ARG, "infinite light %d: image size %d x %d", i, il.image_width, il.image_height);
			auto buf = std::make_unique<DevBuf>();
			KRR_HIP(buf->upload(il.image_rgba, (size_t) il.image_width * il.image_height * 4 * sizeof(float)));
			L.img = buf->as<float4>();
			L.iw  = il.image_width;
			L.ih  = il.image_height;
			wf->imageBufs.push_back(std::move(buf));
		}
		lights.push_back(L);
	}
	std::vector<DevMaterial> mats((size_t) s->n_materials);
	for (int i = 0; i < s->n_materials; i++) {
		const krr_material &m = s->materials[i];
		if (m.type < KRR_MAT_NULL || m.type > KRR_MAT_CONDUCTOR) return krr::fail(KRR_ERR_UNSUPPORTED, "material %d: unsupported type %d", i, m.type);
		mats[i].type		= m.type;
		mats[i].roughness	= 1.f - m.specular[3];
		mats[i].ior			= m.ior;
		mats[i].specTrans	= m.specular_transmission;
		mats[i].anisotropic = m.anisotropic;
		for (int a = 0; a < 3; a++) mats[i].diffuse[a] = m.diffuse[a];
		if (m.diffuse_texture < 0 || m.diffuse_texture > s->n_textures || m.specular_texture < 0 || m.specular_texture > s->n_textures ||
			m.normal_texture < 0 || m.normal_texture > s->n_textures)
			return krr::fail(KRR_ERR_INVALID_ARG, "material %d: texture index out of range", i);
		mats[i].texDiffuse	= m.diffuse_texture - 1;
		mats[i].texSpecular = m.specular_texture - 1;
		mats[i].texNormal	= m.normal_texture - 1;
	}
	// textures
	if (s->n_textures < 0 || (s->n_textures > 0 && !s->textures)) return krr::fail(KRR_ERR_INVALID_ARG, "scene: missing textures");
	wf->textureBufs.clear();
	std::vector<DevTexture> texs((size_t) s->n_textures);
	for (int i = 0; i < s->n_textures; i++) {
		const krr_texture &t = s->textures[i];
		if (t.width < 1 || t.height < 1 || t.width > 65536 || t.height > 65536 || !t.rgba)
			return krr::fail(KRR_ERR_INVALID_ARG, "texture %d: bad image (%d x %d)", i, t.width, t.height);
		auto buf = std::make_unique<DevBuf>();
		KRR_HIP(buf->upload(t.rgba, (size_t) t.width * t.height * 4 * sizeof(float)));
		texs[i] = DevTexture{buf->as<float4>(), t.width, t.height};
		wf->textureBufs.push_back(std::move(buf));
	}
	// emissive textures of mesh lights (DiffuseAreaLight's texture, mesh.cpp:39-59)
	for (int mi = 0; mi < s->n_meshes; mi++) {
		const krr_mesh &m = s->meshes[mi];
		if (m.material < 0) continue;
		const int et = s->materials[m.material].emissive_texture;
		if (et < 0 || et > s->n_textures) return krr::fail(KRR_ERR_INVALID_ARG, "material %d: emissive_texture out of range", m.material);
		if (et == 0 || meshes[mi].lightBase < 0) continue;
		for (int t = 0; t < m.n_triangles; t++) {
			DevLight &L = lights[(size_t) meshes[mi].lightBase + t];
			L.img		= texs[et - 1].img;
			L.iw		= texs[et - 1].w;
			L.ih		= texs[et - 1].h;
			L.hasUV		= m.has_texcoords ? 1 : 0;
		}
	}
	// media
	wf->mediumBufs.clear();
	std::vector<DevMedium> media((size_t) s->n_media);
	for (int i = 0; i < s->n_media; i++) {
		const krr_medium &m = s->media[i];
		DevMedium &d		= media[i];
		std::memset(&d, 0, sizeof(d));
		d.type = m.type;
		for (int a = 0; a < 3; a++) {
			d.sigma_t[a] = m.sigma_t[a];
			d.albedo[a]	 = m.albedo[a];
			d.Le[a]		 = m.Le[a];
		}
		d.g		= m.g;
		d.scale = m.scale;
		if (m.type == KRR_MEDIUM_GRID) {
			for (int a = 0; a < 3; a++) {
				if (m.res[a] < 1 || !(m.bounds_max[a] > m.bounds_min[a])) return krr::fail(KRR_ERR_INVALID_ARG, "medium %d: bad grid", i);
				d.res[a]	  = m.res[a];
				d.bmin[a]	  = m.bounds_min[a];
				d.bmax[a]	  = m.bounds_max[a];
				d.invVoxel[a] = (float) m.res[a] / (m.bounds_max[a] - m.bounds_min[a]);
			}
			if (!m.density) return krr::fail(KRR_ERR_INVALID_ARG, "medium %d: null density", i);
			const size_t nvox = (size_t) m.res[0] * m.res[1] * m.res[2];
			auto db			  = std::make_unique<DevBuf>();
			KRR_HIP(db->upload(m.density, nvox * sizeof(float)));
			std::vector<float> maj;
			buildMajorant(m, d.invVoxel, maj);
			auto mb = std::make_unique<DevBuf>();
			KRR_HIP(mb->upload(maj.data(), maj.size() * sizeof(float)));
			d.density  = db->as<float>();
			d.majorant = mb->as<float>();
			wf->mediumBufs.push_back(std::move(db));
			wf->mediumBufs.push_back(std::move(mb));
		} else if (m.type != KRR_MEDIUM_HOMOGENEOUS) return krr::fail(KRR_ERR_UNSUPPORTED, "medium %d: unsupported type", i);
	}
	float sceneRadius = 1e5f;
	float lo[3] = {0.f, 0.f, 0.f}, hi[3] = {1.f, 1.f, 1.f};
	if (s->n_vertices > 0) {
		for (int a = 0; a < 3; a++) {
			lo[a] = INFINITY;
			hi[a] = -INFINITY;
		}
		for (int64_t v = 0; v < s->n_vertices; v++)
			for (int a = 0; a < 3; a++) {
				lo[a] = std::min(lo[a], s->positions[3 * v + a]);
				hi[a] = std::max(hi[a], s->positions[3 * v + a]);
			}
		float e0 = hi[0] - lo[0], e1 = hi[1] - lo[1], e2 = hi[2] - lo[2];
		sceneRadius = std::sqrt(e0 * e0 + e1 * e1 + e2 * e2);
	}
	KRR_HIP(wf->pos.upload(s->positions, (size_t) s->n_vertices * 3 * sizeof(float)));
	if (s->normals) KRR_HIP(wf->nrm.upload(s->normals, (size_t) s->n_vertices * 3 * sizeof(float)));
	else wf->nrm.release();
	KRR_HIP(wf->idx.upload(s->indices, (size_t) s->n_triangles * 3 * sizeof(int32_t)));
	KRR_HIP(wf->triMesh.upload(triMesh.data(), triMesh.size() * sizeof(int)));
	KRR_HIP(wf->meshes.upload(meshes.data(), meshes.size() * sizeof(DevMesh)));
	KRR_HIP(wf->materials.upload(mats.data(), mats.size() * sizeof(DevMaterial)));
	KRR_HIP(wf->lights.upload(lights.data(), lights.size() * sizeof(DevLight)));
	KRR_HIP(wf->media.upload(media.data(), media.size() * sizeof(DevMedium)));
	if (s->texcoords) KRR_HIP(wf->uv.upload(s->texcoords, (size_t) s->n_vertices * 2 * sizeof(float)));
	else wf->uv.release();
	if (s->tangents) KRR_HIP(wf->tan.upload(s->tangents, (size_t) s->n_vertices * 3 * sizeof(float)));
	else wf->tan.release();
	if (!texs.empty()) KRR_HIP(wf->textures.upload(texs.data(), texs.size() * sizeof(DevTexture)));
	else wf->textures.release();
	int rc = buildAccel(s, wf->accel, wf->leafSize, wf->sahNodeCost, wf->deviceBuilder);
	wf->accel.quantized = wf->quantizedNodes;
	if (rc) return rc;
	DevScene &ds	= wf->ds;
	ds.pos			= wf->pos.as<float>();
	ds.nrm			= s->normals ? wf->nrm.as<float>() : nullptr;
	ds.idx			= wf->idx.as<int>();
	ds.triMesh		= wf->triMesh.as<int>();
	ds.meshes		= wf->meshes.as<DevMesh>();
	ds.materials	= wf->materials.as<DevMaterial>();
	ds.lights		= wf->lights.as<DevLight>();
	ds.nLights		= (int) lights.size();
	ds.media		= wf->media.as<DevMedium>();
	ds.nMedia		= s->n_media;
	ds.uv			= s->texcoords ? wf->uv.as<float>() : nullptr;
	ds.tan			= s->tangents ? wf->tan.as<float>() : nullptr;
	ds.textures		= texs.empty() ? nullptr : wf->textures.as<DevTexture>();
	ds.sceneRadius	= sceneRadius;
	for (int a = 0; a < 3; a++) {
		ds.bmin[a] = lo[a];
		ds.binv[a] = hi[a] > lo[a] ? 1.f / (hi[a] - lo[a]) : 1.f;
	}
	wf->sceneHasMedia = s->n_media > 0;
	wf->sceneTextured = s->n_textures > 0 || wf->accel.alphaTri.p != nullptr;
	for (int i = 0; i < s->n_meshes; i++) wf->sceneTextured |= s->meshes[i].has_tangents != 0;
	for (int i = 0; i < s->n_infinite_lights; i++) wf->sceneTextured |= s->infinite_lights[i].image_rgba != nullptr;
	wf->haveScene	  = true;
	return KRR_OK;
}

// The refill threshold in force: refill_lanes, or with -1 (the default)
// kDefaultRefillLanes for trees of at least kRefillMinNodes nodes and 0 (each
// ray to its end) for smaller ones, whose short traversals do not pay for the
// rounds (36-triangle Cornell box 1024^2: 0.87 ms per frame without refill,
// 1.20-1.32 ms with 4-16).
int refillLanesOf(const krr_wf *wf) {
	if (wf->refillLanes >= 0) return wf->refillLanes;
	return wf->accel.nNodes >= kRefillMinNodes ? kDefaultRefillLanes : 0;
}

// Nodes of the tree the refill launch reads from LDS: lds_nodes, at most the
// tree's node count, none with quantised nodes or without the refill launch.
int ldsNodeCount(const krr_wf *wf) {
	if (wf->ldsNodes <= 0 || refillLanesOf(wf) <= 0 || wf->quantizedNodes) return 0;
	return (int) std::min<int64_t>(wf->ldsNodes, wf->accel.nNodes);
}

// Kernel parameters of one frame (integrator.cpp:205-267 values).
WfParams frameParams(const krr_wf *wf, int W, int H, int N, uint64_t frameIndex) {
	WfParams P;
	P.W			   = W;
	P.H			   = H;
	P.N			   = N;
	P.spp		   = wf->spp;
	P.maxDepth	   = wf->maxDepth;
	P.rr		   = wf->rr;
	P.rrMinDepth   = wf->rrMinDepth;
	P.nee		   = wf->nee ? 1 : 0;
	P.enableMedium = (wf->enableMedium && wf->sceneHasMedia) ? 1 : 0;
	P.enableClamp  = wf->enableClamp ? 1 : 0;
	P.clampMax	   = wf->clampMax;
	P.legacySpp	   = wf->legacySpp ? 1 : 0;
	P.sampleIndex  = (uint32_t) (frameIndex * (uint64_t) wf->spp);
	P.inc		   = ((uint64_t) P.sampleIndex << 1u) | 1u;
	P.refillLanes  = refillLanesOf(wf);
	P.ldsNodes	   = ldsNodeCount(wf);
	P.leafShare	   = wf->leafShare;
	return P;
}

CameraDev cameraDev(const krr_camera &c) {
	CameraDev cam;
	cam.filmH		  = c.film_size[1];
	cam.focalLength	  = c.focal_length;
	cam.focalDistance = c.focal_distance;
	cam.lensRadius	  = c.lens_radius;
	cam.aspect		  = c.aspect_ratio;
	std::memcpy(cam.M, c.transform, sizeof(cam.M));
	cam.medium = c.medium;
	return cam;
}

// Radiance-cache queues: queries of one sample of `queryCap` pixels, and the
// frame's training vertices.
int ensureNrcBuffers(krr_wf *wf, DevBuf &query, size_t queryCap) {
	if (!wf->nrcOn) return KRR_OK;
	if (query.sz < queryCap * 5 * sizeof(float4)) KRR_HIP(query.alloc(queryCap * 5 * sizeof(float4)));
	const size_t b = (size_t) wf->nrcBatch;
	if (!wf->nrcTrainBuf.p) KRR_HIP(wf->nrcTrainBuf.alloc(b * 6 * sizeof(float4)));
	if (!wf->nrcBatchV.p) KRR_HIP(wf->nrcBatchV.alloc(b * 4 * sizeof(float4)));
	if (!wf->nrcBatchT.p) KRR_HIP(wf->nrcBatchT.alloc(b * 3 * sizeof(float)));
	return KRR_OK;
}

int allocFrame(krr_wf *wf) {
	const size_t N = (size_t) wf->W * wf->H;
	for (int q = 0; q < 2; q++) {
		KRR_HIP(wf->rayQ[q][0].alloc(N * 2 * sizeof(float4)));
		KRR_HIP(wf->rayQ[q][1].alloc(N * 4 * sizeof(float4)));
	}
	KRR_HIP(wf->hits.alloc(N * sizeof(float4)));
	KRR_HIP(wf->shadowRay.alloc(N * 2 * sizeof(float4)));
	KRR_HIP(wf->shadowC.alloc(N * sizeof(float4)));
	KRR_HIP(wf->shadowExt.alloc(N * 3 * sizeof(float4)));
	KRR_HIP(wf->L.alloc(N * sizeof(float4)));
	KRR_HIP(wf->pixel.alloc(N * sizeof(float4)));
	KRR_HIP(wf->rng.alloc(N * sizeof(uint64_t)));
	KRR_HIP(wf->channel.alloc(N * sizeof(int)));
	KRR_HIP(wf->plen.alloc(N));
	KRR_HIP(hipMemset(wf->plen.p, 0, N));
	KRR_HIP(wf->order.alloc(N * sizeof(int)));
	{
		const size_t nTiles = (size_t) ((wf->W + 7) / 8) * ((wf->H + 7) / 8), nBlk = (nTiles + kTileBlock - 1) / kTileBlock;
		KRR_HIP(wf->tileKey.alloc(nTiles));
		KRR_HIP(wf->tileCnt.alloc(nTiles * sizeof(int)));
		KRR_HIP(wf->tileHist.alloc((kTileKeys * nBlk + 1) * sizeof(int)));
		KRR_HIP(wf->tileSorted.alloc(nTiles * sizeof(int)));
		KRR_HIP(wf->tileOff.alloc((nTiles + 1) * sizeof(int)));
	}
	hipLaunchKernelGGL(k_order_identity, dim3((unsigned) ((N + 255) / 256)), dim3(256), 0, nullptr, (int) N, wf->order.as<int>());
	KRR_HIP(hipGetLastError());
	KRR_HIP(hipDeviceSynchronize());
	if (!wf->counters.p) {
		KRR_HIP(wf->counters.alloc(sizeof(Counters)));
		KRR_HIP(hipMemset(wf->counters.p, 0, sizeof(Counters)));
	}
	return ensureNrcBuffers(wf, wf->nrcQuery, N);
}

RayQueueDev rayQueue(krr_wf *wf, int q) { return RayQueueDev{wf->rayQ[q][0].as<float4>(), wf->rayQ[q][1].as<float4>()}; }


} // namespace

extern "C" {

const char *krr_last_error(void) { return krr::lastErrorStorage().c_str(); }

int krr_abi_version(void) { return KRR_ABI_VERSION; }

int krr_device_count(void) {
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess) return 0;
	return n;
}

// ---------------------------------------------------------------------------
// Acceleration structure
// ---------------------------------------------------------------------------

int krr_accel_build(const krr_scene_desc *scene, const char *json_opts, krr_accel **out) {
	krr::clearError();
	if (!scene || !out) return krr::fail(KRR_ERR_INVALID_ARG, "krr_accel_build: null argument");
	*out = nullptr;
	int leafSize = 4, stackLdsEntries = kShortStack;
	float nodeCost = kSahNodeCost;
	bool deviceBuilder = true, quantized = false;
	if (json_opts && *json_opts) {
		try {
			Json j			= Json::parse(json_opts);
			leafSize		= (int) j.value("leaf_size", 4.0);
			const std::string b = j.value("builder", "device");
			if (b != "device" && b != "host") throw std::runtime_error("builder must be \"device\" or \"host\"");
			deviceBuilder = b == "device";
			const std::string nf = j.value("node_format", "float");
			if (nf != "quantized" && nf != "float") throw std::runtime_error("node_format must be \"float\" or \"quantized\"");
			quantized		= nf == "quantized";
			stackLdsEntries = (int) j.value("stack_lds_entries", (double) kShortStack);
			nodeCost		= (float) j.value("sah_node_cost", (double) kSahNodeCost);
		} catch (const std::exception &e) {
			return krr::fail(KRR_ERR_BAD_CONFIG, "krr_accel_build: %s", e.what());
		}
		if (leafSize < 1 || leafSize > 8) return krr::fail(KRR_ERR_BAD_CONFIG, "leaf_size must be in [1, 8]");
		if (stackLdsEntries < 0 || stackLdsEntries > 96) return krr::fail(KRR_ERR_BAD_CONFIG, "stack_lds_entries must be in [0, 96]");
		if (!(nodeCost >= 0.f && nodeCost <= 64.f)) return krr::fail(KRR_ERR_BAD_CONFIG, "sah_node_cost must be in [0, 64]");
	}
	if (int rc = validateSceneDesc(scene)) return rc;
	if (!deviceAvailable()) return krr::fail(KRR_ERR_NO_DEVICE, "krr_accel_build: no HIP device");
	for (int m = 0; m < scene->n_meshes; m++)
		if (scene->meshes[m].material >= scene->n_materials) return krr::fail(KRR_ERR_INVALID_ARG, "mesh %d material out of range", m);
	auto a = std::make_unique<krr_accel>();
	int rc = buildAccel(scene, a->a, leafSize, nodeCost, deviceBuilder);
	if (rc) return rc;
	a->a.stackLds  = stackLdsEntries;
	a->a.quantized = quantized;
	KRR_HIP(a->counters.alloc(sizeof(Counters)));
	KRR_HIP(hipMemset(a->counters.p, 0, sizeof(Counters)));
	*out = a.release();
	return KRR_OK;
}

// Both accel trace calls run a persistent grid that shares the accel's work
// counter and stack overflow area: calls on one accel must not overlap.
static constexpr int kAccelBlock = 128;

int krr_accel_trace_closest(krr_accel *accel, const float *rays_dev, int64_t n, krr_hit *hits_dev, void *hip_stream) {
	krr::clearError();
	if (!accel || n < 0 || (n > 0 && (!rays_dev || !hits_dev))) return krr::fail(KRR_ERR_INVALID_ARG, "krr_accel_trace_closest: bad argument");
	if (n > (1 << 30)) return krr::fail(KRR_ERR_CAPACITY, "krr_accel_trace_closest: at most 2^30 rays per call");
	if (n == 0) return KRR_OK;
	const int tb	 = kAccelBlock;
	const size_t lds = stackLds(accel->a, tb);
	if (!accel->gClosest) accel->gClosest = persistentGrid(k_trace_closest<false>, tb, lds, 1 << 20);
	const int grid = std::min(accel->gClosest, (int) ((n + tb - 1) / tb));
	if (int rc = reserveOverflow(accel->ovf, accel->a, (size_t) grid * tb)) return rc;
	hipStream_t st = (hipStream_t) hip_stream;
	Counters *cnt  = accel->counters.as<Counters>();
	hipLaunchKernelGGL(k_accel_begin, dim3(1), dim3(1), 0, st, cnt);
	hipLaunchKernelGGL(k_trace_closest<false>, dim3(grid), dim3(tb), lds, st, accel->a.dev(), (const float4 *) rays_dev, (float4 *) hits_dev,
					   (int) n, accel->ovf.as<int>(), cnt);
	KRR_HIP(hipGetLastError());
	return KRR_OK;
}

} // extern "C"

namespace {

// Device memory of krr_accel_trace_closest_noalloc: a Counters record whose
// work[0] is the persistent grid's ray counter, then the traversal-stack
// overflow of the grid.  Every GPU the library runs on holds its own copy
// until the process ends.  In host code `scratch` names a host-side stand-in:
// only the address hipGetSymbolAddress returns for the current device may
// reach a kernel or a copy.
constexpr size_t kScratchBytes		  = size_t(16) << 20;
constexpr size_t kScratchCounterBytes = (sizeof(Counters) + 255) / 256 * 256;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratchMu;
void *g_scratchDev[kMaxScratchDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device with that device current.
int scratchAddress(void **out) {
	int dev = 0;
	KRR_HIP(hipGetDevice(&dev));
	if (dev < 0 || dev >= kMaxScratchDevices) return krr::fail(KRR_ERR_CAPACITY, "scratch: device %d beyond %d devices", dev, kMaxScratchDevices);
	std::lock_guard<std::mutex> lock(g_scratchMu);
	if (!g_scratchDev[dev]) {
		void *p		 = nullptr;
		hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p) return krr::fail(KRR_ERR_HIP, "scratch: hipGetSymbolAddress: %s", hipGetErrorString(e));
		g_scratchDev[dev] = p;
	}
	*out = g_scratchDev[dev];
	return KRR_OK;
}

} // namespace

extern "C" {

size_t krr_accel_noalloc_scratch_bytes(void) { return kScratchBytes; }

int krr_accel_trace_closest_noalloc(krr_accel *accel, const float *rays_dev, int64_t n, krr_hit *hits_dev, void *hip_stream) {
	krr::clearError();
	if (!accel || n < 0 || (n > 0 && (!rays_dev || !hits_dev)))
		return krr::fail(KRR_ERR_INVALID_ARG, "krr_accel_trace_closest_noalloc: bad argument");
	if (n > (1 << 30)) return krr::fail(KRR_ERR_CAPACITY, "krr_accel_trace_closest_noalloc: at most 2^30 rays per call");
	if (n == 0) return KRR_OK;
	const int tb	 = kAccelBlock;
	const size_t lds = stackLds(accel->a, tb);
	// The resident grid is capped by what the scratch holds: each thread needs
	// its stack entries beyond the LDS part.
	const size_t perBlock = stackOverflowInts(accel->a) * tb * sizeof(int);
	const size_t room	  = kScratchBytes - kScratchCounterBytes;
	if (perBlock > room)
		return krr::fail(KRR_ERR_CAPACITY,
						 "krr_accel_trace_closest_noalloc: a %d-thread block needs %zu B of stack overflow, the scratch holds %zu B "
						 "(raise stack_lds_entries or use krr_accel_trace_closest)",
						 tb, perBlock, room);
	void *base = nullptr;
	if (int rc = scratchAddress(&base)) return rc;
	if (!accel->gClosest) accel->gClosest = persistentGrid(k_trace_closest<false>, tb, lds, 1 << 20);
	int grid = std::min(accel->gClosest, (int) ((n + tb - 1) / tb));
	if (perBlock > 0) grid = std::min(grid, (int) (room / perBlock));
	Counters *cnt  = reinterpret_cast<Counters *>(base);
	int *ovf	   = reinterpret_cast<int *>(reinterpret_cast<unsigned char *>(base) + kScratchCounterBytes);
	hipStream_t st = (hipStream_t) hip_stream;
	// the ray counter starts at zero on every call (a kernel, not a memset)
	hipLaunchKernelGGL(k_accel_begin, dim3(1), dim3(1), 0, st, cnt);
	hipLaunchKernelGGL(k_trace_closest<false>, dim3(grid), dim3(tb), lds, st, accel->a.dev(), (const float4 *) rays_dev, (float4 *) hits_dev,
					   (int) n, ovf, cnt);
	KRR_HIP(hipGetLastError());
	return KRR_OK;
}

int krr_accel_trace_any(krr_accel *accel, const float *rays_dev, int64_t n, int32_t *visible_dev, void *hip_stream) {
	krr::clearError();
	if (!accel || n < 0 || (n > 0 && (!rays_dev || !visible_dev))) return krr::fail(KRR_ERR_INVALID_ARG, "krr_accel_trace_any: bad argument");
	if (n > (1 << 30)) return krr::fail(KRR_ERR_CAPACITY, "krr_accel_trace_any: at most 2^30 rays per call");
	if (n == 0) return KRR_OK;
	const int tb	 = kAccelBlock;
	const size_t lds = stackLds(accel->a, tb);
	if (!accel->gAny) accel->gAny = persistentGrid(k_trace_any<false>, tb, lds, 1 << 20);
	const int grid = std::min(accel->gAny, (int) ((n + tb - 1) / tb));
	if (int rc = reserveOverflow(accel->ovf, accel->a, (size_t) grid * tb)) return rc;
	hipStream_t st = (hipStream_t) hip_stream;
	Counters *cnt  = accel->counters.as<Counters>();
	hipLaunchKernelGGL(k_accel_begin, dim3(1), dim3(1), 0, st, cnt);
	hipLaunchKernelGGL(k_trace_any<false>, dim3(grid), dim3(tb), lds, st, accel->a.dev(), (const float4 *) rays_dev, (int *) visible_dev,
					   (int) n, accel->ovf.as<int>(), cnt);
	KRR_HIP(hipGetLastError());
	return KRR_OK;
}

int krr_accel_info(const krr_accel *accel, int64_t *n_nodes, int32_t *max_depth, int64_t *n_leaf_tris, int64_t *bytes) {
	if (!accel) return krr::fail(KRR_ERR_INVALID_ARG, "krr_accel_info: null accel");
	if (n_nodes) *n_nodes = accel->a.nNodes;
	if (max_depth) *max_depth = accel->a.maxDepth;
	if (n_leaf_tris) *n_leaf_tris = accel->a.nLeafTris;
	if (bytes) *bytes = (int64_t) (accel->a.nodes.sz + accel->a.tris.sz);
	return KRR_OK;
}

int krr_accel_destroy(krr_accel *accel) {
	delete accel;
	return KRR_OK;
}

// ---------------------------------------------------------------------------
// WavefrontPathTracer
// ---------------------------------------------------------------------------

int krr_wf_create(const char *json_params, krr_wf **out) {
	krr::clearError();
	if (!out) return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_create: null out");
	*out	= nullptr;
	auto wf = std::make_unique<krr_wf>();
	int rc	= parseParams(wf.get(), json_params);
	if (rc) return rc;
	if (wf->nrcOn) {
		rc = krr_nrc_create(wf->nrcConfig.c_str(), &wf->nrc);
		if (rc) return rc;
	}
	*out = wf.release();
	return KRR_OK;
}

int krr_wf_get_params(const krr_wf *wf, char *buf, int32_t cap) {
	if (!wf) return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_get_params: null handle");
	Json j = Json::object();
	j.set("nee", Json(wf->nee));
	j.set("enable_medium", Json(wf->enableMedium));
	j.set("max_depth", Json((double) wf->maxDepth));
	j.set("rr", Json((double) wf->rr));
	j.set("enable_clamp", Json(wf->enableClamp));
	j.set("clamp_max", Json((double) wf->clampMax));
	j.set("spp", Json((double) wf->spp));
	j.set("rr_min_depth", Json((double) wf->rrMinDepth));
	j.set("legacy_spp_accumulate", Json(wf->legacySpp));
	j.set("count_traversal", Json(wf->countTraversal));
	j.set("time_kernels", Json(wf->timeKernels));
	j.set("block_size", Json((double) wf->blockSize));
	j.set("trace_block_size", Json((double) wf->traceBlockSize));
	j.set("leaf_size", Json((double) wf->leafSize));
	j.set("wavefront_depths", Json((double) wf->wavefrontDepths));
	j.set("stack_lds_entries", Json((double) wf->stackLdsEntries));
	j.set("refill_lanes", Json((double) wf->refillLanes));
	j.set("refill_lanes_in_force", Json((double) refillLanesOf(wf))); // -1 resolved by the scene set
	j.set("lds_nodes", Json((double) wf->ldsNodes));
	j.set("leaf_share", Json((double) wf->leafShare));
	j.set("sah_node_cost", Json((double) wf->sahNodeCost));
	j.set("occupancy", Json((double) wf->occupancy));
	j.set("bvh_builder", Json(std::string(wf->deviceBuilder ? "device" : "host")));
	j.set("node_format", Json(std::string(wf->quantizedNodes ? "quantized" : "float")));
	j.set("graph", Json(wf->useGraph));
	if (wf->nrcOn) {
		Json n = Json::object();
		n.set("train_mod", Json((double) wf->nrcTrainMod));
		n.set("batch", Json((double) wf->nrcBatch));
		n.set("steps_per_frame", Json((double) wf->nrcSteps));
		n.set("spread", Json((double) wf->nrcSpread));
		j.set("nrc", n);
	} else j.set("nrc", Json(false));
	std::string s = j.dump();
	if (buf && cap > 0) {
		size_t n = std::min((size_t) cap - 1, s.size());
		std::memcpy(buf, s.data(), n);
		buf[n] = 0;
	}
	return (int) s.size();
}

int krr_wf_set_scene(krr_wf *wf, const krr_scene_desc *scene) {
	krr::clearError();
	if (!wf || !scene) return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_set_scene: null argument");
	if (int rc = validateSceneDesc(scene)) return rc;
	if (!deviceAvailable()) return krr::fail(KRR_ERR_NO_DEVICE, "krr_wf_set_scene: no HIP device");
	KRR_HIP(hipDeviceSynchronize());
	wf->haveScene = false;
	wf->sceneVersion++;
	wf->dropGraph();
	int rc		  = uploadScene(wf, scene);
	if (rc) return rc;
	if (wf->W > 0 && wf->H > 0) return allocFrame(wf);
	return KRR_OK;
}

int krr_wf_resize(krr_wf *wf, int32_t width, int32_t height) {
	krr::clearError();
	if (!wf || width < 0 || height < 0) return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_resize: bad argument");
	if ((int64_t) width * height > (int64_t) 1 << 26) return krr::fail(KRR_ERR_CAPACITY, "krr_wf_resize: at most 2^26 pixels per frame");
	if (width > 65536 || height > 65536) return krr::fail(KRR_ERR_CAPACITY, "krr_wf_resize: width and height must be <= 65536");
	if (!deviceAvailable()) return krr::fail(KRR_ERR_NO_DEVICE, "krr_wf_resize: no HIP device");
	KRR_HIP(hipDeviceSynchronize());
	wf->W = width;
	wf->H = height;
	if (width == 0 || height == 0) return KRR_OK;
	return allocFrame(wf);
}

int krr_wf_begin_frame(krr_wf *wf, const krr_camera *camera, uint64_t frame_index) {
	krr::clearError();
	if (!wf || !camera) return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_begin_frame: null argument");
	if (!wf->haveScene || wf->W == 0 || wf->H == 0) return KRR_OK; // integrator.cpp:206 early return
	if (camera->medium >= wf->ds.nMedia) return krr::fail(KRR_ERR_INVALID_ARG, "camera medium out of range");
	wf->cam			 = *camera;
	wf->frameIndex	 = frame_index;
	wf->beginPending = true;
	wf->haveCamera	 = true;
	return KRR_OK;
}

int krr_wf_render(krr_wf *wf, float *film_dev, void *hip_stream) {
	krr::clearError();
	if (!wf) return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_render: null handle");
	if (!wf->haveScene) return krr::fail(KRR_ERR_NO_SCENE, "krr_wf_render: no scene set");
	if (wf->W == 0 || wf->H == 0) return KRR_OK; // integrator.cpp:224 early return
	if (!film_dev) return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_render: null film");
	hipStream_t st = (hipStream_t) hip_stream;
	wf->lastStream = st;
	const int N		   = wf->W * wf->H;
	const WfParams P   = frameParams(wf, wf->W, wf->H, N, wf->frameIndex);
	const CameraDev cam = cameraDev(wf->cam);

	PixelStateDev px{wf->L.as<float4>(), wf->pixel.as<float4>(), wf->rng.as<uint64_t>(), wf->channel.as<int>(), wf->order.as<int>(),
					 wf->plen.as<uint8_t>()};
	ShadowQueueDev sq{wf->shadowRay.as<float4>(), wf->shadowC.as<float4>(), wf->shadowExt.as<float4>()};
	Counters *cnt	   = wf->counters.as<Counters>();
	wf->accel.stackLds = wf->stackLdsEntries; // before dev(): the kernels' LDS stack depth must match lds below
	const TraceBvh bvh = wf->accel.dev();
	const int bs	   = wf->blockSize;
	const unsigned gN  = (unsigned) ((N + bs - 1) / bs);
	const int tb	   = wf->traceBlockSize;
	const int gTmax	   = (N + tb - 1) / tb;
	const size_t lds   = stackLds(wf->accel, tb);
	if ((size_t) (bvh.stackLds + 1) * tb * sizeof(int) != lds) return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_render: stack layout mismatch");
	const bool count   = wf->countTraversal;
	const bool media   = wf->enableMedium && wf->sceneHasMedia;
	const bool plain   = !wf->sceneTextured;
	const bool refill  = refillLanesOf(wf) > 0;
	// persistent grid sizes (resident blocks), computed once per shape
	const uint64_t shapeKey = ((uint64_t) N << 24) ^ ((uint64_t) tb << 12) ^ ((uint64_t) bs << 2) ^ ((uint64_t) wf->accel.stackSize << 50) ^ ((uint64_t) wf->stackLdsEntries << 43) ^
							  (count ? 2 : 0) ^ (media ? 1 : 0) ^ (wf->nrcOn ? 4 : 0) ^ (plain ? 8 : 0) ^ (refill ? 16 : 0) ^
							  ((uint64_t) ldsNodeCount(wf) << 5) ^ ((uint64_t) (wf->occupancy * 1024.f) << 32);
	if (shapeKey != wf->gridKey) {
		// Counting variants use more registers: size each grid for the variant launched.
		wf->gClosest = count ? persistentGrid(k_trace_closest_queue<true>, tb, lds, gTmax) : persistentGrid(k_trace_closest_queue<false>, tb, lds, gTmax);
		if (media) wf->gShadow = count ? persistentGrid(k_shadow_tr<true>, tb, lds, gTmax) : persistentGrid(k_shadow_tr<false>, tb, lds, gTmax);
		else wf->gShadow = count ? persistentGrid(k_trace_shadow<true>, tb, lds, gTmax) : persistentGrid(k_trace_shadow<false>, tb, lds, gTmax);
		wf->gShade = media ? persistentGrid(k_shade<true>, bs, 0, (int) gN) : persistentGrid(k_shade<false>, bs, 0, (int) gN);
		const size_t nodeBytes = (size_t) ldsNodeCount(wf) * sizeof(Bvh4Node);
		wf->gMega = withMega<false>(media, count, false, plain, refill, [&](auto k) { return decltype(k)::grid(tb, lds, gTmax, nodeBytes); });
		if (wf->nrcOn)
			wf->gMegaNrc = withMega<false>(media, count, true, plain, refill, [&](auto k) { return decltype(k)::grid(tb, lds, gTmax, nodeBytes); });
		wf->gMega	 = scaledGrid(wf->gMega, wf->occupancy);
		wf->gMegaNrc = scaledGrid(wf->gMegaNrc, wf->occupancy);
		const int gMax = std::max(std::max(std::max(wf->gClosest, wf->gShadow), wf->gMega), wf->nrcOn ? wf->gMegaNrc : 0);
		if (int rc = reserveOverflow(wf->ovf, wf->accel, (size_t) gMax * tb)) return rc;
		wf->gridKey = shapeKey;
	}
	int *ovf = wf->ovf.as<int>();
	const unsigned gTc = (unsigned) wf->gClosest, gTs = (unsigned) wf->gShadow, gS = (unsigned) wf->gShade;

	const bool timing = wf->timeKernels;
	wf->spans.clear();
	wf->eventsUsed = 0;
	const int renderBegin = timing ? wf->mark(st) : -1;
	auto timed = [&](int kind, auto &&launch) {
		if (!timing) {
			launch();
			return;
		}
		const int b = wf->mark(st);
		launch();
		wf->spans.push_back({kind, b, wf->mark(st)});
	};
	int nrcRc = KRR_OK;
	if (wf->nrcOn && (media || wf->wavefrontDepths != 0))
		return krr::fail(KRR_ERR_UNSUPPORTED, "nrc: needs a scene without media and wavefront_depths = 0");
	hipLaunchKernelGGL(k_reset_totals, dim3(1), dim3(1), 0, st, cnt);
	if (wf->nrcOn) hipLaunchKernelGGL(k_nrc_begin, dim3(1), dim3(1), 0, st, cnt, wf->nrcBatch);
	if (wf->beginPending) {
		hipLaunchKernelGGL(k_begin_frame, dim3(gN), dim3(bs), 0, st, P, px);
		wf->beginPending = false;
	}
	NrcDev nrc{};
	if (wf->nrcOn) {
		nrc.query	   = wf->nrcQuery.as<float4>();
		nrc.train	   = wf->nrcTrainBuf.as<float4>();
		nrc.queryCount = &cnt->nrcQueries;
		nrc.trainCount = &cnt->nrcTrain;
		nrc.candCount  = &cnt->nrcCandidates;
		nrc.keep	   = &cnt->nrcKeep;
		nrc.spread	   = wf->nrcSpread;
		nrc.capQuery   = N;
		nrc.capTrain   = wf->nrcBatch;
		nrc.trainMod   = wf->nrcTrainMod;
		nrc.phase	   = (int) (wf->frameIndex % (uint64_t) wf->nrcTrainMod);
	}
	for (int s = 0; s < wf->spp; s++) {
		// with every depth in the persistent launch, it makes the camera rays itself
		if (wf->wavefrontDepths > 0) hipLaunchKernelGGL(k_camera, dim3(gN), dim3(bs), 0, st, P, cam, px, rayQueue(wf, 0), cnt);
		for (int depth = 0; depth <= wf->maxDepth; depth++) {
			const int cur		 = depth & 1;
			const RayQueueDev cq = rayQueue(wf, cur);
			if (depth >= wf->wavefrontDepths) { // the rest of every path in one launch
				hipLaunchKernelGGL(k_mega_begin, dim3(1), dim3(1), 0, st, cnt);
				if (depth == 0) { // camera paths: the order from the last sample's path lengths
					const int tilesX = (wf->W + 7) / 8, nTiles = tilesX * ((wf->H + 7) / 8), nBlk = (nTiles + kTileBlock - 1) / kTileBlock;
					uint8_t *key = wf->tileKey.as<uint8_t>();
					int *tc = wf->tileCnt.as<int>(), *hist = wf->tileHist.as<int>(), *sorted = wf->tileSorted.as<int>(), *off = wf->tileOff.as<int>();
					hipLaunchKernelGGL(k_tile_keys, dim3(nTiles), dim3(64), 0, st, wf->plen.as<uint8_t>(), wf->W, wf->H, tilesX, key, tc);
					hipLaunchKernelGGL(k_tile_hist, dim3(nBlk), dim3(kTileBlock), 0, st, key, nTiles, nBlk, hist);
					hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(1024), 0, st, hist, kTileKeys * nBlk);
					hipLaunchKernelGGL(k_tile_rank, dim3(nBlk), dim3(kTileBlock), 0, st, key, tc, nTiles, nBlk, hist, sorted, off);
					hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(1024), 0, st, off, nTiles);
					hipLaunchKernelGGL(k_tile_pixels, dim3(nTiles), dim3(64), 0, st, sorted, off, wf->W, wf->H, tilesX, wf->order.as<int>());
				}
				timed(4, [&] {
					const unsigned g = (unsigned) wf->gMega;
					const int camFetch = depth == 0 ? 1 : 0; // k_camera skipped: the kernel makes the camera rays
					const int *cp	 = (const int *) &cnt->rayCount[cur];
					withMega<false>(media, count, wf->nrcOn, plain, refill, [&](auto k) {
						decltype(k)::launch(wf->nrcOn ? (unsigned) wf->gMegaNrc : g, tb, lds, st, P, wf->ds, bvh, cq, cp, px, ovf, cnt,
											wf->nrcOn ? nrc : NrcDev{}, ViewsDev{}, depth, cam, camFetch);
						return 0;
					});
				});
				break;
			}
			hipLaunchKernelGGL(k_depth_begin, dim3(1), dim3(1), 0, st, cnt, cur);
			timed(0, [&] {
				if (count)
					hipLaunchKernelGGL(k_trace_closest_queue<true>, dim3(gTc), dim3(tb), lds, st, bvh, (const float4 *) cq.ray,
									   wf->hits.as<float4>(), (const int *) &cnt->rayCount[cur], ovf, cnt);
				else
					hipLaunchKernelGGL(k_trace_closest_queue<false>, dim3(gTc), dim3(tb), lds, st, bvh, (const float4 *) cq.ray,
									   wf->hits.as<float4>(), (const int *) &cnt->rayCount[cur], ovf, cnt);
			});
			timed(1, [&] {
				if (media)
					hipLaunchKernelGGL(k_shade<true>, dim3(gS), dim3(bs), 0, st, P, wf->ds, cq, (const float4 *) wf->hits.as<float4>(),
									   rayQueue(wf, cur ^ 1), sq, px, cnt, cur, depth);
				else
					hipLaunchKernelGGL(k_shade<false>, dim3(gS), dim3(bs), 0, st, P, wf->ds, cq, (const float4 *) wf->hits.as<float4>(),
									   rayQueue(wf, cur ^ 1), sq, px, cnt, cur, depth);
			});
			if (wf->nee && depth != wf->maxDepth) {
				timed(2, [&] {
					if (media) {
						if (count) hipLaunchKernelGGL(k_shadow_tr<true>, dim3(gTs), dim3(tb), lds, st, P, wf->ds, bvh, sq, px, ovf, cnt);
						else hipLaunchKernelGGL(k_shadow_tr<false>, dim3(gTs), dim3(tb), lds, st, P, wf->ds, bvh, sq, px, ovf, cnt);
					} else {
						if (count) hipLaunchKernelGGL(k_trace_shadow<true>, dim3(gTs), dim3(tb), lds, st, bvh, sq, px, ovf, cnt);
						else hipLaunchKernelGGL(k_trace_shadow<false>, dim3(gTs), dim3(tb), lds, st, bvh, sq, px, ovf, cnt);
					}
				});
			}
		}
		if (wf->nrcOn) { // training targets of this sample, then the cache answers the queries
			timed(5, [&] {
				hipLaunchKernelGGL(k_nrc_targets, dim3(256), dim3(256), 0, st, nrc, (const int *) &cnt->nrcTrainDone, (const float4 *) px.L,
								   wf->nrcBatchV.as<float4>(), wf->nrcBatchT.as<float>());
				hipLaunchKernelGGL(k_nrc_mark_done, dim3(1), dim3(1), 0, st, cnt, wf->nrcBatch, N);
				nrcRc = nrcInferAccumulate(wf->nrc, nrc.query, 5, &cnt->nrcQueries, N, px.L, st);
			});
			if (nrcRc) return nrcRc;
		}
		// the last sample is added to the pixel sums by the film write
		if (s + 1 < wf->spp) hipLaunchKernelGGL(k_accumulate, dim3(gN), dim3(bs), 0, st, N, px);
	}
	if (wf->nrcOn) { // one training step per frame on the recorded vertices
		timed(5, [&] { nrcRc = nrcTrainDevice(wf->nrc, wf->nrcBatchV.as<float>(), wf->nrcBatchT.as<float>(), &cnt->nrcTrainDone, wf->nrcBatch, wf->nrcSteps, st);
		});
		if (nrcRc) return nrcRc;
	}
	hipLaunchKernelGGL(k_film, dim3(gN), dim3(bs), 0, st, P, px, (float4 *) film_dev);
	if (timing) wf->spans.push_back({3, renderBegin, wf->mark(st)});
	KRR_HIP(hipGetLastError());
	wf->stats.paths = (uint64_t) N * wf->spp;
	return KRR_OK;
}

int krr_wf_end_frame(krr_wf *wf) {
	if (!wf) return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_end_frame: null handle");
	return KRR_OK;
}

} // extern "C"

namespace {

ViewsDev viewsDev(krr_wf *wf) {
	return ViewsDev{wf->viewBuf.as<ViewDev>(), wf->viewStart.as<int>(), (int) wf->views.size(), wf->frameBuf.as<FrameDev>(), wf->vTileStart.as<int>()};
}
PixelStateDev viewsPx(krr_wf *wf) {
	return PixelStateDev{wf->vL.as<float4>(), wf->vPixel.as<float4>(), wf->vRng.as<uint64_t>(), wf->vChannel.as<int>(), wf->vOrder.as<int>(),
						 wf->vPlen.as<uint8_t>()};
}

// Launch shape of the multi-view persistent kernel over N paths.
struct ViewsLaunch {
	TraceBvh bvh;
	size_t lds;
	unsigned grid;
	bool media, count, nrc;
};

int viewsLaunch(krr_wf *wf, int N, bool nrc, ViewsLaunch &L) {
	wf->accel.stackLds = wf->stackLdsEntries; // before dev(), as in krr_wf_render
	L.bvh			   = wf->accel.dev();
	const int tb	   = wf->traceBlockSize;
	L.lds			   = stackLds(wf->accel, tb);
	L.count			   = wf->countTraversal;
	L.media			   = wf->enableMedium && wf->sceneHasMedia;
	L.nrc			   = nrc;
	const uint64_t key = ((uint64_t) N << 24) ^ ((uint64_t) tb << 12) ^ ((uint64_t) wf->accel.stackSize << 50) ^
						 ((uint64_t) wf->stackLdsEntries << 43) ^ (L.count ? 2 : 0) ^ (L.media ? 1 : 0) ^ (nrc ? 4 : 0) ^ (wf->sceneTextured ? 0 : 8) ^
						 (refillLanesOf(wf) > 0 ? 16 : 0) ^ ((uint64_t) ldsNodeCount(wf) << 5) ^ ((uint64_t) (wf->occupancy * 1024.f) << 32);
	if (key != wf->viewsGridKey) {
		const int maxBlocks = std::max(1, (N + tb - 1) / tb);
		const size_t nodeBytes = (size_t) ldsNodeCount(wf) * sizeof(Bvh4Node);
		const int g = withMega<true>(L.media, L.count, nrc, !wf->sceneTextured, refillLanesOf(wf) > 0,
									 [&](auto k) { return decltype(k)::grid(tb, L.lds, maxBlocks, nodeBytes); });
		wf->gMegaViews		= scaledGrid(g, wf->occupancy);
		if (int rc = reserveOverflow(wf->ovf, wf->accel, (size_t) wf->gMegaViews * tb)) return rc;
		wf->viewsGridKey = key;
	}
	L.grid = (unsigned) wf->gMegaViews;
	return KRR_OK;
}

// Everything of one multi-view frame after the frame values are set:
// beginFrame of all views, each sample in one persistent launch over all
// views' pixels (plus the cache queries), accumulate, the cache's training,
// the film write of every view.  `trial` stops after the first persistent
// launch and writes no film (autotuner).
int enqueueViews(krr_wf *wf, const WfParams &P, const ViewsLaunch &L, hipStream_t st, bool timing, bool trial) {
	const ViewsDev V	   = viewsDev(wf);
	const PixelStateDev px = viewsPx(wf);
	Counters *cnt		   = wf->counters.as<Counters>();
	const int tb		   = wf->traceBlockSize;
	const unsigned gN	   = (unsigned) ((P.N + wf->blockSize - 1) / wf->blockSize);
	int *ovf			   = wf->ovf.as<int>();
	NrcDev nrc{};
	if (L.nrc) {
		nrc.query	   = wf->vNrcQuery.as<float4>();
		nrc.train	   = wf->nrcTrainBuf.as<float4>();
		nrc.queryCount = &cnt->nrcQueries;
		nrc.trainCount = &cnt->nrcTrain;
		nrc.candCount  = &cnt->nrcCandidates;
		nrc.keep	   = &cnt->nrcKeep;
		nrc.spread	   = wf->nrcSpread;
		nrc.capQuery   = P.N;
		nrc.capTrain   = wf->nrcBatch;
		nrc.trainMod   = wf->nrcTrainMod; // phase comes from the frame values
	}
	auto timed = [&](int kind, auto &&launch) {
		if (!timing) {
			launch();
			return;
		}
		const int b = wf->mark(st);
		launch();
		wf->spans.push_back({kind, b, wf->mark(st)});
	};
	hipLaunchKernelGGL(k_reset_totals, dim3(1), dim3(1), 0, st, cnt);
	if (L.nrc) hipLaunchKernelGGL(k_nrc_begin, dim3(1), dim3(1), 0, st, cnt, wf->nrcBatch);
	hipLaunchKernelGGL(k_begin_views, dim3(gN), dim3(wf->blockSize), 0, st, P, V, px);
	int rc = KRR_OK;
	for (int s = 0; s < wf->spp; s++) {
		hipLaunchKernelGGL(k_mega_begin, dim3(1), dim3(1), 0, st, cnt);
		{ // the order of this sample's camera paths, from the last sample's path lengths (k_tile_keys)
			const int nT = wf->vTiles, nBlk = (nT + kTileBlock - 1) / kTileBlock;
			uint8_t *key = wf->vTileKey.as<uint8_t>();
			int *tc = wf->vTileCnt.as<int>(), *hist = wf->vTileHist.as<int>(), *sorted = wf->vTileSorted.as<int>(), *off = wf->vTileOff.as<int>();
			hipLaunchKernelGGL(k_tile_keys_views, dim3(nT), dim3(64), 0, st, wf->vPlen.as<uint8_t>(), V, key, tc);
			hipLaunchKernelGGL(k_tile_hist, dim3(nBlk), dim3(kTileBlock), 0, st, key, nT, nBlk, hist);
			hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(1024), 0, st, hist, kTileKeys * nBlk);
			hipLaunchKernelGGL(k_tile_rank, dim3(nBlk), dim3(kTileBlock), 0, st, key, tc, nT, nBlk, hist, sorted, off);
			hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(1024), 0, st, off, nT);
			hipLaunchKernelGGL(k_tile_pixels_views, dim3(nT), dim3(64), 0, st, sorted, off, V, wf->vOrder.as<int>());
		}
		timed(4, [&] {
			withMega<true>(L.media, L.count, L.nrc, !wf->sceneTextured, refillLanesOf(wf) > 0, [&](auto k) {
				decltype(k)::launch(L.grid, tb, L.lds, st, P, wf->ds, L.bvh, RayQueueDev{}, nullptr, px, ovf, cnt, nrc, V);
				return 0;
			});
		});
		if (trial) return KRR_OK;
		if (L.nrc) {
			timed(5, [&] {
				hipLaunchKernelGGL(k_nrc_targets, dim3(256), dim3(256), 0, st, nrc, (const int *) &cnt->nrcTrainDone, (const float4 *) px.L,
								   wf->nrcBatchV.as<float4>(), wf->nrcBatchT.as<float>());
				hipLaunchKernelGGL(k_nrc_mark_done, dim3(1), dim3(1), 0, st, cnt, wf->nrcBatch, P.N);
				rc = nrcInferAccumulate(wf->nrc, nrc.query, 5, &cnt->nrcQueries, P.N, px.L, st);
			});
			if (rc) return rc;
		}
		hipLaunchKernelGGL(k_accumulate, dim3(gN), dim3(wf->blockSize), 0, st, P.N, px);
	}
	if (L.nrc) {
		timed(5, [&] { rc = nrcTrainDevice(wf->nrc, wf->nrcBatchV.as<float>(), wf->nrcBatchT.as<float>(), &cnt->nrcTrainDone, wf->nrcBatch, wf->nrcSteps, st); });
		if (rc) return rc;
	}
	hipLaunchKernelGGL(k_film_views, dim3(gN), dim3(wf->blockSize), 0, st, P, V, px);
	return KRR_OK;
}

FrameDev frameDev(const krr_wf *wf, uint64_t frameIndex) {
	FrameDev f;
	f.sampleIndex = (uint32_t) (frameIndex * (uint64_t) wf->spp);
	f.inc		  = ((uint64_t) f.sampleIndex << 1u) | 1u;
	f.phase		  = (int) (frameIndex % (uint64_t) std::max(1, wf->nrcTrainMod));
	return f;
}

} // namespace

extern "C" {

int krr_wf_set_views(krr_wf *wf, const krr_view *views, int32_t n_views) {
	krr::clearError();
	if (!wf || !views || n_views < 1) return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_set_views: need a handle and at least one view");
	if (n_views > 1 << 20) return krr::fail(KRR_ERR_CAPACITY, "krr_wf_set_views: at most 2^20 views");
	std::vector<ViewDev> host((size_t) n_views);
	std::vector<int> start((size_t) n_views + 1), tileStart((size_t) n_views + 1);
	int64_t total = 0, tiles = 0;
	for (int v = 0; v < n_views; v++) {
		const krr_view &kv = views[v];
		if (kv.width < 1 || kv.height < 1 || kv.width > 65536 || kv.height > 65536)
			return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_set_views: view %d size %dx%d outside [1, 65536]", v, kv.width, kv.height);
		if (!kv.film_dev) return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_set_views: view %d has no film", v);
		start[v]	 = (int) total;
		total		+= (int64_t) kv.width * kv.height;
		tileStart[v] = (int) tiles;
		tiles		+= (int64_t) ((kv.width + 7) / 8) * ((kv.height + 7) / 8);
		if (total > (int64_t) 1 << 30) return krr::fail(KRR_ERR_CAPACITY, "krr_wf_set_views: at most 2^30 pixels over all views");
		host[v].cam	 = cameraDev(kv.camera);
		host[v].W	 = kv.width;
		host[v].H	 = kv.height;
		host[v].film = (float4 *) kv.film_dev;
	}
	start[n_views]	   = (int) total;
	tileStart[n_views] = (int) tiles;
	if (!deviceAvailable()) return krr::fail(KRR_ERR_NO_DEVICE, "krr_wf_set_views: no HIP device");
	KRR_HIP(hipDeviceSynchronize());
	wf->dropGraph();
	const size_t N = (size_t) total;
	KRR_HIP(wf->viewBuf.upload(host.data(), host.size() * sizeof(ViewDev)));
	KRR_HIP(wf->viewStart.upload(start.data(), start.size() * sizeof(int)));
	KRR_HIP(wf->vTileStart.upload(tileStart.data(), tileStart.size() * sizeof(int)));
	{ // camera-path order over all views: identity until the first sample's path lengths are in
		const size_t nT = (size_t) tiles, nBlk = (nT + kTileBlock - 1) / kTileBlock;
		KRR_HIP(wf->vPlen.alloc(N));
		KRR_HIP(hipMemset(wf->vPlen.p, 0, N));
		KRR_HIP(wf->vOrder.alloc(N * sizeof(int)));
		KRR_HIP(wf->vTileKey.alloc(nT));
		KRR_HIP(wf->vTileCnt.alloc(nT * sizeof(int)));
		KRR_HIP(wf->vTileHist.alloc((kTileKeys * nBlk + 1) * sizeof(int)));
		KRR_HIP(wf->vTileSorted.alloc(nT * sizeof(int)));
		KRR_HIP(wf->vTileOff.alloc((nT + 1) * sizeof(int)));
		hipLaunchKernelGGL(k_order_identity, dim3((unsigned) ((N + 255) / 256)), dim3(256), 0, nullptr, (int) N, wf->vOrder.as<int>());
		KRR_HIP(hipGetLastError());
		KRR_HIP(hipDeviceSynchronize());
		wf->vTiles = (int) nT;
	}
	if (!wf->frameBuf.p) KRR_HIP(wf->frameBuf.alloc(sizeof(FrameDev)));
	if (wf->vL.sz < N * sizeof(float4)) {
		KRR_HIP(wf->vL.alloc(N * sizeof(float4)));
		KRR_HIP(wf->vPixel.alloc(N * sizeof(float4)));
		KRR_HIP(wf->vRng.alloc(N * sizeof(uint64_t)));
		KRR_HIP(wf->vChannel.alloc(N * sizeof(int)));
	}
	if (!wf->counters.p) {
		KRR_HIP(wf->counters.alloc(sizeof(Counters)));
		KRR_HIP(hipMemset(wf->counters.p, 0, sizeof(Counters)));
	}
	if (int rc = ensureNrcBuffers(wf, wf->vNrcQuery, N)) return rc;
	wf->views.assign(views, views + n_views);
	wf->NV = (int) total;
	wf->viewsVersion++;
	return KRR_OK;
}

int krr_wf_render_views(krr_wf *wf, uint64_t frame_index, void *hip_stream) {
	krr::clearError();
	if (!wf) return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_render_views: null handle");
	if (!wf->haveScene) return krr::fail(KRR_ERR_NO_SCENE, "krr_wf_render_views: no scene set");
	if (wf->NV == 0) return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_render_views: no views set");
	for (const krr_view &v : wf->views)
		if (v.camera.medium >= wf->ds.nMedia) return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_render_views: camera medium out of range");
	const bool media = wf->enableMedium && wf->sceneHasMedia;
	if (wf->nrcOn && media) return krr::fail(KRR_ERR_UNSUPPORTED, "nrc: needs a scene without media");
	hipStream_t st = (hipStream_t) hip_stream;
	wf->lastStream = st;
	const WfParams P = frameParams(wf, 0, 0, wf->NV, frame_index);
	ViewsLaunch L;
	if (int rc = viewsLaunch(wf, wf->NV, wf->nrcOn, L)) return rc;
	hipLaunchKernelGGL(k_set_frame, dim3(1), dim3(1), 0, st, wf->frameBuf.as<FrameDev>(), frameDev(wf, frame_index));
	const bool timing = wf->timeKernels;
	wf->spans.clear();
	wf->eventsUsed = 0;
	if (wf->useGraph && !timing) {
		// every value the recorded launches take is fixed until the views,
		// the scene, the launch shape or the overflow buffer change
		const uint64_t key = wf->viewsVersion * 0x9E3779B97F4A7C15ull ^ wf->sceneVersion * 0xC2B2AE3D27D4EB4Full ^ wf->viewsGridKey ^
							 (uint64_t) (uintptr_t) wf->ovf.p ^ ((uint64_t) wf->blockSize << 56);
		if (!wf->graphExec || key != wf->graphKey) {
			wf->dropGraph();
			if (!wf->capStream) KRR_HIP(hipStreamCreateWithFlags(&wf->capStream, hipStreamNonBlocking));
			KRR_HIP(hipStreamBeginCapture(wf->capStream, hipStreamCaptureModeThreadLocal));
			int rc		 = enqueueViews(wf, P, L, wf->capStream, false, false);
			hipGraph_t g = nullptr;
			hipError_t e = hipStreamEndCapture(wf->capStream, &g);
			if (rc || e != hipSuccess) {
				if (g) (void) hipGraphDestroy(g);
				return rc ? rc : krr::fail(KRR_ERR_HIP, "krr_wf_render_views: graph capture: %s", hipGetErrorString(e));
			}
			wf->graph = g;
			e		  = hipGraphInstantiate(&wf->graphExec, g, nullptr, nullptr, 0);
			if (e != hipSuccess) {
				wf->graphExec = nullptr;
				wf->dropGraph();
				return krr::fail(KRR_ERR_HIP, "krr_wf_render_views: graph instantiate: %s", hipGetErrorString(e));
			}
			wf->graphKey = key;
		}
		KRR_HIP(hipGraphLaunch(wf->graphExec, st));
	} else {
		const int b = timing ? wf->mark(st) : -1;
		if (int rc = enqueueViews(wf, P, L, st, timing, false)) return rc;
		if (timing) wf->spans.push_back({3, b, wf->mark(st)});
	}
	KRR_HIP(hipGetLastError());
	wf->stats.paths = (uint64_t) wf->NV * wf->spp;
	return KRR_OK;
}

int krr_wf_autotune(krr_wf *wf, int64_t trial_paths, char *report, int32_t cap, void *hip_stream) {
	krr::clearError();
	if (!wf) return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_autotune: null handle");
	if (!wf->haveScene) return krr::fail(KRR_ERR_NO_SCENE, "krr_wf_autotune: no scene set");
	const bool views = wf->NV > 0;
	if (!views && (wf->W * wf->H == 0 || !wf->haveCamera))
		return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_autotune: needs views, or a frame size and a camera from krr_wf_begin_frame");
	if (trial_paths <= 0) trial_paths = (int64_t) 1 << 21;
	hipStream_t st	= (hipStream_t) hip_stream;
	wf->lastStream	= st;
	const int total = views ? wf->NV : wf->W * wf->H;
	const int N		= (int) std::min<int64_t>(total, trial_paths);
	const bool media = wf->enableMedium && wf->sceneHasMedia;
	const int saveTb = wf->traceBlockSize, saveLds = wf->stackLdsEntries, saveRefill = wf->refillLanes;
	const float saveOcc = wf->occupancy;
	const bool saveCount = wf->countTraversal;
	wf->countTraversal	 = false;
	hipEvent_t e0 = wf->event(0), e1 = wf->event(1);
	if (!e0 || !e1) return krr::fail(KRR_ERR_HIP, "krr_wf_autotune: no events");
	struct Trial {
		int tb, lds;
		float occ;
		int refill;
		float ms;
	};
	std::vector<Trial> trials;
	int rc = KRR_OK;
	// One candidate: `runs` timed persistent launches over the first N paths
	// of the frame (the views' pixels in order, or the first N pixels of the
	// single view); the fastest run counts.
	auto measure = [&](Trial &t, int runs) -> int {
		wf->traceBlockSize	= t.tb;
		wf->stackLdsEntries = t.lds;
		wf->occupancy		= t.occ;
		wf->refillLanes		= t.refill;
		t.ms				= 1e30f;
		for (int r = 0; r < runs; r++) {
			if (views) {
				const WfParams P = frameParams(wf, 0, 0, N, 1);
				ViewsLaunch L;
				if (int e = viewsLaunch(wf, N, false, L)) return e;
				hipLaunchKernelGGL(k_set_frame, dim3(1), dim3(1), 0, st, wf->frameBuf.as<FrameDev>(), frameDev(wf, 1));
				// begin + one persistent launch, timed from its start
				hipLaunchKernelGGL(k_reset_totals, dim3(1), dim3(1), 0, st, wf->counters.as<Counters>());
				// trials may cover fewer paths than pixels: pixel order, no path-length record
				PixelStateDev tpx = viewsPx(wf);
				tpx.order		  = nullptr;
				tpx.plen		  = nullptr;
				hipLaunchKernelGGL(k_begin_views, dim3((unsigned) ((N + 255) / 256)), dim3(256), 0, st, P, viewsDev(wf), tpx);
				hipLaunchKernelGGL(k_mega_begin, dim3(1), dim3(1), 0, st, wf->counters.as<Counters>());
				KRR_HIP(hipEventRecord(e0, st));
				withMega<true>(L.media, false, false, !wf->sceneTextured, refillLanesOf(wf) > 0, [&](auto k) {
					decltype(k)::launch(L.grid, wf->traceBlockSize, L.lds, st, P, wf->ds, L.bvh, RayQueueDev{}, nullptr, tpx,
										wf->ovf.as<int>(), wf->counters.as<Counters>(), NrcDev{}, viewsDev(wf));
					return 0;
				});
			} else {
				WfParams P = frameParams(wf, wf->W, wf->H, N, wf->frameIndex);
				wf->accel.stackLds = wf->stackLdsEntries;
				const TraceBvh bvh = wf->accel.dev();
				const size_t lds   = stackLds(wf->accel, t.tb);
				const int maxBlocks = std::max(1, (N + t.tb - 1) / t.tb);
				const size_t nodeBytes = (size_t) ldsNodeCount(wf) * sizeof(Bvh4Node);
				int g = withMega<false>(media, false, false, !wf->sceneTextured, refillLanesOf(wf) > 0,
										[&](auto k) { return decltype(k)::grid(t.tb, lds, maxBlocks, nodeBytes); });
				g	  = scaledGrid(g, t.occ);
				if (int e = reserveOverflow(wf->ovf, wf->accel, (size_t) g * t.tb)) return e;
				// trials may cover fewer paths than pixels: pixel order, no miss record
				const PixelStateDev px{wf->L.as<float4>(), wf->pixel.as<float4>(), wf->rng.as<uint64_t>(), wf->channel.as<int>()};
				Counters *cnt = wf->counters.as<Counters>();
				const unsigned gN = (unsigned) ((N + 255) / 256);
				hipLaunchKernelGGL(k_reset_totals, dim3(1), dim3(1), 0, st, cnt);
				hipLaunchKernelGGL(k_begin_frame, dim3(gN), dim3(256), 0, st, P, px);
				hipLaunchKernelGGL(k_camera, dim3(gN), dim3(256), 0, st, P, cameraDev(wf->cam), px, rayQueue(wf, 0), cnt);
				hipLaunchKernelGGL(k_mega_begin, dim3(1), dim3(1), 0, st, cnt);
				KRR_HIP(hipEventRecord(e0, st));
				withMega<false>(media, false, false, !wf->sceneTextured, refillLanesOf(wf) > 0, [&](auto k) {
					decltype(k)::launch((unsigned) g, t.tb, lds, st, P, wf->ds, bvh, rayQueue(wf, 0), (const int *) &cnt->rayCount[0], px,
										wf->ovf.as<int>(), cnt, NrcDev{}, ViewsDev{});
					return 0;
				});
			}
			KRR_HIP(hipEventRecord(e1, st));
			KRR_HIP(hipGetLastError());
			KRR_HIP(hipEventSynchronize(e1));
			float ms = 0.f;
			KRR_HIP(hipEventElapsedTime(&ms, e0, e1));
			if (r > 0 || runs == 1) t.ms = std::min(t.ms, ms); // the first run warms caches
		}
		return KRR_OK;
	};
	const int tbs[]		= {64, 128, 256};
	const int ldss[]	= {8, 16, 24};
	const float occs[]	= {1.f, 0.75f, 0.5f};
	// the first failing candidate ends the search (its error is returned);
	// the launch shapes at the refill threshold in force, then, when
	// refill_lanes is left to the scene (-1), other thresholds at the best
	// shape (scenes without media; the launch without refill wins on small
	// trees, whose traversals are short)
	const int refill0 = refillLanesOf(wf);
	for (int ti = 0; ti < 3 && rc == KRR_OK; ti++)
		for (int li = 0; li < 3 && rc == KRR_OK; li++)
			for (int oi = 0; oi < 3 && rc == KRR_OK; oi++) {
				Trial t{tbs[ti], ldss[li], occs[oi], refill0, 0.f};
				rc = measure(t, 3);
				if (rc == KRR_OK) trials.push_back(t);
			}
	if (rc == KRR_OK && !media && saveRefill < 0) {
		const Trial shape = *std::min_element(trials.begin(), trials.end(), [](const Trial &a, const Trial &b) { return a.ms < b.ms; });
		for (const int r : {0, 4, kDefaultRefillLanes, 12, 16}) {
			if (r == refill0 || rc != KRR_OK) continue;
			Trial t{shape.tb, shape.lds, shape.occ, r, 0.f};
			rc = measure(t, 3);
			if (rc == KRR_OK) trials.push_back(t);
		}
	}
	wf->countTraversal = saveCount;
	if (rc) {
		wf->traceBlockSize	= saveTb;
		wf->stackLdsEntries = saveLds;
		wf->occupancy		= saveOcc;
		wf->refillLanes		= saveRefill;
		return rc;
	}
	const Trial best = *std::min_element(trials.begin(), trials.end(), [](const Trial &a, const Trial &b) { return a.ms < b.ms; });
	wf->traceBlockSize	= best.tb;
	wf->stackLdsEntries = best.lds;
	wf->occupancy		= best.occ;
	wf->refillLanes		= best.refill;
	wf->gridKey = wf->viewsGridKey = ~0ull; // grids and the recorded frame follow the new shape
	wf->dropGraph();
	if (!views) wf->beginPending = true; // the trials used the frame's pixel state: seed it again
	Json j = Json::object(), arr = Json::array();
	for (const Trial &t : trials) {
		Json o = Json::object();
		o.set("trace_block_size", Json((double) t.tb));
		o.set("stack_lds_entries", Json((double) t.lds));
		o.set("occupancy", Json((double) t.occ));
		o.set("refill_lanes", Json((double) t.refill));
		o.set("ms", Json((double) t.ms));
		arr.push(o);
	}
	Json b = Json::object();
	b.set("trace_block_size", Json((double) best.tb));
	b.set("stack_lds_entries", Json((double) best.lds));
	b.set("occupancy", Json((double) best.occ));
	b.set("refill_lanes", Json((double) best.refill));
	b.set("ms", Json((double) best.ms));
	j.set("trial_paths", Json((double) N));
	j.set("candidates", arr);
	j.set("best", b);
	const std::string out = j.dump();
	if (report && cap > 0) {
		const size_t n = std::min((size_t) cap - 1, out.size());
		std::memcpy(report, out.data(), n);
		report[n] = 0;
	}
	return (int) out.size();
}

int krr_wf_stats(krr_wf *wf, krr_stats *out) {
	krr::clearError();
	if (!wf || !out) return krr::fail(KRR_ERR_INVALID_ARG, "krr_wf_stats: null argument");
	std::memset(out, 0, sizeof(*out));
	out->bvh_build_ms = wf->accel.buildSeconds * 1e3;
	if (!wf->counters.p) return KRR_OK;
	KRR_HIP(hipStreamSynchronize(wf->lastStream));
	Counters c;
	KRR_HIP(hipMemcpy(&c, wf->counters.p, sizeof(c), hipMemcpyDeviceToHost));
	out->closest_rays	  = c.closestTotal;
	out->shadow_rays	  = c.shadowTotal;
	out->paths			  = wf->stats.paths;
	out->node_visits	  = c.nodeVisits;
	out->tri_tests		  = c.triTests;
	out->closest_launches	= c.closestLaunches;
	out->shadow_node_visits = c.shadowNodeVisits;
	out->shadow_tri_tests	= c.shadowTriTests;
	out->megakernel_closest_rays = c.megaClosest;
	out->megakernel_shadow_rays	 = c.megaShadow;
	out->megakernel_node_visits = c.megaNodeVisits;
	out->megakernel_tri_tests	 = c.megaTriTests;
	if (wf->timeKernels) {
		double ms[6] = {0, 0, 0, 0, 0, 0};
		for (const auto &sp : wf->spans) {
			float t = 0.f;
			hipEvent_t a = wf->event(sp.begin), b = wf->event(sp.end);
			if (!a || !b) continue;
			KRR_HIP(hipEventSynchronize(b));
			KRR_HIP(hipEventElapsedTime(&t, a, b));
			ms[sp.kind] += t;
		}
		out->closest_trace_ms = ms[0];
		out->shade_ms		  = ms[1];
		out->shadow_trace_ms  = ms[2];
		out->render_ms		  = ms[3];
		out->megakernel_ms	  = ms[4];
		out->nrc_ms			  = ms[5];
	}
	if (wf->nrcOn) {
		out->nrc_queries		= c.nrcQueriesTotal;
		out->nrc_train_vertices = (uint64_t) c.nrcTrain;
		out->nrc_loss			= nrcLastLoss(wf->nrc, wf->lastStream);
		out->nrc_steps			= (uint64_t) krr_nrc_steps(wf->nrc);
	}
	return KRR_OK;
}

int krr_wf_destroy(krr_wf *wf) {
	if (wf) {
		if (deviceAvailable()) (void) hipDeviceSynchronize();
		delete wf;
	}
	return KRR_OK;
}

int krr_error_metric(int32_t metric, const float *img_dev, const float *ref_dev, int64_t n_pixels, double *out, void *hip_stream) {
	krr::clearError();
	if (!out || n_pixels < 0 || (n_pixels > 0 && (!img_dev || !ref_dev))) return krr::fail(KRR_ERR_INVALID_ARG, "krr_error_metric: bad argument");
	if (metric < KRR_METRIC_MSE || metric > KRR_METRIC_RELMSE) return krr::fail(KRR_ERR_INVALID_ARG, "krr_error_metric: unknown metric %d", metric);
	*out = 0.0;
	if (n_pixels == 0) return KRR_OK;
	if (!deviceAvailable()) return krr::fail(KRR_ERR_NO_DEVICE, "krr_error_metric: no HIP device");
	hipStream_t st = (hipStream_t) hip_stream;
	double *acc	   = nullptr;
	KRR_HIP(hipMalloc(&acc, sizeof(double)));
	hipLaunchKernelGGL(k_zero_double, dim3(1), dim3(1), 0, st, acc);
	const unsigned grid = (unsigned) std::min<int64_t>(1024, (n_pixels + 255) / 256);
	hipLaunchKernelGGL(k_metric, dim3(grid), dim3(256), 0, st, metric, (const float4 *) img_dev, (const float4 *) ref_dev, n_pixels, acc);
	hipError_t e = hipGetLastError();
	double sum	 = 0;
	if (e == hipSuccess) e = hipStreamSynchronize(st);
	if (e == hipSuccess) e = hipMemcpy(&sum, acc, sizeof(double), hipMemcpyDeviceToHost);
	(void) hipFree(acc);
	if (e != hipSuccess) return krr::fail(KRR_ERR_HIP, "krr_error_metric: %s", hipGetErrorString(e));
	*out = sum / (double) n_pixels;
	return KRR_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// MegakernelPathTracer (src/render/megakernel/pathtracer.{h,cpp}): the scene
// and acceleration structure live in an inner wavefront pass that never
// allocates frame state; the kernel is in pathtracer.hip.
// ---------------------------------------------------------------------------
struct krr_mk {
	krr_wf *wf = nullptr;
	bool nee = true, timeKernels = false;
	int maxDepth = 10, spp = 1, blockSize = 256;
	float rr = 0.8f, occupancy = 1.f;
	int W = 0, H = 0;
	krr_camera cam{};
	bool haveCamera	 = false;
	uint64_t frameIndex = 0;
	DevBuf rays, ovf;
	int grid		 = 0;
	uint64_t gridKey = ~0ull;
	hipStream_t lastStream = nullptr;
	hipEvent_t ev[2]	   = {nullptr, nullptr};
	bool timed			   = false;
	uint64_t paths		   = 0;
	~krr_mk() {
		for (hipEvent_t e : ev)
			if (e) (void) hipEventDestroy(e);
		if (wf) krr_wf_destroy(wf);
	}
};

namespace {

int parseMkParams(krr_mk *mk, const char *json, std::string &inner) {
	Json in = Json::object();
	if (json && *json) {
		try {
			Json j = Json::parse(json);
			if (!j.isObject()) return krr::fail(KRR_ERR_BAD_CONFIG, "pass params must be a JSON object");
			mk->nee			= j.value("nee", true);
			mk->maxDepth	= (int) j.value("max_depth", 10.0);
			mk->rr			= (float) j.value("rr", 0.8);
			mk->spp			= (int) j.value("spp", 1.0);
			mk->blockSize	= (int) j.value("block_size", 256.0);
			mk->occupancy	= (float) j.value("occupancy", 1.0);
			mk->timeKernels = j.value("time_kernels", false);
			for (const char *k : {"leaf_size", "sah_node_cost", "stack_lds_entries", "bvh_builder", "node_format"})
				if (j.contains(k)) in.set(k, j.at(k));
		} catch (const std::exception &e) {
			return krr::fail(KRR_ERR_BAD_CONFIG, "pass params: %s", e.what());
		}
	}
	if (mk->maxDepth < 0 || mk->maxDepth > 250) return krr::fail(KRR_ERR_BAD_CONFIG, "max_depth must be in [0, 250]");
	if (!(mk->rr > 0.f && mk->rr <= 1.f)) return krr::fail(KRR_ERR_BAD_CONFIG, "rr must be in (0, 1]");
	if (mk->spp < 1 || mk->spp > 1 << 20) return krr::fail(KRR_ERR_BAD_CONFIG, "spp must be >= 1");
	if (mk->blockSize < 64 || mk->blockSize > 256 || mk->blockSize % 64) return krr::fail(KRR_ERR_BAD_CONFIG, "block_size must be 64, 128, 192 or 256");
	if (!(mk->occupancy > 0.f && mk->occupancy <= 1.f)) return krr::fail(KRR_ERR_BAD_CONFIG, "occupancy must be in (0, 1]");
	inner = in.dump();
	return KRR_OK;
}

} // namespace

extern "C" {

int krr_mk_create(const char *json_params, krr_mk **out) {
	krr::clearError();
	if (!out) return krr::fail(KRR_ERR_INVALID_ARG, "krr_mk_create: null out");
	*out	= nullptr;
	auto mk = std::make_unique<krr_mk>();
	std::string inner;
	if (int rc = parseMkParams(mk.get(), json_params, inner)) return rc;
	if (int rc = krr_wf_create(inner.c_str(), &mk->wf)) return rc;
	*out = mk.release();
	return KRR_OK;
}

int krr_mk_get_params(const krr_mk *mk, char *buf, int32_t cap) {
	if (!mk) return krr::fail(KRR_ERR_INVALID_ARG, "krr_mk_get_params: null handle");
	Json j = Json::object();
	j.set("nee", Json(mk->nee));
	j.set("max_depth", Json((double) mk->maxDepth));
	j.set("rr", Json((double) mk->rr));
	j.set("spp", Json((double) mk->spp));
	j.set("block_size", Json((double) mk->blockSize));
	j.set("occupancy", Json((double) mk->occupancy));
	j.set("time_kernels", Json(mk->timeKernels));
	j.set("leaf_size", Json((double) mk->wf->leafSize));
	j.set("sah_node_cost", Json((double) mk->wf->sahNodeCost));
	j.set("stack_lds_entries", Json((double) mk->wf->stackLdsEntries));
	j.set("bvh_builder", Json(std::string(mk->wf->deviceBuilder ? "device" : "host")));
	j.set("node_format", Json(std::string(mk->wf->quantizedNodes ? "quantized" : "float")));
	std::string s = j.dump();
	if (buf && cap > 0) {
		size_t n = std::min((size_t) cap - 1, s.size());
		std::memcpy(buf, s.data(), n);
		buf[n] = 0;
	}
	return (int) s.size();
}

int krr_mk_set_scene(krr_mk *mk, const krr_scene_desc *scene) {
	krr::clearError();
	if (!mk || !scene) return krr::fail(KRR_ERR_INVALID_ARG, "krr_mk_set_scene: null argument");
	if (int rc = validateSceneDesc(scene)) return rc;
	for (int i = 0; i < scene->n_meshes; i++)
		if (scene->meshes[i].material < 0)
			return krr::fail(KRR_ERR_UNSUPPORTED, "krr_mk_set_scene: mesh %d has no material (the megakernel path tracer renders surfaces only)", i);
	mk->gridKey = ~0ull;
	return krr_wf_set_scene(mk->wf, scene);
}

int krr_mk_resize(krr_mk *mk, int32_t width, int32_t height) {
	krr::clearError();
	if (!mk || width < 0 || height < 0) return krr::fail(KRR_ERR_INVALID_ARG, "krr_mk_resize: bad argument");
	if ((int64_t) width * height > (int64_t) 1 << 26) return krr::fail(KRR_ERR_CAPACITY, "krr_mk_resize: at most 2^26 pixels per frame");
	if (width > 65536 || height > 65536) return krr::fail(KRR_ERR_CAPACITY, "krr_mk_resize: width and height must be <= 65536");
	mk->W = width;
	mk->H = height;
	return KRR_OK;
}

int krr_mk_begin_frame(krr_mk *mk, const krr_camera *camera, uint64_t frame_index) {
	krr::clearError();
	if (!mk || !camera) return krr::fail(KRR_ERR_INVALID_ARG, "krr_mk_begin_frame: null argument");
	mk->cam		   = *camera;
	mk->frameIndex = frame_index;
	mk->haveCamera = true;
	return KRR_OK;
}

int krr_mk_render(krr_mk *mk, float *film_dev, void *hip_stream) {
	krr::clearError();
	if (!mk) return krr::fail(KRR_ERR_INVALID_ARG, "krr_mk_render: null handle");
	krr_wf *wf = mk->wf;
	if (!wf->haveScene) return krr::fail(KRR_ERR_NO_SCENE, "krr_mk_render: no scene set");
	if (mk->W == 0 || mk->H == 0) return KRR_OK; // pathtracer.cpp:53 early return
	if (!film_dev) return krr::fail(KRR_ERR_INVALID_ARG, "krr_mk_render: null film");
	if (!mk->haveCamera) return krr::fail(KRR_ERR_INVALID_ARG, "krr_mk_render: no camera (call krr_mk_begin_frame)");
	hipStream_t st	   = (hipStream_t) hip_stream;
	mk->lastStream	   = st;
	wf->accel.stackLds = wf->stackLdsEntries;
	const TraceBvh bvh = wf->accel.dev();
	const int tb	   = mk->blockSize;
	const size_t lds   = stackLds(wf->accel, tb);
	const int tilesX = (mk->W + 7) / 8, tilesY = (mk->H + 7) / 8;
	const int nTiles = tilesX * tilesY;
	const int needed = (nTiles + tb / 64 - 1) / (tb / 64);
	const uint64_t key = ((uint64_t) nTiles << 20) ^ ((uint64_t) tb << 8) ^ ((uint64_t) wf->accel.stackSize << 44) ^
						 ((uint64_t) wf->stackLdsEntries << 36) ^ ((uint64_t) (mk->occupancy * 1024.f) << 52) ^ (wf->sceneTextured ? 1ull : 0ull);
	if (key != mk->gridKey) {
		int dev = 0, cus = 0;
		KRR_HIP(hipGetDevice(&dev));
		KRR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
		const int resident = scaledGrid(pathTracerBlocksPerCU(tb, lds, wf->sceneTextured) * std::max(cus, 1), mk->occupancy);
		mk->grid		   = std::max(1, std::min(needed, resident));
		if (int rc = reserveOverflow(mk->ovf, wf->accel, (size_t) mk->grid * tb)) return rc;
		if (!mk->rays.p) KRR_HIP(mk->rays.alloc(2 * sizeof(unsigned long long)));
		mk->gridKey = key;
	}
	if ((size_t) (bvh.stackLds + 1) * tb * sizeof(int) != lds) return krr::fail(KRR_ERR_INVALID_ARG, "krr_mk_render: stack layout mismatch");
	PtArgs a;
	a.sc		  = wf->ds;
	a.bvh		  = bvh;
	a.cam		  = cameraDev(mk->cam);
	a.W			  = mk->W;
	a.H			  = mk->H;
	a.spp		  = mk->spp;
	a.maxDepth	  = mk->maxDepth;
	a.nee		  = mk->nee ? 1 : 0;
	a.rr		  = mk->rr;
	a.sampleIndex = (uint32_t) mk->frameIndex * 512u;
	a.film		  = (float4 *) film_dev;
	a.ovf		  = mk->ovf.as<int>();
	a.rays		  = mk->rays.as<unsigned long long>();
	a.tilesX	  = tilesX;
	a.nTiles	  = nTiles;
	a.textured	  = wf->sceneTextured;
	mk->timed	  = mk->timeKernels;
	if (mk->timed) {
		for (hipEvent_t &e : mk->ev)
			if (!e) KRR_HIP(hipEventCreate(&e));
		KRR_HIP(hipEventRecord(mk->ev[0], st));
	}
	KRR_HIP(launchPathTracer(a, mk->grid, tb, lds, st));
	if (mk->timed) KRR_HIP(hipEventRecord(mk->ev[1], st));
	mk->paths = (uint64_t) mk->W * mk->H * mk->spp;
	return KRR_OK;
}

int krr_mk_stats(krr_mk *mk, krr_stats *out) {
	krr::clearError();
	if (!mk || !out) return krr::fail(KRR_ERR_INVALID_ARG, "krr_mk_stats: null argument");
	std::memset(out, 0, sizeof(*out));
	out->bvh_build_ms = mk->wf->accel.buildSeconds * 1e3;
	if (!mk->rays.p) return KRR_OK;
	KRR_HIP(hipStreamSynchronize(mk->lastStream));
	unsigned long long r[2] = {0, 0};
	KRR_HIP(hipMemcpy(r, mk->rays.p, sizeof(r), hipMemcpyDeviceToHost));
	out->closest_rays = r[0];
	out->shadow_rays  = r[1];
	out->paths		  = mk->paths;
	if (mk->timed) {
		float t = 0.f;
		KRR_HIP(hipEventSynchronize(mk->ev[1]));
		KRR_HIP(hipEventElapsedTime(&t, mk->ev[0], mk->ev[1]));
		out->render_ms = t;
	}
	return KRR_OK;
}

int krr_mk_destroy(krr_mk *mk) {
	if (mk) {
		if (deviceAvailable()) (void) hipDeviceSynchronize();
		delete mk;
	}
	return KRR_OK;
}

} // extern "C"
