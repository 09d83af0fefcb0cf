/*
 * oracle.h - CPU restatement of KiRaRay's WavefrontPathTracer (test only).
 *
 * TEST INFRASTRUCTURE.  This library is the checker for the HIP tracer in
 * kiraray_amd/: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it.  The product never links or calls it.
 *
 * It restates, stage by stage and with the same RNG consumption order, the
 * reference integrator (src/render/wavefront/integrator.cpp:78-267,
 * device.cu:43-128, medium.cpp:13-153, wavefront.h:34-139) and the shading
 * code it calls, in FP32 (parity mode: decisions identical to the GPU) or
 * FP64 (reference-image mode).  Every pixel runs its path sequentially; the
 * stage order inside one loop depth is the reference's launch order, which is
 * the only order that matters because each pixel owns at most one item per
 * queue per stage (src/render/wavefront/workqueue.h:23-25).
 */
#ifndef KRR_ORACLE_H
#define KRR_ORACLE_H

#include <stdint.h>
#include "krr_wavefront.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_params {
	int32_t nee;				  /* default 1 */
	int32_t enable_medium;		  /* default 1 */
	int32_t max_depth;			  /* default 10 */
	float rr;					  /* default 0.8 */
	int32_t rr_min_depth;		  /* default 0 */
	int32_t enable_clamp;		  /* default 0 */
	float clamp_max;			  /* default 1e3 */
	int32_t spp;				  /* default 1 */
	int32_t legacy_spp_accumulate; /* default 0 */
	int32_t precision;			  /* 32 or 64 */
	int32_t n_threads;			  /* 0 = all OpenMP threads */
	int32_t row_begin, row_end, row_step; /* rendered rows; row_end 0 = height */
	int32_t pixel_step;			  /* render every pixel_step-th pixel of a row (1 = all) */
	int32_t view_id;			  /* multi-view seeding: bits 32..63 of the PCG initstate (0 = reference) */
} oracle_params;

typedef struct oracle_stats {
	uint64_t closest_rays;
	uint64_t shadow_rays;
	uint64_t paths;
	double seconds;
	int32_t threads;
} oracle_stats;

void oracle_default_params(oracle_params *p);

/* Renders one frame (beginFrame + render) into film (float4 W x H, row H-1-y),
 * pixels outside the rendered subset are left untouched. */
int oracle_render(const krr_scene_desc *scene, const krr_camera *camera, const oracle_params *p,
				  uint64_t frame_index, int32_t width, int32_t height, float *film,
				  oracle_stats *stats);

/* MegakernelPathTracer (src/render/megakernel/device.cu:50-195): one loop
 * per sample, power-heuristic MIS, Russian roulette from the first vertex,
 * no media; RNG seeded setPixelSample(pixel, frame_index * 512).  Uses nee,
 * max_depth, rr, spp, precision, n_threads and the row/pixel subset of p;
 * the film holds the mean of the spp samples.  Meshes without material are
 * rejected (KRR_ERR_UNSUPPORTED). */
int oracle_render_megakernel(const krr_scene_desc *scene, const krr_camera *camera, const oracle_params *p,
							 uint64_t frame_index, int32_t width, int32_t height, float *film,
							 oracle_stats *stats);

/* Per-pixel ray counts of the last oracle_render call on this thread are not
 * kept; use the totals in oracle_stats. */

/* PCG32 (src/core/sampler.h:13-91): n outputs of nextUint after setSeed. */
void oracle_pcg32_sequence(uint64_t initstate, uint64_t initseq, int32_t n, uint32_t *out);
/* State after setSeed(initstate, initseq) then advance(delta). */
void oracle_pcg32_advance(uint64_t initstate, uint64_t initseq, int64_t delta, uint64_t *state_out,
						  uint64_t *inc_out);
/* get1D() floats after setPixelSample((x,y), sample_index) + advance(adv). */
void oracle_pixel_samples(int32_t x, int32_t y, uint32_t sample_index, int64_t adv, int32_t n,
						  float *out);
uint32_t oracle_interleave_32bit(uint32_t x, uint32_t y);

/* Brute-force closest hit over every triangle (rays: n x 8 floats
 * {ox,oy,oz,tmax,dx,dy,dz,0}); also the oracle BVH when use_bvh != 0. */
int oracle_trace_closest(const krr_scene_desc *scene, const float *rays, int64_t n, krr_hit *hits,
						 int32_t use_bvh);
int oracle_trace_any(const krr_scene_desc *scene, const float *rays, int64_t n, int32_t *visible,
					 int32_t use_bvh);

/* Camera ray of pixel (x,y) for a given CameraSample (pFilm, pLens, time). */
void oracle_camera_ray(const krr_camera *camera, int32_t x, int32_t y, int32_t width, int32_t height,
					   const float sample[5], float origin[3], float dir[3]);

/* BSDF access for invariant tests.  wo, wi are in the local shading frame
 * (z = normal).  material follows krr_material.  u: up to 3 uniform numbers
 * consumed in the order the BSDF draws them. */
void oracle_bsdf_eval(const krr_material *m, const float wo[3], const float wi[3], float f[3],
					  float *pdf);
/* returns flags (BSDFType bits), writes f, wi, pdf; pdf == 0 means no sample */
int32_t oracle_bsdf_sample(const krr_material *m, const float wo[3], const float u[3], float f[3],
						   float wi[3], float *pdf);
int32_t oracle_bsdf_type(const krr_material *m);
float oracle_fresnel_dielectric(float cos_theta_i, float eta);
void oracle_fresnel_complex(float cos_theta_i, const float eta[3], const float k[3], float out[3]);
/* GGX distribution: D(wh), G1(w), Pdf(wo,wh), Sample(wo,u) (legacy slope sampling). */
float oracle_ggx_D(float ax, float ay, const float wh[3]);
float oracle_ggx_G1(float ax, float ay, const float w[3]);
float oracle_ggx_pdf(float ax, float ay, const float wo[3], const float wh[3]);
void oracle_ggx_sample(float ax, float ay, const float wo[3], const float u[2], float wh[3]);
/* Henyey-Greenstein p(wo,wi) and sample(wo,u) -> wi, pdf. */
float oracle_hg_p(float g, const float wo[3], const float wi[3]);
void oracle_hg_sample(float g, const float wo[3], const float u[2], float wi[3], float *pdf);
/* sampleDiscrete over 3 weights. */
int32_t oracle_sample_discrete3(const float w[3], float u, float *pmf);

/* Majorant grid (64^3 cells) and trilinear density of medium m. */
int oracle_majorant_grid(const krr_medium *m, float *out /* 64^3 */);
float oracle_density(const krr_medium *m, const float p[3]);
/* Majorant segments of the ray o + t d, t in [0, tmax] (sampleRay +
 * MajorantIterator::next, media.h:41-106): returns their number and writes
 * at most cap of them as {tMin, tMax, sigma_maj r, g, b} into seg. */
int32_t oracle_majorant_segments(const krr_medium *m, const float o[3], const float d[3], float tmax, int32_t precision,
								 int32_t cap, double *seg);
/* traceTransmittance (wavefront.h:80-139) of one shadow ray {o, d, tmax,
 * medium} with Ld = 1 and MIS ratios (sr_pu, sr_pl), run n times from one
 * PCG stream (seed): the mean (and mean square) of the radiance it adds,
 * per channel, for the given hero channel. */
int oracle_transmittance(const krr_scene_desc *scene, const float o[3], const float d[3], float tmax, int32_t medium,
						 int32_t channel, float sr_pu, float sr_pl, uint64_t seed, int64_t n, int32_t precision, double out[3],
						 double out_sq[3]);

/* Deterministic elementary functions used on the shading path (checked
 * against libm double in the tests). */
float oracle_dm_sin(float x);
float oracle_dm_cos(float x);
float oracle_dm_log(float x);
float oracle_dm_exp(float x);
float oracle_dm_tan(float x);
float oracle_dm_atan2(float y, float x);

/* relMSE etc. (metrics.cu:13-128) over float4 images, FP64 accumulation. */
double oracle_error_metric(int32_t metric, const float *img, const float *ref, int64_t n_pixels);

#ifdef __cplusplus
}
#endif

#endif
