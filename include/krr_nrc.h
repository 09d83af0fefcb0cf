/*
 * krr_nrc.h - C ABI of the neural radiance cache (gfx950, BF16 MFMA).
 *
 * No counterpart in KiRaRay: the reference has no learned component on this
 * path.  The cache follows Mueller et al., "Real-time Neural Radiance Caching
 * for Path Tracing" (SIGGRAPH 2021): a 5 x 64 ReLU MLP without biases that
 * maps an encoded path vertex (position, outgoing direction, normal,
 * roughness, diffuse and specular albedo) to the RGB radiance leaving it,
 * trained online with Adam on a relative L2 loss.
 *
 * Weights and activations are BF16 on the matrix cores with FP32
 * accumulation; the optimiser keeps FP32 master weights.
 *
 * Vertex records (krr_nrc_vertex) are 16 floats in device memory.  All
 * functions return 0 or a negative krr_status (krr_wavefront.h) with
 * krr_last_error().  Calls on one handle must not overlap.
 */
#ifndef KRR_NRC_H
#define KRR_NRC_H

#include <stdint.h>

#include "krr_wavefront.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Raw path-vertex features, 64 bytes. */
typedef struct krr_nrc_vertex {
	float pos[3];		/* position in the scene box, mapped to [0,1]^3 */
	float dir[2];		/* outgoing direction wo (towards the previous vertex): theta/pi, phi/(2 pi) */
	float normal[2];	/* shading normal: theta/pi, phi/(2 pi) */
	float roughness;	/* 0 = smooth */
	float diffuse[3];	/* diffuse albedo */
	float specular[3];	/* specular albedo */
	float pad[2];
} krr_nrc_vertex;

typedef struct krr_nrc krr_nrc;

/* Network geometry (fixed by this build). */
#define KRR_NRC_WIDTH 64
#define KRR_NRC_HIDDEN_LAYERS 5
#define KRR_NRC_OUTPUTS 3
/* FP32 parameter count: 5 matrices 64x64 (input encoding and hidden layers)
 * plus the 64x3 output matrix = 20672. */
#define KRR_NRC_PARAMS (KRR_NRC_HIDDEN_LAYERS * KRR_NRC_WIDTH * KRR_NRC_WIDTH + KRR_NRC_WIDTH * KRR_NRC_OUTPUTS)

/* json: {"lr": 1e-3, "beta1": 0.9, "beta2": 0.99, "eps": 1e-8,
 *        "loss_eps": 0.01, "seed": 1, "max_batch": 65536}
 * The weights start from a seeded Xavier-uniform draw. */
int krr_nrc_create(const char *json, krr_nrc **out);
int krr_nrc_destroy(krr_nrc *nrc);

/* FP32 master weights, KRR_NRC_PARAMS floats in host memory, layer after
 * layer, each layer row-major [in][out] (in = 64; out = 64, last layer 3). */
int krr_nrc_get_weights(const krr_nrc *nrc, float *host_out);
/* Sets the master weights (and their BF16 copies); resets Adam. */
int krr_nrc_set_weights(krr_nrc *nrc, const float *host_in);

/* The 64-float encoding of n vertex records (tests and tools; the
 * inference and training kernels encode on the fly). */
int krr_nrc_encode(const krr_nrc *nrc, const krr_nrc_vertex *vertices_dev, int64_t n, float *encoded_dev,
				   void *hip_stream);

/* radiance_dev[3i..3i+2] = cache(vertices_dev[i]). */
int krr_nrc_infer(krr_nrc *nrc, const krr_nrc_vertex *vertices_dev, int64_t n, float *radiance_dev, void *hip_stream);

/* Gradient of the mean relative L2 loss over the batch,
 *   L = 1/(3n) sum_i sum_c (y_ic - t_ic)^2 / (sg(y_ic)^2 + loss_eps),
 * with respect to the weights (KRR_NRC_PARAMS floats, same order as the
 * weights), into grad_dev; *loss_out gets L.  n <= max_batch. */
int krr_nrc_gradient(krr_nrc *nrc, const krr_nrc_vertex *vertices_dev, const float *targets_dev, int64_t n,
					 float *grad_dev, float *loss_out, void *hip_stream);

/* One Adam step on the batch: gradient as above, then the update.  *loss_out
 * (may be NULL) gets the loss before the step. */
int krr_nrc_train_step(krr_nrc *nrc, const krr_nrc_vertex *vertices_dev, const float *targets_dev, int64_t n,
					   float *loss_out, void *hip_stream);

/* Adam steps taken since creation or the last krr_nrc_set_weights (the count
 * is kept in device memory; this call waits for the device). */
int64_t krr_nrc_steps(const krr_nrc *nrc);

#ifdef __cplusplus
}
#endif

#endif /* KRR_NRC_H */
