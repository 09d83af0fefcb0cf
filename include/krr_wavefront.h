/*
 * krr_wavefront.h - C ABI of the MI355X (gfx950) wavefront path tracer.
 *
 * This is the drop-in boundary for KiRaRay's WavefrontPathTracer render pass.
 * Every entry point is a plain C function over plain pointers and sizes; the
 * library owns every device allocation it makes, the caller owns the film
 * buffer and the HIP stream it passes in.  Calls on one handle are not
 * thread-safe; separate handles are independent.
 *
 * All functions return 0 (KRR_OK) on success and a negative KRR_ERR_* code on
 * failure; krr_last_error() then describes the failure (thread-local).  The
 * reference logs and exits on every error (src/util/check.h:8-42,
 * src/core/logger.cpp:73-100); this ABI reports instead.
 */
#ifndef KRR_WAVEFRONT_H
#define KRR_WAVEFRONT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KRR_ABI_VERSION 2

enum krr_status {
	KRR_OK				  = 0,
	KRR_ERR_INVALID_ARG	  = -1, /* null handle/pointer, negative size, malformed data */
	KRR_ERR_BAD_CONFIG	  = -2, /* JSON that does not parse or has a wrongly typed key */
	KRR_ERR_NO_SCENE	  = -3, /* render hook called before set_scene */
	KRR_ERR_HIP			  = -4, /* a HIP runtime call failed */
	KRR_ERR_OUT_OF_MEMORY = -5,
	KRR_ERR_CAPACITY	  = -6, /* a size exceeds a documented maximum */
	KRR_ERR_UNSUPPORTED	  = -7, /* valid input the library does not implement */
	KRR_ERR_NO_DEVICE	  = -8, /* no HIP device is visible */
};

/* Material kinds: same numbering as the reference MaterialType
 * (src/core/raytracing.h:24-31).  Disney is not implemented. */
enum krr_material_type {
	KRR_MAT_NULL	   = 0,
	KRR_MAT_DIFFUSE	   = 1,
	KRR_MAT_DIELECTRIC = 2,
	KRR_MAT_CONDUCTOR  = 3,
};

/* Material::MaterialParams under the SpecularGlossiness shading model that the
 * JSON and OBJ importers select (src/scene/krrscene.cpp:204-228,
 * assimp.cpp:222-224): roughness = 1 - specular[3].  Textures
 * (prepareSurfaceInteraction, src/render/shading.h:187-226) replace the
 * parameter they stand for at the hit's texture coordinates: the diffuse
 * texture's rgb the diffuse colour, the specular texture's alpha the
 * glossiness (roughness = 1 - a).  The transmission texture (or, without
 * one, a non-zero transmission constant) gives the material the any-hit
 * alpha test of shading.h:89-113: a hit is ignored with probability
 * 1 - alpha, alpha = 1 - luminance(transmission), decided by a hash of the
 * ray (pbrt.cpp:145-156 stores 1 - opacity there). */
typedef struct krr_material {
	int32_t type;			   /* krr_material_type */
	float diffuse[3];		   /* "diffuse", default (1,1,1) */
	float specular[4];		   /* "specular" rgb, a = 1 - "roughness" */
	float ior;				   /* "eta", default 1.5 */
	float specular_transmission; /* "specular_transmission", default 0 */
	float anisotropic;		   /* "anisotropic", default 0 */
	int32_t diffuse_texture;   /* 0 = none, else 1 + index into krr_scene_desc.textures */
	int32_t specular_texture;  /* 0 = none, else 1 + index into krr_scene_desc.textures */
	int32_t transmission_texture; /* 0 = none, else 1 + index into krr_scene_desc.textures */
	float transmission[3];		  /* constant transmission when there is no texture; 0 = opaque */
	int32_t emissive_texture;	  /* 0 = none, else 1 + index into krr_scene_desc.textures: an emissive
									 mesh of this material emits max(emission) * rgb(uv) (light.h:188-193) */
	int32_t normal_texture;		  /* 0 = none, else 1 + index into krr_scene_desc.textures: tangent-space
									 normal map, n = normalize(T x + B y + N z) with (x, y, z) = 2 rgb(uv) - 1,
									 on meshes with texture coordinates (shading.h:187-203) */
} krr_material;

/* An RGBA float32 image used as a normalised-coordinate texture: height rows
 * of width texels, row 0 at v = 0, sampled bilinearly with wrap-around
 * (texture.cpp:236-243). */
typedef struct krr_texture {
	int32_t width, height;
	const float *rgba; /* host memory, copied by krr_wf_set_scene */
} krr_texture;

/* One mesh instance with identity transform: a contiguous triangle range of
 * the scene arrays (rt::MeshData + rt::InstanceData, src/core/mesh.h:25-56). */
typedef struct krr_mesh {
	int32_t first_triangle;
	int32_t n_triangles;
	int32_t material;	   /* index into materials; -1 = null material (medium interface) */
	int32_t medium_inside;  /* index into media or -1 */
	int32_t medium_outside; /* index into media or -1 */
	int32_t has_normals;	   /* per-vertex normals in krr_scene_desc.normals are valid */
	float emission[3];	   /* Le; any component > 0 makes every triangle a DiffuseAreaLight (scale = max(Le));
							  with the material's emissive_texture the light emits scale * rgb(uv) */
	int32_t has_texcoords;  /* per-vertex texture coordinates in krr_scene_desc.texcoords are valid (else uv = 0) */
	int32_t has_tangents;	/* per-vertex tangents in krr_scene_desc.tangents are valid: the shading frame's
							   tangent is the interpolated one, made orthogonal to n (shading.h:140-145);
							   else getPerpendicular(n) */
} krr_mesh;

enum krr_medium_type {
	KRR_MEDIUM_HOMOGENEOUS = 0, /* HomogeneousMedium (src/render/media.h:108-142) */
	KRR_MEDIUM_GRID		   = 1, /* dense fp32 density grid, stands for NanoVDBMedium<float> */
};

/* Participating medium.  For KRR_MEDIUM_GRID the density value of voxel
 * (i,j,k) sits at index-space point (i,j,k); world p maps to index space as
 * (p - bounds_min) / (bounds_max - bounds_min) * res; lookups are trilinear
 * with zero outside the grid (the NanoVDB SampleFromVoxels<...,1,false>
 * convention used by src/util/volume.h:62,82-86). */
typedef struct krr_medium {
	int32_t type;
	float sigma_t[3];
	float albedo[3];
	float Le[3];		   /* homogeneous emission */
	float g;			   /* Henyey-Greenstein asymmetry */
	float scale;		   /* grid density multiplier */
	int32_t res[3];		   /* grid resolution, x fastest */
	const float *density;  /* host pointer, res[0]*res[1]*res[2] floats */
	float bounds_min[3];
	float bounds_max[3];
} krr_medium;

/* Infinite light (rt::InfiniteLight, src/core/light.h:208-259): radiance
 * scale * tint, times image(worldToLatLong(rotation^T * w)) when an image is
 * given (math_utils.h:133-139: u = atan2(x, -z) / 2pi + 1/2, v = acos(y) / pi).
 * The image is RGBA float32, image_height rows of image_width texels, row 0
 * at v = 0 (the +y pole), sampled bilinearly with wrap-around in both
 * directions as a normalised-coordinate texture.  rotation is row-major 3x3;
 * all zeros means identity.  Light samples are uniform over the sphere, as
 * in the reference. */
typedef struct krr_infinite_light {
	float tint[3];
	float scale;
	float rotation[9];
	int32_t image_width, image_height;
	const float *image_rgba; /* host memory, copied by krr_wf_set_scene; NULL = constant */
} krr_infinite_light;

/* Scene handed to the tracer: flattened world-space triangle soup. */
typedef struct krr_scene_desc {
	int64_t n_vertices;
	const float *positions; /* 3 * n_vertices */
	const float *normals;	/* 3 * n_vertices, may be NULL if no mesh has normals */
	int64_t n_triangles;
	const int32_t *indices; /* 3 * n_triangles, absolute vertex indices */
	int32_t n_meshes;
	const krr_mesh *meshes;
	int32_t n_materials;
	const krr_material *materials;
	int32_t n_media;
	const krr_medium *media;
	int32_t n_infinite_lights;
	const krr_infinite_light *infinite_lights;
	const float *texcoords; /* 2 * n_vertices, may be NULL if no mesh has texture coordinates */
	int32_t n_textures;
	const krr_texture *textures;
	const float *tangents; /* 3 * n_vertices, may be NULL if no mesh has tangents */
} krr_scene_desc;

/* rt::CameraData (src/core/camera.h:20-30).  transform is the camera-to-world
 * affine map, row-major 3x4 (rotation | translation). */
typedef struct krr_camera {
	float film_size[2];	 /* mm, default 42.666667 x 24 */
	float focal_length;	 /* mm, default 21 */
	float focal_distance; /* default 10 */
	float lens_radius;	 /* default 0 (pinhole) */
	float aspect_ratio;	 /* default 1.777777 */
	float shutter_open;
	float shutter_time;
	float transform[12];
	int32_t medium;		 /* medium the camera sits in, -1 = none */
} krr_camera;

/* Counters of the last render() call, read back from the device. */
typedef struct krr_stats {
	uint64_t closest_rays;	 /* rays traced for closest hits, summed over depth */
	uint64_t shadow_rays;	 /* shadow rays traced */
	uint64_t paths;			 /* camera samples started: W*H*spp */
	uint64_t node_visits;	 /* BVH nodes visited by closest-hit rays of the per-depth kernels ("count_traversal") */
	uint64_t tri_tests;		 /* their ray-triangle tests ("count_traversal") */
	uint64_t shadow_node_visits; /* same for shadow rays of the per-depth kernels */
	uint64_t shadow_tri_tests;
	uint64_t closest_launches; /* number of closest-hit trace kernels launched */
	double closest_trace_ms; /* summed HIP-event time of the closest-hit trace kernels ("time_kernels") */
	double shade_ms;		 /* summed time of the shade kernels ("time_kernels") */
	double shadow_trace_ms;	 /* summed time of the shadow trace kernels ("time_kernels") */
	double render_ms;		 /* HIP-event time of the whole render call ("time_kernels") */
	double megakernel_ms;	 /* summed time of the tail megakernel ("time_kernels") */
	uint64_t megakernel_closest_rays; /* closest-hit rays traced inside the tail megakernel (part of closest_rays) */
	uint64_t megakernel_shadow_rays;  /* shadow rays traced inside it (part of shadow_rays) */
	uint64_t megakernel_node_visits; /* BVH node visits of both ("count_traversal") */
	uint64_t megakernel_tri_tests;	 /* their ray-triangle tests ("count_traversal") */
	double nrc_ms;				/* time of the radiance-cache kernels ("time_kernels") */
	uint64_t nrc_queries;		/* radiance-cache queries answered (paths ended by the cache) */
	uint64_t nrc_train_vertices; /* training vertices recorded this frame (at most nrc.batch are used) */
	uint64_t nrc_steps;			/* training steps taken so far */
	double nrc_loss;			/* loss of the last training step (synchronises) */
	double bvh_build_ms;		/* BVH build of the last set_scene (host clock; "bvh_builder" picks the builder) */
} krr_stats;

/* Closest hit record: t along the (unnormalised) ray direction, primitive
 * index into the scene triangle array (-1 = miss), barycentrics of vertex 1
 * and 2 (vertex 0 weight is 1 - b1 - b2), as optixGetTriangleBarycentrics. */
typedef struct krr_hit {
	float t;
	int32_t prim;
	float b1, b2;
} krr_hit;

/* ---------------------------------------------------------------------- */
/* Library                                                                 */
/* ---------------------------------------------------------------------- */

/* Last error message of the calling thread ("" if none). */
const char *krr_last_error(void);
/* KRR_ABI_VERSION the library was built with. */
int krr_abi_version(void);
/* Number of visible HIP devices (0 when none); never fails. */
int krr_device_count(void);

/* ---------------------------------------------------------------------- */
/* Procedural scenes (SceneImporter::import counterpart,                   */
/* src/scene/krrscene.cpp:253-305).  Host-only, no device needed.          */
/* ---------------------------------------------------------------------- */

typedef struct krr_scene krr_scene;

/* json: {"model": [...], "camera": {...}, "media": [...], ...}; see
 * kiraray_amd/csrc/scene.cpp for the procedural model types. */
int krr_scene_create(const char *json, krr_scene **out);
int krr_scene_get_desc(const krr_scene *scene, krr_scene_desc *out);
int krr_scene_get_camera(const krr_scene *scene, int view, krr_camera *out);
/* Number of cameras ("views") the scene defines (multi-view scenes > 1). */
int krr_scene_view_count(const krr_scene *scene);
/* Per-view resolution (width, height) stored in the scene, 0 if unset. */
int krr_scene_view_size(const krr_scene *scene, int view, int32_t *width, int32_t *height);
int krr_scene_destroy(krr_scene *scene);

/* ---------------------------------------------------------------------- */
/* Acceleration structure + launch (OptixBackend counterpart:              */
/* src/core/device/optix.cpp:143-250 build, src/core/device/optix.h:174-187 */
/* launch).                                                                 */
/* ---------------------------------------------------------------------- */

typedef struct krr_accel krr_accel;

/* json_opts (NULL or "" = defaults): leaf_size (4) max triangles per leaf,
 * sah_node_cost (0.25) SAH cost of a node visit in triangle tests,
 * stack_lds_entries (16) traversal stack entries per lane kept in LDS, the
 * deeper ones in a global overflow area, node_format ("float") the nodes the
 * traversal reads: "float" 128-byte nodes or "quantized" 64-byte nodes with
 * 8-bit child boxes (same hits; half the node bytes). */
int krr_accel_build(const krr_scene_desc *scene, const char *json_opts, krr_accel **out);
/* rays_dev: n records of 8 floats {ox,oy,oz,tmax,dx,dy,dz,0} in device memory.
 * Closest hit in (0, tmax); ties in t go to the lowest primitive index.
 * n <= 2^30.  Trace calls on one accel share its work counter and stack
 * overflow area: they must not overlap (same stream, or synchronized). */
int krr_accel_trace_closest(krr_accel *accel, const float *rays_dev, int64_t n, krr_hit *hits_dev,
							void *hip_stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): krr_accel_trace_closest with the same arguments and the
 * same hits, but no hipMalloc, hipFree or allocating library call.  Its work
 * counter and traversal-stack overflow live in a fixed device array of
 * krr_accel_noalloc_scratch_bytes() (16 MiB) that each GPU holds until the
 * process ends; the resident grid is sized to fit it.  KRR_ERR_CAPACITY when
 * one 128-thread block's stack overflow does not fit (raise
 * stack_lds_entries).  All calls on one GPU share that array: calls that
 * overlap on one GPU (any accel, any stream) race. */
int krr_accel_trace_closest_noalloc(krr_accel *accel, const float *rays_dev, int64_t n,
									krr_hit *hits_dev, void *hip_stream);
size_t krr_accel_noalloc_scratch_bytes(void);
/* visible_dev[i] = 1 if no triangle of a non-null material is hit in (0, tmax). */
int krr_accel_trace_any(krr_accel *accel, const float *rays_dev, int64_t n, int32_t *visible_dev,
						void *hip_stream);
/* Structure statistics: node count, max depth, leaf triangle count, bytes. */
int krr_accel_info(const krr_accel *accel, int64_t *n_nodes, int32_t *max_depth,
				   int64_t *n_leaf_tris, int64_t *bytes);
int krr_accel_destroy(krr_accel *accel);

/* ---------------------------------------------------------------------- */
/* WavefrontPathTracer render pass (src/render/wavefront/integrator.h:24-104) */
/* ---------------------------------------------------------------------- */

typedef struct krr_wf krr_wf;

/* json_params: the pass "params" object of the reference (integrator.h:96-103):
 *   nee (true), enable_medium (true), max_depth (10), rr (0.8),
 *   enable_clamp (false), clamp_max (1e3)
 * plus keys this library adds:
 *   spp (1)              samples per pixel per frame (UI-only in the reference)
 *   rr_min_depth (0)     first path depth at which Russian roulette applies
 *   legacy_spp_accumulate (false)  keep the reference's running-sum spp loop
 *   count_traversal (false)  count node visits / triangle tests
 *   time_kernels (false)     time each trace / shade launch with HIP events
 *   block_size (256), trace_block_size (128)  launch shapes
 *   wavefront_depths (0)     depths run as separate wavefront stages (one
 *                            launch per stage per depth); the remaining
 *                            depths of every path run in one persistent
 *                            launch with the same per-path stage order, so
 *                            the image does not depend on this value
 *   leaf_size (4), sah_node_cost (0.25), stack_lds_entries (16),
 *   node_format ("float")    as for krr_accel_build
 *   nrc (false)              neural radiance cache (include/krr_nrc.h): true,
 *                            or {"train_mod": 16, "batch": 65536,
 *                            "steps_per_frame": 4, and the krr_nrc_create
 *                            keys}; the batch is split into steps_per_frame
 *                            consecutive mini-batches, one Adam step each.  Render paths end at the
 *                            first vertex reached through a non-specular
 *                            bounce and take the cache's radiance there;
 *                            pixels with trainHash(pixel) % train_mod ==
 *                            frame % train_mod trace unbiased training paths
 *                            whose vertices train the cache once per frame.
 *                            Needs wavefront_depths = 0 and no media.
 *   occupancy (1.0)          fraction of the resident blocks a persistent
 *                            launch uses (in-flight path capacity)
 *   refill_lanes (-1)        persistent launch without media: a wave's
 *                            traversal round ends once fewer than this many
 *                            of its 64 lanes are still traversing; the
 *                            others shade or take their next ray, and the
 *                            unfinished rays resume in the next round (0: each
 *                            ray runs to its end; 1..64; -1: 8 for trees of
 *                            at least 4096 nodes, else 0); images and ray
 *                            counts do not depend on it
 *   leaf_share (35)          refill launch: an iteration of the traversal
 *                            takes only leaves once this percentage of the
 *                            traversing lanes holds one, else only inner
 *                            nodes (0: both every iteration)
 *   lds_nodes (0)            refill launch with float nodes: the first
 *                            lds_nodes nodes of the tree (its top levels) are
 *                            read from a per-block LDS copy (0..1024)
 *   graph (true)            record the multi-view frame into a HIP graph
 *                            once and replay it (krr_wf_render_views)
 * NULL or "" means all defaults. */
/* Replaces RenderPassFactory::deserialize<WavefrontPathTracer> + from_json
 * (src/core/renderpass.h:202-273, src/render/wavefront/integrator.h:96-103)
 * and WavefrontPathTracer::initialize (integrator.cpp:22-48). */
int krr_wf_create(const char *json_params, krr_wf **out);
/* Uploads the scene and builds the acceleration structure: replaces
 * WavefrontPathTracer::setScene (integrator.cpp:186-203) with the
 * RTScene upload (src/core/device/scene.cpp:28-135). */
int krr_wf_set_scene(krr_wf *wf, const krr_scene_desc *scene);
/* Reallocates the per-pixel queues for a width x height frame: replaces
 * WavefrontPathTracer::resize (integrator.cpp:181-184). */
int krr_wf_resize(krr_wf *wf, int32_t width, int32_t height);
/* Copies the camera and resets per-pixel state; frame_index is the 1-based
 * frame counter of the host loop (src/core/window.cpp:457): replaces
 * WavefrontPathTracer::beginFrame (integrator.cpp:205-221). */
int krr_wf_begin_frame(krr_wf *wf, const krr_camera *camera, uint64_t frame_index);
/* Traces spp samples per pixel and writes RGBA32F into film_dev, W x H
 * float4, row (H-1-y) for pixel row y (src/core/device/cuda.h:42-45):
 * replaces WavefrontPathTracer::render (integrator.cpp:223-267).  Work is
 * queued on hip_stream (0 = null stream) and the call returns without
 * waiting for it. */
int krr_wf_render(krr_wf *wf, float *film_dev, void *hip_stream);
/* RenderPass::endFrame (src/core/renderpass.h:162); nothing to do. */
int krr_wf_end_frame(krr_wf *wf);
/* Counters of the last render; synchronises the stream of that render.  No
 * reference counterpart (the reference shows none of these figures). */
int krr_wf_stats(krr_wf *wf, krr_stats *out);
/* Current params as JSON: replaces to_json (integrator.h:85-94); writes at
 * most cap-1 bytes plus a terminating 0 into buf, returns the full length. */
int krr_wf_get_params(const krr_wf *wf, char *buf, int32_t cap);
/* Frees every device allocation of the pass (~RenderPass,
 * src/core/renderpass.h:146; the reference never frees its queues). */
int krr_wf_destroy(krr_wf *wf);

/* ---------------------------------------------------------------------- */
/* MegakernelPathTracer render pass (src/render/megakernel/pathtracer.h:14-50, */
/* device.cu:50-195): one lane per pixel runs each sample's whole path, with */
/* the power heuristic between light and BSDF sampling, Russian roulette from */
/* the first vertex, no media.  Same scene, camera and film conventions as   */
/* the wavefront pass; an A/B baseline for it.                               */
/* ---------------------------------------------------------------------- */

typedef struct krr_mk krr_mk;

/* json_params: the pass "params" object of the reference (pathtracer.h:37-44):
 *   nee (true), max_depth (10), rr (0.8)
 * plus keys this library adds:
 *   spp (1)                  samples per pixel per frame (the film holds
 *                            their mean; UI-only in the reference)
 *   block_size (256)         threads per block, a multiple of 64
 *   occupancy (1.0)          fraction of the resident blocks launched
 *   time_kernels (false)     time the launch with HIP events
 *   leaf_size, sah_node_cost, stack_lds_entries, bvh_builder, node_format  as for krr_wf_create
 * NULL or "" means all defaults.  Replaces RenderPassFactory::deserialize<
 * MegakernelPathTracer> + from_json (pathtracer.h:45-49). */
int krr_mk_create(const char *json_params, krr_mk **out);
/* Uploads the scene and builds the acceleration structure
 * (MegakernelPathTracer::setScene, pathtracer.cpp:25-39).  Meshes without
 * material (medium interfaces) are rejected with KRR_ERR_UNSUPPORTED. */
int krr_mk_set_scene(krr_mk *mk, const krr_scene_desc *scene);
/* MegakernelPathTracer::resize (pathtracer.cpp:20-23). */
int krr_mk_resize(krr_mk *mk, int32_t width, int32_t height);
/* Camera and 1-based frame index of the next render; the samplers of frame f
 * start at setPixelSample(pixel, f * 512) (device.cu:157). */
int krr_mk_begin_frame(krr_mk *mk, const krr_camera *camera, uint64_t frame_index);
/* MegakernelPathTracer::render (pathtracer.cpp:52-69): RGBA32F, the mean of
 * the spp samples, into film_dev (W x H float4, row H-1-y), queued on
 * hip_stream; returns without waiting. */
int krr_mk_render(krr_mk *mk, float *film_dev, void *hip_stream);
/* closest_rays, shadow_rays, paths, render_ms (with time_kernels) and
 * bvh_build_ms of the last render; synchronises its stream. */
int krr_mk_stats(krr_mk *mk, krr_stats *out);
/* Current params as JSON (to_json, pathtracer.h:37-44); as krr_wf_get_params. */
int krr_mk_get_params(const krr_mk *mk, char *buf, int32_t cap);
/* Frees every device allocation of the pass. */
int krr_mk_destroy(krr_mk *mk);

/* ---------------------------------------------------------------------- */
/* Many camera views in one frame.  No reference counterpart: KiRaRay        */
/* renders one view per frame (integrator.cpp:205-267); this batches views   */
/* of uneven sizes into one persistent launch per sample.                    */
/* ---------------------------------------------------------------------- */

typedef struct krr_view {
	krr_camera camera;
	int32_t width, height; /* 1..65536 */
	float *film_dev;	   /* width x height float4 RGBA32F, row height-1-y, caller-owned */
} krr_view;

/* Copies the view list (cameras, sizes, film pointers) and sizes the pixel
 * state for all views, at most 2^30 pixels in total. */
int krr_wf_set_views(krr_wf *wf, const krr_view *views, int32_t n_views);
/* One frame of every view: beginFrame + render (integrator.cpp:205-267) of
 * each view, all views' pixels in one persistent launch per sample, films
 * written.  View v seeds pixel (x, y) with PCG initstate
 * interleave(x, y) | v << 32, so view 0 renders exactly what krr_wf_render
 * renders for the same camera, size and frame index.  With "graph" the
 * launches are recorded once (per views / scene / launch shape) and replayed;
 * the frame index reaches the kernels through device memory. */
int krr_wf_render_views(krr_wf *wf, uint64_t frame_index, void *hip_stream);
/* Times persistent-launch shapes - trace_block_size {64, 128, 256} x
 * stack_lds_entries {8, 16, 24} x occupancy {1, 0.75, 0.5} - on the first
 * trial_paths (<= 0: 2^21) camera paths of the views (or of the single view,
 * after krr_wf_begin_frame), then (scenes without media, refill_lanes -1)
 * refill_lanes {0, 4, 8, 12, 16} at the fastest shape, and keeps the fastest.  Images do not
 * depend on the choice.  Writes a JSON report (candidates with ms, best) into report
 * (cap bytes) and returns its full length.  No reference counterpart. */
int krr_wf_autotune(krr_wf *wf, int64_t trial_paths, char *report, int32_t cap, void *hip_stream);

/* ---------------------------------------------------------------------- */
/* Error metric (ErrorMeasurePass, src/render/passes/errormeasure/metrics.cu) */
/* ---------------------------------------------------------------------- */

enum krr_metric { KRR_METRIC_MSE = 0, KRR_METRIC_MAPE = 1, KRR_METRIC_SMAPE = 2, KRR_METRIC_RELMSE = 3 };

/* img_dev, ref_dev: n_pixels float4 in device memory; result accumulated in
 * FP64 and written to *out (synchronises hip_stream). */
int krr_error_metric(int32_t metric, const float *img_dev, const float *ref_dev, int64_t n_pixels,
					 double *out, void *hip_stream);

/* ---------------------------------------------------------------------- */
/* Frame accumulation (AccumulatePass, src/render/passes/accumulate/       */
/* accumulate.cu:30-79)                                                     */
/* ---------------------------------------------------------------------- */

enum krr_accum_mode { KRR_ACCUM_SUM = 0, KRR_ACCUM_MOVING_AVERAGE = 1 };

/* Folds frame_dev (n_pixels float4) into accum_dev (n_pixels float4 when
 * precision is 32, double4 when 64), accum_count frames having been folded
 * before; rewrites frame_dev with the accumulated mean.  max_count > 0 stops
 * adding in sum mode after that many frames.  Queued on hip_stream. */
int krr_accumulate(void *accum_dev, float *frame_dev, int64_t n_pixels, int64_t accum_count, int64_t max_count, int32_t mode,
				   int32_t precision, void *hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* KRR_WAVEFRONT_H */
