/*
 * lucid_c.h — C ABI of the MI355X exact order-independent-transparency rasterizer.
 *
 * The entry points replace the reference's LucidRenderer (src/lucid_renderer.h:21-96):
 *
 *   lucid_create       <- LucidRenderer::exConstruct   (src/lucid_renderer.cpp:186-317)
 *   lucid_destroy      <- ~LucidRenderer
 *   lucid_render       <- LucidRenderer::render        (src/lucid_renderer.cpp:319-350),
 *                         with setupInputData/uploadInstances (:352-451) folded in
 *   lucid_read_info    <- the async cmds.download(m_info) (:343-346) consumed by
 *                         verifyInfo/getStats (:629-836), done synchronously here
 *   lucid_verify_info  <- LucidRenderer::verifyInfo    (:629-707)
 *   lucid_read_bin_lists  exposes g_bin_quads/g_bin_tris (bin_dispatcher.glsl:39-41)
 *   lucid_bin_size / lucid_block_size / lucid_max_visible_quads / lucid_max_scene_quads
 *                      <- binSize()/blockSize()/maxVisibleQuads()/maxSceneQuads() (:43-49)
 *
 * Conventions: plain C types only; every call returns an int status (LUCID_OK == 0,
 * negative on error); no exception crosses the boundary. Device pointers are HIP device
 * pointers on the renderer's device; `stream` is a hipStream_t passed as void* (NULL =
 * the default stream). Rendering is asynchronous on `stream`; lucid_finish waits for it.
 */
#ifndef LUCID_C_H
#define LUCID_C_H

#include <stddef.h>
#include <stdint.h>

#include "lucid_structs.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LUCID_OK 0
#define LUCID_ERR_INVALID_ARG (-1)
#define LUCID_ERR_OUT_OF_MEMORY (-2)
#define LUCID_ERR_HIP (-3)
#define LUCID_ERR_NO_DEVICE (-4)
#define LUCID_ERR_CAPACITY (-5)  /* a frame exceeded an internal buffer; buffers were grown */
#define LUCID_ERR_UNSUPPORTED (-6) /* e.g. the backward of a textured frame */

/* LucidRenderOpt bits (src/lucid_renderer.h:10-11). Debug/timer options are accepted
 * and ignored; rocprof replaces them. LUCID_OPT_ALPHA_THRESHOLD ("over" blending only):
 * once every pixel of an 8x8 block (or of one depth range of a long block) has
 * transmittance below 1/128, its later samples are not shaded (shading.glsl:62,282-284;
 * fragment counts stay exact; colours differ from the exact composite by at most 1/128
 * per channel); the backward rejects it. */
#define LUCID_OPT_DEBUG_QUAD_SETUP (1u << 0)
#define LUCID_OPT_DEBUG_BIN_COUNTER (1u << 1)
#define LUCID_OPT_DEBUG_BIN_DISPATCHER (1u << 2)
#define LUCID_OPT_DEBUG_RASTER (1u << 3)
#define LUCID_OPT_TIMERS (1u << 4)
#define LUCID_OPT_ADDITIVE_BLENDING (1u << 5)
#define LUCID_OPT_VISUALIZE_ERRORS (1u << 6)
#define LUCID_OPT_ALPHA_THRESHOLD (1u << 7)
/* Extension: shaded samples stay fp32 instead of the reference's RGBA8 truncation
 * (encodeRGBA8, shared/funcs.glsl:118-121). Used by the differentiable path. */
#define LUCID_OPT_FP32_SAMPLES (1u << 16)
/* Extension: the unsorted baseline. Samples are blended in block-list order with no depth
 * order (no block sort, no ordering window, no exact per-pixel sort except for blocks too
 * long to pre-sort), as hardware alpha blending would (SimpleRenderer,
 * src/simple_renderer.cpp:84-90,124). Bins, lists and fragment counts are the exact
 * frame's; colours are not (unless blending is additive or nothing overlaps). For
 * measuring what the exact order costs; the backward rejects it. */
#define LUCID_OPT_UNSORTED (1u << 17)

typedef struct lucid_renderer lucid_renderer;
typedef struct lucid_batch lucid_batch;

/* Performance parameters of the block rasteriser. Coverage, bins, lists and fragment
 * counts do not depend on them; colours only through the rounding of partial composites
 * (blocks split into depth ranges). The bin size (32) and block size (8) are the
 * reference's and are not tunable: they decide which row starts each scanline. */
typedef struct lucid_tuning {
	int slice_words;      /* entries staged in LDS per step: 64 * slice_words (1 or 2) */
	int presort_min;      /* high-bin blocks with more entries are pre-sorted (64..256) */
	int samples_per_wave; /* target samples per wave of a pre-sorted block (256..65536) */
	int window;           /* per-pixel ordering window of the block rasterisers in frames of
						   * RGBA8 samples with unit lighting and flat colours, and of the
						   * backward's first pass: 1 (one slot; two in the backward) for
						   * lists whose block order is nearly every
						   * pixel's depth order, 4 for more disorder. A pixel the window
						   * cannot order takes the exact per-pixel sort either way. */
} lucid_tuning;

typedef struct lucid_create_info {
	int view_width, view_height; /* at most 2560 x 2048 (src/lucid_renderer.h:27) */
	uint32_t opts;               /* LUCID_OPT_* */
	int device;                  /* HIP device ordinal */
	int64_t max_quads;           /* initial quad capacity (0: 1M); grows on demand */
} lucid_create_info;

/* An RGBA8 texture atlas with its mip chain (the reference's opaque and transparent atlases,
 * Scene::textureAtlasPair, src/scene.cpp:514, bound at shading.glsl:27-28). Level 0 is
 * width x height texels, row-major, row 0 = v 0; level l (l < num_levels <= 16) follows the
 * previous one and is max(1, width >> l) x max(1, height >> l). Sampled as the reference's
 * sampler is set up (src/lucid_base.h:30: linear magnification, minification and mip
 * filtering, repeat addressing) with isotropic LOD selection. */
typedef struct lucid_texture {
	const uint32_t *texels; /* device; NULL = no texture */
	int width, height, num_levels;
} lucid_texture;

typedef struct lucid_frame_inputs {
	/* host arrays (uploaded when they change) */
	const lucid_instance_data *instances; /* InstanceData[num_instances] */
	int num_instances;
	const uint32_t *instance_colors; /* RGBA8 per instance, or NULL (white) */
	const lucid_config *config;      /* LucidConfig (std140 bytes) */
	/* device arrays */
	const uint32_t *quad_indices; /* u32 indices, addressed through InstanceData */
	const float *positions;       /* float3 per vertex */
	int64_t num_verts;
	const uint32_t *colors_rgba8; /* RGBA8 per vertex (reference format), or NULL */
	const float *colors_rgb;      /* float3 per vertex, or NULL (used when colors_rgba8 is NULL) */
	const float *tri_alpha;       /* alpha per triangle id (2 per quad), with colors_rgb */
	const uint32_t *normals;      /* packed 10:10:10 normal per vertex, or NULL */
	/* albedo textures (shading.glsl:125-160), for instances with LUCID_INST_HAS_ALBEDO_TEXTURE:
	 * per-vertex texture coordinates, per-instance uv rectangles (min.xy, size.xy within the
	 * atlas; used with LUCID_INST_HAS_UV_RECT, uploadInstances, src/lucid_renderer.cpp:385)
	 * and the atlases (LUCID_INST_TEX_OPAQUE picks the opaque one, whose alpha reads as 1) */
	const float *tex_coords;              /* device, float2 per vertex, or NULL */
	const lucid_vec4 *instance_uv_rects;  /* host, one per instance, or NULL */
	lucid_texture opaque_texture;
	lucid_texture transparent_texture;
} lucid_frame_inputs;

typedef struct lucid_frame_outputs {
	uint8_t *image_rgba8;   /* device, width*height*4, row-major, row 0 = top; may be NULL */
	float *image_rgba32f;   /* device, width*height*4 float; may be NULL */
	uint32_t *frag_counts;  /* device, width*height fragments per pixel; may be NULL */
} lucid_frame_outputs;

const char *lucid_error_string(int status);

int lucid_create(const lucid_create_info *info, lucid_renderer **out);
void lucid_destroy(lucid_renderer *r);

int lucid_render(lucid_renderer *r, const lucid_frame_inputs *in, const lucid_frame_outputs *out,
				 void *stream);
/* lucid_render with flags. LUCID_RENDER_RECORD_BACKWARD: the frame's rasterisation also
 * records what the backward's first pass computes (each pixel's blend order per depth range,
 * the ranges' fp64 partial composites and the gradient row layout), so that a following
 * lucid_backward_ex(LUCID_BACKWARD_REUSE_FRAME) on the same inputs starts from the record
 * instead of walking the frame again. Needs LUCID_OPT_FP32_SAMPLES, fp32 colours and "over"
 * blending (else LUCID_ERR_UNSUPPORTED). Fragment counts equal lucid_render's; colours too,
 * up to the rounding of pixels the recording pass leaves to the exact per-pixel sort. When
 * the record does not fit the renderer's backward buffers (sized by earlier frames), the
 * frame is rendered without it, the buffers grow, and the backward walks the frame itself.
 * Flags 0: lucid_render. No counterpart in the reference (the differentiable extension). */
#define LUCID_RENDER_RECORD_BACKWARD 1u
int lucid_render_ex(lucid_renderer *r, const lucid_frame_inputs *in,
					const lucid_frame_outputs *out, uint32_t flags, void *stream);
/* 1 if the renderer's last finished frame holds a backward record (rendered with
 * LUCID_RENDER_RECORD_BACKWARD, the record fitted, and no backward buffer has grown since),
 * else 0. No counterpart in the reference. */
int lucid_frame_recorded(const lucid_renderer *r);
/* Adjoint of the fp32-sample forward ("over" blending, fp32 vertex colours colors_rgb and
 * per-triangle alpha tri_alpha): given dL/d(image_rgba32f) (device, width*height float4;
 * the alpha channel is ignored), writes dL/d(colors_rgb) (device, num_verts float3) and
 * dL/d(tri_alpha) (device, one float per triangle id, 2 per quad); either output may be NULL.
 * The frame's fragment lists are recomputed, not stored. Synchronous on `stream`. Returns
 * LUCID_ERR_CAPACITY after growing an internal buffer (call again), LUCID_ERR_UNSUPPORTED
 * for RGBA8 colours, additive blending or the alpha threshold.
 * No counterpart in the reference (which has no backward pass): this is the differentiable
 * extension of the drop-in renderer. */
int lucid_backward(lucid_renderer *r, const lucid_frame_inputs *in, const float *grad_image,
				   float *grad_colors_rgb, float *grad_tri_alpha, void *stream);
/* lucid_backward with flags. LUCID_BACKWARD_REUSE_FRAME: when the renderer's last frame was
 * rendered from the same inputs (same device pointers, counts, instances, config and
 * tuning; their contents must not have changed since) and finished without an overflow,
 * its triangle setup, bins and per-block triangle lists are reused instead of rebuilt; the
 * per-pixel depth-sorted fragment lists are still recomputed. Otherwise (or with flags 0)
 * this is lucid_backward. Results are bit-identical either way.
 * No counterpart in the reference. */
#define LUCID_BACKWARD_REUSE_FRAME 1u
int lucid_backward_ex(lucid_renderer *r, const lucid_frame_inputs *in, const float *grad_image,
					  float *grad_colors_rgb, float *grad_tri_alpha, uint32_t flags, void *stream);

/* Tuning of a renderer (defaults 1, 128, 1024, 1). No counterpart in the reference. */
int lucid_set_tuning(lucid_renderer *r, const lucid_tuning *t);
int lucid_get_tuning(const lucid_renderer *r, lucid_tuning *t);
/* Times every tuning candidate on this frame (`iters` timed frames each, HIP events on
 * `stream`), keeps the fastest in the renderer and returns it in *best (may be NULL).
 * Synchronous. No counterpart in the reference (its parameters are shader constants). */
int lucid_autotune(lucid_renderer *r, const lucid_frame_inputs *in, const lucid_frame_outputs *out,
				   void *stream, int iters, lucid_tuning *best);

/* Batches: one frame of each of n renderers (distinct, same device, created without
 * LUCID_OPT_TIMERS) recorded into one hipGraph, frames spread over forked streams, and
 * launched with a single hipGraphLaunch on `stream`. The graph is recorded again only when
 * a frame's constants or buffers change (new sizes, grown capacities). lucid_batch_finish
 * waits and returns LUCID_ERR_CAPACITY if any frame overflowed (render the batch again).
 * No counterpart in the reference (one scene per frame there). */
int lucid_batch_create(lucid_renderer *const *renderers, int n, lucid_batch **out);
void lucid_batch_destroy(lucid_batch *b);
int lucid_batch_render(lucid_batch *b, const lucid_frame_inputs *ins,
					   const lucid_frame_outputs *outs, void *stream);
int lucid_batch_finish(lucid_batch *b);

/* One forward frame that allocates nothing: no renderer object, no hipMalloc/hipFree and no
 * allocating library call. No counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture set-up, tight serving loops, harnesses
 * that forbid hipMalloc). Same kernels, inputs and outputs as lucid_render on a renderer of
 * width x height with `opts` and `tuning` (NULL: the defaults 1, 128, 1024, 1); results are
 * identical to that renderer's. Every internal buffer lives in a fixed device array of the
 * library (65 MiB, one copy per GPU the library runs on, held until the process ends), so the
 * frame is limited to at most 256 bins of 32 x 32 pixels (e.g. 512 x 512), 65536 quads and
 * 1024 instances (larger: LUCID_ERR_INVALID_ARG), no albedo textures (LUCID_ERR_UNSUPPORTED),
 * and to that array's pools: high bins (1024+ triangles) may hold 131072 triangle
 * references in all (a triangle counts once, a quad twice, per bin it touches; e.g. 20000
 * clustered soup triangles on 256 x 256 take 62358), pre-sorted blocks 2 MiB and deep-pixel
 * spills 1 MiB. A frame that would overflow a pool returns LUCID_ERR_CAPACITY (its outputs
 * are incomplete; render it with lucid_render). Synchronous:
 * the instance table is staged on `stream` and the call waits for the frame. All calls on one
 * GPU share the array: calls that overlap on one GPU (other threads, other streams) race. */
int lucid_render_noalloc(int width, int height, uint32_t opts, const lucid_tuning *tuning,
						 const lucid_frame_inputs *in, const lucid_frame_outputs *out,
						 void *stream);

/* Waits for the last frame. Returns LUCID_ERR_CAPACITY if it overflowed an internal buffer
 * (the buffers have then been grown and the frame must be rendered again). */
int lucid_finish(lucid_renderer *r);

/* Synchronous copy of LucidInfo followed by the 10 per-bin arrays (g_counts);
 * num_u32 must be at least LUCID_INFO_U32_SIZE + 10 * bin_count. */
int lucid_read_info(lucid_renderer *r, uint32_t *dst, size_t num_u32);
/* Checks the prefix-sum invariants of the last frame; returns the number of violations. */
int lucid_verify_info(lucid_renderer *r);
/* Copies the per-bin quad and tri lists of the last frame (host buffers, capacity in entries);
 * num_quads and num_tris receive the totals. */
int lucid_read_bin_lists(lucid_renderer *r, uint32_t *quads, size_t quad_cap, uint32_t *tris,
						 size_t tri_cap, size_t *num_quads, size_t *num_tris);
/* Copies per-triangle setup records of the last frame in the reference's storage order:
 * per triangle 20 u32 = bary0, bary1, scan0, scan1, depth (quad_setup.glsl:305-339),
 * and the per-quad AABB words (0xffffffff = rejected). */
int lucid_read_setup(lucid_renderer *r, uint32_t *tri_records, size_t num_tris,
					 uint32_t *quad_aabbs, size_t num_quads);

/* Per-stage GPU times (ms) of the last frame, measured with HIP events on the frame's
 * stream when the renderer was created with LUCID_OPT_TIMERS; replaces the reference's
 * in-shader timers (LucidRenderOpt::timers, src/lucid_renderer.cpp:582-596,754-762).
 * Stages: 0 setup, 1 binning, 2 row generation of low bins, 3 row generation of high bins
 * plus sorting of long block lists, 4 block rasterisation, 5 exact per-pixel fallback.
 * Returns the number of stages written, or a negative status. */
int lucid_stage_times(lucid_renderer *r, float *ms, int max_stages);

/* Quad generator (host; replaces genQuads/quadNodes/triNeighbours, src/quad_generator.cpp:8-201,
 * as called by Scene::generateQuads, src/scene.cpp:237-247): pairs edge-adjacent triangles
 * (u32 index triples) into quads {v0, v1, v2, v3} = triangles (v0, v1, v2) + (v0, v2, v3), in
 * the original triangle order; unpaired triangles become (a, b, c, c). quads_out needs room
 * for 4 * num_tris indices (or is NULL to count only); *num_quads receives the count.
 * square_weight: the reference's squareness weight (0.5 in its app). */
int lucid_generate_quads(const float *positions, int64_t num_verts, const uint32_t *tris,
						 int64_t num_tris, float square_weight, uint32_t *quads_out,
						 int64_t *num_quads);

int lucid_bin_count(const lucid_renderer *r);
int lucid_bin_size(const lucid_renderer *r);
int lucid_block_size(const lucid_renderer *r);
int64_t lucid_max_visible_quads(const lucid_renderer *r);
int64_t lucid_max_scene_quads(const lucid_renderer *r);

#ifdef __cplusplus
}
#endif

#endif
