/*
 * CPU ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * A single-threaded C restatement of the reference's Lucid compute pipeline
 * (nadult/lucid, the GLSL stages under data/shaders + src/lucid_renderer.cpp). Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker; the product library never links or calls it.
 *
 * Parity status of each function is listed in lucid_oracle.c's header.
 */
#ifndef LUCID_ORACLE_H
#define LUCID_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/lucid_structs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* blend modes */
#define ORACLE_MODE_EXACT 0 /* full per-pixel sort, ties broken by triangle index */
#define ORACLE_MODE_REF3 1  /* the reference's block sort + 3-deep reduction window */
#define ORACLE_MODE_COUNTS 2 /* coverage and bins only: no shading or blending */

/* option bits */
#define ORACLE_OPT_FP32_SAMPLES 0x1u /* no RGBA8 quantisation of shaded samples */
#define ORACLE_OPT_ADDITIVE 0x2u     /* additive blending variant */

/* RGBA8 atlas with its mip chain: level l (max(1, w >> l) x max(1, h >> l) texels, row-major)
 * follows level l - 1. */
typedef struct oracle_texture {
	const uint32_t *texels; /* NULL = none */
	int width, height, num_levels;
} oracle_texture;

typedef struct oracle_inputs {
	int width, height;
	const lucid_config *config;
	const lucid_instance_data *instances;
	int num_instances;
	const uint32_t *instance_colors; /* may be NULL (all white) */
	const uint32_t *quad_indices;
	const float *positions;
	int num_verts;
	const uint32_t *colors_rgba8; /* RGBA8 per vertex, or NULL */
	const float *colors_rgb;      /* fp32 RGB per vertex, or NULL */
	const float *tri_alpha;       /* fp32 alpha per triangle id (2 per quad); with colors_rgb */
	const uint32_t *normals;      /* packed 10:10:10 per vertex, or NULL */
	int mode;
	uint32_t opts;
	/* albedo textures (instances with LUCID_INST_HAS_ALBEDO_TEXTURE) */
	const float *tex_coords;        /* float2 per vertex, or NULL */
	const float *instance_uv_rects; /* float4 per instance (min.xy, size.xy), or NULL */
	oracle_texture textures[2];     /* [0] opaque atlas, [1] transparent atlas */
} oracle_inputs;

/* One covered sample of the exact mode, listed per pixel in blend order (front to back,
 * depth ties by triangle id). */
typedef struct oracle_fragment {
	uint32_t x, y, tri, order; /* order: position in the pixel's sorted list */
	float inv_ray;
	float bary[2];  /* barycentrics of corners 1 and 2 (corner 0: 1 - b0 - b1) */
	float light[3]; /* lighting factor per channel */
} oracle_fragment;

typedef struct oracle_outputs {
	uint32_t *info;        /* LUCID_INFO_U32_SIZE + 10 * bin_count */
	uint32_t *quad_aabb;   /* per input quad; 0xffffffff = rejected */
	uint32_t *tri_records; /* per tri: 20 u32 = bary0, bary1, scan0, scan1, depth (uvec4 each) */
	uint32_t *tri_normals; /* per tri */
	uint32_t *quad_colors; /* per quad: 4 u32 */
	uint32_t *bin_quads;   /* capacity entries */
	uint32_t *bin_tris;    /* capacity entries */
	size_t bin_capacity;
	uint8_t *image;        /* width * height * 4 */
	float *image_f;        /* width * height * 4, may be NULL */
	uint32_t *frag_counts; /* width * height */
	uint32_t *bin_levels;  /* per bin: level used by the rasterizer (low/high), may be NULL */
	uint64_t num_fragments;
	uint32_t num_ref_errors; /* bins the reference would paint red (raster_high limits) */
	oracle_fragment *frags;  /* exact mode: every sample, may be NULL */
	size_t frags_cap;
	size_t num_frags_dumped; /* set by oracle_render (may exceed frags_cap) */
	/* pixels whose samples frags receives: x in [dump_x0, dump_x1), y in [dump_y0, dump_y1);
	 * all pixels when dump_x1 <= dump_x0 */
	int dump_x0, dump_y0, dump_x1, dump_y1;
} oracle_outputs;

/* Runs the whole frame. Returns 0 on success, negative on invalid input or capacity. */
int oracle_render(const oracle_inputs *in, oracle_outputs *out);

/* Encoders restated for unit tests (funcs.glsl:28-33, 52-59, 100-121). */
uint32_t oracle_encode_aabb28(uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1);
uint32_t oracle_encode_rgba8(float r, float g, float b, float a);
uint32_t oracle_encode_normal_uint(float x, float y, float z);
int32_t oracle_f2i(float x);
uint32_t oracle_f2u(float x);

/* textureGrad of one atlas (see lucid_oracle.c), for unit tests: RGBA in out[4]. */
void oracle_texture_grad(const oracle_texture *t, float s, float v, float dsx, float dvx, float dsy,
						 float dvy, float out[4]);

/* Number of worker threads used by oracle_render's raster stage (OpenMP); 1 if built without. */
int oracle_num_threads(void);

#ifdef __cplusplus
}
#endif

#endif
