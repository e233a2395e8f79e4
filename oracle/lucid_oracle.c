/*
 * CPU ORACLE — TEST INFRASTRUCTURE ONLY (see lucid_oracle.h).
 *
 * Single-threaded-per-bin C restatement of the reference's forward pipeline:
 *   setup        data/shaders/quad_setup.glsl:64-340
 *   bin counts   data/shaders/bin_counter.glsl:53-134
 *   categorise   data/shaders/bin_categorizer.glsl:22-89
 *   bin lists    data/shaders/bin_dispatcher.glsl:55-196
 *   raster rows  data/shaders/raster_low.glsl:39-79, raster_high.glsl:54-106,
 *                data/shaders/shared/raster.glsl:116-140, shared/scanline.glsl:13-26
 *   promotion    data/shaders/raster_low.glsl:90-106,229-237
 *   block keys   data/shaders/raster_low.glsl:108-150, raster_high.glsl:146-194,
 *                data/shaders/shared/raster.glsl:142-176
 *   shading      data/shaders/shared/shading.glsl:107-184, shared/funcs.glsl:100-164,261-271
 *   textures     data/shaders/shared/shading.glsl:111-122,125-160 (textureGrad on the atlases with
 *                the sampler of src/lucid_base.h:30, restated from the Vulkan sampling rules:
 *                isotropic LOD, trilinear, repeat), data/shaders/quad_setup.glsl:265-271
 *   reduction    data/shaders/shared/shading.glsl:186-314 (REF3 mode)
 *
 * Arithmetic conventions (shared with the HIP kernels, fixed here because
 * GLSL leaves them to the implementation):
 *   - no floating-point contraction: every a*b+c is a rounded multiply then a rounded add;
 *   - expression order follows the GLSL source left to right; dot() is (x*x+y*y)+z*z;
 *     cross() follows the GLSL specification; mat4*vec4 sums columns 0..3 in order;
 *   - division and sqrt are IEEE correctly rounded; inversesqrt(x) is 1/sqrt(x);
 *   - min/max are IEEE minNum/maxNum (fminf/fmaxf);
 *   - float->int and float->uint conversions truncate and saturate, NaN -> 0
 *     (the behaviour of AMD's v_cvt_i32_f32 / v_cvt_u32_f32).
 *
 * Parity pins (tests/test_oracle.py): encoders against closed forms; coverage
 * against a brute-force pixel-centre edge test; exact blending against stacked
 * screen-aligned layers with closed-form colour and against a float64 per-pixel
 * sort; bin counts against per-quad brute-force rectangles/scanlines and the
 * verifyInfo invariants (src/lucid_renderer.cpp:629-707).
 * Parity unpinned: the reference's own float behaviour on a Vulkan driver (no
 * Vulkan, lavapipe or GLSL compiler exists here), and REF3-mode ordering, which
 * is nondeterministic in the reference (LDS atomics) and is restated with list
 * order as the tie rule.
 */
#include "lucid_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define BIN_SIZE 32
#define BIN_SHIFT 5
#define BIN_MASK 31
#define MAX_BLOCK_TRIS 256        /* raster_low.glsl:17 */
#define MAX_HBLOCK_ROW_TRIS 16384 /* raster_high.glsl:30 */
#define MAX_HBLOCK_TRIS 4096      /* raster_high.glsl:27 */
#define SEGMENT_SIZE 256          /* raster.glsl:71 */
#define LOW_LEVEL_LIMIT 1024      /* bin_categorizer.glsl:352 */
#define INVALID_AABB 0xffffffffu

typedef struct {
	float x, y, z;
} v3;
typedef struct {
	float x, y, z, w;
} v4;

/* ---------------------------------------------------------------- helpers */

int32_t oracle_f2i(float x) {
	if(x != x)
		return 0;
	if(x >= 2147483648.0f)
		return 2147483647;
	if(x <= -2147483648.0f)
		return (int32_t)0x80000000u;
	return (int32_t)x;
}

uint32_t oracle_f2u(float x) {
	if(!(x > 0.0f))
		return 0;
	if(x >= 4294967296.0f)
		return 0xffffffffu;
	return (uint32_t)x;
}

static inline float fmx(float a, float b) { return fmaxf(a, b); }
static inline float fmn(float a, float b) { return fminf(a, b); }
static inline float clampf(float x, float lo, float hi) { return fmn(fmx(x, lo), hi); }
static inline float satf(float x) { return clampf(x, 0.0f, 1.0f); }
static inline int clampi(int x, int lo, int hi) { return x < lo ? lo : x > hi ? hi : x; }
static inline float u2f(uint32_t u) {
	float f;
	memcpy(&f, &u, 4);
	return f;
}
static inline uint32_t f2bits(float f) {
	uint32_t u;
	memcpy(&u, &f, 4);
	return u;
}

static inline v3 v3_make(float x, float y, float z) {
	v3 r = {x, y, z};
	return r;
}
static inline v3 v3_sub(v3 a, v3 b) { return v3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 v3_add(v3 a, v3 b) { return v3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 v3_mul(v3 a, float s) { return v3_make(a.x * s, a.y * s, a.z * s); }
static inline v3 v3_neg(v3 a) { return v3_make(-a.x, -a.y, -a.z); }
static inline float v3_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 v3_cross(v3 a, v3 b) {
	return v3_make(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline int v3_eq(v3 a, v3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
static inline v3 v4_xyz(lucid_vec4 v) { return v3_make(v.x, v.y, v.z); }

/* mat4 * vec4(p, 1.0): columns summed in order */
static inline v4 mat_mul_point(const lucid_mat4 *m, v3 p) {
	v4 r;
	r.x = ((m->col[0].x * p.x + m->col[1].x * p.y) + m->col[2].x * p.z) + m->col[3].x * 1.0f;
	r.y = ((m->col[0].y * p.x + m->col[1].y * p.y) + m->col[2].y * p.z) + m->col[3].y * 1.0f;
	r.z = ((m->col[0].z * p.x + m->col[1].z * p.y) + m->col[2].z * p.z) + m->col[3].z * 1.0f;
	r.w = ((m->col[0].w * p.x + m->col[1].w * p.y) + m->col[2].w * p.z) + m->col[3].w * 1.0f;
	return r;
}

/* ---------------------------------------------------------------- encoders (funcs.glsl) */

uint32_t oracle_encode_aabb28(uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1) {
	return ((x0 & 0x7fu) << 0) | ((y0 & 0x7fu) << 7) | ((x1 & 0x7fu) << 14) |
		   ((y1 & 0x7fu) << 21);
}

uint32_t oracle_encode_rgba8(float r, float g, float b, float a) {
	return oracle_f2u(r * 255.0f) | (oracle_f2u(g * 255.0f) << 8) |
		   (oracle_f2u(b * 255.0f) << 16) | (oracle_f2u(a * 255.0f) << 24);
}

uint32_t oracle_encode_normal_uint(float x, float y, float z) {
	uint32_t ux = oracle_f2u(512.0f + x * 511.0f) & 0x3ffu;
	uint32_t uy = oracle_f2u(512.0f + y * 511.0f) & 0x3ffu;
	uint32_t uz = oracle_f2u(512.0f + z * 511.0f) & 0x3ffu;
	return ux | (uy << 10) | (uz << 20);
}

static inline v4 decode_rgba8(uint32_t c) {
	v4 r;
	r.x = (float)((c >> 0) & 0xffu) * (1.0f / 255.0f);
	r.y = (float)((c >> 8) & 0xffu) * (1.0f / 255.0f);
	r.z = (float)((c >> 16) & 0xffu) * (1.0f / 255.0f);
	r.w = (float)((c >> 24) & 0xffu) * (1.0f / 255.0f);
	return r;
}

static inline v3 decode_normal_uint(uint32_t n) {
	return v3_make(((float)((n >> 0) & 0x3ffu) - 512.0f) * (1.0f / 511.0f),
				   ((float)((n >> 10) & 0x3ffu) - 512.0f) * (1.0f / 511.0f),
				   ((float)((n >> 20) & 0x3ffu) - 512.0f) * (1.0f / 511.0f));
}

/* funcs.glsl:153-164 */
static inline float srgb_to_linear(float c) {
	return c < 0.04045f ? (1.0f / 12.92f) * c : powf((c + 0.055f) * (1.0f / 1.055f), 2.4f);
}
static inline float linear_to_srgb(float c) {
	return c < 0.0031308f ? 12.92f * c : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
}

/* ---------------------------------------------------------------- frame state */

typedef struct frame {
	const oracle_inputs *in;
	int width, height;
	int bin_count_x, bin_count_y, bin_count;
	int num_quads; /* total input quads */
	v3 ray_dir0, dirx, diry, dir0, origin0;
	/* per quad */
	uint32_t *quad_aabb;
	uint32_t *quad_inst;
	/* per tri (2 per quad) records: scan0[4], scan1[4], depth[4], bary0[4], bary1[4] */
	uint32_t *scan;  /* 8 u32 per tri */
	uint32_t *depth; /* 4 u32 per tri */
	uint32_t *bary;  /* 8 u32 per tri */
	uint32_t *tri_normal;
	uint32_t *quad_color;  /* 4 u32 per quad */
	uint32_t *quad_normal; /* 4 u32 per quad (INST_HAS_VERTEX_NORMALS) */
	float *tri_color;     /* fp32 colour mode: 12 floats per tri (rgb0 a | rgb1 _ | rgb2 _) */
	float *tri_tex;       /* textured instances: 6 floats per tri (tex0, tex1 - tex0, tex2 - tex0) */
	/* bins */
	uint32_t *counts; /* 10 * bin_count */
	uint32_t *bin_quads, *bin_tris;
	/* info */
	lucid_info info;
} frame;

/* ---------------------------------------------------------------- setup (quad_setup.glsl) */

static uint32_t vertex_clip_mask(v4 p) {
	return (p.x < -p.w ? 0x001u : 0u) | (p.x > p.w ? 0x002u : 0u) | (p.y < -p.w ? 0x004u : 0u) |
		   (p.y > p.w ? 0x008u : 0u) | (p.z < -p.w ? 0x010u : 0u) | (p.z > p.w ? 0x020u : 0u);
}

/* quad_setup.glsl:77-128 (Blinn's clipped AABB) */
static void clipped_aabb(const v4 v[3], const float inv_w[3], uint32_t clipmask, float aabb[4]) {
	aabb[0] = 1.0f, aabb[1] = 1.0f, aabb[2] = -1.0f, aabb[3] = -1.0f;
	int any_vis = 0;
	uint32_t or_clipmask = clipmask | (clipmask >> 8) | (clipmask >> 16);
	for(int i = 0; i < 3; i++) {
		uint32_t cur = clipmask >> (i * 8);
		if((cur & 0x3u) == 0) {
			any_vis |= 0x1;
			if(v[i].x - aabb[0] * v[i].w < 0.0f)
				aabb[0] = v[i].x * inv_w[i];
			if(v[i].x - aabb[2] * v[i].w > 0.0f)
				aabb[2] = v[i].x * inv_w[i];
		}
		if((cur & 0xcu) == 0) {
			any_vis |= 0x10;
			if(v[i].y - aabb[1] * v[i].w < 0.0f)
				aabb[1] = v[i].y * inv_w[i];
			if(v[i].y - aabb[3] * v[i].w > 0.0f)
				aabb[3] = v[i].y * inv_w[i];
		}
	}
	if((any_vis & 0x0f) == 0) {
		aabb[0] = -1.0f;
		aabb[2] = 1.0f;
	} else if((or_clipmask & 0x3u) != 0) {
		for(int i = 0; i < 3; i++) {
			uint32_t cur = clipmask >> (i * 8);
			if((cur & 0x1u) != 0 && v[i].x - aabb[0] * v[i].w < 0.0f)
				aabb[0] = -1.0f;
			if((cur & 0x2u) != 0 && v[i].x - aabb[2] * v[i].w > 0.0f)
				aabb[2] = 1.0f;
		}
	}
	if((any_vis & 0xf0) == 0) {
		aabb[1] = -1.0f;
		aabb[3] = 1.0f;
	} else if((or_clipmask & 0xcu) != 0) {
		for(int i = 0; i < 3; i++) {
			uint32_t cur = clipmask >> (i * 8);
			if((cur & 0x4u) != 0 && v[i].y - aabb[1] * v[i].w < 0.0f)
				aabb[1] = -1.0f;
			if((cur & 0x8u) != 0 && v[i].y - aabb[3] * v[i].w > 0.0f)
				aabb[3] = 1.0f;
		}
	}
}

static void plain_aabb(v3 a, v3 b, v3 c, float out[4]) {
	out[0] = fmn(fmn(a.x, b.x), c.x);
	out[1] = fmn(fmn(a.y, b.y), c.y);
	out[2] = fmx(fmx(a.x, b.x), c.x);
	out[3] = fmx(fmx(a.y, b.y), c.y);
}

static v3 vertex_load(const frame *f, uint32_t idx) {
	const float *p = f->in->positions + (size_t)idx * 3;
	return v3_make(p[0], p[1], p[2]);
}

/* quad_setup.glsl:274-340 */
static void store_tri(frame *f, uint32_t tri_idx, uint32_t flags_id, v3 tri0, v3 tri1, v3 tri2,
					  uint32_t y_aabb) {
	v3 normal = v3_cross(v3_sub(tri0, tri2), v3_sub(tri1, tri0));
	float multiplier = 1.0f / sqrtf(v3_dot(normal, normal));
	normal = v3_mul(normal, multiplier);
	if((flags_id & LUCID_INST_HAS_VERTEX_NORMALS) == 0)
		f->tri_normal[tri_idx] = oracle_encode_normal_uint(normal.x, normal.y, normal.z);

	v3 edge0 = v3_mul(v3_sub(tri0, tri2), multiplier);
	v3 edge1 = v3_mul(v3_sub(tri1, tri0), multiplier);
	float plane_dist = v3_dot(normal, tri0);
	v3 nrm_tri0 = v3_cross(tri0, normal);
	float param0 = v3_dot(edge0, nrm_tri0);
	float param1 = v3_dot(edge1, nrm_tri0);
	edge0 = v3_cross(normal, edge0);
	edge1 = v3_cross(normal, edge1);
	edge0 = v3_make(v3_dot(edge0, f->dirx), v3_dot(edge0, f->diry), v3_dot(edge0, f->ray_dir0));
	edge1 = v3_make(v3_dot(edge1, f->dirx), v3_dot(edge1, f->diry), v3_dot(edge1, f->ray_dir0));

	v3 pnormal = v3_mul(normal, 1.0f / plane_dist);
	v3 depth_eq =
		v3_make(v3_dot(pnormal, f->dirx), v3_dot(pnormal, f->diry), v3_dot(pnormal, f->ray_dir0));

	uint32_t *d = f->depth + (size_t)tri_idx * 4;
	d[0] = f2bits(depth_eq.x), d[1] = f2bits(depth_eq.y), d[2] = f2bits(depth_eq.z);
	d[3] = flags_id;
	uint32_t *b = f->bary + (size_t)tri_idx * 8;
	b[0] = f2bits(edge0.x), b[1] = f2bits(edge0.y), b[2] = f2bits(edge0.z), b[3] = f2bits(param0);
	b[4] = f2bits(edge1.x), b[5] = f2bits(edge1.y), b[6] = f2bits(edge1.z), b[7] = f2bits(param1);

	v3 nrm0 = v3_cross(tri2, v3_sub(tri1, tri2));
	v3 nrm1 = v3_cross(tri0, v3_sub(tri2, tri0));
	v3 nrm2 = v3_cross(tri1, v3_sub(tri0, tri1));
	float volume = v3_dot(tri0, nrm0);
	if(volume < 0.0f)
		nrm0 = v3_neg(nrm0), nrm1 = v3_neg(nrm1), nrm2 = v3_neg(nrm2);

	v3 nrms[3] = {nrm0, nrm1, nrm2};
	v3 edges[3];
	for(int i = 0; i < 3; i++)
		edges[i] = v3_make(v3_dot(nrms[i], f->dirx), v3_dot(nrms[i], f->diry),
						   v3_dot(nrms[i], f->dir0));
	/* Deviation from the reference: an edge parallel to the pixel rows (edges.x == +-0)
	 * makes the reference's x-intercept form inf - inf = NaN, whose min/max in
	 * rasterBinStep is undefined in GLSL (and drops the edge where NaN is ignored).
	 * It is tilted to edges.x = max(|edges.y| 2^-40, 2^-100) so it keeps bounding rows. */
	for(int i = 0; i < 3; i++)
		if(edges[i].x == 0.0f)
			edges[i].x = fmaxf(fabsf(edges[i].y) * 0x1p-40f, 0x1p-100f);
	float inv_ex[3] = {1.0f / edges[0].x, 1.0f / edges[1].x, 1.0f / edges[2].x};
	float scan_base[3], scan_step[3], scan[3];
	for(int i = 0; i < 3; i++) {
		scan_base[i] = -(edges[i].z * inv_ex[i]);
		scan_step[i] = -(edges[i].y * inv_ex[i]);
	}
	uint32_t x_signs = (edges[0].x < 0.0f ? 1u : 0u) | (edges[1].x < 0.0f ? 2u : 0u) |
					   (edges[2].x < 0.0f ? 4u : 0u);
	uint32_t y_signs = (edges[0].y < 0.0f ? 8u : 0u) | (edges[1].y < 0.0f ? 16u : 0u) |
					   (edges[2].y < 0.0f ? 32u : 0u);
	/* start = (-0.5, 0.5): scan = scan_step * start.y + scan_base - start.x */
	for(int i = 0; i < 3; i++)
		scan[i] = (scan_step[i] * 0.5f + scan_base[i]) - (-0.5f);

	uint32_t *s = f->scan + (size_t)tri_idx * 8;
	s[0] = f2bits(scan[0]), s[1] = f2bits(scan[1]), s[2] = f2bits(scan[2]), s[3] = y_aabb;
	s[4] = f2bits(scan_step[0]), s[5] = f2bits(scan_step[1]), s[6] = f2bits(scan_step[2]);
	s[7] = x_signs | y_signs;
}

/* quad_setup.glsl:136-254 + storeQuad/addVisibleTri (:256-372). Returns 1 if small, 2 if
 * large, 0 if rejected. */
static int process_quad(frame *f, uint32_t quad_id, uint32_t inst_id, const uint32_t vi[4],
						uint32_t flags) {
	const lucid_config *cfg = f->in->config;
	uint32_t v0 = vi[0], v1 = vi[1], v2 = vi[2], v3i = vi[3];
	int cull0 = v0 == v1 || v1 == v2 || v2 == v0;
	int cull1 = v0 == v2 || v2 == v3i || v3i == v0;
	if(cull0 && cull1) {
		f->info.num_rejected_quads[LUCID_REJECTION_TYPE_OTHER]++;
		return 0;
	}
	v3 vws[4] = {vertex_load(f, v0), vertex_load(f, v1), vertex_load(f, v2), vertex_load(f, v3i)};
	cull0 = cull0 || v3_eq(vws[0], vws[1]) || v3_eq(vws[1], vws[2]) || v3_eq(vws[2], vws[0]);
	cull1 = cull1 || v3_eq(vws[0], vws[2]) || v3_eq(vws[2], vws[3]) || v3_eq(vws[3], vws[0]);

	if(cfg->enable_backface_culling != 0) {
		v3 o = f->origin0;
		v3 p0 = v3_sub(vws[0], o), p1 = v3_sub(vws[1], o), p2 = v3_sub(vws[2], o),
		   p3 = v3_sub(vws[3], o);
		v3 nrm0 = v3_cross(p2, v3_sub(p1, p2));
		v3 nrm1 = v3_cross(p3, v3_sub(p2, p3));
		float volume0 = v3_dot(p0, nrm0);
		float volume1 = v3_dot(p0, nrm1);
		cull0 = cull0 || volume0 <= 0.0f;
		cull1 = cull1 || volume1 <= 0.0f;
		if(cull0 && cull1) {
			f->info.num_rejected_quads[LUCID_REJECTION_TYPE_BACKFACE]++;
			return 0;
		}
	}
	uint32_t cull_flags = (cull0 ? 1u : 0u) | (cull1 ? 2u : 0u);

	v4 vndc[4];
	for(int i = 0; i < 4; i++)
		vndc[i] = mat_mul_point(&cfg->view_proj_matrix, vws[i]);
	uint32_t clipmask = vertex_clip_mask(vndc[0]) | (vertex_clip_mask(vndc[1]) << 8) |
						(vertex_clip_mask(vndc[2]) << 16) | (vertex_clip_mask(vndc[3]) << 24);
	uint32_t and_clipmask =
		clipmask & (clipmask >> 8) & (clipmask >> 16) & (clipmask >> 24) & 0xffu;
	uint32_t or_clipmask = clipmask | (clipmask >> 8) | (clipmask >> 16) | (clipmask >> 24);
	if(and_clipmask != 0) {
		f->info.num_rejected_quads[LUCID_REJECTION_TYPE_FRUSTUM]++;
		return 0;
	}

	float aabb0[4], aabb1[4];
	float inv_w[4] = {1.0f / vndc[0].w, 1.0f / vndc[1].w, 1.0f / vndc[2].w, 1.0f / vndc[3].w};
	if((or_clipmask & 0x30u) != 0) {
		v4 t0[3] = {vndc[0], vndc[1], vndc[2]};
		float iw0[3] = {inv_w[0], inv_w[1], inv_w[2]};
		clipped_aabb(t0, iw0, clipmask, aabb0);
		v4 t1[3] = {vndc[0], vndc[2], vndc[3]};
		float iw1[3] = {inv_w[0], inv_w[2], inv_w[3]};
		clipped_aabb(t1, iw1, (clipmask & 0xffu) | ((clipmask & 0xffff0000u) >> 8), aabb1);
	}
	v3 ndc[4];
	for(int i = 0; i < 4; i++)
		ndc[i] = v3_make(vndc[i].x * inv_w[i], vndc[i].y * inv_w[i], vndc[i].z * inv_w[i]);
	if((or_clipmask & 0x30u) == 0) {
		plain_aabb(ndc[0], ndc[1], ndc[2], aabb0);
		plain_aabb(ndc[0], ndc[2], ndc[3], aabb1);
	}

	float sx = (float)f->width * 0.5f, sy = (float)f->height * 0.5f;
	float mul[4] = {sx, sy, sx, sy};
	float aabb[4];
	for(int i = 0; i < 4; i++) {
		aabb0[i] = (aabb0[i] + 1.0f) * mul[i];
		aabb1[i] = (aabb1[i] + 1.0f) * mul[i];
	}
	aabb[0] = fmn(aabb0[0], aabb1[0]);
	aabb[1] = fmn(aabb0[1], aabb1[1]);
	aabb[2] = fmx(aabb0[2], aabb1[2]);
	aabb[3] = fmx(aabb0[3], aabb1[3]);

	if(ceilf(aabb[0] - 0.5001f) == floorf(aabb[2] + 0.5001f) ||
	   ceilf(aabb[1] - 0.5001f) == floorf(aabb[3] + 0.5001f)) {
		f->info.num_rejected_quads[LUCID_REJECTION_TYPE_BETWEEN_SAMPLES]++;
		return 0;
	}

	float maxp[4] = {(float)(f->width - 1), (float)(f->height - 1), (float)(f->width - 1),
					 (float)(f->height - 1)};
	float inset[4] = {0.49f, 0.49f, -0.49f, -0.49f};
	for(int i = 0; i < 4; i++) {
		aabb0[i] = clampf(aabb0[i] + inset[i], 0.0f, maxp[i]);
		aabb1[i] = clampf(aabb1[i] + inset[i], 0.0f, maxp[i]);
		aabb[i] = clampf(aabb[i] + inset[i], 0.0f, maxp[i]);
	}
	uint32_t bin_aabb[4];
	for(int i = 0; i < 4; i++)
		bin_aabb[i] = oracle_f2u(aabb[i]) >> BIN_SHIFT;
	uint32_t enc = oracle_encode_aabb28(bin_aabb[0], bin_aabb[1], bin_aabb[2], bin_aabb[3]) |
				   (cull_flags << 30);
	uint32_t bw = bin_aabb[2] - bin_aabb[0] + 1u, bh = bin_aabb[3] - bin_aabb[1] + 1u;
	int large = (bw * bh) <= 4u ? 0 : 1;

	uint32_t y_aabb[2] = {oracle_f2u(aabb0[1]) | (oracle_f2u(aabb0[3]) << 16),
						  oracle_f2u(aabb1[1]) | (oracle_f2u(aabb1[3]) << 16)};
	f->quad_aabb[quad_id] = enc;
	f->quad_inst[quad_id] = inst_id;

	/* addVisibleTri (:356-372) */
	uint32_t flags_id = flags | (inst_id << 16);
	for(int second = 0; second < 2; second++) {
		if((cull_flags >> second) & 1u)
			continue;
		v3 t0 = v3_sub(vertex_load(f, v0), f->origin0);
		v3 t1 = v3_sub(vertex_load(f, vi[1 + second]), f->origin0);
		v3 t2 = v3_sub(vertex_load(f, vi[2 + second]), f->origin0);
		store_tri(f, quad_id * 2 + second, flags_id, t0, t1, t2, y_aabb[second]);
	}
	/* storeQuad (:256-272): texture coordinates relative to corner 0, vertex colours / normals */
	if((flags & LUCID_INST_HAS_ALBEDO_TEXTURE) && f->in->tex_coords) {
		const float *tc = f->in->tex_coords;
		for(int second = 0; second < 2; second++) {
			float *o = f->tri_tex + ((size_t)quad_id * 2 + second) * 6;
			uint32_t w0 = vi[0], w1 = vi[1 + second], w2 = vi[2 + second];
			float t0x = tc[2 * (size_t)w0], t0y = tc[2 * (size_t)w0 + 1];
			o[0] = t0x, o[1] = t0y;
			o[2] = tc[2 * (size_t)w1] - t0x, o[3] = tc[2 * (size_t)w1 + 1] - t0y;
			o[4] = tc[2 * (size_t)w2] - t0x, o[5] = tc[2 * (size_t)w2 + 1] - t0y;
		}
	}
	if((flags & LUCID_INST_HAS_VERTEX_NORMALS) && f->in->normals) {
		uint32_t *nq = f->quad_normal + (size_t)quad_id * 4;
		for(int k = 0; k < 4; k++)
			nq[k] = f->in->normals[vi[k]];
	}
	if(flags & LUCID_INST_HAS_VERTEX_COLORS) {
		if(f->in->colors_rgba8) {
			uint32_t *c = f->quad_color + (size_t)quad_id * 4;
			for(int k = 0; k < 4; k++)
				c[k] = f->in->colors_rgba8[vi[k]];
		} else if(f->in->colors_rgb) {
			for(int second = 0; second < 2; second++) {
				float *c = f->tri_color + ((size_t)quad_id * 2 + second) * 12;
				uint32_t tv[3] = {vi[0], vi[1 + second], vi[2 + second]};
				for(int k = 0; k < 3; k++) {
					const float *src = f->in->colors_rgb + (size_t)tv[k] * 3;
					c[k * 4 + 0] = src[0], c[k * 4 + 1] = src[1], c[k * 4 + 2] = src[2];
					c[k * 4 + 3] = 0.0f;
				}
				c[3] = f->in->tri_alpha ? f->in->tri_alpha[quad_id * 2 + second] : 1.0f;
			}
		}
	}
	return large ? 2 : 1;
}

/* ---------------------------------------------------------------- binning */

/* scanline.glsl:28-52 + bin_counter.glsl:53-62 */
typedef struct {
	float mn[3], mx[3], step[3];
} scan_params;

static void load_scan_bin(const uint32_t *s, int *min_by, int *max_by, scan_params *p) {
	float scan[3] = {u2f(s[0]), u2f(s[1]), u2f(s[2])};
	float step[3] = {u2f(s[4]), u2f(s[5]), u2f(s[6])};
	uint32_t w = s[3], signs = s[7];
	*min_by = (int)(w & 0xffffu) >> BIN_SHIFT;
	*max_by = (int)(w >> 16) >> BIN_SHIFT;
	float offset = (float)BIN_SIZE - 0.989f;
	float sy = (float)(*min_by * BIN_SIZE) - 0.01f;
	float sx = 0.99f;
	for(int i = 0; i < 3; i++) {
		int xneg = (signs >> i) & 1, yneg = (signs >> (3 + i)) & 1;
		float yoff = yneg ? 0.0f : offset, xoff = xneg ? 0.0f : offset;
		scan[i] = scan[i] + (step[i] * (yoff + sy) - (xoff + sx));
		p->mn[i] = xneg ? -INFINITY : scan[i];
		p->mx[i] = xneg ? scan[i] : INFINITY;
		p->step[i] = step[i] * (float)BIN_SIZE;
	}
}

static void scanline_step(scan_params *p, int *bmin, int *bmax) {
	float xmin = fmx(fmx(p->mn[0], p->mn[1]), p->mn[2]);
	float xmax = fmn(fmn(p->mx[0], p->mx[1]), p->mx[2]);
	for(int i = 0; i < 3; i++)
		p->mn[i] = p->mn[i] + p->step[i], p->mx[i] = p->mx[i] + p->step[i];
	*bmin = oracle_f2i(xmin + 1.0f) >> BIN_SHIFT;
	*bmax = oracle_f2i(xmax) >> BIN_SHIFT;
}

typedef void (*bin_visit_fn)(frame *f, int bin, uint32_t value, int is_tri, void *ctx);

/* Visits every (bin, entry) pair of one visible quad (bin_counter.glsl:64-134,
 * bin_dispatcher.glsl:66-196). */
static void visit_quad_bins(frame *f, uint32_t q, int large, bin_visit_fn fn, void *ctx) {
	uint32_t enc = f->quad_aabb[q];
	int bsx = (int)(enc & 0x7fu), bsy = (int)((enc >> 7) & 0x7fu);
	int bex = (int)((enc >> 14) & 0x7fu), bey = (int)((enc >> 21) & 0x7fu);
	if(!large) {
		uint32_t entry = q | (enc & 0xf0000000u);
		for(int by = bsy; by <= bey; by++)
			for(int bx = bsx; bx <= bex; bx++)
				fn(f, bx + by * f->bin_count_x, entry, 0, ctx);
		return;
	}
	for(int second = 0; second < 2; second++) {
		if((enc >> (30 + second)) & 1u)
			continue;
		uint32_t t = q * 2 + second;
		int tsy, tey;
		scan_params p;
		load_scan_bin(f->scan + (size_t)t * 8, &tsy, &tey, &p);
		for(int by = tsy; by <= tey; by++) {
			int bmin, bmax;
			scanline_step(&p, &bmin, &bmax);
			if(bmin < bsx)
				bmin = bsx;
			if(bmax > bex)
				bmax = bex;
			for(int bx = bmin; bx <= bmax; bx++)
				fn(f, bx + by * f->bin_count_x, t, 1, ctx);
		}
	}
}

static void count_visit(frame *f, int bin, uint32_t value, int is_tri, void *ctx) {
	(void)value, (void)ctx;
	f->counts[(is_tri ? LUCID_COUNTS_TRI_COUNTS : LUCID_COUNTS_QUAD_COUNTS) * f->bin_count + bin]++;
}

typedef struct {
	int overflow;
	size_t capacity;
} dispatch_ctx;

static void dispatch_visit(frame *f, int bin, uint32_t value, int is_tri, void *vctx) {
	dispatch_ctx *ctx = (dispatch_ctx *)vctx;
	uint32_t *temp =
		&f->counts[(is_tri ? LUCID_COUNTS_TRI_OFFSETS_TEMP : LUCID_COUNTS_QUAD_OFFSETS_TEMP) *
					   f->bin_count +
				   bin];
	uint32_t pos = (*temp)++;
	if(pos >= ctx->capacity) {
		ctx->overflow = 1;
		return;
	}
	(is_tri ? f->bin_tris : f->bin_quads)[pos] = value;
}

/* ---------------------------------------------------------------- raster: rows */

typedef struct {
	int8_t mn[32], mx[32]; /* per bin row; mn > mx means empty */
} tri_rows;

/* rasterBinStep (raster.glsl:116-140): one row */
static inline void raster_row(float mn[3], float mx[3], const float step[3], int *rmin,
							  int *rmax) {
	int a = oracle_f2i(fmx(fmx(mn[0], mn[1]), fmx(mn[2], 0.0f)));
	int b = oracle_f2i(fmn(fmn(mx[0], mx[1]), fmn(mx[2], (float)BIN_SIZE))) - 1;
	if(a > b)
		a = BIN_MASK, b = 0;
	for(int i = 0; i < 3; i++)
		mn[i] = mn[i] + step[i], mx[i] = mx[i] + step[i];
	*rmin = a, *rmax = b;
}

/* Row spans of one tri in one bin, rows grouped by `group` (8 for raster_low,
 * 4 for raster_high): raster_low.glsl:39-64, raster_high.glsl:54-90,
 * scanline.glsl:13-26. Returns a bitmask of groups with a non-empty bx_mask. */
static uint32_t tri_bin_rows(const frame *f, uint32_t t, int bin_x, int bin_y, int group,
							 tri_rows *rows, uint32_t group_bx[8]) {
	const uint32_t *s = f->scan + (size_t)t * 8;
	int shift = group == 8 ? 3 : 2;
	int min_g = clampi((int)(s[3] & 0xffffu) - bin_y, 0, BIN_MASK) >> shift;
	int max_g = clampi((int)(s[3] >> 16) - bin_y, 0, BIN_MASK) >> shift;
	float startx = (float)bin_x, starty = (float)(bin_y + min_g * group);
	float mn[3], mx[3], step[3];
	uint32_t signs = s[7];
	for(int i = 0; i < 3; i++) {
		float scan = u2f(s[i]);
		step[i] = u2f(s[4 + i]);
		scan = scan + (step[i] * starty - startx);
		int xneg = (signs >> i) & 1;
		mn[i] = xneg ? -INFINITY : scan;
		mx[i] = xneg ? scan : INFINITY;
	}
	for(int r = 0; r < 32; r++)
		rows->mn[r] = 1, rows->mx[r] = 0;
	for(int g = 0; g < 8; g++)
		group_bx[g] = 0;
	uint32_t groups = 0;
	for(int g = min_g; g <= max_g; g++) {
		uint32_t bx_mask = 0;
		for(int r = 0; r < group; r++) {
			int a, b;
			raster_row(mn, mx, step, &a, &b);
			int row = g * group + r;
			rows->mn[row] = (int8_t)a, rows->mx[row] = (int8_t)b;
			bx_mask |= ((15u << (a >> 3)) & (15u >> (3 - (b >> 3)))) & 15u;
		}
		group_bx[g] = bx_mask;
		if(bx_mask)
			groups |= 1u << g;
		else
			for(int r = 0; r < group; r++)
				rows->mn[g * group + r] = 1, rows->mx[g * group + r] = 0;
	}
	for(int r = 0; r < 32; r++)
		if(rows->mn[r] == BIN_MASK && rows->mx[r] == 0)
			rows->mn[r] = 1, rows->mx[r] = 0;
	return groups;
}

/* ---------------------------------------------------------------- shading (shading.glsl:107-184) */

typedef struct {
	uint32_t enc;  /* RGBA8 encoded sample (0 = skipped), quantised mode */
	float col[4];  /* fp32 sample colour, fp32 mode */
	float inv_ray; /* depth: larger is nearer */
	float bary[2]; /* barycentrics of corners 1 and 2 */
	float light[3]; /* per-channel lighting factor (ambient + sun * light) */
} sample_t;

/* ---------------------------------------------------------------- textures
 *
 * textureGrad (shading.glsl:145-150) with linear magnification, minification and mip
 * filtering and repeat addressing (src/lucid_base.h:30), from the Vulkan sampling rules:
 * scale factors = coordinate derivatives x level-0 size, lambda = log2(max(|d/dx|, |d/dy|))
 * (isotropic), clamped to [0, levels - 1], the two nearest levels blended by its fraction;
 * bilinear at u = s * size - 0.5 with wrapped texel indices. Coordinates are reduced to
 * [0, 1] first (repeat is periodic). */
static void tex_bilinear(const oracle_texture *t, int level, float s, float v, float r[4]) {
	int w = t->width >> level, h = t->height >> level;
	w = w < 1 ? 1 : w;
	h = h < 1 ? 1 : h;
	size_t off = 0;
	for(int l = 0; l < level; l++) {
		int lw = t->width >> l, lh = t->height >> l;
		off += (size_t)(lw < 1 ? 1 : lw) * (size_t)(lh < 1 ? 1 : lh);
	}
	const uint32_t *tx = t->texels + off;
	float su = s - floorf(s), sv = v - floorf(v);
	float u = su * (float)w - 0.5f, uv = sv * (float)h - 0.5f;
	float fu = floorf(u), fv = floorf(uv);
	float au = u - fu, av = uv - fv;
	int x0 = (int)fu, y0 = (int)fv;
	x0 = x0 < 0 ? w - 1 : x0;
	y0 = y0 < 0 ? h - 1 : y0;
	int x1 = x0 + 1 == w ? 0 : x0 + 1, y1 = y0 + 1 == h ? 0 : y0 + 1;
	v4 c00 = decode_rgba8(tx[(size_t)y0 * w + x0]), c10 = decode_rgba8(tx[(size_t)y0 * w + x1]);
	v4 c01 = decode_rgba8(tx[(size_t)y1 * w + x0]), c11 = decode_rgba8(tx[(size_t)y1 * w + x1]);
	float bu = 1.0f - au, bv = 1.0f - av;
	r[0] = (c00.x * bu + c10.x * au) * bv + (c01.x * bu + c11.x * au) * av;
	r[1] = (c00.y * bu + c10.y * au) * bv + (c01.y * bu + c11.y * au) * av;
	r[2] = (c00.z * bu + c10.z * au) * bv + (c01.z * bu + c11.z * au) * av;
	r[3] = (c00.w * bu + c10.w * au) * bv + (c01.w * bu + c11.w * au) * av;
}

void oracle_texture_grad(const oracle_texture *t, float s, float v, float dsx, float dvx, float dsy,
						 float dvy, float out[4]) {
	float fw = (float)t->width, fh = (float)t->height;
	float ux = dsx * fw, vx = dvx * fh, uy = dsy * fw, vy = dvy * fh;
	float rho = fmx(sqrtf(ux * ux + vx * vx), sqrtf(uy * uy + vy * vy));
	float lambda = log2f(rho);
	if(!(lambda > 0.0f)) {
		tex_bilinear(t, 0, s, v, out);
		return;
	}
	float d = fmn(lambda, (float)(t->num_levels - 1));
	float fl = floorf(d);
	int lo = (int)fl;
	float delta = d - fl;
	tex_bilinear(t, lo, s, v, out);
	if(delta > 0.0f) {
		float h[4];
		int hi = lo + 1 < t->num_levels - 1 ? lo + 1 : t->num_levels - 1;
		tex_bilinear(t, hi, s, v, h);
		float nd = 1.0f - delta;
		for(int k = 0; k < 4; k++)
			out[k] = out[k] * nd + h[k] * delta;
	}
}

static void shade_sample(const frame *f, int px_i, int py_i, uint32_t t, sample_t *out) {
	const lucid_config *cfg = f->in->config;
	float px = (float)px_i, py = (float)py_i;
	const uint32_t *d = f->depth + (size_t)t * 4;
	const uint32_t *b = f->bary + (size_t)t * 8;
	float de[3] = {u2f(d[0]), u2f(d[1]), u2f(d[2])};
	uint32_t flags = d[3] & 0xffffu, inst = d[3] >> 16;
	float e0q[3] = {u2f(b[0]), u2f(b[1]), u2f(b[2])};
	float e1q[3] = {u2f(b[4]), u2f(b[5]), u2f(b[6])};
	float bp[2] = {u2f(b[3]), u2f(b[7])};

	float inv_ray_pos = de[0] * px + (de[1] * py + de[2]);
	out->inv_ray = inv_ray_pos;
	float ray_pos = 1.0f / inv_ray_pos;
	float e0 = e0q[0] * px + (e0q[1] * py + e0q[2]);
	float e1 = e1q[0] * px + (e1q[1] * py + e1q[2]);
	float bary0 = e0 * ray_pos, bary1 = e1 * ray_pos;
	/* shading.glsl:111-122: barycentrics one pixel right and one pixel down */
	int textured = (flags & LUCID_INST_HAS_ALBEDO_TEXTURE) && f->in->tex_coords;
	float bdx[2] = {0.0f, 0.0f}, bdy[2] = {0.0f, 0.0f};
	if(textured) {
		float ray_posx = 1.0f / (inv_ray_pos + de[0]);
		float ray_posy = 1.0f / (inv_ray_pos + de[1]);
		bdx[0] = (e0 + e0q[0]) * ray_posx - bary0;
		bdx[1] = (e1 + e1q[0]) * ray_posx - bary1;
		bdy[0] = (e0 + e0q[1]) * ray_posy - bary0;
		bdy[1] = (e1 + e1q[1]) * ray_posy - bary1;
	}
	bary0 = bary0 - bp[0];
	bary1 = bary1 - bp[1];
	out->bary[0] = bary0, out->bary[1] = bary1;
	out->light[0] = out->light[1] = out->light[2] = 0.0f;

	float color[4] = {1.0f, 1.0f, 1.0f, 1.0f};
	if((flags & LUCID_INST_HAS_COLOR) && f->in->instance_colors) {
		v4 ic = decode_rgba8(f->in->instance_colors[inst]);
		color[0] = ic.x, color[1] = ic.y, color[2] = ic.z, color[3] = ic.w;
	}
	if(textured) {
		/* shading.glsl:131-159 */
		const float *tt = f->tri_tex + (size_t)t * 6;
		float s = bary0 * tt[2] + (bary1 * tt[4] + tt[0]);
		float v = bary0 * tt[3] + (bary1 * tt[5] + tt[1]);
		float dsx = bdx[0] * tt[2] + bdx[1] * tt[4], dvx = bdx[0] * tt[3] + bdx[1] * tt[5];
		float dsy = bdy[0] * tt[2] + bdy[1] * tt[4], dvy = bdy[0] * tt[3] + bdy[1] * tt[5];
		if((flags & LUCID_INST_HAS_UV_RECT) && f->in->instance_uv_rects) {
			const float *rc = f->in->instance_uv_rects + (size_t)inst * 4;
			s = rc[2] * (s - floorf(s)) + rc[0];
			v = rc[3] * (v - floorf(v)) + rc[1];
			dsx = dsx * rc[2], dsy = dsy * rc[2];
			dvx = dvx * rc[3], dvy = dvy * rc[3];
		}
		int opaque = (flags & LUCID_INST_TEX_OPAQUE) != 0;
		float tc[4];
		oracle_texture_grad(&f->in->textures[opaque ? 0 : 1], s, v, dsx, dvx, dsy, dvy, tc);
		if(opaque)
			tc[3] = 1.0f;
		for(int k = 0; k < 4; k++)
			color[k] = color[k] * tc[k];
	}
	int fp32 = (f->in->opts & ORACLE_OPT_FP32_SAMPLES) != 0;
	if(flags & LUCID_INST_HAS_VERTEX_COLORS) {
		float w0 = (1.0f - bary0) - bary1;
		if(f->in->colors_rgba8) {
			const uint32_t *qc = f->quad_color + (size_t)(t >> 1) * 4;
			uint32_t s = t & 1;
			v4 c0 = decode_rgba8(qc[0]), c1 = decode_rgba8(qc[1 + s]), c2 = decode_rgba8(qc[2 + s]);
			color[0] = color[0] * (w0 * c0.x + (bary0 * c1.x + bary1 * c2.x));
			color[1] = color[1] * (w0 * c0.y + (bary0 * c1.y + bary1 * c2.y));
			color[2] = color[2] * (w0 * c0.z + (bary0 * c1.z + bary1 * c2.z));
			color[3] = color[3] * (w0 * c0.w + (bary0 * c1.w + bary1 * c2.w));
		} else {
			const float *c = f->tri_color + (size_t)t * 12;
			for(int k = 0; k < 3; k++)
				color[k] = color[k] * (w0 * c[k] + (bary0 * c[4 + k] + bary1 * c[8 + k]));
			color[3] = color[3] * c[3];
		}
	}
	if(color[3] == 0.0f) {
		out->enc = 0;
		out->col[0] = out->col[1] = out->col[2] = out->col[3] = 0.0f;
		return;
	}
	v3 normal;
	if(flags & LUCID_INST_HAS_VERTEX_NORMALS) {
		/* shading.glsl:173-178 */
		const uint32_t *nq = f->quad_normal + (size_t)(t >> 1) * 4;
		uint32_t s = t & 1;
		v3 n0 = decode_normal_uint(nq[0]), n1 = decode_normal_uint(nq[1 + s]),
		   n2 = decode_normal_uint(nq[2 + s]);
		n1 = v3_sub(n1, n0);
		n2 = v3_sub(n2, n0);
		normal = v3_make(bary0 * n1.x + (bary1 * n2.x + n0.x), bary0 * n1.y + (bary1 * n2.y + n0.y),
						 bary0 * n1.z + (bary1 * n2.z + n0.z));
	} else {
		normal = decode_normal_uint(f->tri_normal[t]);
	}
	v3 sun_dir = v4_xyz(cfg->lighting.sun_dir);
	float light_value = fmx(0.0f, v3_dot(v3_neg(sun_dir), normal) * 0.7f + 0.3f);
	/* finalShading (funcs.glsl:261-271) */
	for(int k = 0; k < 3; k++) {
		float amb_c = (&cfg->lighting.ambient_color.x)[k];
		float sun_c = (&cfg->lighting.sun_color.x)[k];
		float amb = amb_c * cfg->lighting.ambient_power;
		float dif = (sun_c * cfg->lighting.sun_power) * light_value;
		out->light[k] = amb + dif;
		float lin = srgb_to_linear(color[k]);
		color[k] = satf(linear_to_srgb(lin * (amb + dif)));
	}
	if(fp32) {
		out->enc = 1;
		for(int k = 0; k < 4; k++)
			out->col[k] = color[k];
	} else {
		out->enc = oracle_encode_rgba8(color[0], color[1], color[2], color[3]);
		v4 dc = decode_rgba8(out->enc);
		out->col[0] = dc.x, out->col[1] = dc.y, out->col[2] = dc.z, out->col[3] = dc.w;
	}
}

/* ---------------------------------------------------------------- blending */

typedef struct {
	uint32_t near_key; /* ~ord(inv_ray): smaller is nearer */
	uint32_t tri;
	uint32_t enc;
	uint32_t pix; /* pixel id inside the bin: row * 32 + x */
	float col[4];
	float inv_ray, bary[2], light[3];
} frag_t;

static inline uint32_t ord_f32(float f) {
	uint32_t u = f2bits(f);
	return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

static int frag_cmp(const void *pa, const void *pb) {
	const frag_t *a = (const frag_t *)pa, *b = (const frag_t *)pb;
	if(a->near_key != b->near_key)
		return a->near_key < b->near_key ? -1 : 1;
	return a->tri < b->tri ? -1 : a->tri > b->tri ? 1 : 0;
}

typedef struct {
	float c[3];
	float t;
} blend_state;

static inline void blend_one(blend_state *s, const float col[4], int additive) {
	if(additive) {
		for(int k = 0; k < 3; k++)
			s->c[k] = s->c[k] + col[k] * col[3];
	} else {
		for(int k = 0; k < 3; k++)
			s->c[k] = s->c[k] + (col[k] * col[3]) * s->t;
		s->t = s->t * (1.0f - col[3]);
	}
}

static void write_pixel(const frame *f, oracle_outputs *out, int px, int py, blend_state *s) {
	const lucid_config *cfg = f->in->config;
	float bg[3] = {cfg->background_color.x, cfg->background_color.y, cfg->background_color.z};
	float c[3];
	for(int k = 0; k < 3; k++)
		c[k] = satf(s->c[k] + s->t * bg[k]);
	size_t o = ((size_t)py * f->width + px) * 4;
	for(int k = 0; k < 3; k++)
		out->image[o + k] = (uint8_t)oracle_f2u(c[k] * 255.0f + 0.5f);
	out->image[o + 3] = 255;
	if(out->image_f) {
		for(int k = 0; k < 3; k++)
			out->image_f[o + k] = c[k];
		out->image_f[o + 3] = 1.0f;
	}
}

/* REF3: the reference's 3-deep reduction window (shading.glsl:186-314). */
typedef struct {
	float depth[3];
	uint32_t color[3];
	float out_trans;
	float out_color[3];
} ref3_ctx;

static void ref3_init(ref3_ctx *c) {
	for(int i = 0; i < 3; i++)
		c->depth[i] = 999999999.0f, c->color[i] = 0;
	c->out_trans = 1.0f;
	c->out_color[0] = c->out_color[1] = c->out_color[2] = 0.0f;
}

static void ref3_swap(ref3_ctx *c, int i, int j) {
	float d = c->depth[i];
	c->depth[i] = c->depth[j], c->depth[j] = d;
	uint32_t k = c->color[i];
	c->color[i] = c->color[j], c->color[j] = k;
}

static void ref3_blend(ref3_ctx *c, uint32_t enc, int additive) {
	v4 cc = decode_rgba8(enc);
	float col[4] = {cc.x, cc.y, cc.z, cc.w};
	blend_state s = {{c->out_color[0], c->out_color[1], c->out_color[2]}, c->out_trans};
	blend_one(&s, col, additive);
	for(int k = 0; k < 3; k++)
		c->out_color[k] = s.c[k];
	c->out_trans = s.t;
}

static void ref3_reduce(ref3_ctx *c, uint32_t color, float depth, int additive) {
	if(depth > c->depth[0]) {
		uint32_t tc = color;
		color = c->color[0], c->color[0] = tc;
		float td = depth;
		depth = c->depth[0], c->depth[0] = td;
		if(c->depth[0] > c->depth[1]) {
			ref3_swap(c, 0, 1);
			if(c->depth[1] > c->depth[2])
				ref3_swap(c, 1, 2);
		}
	}
	c->depth[2] = c->depth[1], c->depth[1] = c->depth[0], c->depth[0] = depth;
	if(c->color[2] != 0)
		ref3_blend(c, c->color[2], additive);
	c->color[2] = c->color[1], c->color[1] = c->color[0], c->color[0] = color;
}

static void ref3_finish(ref3_ctx *c, blend_state *s, int additive) {
	for(int i = 2; i >= 0; i--)
		if(c->color[i] != 0)
			ref3_blend(c, c->color[i], additive);
	s->c[0] = c->out_color[0], s->c[1] = c->out_color[1], s->c[2] = c->out_color[2];
	s->t = c->out_trans;
}

/* ---------------------------------------------------------------- raster: one bin */

typedef struct {
	uint32_t *tris;
	tri_rows *rows;
	uint32_t (*gbx)[8];
	frag_t *frags;
} bin_scratch;

static int ensure(void **ptr, size_t *cap, size_t need, size_t elem) {
	if(need <= *cap)
		return 0;
	size_t n = *cap ? *cap : 1024;
	while(n < need)
		n *= 2;
	void *p = realloc(*ptr, n * elem);
	if(!p)
		return -1;
	*ptr = p;
	*cap = n;
	return 0;
}

typedef struct {
	uint64_t frags;
	uint32_t ref_errors;
	int level;
} bin_result;

static inline float inverse_sqrt(float x) { return 1.0f / sqrtf(x); }

/* rasterHalfBlockCentroid (raster.glsl:142-150) over rows r0..r0+3 */
static void hblock_centroid(const tri_rows *rw, int r0, int startx, float *cx, float *cy,
							uint32_t *num) {
	float cpx[4], cpy[4];
	static const float ys[4] = {1.0f, 3.0f, 5.0f, 7.0f};
	uint32_t n = 0;
	for(int i = 0; i < 4; i++) {
		int mn = rw->mn[r0 + i], mx = rw->mx[r0 + i];
		if(mn > mx) /* empty rows are encoded (31, 0) in the reference */
			mn = BIN_MASK, mx = 0;
		int xmin = mn - startx;
		if(xmin < 0)
			xmin = 0;
		int xmax = mx - startx;
		if(xmax > 7)
			xmax = 7;
		int count = xmax - xmin + 1;
		if(count < 0)
			count = 0;
		cpx[i] = (float)(xmin * 2 + count) * (float)count;
		cpy[i] = ys[i] * (float)count;
		n += (uint32_t)count;
	}
	*cx = ((cpx[0] + cpx[1]) + cpx[2]) + cpx[3];
	*cy = ((cpy[0] + cpy[1]) + cpy[2]) + cpy[3];
	*num = n;
}

static uint32_t block_depth(const frame *f, float cx, float cy, uint32_t t, float range) {
	const uint32_t *d = f->depth + (size_t)t * 4;
	float ray_pos = u2f(d[0]) * cx + (u2f(d[1]) * cy + u2f(d[2]));
	float depth = range * satf(inverse_sqrt(ray_pos + 1.0f));
	return oracle_f2u(depth);
}

static int cmp_u64(const void *a, const void *b) {
	uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
	return x < y ? -1 : x > y ? 1 : 0;
}

/* REF3 ordering of one half-block stream and its reduction. `order` lists bin-tri
 * indices (into sc->tris) in the reference's processing order. */
static void ref3_hblock(const frame *f, oracle_outputs *out, int bin_x, int bin_y, int hbx,
						int hby, const uint32_t *order, int count, bin_scratch *sc, int red) {
	int additive = (f->in->opts & ORACLE_OPT_ADDITIVE) != 0;
	ref3_ctx ctx[32];
	for(int p = 0; p < 32; p++)
		ref3_init(&ctx[p]);
	uint32_t sample_idx = 0;
	for(int i = 0; i < count; i++) {
		int li = (int)order[i];
		const tri_rows *rw = &sc->rows[li];
		uint32_t t = sc->tris[li];
		for(int ry = 0; ry < 4; ry++) {
			int row = hby * 4 + ry;
			int mn = rw->mn[row], mx = rw->mx[row];
			for(int x = mn; x <= mx; x++) {
				if(x < hbx * 8 || x >= hbx * 8 + 8)
					continue;
				int px = bin_x + x, py = bin_y + row;
				int pid = (x - hbx * 8) + ry * 8;
				sample_t s;
				shade_sample(f, px, py, t, &s);
				ref3_reduce(&ctx[pid], s.enc, s.inv_ray, additive);
				sample_idx++;
				if(sample_idx % SEGMENT_SIZE == 0)
					for(int p = 0; p < 32; p++)
						for(int k = 0; k < 3; k++)
							ctx[p].out_color[k] = satf(ctx[p].out_color[k]);
			}
		}
	}
	if(sample_idx % SEGMENT_SIZE != 0)
		for(int p = 0; p < 32; p++)
			for(int k = 0; k < 3; k++)
				ctx[p].out_color[k] = satf(ctx[p].out_color[k]);
	for(int p = 0; p < 32; p++) {
		int px = bin_x + hbx * 8 + (p & 7), py = bin_y + hby * 4 + (p >> 3);
		if(px >= f->width || py >= f->height)
			continue;
		blend_state s;
		ref3_finish(&ctx[p], &s, additive);
		if(red) {
			size_t o = ((size_t)py * f->width + px) * 4;
			out->image[o + 0] = 255, out->image[o + 1] = 0, out->image[o + 2] = 0;
			out->image[o + 3] = 0;
			if(out->image_f)
				out->image_f[o + 0] = 1.0f, out->image_f[o + 1] = out->image_f[o + 2] = 0.0f,
								 out->image_f[o + 3] = 0.0f;
			continue;
		}
		write_pixel(f, out, px, py, &s);
	}
}

static int raster_bin(const frame *f, oracle_outputs *out, int bin, int level, bin_scratch *sc,
					  size_t *tri_cap, size_t *rows_cap, size_t *frag_cap, bin_result *res) {
	int bin_x = (bin % f->bin_count_x) * BIN_SIZE, bin_y = (bin / f->bin_count_x) * BIN_SIZE;
	uint32_t qc = f->counts[LUCID_COUNTS_QUAD_COUNTS * f->bin_count + bin];
	uint32_t qo = f->counts[LUCID_COUNTS_QUAD_OFFSETS * f->bin_count + bin];
	uint32_t tc = f->counts[LUCID_COUNTS_TRI_COUNTS * f->bin_count + bin];
	uint32_t to = f->counts[LUCID_COUNTS_TRI_OFFSETS * f->bin_count + bin];
	size_t ntri = 0;
	if(ensure((void **)&sc->tris, tri_cap, (size_t)qc * 2 + tc + 1, sizeof(uint32_t)))
		return -1;
	for(uint32_t i = 0; i < qc; i++) {
		uint32_t e = f->bin_quads[qo + i], q = e & 0x0fffffffu;
		for(int s = 0; s < 2; s++)
			if(!((e >> (30 + s)) & 1u))
				sc->tris[ntri++] = q * 2 + (uint32_t)s;
	}
	for(uint32_t i = 0; i < tc; i++)
		sc->tris[ntri++] = f->bin_tris[to + i];

	size_t cap_before = *rows_cap;
	if(ensure((void **)&sc->rows, rows_cap, ntri + 1, sizeof(tri_rows)))
		return -1;
	if(*rows_cap != cap_before || !sc->gbx) {
		void *p = realloc(sc->gbx, (*rows_cap) * sizeof(uint32_t[8]));
		if(!p)
			return -1;
		sc->gbx = (uint32_t(*)[8])p;
	}

	/* raster_low: 8-row groups; promotion if an 8x8 block has more than 256 tris */
	int group = 8;
	if(level == LUCID_BIN_LEVEL_LOW) {
		uint32_t block_counts[16] = {0};
		for(size_t i = 0; i < ntri; i++) {
			uint32_t groups = tri_bin_rows(f, sc->tris[i], bin_x, bin_y, 8, &sc->rows[i], sc->gbx[i]);
			for(int by = 0; by < 4; by++)
				if(groups & (1u << by))
					for(int bx = 0; bx < 4; bx++)
						if(sc->gbx[i][by] & (1u << bx))
							block_counts[by * 4 + bx]++;
		}
		for(int b = 0; b < 16; b++)
			if(block_counts[b] > MAX_BLOCK_TRIS)
				level = LUCID_BIN_LEVEL_HIGH;
	}
	if(level == LUCID_BIN_LEVEL_HIGH) {
		group = 4;
		for(size_t i = 0; i < ntri; i++)
			tri_bin_rows(f, sc->tris[i], bin_x, bin_y, 4, &sc->rows[i], sc->gbx[i]);
	}
	res->level = level;

	/* fragments */
	size_t nfr = 0;
	for(size_t i = 0; i < ntri; i++)
		for(int r = 0; r < 32; r++)
			if(sc->rows[i].mn[r] <= sc->rows[i].mx[r])
				nfr += (size_t)(sc->rows[i].mx[r] - sc->rows[i].mn[r] + 1);
	if(ensure((void **)&sc->frags, frag_cap, nfr + 1, sizeof(frag_t)))
		return -1;

	int fp32 = (f->in->opts & ORACLE_OPT_FP32_SAMPLES) != 0;
	int additive = (f->in->opts & ORACLE_OPT_ADDITIVE) != 0;
	uint32_t counts[32 * 32];
	memset(counts, 0, sizeof(counts));
	size_t n = 0;
	for(size_t i = 0; i < ntri; i++) {
		uint32_t t = sc->tris[i];
		for(int r = 0; r < 32; r++) {
			int py = bin_y + r;
			for(int x = sc->rows[i].mn[r]; x <= sc->rows[i].mx[r]; x++) {
				int px = bin_x + x;
				if(px >= f->width || py >= f->height)
					continue;
				counts[r * 32 + x]++;
				if(f->in->mode != ORACLE_MODE_EXACT)
					continue;
				sample_t s;
				shade_sample(f, px, py, t, &s);
				frag_t *fr = &sc->frags[n++];
				fr->near_key = ~ord_f32(s.inv_ray);
				fr->tri = t;
				fr->enc = s.enc;
				fr->pix = (uint32_t)(r * 32 + x);
				memcpy(fr->col, s.col, sizeof(fr->col));
				fr->inv_ray = s.inv_ray;
				fr->bary[0] = s.bary[0], fr->bary[1] = s.bary[1];
				memcpy(fr->light, s.light, sizeof(fr->light));
			}
		}
	}
	uint64_t total = 0;
	for(int p = 0; p < 32 * 32; p++) {
		int px = bin_x + (p & 31), py = bin_y + (p >> 5);
		if(px >= f->width || py >= f->height)
			continue;
		out->frag_counts[(size_t)py * f->width + px] = counts[p];
		total += counts[p];
	}
	res->frags = total;

	if(f->in->mode == ORACLE_MODE_COUNTS)
		return 0;
	if(f->in->mode == ORACLE_MODE_EXACT) {
		/* bucket by pixel, then sort each pixel's list by (near_key, tri) */
		uint32_t start[32 * 32 + 1];
		start[0] = 0;
		for(int p = 0; p < 32 * 32; p++)
			start[p + 1] = start[p] + counts[p];
		frag_t *sorted = (frag_t *)malloc((n + 1) * sizeof(frag_t));
		if(!sorted)
			return -1;
		uint32_t fill[32 * 32];
		memcpy(fill, start, sizeof(fill));
		for(size_t i = 0; i < n; i++)
			sorted[fill[sc->frags[i].pix]++] = sc->frags[i];
		for(int p = 0; p < 32 * 32; p++) {
			int px = bin_x + (p & 31), py = bin_y + (p >> 5);
			if(px >= f->width || py >= f->height)
				continue;
			frag_t *lst = sorted + start[p];
			qsort(lst, counts[p], sizeof(frag_t), frag_cmp);
			const int dump_all = out->dump_x1 <= out->dump_x0;
			if(out->frags && (dump_all || (px >= out->dump_x0 && px < out->dump_x1 &&
										   py >= out->dump_y0 && py < out->dump_y1))) {
				size_t at;
#ifdef _OPENMP
#pragma omp atomic capture
#endif
				{
					at = out->num_frags_dumped;
					out->num_frags_dumped += counts[p];
				}
				for(uint32_t k = 0; k < counts[p] && at + k < out->frags_cap; k++) {
					oracle_fragment *d = &out->frags[at + k];
					d->x = (uint32_t)px, d->y = (uint32_t)py, d->tri = lst[k].tri;
					d->inv_ray = lst[k].inv_ray;
					d->bary[0] = lst[k].bary[0], d->bary[1] = lst[k].bary[1];
					memcpy(d->light, lst[k].light, sizeof(d->light));
					d->order = k;
				}
			}
			blend_state s = {{0.0f, 0.0f, 0.0f}, 1.0f};
			for(uint32_t k = 0; k < counts[p]; k++) {
				if(fp32 ? lst[k].col[3] == 0.0f : lst[k].enc == 0)
					continue;
				blend_one(&s, lst[k].col, additive);
			}
			write_pixel(f, out, px, py, &s);
		}
		free(sorted);
		return 0;
	}

	/* REF3: reproduce the reference's block/half-block ordering. */
	uint32_t *order = (uint32_t *)malloc((ntri + 1) * sizeof(uint32_t));
	uint32_t *row_idx = (uint32_t *)malloc((ntri + 1) * sizeof(uint32_t));
	uint64_t *keys = (uint64_t *)malloc((ntri + 1) * sizeof(uint64_t));
	if(!order || !row_idx || !keys) {
		free(order), free(row_idx), free(keys);
		return -1;
	}
	if(group == 8) {
		for(int by = 0; by < 4; by++) {
			/* row_tri_idx: rank among tris with a non-empty record in this block row */
			uint32_t rc = 0;
			for(size_t i = 0; i < ntri; i++)
				row_idx[i] = sc->gbx[i][by] != 0 ? rc++ : 0xffffffffu;
			for(int bx = 0; bx < 4; bx++) {
				int cnt = 0;
				for(size_t i = 0; i < ntri; i++)
					if(row_idx[i] != 0xffffffffu && (sc->gbx[i][by] & (1u << bx)))
						order[cnt++] = (uint32_t)i;
				if(cnt > 3) {
					float block_x = (float)(bin_x + bx * 8), block_y = (float)(bin_y + by * 8);
					for(int k = 0; k < cnt; k++) {
						const tri_rows *rw = &sc->rows[order[k]];
						float cx0, cy0, cx1, cy1;
						uint32_t n0, n1;
						hblock_centroid(rw, by * 8, bx * 8, &cx0, &cy0, &n0);
						hblock_centroid(rw, by * 8 + 4, bx * 8, &cx1, &cy1, &n1);
						float cx = cx0 + cx1, cy = cy0 + cy1;
						float scale = 0.5f / (float)(n0 + n1);
						cx = cx * scale + block_x, cy = cy * scale + block_y;
						uint32_t depth = block_depth(f, cx, cy, sc->tris[order[k]], (float)0x3ffffe);
						uint32_t key = row_idx[order[k]] | (depth << 10);
						keys[k] = ((uint64_t)key << 32) | order[k];
					}
					qsort(keys, (size_t)cnt, sizeof(uint64_t), cmp_u64);
					for(int k = 0; k < cnt; k++)
						order[k] = (uint32_t)(keys[k] & 0xffffffffu);
				}
				ref3_hblock(f, out, bin_x, bin_y, bx, by * 2, order, cnt, sc, 0);
				ref3_hblock(f, out, bin_x, bin_y, bx, by * 2 + 1, order, cnt, sc, 0);
			}
		}
	} else {
		/* raster_high: per half-block rows of 4, 18-bit key with overflow (raster_high.glsl:187-190) */
		int red = 0;
		for(int rby = 0; rby < 8; rby++) {
			uint32_t rc = 0;
			for(size_t i = 0; i < ntri; i++)
				row_idx[i] = sc->gbx[i][rby] != 0 ? rc++ : 0xffffffffu;
			if(rc > MAX_HBLOCK_ROW_TRIS)
				red = 1;
			for(int hbx = 0; hbx < 4; hbx++) {
				int cnt = 0;
				for(size_t i = 0; i < ntri; i++)
					if(row_idx[i] != 0xffffffffu && (sc->gbx[i][rby] & (1u << hbx)))
						cnt++;
				if(cnt > MAX_HBLOCK_TRIS)
					red = 1;
			}
		}
		if(red)
			res->ref_errors++;
		for(int rby = 0; rby < 8; rby++) {
			uint32_t rc = 0;
			for(size_t i = 0; i < ntri; i++)
				row_idx[i] = sc->gbx[i][rby] != 0 ? rc++ : 0xffffffffu;
			for(int hbx = 0; hbx < 4; hbx++) {
				int cnt = 0;
				for(size_t i = 0; i < ntri; i++)
					if(row_idx[i] != 0xffffffffu && (sc->gbx[i][rby] & (1u << hbx)))
						order[cnt++] = (uint32_t)i;
				float block_x = (float)(bin_x + hbx * 8), block_y = (float)(bin_y + rby * 4);
				for(int k = 0; k < cnt; k++) {
					const tri_rows *rw = &sc->rows[order[k]];
					float cx, cy;
					uint32_t nf;
					hblock_centroid(rw, rby * 4, hbx * 8, &cx, &cy, &nf);
					float scale = 0.5f / (float)nf;
					cx = cx * scale + block_x, cy = cy * scale + block_y;
					uint32_t depth = block_depth(f, cx, cy, sc->tris[order[k]], (float)0x7fffe);
					uint32_t key = row_idx[order[k]] | (depth << 14);
					keys[k] = ((uint64_t)key << 32) | order[k];
				}
				qsort(keys, (size_t)cnt, sizeof(uint64_t), cmp_u64);
				for(int k = 0; k < cnt; k++)
					order[k] = (uint32_t)(keys[k] & 0xffffffffu);
				ref3_hblock(f, out, bin_x, bin_y, hbx, rby, order, cnt, sc, red);
			}
		}
	}
	free(order), free(row_idx), free(keys);
	return 0;
}

/* ---------------------------------------------------------------- driver */

int oracle_num_threads(void) {
#ifdef _OPENMP
	return omp_get_max_threads();
#else
	return 1;
#endif
}

int oracle_render(const oracle_inputs *in, oracle_outputs *out) {
	if(!in || !out || !in->config || in->width <= 0 || in->height <= 0 || in->num_instances < 0)
		return -1;
	frame fr;
	memset(&fr, 0, sizeof(fr));
	frame *f = &fr;
	f->in = in;
	f->width = in->width, f->height = in->height;
	f->bin_count_x = (in->width + BIN_SIZE - 1) / BIN_SIZE;
	f->bin_count_y = (in->height + BIN_SIZE - 1) / BIN_SIZE;
	f->bin_count = f->bin_count_x * f->bin_count_y;

	int64_t *prefix = (int64_t *)malloc(sizeof(int64_t) * (size_t)(in->num_instances + 1));
	if(!prefix)
		return -2;
	prefix[0] = 0;
	for(int i = 0; i < in->num_instances; i++)
		prefix[i + 1] = prefix[i] + in->instances[i].num_quads;
	int64_t nq = prefix[in->num_instances];
	if(nq >= (1 << 28)) {
		free(prefix);
		return -1;
	}
	f->num_quads = (int)nq;

	const lucid_frustum *fru = &in->config->frustum;
	f->dir0 = v4_xyz(fru->ws_dir0);
	f->dirx = v4_xyz(fru->ws_dirx);
	f->diry = v4_xyz(fru->ws_diry);
	f->origin0 = v4_xyz(fru->ws_origin0);
	f->ray_dir0 = v3_add(f->dir0, v3_mul(v3_add(f->dirx, f->diry), 0.5f));

	size_t Q = (size_t)nq, T = Q * 2;
	f->quad_aabb = out->quad_aabb;
	f->quad_inst = (uint32_t *)calloc(Q + 1, sizeof(uint32_t));
	f->scan = (uint32_t *)calloc(T * 8 + 8, sizeof(uint32_t));
	f->depth = (uint32_t *)calloc(T * 4 + 4, sizeof(uint32_t));
	f->bary = (uint32_t *)calloc(T * 8 + 8, sizeof(uint32_t));
	f->tri_normal = out->tri_normals;
	f->quad_color = out->quad_colors;
	f->tri_color = (float *)calloc(T * 12 + 12, sizeof(float));
	f->quad_normal = (uint32_t *)calloc(Q * 4 + 4, sizeof(uint32_t));
	f->tri_tex = (float *)calloc(in->tex_coords ? T * 6 + 6 : 6, sizeof(float));
	f->counts = out->info + LUCID_INFO_U32_SIZE;
	f->bin_quads = out->bin_quads;
	f->bin_tris = out->bin_tris;
	if(!f->quad_inst || !f->scan || !f->depth || !f->bary || !f->tri_color || !f->quad_normal ||
	   !f->tri_tex) {
		free(prefix), free(f->quad_inst), free(f->scan), free(f->depth), free(f->bary),
			free(f->tri_color), free(f->quad_normal), free(f->tri_tex);
		return -2;
	}
	memset(out->info, 0, sizeof(uint32_t) * (LUCID_INFO_U32_SIZE + 10 * (size_t)f->bin_count));
	for(size_t q = 0; q < Q; q++)
		out->quad_aabb[q] = INVALID_AABB;

	/* ---- setup */
	int *kind = (int *)calloc(Q + 1, sizeof(int));
	for(int i = 0; i < in->num_instances; i++) {
		const lucid_instance_data *ins = &in->instances[i];
		for(int k = 0; k < ins->num_quads; k++) {
			uint32_t q = (uint32_t)(prefix[i] + k);
			const uint32_t *idx = in->quad_indices + (size_t)ins->index_offset + (size_t)k * 4;
			uint32_t vi[4];
			for(int j = 0; j < 4; j++)
				vi[j] = idx[j] + (uint32_t)ins->vertex_offset;
			kind[q] = process_quad(f, q, (uint32_t)i, vi, ins->flags);
			if(kind[q] == 1)
				f->info.num_visible_quads[0]++;
			else if(kind[q] == 2)
				f->info.num_visible_quads[1]++;
		}
	}
	f->info.num_input_quads = (int32_t)nq;

	/* ---- bin counts, offsets (bin_categorizer.glsl:22-53), lists */
	for(size_t q = 0; q < Q; q++)
		if(kind[q])
			visit_quad_bins(f, (uint32_t)q, kind[q] == 2, count_visit, NULL);
	uint32_t *cnt = f->counts;
	int NB = f->bin_count;
	uint32_t acc_q = 0, acc_t = 0;
	for(int b = 0; b < NB; b++) {
		cnt[LUCID_COUNTS_QUAD_OFFSETS * NB + b] = acc_q;
		cnt[LUCID_COUNTS_QUAD_OFFSETS_TEMP * NB + b] = acc_q;
		acc_q += cnt[LUCID_COUNTS_QUAD_COUNTS * NB + b];
		cnt[LUCID_COUNTS_TRI_OFFSETS * NB + b] = acc_t;
		cnt[LUCID_COUNTS_TRI_OFFSETS_TEMP * NB + b] = acc_t;
		acc_t += cnt[LUCID_COUNTS_TRI_COUNTS * NB + b];
	}
	int rc = 0;
	if(acc_q > out->bin_capacity || acc_t > out->bin_capacity) {
		rc = -3;
		goto done;
	}
	{
		dispatch_ctx dctx = {0, out->bin_capacity};
		for(size_t q = 0; q < Q; q++)
			if(kind[q])
				visit_quad_bins(f, (uint32_t)q, kind[q] == 2, dispatch_visit, &dctx);
	}
	/* categorizeBins (bin_categorizer.glsl:57-89) */
	{
		int nlow = 0, nhigh = 0, nempty = 0;
		for(int b = 0; b < NB; b++) {
			uint32_t nt = cnt[LUCID_COUNTS_TRI_COUNTS * NB + b] +
						  cnt[LUCID_COUNTS_QUAD_COUNTS * NB + b] * 2;
			if(nt == 0)
				nempty++;
			else if(nt < LOW_LEVEL_LIMIT)
				cnt[LUCID_COUNTS_LOW_BINS * NB + nlow++] = (uint32_t)b;
			else
				cnt[LUCID_COUNTS_HIGH_BINS * NB + nhigh++] = (uint32_t)b;
		}
		f->info.bin_level_counts[LUCID_BIN_LEVEL_EMPTY] = nempty;
		f->info.bin_level_counts[LUCID_BIN_LEVEL_LOW] = nlow;
		f->info.bin_level_counts[LUCID_BIN_LEVEL_HIGH] = nhigh;
	}

	/* ---- raster: background everywhere, then every non-empty bin */
	{
		blend_state bgs = {{0.0f, 0.0f, 0.0f}, 1.0f};
		for(int py = 0; py < f->height; py++)
			for(int px = 0; px < f->width; px++) {
				write_pixel(f, out, px, py, &bgs);
				out->frag_counts[(size_t)py * f->width + px] = 0;
			}
	}
	{
		int nlow = f->info.bin_level_counts[LUCID_BIN_LEVEL_LOW];
		int nhigh = f->info.bin_level_counts[LUCID_BIN_LEVEL_HIGH];
		int njobs = nlow + nhigh;
		uint64_t frags = 0;
		uint32_t ref_errors = 0, promoted = 0;
		int err = 0;
#ifdef _OPENMP
#pragma omp parallel reduction(+ : frags, ref_errors, promoted) reduction(| : err)
#endif
		{
			bin_scratch sc;
			memset(&sc, 0, sizeof(sc));
			size_t tri_cap = 0, rows_cap = 0, frag_cap = 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
			for(int j = 0; j < njobs; j++) {
				int level = j < nlow ? LUCID_BIN_LEVEL_LOW : LUCID_BIN_LEVEL_HIGH;
				int bin = (int)cnt[(j < nlow ? LUCID_COUNTS_LOW_BINS : LUCID_COUNTS_HIGH_BINS) * NB +
								   (j < nlow ? j : j - nlow)];
				bin_result r = {0, 0, level};
				if(raster_bin(f, out, bin, level, &sc, &tri_cap, &rows_cap, &frag_cap, &r))
					err |= 1;
				frags += r.frags;
				ref_errors += r.ref_errors;
				if(level == LUCID_BIN_LEVEL_LOW && r.level == LUCID_BIN_LEVEL_HIGH)
					promoted++;
				if(out->bin_levels)
					out->bin_levels[bin] = (uint32_t)r.level;
			}
			free(sc.tris), free(sc.rows), free(sc.gbx), free(sc.frags);
		}
		if(err) {
			rc = -2;
			goto done;
		}
		out->num_fragments = frags;
		out->num_ref_errors = ref_errors;
		f->info.stats[0] = (uint32_t)frags;
		/* promoted bins are appended to the high list (raster_low.glsl:229-237) */
		f->info.bin_level_counts[LUCID_BIN_LEVEL_HIGH] += (int32_t)promoted;
		if(promoted) {
			int k = nhigh;
			for(int j = 0; j < nlow; j++) {
				int bin = (int)cnt[LUCID_COUNTS_LOW_BINS * NB + j];
				if(out->bin_levels && out->bin_levels[bin] == LUCID_BIN_LEVEL_HIGH)
					cnt[LUCID_COUNTS_HIGH_BINS * NB + k++] = (uint32_t)bin;
			}
		}
	}

	/* ---- outputs: tri records in the reference's arena order */
	if(out->tri_records) {
		for(size_t t = 0; t < T; t++) {
			uint32_t *o = out->tri_records + t * 20;
			memcpy(o + 0, f->bary + t * 8, 8 * sizeof(uint32_t));
			memcpy(o + 8, f->scan + t * 8, 8 * sizeof(uint32_t));
			memcpy(o + 16, f->depth + t * 4, 4 * sizeof(uint32_t));
		}
	}
	memcpy(out->info, &f->info, sizeof(lucid_info));

done:
	free(kind);
	free(prefix);
	free(f->quad_inst), free(f->scan), free(f->depth), free(f->bary), free(f->tri_color);
	free(f->quad_normal);
	free(f->tri_tex);
	return rc;
}
