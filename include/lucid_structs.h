/*
 * Byte-exact C mirrors of the data layouts the Lucid compute rasterizer
 * exchanges between host and device.
 *
 * Reference layouts (GLSL std430/std140, shared with C++ through
 * src/shader_structs.h:5-9):
 *   InstanceData  data/shaders/shared/structures.glsl:9-14
 *   Lighting      data/shaders/shared/structures.glsl:16-21
 *   Frustum       data/shaders/shared/structures.glsl:23-28
 *   LucidInfo     data/shaders/shared/structures.glsl:64-99
 *   LucidConfig   data/shaders/shared/structures.glsl:101-111 (std140 UBO)
 *   constants     data/shaders/shared/definitions.glsl:72-137
 *
 * The structs are plain C so that the C ABI (lucid_c.h), the C oracle and
 * ctypes/numpy on the Python side all read the same bytes.
 */
#ifndef LUCID_STRUCTS_H
#define LUCID_STRUCTS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* definitions.glsl:72-91 */
#define LUCID_BIN_LEVELS_COUNT 5
#define LUCID_REJECTION_TYPE_COUNT 4
#define LUCID_TIMERS_COUNT 8
#define LUCID_STATS_COUNT 4

#define LUCID_BIN_LEVEL_EMPTY 0
#define LUCID_BIN_LEVEL_MICRO 1
#define LUCID_BIN_LEVEL_LOW 2
#define LUCID_BIN_LEVEL_MEDIUM 3
#define LUCID_BIN_LEVEL_HIGH 4

/* Instance flags; map 1:1 to DrawCallOpts (definitions.glsl:84-94, lucid_base.h:140-142) */
#define LUCID_INST_HAS_VERTEX_COLORS 0x001u
#define LUCID_INST_HAS_VERTEX_TEX_COORDS 0x002u
#define LUCID_INST_HAS_VERTEX_NORMALS 0x004u
#define LUCID_INST_IS_OPAQUE 0x008u
#define LUCID_INST_TEX_OPAQUE 0x010u
#define LUCID_INST_HAS_UV_RECT 0x020u
#define LUCID_INST_HAS_ALBEDO_TEXTURE 0x040u
#define LUCID_INST_HAS_NORMAL_TEXTURE 0x080u
#define LUCID_INST_HAS_PBR_TEXTURE 0x100u
#define LUCID_INST_HAS_COLOR 0x200u

/* Rejection reasons during quad setup (definitions.glsl:96-100) */
#define LUCID_REJECTION_TYPE_OTHER 0
#define LUCID_REJECTION_TYPE_BACKFACE 1
#define LUCID_REJECTION_TYPE_FRUSTUM 2
#define LUCID_REJECTION_TYPE_BETWEEN_SAMPLES 3

#define LUCID_INFO_MAX_DISPATCHES 256

/* Per-bin counter arrays stored right after LucidInfo (definitions.glsl:116-129).
 * Index k selects array k; each array has BIN_COUNT entries. */
#define LUCID_COUNTS_QUAD_COUNTS 0
#define LUCID_COUNTS_QUAD_OFFSETS 1
#define LUCID_COUNTS_QUAD_OFFSETS_TEMP 2
#define LUCID_COUNTS_TRI_COUNTS 3
#define LUCID_COUNTS_TRI_OFFSETS 4
#define LUCID_COUNTS_TRI_OFFSETS_TEMP 5
#define LUCID_COUNTS_MICRO_BINS 6
#define LUCID_COUNTS_LOW_BINS 7
#define LUCID_COUNTS_MEDIUM_BINS 8
#define LUCID_COUNTS_HIGH_BINS 9
#define LUCID_COUNTS_ARRAYS 10

typedef struct lucid_instance_data {
	int32_t index_offset;  /* offset (in u32 indices) of the instance's first quad */
	int32_t vertex_offset; /* added to every index of the instance */
	int32_t num_quads;     /* <= 1024 in the reference (lucid_renderer.h:29) */
	uint32_t flags;        /* LUCID_INST_* */
} lucid_instance_data;

typedef struct lucid_vec4 {
	float x, y, z, w;
} lucid_vec4;

typedef struct lucid_mat4 {
	lucid_vec4 col[4]; /* column-major, like GLSL mat4 */
} lucid_mat4;

typedef struct lucid_lighting {
	lucid_vec4 ambient_color;
	lucid_vec4 sun_color;
	lucid_vec4 sun_dir;
	float sun_power, ambient_power;
	float _pad[2]; /* std140 rounds the struct up to 16 bytes */
} lucid_lighting;

/* All vectors in world space. ws_dirx/ws_diry are per-pixel ray increments. */
typedef struct lucid_frustum {
	lucid_vec4 ws_origins[4], ws_dirs[4];
	lucid_vec4 ws_origin0, ws_dir0;
	lucid_vec4 ws_dirx, ws_diry;
} lucid_frustum;

typedef struct lucid_config {
	lucid_frustum frustum;
	lucid_mat4 view_proj_matrix;
	lucid_lighting lighting;
	lucid_vec4 background_color;
	uint32_t enable_backface_culling;
	int32_t num_instances;
	int32_t instance_packet_size;
	uint32_t _pad;
} lucid_config;

typedef struct lucid_info {
	int32_t num_input_quads;
	int32_t num_visible_quads[2]; /* [0]: small quads, [1]: large quads */
	int32_t num_counted_quads[2];
	int32_t bin_level_counts[LUCID_BIN_LEVELS_COUNT];
	uint32_t a_small_bins, a_high_bins;
	uint32_t a_setup_work_groups;
	uint32_t a_dummy_counter;
	uint32_t num_binning_dispatches[3];
	uint32_t bin_level_dispatches[LUCID_BIN_LEVELS_COUNT][3];
	uint32_t num_rejected_quads[LUCID_REJECTION_TYPE_COUNT];
	uint32_t setup_timers[LUCID_TIMERS_COUNT];
	uint32_t raster_timers[LUCID_TIMERS_COUNT];
	uint32_t bin_dispatcher_timers[LUCID_TIMERS_COUNT];
	uint32_t stats[LUCID_STATS_COUNT];
	int32_t dispatcher_first_batch[2][LUCID_INFO_MAX_DISPATCHES];
	int32_t dispatcher_num_batches[2][LUCID_INFO_MAX_DISPATCHES];
	int32_t temp[64];
} lucid_info;

#define LUCID_INFO_U32_SIZE 1152

#if defined(__cplusplus)
#define LUCID_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define LUCID_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

LUCID_STATIC_ASSERT(sizeof(lucid_instance_data) == 16, "InstanceData is 16 bytes");
LUCID_STATIC_ASSERT(sizeof(lucid_lighting) == 64, "Lighting is 64 bytes in std140");
LUCID_STATIC_ASSERT(sizeof(lucid_frustum) == 192, "Frustum is 192 bytes");
LUCID_STATIC_ASSERT(offsetof(lucid_config, view_proj_matrix) == 192, "LucidConfig layout");
LUCID_STATIC_ASSERT(offsetof(lucid_config, lighting) == 256, "LucidConfig layout");
LUCID_STATIC_ASSERT(offsetof(lucid_config, background_color) == 320, "LucidConfig layout");
LUCID_STATIC_ASSERT(offsetof(lucid_config, enable_backface_culling) == 336, "LucidConfig layout");
LUCID_STATIC_ASSERT(offsetof(lucid_config, instance_packet_size) == 344, "LucidConfig layout");
LUCID_STATIC_ASSERT(sizeof(lucid_config) == 352, "LucidConfig is 352 bytes");
LUCID_STATIC_ASSERT(sizeof(lucid_info) == LUCID_INFO_U32_SIZE * 4, "LucidInfo is 1152 u32");
LUCID_STATIC_ASSERT(offsetof(lucid_info, num_binning_dispatches) == 14 * 4, "LucidInfo layout");
LUCID_STATIC_ASSERT(offsetof(lucid_info, bin_level_dispatches) == 17 * 4, "LucidInfo layout");
LUCID_STATIC_ASSERT(offsetof(lucid_info, num_rejected_quads) == 32 * 4, "LucidInfo layout");
LUCID_STATIC_ASSERT(offsetof(lucid_info, stats) == 60 * 4, "LucidInfo layout");
LUCID_STATIC_ASSERT(offsetof(lucid_info, dispatcher_first_batch) == 64 * 4, "LucidInfo layout");
LUCID_STATIC_ASSERT(offsetof(lucid_info, temp) == 1088 * 4, "LucidInfo layout");

#ifdef __cplusplus
}
#endif

#endif
