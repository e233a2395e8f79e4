// Device-side definitions shared by the Lucid HIP kernels (gfx950, wave64).
//
// Arithmetic conventions (they make bin counts, bin lists and per-pixel
// fragment counts bit-reproducible on the CPU):
//   * build with -ffp-contract=off: no fused multiply-add anywhere;
//   * expressions keep the GLSL source's left-to-right order;
//   * division and sqrt are IEEE correctly rounded (HIP's default for fp32);
//   * min/max are IEEE minNum/maxNum (v_min_f32/v_max_f32 in IEEE mode);
//   * float->int/uint conversions truncate, saturate and map NaN to 0.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lucid_c.h"
#include "../../include/lucid_structs.h"

namespace lucid {

constexpr int kBinSize = 32;
constexpr int kBinShift = 5;
constexpr int kBinMask = 31;
constexpr int kBlockSize = 8;
constexpr uint32_t kInvalidAabb = 0xffffffffu;
constexpr int kLowLevelLimit = 1024; // bin_categorizer.glsl:352
constexpr int kMaxBlockTris = 256;   // raster_low.glsl:17 (promotion threshold)
constexpr int kLowRowCap = 1024;     // records per block row of a low bin (< 1024 tris)

// Internal control words (separate from LucidInfo, whose layout is the reference's).
enum CtrlWord : int {
	kCtrlRecorded = 0,      // 1: this frame's rasterisation records the backward's pass A
	kCtrlRecordCursor = 1,  // high-bin record pool size the frame needs (uint4 units)
	kCtrlFallbackCount = 2, // pixels in the exact-fallback queue
	kCtrlOverflow = 3,      // bit 0 bin lists, 1 record pool, 2 fallback queue, 3 spill pool
	kCtrlSpillCursor = 4,   // bump allocator of the deep-pixel spill pool (u64 units)
	kCtrlTotalBinQuads = 5,
	kCtrlTotalBinTris = 6,
	kCtrlFallbackPixels = 7,
	kCtrlListCursor = 8,    // high-bin block-list space the frame needs (u32 units)
	kCtrlOverflowCarry = 9, // overflow bits of earlier frames not yet seen by the host
	kCtrlSortedCursor = 10, // pre-sorted pool size a frame needed, after an overflow (u32 units)
	kCtrlBwdRows = 11,      // gradient rows of the block kernel (set by the row scan)
	kCtrlBwdExtra = 12,     // bump allocator of per-fragment gradient rows of exact pixels
	kCtrlBwdRanges = 13,    // depth ranges of the backward block kernels (set by the row scan)
	kCtrlBwdGradMax = 14,   // max |dL/dpixel| (float bits) over the gradient image
	kCtrlTopoChanged = 15,  // backward: the corner -> vertex map differs from the cached one
	kCtrlWords = 16,
	// device-only words (not copied to the host) up to kStatStripeBase:
	kCtrlBigSorts = 16, // blocks queued in big_queue for the long-list sort pass
};

// Frame statistics summed per block (fragments, list entries) go to kStatStripes counters
// per statistic, one 128-byte line each, after the control words in the same buffer: the
// block kernels add to stripe (slot % kStatStripes), so that tens of thousands of
// work-groups do not serialise on one address; the fallback kernel folds them into
// LucidInfo once.
constexpr int kStatStripes = 64;
constexpr int kStatStride = 32; // u32 words per stripe
constexpr int kStatStripeBase = 32;
__device__ __forceinline__ uint32_t *stat_stripe(uint32_t *ctrl, int stat, uint32_t stripe) {
	return ctrl + kStatStripeBase + (stat * kStatStripes + (int)stripe) * kStatStride;
}
// The pre-sorted block pool is split into kSortStripes equal regions with a bump cursor
// each (block slot % kSortStripes picks one), for the same reason.
constexpr int kSortStripes = 16;
constexpr int kSortCursorBase = kStatStripeBase + 2 * kStatStripes * kStatStride;
__device__ __forceinline__ uint32_t *sort_cursor(uint32_t *ctrl, uint32_t stripe) {
	return ctrl + kSortCursorBase + (int)stripe * kStatStride;
}
// The large-quad list likewise: setup work-group w appends to sub-list w % kLargeStripes,
// whose capacity (large_stripe_cap) holds every quad of those work-groups; the binning
// kernels index the sub-lists as one list, in stripe order.
constexpr int kLargeStripes = 16;
constexpr int kSetupThreads = 256;
constexpr int kLargeCursorBase = kSortCursorBase + kSortStripes * kStatStride;
constexpr int kCtrlAllocWords = kLargeCursorBase + kLargeStripes * kStatStride;
__device__ __forceinline__ uint32_t *large_cursor(uint32_t *ctrl, uint32_t stripe) {
	return ctrl + kLargeCursorBase + (int)stripe * kStatStride;
}
__host__ __device__ constexpr uint32_t large_stripe_cap(uint32_t num_quads) {
	const uint32_t wgs = (num_quads + kSetupThreads - 1) / kSetupThreads;
	return (wgs + kLargeStripes - 1) / kLargeStripes * kSetupThreads;
}

// Overflow bits of kCtrlOverflow
constexpr uint32_t kOvfBinLists = 1u, kOvfRecords = 2u, kOvfFallback = 4u, kOvfSpill = 8u,
				   kOvfLists = 16u, kOvfSorted = 32u, kOvfGradRows = 64u;

// Buffers of the backward pass (passed by value beside FrameBufs).
constexpr int kGradWords = 10; // per row: corner 0 rgb, corner 1 rgb, corner 2 rgb, alpha
// Start state of one (depth range, pixel) for pass B: transmittance before the range and
// q = g . (out - C_before) (fp64), with g the masked dL/d(pixel rgb).
struct BwdStart {
	double t, q;
	float gr, gg, gb, pad;
};

// What the backward's pass A leaves for one (depth range, pixel) when a forward frame
// records it: T before the range and, per channel, D = (out - composite before the range)
// in fp64, so that the start state is q = sum_c g_c D_c once dL/dpixel is known.
struct BwdRecord {
	double t, d[3];
};

struct BwdBufs {
	const float4 *grad_image; // dL/d(rgba32f image); alpha ignored
	uint32_t *row_base;       // bin_count * 16: first gradient row of each block
	uint32_t *range_base;     // bin_count * 16: first depth range of each block
	float *rows;              // kGradWords floats per row
	uint32_t *row_tri;        // triangle id of each row
	uint64_t rows_cap;
	BwdStart *start;          // 64 per depth range
	uint64_t start_cap;       // depth ranges
	uint8_t *seq;             // 64 bytes per gradient row: recorded blend order
	uint64_t seq_rows_cap;    // gradient rows
	BwdRecord *rec;           // 64 per depth range (start_cap): pass A recorded by a forward
	uint8_t *pix_flags;       // width * height, with rec: clamp bits 0-2, bit 3 = exact pixel
};

// Scratch of the gradient reduction (ntri triangle ids, nrows gradient rows).
struct ReduceBufs {
	uint32_t *counts, *offsets; // ntri + 1 each: rows per triangle, exclusive scan
	uint32_t *grouped;          // nrows: row indices grouped by triangle
	float *tri_grad;            // ntri * kGradWords
	uint32_t *vkeys, *vsorted, *vcorners, *viota; // 3 ntri each: corners grouped by vertex
	void *temp;
	size_t temp_bytes;
};

// Per-block slot flags (blk_flags)
constexpr uint32_t kSlotLevelMask = 0xffu; // LUCID_BIN_LEVEL_*; 0 = empty bin
constexpr uint32_t kSlotPresorted = 1u << 8;
constexpr uint32_t kSlotPromoted = 1u << 9;  // low bin handed to the high path
constexpr uint32_t kSlotWhole = 1u << 10;    // list too long: every pixel takes the fallback
constexpr uint32_t kSlotPartShift = 12;      // log2 of the waves sharing a pre-sorted block
constexpr int kRegSortCap = 256;  // lists sorted in registers by the raster wave itself
constexpr int kBigSortCap = 8192; // lists sorted in LDS by high_finalize_kernel
constexpr int kRowgenSplit = 8;   // work-groups per high bin in rowgen_high

enum ColorMode : int { kColorNone = 0, kColorRGBA8 = 1, kColorF32 = 2 };

// Per-frame constants (passed by value to every kernel).
struct FrameConst {
	int width, height;
	int bin_count_x, bin_count_y, bin_count;
	int num_quads, num_instances;
	int num_bin_wgs; // work-groups of the bin counter/dispatcher
	uint32_t opts;
	uint32_t backface;
	// tuning (results do not depend on it)
	uint32_t presort_min;      // high-bin blocks with more entries are pre-sorted (<= kRegSortCap)
	uint32_t samples_per_wave; // target samples per wave of a pre-sorted block
	int slice_words;           // block rasteriser slice width: 1 or 2 words of 64 entries
	int window;                // ordering window of flat-colour frames (tuning): 1 or 4
	int color_mode;
	int has_normals;
	// 1 when an earlier frame of the renderer may still be running: the clear kernel then
	// moves its overflow bits to kCtrlOverflowCarry instead of dropping them, so that they
	// are reported with this frame's status
	uint32_t keep_overflow;
	int has_textures; // some instance samples an albedo texture (setup stores texture coordinates)
	// ambient 1 and no sun light: every lighting factor of the frame is exactly 1, and
	// finalShading's sRGB round trip L2S(S2L(x) * 1) is the identity
	int identity_light;
	float vp[16]; // column-major view-projection matrix
	float origin[3], dir0[3], dirx[3], diry[3], ray_dir0[3];
	float ambient[3], sun_color[3], sun_dir[3];
	float sun_power, ambient_power;
	float background[4];
};

// An RGBA8 texture atlas and its mip chain (lucid_texture): texels of level l start at
// offset[l] and form a max(1, width >> l) x max(1, height >> l) row-major image.
constexpr int kMaxTexLevels = 16;
struct TexDesc {
	const uint32_t *texels;
	int width, height, levels;
	uint32_t offset[kMaxTexLevels];
};

// Device buffers of one frame (passed by value).
struct FrameBufs {
	// inputs
	const lucid_instance_data *instances;
	const int *inst_prefix; // num_instances + 1
	const uint32_t *inst_colors;
	const uint32_t *quad_indices;
	const float *positions;
	const uint32_t *colors_rgba8;
	const float *colors_rgb;
	const float *tri_alpha;
	const uint32_t *normals;
	const float *tex_coords;    // float2 per vertex (textured instances)
	const float4 *inst_uv_rects; // per instance: min.xy, size.xy of the texture in its atlas
	TexDesc tex[2];              // 0: opaque atlas, 1: transparent atlas
	// info + control
	uint32_t *info;   // LucidInfo followed by 10 * bin_count counters
	uint32_t *ctrl;   // kCtrlWords
	// setup records; per-triangle arrays are indexed by tri_rec(): the first triangles of
	// all quads, then the second ones (soup scenes cull every second triangle, so their
	// records stay dense)
	uint32_t tri_half;    // quad capacity: record index of triangle 2 q + 1 is q + tri_half
	uint32_t *quad_aabb;  // per quad
	uint4 *tri_scan;      // 2 per tri
	uint4 *tri_depth;     // 1 per tri
	uint4 *tri_shade;     // kRecStride per tri: the shading record the rasterisers stage
	uint4 *tri_tex;       // 2 per tri of a textured instance: tex0.xy, tex1 - tex0 | tex2 - tex0
	uint32_t *large_list; // quads whose bin AABB covers more than 4 bins (kLargeStripes sub-lists)
	// binning
	uint32_t *wg_hist; // num_bin_wgs * 2 * bin_count
	uint32_t *bin_quads, *bin_tris;
	uint64_t bin_quads_cap, bin_tris_cap;
	// raster
	uint4 *low_records;  // bin_count * 4 * kLowRowCap
	uint4 *high_records; // pool
	uint64_t high_records_cap;
	uint4 *blk_slot;     // bin_count * 16: list offset, list length, record offset, record count
	uint32_t *blk_flags; // bin_count * 16
	uint32_t *blk_sorted; // bin_count * 16: offset of the pre-sorted arrays
	uint32_t *big_queue;  // bin_count * 16: blocks left to high_finalize's long-list pass
	uint32_t *blk_lists;  // low bins: fixed bin_count * 16 * kRegSortCap; then high bump region
	uint64_t blk_lists_cap;
	uint32_t *blk_count;  // bin_count * 16: list lengths of high bins (atomic counters)
	uint32_t *band_count; // bin_count * 4: record counts of high-bin bands (atomic counters)
	uint32_t *sorted_pool;
	uint64_t sorted_cap;
	uint4 *fallback_queue; // one {slot, px, py, whole} per pixel
	uint32_t fallback_cap;
	uint64_t *spill; // deep-pixel spill pool
	uint64_t spill_cap;
	// outputs
	uint8_t *image;
	float *image_f;
	uint32_t *frag_counts;
};

// ------------------------------------------------------------------ batched launches
//
// A batch renders many frames (scenes of different sizes and resolutions) with one launch
// per pipeline stage: the stage's grid is the concatenation of every frame's own grid, and a
// work-group finds its frame and its position within that frame's grid (WgPos) from an
// exclusive prefix of the per-frame work-group counts. Every stage kernel is written once,
// as a body taking WgPos, with a single-frame wrapper (the frame's constants and buffers as
// kernel arguments, WgPos = blockIdx.x / gridDim.x) and a batched wrapper.
struct WgPos {
	uint32_t id, count;
};

// Work-groups go to the 8 XCDs (each with its own L2) in turn: work-group w runs on XCD
// w % 8. bin_run_id renumbers a grid so that each run of 16 consecutive ids (the blocks of
// one bin, which gather the same triangles' records) runs on one XCD, one after another,
// while the runs still go to the XCDs in turn (so dense bins stay spread over the chip).
// Ids from the last whole round of 8 runs on keep their place; a grid whose per-frame
// ranges are multiples of 16 keeps every run inside one frame.
__device__ __forceinline__ uint32_t bin_run_id(uint32_t w, uint32_t count) {
	constexpr uint32_t kXcds = 8, kRun = 16;
	const uint32_t whole = count / (kXcds * kRun) * (kXcds * kRun);
	if(w >= whole)
		return w;
	const uint32_t x = w % kXcds, k = w / kXcds;
	return ((k / kRun) * kXcds + x) * kRun + k % kRun;
}

struct FrameDesc {
	FrameConst c;
	FrameBufs b;
};

enum BatchStage : int {
	kBsClear,
	kBsSetup,
	kBsBinCount,
	kBsBinScan,
	kBsBinCategorize,
	kBsBinDispatch,
	kBsRowgenLow,
	kBsHighPrep,
	kBsRowgenHigh,
	kBsHighFinalize,
	kBsHighFinalizeLong, // persistent: a few work-groups per frame over its queue
	kBsBlockRaster,
	kBsBlockRasterBig,
	kBsFallback,
	kBsCount
};

// One stage of a batched launch: frame f owns work-groups [pre[f], pre[f + 1]).
struct StageBatch {
	const FrameDesc *frames;
	const uint32_t *pre;
	uint32_t num_frames;
};

// The frame of work-group wg: the last frame whose range starts at or before wg (frames
// with no work-groups in this stage have empty ranges and are passed over).
__device__ __forceinline__ uint32_t batch_frame(const StageBatch &sb, uint32_t wg) {
	uint32_t lo = 0, hi = sb.num_frames - 1;
	while(lo < hi) {
		const uint32_t mid = (lo + hi + 1) >> 1;
		if(sb.pre[mid] <= wg)
			lo = mid;
		else
			hi = mid - 1;
	}
	return lo;
}

// A batched launch: the frames (device array), the per-stage prefixes (device,
// kBsCount x (num_frames + 1)), and what selects the kernels (shared by every frame).
// Shading modes of the block rasterisers: general shading records, the same with albedo
// textures, or flat colour planes (RGBA8 samples, unit lighting, vertex colours absent or
// fp32 with a per-triangle alpha, no textures: a sample's colour is then an affine function
// of the pixel divided by the inverse depth).
// kShadeFlatNarrow: flat records with the narrow ordering windows (tuning window 1).
enum ShadeMode : int { kShadeGeneral = 0, kShadeTextured = 1, kShadeFlat = 2, kShadeFlatNarrow = 3 };

inline int frame_shade_mode(const FrameConst &c) {
	if(c.has_textures)
		return kShadeTextured;
	const bool fp32 = (c.opts & LUCID_OPT_FP32_SAMPLES) != 0;
	const bool flat_colors = c.color_mode == kColorNone || c.color_mode == kColorF32;
	if(fp32 || !c.identity_light || !flat_colors)
		return kShadeGeneral;
	return c.window == 1 ? kShadeFlatNarrow : kShadeFlat;
}

struct BatchPlan {
	const FrameDesc *frames;
	const uint32_t *pre;
	uint32_t num_frames;
	uint32_t total[kBsCount]; // work-groups of each stage over all frames
	size_t bin_lds;           // dynamic LDS of the bin kernels: the largest frame's
	bool fp32;
	int mode; // ShadeMode of the block kernels: textured if a frame is, flat if every frame is
	int slice_words;
	uint32_t *status; // device, kCtrlWords per frame: each frame's control words after it
	__host__ const uint32_t *stage_pre(int s) const { return pre + (size_t)s * (num_frames + 1); }
};

// Record index of triangle t (2 q + s) in the per-triangle setup arrays.
__device__ __forceinline__ uint32_t tri_rec(const FrameBufs &b, uint32_t t) {
	return (t >> 1) + (t & 1u) * b.tri_half;
}

// ------------------------------------------------------------------ conversions

__device__ __forceinline__ int f2i(float x) {
	// v_cvt_i32_f32 semantics: truncate, saturate, NaN -> 0. Branch-free: the clamp keeps the
	// conversion in range (2^31 - 128 is the largest float below 2^31).
	const int v = (int)fminf(fmaxf(x, -2147483648.0f), 2147483520.0f);
	return x != x ? 0 : x >= 2147483648.0f ? 2147483647 : v;
}

// Saturating float -> uint32 (truncating): NaN and x <= 0 give 0, x >= 2^32 gives 2^32 - 1.
// Branch-free: the clamp keeps the conversion in range (no float lies strictly between
// 2^32 - 256 and 2^32, so the upper clamp changes no x below 2^32).
__device__ __forceinline__ uint32_t f2u(float x) {
	const uint32_t u = (uint32_t)fminf(fmaxf(x, 0.0f), 4294967040.0f);
	return x >= 4294967296.0f ? 0xffffffffu : u;
}

__device__ __forceinline__ float bits2f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t f2bits(float f) { return __float_as_uint(f); }

__device__ __forceinline__ float fmx(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ float fmn(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float sat(float x) { return fmn(fmx(x, 0.0f), 1.0f); }

// Monotone map of float order onto uint order (for sort keys).
__device__ __forceinline__ uint32_t ord_f32(float f) {
	uint32_t u = f2bits(f);
	return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
// ~ord_f32(f) (reversed order) as one select: u for negative f, else u ^ 0x7fffffff.
__device__ __forceinline__ uint32_t inv_ord_f32(float f) {
	const uint32_t u = f2bits(f);
	return (int32_t)u < 0 ? u : u ^ 0x7fffffffu;
}

// ------------------------------------------------------------------ small vectors

struct V3 {
	float x, y, z;
};
__device__ __forceinline__ V3 mk3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 sub3(V3 a, V3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 add3(V3 a, V3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 mul3(V3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 neg3(V3 a) { return mk3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
	return mk3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
__device__ __forceinline__ V3 ld3(const float v[3]) { return mk3(v[0], v[1], v[2]); }

// ------------------------------------------------------------------ encoders

__device__ __forceinline__ uint32_t encode_aabb28(uint32_t x0, uint32_t y0, uint32_t x1,
												  uint32_t y1) {
	return (x0 & 0x7fu) | ((y0 & 0x7fu) << 7) | ((x1 & 0x7fu) << 14) | ((y1 & 0x7fu) << 21);
}

__device__ __forceinline__ uint32_t encode_rgba8(float r, float g, float b, float a) {
	return f2u(r * 255.0f) | (f2u(g * 255.0f) << 8) | (f2u(b * 255.0f) << 16) |
		   (f2u(a * 255.0f) << 24);
}

// encode_rgba8 of colour channels already in [0, 1] (shaded samples are saturated): the
// truncating conversion needs no clamp there; alpha keeps the saturating one
__device__ __forceinline__ uint32_t encode_rgba8_unit_rgb(float r, float g, float b, float a) {
	return (uint32_t)(r * 255.0f) | ((uint32_t)(g * 255.0f) << 8) |
		   ((uint32_t)(b * 255.0f) << 16) | (f2u(a * 255.0f) << 24);
}

__device__ __forceinline__ float4 decode_rgba8(uint32_t c) {
	return make_float4((float)(c & 0xffu) * (1.0f / 255.0f),
					   (float)((c >> 8) & 0xffu) * (1.0f / 255.0f),
					   (float)((c >> 16) & 0xffu) * (1.0f / 255.0f),
					   (float)(c >> 24) * (1.0f / 255.0f));
}

__device__ __forceinline__ uint32_t encode_normal(V3 n) {
	uint32_t x = f2u(512.0f + n.x * 511.0f) & 0x3ffu;
	uint32_t y = f2u(512.0f + n.y * 511.0f) & 0x3ffu;
	uint32_t z = f2u(512.0f + n.z * 511.0f) & 0x3ffu;
	return x | (y << 10) | (z << 20);
}

__device__ __forceinline__ V3 decode_normal(uint32_t n) {
	return mk3(((float)(n & 0x3ffu) - 512.0f) * (1.0f / 511.0f),
			   ((float)((n >> 10) & 0x3ffu) - 512.0f) * (1.0f / 511.0f),
			   ((float)((n >> 20) & 0x3ffu) - 512.0f) * (1.0f / 511.0f));
}

// ------------------------------------------------------------------ wave helpers

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ void wave_sync_lds() {
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t wave_inclusive_add(uint32_t v) {
	int l = lane_id();
#pragma unroll
	for(int d = 1; d < 64; d <<= 1) {
		uint32_t t = __shfl_up(v, d, 64);
		if(l >= d)
			v += t;
	}
	return v;
}

__device__ __forceinline__ uint32_t wave_inclusive_max(uint32_t v) {
	int l = lane_id();
#pragma unroll
	for(int d = 1; d < 64; d <<= 1) {
		uint32_t t = __shfl_up(v, d, 64);
		if(l >= d)
			v = v > t ? v : t;
	}
	return v;
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Index of the k-th (0-based) set bit of m (k < popcount(m)).
__device__ __forceinline__ int select_bit64(uint64_t m, uint32_t k) {
	int pos = 0;
	uint32_t lo = (uint32_t)m;
	uint32_t c = __popc(lo);
	uint32_t w = lo;
	if(k >= c) {
		k -= c;
		pos = 32;
		w = (uint32_t)(m >> 32);
	}
	c = __popc(w & 0xffffu);
	if(k >= c) {
		k -= c, pos += 16, w >>= 16;
	}
	c = __popc(w & 0xffu);
	if(k >= c) {
		k -= c, pos += 8, w >>= 8;
	}
	c = __popc(w & 0xfu);
	if(k >= c) {
		k -= c, pos += 4, w >>= 4;
	}
	c = __popc(w & 0x3u);
	if(k >= c) {
		k -= c, pos += 2, w >>= 2;
	}
	c = w & 1u;
	if(k >= c)
		pos += 1;
	return pos;
}

// ------------------------------------------------------------------ scanline rows

// Scanline state of one triangle inside one bin (scanline.glsl:13-26).
struct RowScan {
	float mn[3], mx[3], step[3];
};

// Starts the scan at absolute row `start_row` for a bin whose left edge is bin_x.
__device__ __forceinline__ void row_scan_init(RowScan &s, uint4 scan0, uint4 scan1, float startx,
											  float starty) {
	uint32_t signs = scan1.w;
	float sc[3] = {bits2f(scan0.x), bits2f(scan0.y), bits2f(scan0.z)};
	float st[3] = {bits2f(scan1.x), bits2f(scan1.y), bits2f(scan1.z)};
#pragma unroll
	for(int i = 0; i < 3; i++) {
		float v = sc[i] + (st[i] * starty - startx);
		bool xneg = (signs >> i) & 1u;
		s.mn[i] = xneg ? -__builtin_huge_valf() : v;
		s.mx[i] = xneg ? v : __builtin_huge_valf();
		s.step[i] = st[i];
	}
}

// One row of rasterBinStep (raster.glsl:116-140): returns the span packed as
// min | max << 5 (empty rows are (31, 0)) and advances the state.
__device__ __forceinline__ uint32_t row_scan_step(RowScan &s) {
	int a = f2i(fmx(fmx(s.mn[0], s.mn[1]), fmx(s.mn[2], 0.0f)));
	int b = f2i(fmn(fmn(s.mx[0], s.mx[1]), fmn(s.mx[2], (float)kBinSize))) - 1;
	if(a > b)
		a = kBinMask, b = 0;
#pragma unroll
	for(int i = 0; i < 3; i++) {
		s.mn[i] = s.mn[i] + s.step[i];
		s.mx[i] = s.mx[i] + s.step[i];
	}
	return (uint32_t)a | ((uint32_t)b << 5);
}

__device__ __forceinline__ uint32_t span_bx_mask(uint32_t span) {
	uint32_t a = span & 31u, b = (span >> 5) & 31u;
	return ((15u << (a >> 3)) & (15u >> (3u - (b >> 3)))) & 15u;
}

// Block-row record: 8 row spans (10 bits each) of one triangle in one 8-row band.
//   x = triangle id, y = rows 0-2, z = rows 3-5, w = rows 6-7 | bx_mask << 20
__device__ __forceinline__ uint32_t record_span(uint4 rec, int row) {
	uint32_t w = row < 3 ? rec.y : row < 6 ? rec.z : rec.w;
	int sh = (row < 3 ? row : row < 6 ? row - 3 : row - 6) * 10;
	return (w >> sh) & 0x3ffu;
}

__device__ __forceinline__ uint32_t record_bx_mask(uint4 rec) { return (rec.w >> 20) & 15u; }

// 64-bit coverage mask of block column bx (bit y*8 + x) from a record.
__device__ __forceinline__ uint64_t record_block_mask(uint4 rec, int bx) {
	uint64_t m = 0;
	int x0 = bx * 8;
#pragma unroll
	for(int r = 0; r < 8; r++) {
		uint32_t sp = record_span(rec, r);
		int a = (int)(sp & 31u) - x0, b = (int)((sp >> 5) & 31u) - x0;
		a = a < 0 ? 0 : a;
		b = b > 7 ? 7 : b;
		if(b >= a) {
			uint32_t bits = ((1u << (b - a + 1)) - 1u) << a;
			m |= (uint64_t)bits << (r * 8);
		}
	}
	return m;
}

// ------------------------------------------------------------------ shading records

// GLSL specifies pow(x, y) with the precision of exp2(y * log2(x)); the hardware log2/exp2
// evaluate exactly that.
__device__ __forceinline__ float glsl_pow(float x, float y) {
	return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}
// Lighting of finalShading (funcs.glsl:261-271): light = max(0, dot(-sun_dir, normal) *
// 0.7 + 0.3), and per channel the factor ambient + sun * light.
__device__ __forceinline__ float light_value(const FrameConst &c, V3 normal) {
	return fmx(0.0f, dot3(neg3(ld3(c.sun_dir)), normal) * 0.7f + 0.3f);
}
__device__ __forceinline__ float light_factor(const FrameConst &c, int k, float light) {
	const float amb = c.ambient[k] * c.ambient_power;
	const float dif = (c.sun_color[k] * c.sun_power) * light;
	return amb + dif;
}

// Shading record of one triangle, 7 x uint4 (the rasteriser stages one per list entry in LDS):
//   q0  depth plane x, y, z | flags | instance << 16
//   q1  barycentric plane 0 x, y, z | offset;  q2 the same for plane 1
//   q3  RGBA8 colours: the triangle's three vertex colours | 0
//       fp32 colours:  c0.rgb | triangle alpha
//   q4  c1.rgb | L0;  q5 c2.rgb | L1
//   q6  instance colour (RGBA8; white without one) | L2 | triangle id | light
// A flat-shaded triangle keeps its light value and, in L0-L2, fp = f^(1/2.4) of its three
// lighting factors f; with LUCID_INST_HAS_VERTEX_NORMALS, L0-L2 are the packed vertex
// normals (light is then per pixel).
constexpr int kRecQuads = 7;
constexpr int kRecStride = 8; // uint4 per record in tri_shade: one 128-byte line

// Composes the record of triangle t = 2 q + s from what setup computed for it: its depth
// record d, barycentric planes, packed face normal, and the quad's vertex ids vq (the
// triangle uses vq[0], vq[1 + s], vq[2 + s]). Setup stores it in tri_shade once per frame.
__device__ __forceinline__ void make_shading_record(const FrameConst &c, const FrameBufs &b,
													uint32_t t, uint4 d, uint4 bary0, uint4 bary1,
													uint32_t packed_normal,
													const uint32_t (&vq)[4],
													uint4 (&q)[kRecQuads]) {
	// every slot is computed into a named local and stored once, so that no part of q
	// is written under a branch (which would push q to scratch memory)
	const uint32_t flags = d.w & 0xffffu, inst = d.w >> 16;
	const bool s = t & 1u;
	const uint32_t v0 = vq[0], v1 = s ? vq[2] : vq[1], v2 = s ? vq[3] : vq[2];
	const bool vcol = (flags & LUCID_INST_HAS_VERTEX_COLORS) != 0;
	uint4 q3 = make_uint4(0u, 0u, 0u, 0u), q4 = q3, q5 = q3;
	if(vcol && c.color_mode == kColorRGBA8)
		q3 = make_uint4(b.colors_rgba8[v0], b.colors_rgba8[v1], b.colors_rgba8[v2], 0u);
	if(vcol && c.color_mode == kColorF32) {
		const float *c0 = b.colors_rgb + (size_t)v0 * 3, *c1 = b.colors_rgb + (size_t)v1 * 3,
					*c2 = b.colors_rgb + (size_t)v2 * 3;
		const float alpha = b.tri_alpha ? b.tri_alpha[t] : 1.0f;
		q3 = make_uint4(f2bits(c0[0]), f2bits(c0[1]), f2bits(c0[2]), f2bits(alpha));
		q4 = make_uint4(f2bits(c1[0]), f2bits(c1[1]), f2bits(c1[2]), 0u);
		q5 = make_uint4(f2bits(c2[0]), f2bits(c2[1]), f2bits(c2[2]), 0u);
	}
	const uint32_t icol =
		((flags & LUCID_INST_HAS_COLOR) && b.inst_colors) ? b.inst_colors[inst] : 0xffffffffu;
	uint32_t l0, l1, l2, lv = 0u;
	if(flags & LUCID_INST_HAS_VERTEX_NORMALS) {
		const bool have = b.normals != nullptr;
		l0 = have ? b.normals[v0] : 0u, l1 = have ? b.normals[v1] : 0u;
		l2 = have ? b.normals[v2] : 0u;
	} else {
		// the packed normal, as the reference stores and shades it
		const float light = light_value(c, decode_normal(packed_normal));
		lv = f2bits(light);
		l0 = f2bits(glsl_pow(light_factor(c, 0, light), 1.0f / 2.4f));
		l1 = f2bits(glsl_pow(light_factor(c, 1, light), 1.0f / 2.4f));
		l2 = f2bits(glsl_pow(light_factor(c, 2, light), 1.0f / 2.4f));
	}
	q4.w = l0;
	q5.w = l1;
	q[0] = d;
	q[1] = bary0;
	q[2] = bary1;
	q[3] = q3;
	q[4] = q4;
	q[5] = q5;
	q[6] = make_uint4(icol, l2, t, lv);
}

// ------------------------------------------------------------------ gradient fixed point

// The backward pass sums gradients as 64-bit fixed point (integer sums are associative, so
// a sum does not depend on the order its terms arrive in). The scale 2^k is fixed per
// frame from max |dL/dpixel| over the image and a bound on one
// fragment's contribution (|d sample / d colour| <= 12.92 * 2.4 * lighting, |d out / d a|
// <= 2 per channel): one contribution stays below 2^50 units and a row of at most 64 below
// 2^56, a resolution of ~1e-15 of the largest possible contribution.
__device__ __forceinline__ int grad_fixed_exp(const FrameConst &c, float gmax) {
	float lmax = 0.0f; // bound on the lighting factor (light in [0, 1])
#pragma unroll
	for(int k = 0; k < 3; k++)
		lmax = fmaxf(lmax, fabsf(c.ambient[k] * c.ambient_power) +
							   fabsf(c.sun_color[k] * c.sun_power));
	const double bound = 64.0 * (double)gmax * fmax(32.0 * (double)lmax, 8.0);
	int ex = 0;
	(void)frexp(bound, &ex); // bound < 2^ex: one contribution < 2^(ex - 6)
	int k = 56 - ex;         // one contribution < 2^50 units, a row < 2^56
	return k < -1000 ? -1000 : k > 1000 ? 1000 : k;
}

// round(v scale) as a two's-complement 64-bit integer, for |v scale| < 2^51: added to
// 1.5 * 2^52 in one rounding (fma), the integer lands in the low mantissa bits of a double
// whose exponent is fixed, so the integer is the bit pattern minus that of 1.5 * 2^52.
__device__ __forceinline__ unsigned long long to_fixed(float v, double scale) {
	const double y = __builtin_fma((double)v, scale, 0x1.8p52);
	return (unsigned long long)__double_as_longlong(y) - 0x4338000000000000ull;
}


} // namespace lucid
