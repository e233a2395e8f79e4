// Host driver of the MI355X Lucid rasterizer and its C ABI (include/lucid_c.h).
//
// Mirrors LucidRenderer (src/lucid_renderer.h:21-96, src/lucid_renderer.cpp:186-707):
//   * construction sizes the bin grid from the view size (:186-232) and owns every internal
//     device buffer (:238-313);
//   * render() uploads instance data when it changes (:352-429), builds the per-frame
//     constants from LucidConfig (:431-451) and enqueues the stage kernels on one stream
//     (:319-350): clear, setup, binning (count, scan, categorise, dispatch), raster_low,
//     raster_high, exact fallback;
//   * verifyInfo/getStats read LucidInfo + per-bin counters back (:629-836).
// No CUDA compatibility layer: plain HIP runtime calls for gfx950.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/lucid_c.h"
#include "lucid_device.h"

namespace lucid {
hipError_t launch_setup(const FrameConst &c, const FrameBufs &b, uint32_t n_info,
						hipStream_t stream);
hipError_t launch_binning(const FrameConst &c, const FrameBufs &b, hipStream_t stream);
hipError_t launch_setup_batch(const BatchPlan &p, hipStream_t stream);
hipError_t launch_binning_batch(const BatchPlan &p, hipStream_t stream);
hipError_t launch_raster_batch(const BatchPlan &p, hipStream_t stream);
hipError_t launch_raster(const FrameConst &c, const FrameBufs &b, hipStream_t stream,
						 hipEvent_t *events);
hipError_t launch_backward_prepare(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
								   hipStream_t stream);
hipError_t launch_backward_reuse(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
								 hipStream_t stream);
hipError_t launch_backward_order(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
								 hipStream_t stream);
hipError_t launch_backward_rows(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
								hipStream_t stream);
hipError_t launch_grad_max(const FrameConst &c, const BwdBufs &g, uint32_t *ctrl,
						   hipStream_t stream);
hipError_t launch_raster_recording(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
								   hipStream_t stream, hipEvent_t *events);
hipError_t launch_backward_recorded(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
									hipStream_t stream);
hipError_t launch_backward_recorded_order(const FrameConst &c, const FrameBufs &b,
										  const BwdBufs &g, hipStream_t stream);
size_t reduce_temp_bytes(uint32_t npairs, uint32_t nscan);
hipError_t reduce_check_corners(const FrameConst &c, const FrameBufs &b, uint32_t ntri,
								uint32_t *vkeys, uint32_t *changed, hipStream_t stream);
hipError_t reduce_group_rows(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
							 const ReduceBufs &r, uint32_t ntri, hipStream_t stream);
hipError_t reduce_sum(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
					  const ReduceBufs &r, uint32_t ntri, uint32_t nverts, bool regroup,
					  float *grad_colors, float *grad_alpha, hipStream_t stream);
hipError_t reduce_vertices(const ReduceBufs &r, uint32_t ntri, uint32_t nverts, bool regroup,
						   float *grad_colors, float *grad_alpha, hipStream_t stream);
} // namespace lucid

using namespace lucid;

namespace {

constexpr int kMaxWidth = 2560, kMaxHeight = 2048;   // src/lucid_renderer.h:27
constexpr int kMaxInstances = 64 * 1024;            // src/lucid_renderer.h:28
constexpr int64_t kMaxQuads = (int64_t)1 << 28;     // quad ids share a word with 4 flag bits
constexpr int kNumBinWgs = 256; // bin_wg_scan_kernel assumes at most 256 (16 rows per part)

struct DevBuf {
	void *ptr = nullptr;
	size_t bytes = 0;
	~DevBuf() { release(); }
	void release() {
		if(ptr)
			(void)hipFree(ptr);
		ptr = nullptr, bytes = 0;
	}
	// Grows (never shrinks); contents are not preserved.
	hipError_t reserve(size_t need) {
		if(need <= bytes)
			return hipSuccess;
		release();
		hipError_t e = hipMalloc(&ptr, need);
		if(e != hipSuccess) {
			ptr = nullptr;
			return e;
		}
		bytes = need;
		return hipSuccess;
	}
	template <class T> T *as() const { return (T *)ptr; }
};

} // namespace

struct lucid_renderer {
	int device = 0;
	int width = 0, height = 0;
	uint32_t opts = 0;
	lucid_tuning tuning = {1, 128, 1024, 1};
	int bin_count_x = 0, bin_count_y = 0, bin_count = 0;

	DevBuf info, ctrl;
	DevBuf instances, inst_prefix, inst_colors, inst_uv_rects;
	DevBuf quad_aabb, tri_scan, tri_depth, tri_shade, large_list;
	DevBuf tri_tex;            // texture coordinates per triangle, once a frame is textured
	bool tex_enabled = false;
	DevBuf wg_hist, bin_quads, bin_tris;
	DevBuf low_records, high_records, fallback_queue, spill;
	DevBuf blk_slot, blk_flags, blk_sorted, blk_lists, blk_count, band_count, sorted_pool;
	// backward
	DevBuf grad_row_base, grad_range_base, grad_rows, grad_row_tri, grad_start, grad_seq,
		red_counts, red_offsets, red_grouped, tri_grad, red_temp;
	DevBuf grad_rec, grad_pix_flags; // pass A as recorded by a forward frame
	uint64_t grad_rows_cap = 0, grad_seq_rows_cap = 0, grad_start_cap = 0;
	// corners grouped by vertex (depends on the mesh only): the corner -> vertex map of the
	// last backward pass, and the corners sorted by vertex
	DevBuf vtx_keys, vtx_sorted, vtx_corners, vtx_iota;
	uint64_t vtx_ntri = 0, vtx_nverts = 0;
	bool vtx_valid = false;
	// helper stream of the backward pass (the row grouping runs beside pass B)
	hipStream_t aux = nullptr;
	hipEvent_t aux_fork = nullptr, aux_join = nullptr, aux_scale = nullptr, aux_order = nullptr;
	// The setup records, bins and block lists in the buffers, and the inputs they were built
	// from (valid once the frame that built them finished without an overflow).
	struct Retained {
		bool valid = false, pending = false;
		// record: the frame ran the recording pass A; recorded: its gate admitted it, so the
		// recorded order, start records and gradient row layout belong to this frame
		bool record = false, recorded = false;
		lucid_frame_inputs in;
		lucid_config cfg;
		lucid_tuning tuning;
		std::vector<lucid_instance_data> instances;
	} retained;

	int64_t quad_cap = 0;
	uint64_t bin_quads_cap = 0, bin_tris_cap = 0, high_records_cap = 0, spill_cap = 0;
	uint64_t blk_lists_cap = 0, sorted_cap = 0;
	uint32_t fallback_cap = 0;

	std::vector<lucid_instance_data> host_instances;
	std::vector<uint32_t> host_inst_colors;
	bool inst_colors_null = true;
	std::vector<lucid_vec4> host_uv_rects;
	bool uv_rects_null = true;

	uint32_t *status = nullptr; // pinned copy of the control words of the last frame
	hipEvent_t done = nullptr;
	static constexpr int kStages = 6;
	hipEvent_t stage_ev[kStages + 1] = {};
	bool timers = false;
	bool frame_pending = false;
	// a frame that was never waited for overflowed; lucid_finish reports it
	bool unreported_overflow = false;
	int64_t last_num_quads = 0;
	hipStream_t last_stream = nullptr;

	~lucid_renderer() {
		if(done)
			(void)hipEventDestroy(done);
		if(aux_fork)
			(void)hipEventDestroy(aux_fork);
		if(aux_join)
			(void)hipEventDestroy(aux_join);
		if(aux_scale)
			(void)hipEventDestroy(aux_scale);
		if(aux_order)
			(void)hipEventDestroy(aux_order);
		if(aux)
			(void)hipStreamDestroy(aux);
		for(hipEvent_t e : stage_ev)
			if(e)
				(void)hipEventDestroy(e);
		if(status)
			(void)hipHostFree(status);
	}
};

namespace {

int hip_status(hipError_t e) {
	if(e == hipSuccess)
		return LUCID_OK;
	if(e == hipErrorOutOfMemory)
		return LUCID_ERR_OUT_OF_MEMORY;
	return LUCID_ERR_HIP;
}

#define LUCID_CHECK(expr)                                                                          \
	do {                                                                                           \
		hipError_t _e = (expr);                                                                    \
		if(_e != hipSuccess)                                                                       \
			return hip_status(_e);                                                                 \
	} while(0)

int ensure_quad_capacity(lucid_renderer *r, int64_t nq) {
	if(nq <= r->quad_cap)
		return LUCID_OK;
	int64_t cap = std::max<int64_t>(nq, std::max<int64_t>(r->quad_cap * 3 / 2, 1024));
	int64_t tris = cap * 2;
	LUCID_CHECK(r->quad_aabb.reserve(cap * 4));
	LUCID_CHECK(r->tri_scan.reserve(tris * 32));
	LUCID_CHECK(r->tri_depth.reserve(tris * 16));
	LUCID_CHECK(r->tri_shade.reserve(tris * kRecStride * 16));
	if(r->tex_enabled)
		LUCID_CHECK(r->tri_tex.reserve(tris * 32));
	LUCID_CHECK(r->large_list.reserve((size_t)kLargeStripes * large_stripe_cap((uint32_t)cap) * 4));
	// small quads cover at most 4 bins: 4 entries per quad is an exact bound
	uint64_t bq = (uint64_t)cap * 4;
	LUCID_CHECK(r->bin_quads.reserve(bq * 4));
	r->bin_quads_cap = bq;
	uint64_t bt = std::max<uint64_t>(r->bin_tris_cap, (uint64_t)tris * 2 + 65536);
	LUCID_CHECK(r->bin_tris.reserve(bt * 4));
	r->bin_tris_cap = bt;
	r->quad_cap = cap;
	return LUCID_OK;
}

int grow_after_overflow(lucid_renderer *r) {
	// this frame's overflow bits and those of earlier frames it was queued behind
	uint32_t flags = r->status[kCtrlOverflow] | r->status[kCtrlOverflowCarry];
	if(!flags)
		return LUCID_OK;
	if(flags & kOvfBinLists) {
		uint64_t need = std::max<uint64_t>(r->status[kCtrlTotalBinTris], r->bin_tris_cap * 2);
		LUCID_CHECK(r->bin_tris.reserve(need * 4));
		r->bin_tris_cap = need;
	}
	if(flags & kOvfRecords) {
		uint64_t need = std::max<uint64_t>((uint64_t)r->status[kCtrlRecordCursor],
										   r->high_records_cap * 2);
		LUCID_CHECK(r->high_records.reserve(need * 16));
		r->high_records_cap = need;
	}
	if(flags & kOvfSpill) {
		uint64_t need = std::max<uint64_t>((uint64_t)r->status[kCtrlSpillCursor], r->spill_cap * 2);
		LUCID_CHECK(r->spill.reserve(need * 8));
		r->spill_cap = need;
	}
	if(flags & kOvfLists) {
		uint64_t low = (uint64_t)r->bin_count * 16 * kRegSortCap;
		uint64_t need = std::max<uint64_t>(low + (uint64_t)r->status[kCtrlListCursor],
										   r->blk_lists_cap * 2);
		LUCID_CHECK(r->blk_lists.reserve(need * 4));
		r->blk_lists_cap = need;
	}
	if(flags & kOvfSorted) {
		uint64_t need = std::max<uint64_t>((uint64_t)r->status[kCtrlSortedCursor], r->sorted_cap * 2);
		LUCID_CHECK(r->sorted_pool.reserve(need * 4));
		r->sorted_cap = need;
	}
	if(flags & kOvfFallback) {
		uint64_t need = std::max<uint64_t>((uint64_t)r->status[kCtrlFallbackCount],
										   (uint64_t)r->fallback_cap * 2);
		LUCID_CHECK(r->fallback_queue.reserve(need * 16));
		r->fallback_cap = (uint32_t)need;
	}
	if(flags & kOvfGradRows) {
		uint64_t need = std::max<uint64_t>(
			(uint64_t)r->status[kCtrlBwdRows] + r->status[kCtrlBwdExtra], r->grad_rows_cap * 2);
		LUCID_CHECK(r->grad_rows.reserve(need * kGradWords * 4));
		LUCID_CHECK(r->grad_row_tri.reserve(need * 4));
		r->grad_rows_cap = need;
	}
	return LUCID_ERR_CAPACITY;
}

bool tuning_valid(const lucid_tuning &t) {
	return (t.slice_words == 1 || t.slice_words == 2) && t.presort_min >= 64 &&
		   t.presort_min <= kRegSortCap && t.samples_per_wave >= 256 && t.samples_per_wave <= 65536 &&
		   (t.window == 1 || t.window == 4);
}

// Per-frame constants of a width x height frame (bin grid, options, tuning, LucidConfig).
void fill_const(int width, int height, uint32_t opts, const lucid_tuning &tuning,
				const lucid_config &cfg, int64_t nq, int ninst, int color_mode, bool has_normals,
				FrameConst &c) {
	memset(&c, 0, sizeof(c));
	c.width = width, c.height = height;
	c.bin_count_x = (width + kBinSize - 1) / kBinSize;
	c.bin_count_y = (height + kBinSize - 1) / kBinSize;
	c.bin_count = c.bin_count_x * c.bin_count_y;
	c.num_quads = (int)nq;
	c.num_instances = ninst;
	c.num_bin_wgs = kNumBinWgs;
	c.opts = opts;
	c.backface = cfg.enable_backface_culling;
	c.presort_min = (uint32_t)tuning.presort_min;
	c.samples_per_wave = (uint32_t)tuning.samples_per_wave;
	c.slice_words = tuning.slice_words;
	c.window = tuning.window;
	c.color_mode = color_mode;
	c.has_normals = has_normals ? 1 : 0;
	for(int col = 0; col < 4; col++) {
		const lucid_vec4 &v = cfg.view_proj_matrix.col[col];
		c.vp[col * 4 + 0] = v.x, c.vp[col * 4 + 1] = v.y, c.vp[col * 4 + 2] = v.z,
					   c.vp[col * 4 + 3] = v.w;
	}
	const lucid_frustum &f = cfg.frustum;
	float o[3] = {f.ws_origin0.x, f.ws_origin0.y, f.ws_origin0.z};
	float d0[3] = {f.ws_dir0.x, f.ws_dir0.y, f.ws_dir0.z};
	float dx[3] = {f.ws_dirx.x, f.ws_dirx.y, f.ws_dirx.z};
	float dy[3] = {f.ws_diry.x, f.ws_diry.y, f.ws_diry.z};
	for(int i = 0; i < 3; i++) {
		c.origin[i] = o[i], c.dir0[i] = d0[i], c.dirx[i] = dx[i], c.diry[i] = dy[i];
		// quad_setup.glsl:397-398: ws_dir0 + (ws_dirx + ws_diry) * 0.5
		c.ray_dir0[i] = d0[i] + (dx[i] + dy[i]) * 0.5f;
	}
	const lucid_lighting &l = cfg.lighting;
	c.ambient[0] = l.ambient_color.x, c.ambient[1] = l.ambient_color.y, c.ambient[2] = l.ambient_color.z;
	c.sun_color[0] = l.sun_color.x, c.sun_color[1] = l.sun_color.y, c.sun_color[2] = l.sun_color.z;
	c.sun_dir[0] = l.sun_dir.x, c.sun_dir[1] = l.sun_dir.y, c.sun_dir[2] = l.sun_dir.z;
	c.sun_power = l.sun_power, c.ambient_power = l.ambient_power;
	bool ident = true;
	for(int k = 0; k < 3; k++) // the lighting factor is ambient + sun * light (light_factor)
		ident = ident && c.ambient[k] * c.ambient_power == 1.0f && c.sun_color[k] * c.sun_power == 0.0f;
	c.identity_light = ident ? 1 : 0;
	c.background[0] = cfg.background_color.x, c.background[1] = cfg.background_color.y;
	c.background[2] = cfg.background_color.z, c.background[3] = cfg.background_color.w;
}

} // namespace

extern "C" {

const char *lucid_error_string(int status) {
	switch(status) {
	case LUCID_OK: return "ok";
	case LUCID_ERR_INVALID_ARG: return "invalid argument";
	case LUCID_ERR_OUT_OF_MEMORY: return "out of device memory";
	case LUCID_ERR_HIP: return "HIP runtime error";
	case LUCID_ERR_NO_DEVICE: return "no HIP device";
	case LUCID_ERR_CAPACITY: return "internal capacity exceeded; buffers grown, render again";
	case LUCID_ERR_UNSUPPORTED: return "unsupported input for this entry point";
	default: return "unknown error";
	}
}

int lucid_create(const lucid_create_info *ci, lucid_renderer **out) {
	if(!ci || !out || ci->view_width <= 0 || ci->view_height <= 0 || ci->view_width > kMaxWidth ||
	   ci->view_height > kMaxHeight || ci->max_quads < 0 || ci->max_quads > kMaxQuads)
		return LUCID_ERR_INVALID_ARG;
	*out = nullptr;
	int ndev = 0;
	if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
		return LUCID_ERR_NO_DEVICE;
	if(ci->device < 0 || ci->device >= ndev)
		return LUCID_ERR_INVALID_ARG;
	LUCID_CHECK(hipSetDevice(ci->device));
	lucid_renderer *r = new(std::nothrow) lucid_renderer();
	if(!r)
		return LUCID_ERR_OUT_OF_MEMORY;
	r->device = ci->device;
	r->width = ci->view_width, r->height = ci->view_height;
	r->opts = ci->opts;
	r->bin_count_x = (r->width + kBinSize - 1) / kBinSize;
	r->bin_count_y = (r->height + kBinSize - 1) / kBinSize;
	r->bin_count = r->bin_count_x * r->bin_count_y;
	const int NB = r->bin_count;
	int rc = LUCID_OK;
	auto chk = [&](hipError_t e) {
		if(e != hipSuccess && rc == LUCID_OK)
			rc = hip_status(e);
	};
	chk(r->info.reserve(((size_t)LUCID_INFO_U32_SIZE + 10 * (size_t)NB) * 4));
	chk(r->ctrl.reserve(kCtrlAllocWords * 4));
	chk(r->wg_hist.reserve((size_t)kNumBinWgs * 2 * NB * 4));
	chk(r->low_records.reserve((size_t)NB * 4 * kLowRowCap * 16));
	// one 16-byte item per pixel sent to the exact sort; grows after an overflow
	r->fallback_cap = (uint32_t)NB * 16 * 8;
	chk(r->fallback_queue.reserve((size_t)r->fallback_cap * 16));
	r->high_records_cap = 1u << 22;
	chk(r->high_records.reserve(r->high_records_cap * 16));
	r->spill_cap = 1u << 20;
	chk(r->spill.reserve(r->spill_cap * 8));
	chk(r->blk_slot.reserve((size_t)NB * 16 * 16));
	chk(r->blk_flags.reserve((size_t)NB * 16 * 4));
	chk(r->blk_sorted.reserve((size_t)NB * 16 * 4 * 2)); // then the long-list queue
	chk(r->blk_count.reserve((size_t)NB * 16 * 4));
	chk(r->band_count.reserve((size_t)NB * 4 * 4));
	r->blk_lists_cap = (uint64_t)NB * 16 * kRegSortCap + (1u << 22);
	chk(r->blk_lists.reserve(r->blk_lists_cap * 4));
	r->sorted_cap = 1u << 22;
	chk(r->sorted_pool.reserve(r->sorted_cap * 4));
	chk(hipHostMalloc((void **)&r->status, kCtrlWords * 4, hipHostMallocDefault));
	chk(hipEventCreateWithFlags(&r->done, hipEventDisableTiming));
	r->timers = (ci->opts & LUCID_OPT_TIMERS) != 0;
	if(r->timers)
		for(hipEvent_t &e : r->stage_ev)
			chk(hipEventCreate(&e));
	if(rc == LUCID_OK) {
		int e = ensure_quad_capacity(r, ci->max_quads > 0 ? ci->max_quads : (1 << 20));
		if(e != LUCID_OK)
			rc = e;
	}
	if(rc != LUCID_OK) {
		delete r;
		return rc;
	}
	memset(r->status, 0, kCtrlWords * 4);
	*out = r;
	return LUCID_OK;
}

void lucid_destroy(lucid_renderer *r) {
	if(!r)
		return;
	(void)hipSetDevice(r->device);
	if(r->frame_pending)
		(void)hipEventSynchronize(r->done);
	delete r;
}

namespace {

// Checks the instance table and input arrays of a frame (uploadInstances,
// src/lucid_renderer.cpp:352-429): writes the exclusive prefix of the instances' quad counts
// (num_instances + 1 entries), the frame's quad count and its colour mode.
// A texture atlas descriptor is usable: texels, a positive size, 1 to 16 levels, no level
// below 1 x 1 before the last.
bool texture_valid(const lucid_texture &t) {
	if(!t.texels || t.width <= 0 || t.height <= 0 || t.num_levels < 1 ||
	   t.num_levels > kMaxTexLevels)
		return false;
	const int top = std::max(t.width, t.height);
	int levels = 1;
	while(levels < kMaxTexLevels && (top >> levels) > 0)
		levels++;
	return t.num_levels <= levels;
}

// Some instance of the frame samples an albedo texture.
bool any_textured(const lucid_frame_inputs *in) {
	for(int i = 0; i < in->num_instances && in->instances; i++)
		if(in->instances[i].flags & LUCID_INST_HAS_ALBEDO_TEXTURE)
			return true;
	return false;
}

// Descriptor of an atlas for the kernels: the first texel of each mip level.
TexDesc tex_desc(const lucid_texture &t) {
	TexDesc d;
	memset(&d, 0, sizeof(d));
	if(!texture_valid(t))
		return d;
	d.texels = t.texels, d.width = t.width, d.height = t.height, d.levels = t.num_levels;
	uint32_t off = 0;
	for(int l = 0; l < t.num_levels; l++) {
		d.offset[l] = off;
		off += (uint32_t)std::max(1, t.width >> l) * (uint32_t)std::max(1, t.height >> l);
	}
	return d;
}

int validate_inputs(const lucid_frame_inputs *in, int *prefix, int64_t &nq, int &color_mode,
					bool &textured) {
	if(!in->config || in->num_instances < 0 || in->num_instances > kMaxInstances ||
	   (in->num_instances > 0 && !in->instances))
		return LUCID_ERR_INVALID_ARG;
	nq = 0;
	textured = false;
	bool has_colors_flag = false;
	for(int i = 0; i < in->num_instances; i++) {
		const lucid_instance_data &ins = in->instances[i];
		if(ins.num_quads < 0 || ins.index_offset < 0)
			return LUCID_ERR_INVALID_ARG;
		if(ins.flags & LUCID_INST_HAS_ALBEDO_TEXTURE) {
			const lucid_texture &t = (ins.flags & LUCID_INST_TEX_OPAQUE) ? in->opaque_texture
																		  : in->transparent_texture;
			if(!in->tex_coords || !texture_valid(t) ||
			   ((ins.flags & LUCID_INST_HAS_UV_RECT) && !in->instance_uv_rects))
				return LUCID_ERR_INVALID_ARG;
			textured = true;
		}
		if(ins.flags & LUCID_INST_HAS_VERTEX_COLORS)
			has_colors_flag = true;
		prefix[i] = (int)nq;
		nq += ins.num_quads;
		if(nq >= kMaxQuads)
			return LUCID_ERR_INVALID_ARG;
	}
	prefix[in->num_instances] = (int)nq;
	if(nq > 0 && (!in->quad_indices || !in->positions))
		return LUCID_ERR_INVALID_ARG;
	color_mode = kColorNone;
	if(in->colors_rgba8)
		color_mode = kColorRGBA8;
	else if(in->colors_rgb)
		color_mode = kColorF32;
	if(has_colors_flag && color_mode == kColorNone)
		return LUCID_ERR_INVALID_ARG;
	return LUCID_OK;
}

// Validates the frame inputs, uploads instance data when it changed, grows capacities and
// fills the per-frame constants and buffer table (shared by the forward and backward).
int prepare_frame(lucid_renderer *r, const lucid_frame_inputs *in, const lucid_frame_outputs *out,
				  hipStream_t stream, FrameConst &c, FrameBufs &b, int64_t &nq_out) {
	if(!in->config || in->num_instances < 0 || in->num_instances > kMaxInstances ||
	   (in->num_instances > 0 && !in->instances))
		return LUCID_ERR_INVALID_ARG;
	LUCID_CHECK(hipSetDevice(r->device));

	// --- instances (uploadInstances, src/lucid_renderer.cpp:352-429)
	std::vector<int> prefix(in->num_instances + 1, 0);
	int64_t nq = 0;
	int color_mode = kColorNone;
	bool textured = false;
	int ve = validate_inputs(in, prefix.data(), nq, color_mode, textured);
	if(ve != LUCID_OK)
		return ve;

	bool inst_changed = r->host_instances.size() != (size_t)in->num_instances ||
						(in->num_instances &&
						 memcmp(r->host_instances.data(), in->instances,
								sizeof(lucid_instance_data) * in->num_instances) != 0);
	bool colors_null = in->instance_colors == nullptr;
	bool colors_changed = colors_null != r->inst_colors_null ||
						  (!colors_null &&
						   (r->host_inst_colors.size() != (size_t)in->num_instances ||
							memcmp(r->host_inst_colors.data(), in->instance_colors,
								   4 * (size_t)in->num_instances) != 0));
	const bool rects_null = in->instance_uv_rects == nullptr;
	const bool rects_changed = rects_null != r->uv_rects_null ||
							   (!rects_null &&
								(r->host_uv_rects.size() != (size_t)in->num_instances ||
								 memcmp(r->host_uv_rects.data(), in->instance_uv_rects,
										sizeof(lucid_vec4) * (size_t)in->num_instances) != 0));
	const bool tex_alloc = textured && !r->tex_enabled;
	// An earlier frame that has not been waited for shares every buffer with this one. When
	// it has finished (or must finish first, because it reads the instance table replaced or
	// the buffers reallocated below), its overflow is taken now: the buffers grow before this
	// frame, and the next lucid_finish reports it. While it still runs, this frame is queued
	// behind it on `stream`, and its overflow bits reach the host with this frame's status.
	if(r->frame_pending) {
		if(inst_changed || colors_changed || rects_changed || tex_alloc || nq > r->quad_cap)
			LUCID_CHECK(hipEventSynchronize(r->done));
		if(hipEventQuery(r->done) == hipSuccess) {
			r->frame_pending = false;
			int g = grow_after_overflow(r);
			if(g != LUCID_OK && g != LUCID_ERR_CAPACITY)
				return g;
			if(g == LUCID_ERR_CAPACITY)
				r->unreported_overflow = true;
			memset(r->status, 0, kCtrlWords * 4);
		} else if(stream != r->last_stream) {
			LUCID_CHECK(hipStreamWaitEvent(stream, r->done, 0));
		}
	}
	int e = ensure_quad_capacity(r, nq);
	if(e != LUCID_OK)
		return e;
	if(tex_alloc) {
		r->tex_enabled = true;
		LUCID_CHECK(r->tri_tex.reserve((size_t)r->quad_cap * 2 * 32));
	}
	if(rects_changed) {
		if(!rects_null && in->num_instances) {
			LUCID_CHECK(r->inst_uv_rects.reserve(sizeof(lucid_vec4) * (size_t)in->num_instances));
			LUCID_CHECK(hipMemcpy(r->inst_uv_rects.ptr, in->instance_uv_rects,
								  sizeof(lucid_vec4) * (size_t)in->num_instances,
								  hipMemcpyHostToDevice));
			r->host_uv_rects.assign(in->instance_uv_rects,
									in->instance_uv_rects + in->num_instances);
		} else {
			r->host_uv_rects.clear();
		}
		r->uv_rects_null = rects_null;
	}

	if(inst_changed || colors_changed) {
		size_t n = std::max(1, in->num_instances);
		LUCID_CHECK(r->instances.reserve(n * sizeof(lucid_instance_data)));
		LUCID_CHECK(r->inst_prefix.reserve((n + 1) * 4));
		LUCID_CHECK(r->inst_colors.reserve(n * 4));
		if(in->num_instances) {
			LUCID_CHECK(hipMemcpy(r->instances.ptr, in->instances,
								  sizeof(lucid_instance_data) * in->num_instances,
								  hipMemcpyHostToDevice));
			if(!colors_null)
				LUCID_CHECK(hipMemcpy(r->inst_colors.ptr, in->instance_colors,
									  4 * (size_t)in->num_instances, hipMemcpyHostToDevice));
		}
		LUCID_CHECK(hipMemcpy(r->inst_prefix.ptr, prefix.data(), 4 * prefix.size(),
							  hipMemcpyHostToDevice));
		r->host_instances.assign(in->instances, in->instances + in->num_instances);
		r->inst_colors_null = colors_null;
		if(!colors_null)
			r->host_inst_colors.assign(in->instance_colors, in->instance_colors + in->num_instances);
	}

	fill_const(r->width, r->height, r->opts, r->tuning, *in->config, nq, in->num_instances,
			   color_mode, in->normals != nullptr, c);
	c.keep_overflow = r->frame_pending ? 1u : 0u;
	c.has_textures = textured ? 1 : 0;
	memset(&b, 0, sizeof(b));
	b.instances = r->instances.as<lucid_instance_data>();
	b.inst_prefix = r->inst_prefix.as<int>();
	b.inst_colors = colors_null ? nullptr : r->inst_colors.as<uint32_t>();
	b.quad_indices = in->quad_indices;
	b.positions = in->positions;
	b.colors_rgba8 = in->colors_rgba8;
	b.colors_rgb = in->colors_rgb;
	b.tri_alpha = in->tri_alpha;
	b.normals = in->normals;
	b.tex_coords = in->tex_coords;
	b.inst_uv_rects = rects_null ? nullptr : r->inst_uv_rects.as<float4>();
	b.tex[0] = tex_desc(in->opaque_texture);
	b.tex[1] = tex_desc(in->transparent_texture);
	b.tri_tex = r->tri_tex.as<uint4>();
	b.info = r->info.as<uint32_t>();
	b.ctrl = r->ctrl.as<uint32_t>();
	b.tri_half = (uint32_t)r->quad_cap;
	b.quad_aabb = r->quad_aabb.as<uint32_t>();
	b.tri_scan = r->tri_scan.as<uint4>();
	b.tri_depth = r->tri_depth.as<uint4>();
	b.tri_shade = r->tri_shade.as<uint4>();
	b.large_list = r->large_list.as<uint32_t>();
	b.wg_hist = r->wg_hist.as<uint32_t>();
	b.bin_quads = r->bin_quads.as<uint32_t>();
	b.bin_tris = r->bin_tris.as<uint32_t>();
	b.bin_quads_cap = r->bin_quads_cap;
	b.bin_tris_cap = r->bin_tris_cap;
	b.low_records = r->low_records.as<uint4>();
	b.high_records = r->high_records.as<uint4>();
	b.high_records_cap = r->high_records_cap;
	b.fallback_queue = r->fallback_queue.as<uint4>();
	b.fallback_cap = r->fallback_cap;
	b.spill = r->spill.as<uint64_t>();
	b.spill_cap = r->spill_cap;
	b.blk_slot = r->blk_slot.as<uint4>();
	b.blk_flags = r->blk_flags.as<uint32_t>();
	b.blk_sorted = r->blk_sorted.as<uint32_t>();
	b.big_queue = b.blk_sorted + (size_t)r->bin_count * 16;
	b.blk_lists = r->blk_lists.as<uint32_t>();
	b.blk_lists_cap = r->blk_lists_cap;
	b.blk_count = r->blk_count.as<uint32_t>();
	b.band_count = r->band_count.as<uint32_t>();
	b.sorted_pool = r->sorted_pool.as<uint32_t>();
	b.sorted_cap = r->sorted_cap;
	b.image = out ? out->image_rgba8 : nullptr;
	b.image_f = out ? out->image_rgba32f : nullptr;
	b.frag_counts = out ? out->frag_counts : nullptr;
	nq_out = nq;
	return LUCID_OK;
}

} // namespace

namespace {

// Backward buffers at their current capacities (first use: optimistic sizes; a frame that
// needs more reports it and they grow), described in g.
int backward_buffers(lucid_renderer *r, const float *grad_image, BwdBufs &g) {
	const size_t slots = (size_t)r->bin_count * 16;
	LUCID_CHECK(r->grad_row_base.reserve(slots * 4));
	LUCID_CHECK(r->grad_range_base.reserve(slots * 4));
	if(r->grad_rows_cap == 0) {
		r->grad_rows_cap = 1u << 20;
		LUCID_CHECK(r->grad_rows.reserve(r->grad_rows_cap * kGradWords * 4));
		LUCID_CHECK(r->grad_row_tri.reserve(r->grad_rows_cap * 4));
	}
	if(r->grad_seq_rows_cap == 0) {
		r->grad_seq_rows_cap = 1u << 16;
		r->grad_start_cap = 1u << 12;
		LUCID_CHECK(r->grad_seq.reserve(r->grad_seq_rows_cap * 64));
		LUCID_CHECK(r->grad_start.reserve(r->grad_start_cap * 64 * sizeof(BwdStart)));
	}
	LUCID_CHECK(r->grad_rec.reserve(r->grad_start_cap * 64 * sizeof(BwdRecord)));
	LUCID_CHECK(r->grad_pix_flags.reserve((size_t)r->width * r->height));
	g.grad_image = (const float4 *)grad_image;
	g.row_base = r->grad_row_base.as<uint32_t>();
	g.range_base = r->grad_range_base.as<uint32_t>();
	g.rows = r->grad_rows.as<float>();
	g.row_tri = r->grad_row_tri.as<uint32_t>();
	g.rows_cap = r->grad_rows_cap;
	g.start = r->grad_start.as<BwdStart>();
	g.start_cap = r->grad_start_cap;
	g.seq = r->grad_seq.as<uint8_t>();
	g.seq_rows_cap = r->grad_seq_rows_cap;
	g.rec = r->grad_rec.as<BwdRecord>();
	g.pix_flags = r->grad_pix_flags.as<uint8_t>();
	return LUCID_OK;
}

// Grows the backward buffers to a frame's needs: block_rows gradient rows of the block
// kernels (each with its 64-byte recorded order), `ranges` depth ranges (start states and
// records) and `extra` rows of exactly sorted pixels. Contents are not kept, so a recorded
// pass A no longer stands after a growth.
int grow_backward(lucid_renderer *r, uint64_t block_rows, uint64_t ranges, uint64_t extra) {
	if(block_rows > r->grad_seq_rows_cap) {
		r->grad_seq_rows_cap = block_rows + block_rows / 8;
		LUCID_CHECK(r->grad_seq.reserve(r->grad_seq_rows_cap * 64));
	}
	if(ranges > r->grad_start_cap) {
		r->grad_start_cap = ranges + ranges / 8;
		LUCID_CHECK(r->grad_start.reserve(r->grad_start_cap * 64 * sizeof(BwdStart)));
		LUCID_CHECK(r->grad_rec.reserve(r->grad_start_cap * 64 * sizeof(BwdRecord)));
	}
	const uint64_t want = block_rows + std::max<uint64_t>(std::max<uint64_t>(1u << 16, block_rows / 8),
														  extra + extra / 8);
	if(want > r->grad_rows_cap) {
		LUCID_CHECK(r->grad_rows.reserve(want * kGradWords * 4));
		LUCID_CHECK(r->grad_row_tri.reserve(want * 4));
		r->grad_rows_cap = want;
	}
	r->retained.recorded = false;
	return LUCID_OK;
}

// Records the inputs the buffers are being built from; `pending` until the frame finishes.
void retain_frame(lucid_renderer *r, const lucid_frame_inputs *in, bool pending) {
	lucid_renderer::Retained &k = r->retained;
	k.in = *in;
	k.cfg = *in->config;
	k.tuning = r->tuning;
	k.instances.assign(in->instances, in->instances + std::max(in->num_instances, 0));
	k.in.instances = nullptr;
	k.in.config = nullptr;
	k.pending = pending;
	k.valid = !pending;
	k.record = k.recorded = false;
}

// The buffers hold the setup, bins and block lists of exactly these inputs.
bool retained_matches(const lucid_renderer *r, const lucid_frame_inputs *in) {
	const lucid_renderer::Retained &k = r->retained;
	if(!k.valid || k.pending || in->num_instances != k.in.num_instances ||
	   in->instance_colors != k.in.instance_colors || in->quad_indices != k.in.quad_indices ||
	   in->positions != k.in.positions || in->num_verts != k.in.num_verts ||
	   in->colors_rgba8 != k.in.colors_rgba8 || in->colors_rgb != k.in.colors_rgb ||
	   in->tri_alpha != k.in.tri_alpha || in->normals != k.in.normals)
		return false;
	if(memcmp(&k.cfg, in->config, sizeof(lucid_config)) != 0 ||
	   memcmp(&k.tuning, &r->tuning, sizeof(lucid_tuning)) != 0)
		return false;
	return in->num_instances <= 0 ||
		   memcmp(k.instances.data(), in->instances,
				  sizeof(lucid_instance_data) * (size_t)in->num_instances) == 0;
}

} // namespace

int lucid_render(lucid_renderer *r, const lucid_frame_inputs *in, const lucid_frame_outputs *out,
				 void *stream_ptr) {
	return lucid_render_ex(r, in, out, 0u, stream_ptr);
}

int lucid_render_ex(lucid_renderer *r, const lucid_frame_inputs *in,
					const lucid_frame_outputs *out, uint32_t flags, void *stream_ptr) {
	if(!r || !in || !out || (flags & ~LUCID_RENDER_RECORD_BACKWARD))
		return LUCID_ERR_INVALID_ARG;
	const bool record = (flags & LUCID_RENDER_RECORD_BACKWARD) != 0;
	// the recorded pass is the backward's: fp32 samples of fp32 colours, "over" blending
	if(record && (!(r->opts & LUCID_OPT_FP32_SAMPLES) || !in->colors_rgb || !in->tri_alpha ||
				  in->colors_rgba8 ||
				  (r->opts & (LUCID_OPT_ADDITIVE_BLENDING | LUCID_OPT_ALPHA_THRESHOLD |
							  LUCID_OPT_UNSORTED)) ||
				  any_textured(in)))
		return LUCID_ERR_UNSUPPORTED;
	hipStream_t stream = (hipStream_t)stream_ptr;
	FrameConst c;
	FrameBufs b;
	int64_t nq = 0;
	int pe = prepare_frame(r, in, out, stream, c, b, nq);
	if(pe != LUCID_OK)
		return pe;
	BwdBufs g;
	if(record) {
		pe = backward_buffers(r, nullptr, g);
		if(pe != LUCID_OK)
			return pe;
	}
	r->retained.valid = r->retained.pending = false; // the buffers are rebuilt below
	(void)hipGetLastError(); // launches below report their own status, not an older call's

	uint32_t n_info = LUCID_INFO_U32_SIZE + 10u * (uint32_t)r->bin_count;
	if(r->timers)
		LUCID_CHECK(hipEventRecord(r->stage_ev[0], stream));
	LUCID_CHECK(launch_setup(c, b, n_info, stream));
	if(r->timers)
		LUCID_CHECK(hipEventRecord(r->stage_ev[1], stream));
	LUCID_CHECK(launch_binning(c, b, stream));
	if(r->timers)
		LUCID_CHECK(hipEventRecord(r->stage_ev[2], stream));
	if(record)
		LUCID_CHECK(launch_raster_recording(c, b, g, stream, r->timers ? r->stage_ev + 3 : nullptr));
	else
		LUCID_CHECK(launch_raster(c, b, stream, r->timers ? r->stage_ev + 3 : nullptr));
	LUCID_CHECK(hipMemcpyAsync(r->status, r->ctrl.ptr, kCtrlWords * 4, hipMemcpyDeviceToHost,
							   stream));
	LUCID_CHECK(hipEventRecord(r->done, stream));
	r->frame_pending = true;
	r->last_num_quads = nq;
	r->last_stream = stream;
	retain_frame(r, in, true);
	r->retained.record = record;
	return LUCID_OK;
}

int lucid_frame_recorded(const lucid_renderer *r) {
	return r && r->retained.valid && !r->retained.pending && r->retained.recorded ? 1 : 0;
}

int lucid_backward(lucid_renderer *r, const lucid_frame_inputs *in, const float *grad_image,
				   float *grad_colors_rgb, float *grad_tri_alpha, void *stream_ptr) {
	return lucid_backward_ex(r, in, grad_image, grad_colors_rgb, grad_tri_alpha, 0u, stream_ptr);
}

int lucid_backward_ex(lucid_renderer *r, const lucid_frame_inputs *in, const float *grad_image,
					  float *grad_colors_rgb, float *grad_tri_alpha, uint32_t flags,
					  void *stream_ptr) {
	if(!r || !in || !grad_image || (!grad_colors_rgb && !grad_tri_alpha) ||
	   (flags & ~LUCID_BACKWARD_REUSE_FRAME))
		return LUCID_ERR_INVALID_ARG;
	// the adjoint is of the fp32-colour, fp32-sample, exact "over" forward, untextured
	if(!in->colors_rgb || !in->tri_alpha || in->colors_rgba8 ||
	   (r->opts & (LUCID_OPT_ADDITIVE_BLENDING | LUCID_OPT_ALPHA_THRESHOLD | LUCID_OPT_UNSORTED)) ||
	   any_textured(in))
		return LUCID_ERR_UNSUPPORTED;
	hipStream_t stream = (hipStream_t)stream_ptr;
	LUCID_CHECK(hipSetDevice(r->device));
	if(r->frame_pending) {
		int f = lucid_finish(r);
		if(f != LUCID_OK && f != LUCID_ERR_CAPACITY)
			return f;
	}
	FrameConst c;
	FrameBufs b;
	int64_t nq = 0;
	int pe = prepare_frame(r, in, nullptr, stream, c, b, nq);
	if(pe != LUCID_OK)
		return pe;
	const uint64_t ntri = 2 * (uint64_t)nq;
	const uint64_t nverts = (uint64_t)std::max<int64_t>(in->num_verts, 0);
	if(3 * ntri >= (1ull << 32))
		return LUCID_ERR_UNSUPPORTED;
	// corner -> vertex map, checked against the one the cached vertex grouping was built from
	const bool vtx_same_sizes = r->vtx_valid && r->vtx_ntri == ntri && r->vtx_nverts == nverts;
	LUCID_CHECK(r->vtx_keys.reserve(std::max<uint64_t>(3 * ntri, 1) * 4));
	LUCID_CHECK(r->vtx_sorted.reserve(std::max<uint64_t>(3 * ntri, 1) * 4));
	LUCID_CHECK(r->vtx_corners.reserve(std::max<uint64_t>(3 * ntri, 1) * 4));
	LUCID_CHECK(r->vtx_iota.reserve(std::max<uint64_t>(3 * ntri, 1) * 4));
	BwdBufs g;
	pe = backward_buffers(r, grad_image, g);
	if(pe != LUCID_OK)
		return pe;

	if(!r->aux) {
		LUCID_CHECK(hipStreamCreateWithFlags(&r->aux, hipStreamNonBlocking));
		LUCID_CHECK(hipEventCreateWithFlags(&r->aux_fork, hipEventDisableTiming));
		LUCID_CHECK(hipEventCreateWithFlags(&r->aux_join, hipEventDisableTiming));
		LUCID_CHECK(hipEventCreateWithFlags(&r->aux_scale, hipEventDisableTiming));
		LUCID_CHECK(hipEventCreateWithFlags(&r->aux_order, hipEventDisableTiming));
	}
	LUCID_CHECK(r->red_counts.reserve((ntri + 1) * 4));
	LUCID_CHECK(r->red_offsets.reserve((ntri + 1) * 4));
	LUCID_CHECK(r->red_grouped.reserve(r->grad_rows_cap * 4));
	LUCID_CHECK(r->tri_grad.reserve(std::max<uint64_t>(ntri, 1) * kGradWords * 4));
	const size_t temp_bytes =
		std::max<size_t>(reduce_temp_bytes((uint32_t)(3 * ntri), (uint32_t)(ntri + 1)), 16);
	LUCID_CHECK(r->red_temp.reserve(temp_bytes));
	ReduceBufs rb;
	rb.counts = r->red_counts.as<uint32_t>();
	rb.offsets = r->red_offsets.as<uint32_t>();
	rb.grouped = r->red_grouped.as<uint32_t>();
	rb.tri_grad = r->tri_grad.as<float>();
	rb.vkeys = r->vtx_keys.as<uint32_t>();
	rb.vsorted = r->vtx_sorted.as<uint32_t>();
	rb.vcorners = r->vtx_corners.as<uint32_t>();
	rb.viota = r->vtx_iota.as<uint32_t>();
	rb.temp = r->red_temp.ptr;
	rb.temp_bytes = r->red_temp.bytes;

	// Forward stages (the fragment lists are recomputed), the gradient row layout, the
	// recorded blend order (pass A) and the exact per-pixel adjoints; then the block rows
	// (pass B) while the helper stream groups the rows by triangle; then the sums. Buffers
	// have the sizes of the previous pass: a frame that needs more reports an overflow, and
	// the buffers grow for the caller's next attempt.
	const uint32_t n_info = LUCID_INFO_U32_SIZE + 10u * (uint32_t)r->bin_count;
	const bool regroup = !vtx_same_sizes;
	r->vtx_valid = false; // vtx_keys is rewritten below
	const bool reuse = (flags & LUCID_BACKWARD_REUSE_FRAME) && retained_matches(r, in);
	// a frame rendered with LUCID_RENDER_RECORD_BACKWARD already holds pass A's results
	const bool recorded = reuse && r->retained.recorded;
	if(recorded) {
		LUCID_CHECK(launch_backward_recorded(c, b, g, stream));
	} else if(reuse) {
		LUCID_CHECK(launch_backward_reuse(c, b, g, stream));
	} else {
		r->retained.valid = r->retained.pending = false; // the buffers are rebuilt below
		LUCID_CHECK(launch_setup(c, b, n_info, stream));
		LUCID_CHECK(launch_binning(c, b, stream));
		LUCID_CHECK(launch_backward_prepare(c, b, g, stream));
	}
	// The helper stream takes the gradient scale and the corner check beside pass A, then the
	// row grouping beside pass B (which needs the scale). Pass A is submitted first: the GPU
	// would otherwise wait for the host to submit the helper stream's work before it.
	LUCID_CHECK(hipEventRecord(r->aux_fork, stream));
	if(recorded)
		LUCID_CHECK(launch_backward_recorded_order(c, b, g, stream));
	else
		LUCID_CHECK(launch_backward_order(c, b, g, stream));
	LUCID_CHECK(hipEventRecord(r->aux_order, stream));
	LUCID_CHECK(hipStreamWaitEvent(r->aux, r->aux_fork, 0));
	LUCID_CHECK(launch_grad_max(c, g, b.ctrl, r->aux));
	LUCID_CHECK(reduce_check_corners(c, b, (uint32_t)ntri, rb.vkeys,
									 r->ctrl.as<uint32_t>() + kCtrlTopoChanged, r->aux));
	LUCID_CHECK(hipEventRecord(r->aux_scale, r->aux));
	// pass B is submitted before the grouping beside it: submitted the other way round, the
	// grouping's large grids took the CUs first and pass B started only after them
	LUCID_CHECK(hipStreamWaitEvent(stream, r->aux_scale, 0));
	LUCID_CHECK(launch_backward_rows(c, b, g, stream));
	LUCID_CHECK(hipStreamWaitEvent(r->aux, r->aux_order, 0));
	LUCID_CHECK(reduce_group_rows(c, b, g, rb, (uint32_t)ntri, r->aux));
	LUCID_CHECK(hipEventRecord(r->aux_join, r->aux));
	LUCID_CHECK(hipStreamWaitEvent(stream, r->aux_join, 0));
	LUCID_CHECK(reduce_sum(c, b, g, rb, (uint32_t)ntri, (uint32_t)nverts, regroup,
						   grad_colors_rgb, grad_tri_alpha, stream));
	LUCID_CHECK(hipMemcpyAsync(r->status, r->ctrl.ptr, kCtrlWords * 4, hipMemcpyDeviceToHost,
							   stream));
	LUCID_CHECK(hipStreamSynchronize(stream));
	r->last_num_quads = nq;
	r->last_stream = stream;
	const uint64_t block_rows = r->status[kCtrlBwdRows], ranges = r->status[kCtrlBwdRanges];
	if(r->status[kCtrlOverflow]) {
		// the recorded order and start states take block_rows and ranges (sizes of this
		// frame, known once its row scan ran); the rows add room for the exact pixels
		pe = grow_backward(r, block_rows, ranges, r->status[kCtrlBwdExtra]);
		if(pe != LUCID_OK)
			return pe;
		r->status[kCtrlOverflow] &= ~kOvfGradRows; // grown above
		// a reused frame stays valid when only the backward's own buffers overflowed
		const uint32_t own = kOvfFallback | kOvfSpill;
		if(!reuse || (r->status[kCtrlOverflow] & ~own))
			r->retained.valid = false;
		int gr = grow_after_overflow(r);
		memset(r->status, 0, kCtrlWords * 4);
		return gr == LUCID_OK ? LUCID_ERR_CAPACITY : gr; // this pass's results are incomplete
	}
	if(!reuse)
		retain_frame(r, in, false); // the buffers now hold this frame's setup and lists
	const bool changed = r->status[kCtrlTopoChanged] != 0;
	memset(r->status, 0, kCtrlWords * 4);
	bool grouped = grad_colors_rgb != nullptr && regroup;
	if(changed && !regroup && grad_colors_rgb) {
		// same sizes, different mesh: the vertex sums above used the old grouping
		LUCID_CHECK(reduce_vertices(rb, (uint32_t)ntri, (uint32_t)nverts, true, grad_colors_rgb,
									nullptr, stream)); // (the alpha gradients stand)
		LUCID_CHECK(hipStreamSynchronize(stream));
		grouped = true;
	}
	// the grouping matches vtx_keys if it was just built, or if nothing changed
	r->vtx_valid = grouped || (!regroup && !changed);
	r->vtx_ntri = ntri, r->vtx_nverts = nverts;
	return LUCID_OK;
}

int lucid_set_tuning(lucid_renderer *r, const lucid_tuning *t) {
	if(!r || !t || !tuning_valid(*t))
		return LUCID_ERR_INVALID_ARG;
	r->tuning = *t;
	return LUCID_OK;
}

int lucid_get_tuning(const lucid_renderer *r, lucid_tuning *t) {
	if(!r || !t)
		return LUCID_ERR_INVALID_ARG;
	*t = r->tuning;
	return LUCID_OK;
}

int lucid_autotune(lucid_renderer *r, const lucid_frame_inputs *in, const lucid_frame_outputs *out,
				   void *stream_ptr, int iters, lucid_tuning *best_out) {
	if(!r || !in || !out || iters < 1)
		return LUCID_ERR_INVALID_ARG;
	hipStream_t stream = (hipStream_t)stream_ptr;
	LUCID_CHECK(hipSetDevice(r->device));
	const lucid_tuning saved = r->tuning;
	// one frame at the current tuning settles capacities and uploads
	for(int k = 0; k < 4; k++) {
		int e = lucid_render(r, in, out, stream);
		if(e != LUCID_OK)
			return e;
		e = lucid_finish(r);
		if(e == LUCID_OK)
			break;
		if(e != LUCID_ERR_CAPACITY)
			return e;
	}
	hipEvent_t e0, e1;
	LUCID_CHECK(hipEventCreate(&e0));
	LUCID_CHECK(hipEventCreate(&e1));
	const int slice_words[] = {1, 2};
	const int presort[] = {64, 128, 256};
	const int spw[] = {1024, 2048, 4096};
	const int windows[] = {1, 4};
	lucid_tuning best = saved;
	float best_ms = 1e30f;
	int rc = LUCID_OK;
	for(int win : windows)
		for(int sw : slice_words)
			for(int pm : presort)
				for(int sp : spw) {
					r->tuning = {sw, pm, sp, win};
					int e = LUCID_OK, f = LUCID_OK;
					// A candidate that overflows a buffer (e.g. a larger pre-sorted pool) grows it
					// on its untimed first frame and is timed again with the grown buffers.
					for(int attempt = 0; attempt < 3; attempt++) {
						e = lucid_render(r, in, out, stream); // untimed: settles this candidate's sizes
						f = e == LUCID_OK ? lucid_finish(r) : LUCID_OK;
						if(e != LUCID_OK || f != LUCID_ERR_CAPACITY)
							break;
					}
					if(e == LUCID_OK && f == LUCID_OK)
						e = hip_status(hipEventRecord(e0, stream));
					for(int k = 0; k < iters && e == LUCID_OK && f == LUCID_OK; k++)
						e = lucid_render(r, in, out, stream);
					if(e == LUCID_OK && f == LUCID_OK)
						e = hip_status(hipEventRecord(e1, stream));
					int f2 = lucid_finish(r);
					if(f == LUCID_OK)
						f = f2;
					if(e != LUCID_OK || f != LUCID_OK) {
						rc = e != LUCID_OK ? e : f;
						if(rc == LUCID_ERR_CAPACITY)
							continue; // still overflowing after growth: skipped, the others still run
						break;
					}
					float ms = 0.0f;
					if(hipEventSynchronize(e1) == hipSuccess &&
					   hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < best_ms)
						best_ms = ms, best = r->tuning;
				}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	r->tuning = best_ms < 1e30f ? best : saved;
	if(best_out)
		*best_out = r->tuning;
	return best_ms < 1e30f ? LUCID_OK : (rc != LUCID_OK ? rc : LUCID_ERR_HIP);
}

// ------------------------------------------------------------------ batches

struct lucid_batch {
	int device = 0;
	std::vector<lucid_renderer *> rs;
	std::vector<FrameDesc> frames;   // the frames of the recorded launch (host copy)
	std::vector<uint32_t> pre;       // kBsCount x (n + 1) work-group prefixes
	DevBuf frames_dev, pre_dev, status_dev;
	uint32_t *status = nullptr;      // pinned: kCtrlWords per frame after the last launch
	BatchPlan plan = {};
	hipEvent_t done = nullptr;
	hipStream_t capture_stream = nullptr;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	bool pending = false;
	~lucid_batch() {
		if(exec)
			(void)hipGraphExecDestroy(exec);
		if(graph)
			(void)hipGraphDestroy(graph);
		if(done)
			(void)hipEventDestroy(done);
		if(capture_stream)
			(void)hipStreamDestroy(capture_stream);
		if(status)
			(void)hipHostFree(status);
	}
};

namespace {

constexpr uint32_t kBatchFallbackWgs = 64; // persistent exact-fallback work-groups per frame
constexpr uint32_t kBatchLongSortWgs = 16; // persistent long-list sort work-groups per frame
constexpr int kScanColsHost = 16;         // columns per work-group of bin_wg_scan_kernel

// Work-groups of each pipeline stage for one frame (the grids launch_setup, launch_binning and
// launch_raster use for it alone; the exact fallback is persistent and takes fewer).
void stage_wgs(const FrameConst &c, uint32_t w[kBsCount]) {
	const uint32_t nb = (uint32_t)c.bin_count;
	const uint32_t n_info = LUCID_INFO_U32_SIZE + 10u * nb;
	const uint32_t n_clear = std::max(n_info, (uint32_t)kCtrlAllocWords);
	w[kBsClear] = (n_clear + 255) / 256;
	w[kBsSetup] = c.num_quads > 0 ? ((uint32_t)c.num_quads + kSetupThreads - 1) / kSetupThreads : 0;
	w[kBsBinCount] = (uint32_t)c.num_bin_wgs;
	w[kBsBinScan] = (2 * nb + kScanColsHost - 1) / kScanColsHost;
	w[kBsBinCategorize] = 1;
	w[kBsBinDispatch] = (uint32_t)c.num_bin_wgs;
	w[kBsRowgenLow] = nb;
	w[kBsHighPrep] = 1;
	w[kBsRowgenHigh] = nb * kRowgenSplit;
	w[kBsHighFinalize] = nb * 16;
	w[kBsHighFinalizeLong] = kBatchLongSortWgs;
	w[kBsBlockRaster] = nb * 16;
	w[kBsBlockRasterBig] = nb * 16;
	w[kBsFallback] = kBatchFallbackWgs;
}

// Records the batch's launches into one graph: one launch per stage for all frames, then
// the frames' control words to pinned host memory.
int capture_batch(lucid_batch *bt) {
	if(bt->exec)
		(void)hipGraphExecDestroy(bt->exec), bt->exec = nullptr;
	if(bt->graph)
		(void)hipGraphDestroy(bt->graph), bt->graph = nullptr;
	hipStream_t s = bt->capture_stream;
	const size_t n = bt->rs.size();
	LUCID_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
	hipError_t err = launch_setup_batch(bt->plan, s);
	if(err == hipSuccess)
		err = launch_binning_batch(bt->plan, s);
	if(err == hipSuccess)
		err = launch_raster_batch(bt->plan, s);
	if(err == hipSuccess)
		err = hipMemcpyAsync(bt->status, bt->plan.status, n * kCtrlWords * 4, hipMemcpyDeviceToHost, s);
	hipGraph_t g = nullptr;
	hipError_t end = hipStreamEndCapture(s, &g);
	if(err != hipSuccess || end != hipSuccess) {
		if(g)
			(void)hipGraphDestroy(g);
		return hip_status(err != hipSuccess ? err : end);
	}
	bt->graph = g;
	LUCID_CHECK(hipGraphInstantiate(&bt->exec, g, nullptr, nullptr, 0));
	return LUCID_OK;
}
} // namespace

int lucid_batch_create(lucid_renderer *const *renderers, int n, lucid_batch **out) {
	if(!renderers || n <= 0 || !out)
		return LUCID_ERR_INVALID_ARG;
	for(int i = 0; i < n; i++)
		if(!renderers[i] || renderers[i]->device != renderers[0]->device || renderers[i]->timers ||
		   ((renderers[i]->opts ^ renderers[0]->opts) & LUCID_OPT_FP32_SAMPLES))
			return LUCID_ERR_INVALID_ARG; // one device and sample format; timers are per renderer
	for(int i = 0; i < n; i++)
		for(int j = 0; j < i; j++)
			if(renderers[i] == renderers[j])
				return LUCID_ERR_INVALID_ARG; // frames run concurrently: one renderer each
	lucid_batch *bt = new(std::nothrow) lucid_batch;
	if(!bt)
		return LUCID_ERR_OUT_OF_MEMORY;
	bt->device = renderers[0]->device;
	bt->rs.assign(renderers, renderers + n);
	int rc = LUCID_OK;
	if(hipSetDevice(bt->device) != hipSuccess ||
	   hipStreamCreateWithFlags(&bt->capture_stream, hipStreamNonBlocking) != hipSuccess ||
	   hipEventCreateWithFlags(&bt->done, hipEventDisableTiming) != hipSuccess ||
	   hipHostMalloc((void **)&bt->status, (size_t)n * kCtrlWords * 4, hipHostMallocDefault) != hipSuccess)
		rc = LUCID_ERR_HIP;
	if(rc == LUCID_OK) {
		rc = hip_status(bt->status_dev.reserve((size_t)n * kCtrlWords * 4));
		if(rc == LUCID_OK)
			rc = hip_status(bt->pre_dev.reserve((size_t)kBsCount * (n + 1) * 4));
		if(rc == LUCID_OK)
			rc = hip_status(bt->frames_dev.reserve((size_t)n * sizeof(FrameDesc)));
	}
	if(rc != LUCID_OK) {
		delete bt;
		return rc;
	}
	memset(bt->status, 0, (size_t)n * kCtrlWords * 4);
	*out = bt;
	return LUCID_OK;
}

void lucid_batch_destroy(lucid_batch *bt) {
	if(!bt)
		return;
	(void)hipSetDevice(bt->device);
	if(bt->pending)
		(void)hipEventSynchronize(bt->done);
	delete bt;
}

int lucid_batch_render(lucid_batch *bt, const lucid_frame_inputs *ins,
					   const lucid_frame_outputs *outs, void *stream_ptr) {
	if(!bt || !ins || !outs)
		return LUCID_ERR_INVALID_ARG;
	hipStream_t stream = (hipStream_t)stream_ptr;
	LUCID_CHECK(hipSetDevice(bt->device));
	if(bt->pending) {
		int f = lucid_batch_finish(bt);
		if(f != LUCID_OK && f != LUCID_ERR_CAPACITY)
			return f;
	}
	const size_t n = bt->rs.size();
	std::vector<FrameDesc> frames(n);
	bool tex = false, flat = true;
	for(size_t i = 0; i < n; i++) {
		lucid_renderer *r = bt->rs[i];
		if(r->frame_pending) {
			int f = lucid_finish(r);
			if(f != LUCID_OK && f != LUCID_ERR_CAPACITY)
				return f;
		}
		int64_t nq = 0;
		int e = prepare_frame(r, &ins[i], &outs[i], stream, frames[i].c, frames[i].b, nq);
		if(e != LUCID_OK)
			return e;
		// every frame runs the block kernels of the first renderer's slice width (a tuning
		// knob: results do not depend on it)
		frames[i].c.slice_words = bt->rs[0]->tuning.slice_words;
		tex = tex || frames[i].c.has_textures;
		flat = flat && frame_shade_mode(frames[i].c) >= kShadeFlat; // flat records, either window
		r->last_num_quads = nq;
		r->last_stream = stream;
		r->retained.valid = r->retained.pending = false;
	}
	// per-stage work-group prefixes: frame i owns [pre[s][i], pre[s][i + 1]) of stage s
	std::vector<uint32_t> pre((size_t)kBsCount * (n + 1), 0);
	BatchPlan plan;
	memset(&plan, 0, sizeof(plan)); // compared bytewise (padding included) below
	size_t bin_lds = 0;
	for(size_t i = 0; i < n; i++) {
		uint32_t w[kBsCount];
		stage_wgs(frames[i].c, w);
		for(int st = 0; st < kBsCount; st++)
			pre[(size_t)st * (n + 1) + i + 1] = pre[(size_t)st * (n + 1) + i] + w[st];
		bin_lds = std::max(bin_lds, (size_t)2 * frames[i].c.bin_count * sizeof(uint32_t));
	}
	for(int st = 0; st < kBsCount; st++)
		plan.total[st] = pre[(size_t)st * (n + 1) + n];
	plan.frames = bt->frames_dev.as<FrameDesc>();
	plan.pre = bt->pre_dev.as<uint32_t>();
	plan.num_frames = (uint32_t)n;
	plan.bin_lds = bin_lds;
	plan.fp32 = (bt->rs[0]->opts & LUCID_OPT_FP32_SAMPLES) != 0;
	plan.mode = tex ? kShadeTextured
					: flat ? (bt->rs[0]->tuning.window == 1 ? kShadeFlatNarrow : kShadeFlat)
						   : kShadeGeneral;
	plan.slice_words = bt->rs[0]->tuning.slice_words;
	plan.status = bt->status_dev.as<uint32_t>();
	// the launches read the frames' constants and buffers from device memory, so a frame
	// whose inputs change only needs them uploaded; the graph is recorded again only when a
	// grid, the LDS size or a kernel choice changes
	if(frames.size() != bt->frames.size() ||
	   memcmp(frames.data(), bt->frames.data(), n * sizeof(FrameDesc)) != 0) {
		LUCID_CHECK(hipMemcpy(bt->frames_dev.ptr, frames.data(), n * sizeof(FrameDesc),
							  hipMemcpyHostToDevice));
		bt->frames = frames;
	}
	if(pre != bt->pre) {
		LUCID_CHECK(hipMemcpy(bt->pre_dev.ptr, pre.data(), pre.size() * 4, hipMemcpyHostToDevice));
		bt->pre = pre;
	}
	const bool same = bt->exec && memcmp(&plan, &bt->plan, sizeof(BatchPlan)) == 0;
	bt->plan = plan;
	if(!same) {
		int e = capture_batch(bt);
		if(e != LUCID_OK)
			return e;
	}
	LUCID_CHECK(hipGraphLaunch(bt->exec, stream));
	LUCID_CHECK(hipEventRecord(bt->done, stream));
	bt->pending = true;
	return LUCID_OK;
}

int lucid_batch_finish(lucid_batch *bt) {
	if(!bt)
		return LUCID_ERR_INVALID_ARG;
	LUCID_CHECK(hipSetDevice(bt->device));
	if(!bt->pending)
		return LUCID_OK;
	LUCID_CHECK(hipEventSynchronize(bt->done));
	bt->pending = false;
	int rc = LUCID_OK;
	for(size_t i = 0; i < bt->rs.size(); i++) {
		lucid_renderer *r = bt->rs[i];
		memcpy(r->status, bt->status + i * kCtrlWords, kCtrlWords * 4);
		int g = grow_after_overflow(r);
		memset(r->status, 0, kCtrlWords * 4);
		if(g != LUCID_OK && rc == LUCID_OK)
			rc = g;
	}
	return rc;
}

} // extern "C"

// ------------------------------------------------------------------ allocation-free frames
//
// lucid_render_noalloc renders one frame without any allocation: every internal buffer of
// the frame is carved from the file-scope device array `scratch` below, whose size is fixed
// by the limits that follow. Each GPU the library runs on holds its own copy of the array
// (allocated with the code object) until the process ends. Host code never passes the
// array's name to a kernel or copy: it is a host-side stand-in, not a device address; the
// device address comes from hipGetSymbolAddress, once per device.
namespace {

constexpr int kNaMaxBins = 256;                        // e.g. 512 x 512 or 1024 x 256 pixels
constexpr int64_t kNaMaxQuads = (int64_t)1 << 16;
constexpr int kNaMaxInstances = 1024;
constexpr uint64_t kNaHighRecords = (uint64_t)1 << 19; // uint4 records of high bins
constexpr uint64_t kNaListsExtra = (uint64_t)1 << 21;  // u32 list entries of high bins
constexpr uint64_t kNaSorted = (uint64_t)1 << 19;      // u32 words of pre-sorted blocks
constexpr uint64_t kNaSpill = (uint64_t)1 << 17;       // u64 keys of deep pixels
constexpr uint32_t kNaFallback = (uint32_t)kNaMaxBins * 16 * 8; // pixels queued for the exact sort
constexpr uint64_t kNaBinQuads = (uint64_t)kNaMaxQuads * 4;
constexpr uint64_t kNaBinTris = (uint64_t)kNaMaxQuads * 4 + 65536;
constexpr int kNaMaxDevices = 64;

struct NaLayout {
	size_t info, ctrl, instances, inst_prefix, inst_colors, quad_aabb, tri_scan, tri_depth,
		tri_shade, large_list, wg_hist, bin_quads, bin_tris, low_records, high_records, fallback,
		spill, blk_slot, blk_flags, blk_sorted, blk_lists, blk_count, band_count, sorted_pool,
		total;
};

// Byte offsets of the frame's buffers in `scratch` (each 256-byte aligned), sized as
// lucid_create and ensure_quad_capacity size a renderer's, for the limits above.
constexpr NaLayout na_layout() {
	NaLayout l{};
	size_t o = 0;
	const size_t nb = kNaMaxBins, q = (size_t)kNaMaxQuads, t = 2 * q;
	size_t *fields[] = {&l.info,      &l.ctrl,         &l.instances,  &l.inst_prefix, &l.inst_colors,
						&l.quad_aabb, &l.tri_scan,     &l.tri_depth,  &l.tri_shade,   &l.large_list,
						&l.wg_hist,   &l.bin_quads,    &l.bin_tris,   &l.low_records, &l.high_records,
						&l.fallback,  &l.spill,        &l.blk_slot,   &l.blk_flags,   &l.blk_sorted,
						&l.blk_lists, &l.blk_count,    &l.band_count, &l.sorted_pool};
	const size_t bytes[] = {
		((size_t)LUCID_INFO_U32_SIZE + 10 * nb) * 4,
		(size_t)kCtrlAllocWords * 4,
		(size_t)kNaMaxInstances * sizeof(lucid_instance_data),
		((size_t)kNaMaxInstances + 1) * 4,
		(size_t)kNaMaxInstances * 4,
		q * 4,
		t * 32,
		t * 16,
		t * kRecStride * 16,
		(size_t)kLargeStripes * large_stripe_cap((uint32_t)q) * 4,
		(size_t)kNumBinWgs * 2 * nb * 4,
		kNaBinQuads * 4,
		kNaBinTris * 4,
		nb * 4 * kLowRowCap * 16,
		kNaHighRecords * 16,
		(size_t)kNaFallback * 16,
		kNaSpill * 8,
		nb * 16 * 16,
		nb * 16 * 4,
		nb * 16 * 4 * 2, // blk_sorted, then the long-list queue
		(nb * 16 * kRegSortCap + kNaListsExtra) * 4,
		nb * 16 * 4,
		nb * 4 * 4,
		kNaSorted * 4,
	};
	static_assert(sizeof(fields) / sizeof(fields[0]) == sizeof(bytes) / sizeof(bytes[0]),
				  "one size per buffer");
	for(size_t i = 0; i < sizeof(bytes) / sizeof(bytes[0]); i++) {
		*fields[i] = o;
		o = (o + bytes[i] + 255) & ~(size_t)255;
	}
	l.total = o;
	return l;
}

constexpr NaLayout kNaLayout = na_layout();
static_assert(kNaLayout.total <= ((size_t)72 << 20), "the allocation-free scratch stays modest");

} // namespace

// The frame scratch of lucid_render_noalloc (one copy per device).
static __device__ __align__(256) uint8_t scratch[kNaLayout.total];

namespace {

// Device address of `scratch` per device, looked up on the first call on that device.
uint8_t *g_scratch_ptr[kNaMaxDevices] = {};

int noalloc_scratch(int device, uint8_t **out) {
	if(device < 0 || device >= kNaMaxDevices)
		return LUCID_ERR_INVALID_ARG;
	uint8_t *p = __atomic_load_n(&g_scratch_ptr[device], __ATOMIC_ACQUIRE);
	if(!p) {
		void *addr = nullptr;
		LUCID_CHECK(hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)));
		p = (uint8_t *)addr;
		__atomic_store_n(&g_scratch_ptr[device], p, __ATOMIC_RELEASE);
	}
	*out = p;
	return LUCID_OK;
}

} // namespace

extern "C" {

int lucid_render_noalloc(int width, int height, uint32_t opts, const lucid_tuning *tuning,
						 const lucid_frame_inputs *in, const lucid_frame_outputs *out,
						 void *stream_ptr) {
	if(!in || !out || !in->config || width <= 0 || height <= 0 || width > kMaxWidth ||
	   height > kMaxHeight)
		return LUCID_ERR_INVALID_ARG;
	const int nbx = (width + kBinSize - 1) / kBinSize, nby = (height + kBinSize - 1) / kBinSize;
	if(nbx * nby > kNaMaxBins || in->num_instances > kNaMaxInstances)
		return LUCID_ERR_INVALID_ARG;
	lucid_tuning t = {1, 128, 1024, 1};
	if(tuning) {
		t = *tuning;
		if(!tuning_valid(t))
			return LUCID_ERR_INVALID_ARG;
	}
	int prefix[kNaMaxInstances + 1];
	int64_t nq = 0;
	int color_mode = kColorNone;
	if(in->num_instances < 0)
		return LUCID_ERR_INVALID_ARG;
	bool textured = false;
	int ve = validate_inputs(in, prefix, nq, color_mode, textured);
	if(ve != LUCID_OK)
		return ve;
	if(nq > kNaMaxQuads)
		return LUCID_ERR_INVALID_ARG;
	if(textured)
		return LUCID_ERR_UNSUPPORTED; // no texture-coordinate records in the fixed array
	int device = 0;
	LUCID_CHECK(hipGetDevice(&device));
	uint8_t *base = nullptr;
	int se = noalloc_scratch(device, &base);
	if(se != LUCID_OK)
		return se;
	hipStream_t stream = (hipStream_t)stream_ptr;
	const NaLayout &L = kNaLayout;

	FrameConst c;
	fill_const(width, height, opts & ~(uint32_t)LUCID_OPT_TIMERS, t, *in->config, nq,
			   in->num_instances, color_mode, in->normals != nullptr, c);
	c.keep_overflow = 0u;
	FrameBufs b;
	memset(&b, 0, sizeof(b));
	b.instances = (const lucid_instance_data *)(base + L.instances);
	b.inst_prefix = (const int *)(base + L.inst_prefix);
	b.inst_colors = in->instance_colors ? (const uint32_t *)(base + L.inst_colors) : nullptr;
	b.quad_indices = in->quad_indices;
	b.positions = in->positions;
	b.colors_rgba8 = in->colors_rgba8;
	b.colors_rgb = in->colors_rgb;
	b.tri_alpha = in->tri_alpha;
	b.normals = in->normals;
	b.info = (uint32_t *)(base + L.info);
	b.ctrl = (uint32_t *)(base + L.ctrl);
	b.tri_half = (uint32_t)kNaMaxQuads;
	b.quad_aabb = (uint32_t *)(base + L.quad_aabb);
	b.tri_scan = (uint4 *)(base + L.tri_scan);
	b.tri_depth = (uint4 *)(base + L.tri_depth);
	b.tri_shade = (uint4 *)(base + L.tri_shade);
	b.large_list = (uint32_t *)(base + L.large_list);
	b.wg_hist = (uint32_t *)(base + L.wg_hist);
	b.bin_quads = (uint32_t *)(base + L.bin_quads);
	b.bin_tris = (uint32_t *)(base + L.bin_tris);
	b.bin_quads_cap = kNaBinQuads;
	b.bin_tris_cap = kNaBinTris;
	b.low_records = (uint4 *)(base + L.low_records);
	b.high_records = (uint4 *)(base + L.high_records);
	b.high_records_cap = kNaHighRecords;
	b.fallback_queue = (uint4 *)(base + L.fallback);
	b.fallback_cap = kNaFallback;
	b.spill = (uint64_t *)(base + L.spill);
	b.spill_cap = kNaSpill;
	b.blk_slot = (uint4 *)(base + L.blk_slot);
	b.blk_flags = (uint32_t *)(base + L.blk_flags);
	b.blk_sorted = (uint32_t *)(base + L.blk_sorted);
	b.big_queue = b.blk_sorted + (size_t)kNaMaxBins * 16;
	b.blk_lists = (uint32_t *)(base + L.blk_lists);
	b.blk_lists_cap = (uint64_t)kNaMaxBins * 16 * kRegSortCap + kNaListsExtra;
	b.blk_count = (uint32_t *)(base + L.blk_count);
	b.band_count = (uint32_t *)(base + L.band_count);
	b.sorted_pool = (uint32_t *)(base + L.sorted_pool);
	b.sorted_cap = kNaSorted;
	b.image = out->image_rgba8;
	b.image_f = out->image_rgba32f;
	b.frag_counts = out->frag_counts;

	(void)hipGetLastError(); // launches below report their own status, not an older call's
	// the instance table, its quad prefix and the instance colours (uploadInstances,
	// src/lucid_renderer.cpp:352-429), staged from host memory on the frame's stream
	const size_t ninst = (size_t)in->num_instances;
	if(ninst)
		LUCID_CHECK(hipMemcpyAsync(base + L.instances, in->instances,
								   ninst * sizeof(lucid_instance_data), hipMemcpyHostToDevice,
								   stream));
	LUCID_CHECK(hipMemcpyAsync(base + L.inst_prefix, prefix, (ninst + 1) * 4,
							   hipMemcpyHostToDevice, stream));
	if(ninst && in->instance_colors)
		LUCID_CHECK(hipMemcpyAsync(base + L.inst_colors, in->instance_colors, ninst * 4,
								   hipMemcpyHostToDevice, stream));
	const uint32_t n_info = LUCID_INFO_U32_SIZE + 10u * (uint32_t)c.bin_count;
	LUCID_CHECK(launch_setup(c, b, n_info, stream));
	LUCID_CHECK(launch_binning(c, b, stream));
	LUCID_CHECK(launch_raster(c, b, stream, nullptr));
	uint32_t status[kCtrlWords];
	LUCID_CHECK(hipMemcpyAsync(status, b.ctrl, sizeof(status), hipMemcpyDeviceToHost, stream));
	LUCID_CHECK(hipStreamSynchronize(stream));
	return status[kCtrlOverflow] ? LUCID_ERR_CAPACITY : LUCID_OK;
}

int lucid_finish(lucid_renderer *r) {
	if(!r)
		return LUCID_ERR_INVALID_ARG;
	LUCID_CHECK(hipSetDevice(r->device));
	if(!r->frame_pending)
		return LUCID_OK;
	LUCID_CHECK(hipEventSynchronize(r->done));
	r->frame_pending = false;
	const bool recorded = r->status[kCtrlRecorded] != 0;
	const uint64_t block_rows = r->status[kCtrlBwdRows], ranges = r->status[kCtrlBwdRanges];
	int g = grow_after_overflow(r);
	if(r->unreported_overflow && g == LUCID_OK)
		g = LUCID_ERR_CAPACITY;
	r->unreported_overflow = false;
	memset(r->status, 0, kCtrlWords * 4);
	if(r->retained.pending) {
		r->retained.pending = false;
		r->retained.valid = g == LUCID_OK;
		if(g == LUCID_OK && r->retained.record && !recorded) {
			// rendered without the record: size the backward buffers for this frame
			int e = grow_backward(r, block_rows, ranges, 0);
			if(e != LUCID_OK)
				return e;
		}
		r->retained.recorded = r->retained.valid && r->retained.record && recorded;
	}
	return g;
}

int lucid_read_info(lucid_renderer *r, uint32_t *dst, size_t num_u32) {
	if(!r || !dst)
		return LUCID_ERR_INVALID_ARG;
	size_t n = (size_t)LUCID_INFO_U32_SIZE + 10 * (size_t)r->bin_count;
	if(num_u32 < n)
		return LUCID_ERR_INVALID_ARG;
	LUCID_CHECK(hipSetDevice(r->device));
	LUCID_CHECK(hipStreamSynchronize(r->last_stream));
	LUCID_CHECK(hipMemcpy(dst, r->info.ptr, n * 4, hipMemcpyDeviceToHost));
	return LUCID_OK;
}

int lucid_verify_info(lucid_renderer *r) {
	if(!r)
		return LUCID_ERR_INVALID_ARG;
	std::vector<uint32_t> buf((size_t)LUCID_INFO_U32_SIZE + 10 * (size_t)r->bin_count);
	int e = lucid_read_info(r, buf.data(), buf.size());
	if(e != LUCID_OK)
		return e;
	const int NB = r->bin_count;
	const uint32_t *cnt = buf.data() + LUCID_INFO_U32_SIZE;
	int errors = 0;
	// src/lucid_renderer.cpp:650-680
	for(int k = 0; k < 2; k++) {
		const uint32_t *counts = cnt + (k ? LUCID_COUNTS_TRI_COUNTS : LUCID_COUNTS_QUAD_COUNTS) * NB;
		const uint32_t *offs = cnt + (k ? LUCID_COUNTS_TRI_OFFSETS : LUCID_COUNTS_QUAD_OFFSETS) * NB;
		const uint32_t *temp =
			cnt + (k ? LUCID_COUNTS_TRI_OFFSETS_TEMP : LUCID_COUNTS_QUAD_OFFSETS_TEMP) * NB;
		for(int i = 0; i < NB; i++) {
			if(i > 0 && offs[i] != offs[i - 1] + counts[i - 1])
				errors++;
			if(temp[i] != offs[i] + counts[i])
				errors++;
		}
	}
	return errors;
}

int lucid_read_bin_lists(lucid_renderer *r, uint32_t *quads, size_t quad_cap, uint32_t *tris,
						 size_t tri_cap, size_t *num_quads, size_t *num_tris) {
	if(!r)
		return LUCID_ERR_INVALID_ARG;
	LUCID_CHECK(hipSetDevice(r->device));
	LUCID_CHECK(hipStreamSynchronize(r->last_stream));
	uint32_t ctrl[kCtrlWords];
	LUCID_CHECK(hipMemcpy(ctrl, r->ctrl.ptr, sizeof(ctrl), hipMemcpyDeviceToHost));
	size_t nq = ctrl[kCtrlTotalBinQuads], nt = ctrl[kCtrlTotalBinTris];
	if(num_quads)
		*num_quads = nq;
	if(num_tris)
		*num_tris = nt;
	if(ctrl[kCtrlOverflow] & 1u)
		return LUCID_ERR_CAPACITY;
	if(quads && quad_cap >= nq && nq)
		LUCID_CHECK(hipMemcpy(quads, r->bin_quads.ptr, nq * 4, hipMemcpyDeviceToHost));
	if(tris && tri_cap >= nt && nt)
		LUCID_CHECK(hipMemcpy(tris, r->bin_tris.ptr, nt * 4, hipMemcpyDeviceToHost));
	if((quads && quad_cap < nq) || (tris && tri_cap < nt))
		return LUCID_ERR_INVALID_ARG;
	return LUCID_OK;
}

int lucid_read_setup(lucid_renderer *r, uint32_t *tri_records, size_t num_tris,
					 uint32_t *quad_aabbs, size_t num_quads) {
	if(!r)
		return LUCID_ERR_INVALID_ARG;
	size_t q = (size_t)r->last_num_quads;
	if((tri_records && num_tris < 2 * q) || (quad_aabbs && num_quads < q))
		return LUCID_ERR_INVALID_ARG;
	LUCID_CHECK(hipSetDevice(r->device));
	LUCID_CHECK(hipStreamSynchronize(r->last_stream));
	if(quad_aabbs && q)
		LUCID_CHECK(hipMemcpy(quad_aabbs, r->quad_aabb.ptr, q * 4, hipMemcpyDeviceToHost));
	if(tri_records && q) {
		// record arrays hold the first triangles of all quads, then the second ones
		const size_t half = (size_t)r->quad_cap;
		// barycentric planes are quads 1-2 of each shading record (kRecStride uint4)
		const size_t rec_words = (size_t)kRecStride * 4;
		std::vector<uint32_t> scan(2 * q * 8), depth(2 * q * 4), shade(2 * q * rec_words);
		for(size_t h = 0; h < 2; h++) {
			const size_t src = h * half, dst = h * q;
			LUCID_CHECK(hipMemcpy(&scan[dst * 8], (char *)r->tri_scan.ptr + src * 32, q * 32,
								  hipMemcpyDeviceToHost));
			LUCID_CHECK(hipMemcpy(&depth[dst * 4], (char *)r->tri_depth.ptr + src * 16, q * 16,
								  hipMemcpyDeviceToHost));
			LUCID_CHECK(hipMemcpy(&shade[dst * rec_words],
								  (char *)r->tri_shade.ptr + src * rec_words * 4,
								  q * rec_words * 4, hipMemcpyDeviceToHost));
		}
		for(size_t i = 0; i < 2 * q; i++) {
			const size_t k = (i >> 1) + (i & 1) * q;
			uint32_t *o = tri_records + i * 20;
			memcpy(o, &shade[k * rec_words + 4], 32);
			memcpy(o + 8, &scan[k * 8], 32);
			memcpy(o + 16, &depth[k * 4], 16);
		}
	}
	return LUCID_OK;
}

int lucid_stage_times(lucid_renderer *r, float *ms, int max_stages) {
	if(!r || !ms || max_stages < 0)
		return LUCID_ERR_INVALID_ARG;
	if(!r->timers)
		return 0;
	LUCID_CHECK(hipSetDevice(r->device));
	LUCID_CHECK(hipEventSynchronize(r->stage_ev[lucid_renderer::kStages]));
	int n = std::min(max_stages, (int)lucid_renderer::kStages);
	for(int i = 0; i < n; i++)
		LUCID_CHECK(hipEventElapsedTime(&ms[i], r->stage_ev[i], r->stage_ev[i + 1]));
	return n;
}

int lucid_bin_count(const lucid_renderer *r) { return r ? r->bin_count : 0; }
int lucid_bin_size(const lucid_renderer *r) { return r ? kBinSize : 0; }
int lucid_block_size(const lucid_renderer *r) { return r ? kBlockSize : 0; }
int64_t lucid_max_visible_quads(const lucid_renderer *r) { return r ? r->quad_cap : 0; }
// src/lucid_renderer.h:49 allows 5/2 of the visible-quad arena; here ids index input quads.
int64_t lucid_max_scene_quads(const lucid_renderer *r) { return r ? kMaxQuads - 1 : 0; }

} // extern "C"
