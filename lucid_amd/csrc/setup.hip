// Quad/triangle setup (≙ data/shaders/quad_setup.glsl).
//
// One thread per input quad. Differences from the reference, by design:
//   * quad ids are global input positions (instance order), so records are written in place
//     instead of being compacted through atomics (quad_setup.glsl:415-444): ids are
//     deterministic and the per-bin lists refer to input quads;
//   * large quads (bin AABB > 4 bins) are appended to a work list for the binning kernels
//     instead of being packed from the back of the visible-quad arena.
// The arithmetic of processInputQuad (:136-254) and storeTri (:274-340) is kept term by term.
#include "lucid_device.h"

namespace lucid {

__device__ __forceinline__ float4 mul_point(const float vp[16], V3 p) {
	float4 r;
	r.x = ((vp[0] * p.x + vp[4] * p.y) + vp[8] * p.z) + vp[12] * 1.0f;
	r.y = ((vp[1] * p.x + vp[5] * p.y) + vp[9] * p.z) + vp[13] * 1.0f;
	r.z = ((vp[2] * p.x + vp[6] * p.y) + vp[10] * p.z) + vp[14] * 1.0f;
	r.w = ((vp[3] * p.x + vp[7] * p.y) + vp[11] * p.z) + vp[15] * 1.0f;
	return r;
}

__device__ __forceinline__ uint32_t clip_mask(float4 p) {
	return (p.x < -p.w ? 0x001u : 0u) | (p.x > p.w ? 0x002u : 0u) | (p.y < -p.w ? 0x004u : 0u) |
		   (p.y > p.w ? 0x008u : 0u) | (p.z < -p.w ? 0x010u : 0u) | (p.z > p.w ? 0x020u : 0u);
}

// Blinn's screen coverage of a triangle crossing the near plane (quad_setup.glsl:77-128).
__device__ void clipped_aabb(float4 a, float4 b, float4 c, float iwa, float iwb, float iwc,
							 uint32_t clipmask, float box[4]) {
	float4 v[3] = {a, b, c};
	float iw[3] = {iwa, iwb, iwc};
	box[0] = 1.0f, box[1] = 1.0f, box[2] = -1.0f, box[3] = -1.0f;
	int any_vis = 0;
	uint32_t orm = clipmask | (clipmask >> 8) | (clipmask >> 16);
#pragma unroll
	for(int i = 0; i < 3; i++) {
		uint32_t cm = clipmask >> (i * 8);
		if((cm & 0x3u) == 0) {
			any_vis |= 0x1;
			if(v[i].x - box[0] * v[i].w < 0.0f)
				box[0] = v[i].x * iw[i];
			if(v[i].x - box[2] * v[i].w > 0.0f)
				box[2] = v[i].x * iw[i];
		}
		if((cm & 0xcu) == 0) {
			any_vis |= 0x10;
			if(v[i].y - box[1] * v[i].w < 0.0f)
				box[1] = v[i].y * iw[i];
			if(v[i].y - box[3] * v[i].w > 0.0f)
				box[3] = v[i].y * iw[i];
		}
	}
	if((any_vis & 0x0f) == 0) {
		box[0] = -1.0f, box[2] = 1.0f;
	} else if(orm & 0x3u) {
#pragma unroll
		for(int i = 0; i < 3; i++) {
			uint32_t cm = clipmask >> (i * 8);
			if((cm & 0x1u) && v[i].x - box[0] * v[i].w < 0.0f)
				box[0] = -1.0f;
			if((cm & 0x2u) && v[i].x - box[2] * v[i].w > 0.0f)
				box[2] = 1.0f;
		}
	}
	if((any_vis & 0xf0) == 0) {
		box[1] = -1.0f, box[3] = 1.0f;
	} else if(orm & 0xcu) {
#pragma unroll
		for(int i = 0; i < 3; i++) {
			uint32_t cm = clipmask >> (i * 8);
			if((cm & 0x4u) && v[i].y - box[1] * v[i].w < 0.0f)
				box[1] = -1.0f;
			if((cm & 0x8u) && v[i].y - box[3] * v[i].w > 0.0f)
				box[3] = 1.0f;
		}
	}
}

__device__ __forceinline__ V3 load_pos(const float *pos, uint32_t v) {
	const float *p = pos + (size_t)v * 3;
	return mk3(p[0], p[1], p[2]);
}

// storeTri (quad_setup.glsl:274-340). The depth record and scan parameters are stored for
// the binning and row kernels; the depth record, barycentric planes and packed normal are
// also returned for the triangle's shading record.
__device__ void store_tri(const FrameConst &c, const FrameBufs &b, uint32_t t, uint32_t flags_id,
						  V3 tri0, V3 tri1, V3 tri2, uint32_t y_aabb, uint4 &depth_out,
						  uint4 &bary0_out, uint4 &bary1_out, uint32_t &normal_out) {
	V3 dirx = ld3(c.dirx), diry = ld3(c.diry), dir0 = ld3(c.dir0), rd0 = ld3(c.ray_dir0);
	V3 normal = cross3(sub3(tri0, tri2), sub3(tri1, tri0));
	float mult = 1.0f / sqrtf(dot3(normal, normal));
	normal = mul3(normal, mult);
	normal_out = encode_normal(normal);

	V3 edge0 = mul3(sub3(tri0, tri2), mult);
	V3 edge1 = mul3(sub3(tri1, tri0), mult);
	float plane_dist = dot3(normal, tri0);
	V3 nrm_tri0 = cross3(tri0, normal);
	float param0 = dot3(edge0, nrm_tri0);
	float param1 = dot3(edge1, nrm_tri0);
	edge0 = cross3(normal, edge0);
	edge1 = cross3(normal, edge1);
	edge0 = mk3(dot3(edge0, dirx), dot3(edge0, diry), dot3(edge0, rd0));
	edge1 = mk3(dot3(edge1, dirx), dot3(edge1, diry), dot3(edge1, rd0));
	V3 pn = mul3(normal, 1.0f / plane_dist);
	V3 deq = mk3(dot3(pn, dirx), dot3(pn, diry), dot3(pn, rd0));

	depth_out = make_uint4(f2bits(deq.x), f2bits(deq.y), f2bits(deq.z), flags_id);
	b.tri_depth[tri_rec(b, t)] = depth_out;
	bary0_out = make_uint4(f2bits(edge0.x), f2bits(edge0.y), f2bits(edge0.z), f2bits(param0));
	bary1_out = make_uint4(f2bits(edge1.x), f2bits(edge1.y), f2bits(edge1.z), f2bits(param1));

	V3 n0 = cross3(tri2, sub3(tri1, tri2));
	V3 n1 = cross3(tri0, sub3(tri2, tri0));
	V3 n2 = cross3(tri1, sub3(tri0, tri1));
	if(dot3(tri0, n0) < 0.0f)
		n0 = neg3(n0), n1 = neg3(n1), n2 = neg3(n2);
	V3 e[3] = {mk3(dot3(n0, dirx), dot3(n0, diry), dot3(n0, dir0)),
			   mk3(dot3(n1, dirx), dot3(n1, diry), dot3(n1, dir0)),
			   mk3(dot3(n2, dirx), dot3(n2, diry), dot3(n2, dir0))};
	float base[3], step[3], scan[3];
#pragma unroll
	for(int i = 0; i < 3; i++) {
		// An edge parallel to the pixel rows has e.x == +-0, and its x-intercept form would
		// be inf - inf = NaN, which bounds nothing (rows past the edge would fill with the
		// other two edges' spans). Tilted to e.x = max(|e.y| 2^-40, 2^-100) > 0, the
		// intercept runs from about -2^40 to +2^40 pixels across the edge's row position
		// and still empties the rows beyond it.
		if(e[i].x == 0.0f)
			e[i].x = fmaxf(fabsf(e[i].y) * 0x1p-40f, 0x1p-100f);
		float inv_ex = 1.0f / e[i].x;
		base[i] = -(e[i].z * inv_ex);
		step[i] = -(e[i].y * inv_ex);
		scan[i] = (step[i] * 0.5f + base[i]) - (-0.5f);
	}
	uint32_t signs = (e[0].x < 0.0f ? 1u : 0u) | (e[1].x < 0.0f ? 2u : 0u) |
					 (e[2].x < 0.0f ? 4u : 0u) | (e[0].y < 0.0f ? 8u : 0u) |
					 (e[1].y < 0.0f ? 16u : 0u) | (e[2].y < 0.0f ? 32u : 0u);
	b.tri_scan[2 * (size_t)tri_rec(b, t) + 0] = make_uint4(f2bits(scan[0]), f2bits(scan[1]), f2bits(scan[2]), y_aabb);
	b.tri_scan[2 * (size_t)tri_rec(b, t) + 1] = make_uint4(f2bits(step[0]), f2bits(step[1]), f2bits(step[2]), signs);
}

// Returns 0 small, 1 large, 2 + rejection type when rejected.
__device__ int process_quad(const FrameConst &c, const FrameBufs &b, uint32_t q, uint32_t inst,
							uint32_t flags, const uint32_t vi[4]) {
	bool cull0 = vi[0] == vi[1] || vi[1] == vi[2] || vi[2] == vi[0];
	bool cull1 = vi[0] == vi[2] || vi[2] == vi[3] || vi[3] == vi[0];
	if(cull0 && cull1)
		return 2 + LUCID_REJECTION_TYPE_OTHER;
	V3 p[4];
#pragma unroll
	for(int i = 0; i < 4; i++)
		p[i] = load_pos(b.positions, vi[i]);
	auto eq = [](V3 a, V3 b2) { return a.x == b2.x && a.y == b2.y && a.z == b2.z; };
	cull0 = cull0 || eq(p[0], p[1]) || eq(p[1], p[2]) || eq(p[2], p[0]);
	cull1 = cull1 || eq(p[0], p[2]) || eq(p[2], p[3]) || eq(p[3], p[0]);
	if(c.backface) {
		V3 o = ld3(c.origin);
		V3 p0 = sub3(p[0], o), p1 = sub3(p[1], o), p2 = sub3(p[2], o), p3 = sub3(p[3], o);
		float vol0 = dot3(p0, cross3(p2, sub3(p1, p2)));
		float vol1 = dot3(p0, cross3(p3, sub3(p2, p3)));
		cull0 = cull0 || vol0 <= 0.0f;
		cull1 = cull1 || vol1 <= 0.0f;
		if(cull0 && cull1)
			return 2 + LUCID_REJECTION_TYPE_BACKFACE;
	}
	uint32_t cull_flags = (cull0 ? 1u : 0u) | (cull1 ? 2u : 0u);

	float4 v[4];
#pragma unroll
	for(int i = 0; i < 4; i++)
		v[i] = mul_point(c.vp, p[i]);
	uint32_t cm = clip_mask(v[0]) | (clip_mask(v[1]) << 8) | (clip_mask(v[2]) << 16) |
				  (clip_mask(v[3]) << 24);
	uint32_t andm = cm & (cm >> 8) & (cm >> 16) & (cm >> 24) & 0xffu;
	uint32_t orm = cm | (cm >> 8) | (cm >> 16) | (cm >> 24);
	if(andm != 0)
		return 2 + LUCID_REJECTION_TYPE_FRUSTUM;

	float iw[4] = {1.0f / v[0].w, 1.0f / v[1].w, 1.0f / v[2].w, 1.0f / v[3].w};
	float a0[4], a1[4];
	if(orm & 0x30u) {
		clipped_aabb(v[0], v[1], v[2], iw[0], iw[1], iw[2], cm, a0);
		clipped_aabb(v[0], v[2], v[3], iw[0], iw[2], iw[3], (cm & 0xffu) | ((cm & 0xffff0000u) >> 8), a1);
	} else {
		float nx[4], ny[4];
#pragma unroll
		for(int i = 0; i < 4; i++)
			nx[i] = v[i].x * iw[i], ny[i] = v[i].y * iw[i];
		a0[0] = fmn(fmn(nx[0], nx[1]), nx[2]);
		a0[1] = fmn(fmn(ny[0], ny[1]), ny[2]);
		a0[2] = fmx(fmx(nx[0], nx[1]), nx[2]);
		a0[3] = fmx(fmx(ny[0], ny[1]), ny[2]);
		a1[0] = fmn(fmn(nx[0], nx[2]), nx[3]);
		a1[1] = fmn(fmn(ny[0], ny[2]), ny[3]);
		a1[2] = fmx(fmx(nx[0], nx[2]), nx[3]);
		a1[3] = fmx(fmx(ny[0], ny[2]), ny[3]);
	}
	float hx = (float)c.width * 0.5f, hy = (float)c.height * 0.5f;
	float mul[4] = {hx, hy, hx, hy};
	float box[4];
#pragma unroll
	for(int i = 0; i < 4; i++) {
		a0[i] = (a0[i] + 1.0f) * mul[i];
		a1[i] = (a1[i] + 1.0f) * mul[i];
	}
	box[0] = fmn(a0[0], a1[0]), box[1] = fmn(a0[1], a1[1]);
	box[2] = fmx(a0[2], a1[2]), box[3] = fmx(a0[3], a1[3]);
	if(ceilf(box[0] - 0.5001f) == floorf(box[2] + 0.5001f) ||
	   ceilf(box[1] - 0.5001f) == floorf(box[3] + 0.5001f))
		return 2 + LUCID_REJECTION_TYPE_BETWEEN_SAMPLES;

	float maxp[4] = {(float)(c.width - 1), (float)(c.height - 1), (float)(c.width - 1),
					 (float)(c.height - 1)};
	float inset[4] = {0.49f, 0.49f, -0.49f, -0.49f};
#pragma unroll
	for(int i = 0; i < 4; i++) {
		a0[i] = fmn(fmx(a0[i] + inset[i], 0.0f), maxp[i]);
		a1[i] = fmn(fmx(a1[i] + inset[i], 0.0f), maxp[i]);
		box[i] = fmn(fmx(box[i] + inset[i], 0.0f), maxp[i]);
	}
	uint32_t bb[4];
#pragma unroll
	for(int i = 0; i < 4; i++)
		bb[i] = f2u(box[i]) >> kBinShift;
	uint32_t enc = encode_aabb28(bb[0], bb[1], bb[2], bb[3]) | (cull_flags << 30);
	uint32_t area = (bb[2] - bb[0] + 1u) * (bb[3] - bb[1] + 1u);
	uint32_t y_aabb[2] = {f2u(a0[1]) | (f2u(a0[3]) << 16), f2u(a1[1]) | (f2u(a1[3]) << 16)};
	b.quad_aabb[q] = enc;

	uint32_t flags_id = flags | (inst << 16);
	V3 o = ld3(c.origin);
	const uint32_t vq[4] = {vi[0], vi[1], vi[2], vi[3]};
#pragma unroll
	for(int s = 0; s < 2; s++) {
		if((cull_flags >> s) & 1u)
			continue; // no record is read for a culled triangle
		const uint32_t t = 2 * q + s;
		uint4 d, b0, b1;
		uint32_t nrm;
		store_tri(c, b, t, flags_id, sub3(p[0], o), sub3(p[1 + s], o), sub3(p[2 + s], o),
				  y_aabb[s], d, b0, b1, nrm);
		// shading record: depth and barycentric planes, colours (storeQuad,
		// quad_setup.glsl:256-272) and lighting, one 128-byte line
		uint4 rec[kRecQuads];
		make_shading_record(c, b, t, d, b0, b1, nrm, vq, rec);
		uint4 *dst = b.tri_shade + (size_t)kRecStride * tri_rec(b, t);
#pragma unroll
		for(int j = 0; j < kRecQuads; j++)
			dst[j] = rec[j];
		if(c.has_textures && (flags & LUCID_INST_HAS_ALBEDO_TEXTURE)) {
			// texture coordinates of the triangle's corners, relative to corner 0
			// (storeQuad, quad_setup.glsl:265-271)
			const float *tc = b.tex_coords;
			const uint32_t w0 = vq[0], w1 = vq[1 + s], w2 = vq[2 + s];
			const float t0x = tc[2 * (size_t)w0], t0y = tc[2 * (size_t)w0 + 1];
			const float d1x = tc[2 * (size_t)w1] - t0x, d1y = tc[2 * (size_t)w1 + 1] - t0y;
			const float d2x = tc[2 * (size_t)w2] - t0x, d2y = tc[2 * (size_t)w2 + 1] - t0y;
			uint4 *tt = b.tri_tex + 2 * (size_t)tri_rec(b, t);
			tt[0] = make_uint4(f2bits(t0x), f2bits(t0y), f2bits(d1x), f2bits(d1y));
			tt[1] = make_uint4(f2bits(d2x), f2bits(d2y), 0u, 0u);
		}
	}
	return area <= 4u ? 0 : 1;
}

// Instance of quad q: the last instance whose first quad (inst_prefix) is at most q
// (instances without quads share their successor's prefix and are skipped).
__device__ __forceinline__ int find_instance(const FrameConst &c, const FrameBufs &b, uint32_t q) {
	int lo = 0, hi = c.num_instances - 1;
	while(lo < hi) {
		int mid = (lo + hi + 1) >> 1;
		if((uint32_t)b.inst_prefix[mid] <= q)
			lo = mid;
		else
			hi = mid - 1;
	}
	return lo;
}

__device__ __forceinline__ void setup_kernel_body(const FrameConst &c, const FrameBufs &b, WgPos wg) {
	__shared__ uint32_t s_counts[6]; // small, large, rejected x4
	__shared__ int s_inst[2];         // instances of the work-group's first and last quads
	__shared__ uint32_t s_prefix[kSetupThreads]; // their prefixes, when they span few instances
	if(threadIdx.x < 6)
		s_counts[threadIdx.x] = 0;
	const uint32_t q0 = wg.id * (uint32_t)kSetupThreads;
	const uint32_t nq = (uint32_t)c.num_quads;
	// the two searches over the whole instance table run once per work-group (one lane of
	// each of two waves); each quad then searches the few instances of its work-group in LDS
	if(q0 < nq && (threadIdx.x == 0 || threadIdx.x == 64))
		s_inst[threadIdx.x >> 6] = find_instance(c, b, threadIdx.x == 0 ? q0 : min(q0 + kSetupThreads - 1, nq - 1));
	__syncthreads();
	const int i0 = q0 < nq ? s_inst[0] : 0, span = q0 < nq ? s_inst[1] - i0 + 1 : 0;
	const bool local = span <= kSetupThreads;
	if(local && (int)threadIdx.x < span)
		s_prefix[threadIdx.x] = (uint32_t)b.inst_prefix[i0 + threadIdx.x];
	__syncthreads();

	uint32_t q = q0 + threadIdx.x;
	int kind = -1;
	if(q < nq) {
		int lo;
		if(local) {
			int l = 0, h = span - 1;
			while(l < h) {
				const int mid = (l + h + 1) >> 1;
				if(s_prefix[mid] <= q)
					l = mid;
				else
					h = mid - 1;
			}
			lo = i0 + l;
		} else {
			lo = find_instance(c, b, q);
		}
		lucid_instance_data ins = b.instances[lo];
		uint32_t k = q - (local ? s_prefix[lo - i0] : (uint32_t)b.inst_prefix[lo]);
		const uint32_t *ix = b.quad_indices + (size_t)ins.index_offset + (size_t)k * 4;
		uint32_t vi[4];
#pragma unroll
		for(int j = 0; j < 4; j++)
			vi[j] = ix[j] + (uint32_t)ins.vertex_offset;
		kind = process_quad(c, b, q, (uint32_t)lo, ins.flags, vi);
		if(kind >= 2)
			b.quad_aabb[q] = kInvalidAabb;
	}
	// wave-aggregated counters and this work-group's large-quad sub-list
	uint64_t large = ballot(kind == 1);
	const uint32_t stripe = wg.id % kLargeStripes;
	uint32_t base = 0;
	if(lane_id() == 0 && large)
		base = atomicAdd(large_cursor(b.ctrl, stripe), (uint32_t)__popcll(large));
	base = __shfl(base, 0, 64);
	if(kind == 1) {
		uint32_t rank = (uint32_t)__popcll(large & ((1ull << lane_id()) - 1ull));
		b.large_list[(size_t)stripe * large_stripe_cap((uint32_t)c.num_quads) + base + rank] = q;
	}
#pragma unroll
	for(int k = 0; k < 6; k++) {
		uint64_t m = ballot(kind == k);
		if(lane_id() == 0 && m)
			atomicAdd(&s_counts[k], (uint32_t)__popcll(m));
	}
	__syncthreads();
	if(threadIdx.x < 6 && s_counts[threadIdx.x]) {
		lucid_info *info = (lucid_info *)b.info;
		uint32_t *dst = threadIdx.x < 2 ? (uint32_t *)&info->num_visible_quads[threadIdx.x]
										 : &info->num_rejected_quads[threadIdx.x - 2];
		atomicAdd(dst, s_counts[threadIdx.x]);
	}
}

__global__ __launch_bounds__(kSetupThreads) void setup_kernel(FrameConst c, FrameBufs b) {
	setup_kernel_body(c, b, WgPos{blockIdx.x, gridDim.x});
}
__global__ __launch_bounds__(kSetupThreads) void setup_kernel_batch(const FrameDesc *__restrict__ frames,
		const uint32_t *__restrict__ pre, uint32_t num_frames) {
	const StageBatch sb{frames, pre, num_frames};
	const uint32_t f = batch_frame(sb, blockIdx.x);
	const FrameDesc &d = sb.frames[f];
	setup_kernel_body(d.c, d.b, WgPos{blockIdx.x - sb.pre[f], sb.pre[f + 1] - sb.pre[f]});
}

// Per-frame reset of LucidInfo + per-bin counters and the control words
// (src/lucid_renderer.cpp:437 clears the same region with a fill).
__device__ __forceinline__ void clear_body(uint32_t *info, uint32_t n_info, uint32_t *ctrl,
										   uint32_t num_quads, uint32_t keep_overflow, uint32_t wg) {
	uint32_t i = wg * 256u + threadIdx.x;
	if(i < n_info)
		info[i] = i == 0 ? num_quads : 0u;
	// control words and statistics stripes; the overflow bits of an unfinished earlier
	// frame move to the carry word (which no kernel reads), so that they reach the host
	// with this frame's status instead of being lost when its status copy replaces theirs
	if(i == (uint32_t)kCtrlOverflow) {
		ctrl[kCtrlOverflowCarry] = keep_overflow ? ctrl[kCtrlOverflowCarry] | ctrl[kCtrlOverflow] : 0u;
		ctrl[kCtrlOverflow] = 0u;
	} else if(i < (uint32_t)kCtrlAllocWords && i != (uint32_t)kCtrlOverflowCarry) {
		ctrl[i] = 0u;
	}
}

__global__ __launch_bounds__(256) void clear_kernel(uint32_t *info, uint32_t n_info, uint32_t *ctrl,
													uint32_t num_quads, uint32_t keep_overflow) {
	clear_body(info, n_info, ctrl, num_quads, keep_overflow, blockIdx.x);
}

__global__ __launch_bounds__(256) void clear_kernel_batch(const FrameDesc *__restrict__ frames,
		const uint32_t *__restrict__ pre, uint32_t num_frames) {
	const StageBatch sb{frames, pre, num_frames};
	const uint32_t f = batch_frame(sb, blockIdx.x);
	const FrameDesc &d = sb.frames[f];
	clear_body(d.b.info, LUCID_INFO_U32_SIZE + 10u * (uint32_t)d.c.bin_count, d.b.ctrl,
			   (uint32_t)d.c.num_quads, d.c.keep_overflow, blockIdx.x - sb.pre[f]);
}

hipError_t launch_setup_batch(const BatchPlan &p, hipStream_t stream) {
	hipLaunchKernelGGL(clear_kernel_batch, dim3(p.total[kBsClear]), dim3(256), 0, stream,
					   p.frames, p.stage_pre(kBsClear), p.num_frames);
	if(p.total[kBsSetup])
		hipLaunchKernelGGL(setup_kernel_batch, dim3(p.total[kBsSetup]), dim3(kSetupThreads), 0,
						   stream, p.frames, p.stage_pre(kBsSetup), p.num_frames);
	return hipGetLastError();
}

hipError_t launch_setup(const FrameConst &c, const FrameBufs &b, uint32_t n_info,
						hipStream_t stream) {
	uint32_t n = n_info > (uint32_t)kCtrlAllocWords ? n_info : (uint32_t)kCtrlAllocWords;
	hipLaunchKernelGGL(clear_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, b.info, n_info,
					   b.ctrl, (uint32_t)c.num_quads, c.keep_overflow);
	if(c.num_quads > 0)
		hipLaunchKernelGGL(setup_kernel, dim3((c.num_quads + kSetupThreads - 1) / kSetupThreads),
						   dim3(kSetupThreads), 0, stream, c, b);
	return hipGetLastError();
}

} // namespace lucid
