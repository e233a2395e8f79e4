// Binning: per-bin counts, offsets/categories and per-bin quad/tri lists
// (≙ data/shaders/bin_counter.glsl, bin_categorizer.glsl, bin_dispatcher.glsl).
//
// MI355X design: a fixed grid of work-groups each owns a static slice of the quad range and
// of the large-quad list (no work stealing, so no batch linked lists, bin_counter.glsl:141-242).
// Each work-group builds an LDS histogram, the scan kernels turn the per-(work-group, bin)
// histograms into exclusive offsets, and the dispatcher replays its slice into its reserved
// ranges. Counts and offsets are deterministic; the order inside one work-group's range of a
// bin comes from LDS atomics and is not (lists are compared as multisets).
#include "lucid_device.h"

namespace lucid {

constexpr int kBinThreads = 512;
constexpr int kBinPrefetch = 8; // quad AABBs a thread loads before it counts or places them

// loadScanlineParamsBin + scanlineStep (scanline.glsl:28-52, bin_counter.glsl:53-62)
struct BinScan {
	float mn[3], mx[3], step[3];
	int min_by, max_by;
};

__device__ __forceinline__ void bin_scan_init(BinScan &s, uint4 s0, uint4 s1) {
	s.min_by = (int)(s0.w & 0xffffu) >> kBinShift;
	s.max_by = (int)(s0.w >> 16) >> kBinShift;
	float offset = (float)kBinSize - 0.989f;
	float sy = (float)(s.min_by * kBinSize) - 0.01f;
	float sx = 0.99f;
	float sc[3] = {bits2f(s0.x), bits2f(s0.y), bits2f(s0.z)};
	float st[3] = {bits2f(s1.x), bits2f(s1.y), bits2f(s1.z)};
#pragma unroll
	for(int i = 0; i < 3; i++) {
		bool xneg = (s1.w >> i) & 1u, yneg = (s1.w >> (3 + i)) & 1u;
		float yoff = yneg ? 0.0f : offset, xoff = xneg ? 0.0f : offset;
		float v = sc[i] + (st[i] * (yoff + sy) - (xoff + sx));
		s.mn[i] = xneg ? -__builtin_huge_valf() : v;
		s.mx[i] = xneg ? v : __builtin_huge_valf();
		s.step[i] = st[i] * (float)kBinSize;
	}
}

__device__ __forceinline__ void bin_scan_step(BinScan &s, int &bmin, int &bmax) {
	float xmin = fmx(fmx(s.mn[0], s.mn[1]), s.mn[2]);
	float xmax = fmn(fmn(s.mx[0], s.mx[1]), s.mx[2]);
#pragma unroll
	for(int i = 0; i < 3; i++)
		s.mn[i] = s.mn[i] + s.step[i], s.mx[i] = s.mx[i] + s.step[i];
	bmin = f2i(xmin + 1.0f) >> kBinShift;
	bmax = f2i(xmax) >> kBinShift;
}

__device__ __forceinline__ bool small_quad(uint32_t a, int &bsx, int &bsy, int &bex, int &bey) {
	bsx = (int)(a & 0x7fu), bsy = (int)((a >> 7) & 0x7fu);
	bex = (int)((a >> 14) & 0x7fu), bey = (int)((a >> 21) & 0x7fu);
	uint32_t area = ((uint32_t)(bex - bsx) + 1u) * ((uint32_t)(bey - bsy) + 1u);
	return area <= 4u;
}

struct Slice {
	uint32_t q0, q1, l0, l1;
	uint32_t lstart[kLargeStripes + 1]; // flat index of each large sub-list's first entry
	uint32_t lcap;
};

__device__ __forceinline__ Slice wg_slice(const FrameConst &c, const FrameBufs &b, uint32_t wg) {
	Slice s;
	uint32_t nq = (uint32_t)c.num_quads, nw = (uint32_t)c.num_bin_wgs;
	uint32_t cq = (nq + nw - 1) / nw;
	s.q0 = min(nq, wg * cq), s.q1 = min(nq, s.q0 + cq);
	uint32_t nl = 0;
#pragma unroll
	for(int k = 0; k < kLargeStripes; k++) {
		s.lstart[k] = nl;
		nl += *large_cursor(b.ctrl, k);
	}
	s.lstart[kLargeStripes] = nl;
	s.lcap = large_stripe_cap(nq);
	uint32_t cl = (nl + nw - 1) / nw;
	s.l0 = min(nl, wg * cl), s.l1 = min(nl, s.l0 + cl);
	return s;
}

// Quad of flat large-list entry i (< the total of the sub-lists).
__device__ __forceinline__ uint32_t large_at(const FrameBufs &b, const Slice &s, uint32_t i) {
	// the starts do not decrease, so the stripes that start at or before i are a prefix:
	// k counts them and base ends at the last one's start (static indices only, so the
	// starts stay in registers)
	int k = 0;
	uint32_t base = 0;
#pragma unroll
	for(int j = 1; j < kLargeStripes; j++) {
		const bool ge = i >= s.lstart[j];
		k += ge ? 1 : 0;
		base = ge ? s.lstart[j] : base;
	}
	return b.large_list[(size_t)k * s.lcap + (i - base)];
}

// bin_counter.glsl: per-work-group histograms of small quads and large tris.
__device__ __forceinline__ void bin_count_kernel_body(const FrameConst &c, const FrameBufs &b, WgPos wg) {
	extern __shared__ uint32_t s_hist[]; // [2][bin_count]
	const int NB = c.bin_count, NBX = c.bin_count_x;
	uint32_t *hq = s_hist, *ht = s_hist + NB;
	for(int i = threadIdx.x; i < 2 * NB; i += kBinThreads)
		s_hist[i] = 0;
	__syncthreads();
	Slice sl = wg_slice(c, b, wg.id);

	// small quads: +1 over the bin rectangle (bin_counter.glsl:64-80); a thread loads
	// kBinPrefetch AABBs at a time, so that their loads are in flight together
	for(uint32_t q0 = sl.q0; q0 < sl.q1; q0 += kBinThreads * kBinPrefetch) {
		uint32_t av[kBinPrefetch];
#pragma unroll
		for(int k = 0; k < kBinPrefetch; k++) {
			const uint32_t q = q0 + k * kBinThreads + threadIdx.x;
			av[k] = q < sl.q1 ? b.quad_aabb[q] : kInvalidAabb;
		}
#pragma unroll
		for(int k = 0; k < kBinPrefetch; k++) {
			int bsx, bsy, bex, bey;
			if(av[k] == kInvalidAabb || !small_quad(av[k], bsx, bsy, bex, bey))
				continue;
			for(int by = bsy; by <= bey; by++)
				for(int bx = bsx; bx <= bex; bx++)
					atomicAdd(&hq[by * NBX + bx], 1u);
		}
	}
	// large tris: per bin row +1/-1 marks (bin_counter.glsl:112-134)
	for(uint32_t i = sl.l0 * 2 + threadIdx.x; i < sl.l1 * 2; i += kBinThreads) {
		uint32_t q = large_at(b, sl, i >> 1), s = i & 1u;
		uint32_t a = b.quad_aabb[q];
		if((a >> (30 + s)) & 1u)
			continue;
		uint32_t t = 2 * q + s;
		int bsx = (int)(a & 0x7fu), bex = (int)((a >> 14) & 0x7fu);
		BinScan bs;
		bin_scan_init(bs, b.tri_scan[2 * (size_t)tri_rec(b, t)], b.tri_scan[2 * (size_t)tri_rec(b, t) + 1]);
		for(int by = bs.min_by; by <= bs.max_by; by++) {
			int bmin, bmax;
			bin_scan_step(bs, bmin, bmax);
			bmin = max(bmin, bsx), bmax = min(bmax, bex);
			if(bmax >= bmin) {
				atomicAdd(&ht[by * NBX + bmin], 1u);
				if(bmax + 1 < NBX)
					atomicAdd(&ht[by * NBX + bmax + 1], 0xffffffffu);
			}
		}
	}
	__syncthreads();
	// prefix over each bin row turns the marks into counts
	int wave = threadIdx.x >> 6, nwaves = kBinThreads / 64;
	for(int by = wave; by < c.bin_count_y; by += nwaves) {
		uint32_t carry = 0;
		for(int bx0 = 0; bx0 < NBX; bx0 += 64) {
			int bx = bx0 + lane_id();
			uint32_t v = bx < NBX ? ht[by * NBX + bx] : 0u;
			uint32_t inc = wave_inclusive_add(v) + carry;
			if(bx < NBX)
				ht[by * NBX + bx] = inc;
			carry = __shfl(inc, 63, 64);
		}
	}
	__syncthreads();
	uint32_t *dst = b.wg_hist + (size_t)wg.id * 2 * NB;
	for(int i = threadIdx.x; i < 2 * NB; i += kBinThreads)
		dst[i] = s_hist[i];
}

__global__ __launch_bounds__(kBinThreads) void bin_count_kernel(FrameConst c, FrameBufs b) {
	bin_count_kernel_body(c, b, WgPos{blockIdx.x, gridDim.x});
}
__global__ __launch_bounds__(kBinThreads) void bin_count_kernel_batch(const FrameDesc *__restrict__ frames,
		const uint32_t *__restrict__ pre, uint32_t num_frames) {
	const StageBatch sb{frames, pre, num_frames};
	const uint32_t f = batch_frame(sb, blockIdx.x);
	const FrameDesc &d = sb.frames[f];
	bin_count_kernel_body(d.c, d.b, WgPos{blockIdx.x - sb.pre[f], sb.pre[f + 1] - sb.pre[f]});
}

// Per bin: totals over work-groups and each work-group's exclusive offset inside the bin.
// Exclusive scan of the per-work-group bin histograms down each (kind, bin) column, and the
// column totals (the bin counts). A work-group takes 16 columns; its 16 row-parts each sum
// a run of rows with independent loads, an LDS scan joins the parts, then each part writes
// its exclusive prefixes.
constexpr int kScanCols = 16, kScanParts = 16;
__device__ __forceinline__ void bin_wg_scan_kernel_body(const FrameConst &c, const FrameBufs &b, WgPos wg) {
	const int NB = c.bin_count, nw = c.num_bin_wgs;
	const int col_l = threadIdx.x % kScanCols, part = threadIdx.x / kScanCols;
	const int col = wg.id * kScanCols + col_l; // 0 .. 2 NB: quads then tris
	const int rows = (nw + kScanParts - 1) / kScanParts; // <= 16: num_bin_wgs <= 256
	const int w0 = part * rows, w1 = min(nw, w0 + rows);
	__shared__ uint32_t s_part[kScanParts][kScanCols];
	uint32_t sum = 0;
	if(col < 2 * NB)
		for(int w = w0; w < w1; w++)
			sum += b.wg_hist[(size_t)w * 2 * NB + col];
	s_part[part][col_l] = sum;
	__syncthreads();
	if(part == 0) {
		uint32_t run = 0;
		for(int p = 0; p < kScanParts; p++) {
			uint32_t v = s_part[p][col_l];
			s_part[p][col_l] = run;
			run += v;
		}
		if(col < 2 * NB) {
			const int k = col / NB, bin = col - k * NB;
			uint32_t *counts = b.info + LUCID_INFO_U32_SIZE;
			counts[(k == 0 ? LUCID_COUNTS_QUAD_COUNTS : LUCID_COUNTS_TRI_COUNTS) * NB + bin] = run;
		}
	}
	__syncthreads();
	if(col >= 2 * NB)
		return;
	// all loads first (independent), then the stores: a part spans at most kMaxRows rows
	constexpr int kMaxRows = 16;
	uint32_t v[kMaxRows];
#pragma unroll
	for(int j = 0; j < kMaxRows; j++)
		v[j] = w0 + j < w1 ? b.wg_hist[(size_t)(w0 + j) * 2 * NB + col] : 0u;
	uint32_t run = s_part[part][col_l];
#pragma unroll
	for(int j = 0; j < kMaxRows; j++) {
		if(w0 + j < w1)
			b.wg_hist[(size_t)(w0 + j) * 2 * NB + col] = run;
		run += v[j];
	}
}

__global__ __launch_bounds__(kScanCols * kScanParts) void bin_wg_scan_kernel(FrameConst c, FrameBufs b) {
	bin_wg_scan_kernel_body(c, b, WgPos{blockIdx.x, gridDim.x});
}
__global__ __launch_bounds__(kScanCols * kScanParts) void bin_wg_scan_kernel_batch(const FrameDesc *__restrict__ frames,
		const uint32_t *__restrict__ pre, uint32_t num_frames) {
	const StageBatch sb{frames, pre, num_frames};
	const uint32_t f = batch_frame(sb, blockIdx.x);
	const FrameDesc &d = sb.frames[f];
	bin_wg_scan_kernel_body(d.c, d.b, WgPos{blockIdx.x - sb.pre[f], sb.pre[f + 1] - sb.pre[f]});
}

// bin_categorizer.glsl: exclusive offsets over bins, bin levels and level lists.
__device__ __forceinline__ void bin_categorize_kernel_body(const FrameConst &c, const FrameBufs &b, WgPos wg) {
	const int NB = c.bin_count;
	uint32_t *counts = b.info + LUCID_INFO_U32_SIZE;
	lucid_info *info = (lucid_info *)b.info;
	__shared__ uint32_t s_wave[16][4];
	__shared__ uint32_t s_carry[4];
	if(threadIdx.x < 4)
		s_carry[threadIdx.x] = 0;
	__syncthreads();
	int wave = threadIdx.x >> 6;
	for(int base = 0; base < NB; base += 1024) {
		int bin = base + threadIdx.x;
		uint32_t qc = bin < NB ? counts[LUCID_COUNTS_QUAD_COUNTS * NB + bin] : 0u;
		uint32_t tc = bin < NB ? counts[LUCID_COUNTS_TRI_COUNTS * NB + bin] : 0u;
		uint32_t ntris = tc + 2 * qc;
		uint32_t vals[4] = {qc, tc, (bin < NB && ntris > 0 && ntris < (uint32_t)kLowLevelLimit) ? 1u : 0u,
							(bin < NB && ntris >= (uint32_t)kLowLevelLimit) ? 1u : 0u};
		uint32_t inc[4];
#pragma unroll
		for(int k = 0; k < 4; k++) {
			inc[k] = wave_inclusive_add(vals[k]);
			if(lane_id() == 63)
				s_wave[wave][k] = inc[k];
		}
		__syncthreads();
		if(threadIdx.x < 4) {
			uint32_t run = s_carry[threadIdx.x];
			for(int w = 0; w < 16; w++) {
				uint32_t v = s_wave[w][threadIdx.x];
				s_wave[w][threadIdx.x] = run;
				run += v;
			}
			s_carry[threadIdx.x] = run;
		}
		__syncthreads();
		if(bin < NB) {
			uint32_t ex[4];
#pragma unroll
			for(int k = 0; k < 4; k++)
				ex[k] = s_wave[wave][k] + inc[k] - vals[k];
			counts[LUCID_COUNTS_QUAD_OFFSETS * NB + bin] = ex[0];
			counts[LUCID_COUNTS_QUAD_OFFSETS_TEMP * NB + bin] = ex[0] + qc;
			counts[LUCID_COUNTS_TRI_OFFSETS * NB + bin] = ex[1];
			counts[LUCID_COUNTS_TRI_OFFSETS_TEMP * NB + bin] = ex[1] + tc;
			if(vals[2])
				counts[LUCID_COUNTS_LOW_BINS * NB + ex[2]] = (uint32_t)bin;
			if(vals[3])
				counts[LUCID_COUNTS_HIGH_BINS * NB + ex[3]] = (uint32_t)bin;
		}
		__syncthreads();
	}
	if(threadIdx.x == 0) {
		uint32_t nq = s_carry[0], nt = s_carry[1], nlow = s_carry[2], nhigh = s_carry[3];
		b.ctrl[kCtrlTotalBinQuads] = nq;
		b.ctrl[kCtrlTotalBinTris] = nt;
		if(nq > b.bin_quads_cap || nt > b.bin_tris_cap)
			atomicOr(&b.ctrl[kCtrlOverflow], 1u);
		info->bin_level_counts[LUCID_BIN_LEVEL_EMPTY] = (int32_t)((uint32_t)NB - nlow - nhigh);
		info->bin_level_counts[LUCID_BIN_LEVEL_LOW] = (int32_t)nlow;
		info->bin_level_counts[LUCID_BIN_LEVEL_HIGH] = (int32_t)nhigh;
		info->bin_level_dispatches[LUCID_BIN_LEVEL_LOW][0] = nlow;
		info->bin_level_dispatches[LUCID_BIN_LEVEL_HIGH][0] = nhigh;
		for(int k = 0; k < LUCID_BIN_LEVELS_COUNT; k++)
			info->bin_level_dispatches[k][1] = info->bin_level_dispatches[k][2] = 1;
		info->num_binning_dispatches[0] = (uint32_t)c.num_bin_wgs;
		info->num_binning_dispatches[1] = info->num_binning_dispatches[2] = 1;
	}
}

__global__ __launch_bounds__(1024) void bin_categorize_kernel(FrameConst c, FrameBufs b) {
	bin_categorize_kernel_body(c, b, WgPos{blockIdx.x, gridDim.x});
}
__global__ __launch_bounds__(1024) void bin_categorize_kernel_batch(const FrameDesc *__restrict__ frames,
		const uint32_t *__restrict__ pre, uint32_t num_frames) {
	const StageBatch sb{frames, pre, num_frames};
	const uint32_t f = batch_frame(sb, blockIdx.x);
	const FrameDesc &d = sb.frames[f];
	bin_categorize_kernel_body(d.c, d.b, WgPos{blockIdx.x - sb.pre[f], sb.pre[f + 1] - sb.pre[f]});
}

// bin_dispatcher.glsl: replay the work-group's slice into its reserved ranges.
__device__ __forceinline__ void bin_dispatch_kernel_body(const FrameConst &c, const FrameBufs &b, WgPos wg) {
	if(b.ctrl[kCtrlOverflow] & 1u)
		return;
	extern __shared__ uint32_t s_pos[]; // [2][bin_count]
	const int NB = c.bin_count, NBX = c.bin_count_x;
	const uint32_t *counts = b.info + LUCID_INFO_U32_SIZE;
	const uint32_t *wgo = b.wg_hist + (size_t)wg.id * 2 * NB;
	for(int i = threadIdx.x; i < NB; i += kBinThreads) {
		s_pos[i] = counts[LUCID_COUNTS_QUAD_OFFSETS * NB + i] + wgo[i];
		s_pos[NB + i] = counts[LUCID_COUNTS_TRI_OFFSETS * NB + i] + wgo[NB + i];
	}
	__syncthreads();
	uint32_t *pq = s_pos, *pt = s_pos + NB;
	Slice sl = wg_slice(c, b, wg.id);
	for(uint32_t q0 = sl.q0; q0 < sl.q1; q0 += kBinThreads * kBinPrefetch) {
		uint32_t av[kBinPrefetch];
#pragma unroll
		for(int k = 0; k < kBinPrefetch; k++) {
			const uint32_t q = q0 + k * kBinThreads + threadIdx.x;
			av[k] = q < sl.q1 ? b.quad_aabb[q] : kInvalidAabb;
		}
#pragma unroll
		for(int k = 0; k < kBinPrefetch; k++) {
			const uint32_t q = q0 + k * kBinThreads + threadIdx.x;
			int bsx, bsy, bex, bey;
			if(av[k] == kInvalidAabb || !small_quad(av[k], bsx, bsy, bex, bey))
				continue;
			const uint32_t entry = q | (av[k] & 0xf0000000u);
			for(int by = bsy; by <= bey; by++)
				for(int bx = bsx; bx <= bex; bx++)
					b.bin_quads[atomicAdd(&pq[by * NBX + bx], 1u)] = entry;
		}
	}
	for(uint32_t i = sl.l0 * 2 + threadIdx.x; i < sl.l1 * 2; i += kBinThreads) {
		uint32_t q = large_at(b, sl, i >> 1), s = i & 1u;
		uint32_t a = b.quad_aabb[q];
		if((a >> (30 + s)) & 1u)
			continue;
		uint32_t t = 2 * q + s;
		int bsx = (int)(a & 0x7fu), bex = (int)((a >> 14) & 0x7fu);
		BinScan bs;
		bin_scan_init(bs, b.tri_scan[2 * (size_t)tri_rec(b, t)], b.tri_scan[2 * (size_t)tri_rec(b, t) + 1]);
		for(int by = bs.min_by; by <= bs.max_by; by++) {
			int bmin, bmax;
			bin_scan_step(bs, bmin, bmax);
			bmin = max(bmin, bsx), bmax = min(bmax, bex);
			for(int bx = bmin; bx <= bmax; bx++)
				b.bin_tris[atomicAdd(&pt[by * NBX + bx], 1u)] = t;
		}
	}
}

__global__ __launch_bounds__(kBinThreads) void bin_dispatch_kernel(FrameConst c, FrameBufs b) {
	bin_dispatch_kernel_body(c, b, WgPos{blockIdx.x, gridDim.x});
}
__global__ __launch_bounds__(kBinThreads) void bin_dispatch_kernel_batch(const FrameDesc *__restrict__ frames,
		const uint32_t *__restrict__ pre, uint32_t num_frames) {
	const StageBatch sb{frames, pre, num_frames};
	const uint32_t f = batch_frame(sb, blockIdx.x);
	const FrameDesc &d = sb.frames[f];
	bin_dispatch_kernel_body(d.c, d.b, WgPos{blockIdx.x - sb.pre[f], sb.pre[f + 1] - sb.pre[f]});
}

hipError_t launch_binning_batch(const BatchPlan &p, hipStream_t stream) {
	hipLaunchKernelGGL(bin_count_kernel_batch, dim3(p.total[kBsBinCount]), dim3(kBinThreads),
					   p.bin_lds, stream, p.frames, p.stage_pre(kBsBinCount), p.num_frames);
	hipLaunchKernelGGL(bin_wg_scan_kernel_batch, dim3(p.total[kBsBinScan]),
					   dim3(kScanCols * kScanParts), 0, stream, p.frames, p.stage_pre(kBsBinScan), p.num_frames);
	hipLaunchKernelGGL(bin_categorize_kernel_batch, dim3(p.total[kBsBinCategorize]), dim3(1024), 0,
					   stream, p.frames, p.stage_pre(kBsBinCategorize), p.num_frames);
	hipLaunchKernelGGL(bin_dispatch_kernel_batch, dim3(p.total[kBsBinDispatch]), dim3(kBinThreads),
					   p.bin_lds, stream, p.frames, p.stage_pre(kBsBinDispatch), p.num_frames);
	return hipGetLastError();
}

hipError_t launch_binning(const FrameConst &c, const FrameBufs &b, hipStream_t stream) {
	size_t lds = (size_t)2 * c.bin_count * sizeof(uint32_t);
	hipLaunchKernelGGL(bin_count_kernel, dim3(c.num_bin_wgs), dim3(kBinThreads), lds, stream, c, b);
	hipLaunchKernelGGL(bin_wg_scan_kernel, dim3((2 * c.bin_count + kScanCols - 1) / kScanCols),
					   dim3(kScanCols * kScanParts), 0, stream, c, b);
	hipLaunchKernelGGL(bin_categorize_kernel, dim3(1), dim3(1024), 0, stream, c, b);
	hipLaunchKernelGGL(bin_dispatch_kernel, dim3(c.num_bin_wgs), dim3(kBinThreads), lds, stream, c,
					   b);
	return hipGetLastError();
}

} // namespace lucid
