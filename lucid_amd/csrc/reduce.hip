// Gradient reduction of the backward pass: the gradient rows (one per block list entry and
// one per fragment of an exactly sorted pixel) are grouped by triangle id with a counting
// sort (atomics only count and place row indices), and each triangle's group is summed in
// 64-bit fixed point, whose sums do not depend on the order of the rows. The per-vertex
// sums then run over the triangle corners grouped by vertex (a stable rocPRIM radix sort of
// the corner -> vertex map, run only when the mesh changes and kept otherwise). No gradient
// value goes through a global atomic, and the result does not depend on scheduling.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "lucid_device.h"

namespace lucid {

namespace {

__global__ void iota_kernel(uint32_t *v, uint32_t n) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i < n)
		v[i] = i;
}

__global__ void zero_kernel(uint32_t *p, size_t n) {
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if(i < n)
		p[i] = 0u;
}

// Gradient rows of the frame (block rows, then the rows of exactly sorted pixels), read on
// the device so that the grouping can be enqueued before the counts are known; 0 after an
// overflow (the pass is then repeated with larger buffers).
__device__ __forceinline__ uint32_t frame_rows(const uint32_t *ctrl, const BwdBufs &g) {
	if(ctrl[kCtrlOverflow])
		return 0u;
	const uint64_t n = (uint64_t)ctrl[kCtrlBwdRows] + ctrl[kCtrlBwdExtra];
	return (uint32_t)(n < g.rows_cap ? n : g.rows_cap);
}

__global__ void row_hist_kernel(const uint32_t *ctrl, BwdBufs g, uint32_t ntri, uint32_t *counts) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i < frame_rows(ctrl, g)) {
		const uint32_t t = g.row_tri[i];
		if(t < ntri)
			atomicAdd(&counts[t], 1u);
	}
}

// Places every row index in its triangle's group; counts run back down to zero.
__global__ void row_scatter_kernel(const uint32_t *ctrl, BwdBufs g, uint32_t ntri,
								   const uint32_t *offsets, uint32_t *counts, uint32_t *grouped) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i < frame_rows(ctrl, g)) {
		const uint32_t t = g.row_tri[i];
		if(t < ntri)
			grouped[offsets[t] + atomicSub(&counts[t], 1u) - 1u] = i;
	}
}

// One thread per triangle: the sum of its group of rows in fixed point at the frame's scale
// 2^(k - 6) (a row is below 2^50 units there), lowered by one bit per doubling of the group
// beyond 4096 rows so that the sum stays below 2^62. The alpha component also goes to
// grad_alpha (when given).
__global__ void tri_group_sum_kernel(FrameConst c, const uint32_t *ctrl, const uint32_t *offsets,
									 uint32_t ntri, const uint32_t *grouped, const float *rows,
									 float *tri_grad, float *grad_alpha) {
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	if(t >= ntri)
		return;
	const uint32_t o0 = offsets[t], o1 = offsets[t + 1];
	const uint32_t n = o1 - o0;
	int k = grad_fixed_exp(c, __uint_as_float(ctrl[kCtrlBwdGradMax])) - 6;
	for(uint32_t cap = 4096; cap < n; cap <<= 1)
		k--;
	const double scale = ldexp(1.0, k);
	unsigned long long acc[kGradWords];
#pragma unroll
	for(int w = 0; w < kGradWords; w++)
		acc[w] = 0ull;
	uint32_t j = o0;
	for(; j + 2 <= o1; j += 2) { // two independent row loads in flight
		const float *r0 = rows + (size_t)grouped[j] * kGradWords;
		const float *r1 = rows + (size_t)grouped[j + 1] * kGradWords;
		float a[kGradWords], b[kGradWords];
#pragma unroll
		for(int w = 0; w < kGradWords; w++)
			a[w] = r0[w], b[w] = r1[w];
#pragma unroll
		for(int w = 0; w < kGradWords; w++)
			acc[w] += to_fixed(a[w], scale) + to_fixed(b[w], scale);
	}
	if(j < o1) {
		const float *r0 = rows + (size_t)grouped[j] * kGradWords;
#pragma unroll
		for(int w = 0; w < kGradWords; w++)
			acc[w] += to_fixed(r0[w], scale);
	}
	float out[kGradWords];
#pragma unroll
	for(int w = 0; w < kGradWords; w++) {
		out[w] = (float)ldexp((double)(long long)acc[w], -k);
		tri_grad[(size_t)t * kGradWords + w] = out[w];
	}
	if(grad_alpha)
		grad_alpha[t] = out[9];
}

// Vertex of each triangle corner: triangle 2 q + s of quad q uses quad vertices
// (0, 1 + s, 2 + s), quads addressed through the instances (quad_setup.glsl:136-160).
// Writes the corner -> vertex map into `vkeys` and sets *changed when it differs from what
// `vkeys` held (the vertex grouping of the previous backward pass is then stale).
__global__ void corner_check_kernel(FrameConst c, FrameBufs b, uint32_t ntri, uint32_t *vkeys,
									uint32_t *changed) {
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	bool diff = false;
	if(t < ntri) {
		const uint32_t q = t >> 1, s = t & 1u;
		int lo = 0, hi = c.num_instances - 1;
		while(lo < hi) {
			const int mid = (lo + hi + 1) >> 1;
			if((uint32_t)b.inst_prefix[mid] <= q)
				lo = mid;
			else
				hi = mid - 1;
		}
		const lucid_instance_data ins = b.instances[lo];
		const uint32_t k = q - (uint32_t)b.inst_prefix[lo];
		const uint32_t *ix = b.quad_indices + (size_t)ins.index_offset + (size_t)k * 4;
		const uint32_t v[3] = {ix[0] + (uint32_t)ins.vertex_offset,
							   ix[1 + s] + (uint32_t)ins.vertex_offset,
							   ix[2 + s] + (uint32_t)ins.vertex_offset};
#pragma unroll
		for(int i = 0; i < 3; i++)
			if(vkeys[3 * t + i] != v[i]) {
				vkeys[3 * t + i] = v[i];
				diff = true;
			}
	}
	if(__ballot(diff) != 0 && (threadIdx.x & 63) == 0)
		atomicOr(changed, 1u);
}

// One thread per sorted corner position; the first position of each vertex sums its run of
// corners in sorted (stable, corner id) order.
__global__ void vertex_segment_sum_kernel(const uint32_t *keys, const uint32_t *corners,
										  uint32_t n, const float *tri_grad, uint32_t nverts,
										  float *grad_colors) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n)
		return;
	const uint32_t k = keys[i];
	if(i > 0 && keys[i - 1] == k)
		return;
	float sum[3] = {0.0f, 0.0f, 0.0f};
	for(uint32_t j = i; j < n && keys[j] == k; j++) {
		const uint32_t p = corners[j], t = p / 3, corner = p - 3 * t;
		const float *g = tri_grad + (size_t)t * kGradWords + corner * 3;
		sum[0] += g[0], sum[1] += g[1], sum[2] += g[2];
	}
	if(k < nverts) {
		grad_colors[(size_t)k * 3 + 0] = sum[0];
		grad_colors[(size_t)k * 3 + 1] = sum[1];
		grad_colors[(size_t)k * 3 + 2] = sum[2];
	}
}

__global__ void alpha_grad_kernel(const float *tri_grad, uint32_t ntri, float *grad_alpha) {
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	if(t < ntri)
		grad_alpha[t] = tri_grad[(size_t)t * kGradWords + 9];
}

// Exclusive prefix sum of u32 counts in three launches: tile totals, their exclusive scan
// (one work-group), then each tile's scan plus its offset. A tile is kScanTile consecutive
// counts, kScanItems per thread in thread order.
constexpr int kScanThreads = 256, kScanItems = 16, kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t scan_load(const uint32_t *in, size_t n, size_t i) {
	return i < n ? in[i] : 0u;
}

__global__ __launch_bounds__(kScanThreads) void scan_tile_sums_kernel(const uint32_t *in, size_t n,
																		uint32_t *tile_sums) {
	__shared__ uint32_t s_wave[kScanThreads / 64];
	const size_t base = (size_t)blockIdx.x * kScanTile;
	uint32_t sum = 0;
#pragma unroll
	for(int k = 0; k < kScanItems; k++) // coalesced: item k of the tile's k-th row
		sum += scan_load(in, n, base + (size_t)k * kScanThreads + threadIdx.x);
	sum = __shfl(wave_inclusive_add(sum), 63, 64);
	if(lane_id() == 0)
		s_wave[threadIdx.x >> 6] = sum;
	__syncthreads();
	if(threadIdx.x == 0) {
		uint32_t t = 0;
		for(int w = 0; w < kScanThreads / 64; w++)
			t += s_wave[w];
		tile_sums[blockIdx.x] = t;
	}
}

// In place: tile_sums[i] becomes the sum of the tiles before i.
__global__ __launch_bounds__(1024) void scan_tile_offsets_kernel(uint32_t *tile_sums, uint32_t ntiles) {
	__shared__ uint32_t s_wave[16];
	__shared__ uint32_t s_carry;
	if(threadIdx.x == 0)
		s_carry = 0;
	__syncthreads();
	for(uint32_t base = 0; base < ntiles; base += 1024) {
		const uint32_t i = base + threadIdx.x;
		const uint32_t v = i < ntiles ? tile_sums[i] : 0u;
		const uint32_t inc = wave_inclusive_add(v);
		if(lane_id() == 63)
			s_wave[threadIdx.x >> 6] = inc;
		__syncthreads();
		uint32_t before = s_carry, total = 0;
		for(int w = 0; w < 16; w++) {
			const uint32_t t = s_wave[w];
			before += w < (int)(threadIdx.x >> 6) ? t : 0u;
			total += t;
		}
		if(i < ntiles)
			tile_sums[i] = before + inc - v;
		__syncthreads(); // s_wave and s_carry are read before they change
		if(threadIdx.x == 0)
			s_carry += total;
		__syncthreads();
	}
}

__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(const uint32_t *in, size_t n,
																	const uint32_t *tile_offsets,
																	uint32_t *out) {
	__shared__ uint32_t s_tile[kScanTile];
	__shared__ uint32_t s_wave[kScanThreads / 64];
	const size_t base = (size_t)blockIdx.x * kScanTile;
#pragma unroll
	for(int k = 0; k < kScanItems; k++) { // coalesced into LDS, then read in thread order
		const int j = k * kScanThreads + (int)threadIdx.x;
		s_tile[j] = scan_load(in, n, base + (size_t)j);
	}
	__syncthreads();
	uint32_t v[kScanItems], sum = 0;
#pragma unroll
	for(int k = 0; k < kScanItems; k++) {
		// rotated reads: thread t reads its items starting at item t % 16 (no bank conflicts)
		const int kk = (k + (int)threadIdx.x) & (kScanItems - 1);
		v[kk] = s_tile[threadIdx.x * kScanItems + kk];
	}
#pragma unroll
	for(int k = 0; k < kScanItems; k++)
		sum += v[k];
	const uint32_t inc = wave_inclusive_add(sum);
	if(lane_id() == 63)
		s_wave[threadIdx.x >> 6] = inc;
	__syncthreads();
	uint32_t run = tile_offsets[blockIdx.x] + inc - sum;
	for(int w = 0; w < (int)(threadIdx.x >> 6); w++)
		run += s_wave[w];
#pragma unroll
	for(int k = 0; k < kScanItems; k++) {
		const uint32_t x = v[k];
		v[k] = run;
		run += x;
	}
	__syncthreads(); // every thread has read its items
#pragma unroll
	for(int k = 0; k < kScanItems; k++) {
		const int kk = (k + (int)threadIdx.x) & (kScanItems - 1);
		s_tile[threadIdx.x * kScanItems + kk] = v[kk];
	}
	__syncthreads();
#pragma unroll
	for(int k = 0; k < kScanItems; k++) {
		const int j = k * kScanThreads + (int)threadIdx.x;
		if(base + (size_t)j < n)
			out[base + (size_t)j] = s_tile[j];
	}
}

inline uint32_t scan_tiles(size_t n) { return (uint32_t)((n + kScanTile - 1) / kScanTile); }

// out[i] = in[0] + ... + in[i - 1] for i < n; `temp` holds scan_tiles(n) words.
hipError_t exclusive_sum(const uint32_t *in, uint32_t *out, size_t n, uint32_t *temp,
						 hipStream_t stream) {
	const uint32_t nt = scan_tiles(n);
	if(!nt)
		return hipSuccess;
	hipLaunchKernelGGL(scan_tile_sums_kernel, dim3(nt), dim3(kScanThreads), 0, stream, in, n, temp);
	hipLaunchKernelGGL(scan_tile_offsets_kernel, dim3(1), dim3(1024), 0, stream, temp, nt);
	hipLaunchKernelGGL(scan_apply_kernel, dim3(nt), dim3(kScanThreads), 0, stream, in, n, temp, out);
	return hipGetLastError();
}

int key_bits(uint32_t max_key) {
	int bits = 1;
	while(bits < 32 && (max_key >> bits) != 0)
		bits++;
	return bits;
}

inline dim3 grid_for(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

} // namespace

// Temporary storage of the reduction: one stable key/value radix sort of `npairs` pairs or
// one exclusive scan of `nscan` counts.
size_t reduce_temp_bytes(uint32_t npairs, uint32_t nscan) {
	size_t sort = 0;
	(void)rocprim::radix_sort_pairs((void *)nullptr, sort, (const uint32_t *)nullptr,
									(uint32_t *)nullptr, (const uint32_t *)nullptr,
									(uint32_t *)nullptr, npairs, 0, 32);
	const size_t scan = (size_t)scan_tiles(nscan) * 4;
	return sort > scan ? sort : scan;
}

// Corner -> vertex map of the frame's triangles into vkeys (3 per triangle); sets *changed
// when it differs from the map already there.
hipError_t reduce_check_corners(const FrameConst &c, const FrameBufs &b, uint32_t ntri,
								uint32_t *vkeys, uint32_t *changed, hipStream_t stream) {
	if(ntri)
		hipLaunchKernelGGL(corner_check_kernel, grid_for(ntri), dim3(256), 0, stream, c, b, ntri,
						   vkeys, changed);
	return hipGetLastError();
}

hipError_t reduce_vertices(const ReduceBufs &r, uint32_t ntri, uint32_t nverts, bool regroup,
						   float *grad_colors, float *grad_alpha, hipStream_t stream);

// Groups the gradient rows by triangle (counts, offsets, grouped row indices); the row
// count is read on the device, `rows_cap` bounds the grid.
hipError_t reduce_group_rows(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
							 const ReduceBufs &r, uint32_t ntri, hipStream_t stream) {
	hipLaunchKernelGGL(zero_kernel, grid_for((size_t)ntri + 1), dim3(256), 0, stream, r.counts,
					   (size_t)ntri + 1);
	if(g.rows_cap)
		hipLaunchKernelGGL(row_hist_kernel, grid_for(g.rows_cap), dim3(256), 0, stream, b.ctrl, g,
						   ntri, r.counts);
	if(r.temp_bytes < (size_t)scan_tiles((size_t)ntri + 1) * 4)
		return hipErrorInvalidValue;
	hipError_t e = exclusive_sum(r.counts, r.offsets, (size_t)ntri + 1, (uint32_t *)r.temp, stream);
	if(e != hipSuccess)
		return e;
	if(g.rows_cap)
		hipLaunchKernelGGL(row_scatter_kernel, grid_for(g.rows_cap), dim3(256), 0, stream, b.ctrl,
						   g, ntri, r.offsets, r.counts, r.grouped);
	return hipGetLastError();
}

// Sums the grouped rows into dL/d(vertex rgb) (3 floats per vertex) and dL/d(triangle
// alpha) (per triangle id). The corners grouped by vertex (r.vkeys sorted into r.vsorted /
// r.vcorners) depend only on the mesh: they are regrouped only when `regroup` is set.
hipError_t reduce_sum(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
					  const ReduceBufs &r, uint32_t ntri, uint32_t nverts, bool regroup,
					  float *grad_colors, float *grad_alpha, hipStream_t stream) {
	if(ntri)
		hipLaunchKernelGGL(tri_group_sum_kernel, grid_for(ntri), dim3(256), 0, stream, c, b.ctrl,
						   r.offsets, ntri, r.grouped, g.rows, r.tri_grad, grad_alpha);
	// (the alpha gradients are written above)
	return reduce_vertices(r, ntri, nverts, regroup, grad_colors, nullptr, stream);
}

// Per-vertex colour sums (and the alpha copy) from r.tri_grad.
hipError_t reduce_vertices(const ReduceBufs &r, uint32_t ntri, uint32_t nverts, bool regroup,
						   float *grad_colors, float *grad_alpha, hipStream_t stream) {
	if(grad_colors) {
		hipLaunchKernelGGL(zero_kernel, grid_for((size_t)nverts * 3), dim3(256), 0, stream,
						   (uint32_t *)grad_colors, (size_t)nverts * 3);
		const uint32_t np = 3 * ntri;
		if(regroup && np) {
			hipLaunchKernelGGL(iota_kernel, grid_for(np), dim3(256), 0, stream, r.viota, np);
			size_t bytes = r.temp_bytes;
			hipError_t e = rocprim::radix_sort_pairs(
				r.temp, bytes, (const uint32_t *)r.vkeys, r.vsorted, (const uint32_t *)r.viota,
				r.vcorners, np, 0, key_bits(nverts), stream);
			if(e != hipSuccess)
				return e;
		}
		if(np)
			hipLaunchKernelGGL(vertex_segment_sum_kernel, grid_for(np), dim3(256), 0, stream,
							   r.vsorted, r.vcorners, np, r.tri_grad, nverts, grad_colors);
	}
	if(grad_alpha && ntri)
		hipLaunchKernelGGL(alpha_grad_kernel, grid_for(ntri), dim3(256), 0, stream, r.tri_grad,
						   ntri, grad_alpha);
	return hipGetLastError();
}

} // namespace lucid
