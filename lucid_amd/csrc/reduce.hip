// Gradient reduction of the backward pass: per-(block, entry) and per-fragment gradient rows
// are summed per triangle, then per vertex, by stable radix sorts on the triangle / vertex
// id (rocPRIM) followed by segmented sums in a fixed order. Rows of one key are summed in
// ascending row order, so the result does not depend on scheduling, and no gradient value
// goes through a global atomic.
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

__global__ void zero_kernel(float *p, size_t n) {
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if(i < n)
		p[i] = 0.0f;
}

// One thread per sorted position; the first position of each key sums its run of rows.
__global__ void tri_segment_sum_kernel(const uint32_t *keys, const uint32_t *rows_sorted,
									   uint32_t n, const float *rows, float *tri_grad) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n)
		return;
	const uint32_t k = keys[i];
	if(i > 0 && keys[i - 1] == k)
		return;
	float sum[kGradWords];
#pragma unroll
	for(int w = 0; w < kGradWords; w++)
		sum[w] = 0.0f;
	for(uint32_t j = i; j < n && keys[j] == k; j++) {
		const float *r = rows + (size_t)rows_sorted[j] * kGradWords;
#pragma unroll
		for(int w = 0; w < kGradWords; w++)
			sum[w] += r[w];
	}
#pragma unroll
	for(int w = 0; w < kGradWords; w++)
		tri_grad[(size_t)k * kGradWords + w] = sum[w];
}

// Vertex of each triangle corner: triangle 2 q + s of quad q uses quad vertices
// (0, 1 + s, 2 + s), quads addressed through the instances (quad_setup.glsl:136-160).
__global__ void corner_pairs_kernel(FrameConst c, FrameBufs b, uint32_t ntri, uint32_t *vkeys,
									uint32_t *vvals) {
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	if(t >= ntri)
		return;
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
	const uint32_t v0 = ix[0] + (uint32_t)ins.vertex_offset;
	const uint32_t v1 = ix[1 + s] + (uint32_t)ins.vertex_offset;
	const uint32_t v2 = ix[2 + s] + (uint32_t)ins.vertex_offset;
	vkeys[3 * t + 0] = v0, vkeys[3 * t + 1] = v1, vkeys[3 * t + 2] = v2;
	vvals[3 * t + 0] = 3 * t + 0, vvals[3 * t + 1] = 3 * t + 1, vvals[3 * t + 2] = 3 * t + 2;
}

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

int key_bits(uint32_t max_key) {
	int bits = 1;
	while(bits < 32 && (max_key >> bits) != 0)
		bits++;
	return bits;
}

inline dim3 grid_for(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

} // namespace

// Temporary storage of one stable key/value radix sort of n pairs.
size_t reduce_sort_temp_bytes(uint32_t n) {
	size_t bytes = 0;
	(void)rocprim::radix_sort_pairs((void *)nullptr, bytes, (const uint32_t *)nullptr,
									(uint32_t *)nullptr, (const uint32_t *)nullptr,
									(uint32_t *)nullptr, n, 0, 32);
	return bytes;
}

// Sums the gradient rows into dL/d(vertex rgb) (3 floats per vertex) and dL/d(triangle
// alpha) (per triangle id). Scratch: keys/vals of max(nrows, 3 ntri) pairs, tri_grad of
// ntri * kGradWords floats, `temp` of reduce_sort_temp_bytes bytes.
hipError_t reduce_gradients(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
							uint32_t nrows, uint32_t ntri, uint32_t nverts, uint32_t *keys_out,
							uint32_t *vals_in, uint32_t *vals_out, uint32_t *keys_in2,
							float *tri_grad, void *temp, size_t temp_bytes, float *grad_colors,
							float *grad_alpha, hipStream_t stream) {
	hipError_t e;
	hipLaunchKernelGGL(zero_kernel, grid_for((size_t)ntri * kGradWords), dim3(256), 0, stream,
					   tri_grad, (size_t)ntri * kGradWords);
	if(nrows) {
		hipLaunchKernelGGL(iota_kernel, grid_for(nrows), dim3(256), 0, stream, vals_in, nrows);
		size_t bytes = temp_bytes;
		e = rocprim::radix_sort_pairs(temp, bytes, (const uint32_t *)g.row_tri, keys_out,
									  (const uint32_t *)vals_in, vals_out, nrows, 0,
									  key_bits(ntri), stream);
		if(e != hipSuccess)
			return e;
		hipLaunchKernelGGL(tri_segment_sum_kernel, grid_for(nrows), dim3(256), 0, stream,
						   keys_out, vals_out, nrows, g.rows, tri_grad);
	}
	if(grad_colors) {
		hipLaunchKernelGGL(zero_kernel, grid_for((size_t)nverts * 3), dim3(256), 0, stream,
						   grad_colors, (size_t)nverts * 3);
		const uint32_t np = 3 * ntri;
		hipLaunchKernelGGL(corner_pairs_kernel, grid_for(ntri), dim3(256), 0, stream, c, b, ntri,
						   keys_in2, vals_in);
		size_t bytes = temp_bytes;
		e = rocprim::radix_sort_pairs(temp, bytes, (const uint32_t *)keys_in2, keys_out,
									  (const uint32_t *)vals_in, vals_out, np, 0,
									  key_bits(nverts), stream);
		if(e != hipSuccess)
			return e;
		hipLaunchKernelGGL(vertex_segment_sum_kernel, grid_for(np), dim3(256), 0, stream,
						   keys_out, vals_out, np, tri_grad, nverts, grad_colors);
	}
	if(grad_alpha)
		hipLaunchKernelGGL(alpha_grad_kernel, grid_for(ntri), dim3(256), 0, stream, tri_grad, ntri,
						   grad_alpha);
	return hipGetLastError();
}

} // namespace lucid
