// Tile rasterisation, per-pixel depth ordering and front-to-back blending
// (≙ data/shaders/raster_low.glsl, raster_high.glsl, shared/raster.glsl, shared/shading.glsl).
//
// Work decomposition (wave64, one lane per pixel of an 8x8 block):
//   raster_low   one 256-thread work-group per bin with fewer than 1024 triangles. Builds
//                block-row records with the reference's 8-row scanline rule, promotes the bin
//                to raster_high when an 8x8 block holds more than 256 triangles
//                (raster_low.glsl:90-106,229-237), otherwise each wave rasterises 4 blocks.
//   raster_high  one 256-thread work-group per (high bin, 8-row band) with the 4-row rule of
//                raster_high.glsl:54-90; each wave rasterises one 8x8 block.
//   fallback     exact per-pixel sort for pixels whose speculative ordering was violated.
//
// Exact ordering: a block's triangles are sorted by the depth at their covered-pixel centroid
// (wave bitonic in registers / LDS). Each pixel lane then consumes its fragments in that
// order through a K-entry insertion window keyed by (depth at the pixel, triangle id) and
// checks every arrival against the last emitted key. If no arrival is ever smaller than the
// last emitted key, the emitted sequence is exactly the per-pixel sorted order; otherwise the
// pixel is queued for the fallback kernel, which sorts the pixel's full fragment list
// (LDS bitonic, global-memory spill for deep pixels). Ties in depth are ordered by triangle id.
#include "lucid_device.h"

namespace lucid {

constexpr int kRasterThreads = 256;
constexpr int kWindow = 4;
constexpr int kLowCap = 256;    // == kMaxBlockTris
constexpr int kHighCap = 1024;  // block lists above this go to the fallback kernel
constexpr int kFallbackCap = 512; // fragments per pixel sorted in LDS by the fallback kernel

// ------------------------------------------------------------------ shading

__device__ __forceinline__ float srgb_to_linear(float x) {
	return x < 0.04045f ? (1.0f / 12.92f) * x : powf((x + 0.055f) * (1.0f / 1.055f), 2.4f);
}
__device__ __forceinline__ float linear_to_srgb(float x) {
	return x < 0.0031308f ? 12.92f * x : 1.055f * powf(x, 1.0f / 2.4f) - 0.055f;
}

struct Shaded {
	uint32_t enc; // RGBA8 sample (quantised mode); 0 = contributes nothing
	float4 col;   // fp32 sample (fp32 mode)
};

__device__ __forceinline__ float sample_inv_depth(const FrameBufs &b, uint32_t t, float px,
												  float py) {
	uint4 d = b.tri_depth[t];
	return bits2f(d.x) * px + (bits2f(d.y) * py + bits2f(d.z));
}

// shadeSample (shading.glsl:107-184) without textures.
template <bool FP32>
__device__ __forceinline__ Shaded shade_sample(const FrameConst &c, const FrameBufs &b, uint32_t t,
											   float px, float py, float inv_ray) {
	uint4 d = b.tri_depth[t];
	uint32_t flags = d.w & 0xffffu, inst = d.w >> 16;
	uint4 b0 = b.tri_bary[2 * (size_t)t], b1 = b.tri_bary[2 * (size_t)t + 1];
	float ray_pos = 1.0f / inv_ray;
	float e0 = bits2f(b0.x) * px + (bits2f(b0.y) * py + bits2f(b0.z));
	float e1 = bits2f(b1.x) * px + (bits2f(b1.y) * py + bits2f(b1.z));
	float bary0 = e0 * ray_pos, bary1 = e1 * ray_pos;
	bary0 = bary0 - bits2f(b0.w);
	bary1 = bary1 - bits2f(b1.w);

	float4 color = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
	if((flags & LUCID_INST_HAS_COLOR) && b.inst_colors)
		color = decode_rgba8(b.inst_colors[inst]);
	if(flags & LUCID_INST_HAS_VERTEX_COLORS) {
		float w0 = (1.0f - bary0) - bary1;
		if(c.color_mode == kColorRGBA8) {
			uint4 qc = b.quad_color[t >> 1];
			bool s = t & 1u;
			float4 c0 = decode_rgba8(qc.x), c1 = decode_rgba8(s ? qc.z : qc.y),
				   c2 = decode_rgba8(s ? qc.w : qc.z);
			color.x = color.x * (w0 * c0.x + (bary0 * c1.x + bary1 * c2.x));
			color.y = color.y * (w0 * c0.y + (bary0 * c1.y + bary1 * c2.y));
			color.z = color.z * (w0 * c0.z + (bary0 * c1.z + bary1 * c2.z));
			color.w = color.w * (w0 * c0.w + (bary0 * c1.w + bary1 * c2.w));
		} else if(c.color_mode == kColorF32) {
			float4 c0 = b.tri_color[3 * (size_t)t], c1 = b.tri_color[3 * (size_t)t + 1],
				   c2 = b.tri_color[3 * (size_t)t + 2];
			color.x = color.x * (w0 * c0.x + (bary0 * c1.x + bary1 * c2.x));
			color.y = color.y * (w0 * c0.y + (bary0 * c1.y + bary1 * c2.y));
			color.z = color.z * (w0 * c0.z + (bary0 * c1.z + bary1 * c2.z));
			color.w = color.w * c0.w;
		}
	}
	Shaded out;
	if(color.w == 0.0f) {
		out.enc = 0;
		out.col = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		return out;
	}
	V3 normal;
	if(flags & LUCID_INST_HAS_VERTEX_NORMALS) {
		uint4 nq = b.quad_normal[t >> 1];
		bool s = t & 1u;
		V3 n0 = decode_normal(nq.x), n1 = decode_normal(s ? nq.z : nq.y),
		   n2 = decode_normal(s ? nq.w : nq.z);
		n1 = sub3(n1, n0);
		n2 = sub3(n2, n0);
		normal = mk3(bary0 * n1.x + (bary1 * n2.x + n0.x), bary0 * n1.y + (bary1 * n2.y + n0.y),
					 bary0 * n1.z + (bary1 * n2.z + n0.z));
	} else {
		normal = decode_normal(b.tri_normal[t]);
	}
	float light = fmx(0.0f, dot3(neg3(ld3(c.sun_dir)), normal) * 0.7f + 0.3f);
	float rgb[3] = {color.x, color.y, color.z};
#pragma unroll
	for(int k = 0; k < 3; k++) {
		float amb = c.ambient[k] * c.ambient_power;
		float dif = (c.sun_color[k] * c.sun_power) * light;
		rgb[k] = sat(linear_to_srgb(srgb_to_linear(rgb[k]) * (amb + dif)));
	}
	if(FP32) {
		out.enc = 1;
		out.col = make_float4(rgb[0], rgb[1], rgb[2], color.w);
	} else {
		out.enc = encode_rgba8(rgb[0], rgb[1], rgb[2], color.w);
		out.col = decode_rgba8(out.enc);
	}
	return out;
}

__device__ __forceinline__ uint64_t pixel_key(float inv_ray, uint32_t t) {
	return ((uint64_t)(~ord_f32(inv_ray)) << 32) | t;
}

// ------------------------------------------------------------------ per-pixel state

template <bool FP32>
struct PixelState {
	uint64_t wk[kWindow];
	float4 wc[kWindow];
	int n;
	uint64_t last;
	bool bad;
	float cr, cg, cb, tr;

	__device__ __forceinline__ void init() {
#pragma unroll
		for(int i = 0; i < kWindow; i++)
			wk[i] = ~0ull, wc[i] = make_float4(0, 0, 0, 0);
		n = 0, last = 0, bad = false;
		cr = cg = cb = 0.0f, tr = 1.0f;
	}

	__device__ __forceinline__ void emit(float4 col, bool additive) {
		if(additive) {
			cr = cr + col.x * col.w;
			cg = cg + col.y * col.w;
			cb = cb + col.z * col.w;
		} else {
			cr = cr + (col.x * col.w) * tr;
			cg = cg + (col.y * col.w) * tr;
			cb = cb + (col.z * col.w) * tr;
			tr = tr * (1.0f - col.w);
		}
	}

	__device__ __forceinline__ void insert(uint64_t key, float4 col, bool additive) {
		if(key < last) {
			bad = true;
			return;
		}
		if(n < kWindow) {
#pragma unroll
			for(int i = 0; i < kWindow; i++)
				if(i == n)
					wk[i] = key, wc[i] = col;
			n++;
			return;
		}
		uint64_t mk = key;
		int mi = -1;
#pragma unroll
		for(int i = 0; i < kWindow; i++)
			if(wk[i] < mk)
				mk = wk[i], mi = i;
		float4 out = col;
#pragma unroll
		for(int i = 0; i < kWindow; i++)
			if(i == mi) {
				out = wc[i];
				wk[i] = key, wc[i] = col;
			}
		emit(out, additive);
		last = mk;
	}

	__device__ __forceinline__ void flush(bool additive) {
#pragma unroll
		for(int r = 0; r < kWindow; r++) {
			uint64_t mk = ~0ull;
			int mi = -1;
#pragma unroll
			for(int i = 0; i < kWindow; i++)
				if(wk[i] < mk)
					mk = wk[i], mi = i;
			if(mi >= 0) {
				float4 out = make_float4(0, 0, 0, 0);
#pragma unroll
				for(int i = 0; i < kWindow; i++)
					if(i == mi)
						out = wc[i], wk[i] = ~0ull;
				emit(out, additive);
			}
		}
	}
};

__device__ __forceinline__ void write_pixel(const FrameConst &c, const FrameBufs &b, int px,
											int py, float cr, float cg, float cb, float tr,
											uint32_t count) {
	if(px >= c.width || py >= c.height)
		return;
	float r = sat(cr + tr * c.background[0]);
	float g = sat(cg + tr * c.background[1]);
	float bl = sat(cb + tr * c.background[2]);
	size_t o = (size_t)py * c.width + px;
	if(b.image) {
		uint32_t v = f2u(r * 255.0f + 0.5f) | (f2u(g * 255.0f + 0.5f) << 8) |
					 (f2u(bl * 255.0f + 0.5f) << 16) | (255u << 24);
		((uint32_t *)b.image)[o] = v;
	}
	if(b.image_f)
		((float4 *)b.image_f)[o] = make_float4(r, g, bl, 1.0f);
	if(b.frag_counts)
		b.frag_counts[o] = count;
}

// ------------------------------------------------------------------ wave sorts

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
	uint32_t lo = __shfl((uint32_t)v, src, 64), hi = __shfl((uint32_t)(v >> 32), src, 64);
	return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
	uint32_t lo = __shfl_xor((uint32_t)v, m, 64), hi = __shfl_xor((uint32_t)(v >> 32), m, 64);
	return ((uint64_t)hi << 32) | lo;
}

// Ascending bitonic sort of 64*E keys held in registers; element i = e * 64 + lane.
template <int E>
__device__ __forceinline__ void bitonic_regs(uint64_t (&v)[E]) {
	const int lane = lane_id();
#pragma unroll
	for(int k = 2; k <= 64 * E; k <<= 1) {
#pragma unroll
		for(int j = k >> 1; j > 0; j >>= 1) {
			if(j >= 64) {
				const int jr = j / 64;
#pragma unroll
				for(int e = 0; e < E; e++) {
					if((e & jr) == 0) {
						const int pe = e | jr;
						bool asc = (((e * 64) & k) == 0);
						uint64_t a = v[e], bb = v[pe];
						bool sw = (a > bb) == asc;
						v[e] = sw ? bb : a;
						v[pe] = sw ? a : bb;
					}
				}
			} else {
#pragma unroll
				for(int e = 0; e < E; e++) {
					int i = e * 64 + lane;
					uint64_t p = shfl_xor64(v[e], j);
					bool asc = (i & k) == 0;
					bool lower = (lane & j) == 0;
					uint64_t mn = v[e] < p ? v[e] : p, mx = v[e] < p ? p : v[e];
					v[e] = (asc == lower) ? mn : mx;
				}
			}
		}
	}
}

// Ascending bitonic sort of P (power of two, >= 2) keys in LDS by one wave.
__device__ void bitonic_lds(uint64_t *key, uint32_t P) {
	const int lane = lane_id();
	for(uint32_t k = 2; k <= P; k <<= 1) {
		for(uint32_t j = k >> 1; j > 0; j >>= 1) {
			for(uint32_t i = lane; i < P / 2; i += 64) {
				uint32_t lo = ((i & ~(j - 1)) << 1) | (i & (j - 1));
				uint32_t hi = lo | j;
				bool asc = (lo & k) == 0;
				uint64_t a = key[lo], bb = key[hi];
				if((a > bb) == asc) {
					key[lo] = bb;
					key[hi] = a;
				}
			}
			wave_sync_lds();
		}
	}
}

// Same on global memory (deep pixels); agent-scope fences make each stage visible.
__device__ void bitonic_global(uint64_t *key, uint32_t P) {
	const int lane = lane_id();
	for(uint32_t k = 2; k <= P; k <<= 1) {
		for(uint32_t j = k >> 1; j > 0; j >>= 1) {
			for(uint32_t i = lane; i < P / 2; i += 64) {
				uint32_t lo = ((i & ~(j - 1)) << 1) | (i & (j - 1));
				uint32_t hi = lo | j;
				bool asc = (lo & k) == 0;
				uint64_t a = __hip_atomic_load(&key[lo], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				uint64_t bb = __hip_atomic_load(&key[hi], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				if((a > bb) == asc) {
					__hip_atomic_store(&key[lo], bb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
					__hip_atomic_store(&key[hi], a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				}
			}
			__threadfence();
			__builtin_amdgcn_wave_barrier();
		}
	}
}

__device__ __forceinline__ uint32_t next_pow2(uint32_t x) {
	return x <= 1 ? 1u : 1u << (32 - __clz(x - 1));
}

// ------------------------------------------------------------------ block raster (one wave)

struct WaveLds {
	uint64_t *key; // [cap]: sort keys, then block masks in sorted order
	uint32_t *tri; // [cap]
	uint32_t *prefix; // [cap + 1]
	uint32_t *own;  // [64]
	uint64_t *pm;   // [64]
};

struct BlockRef {
	int bin, by, bx;    // bin id, block row/col inside the bin
	int px0, py0;       // absolute pixel of the block's top-left corner
	const uint4 *recs;  // block-row records
	uint32_t rec_offset;
	uint32_t nrec;
	int level;          // LUCID_BIN_LEVEL_LOW or _HIGH (selects the record buffer)
};

// Centroid key of a record inside block column bx, plus its coverage mask.
__device__ __forceinline__ uint64_t block_key(const FrameBufs &b, uint4 rec, int bx, int px0,
											  int py0, uint32_t idx) {
	int n = 0, sx = 0, sy = 0;
	int x0 = bx * 8;
#pragma unroll
	for(int r = 0; r < 8; r++) {
		uint32_t sp = record_span(rec, r);
		int a = (int)(sp & 31u) - x0, e = (int)((sp >> 5) & 31u) - x0;
		a = a < 0 ? 0 : a;
		e = e > 7 ? 7 : e;
		int cnt = e >= a ? e - a + 1 : 0;
		n += cnt;
		sx += (a + e) * cnt;
		sy += r * cnt;
	}
	float inv_n = 1.0f / (float)(n > 0 ? n : 1);
	float cx = (float)px0 + (float)sx * 0.5f * inv_n;
	float cy = (float)py0 + (float)sy * inv_n;
	float inv = sample_inv_depth(b, rec.x, cx, cy);
	return ((uint64_t)(~ord_f32(inv)) << 32) | idx;
}

// `whole` marks a block whose fragments were not counted by the raster kernel.
__device__ __forceinline__ void enqueue_fallback(const FrameBufs &b, const BlockRef &br,
												 uint64_t mask, bool whole = false) {
	uint32_t slot = atomicAdd(&b.ctrl[kCtrlFallbackCount], 1u);
	if(slot >= b.fallback_cap) {
		atomicOr(&b.ctrl[kCtrlOverflow], 4u);
		return;
	}
	b.fallback_queue[2 * slot] =
		make_uint4((uint32_t)br.bin,
				   (uint32_t)(br.by * 4 + br.bx) | ((uint32_t)br.level << 8) | (whole ? 1u << 16 : 0u),
				   br.rec_offset, br.nrec);
	b.fallback_queue[2 * slot + 1] =
		make_uint4((uint32_t)mask, (uint32_t)(mask >> 32), (uint32_t)br.px0, (uint32_t)br.py0);
}

// Rasterises one 8x8 block with one wave. `sel` lists the record indices of the block
// (m entries, m <= cap). Returns the number of fragments (for statistics).
template <bool FP32, int RegE>
__device__ uint32_t raster_block(const FrameConst &c, const FrameBufs &b, const WaveLds &L,
								 const BlockRef &br, const uint16_t *sel16, const uint32_t *sel32,
								 uint32_t m) {
	const int lane = lane_id();
	const bool additive = (c.opts & (1u << 5)) != 0;
	// --- 1. keys + sort
	uint32_t P = next_pow2(m < 64 ? 64 : m);
	if(RegE > 0 && P <= 64u * RegE) {
		// register sort; pick the smallest E that holds P
		auto load = [&](uint32_t j) -> uint64_t {
			if(j >= m)
				return ~0ull;
			uint32_t ri = sel16 ? sel16[j] : sel32[j];
			return block_key(b, br.recs[ri], br.bx, br.px0, br.py0, j);
		};
		if(P <= 64) {
			uint64_t v[1] = {load(lane)};
			bitonic_regs<1>(v);
			L.key[lane] = v[0];
		} else if(P <= 128) {
			uint64_t v[2] = {load(lane), load(64 + lane)};
			bitonic_regs<2>(v);
			L.key[lane] = v[0], L.key[64 + lane] = v[1];
		} else {
			uint64_t v[4] = {load(lane), load(64 + lane), load(128 + lane), load(192 + lane)};
			bitonic_regs<4>(v);
#pragma unroll
			for(int e = 0; e < 4; e++)
				L.key[e * 64 + lane] = v[e];
		}
	} else {
		for(uint32_t j = lane; j < P; j += 64) {
			uint64_t k = ~0ull;
			if(j < m) {
				uint32_t ri = sel16 ? sel16[j] : sel32[j];
				k = block_key(b, br.recs[ri], br.bx, br.px0, br.py0, j);
			}
			L.key[j] = k;
		}
		wave_sync_lds();
		bitonic_lds(L.key, P);
	}
	wave_sync_lds();
	// --- 2. sorted masks, tri ids and sample prefix sums
	uint32_t carry = 0;
	for(uint32_t base = 0; base < m; base += 64) {
		uint32_t i = base + lane;
		uint32_t cnt = 0;
		if(i < m) {
			uint32_t j = (uint32_t)L.key[i];
			uint32_t ri = sel16 ? sel16[j] : sel32[j];
			uint4 rec = br.recs[ri];
			uint64_t mask = record_block_mask(rec, br.bx);
			cnt = (uint32_t)__popcll(mask);
			L.key[i] = mask;
			L.tri[i] = rec.x;
		}
		uint32_t inc = wave_inclusive_add(cnt);
		if(i < m)
			L.prefix[i] = carry + inc - cnt;
		carry += __shfl(inc, 63, 64);
	}
	if(lane == 0)
		L.prefix[m] = carry;
	wave_sync_lds();
	const uint32_t total = carry;

	// --- 3. sample chunks
	PixelState<FP32> ps;
	ps.init();
	uint32_t count = 0;
	uint32_t e_cur = 0;
	for(uint32_t s0 = 0; s0 < total; s0 += 64) {
		L.own[lane] = 0;
		L.pm[lane] = 0;
		wave_sync_lds();
		uint32_t e = e_cur + lane;
		if(e < m) {
			uint32_t st = L.prefix[e];
			if(st < s0 + 64u)
				L.own[st > s0 ? st - s0 : 0] = e;
		}
		wave_sync_lds();
		uint32_t own = wave_inclusive_max(L.own[lane]);
		uint32_t s = s0 + lane;
		bool valid = s < total;
		uint64_t key = ~0ull;
		float4 col = make_float4(0, 0, 0, 0);
		uint32_t pix = 0;
		bool contributes = false;
		if(valid) {
			uint32_t k = s - L.prefix[own];
			uint64_t mask = L.key[own];
			pix = (uint32_t)select_bit64(mask, k);
			uint32_t t = L.tri[own];
			float px = (float)(br.px0 + (int)(pix & 7u)), py = (float)(br.py0 + (int)(pix >> 3));
			float inv = sample_inv_depth(b, t, px, py);
			Shaded sh = shade_sample<FP32>(c, b, t, px, py, inv);
			contributes = FP32 ? sh.col.w != 0.0f : sh.enc != 0;
			key = pixel_key(inv, t);
			col = sh.col;
			atomicOr((unsigned long long *)&L.pm[pix], 1ull << lane);
		}
		uint32_t own63 = __shfl(own, 63, 64);
		wave_sync_lds();
		uint64_t bits = L.pm[lane];
		count += (uint32_t)__popcll(bits);
		// only contributing samples take part in ordering
		uint64_t contrib = ballot(contributes);
		bits &= contrib;
		while(ballot(bits != 0)) {
			int src = bits ? (int)__ffsll((unsigned long long)bits) - 1 : lane;
			uint64_t k2 = shfl64(key, src);
			float4 c2;
			c2.x = __shfl(col.x, src, 64);
			c2.y = __shfl(col.y, src, 64);
			c2.z = __shfl(col.z, src, 64);
			c2.w = __shfl(col.w, src, 64);
			if(bits) {
				bits &= bits - 1;
				ps.insert(k2, c2, additive);
			}
		}
		e_cur = (L.prefix[own63 + 1] <= s0 + 64u) ? own63 + 1 : own63;
	}
	ps.flush(additive);
	int px = br.px0 + (lane & 7), py = br.py0 + (lane >> 3);
	write_pixel(c, b, px, py, ps.cr, ps.cg, ps.cb, ps.tr, count);
	bool on_screen = px < c.width && py < c.height;
	uint64_t bad = ballot(ps.bad && on_screen);
	if(bad && lane == 0)
		enqueue_fallback(b, br, bad);
	// fragments of on-screen pixels (stats[0] counts what reaches the image)
	uint32_t visible = on_screen ? count : 0u;
	return __shfl(wave_inclusive_add(visible), 63, 64);
}

// ------------------------------------------------------------------ record generation

// Spans of rows [row_lo, row_hi] (inclusive, inside the bin) for a triangle whose scan
// starts at row `start` (group start of the reference's rule), packed as block-row records.
struct RowGen {
	RowScan scan;
	int row; // next row the scan state corresponds to
};

__device__ __forceinline__ uint4 make_record(uint32_t t, const uint32_t sp[8], uint32_t bxm) {
	return make_uint4(t, sp[0] | (sp[1] << 10) | (sp[2] << 20), sp[3] | (sp[4] << 10) | (sp[5] << 20),
					  sp[6] | (sp[7] << 10) | (bxm << 20));
}

__device__ __forceinline__ uint32_t bin_tri_at(const FrameBufs &b, uint32_t i, uint32_t qc,
											   uint32_t qo, uint32_t to) {
	if(i < 2 * qc) {
		uint32_t e = b.bin_quads[qo + (i >> 1)];
		uint32_t s = i & 1u;
		if((e >> (30 + s)) & 1u)
			return 0xffffffffu;
		return 2 * (e & 0x0fffffffu) + s;
	}
	return b.bin_tris[to + (i - 2 * qc)];
}

__device__ __forceinline__ void background_bin(const FrameConst &c, const FrameBufs &b, int bx0,
											   int by0) {
	for(int i = threadIdx.x; i < 32 * 32; i += kRasterThreads)
		write_pixel(c, b, bx0 + (i & 31), by0 + (i >> 5), 0.0f, 0.0f, 0.0f, 1.0f, 0u);
}

// ------------------------------------------------------------------ raster_low

template <bool FP32>
__global__ __launch_bounds__(kRasterThreads) void raster_low_kernel(FrameConst c, FrameBufs b) {
	if(b.ctrl[kCtrlOverflow] & 1u)
		return;
	const int NB = c.bin_count;
	const int bin = blockIdx.x;
	const uint32_t *counts = b.info + LUCID_INFO_U32_SIZE;
	const uint32_t qc = counts[LUCID_COUNTS_QUAD_COUNTS * NB + bin];
	const uint32_t tc = counts[LUCID_COUNTS_TRI_COUNTS * NB + bin];
	const uint32_t qo = counts[LUCID_COUNTS_QUAD_OFFSETS * NB + bin];
	const uint32_t to = counts[LUCID_COUNTS_TRI_OFFSETS * NB + bin];
	const uint32_t nlevel = tc + 2 * qc;
	const int bin_x = (bin % c.bin_count_x) * kBinSize, bin_y = (bin / c.bin_count_x) * kBinSize;
	if(nlevel == 0) {
		background_bin(c, b, bin_x, bin_y);
		return;
	}
	if(nlevel >= (uint32_t)kLowLevelLimit)
		return;

	__shared__ uint32_t s_row_n[4];
	__shared__ uint32_t s_blk_n[16];
	__shared__ uint16_t s_blk_idx[16][kLowCap];
	__shared__ uint64_t s_key[4][kLowCap];
	__shared__ uint32_t s_tri[4][kLowCap];
	__shared__ uint32_t s_prefix[4][kLowCap + 1];
	__shared__ uint32_t s_own[4][64];
	__shared__ uint64_t s_pm[4][64];
	__shared__ uint32_t s_stats[2];
	if(threadIdx.x < 16)
		s_blk_n[threadIdx.x] = 0;
	if(threadIdx.x < 4)
		s_row_n[threadIdx.x] = 0;
	if(threadIdx.x < 2)
		s_stats[threadIdx.x] = 0;
	__syncthreads();

	uint4 *recs = b.low_records + (size_t)bin * 4 * kLowRowCap;
	// --- block-row records, 8-row rule (raster_low.glsl:39-79)
	for(uint32_t i = threadIdx.x; i < nlevel; i += kRasterThreads) {
		uint32_t t = bin_tri_at(b, i, qc, qo, to);
		if(t == 0xffffffffu)
			continue;
		uint4 s0 = b.tri_scan[2 * (size_t)t], s1 = b.tri_scan[2 * (size_t)t + 1];
		int min_by = min(max((int)(s0.w & 0xffffu) - bin_y, 0), kBinMask) >> 3;
		int max_by = min(max((int)(s0.w >> 16) - bin_y, 0), kBinMask) >> 3;
		RowScan rs;
		row_scan_init(rs, s0, s1, (float)bin_x, (float)(bin_y + min_by * 8));
		for(int by = min_by; by <= max_by; by++) {
			uint32_t sp[8], bxm = 0;
#pragma unroll
			for(int r = 0; r < 8; r++) {
				sp[r] = row_scan_step(rs);
				bxm |= span_bx_mask(sp[r]);
			}
			if(!bxm)
				continue;
			uint32_t idx = atomicAdd(&s_row_n[by], 1u);
			recs[by * kLowRowCap + idx] = make_record(t, sp, bxm);
			for(int bx = 0; bx < 4; bx++)
				if(bxm & (1u << bx)) {
					uint32_t k = atomicAdd(&s_blk_n[by * 4 + bx], 1u);
					if(k < (uint32_t)kLowCap)
						s_blk_idx[by * 4 + bx][k] = (uint16_t)idx;
				}
		}
	}
	__syncthreads();
	bool promote = false;
	for(int k = 0; k < 16; k++)
		promote |= s_blk_n[k] > (uint32_t)kMaxBlockTris;
	if(promote) {
		if(threadIdx.x == 0) {
			lucid_info *info = (lucid_info *)b.info;
			uint32_t id = atomicAdd((uint32_t *)&info->bin_level_counts[LUCID_BIN_LEVEL_HIGH], 1u);
			b.info[LUCID_INFO_U32_SIZE + LUCID_COUNTS_HIGH_BINS * NB + id] = (uint32_t)bin;
		}
		return;
	}
	const int w = threadIdx.x >> 6;
	WaveLds L{s_key[w], s_tri[w], s_prefix[w], s_own[w], s_pm[w]};
	uint32_t frags = 0, pairs = 0;
	for(int by = 0; by < 4; by++) {
		BlockRef br;
		br.bin = bin, br.by = by, br.bx = w;
		br.px0 = bin_x + w * 8, br.py0 = bin_y + by * 8;
		br.recs = recs + by * kLowRowCap;
		br.rec_offset = (uint32_t)((size_t)bin * 4 * kLowRowCap + by * kLowRowCap);
		br.nrec = s_row_n[by];
		br.level = LUCID_BIN_LEVEL_LOW;
		uint32_t m = s_blk_n[by * 4 + w];
		frags += raster_block<FP32, 4>(c, b, L, br, s_blk_idx[by * 4 + w], nullptr, m);
		pairs += m;
	}
	if(lane_id() == 0) {
		atomicAdd(&s_stats[0], frags);
		atomicAdd(&s_stats[1], pairs);
	}
	__syncthreads();
	if(threadIdx.x == 0) {
		lucid_info *info = (lucid_info *)b.info;
		atomicAdd(&info->stats[0], s_stats[0]);
		atomicAdd(&info->stats[1], s_stats[1]);
	}
}

// ------------------------------------------------------------------ raster_high

template <bool FP32>
__global__ __launch_bounds__(kRasterThreads) void raster_high_kernel(FrameConst c, FrameBufs b) {
	if(b.ctrl[kCtrlOverflow] & 1u)
		return;
	const int NB = c.bin_count;
	const lucid_info *info = (const lucid_info *)b.info;
	const uint32_t slot = blockIdx.x >> 2;
	const int by = blockIdx.x & 3;
	if(slot >= (uint32_t)info->bin_level_counts[LUCID_BIN_LEVEL_HIGH])
		return;
	const uint32_t *counts = b.info + LUCID_INFO_U32_SIZE;
	const int bin = (int)counts[LUCID_COUNTS_HIGH_BINS * NB + slot];
	const uint32_t qc = counts[LUCID_COUNTS_QUAD_COUNTS * NB + bin];
	const uint32_t tc = counts[LUCID_COUNTS_TRI_COUNTS * NB + bin];
	const uint32_t qo = counts[LUCID_COUNTS_QUAD_OFFSETS * NB + bin];
	const uint32_t to = counts[LUCID_COUNTS_TRI_OFFSETS * NB + bin];
	const uint32_t nlevel = tc + 2 * qc;
	const int bin_x = (bin % c.bin_count_x) * kBinSize, bin_y = (bin / c.bin_count_x) * kBinSize;

	__shared__ uint32_t s_row_n;
	__shared__ uint32_t s_blk_n[4];
	__shared__ uint32_t s_base;
	__shared__ uint32_t s_blk_idx[4][kHighCap];
	__shared__ uint64_t s_key[4][kHighCap];
	__shared__ uint32_t s_tri[4][kHighCap];
	__shared__ uint32_t s_prefix[4][kHighCap + 1];
	__shared__ uint32_t s_own[4][64];
	__shared__ uint64_t s_pm[4][64];
	__shared__ uint32_t s_stats[2];
	if(threadIdx.x < 4)
		s_blk_n[threadIdx.x] = 0;
	if(threadIdx.x < 2)
		s_stats[threadIdx.x] = 0;
	if(threadIdx.x == 0) {
		s_row_n = 0;
		uint32_t base = atomicAdd(&b.ctrl[kCtrlRecordCursor], nlevel);
		if((uint64_t)base + nlevel > b.high_records_cap) {
			atomicOr(&b.ctrl[kCtrlOverflow], 2u);
			base = 0xffffffffu;
		}
		s_base = base;
	}
	__syncthreads();
	if(s_base == 0xffffffffu)
		return;
	uint4 *recs = b.high_records + s_base;
	const int band0 = by * 8;
	// --- records of this 8-row band with the 4-row rule (raster_high.glsl:54-90)
	for(uint32_t i = threadIdx.x; i < nlevel; i += kRasterThreads) {
		uint32_t t = bin_tri_at(b, i, qc, qo, to);
		if(t == 0xffffffffu)
			continue;
		uint4 s0 = b.tri_scan[2 * (size_t)t], s1 = b.tri_scan[2 * (size_t)t + 1];
		int min_g = min(max((int)(s0.w & 0xffffu) - bin_y, 0), kBinMask) >> 2;
		int max_g = min(max((int)(s0.w >> 16) - bin_y, 0), kBinMask) >> 2;
		int row_lo = min_g * 4, row_hi = max_g * 4 + 3;
		if(row_hi < band0 || row_lo > band0 + 7)
			continue;
		RowScan rs;
		row_scan_init(rs, s0, s1, (float)bin_x, (float)(bin_y + row_lo));
		uint32_t sp[8], bxm = 0;
#pragma unroll
		for(int r = 0; r < 8; r++)
			sp[r] = kBinMask; // empty (31, 0)
		int last = min(row_hi, band0 + 7);
		for(int row = row_lo; row <= last; row++) {
			uint32_t v = row_scan_step(rs);
			if(row >= band0) {
				int r = row - band0;
#pragma unroll
				for(int k = 0; k < 8; k++)
					if(k == r)
						sp[k] = v;
				bxm |= span_bx_mask(v);
			}
		}
		if(!bxm)
			continue;
		uint32_t idx = atomicAdd(&s_row_n, 1u);
		recs[idx] = make_record(t, sp, bxm);
		for(int bx = 0; bx < 4; bx++)
			if(bxm & (1u << bx)) {
				uint32_t k = atomicAdd(&s_blk_n[bx], 1u);
				if(k < (uint32_t)kHighCap)
					s_blk_idx[bx][k] = idx;
			}
	}
	__syncthreads();
	const int w = threadIdx.x >> 6;
	BlockRef br;
	br.bin = bin, br.by = by, br.bx = w;
	br.px0 = bin_x + w * 8, br.py0 = bin_y + band0;
	br.recs = recs;
	br.rec_offset = s_base;
	br.nrec = s_row_n;
	br.level = LUCID_BIN_LEVEL_HIGH;
	uint32_t m = s_blk_n[w];
	uint32_t frags = 0;
	if(m > (uint32_t)kHighCap) {
		// too many triangles for LDS: every pixel of the block takes the exact per-pixel path
		if(lane_id() == 0)
			enqueue_fallback(b, br, ~0ull, true);
	} else {
		WaveLds L{s_key[w], s_tri[w], s_prefix[w], s_own[w], s_pm[w]};
		frags = raster_block<FP32, 4>(c, b, L, br, nullptr, s_blk_idx[w], m);
	}
	if(lane_id() == 0) {
		atomicAdd(&s_stats[0], frags);
		atomicAdd(&s_stats[1], m);
	}
	__syncthreads();
	if(threadIdx.x == 0) {
		lucid_info *inf = (lucid_info *)b.info;
		atomicAdd(&inf->stats[0], s_stats[0]);
		atomicAdd(&inf->stats[1], s_stats[1]);
	}
}

// ------------------------------------------------------------------ exact fallback

// Collects the keys of one pixel's fragments into `dst` (LDS or global), returns the count.
__device__ uint32_t gather_pixel_keys(const FrameBufs &b, const uint4 *recs, uint32_t nrec, int row,
									  int xb, float px, float py, uint64_t *dst, uint32_t cap,
									  bool global_dst) {
	const int lane = lane_id();
	uint32_t n = 0;
	for(uint32_t base = 0; base < nrec; base += 64) {
		uint32_t i = base + lane;
		bool hit = false;
		uint64_t key = 0;
		if(i < nrec) {
			uint4 rec = recs[i];
			uint32_t sp = record_span(rec, row);
			int a = (int)(sp & 31u), e = (int)((sp >> 5) & 31u);
			hit = a <= xb && xb <= e;
			if(hit)
				key = pixel_key(sample_inv_depth(b, rec.x, px, py), rec.x);
		}
		uint64_t bal = ballot(hit);
		if(hit) {
			uint32_t pos = n + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
			if(pos < cap) {
				if(global_dst)
					__hip_atomic_store(&dst[pos], key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				else
					dst[pos] = key;
			}
		}
		n += (uint32_t)__popcll(bal);
	}
	return n;
}

template <bool FP32>
__device__ void blend_sorted(const FrameConst &c, const FrameBufs &b, const uint64_t *keys,
							 uint32_t n, float px, float py, bool global_src, float &cr, float &cg,
							 float &cb, float &tr) {
	const int lane = lane_id();
	const bool additive = (c.opts & (1u << 5)) != 0;
	cr = cg = cb = 0.0f, tr = 1.0f;
	for(uint32_t base = 0; base < n; base += 64) {
		uint32_t i = base + lane;
		float4 col = make_float4(0, 0, 0, 0);
		bool contributes = false;
		if(i < n) {
			uint64_t k = global_src
							 ? __hip_atomic_load(&keys[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
							 : keys[i];
			uint32_t t = (uint32_t)k;
			float inv = sample_inv_depth(b, t, px, py);
			Shaded sh = shade_sample<FP32>(c, b, t, px, py, inv);
			col = sh.col;
			contributes = FP32 ? sh.col.w != 0.0f : sh.enc != 0;
		}
		uint32_t cnt = min(64u, n - base);
		for(uint32_t j = 0; j < cnt; j++) {
			float4 c2;
			c2.x = __shfl(col.x, j, 64);
			c2.y = __shfl(col.y, j, 64);
			c2.z = __shfl(col.z, j, 64);
			c2.w = __shfl(col.w, j, 64);
			int ok = __shfl((int)contributes, j, 64);
			if(!ok)
				continue;
			if(additive) {
				cr = cr + c2.x * c2.w, cg = cg + c2.y * c2.w, cb = cb + c2.z * c2.w;
			} else {
				cr = cr + (c2.x * c2.w) * tr;
				cg = cg + (c2.y * c2.w) * tr;
				cb = cb + (c2.z * c2.w) * tr;
				tr = tr * (1.0f - c2.w);
			}
		}
	}
}

template <bool FP32>
__global__ __launch_bounds__(kRasterThreads) void fallback_kernel(FrameConst c, FrameBufs b) {
	uint32_t nq = min(b.ctrl[kCtrlFallbackCount], b.fallback_cap);
	if(blockIdx.x >= nq)
		return;
	__shared__ uint64_t s_keys[4][kFallbackCap];
	uint4 q0 = b.fallback_queue[2 * blockIdx.x], q1 = b.fallback_queue[2 * blockIdx.x + 1];
	int blk = (int)(q0.y & 0xffu), level = (int)((q0.y >> 8) & 0xffu);
	bool whole = (q0.y >> 16) & 1u;
	int bx = blk & 3;
	const uint4 *recs = (level == LUCID_BIN_LEVEL_LOW ? b.low_records : b.high_records) + q0.z;
	uint32_t nrec = q0.w;
	uint64_t mask = (uint64_t)q1.x | ((uint64_t)q1.y << 32);
	int px0 = (int)q1.z, py0 = (int)q1.w;
	const int w = threadIdx.x >> 6, lane = lane_id();
	uint32_t pixels = 0, frags = 0;
	int rank = 0;
	for(int p = 0; p < 64; p++) {
		if(!((mask >> p) & 1ull))
			continue;
		if((rank++ & 3) != w)
			continue;
		int px = px0 + (p & 7), py = py0 + (p >> 3);
		if(px >= c.width || py >= c.height)
			continue;
		int row = p >> 3, xb = bx * 8 + (p & 7);
		float fx = (float)px, fy = (float)py;
		uint64_t *keys = s_keys[w];
		uint32_t n = gather_pixel_keys(b, recs, nrec, row, xb, fx, fy, keys, kFallbackCap, false);
		float cr, cg, cb, tr;
		if(n <= (uint32_t)kFallbackCap) {
			uint32_t P = next_pow2(n < 2 ? 2 : n);
			for(uint32_t i = n + lane; i < P; i += 64)
				keys[i] = ~0ull;
			wave_sync_lds();
			bitonic_lds(keys, P);
			blend_sorted<FP32>(c, b, keys, n, fx, fy, false, cr, cg, cb, tr);
		} else {
			// deep pixel: spill its keys to global memory
			uint32_t P = next_pow2(n);
			uint32_t off = 0;
			if(lane == 0) {
				off = (uint32_t)atomicAdd(&b.ctrl[kCtrlSpillCursor], P);
				if((uint64_t)off + P > b.spill_cap) {
					atomicOr(&b.ctrl[kCtrlOverflow], 8u);
					off = 0xffffffffu;
				}
			}
			off = __shfl(off, 0, 64);
			if(off == 0xffffffffu)
				continue;
			uint64_t *g = b.spill + off;
			gather_pixel_keys(b, recs, nrec, row, xb, fx, fy, g, P, true);
			for(uint32_t i = n + lane; i < P; i += 64)
				__hip_atomic_store(&g[i], ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			__threadfence();
			__builtin_amdgcn_wave_barrier();
			bitonic_global(g, P);
			blend_sorted<FP32>(c, b, g, n, fx, fy, true, cr, cg, cb, tr);
		}
		if(lane == 0)
			write_pixel(c, b, px, py, cr, cg, cb, tr, n);
		pixels++;
		frags += n;
	}
	if(lane == 0 && pixels) {
		lucid_info *info = (lucid_info *)b.info;
		atomicAdd(&b.ctrl[kCtrlFallbackPixels], pixels);
		// stats[2]: pixels resolved by the exact per-pixel sort (the reference counts the
		// pixels its 3-deep window gets wrong there, shading.glsl:258-265)
		atomicAdd(&info->stats[2], pixels);
		if(whole) {
			atomicAdd(&info->stats[0], frags);
			if(w == 0)
				atomicAdd(&info->stats[3], 1u); // stats[3]: blocks too large for LDS
		}
	}
}

// ------------------------------------------------------------------ launch

template <bool FP32>
static void launch_raster_t(const FrameConst &c, const FrameBufs &b, hipStream_t stream,
							hipEvent_t *ev) {
	hipLaunchKernelGGL(raster_low_kernel<FP32>, dim3(c.bin_count), dim3(kRasterThreads), 0, stream,
					   c, b);
	if(ev)
		(void)hipEventRecord(ev[0], stream);
	hipLaunchKernelGGL(raster_high_kernel<FP32>, dim3(c.bin_count * 4), dim3(kRasterThreads), 0,
					   stream, c, b);
	if(ev)
		(void)hipEventRecord(ev[1], stream);
	hipLaunchKernelGGL(fallback_kernel<FP32>, dim3(b.fallback_cap), dim3(kRasterThreads), 0, stream,
					   c, b);
	if(ev)
		(void)hipEventRecord(ev[2], stream);
}

// events (optional): recorded after raster_low, raster_high and the fallback kernel.
hipError_t launch_raster(const FrameConst &c, const FrameBufs &b, hipStream_t stream,
						 hipEvent_t *events) {
	if(c.opts & (1u << 16))
		launch_raster_t<true>(c, b, stream, events);
	else
		launch_raster_t<false>(c, b, stream, events);
	return hipGetLastError();
}

} // namespace lucid
