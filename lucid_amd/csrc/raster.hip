// Tile rasterisation, per-pixel depth ordering and front-to-back blending
// (≙ data/shaders/raster_low.glsl, raster_high.glsl, shared/raster.glsl, shared/shading.glsl).
//
// Kernels (wave64; an 8x8 block of pixels is one wave, one lane per pixel):
//   rowgen_low   one work-group per bin with fewer than 1024 triangles: block-row records with
//                the reference's 8-row scanline rule and per-block index lists; promotes the bin
//                to the high path when an 8x8 block holds more than 256 triangles
//                (raster_low.glsl:39-106,229-237).
//   rowgen_high  one work-group per (high bin, 8-row band) with the 4-row rule of
//                raster_high.glsl:54-90.
//   high_prep / high_finalize  allocation for high bins; blocks with more than 256 triangles
//                are sorted by centroid depth (LDS bitonic, whole work-group) and split into
//                pixel-row parts so that several waves share a dense block.
//   block_raster one wave per block: sorts short lists itself in registers, expands coverage
//                masks into dense 64-sample chunks, shades, routes samples to pixel lanes and
//                blends front to back (block_raster_big: one wave per part of a sorted block).
//   fallback     exact per-pixel sort for pixels whose speculative order was violated.
//
// Exact ordering: triangles of a block are pre-sorted by the depth at their covered-pixel
// centroid. Each pixel lane consumes its fragments in that order through a K-entry insertion
// window keyed by (depth at the pixel, triangle id) and checks every arrival against the last
// emitted key. If no arrival is ever smaller than the last emitted key, the emitted sequence is
// exactly the per-pixel sorted order; otherwise the pixel is queued for the fallback kernel,
// which sorts its full fragment list (LDS bitonic, global-memory spill for deep pixels).
// Ties in depth are ordered by triangle id.
#include "lucid_device.h"

namespace lucid {

constexpr int kRowgenThreads = 256;
constexpr int kSortThreads = 256;
constexpr int kFallbackThreads = 256;
constexpr int kWindow = 4;
// Narrow ordering windows (tuning window 1): one slot in the register-sorted blocks and in
// the pre-sorted ones (whose depth ranks keep 32 - log2 of the padded list length bits).
constexpr int kNarrowWindow = 1, kNarrowBigWindow = 1;
constexpr int kFallbackCap = 512; // fragments per pixel sorted in LDS by the fallback kernel

// ------------------------------------------------------------------ shading

__device__ __forceinline__ float srgb_to_linear(float x) {
	return x < 0.04045f ? (1.0f / 12.92f) * x : glsl_pow((x + 0.055f) * (1.0f / 1.055f), 2.4f);
}
__device__ __forceinline__ float linear_to_srgb(float x) {
	return x < 0.0031308f ? 12.92f * x : 1.055f * glsl_pow(x, 1.0f / 2.4f) - 0.055f;
}

__device__ __forceinline__ float sample_inv_depth(const FrameBufs &b, uint32_t t, float px,
												  float py) {
	uint4 d = b.tri_depth[tri_rec(b, t)];
	return bits2f(d.x) * px + (bits2f(d.y) * py + bits2f(d.z));
}

// Shaded sample colour: RGBA8 word in the reference's quantised mode (0 = contributes
// nothing, shading.glsl:167-168,274) or fp32 with LUCID_OPT_FP32_SAMPLES.
template <bool FP32> struct SampleColor;
template <> struct SampleColor<false> {
	uint32_t enc;
	// component-wise select (keeps struct selects in registers)
	static __device__ __forceinline__ SampleColor sel(bool p, SampleColor a, SampleColor b) {
		SampleColor o;
		o.enc = p ? a.enc : b.enc;
		return o;
	}
	__device__ __forceinline__ bool contributes() const { return enc != 0; }
	__device__ __forceinline__ float4 rgba() const { return decode_rgba8(enc); }
	__device__ __forceinline__ void zero() { enc = 0; }
	__device__ __forceinline__ SampleColor shfl(int src) const {
		SampleColor o;
		o.enc = __shfl(enc, src, 64);
		return o;
	}
};
template <> struct SampleColor<true> {
	float4 col;
	static __device__ __forceinline__ SampleColor sel(bool p, SampleColor a, SampleColor b) {
		SampleColor o;
		o.col.x = p ? a.col.x : b.col.x;
		o.col.y = p ? a.col.y : b.col.y;
		o.col.z = p ? a.col.z : b.col.z;
		o.col.w = p ? a.col.w : b.col.w;
		return o;
	}
	__device__ __forceinline__ bool contributes() const { return col.w != 0.0f; }
	__device__ __forceinline__ float4 rgba() const { return col; }
	__device__ __forceinline__ void zero() { col = make_float4(0, 0, 0, 0); }
	__device__ __forceinline__ SampleColor shfl(int src) const {
		SampleColor o;
		o.col.x = __shfl(col.x, src, 64), o.col.y = __shfl(col.y, src, 64);
		o.col.z = __shfl(col.z, src, 64), o.col.w = __shfl(col.w, src, 64);
		return o;
	}
};

// sRGB round trip of finalShading, s = L2S(S2L(x) * f), unclamped. fp = f^(1/2.4).
// Above both linear segments the two powers cancel: L2S(S2L(x) f) = (x + 0.055) fp - 0.055
// exactly (the threshold S2L(x) f >= 0.0031308 is (x + 0.055) fp >= kSrgbHiU); below both,
// s = 12.92 ((x / 12.92) f) with the reference's own operations. Only lanes in neither
// case (or in a NaN case) evaluate the powers, in a branch the wave skips when none is.
constexpr float kSrgbHiU = 0.09544991f; // 1.055 * 0.0031308^(1/2.4)
struct LitChannel {
	float s;     // L2S(S2L(x) f) before the clamp
	float deriv; // ds/dx
};
template <bool DERIV>
__device__ __forceinline__ LitChannel lit_channel(float x, float f, float fp) {
	const float u = (x + 0.055f) * fp;
	const float lin_lo = (1.0f / 12.92f) * x;
	const bool hi = x >= 0.04045f && u >= kSrgbHiU;
	const bool lo = x < 0.04045f && lin_lo * f < 0.0031308f;
	LitChannel o;
	o.s = hi ? u - 0.055f : 12.92f * (lin_lo * f);
	o.deriv = hi ? fp : f;
	if(!hi && !lo) {
		const float lin = srgb_to_linear(x);
		const float y = lin * f;
		o.s = linear_to_srgb(y);
		if constexpr(DERIV) {
			// S2L'(x) = 2.4 S2L(x) / (x + 0.055), L2S'(y) = (L2S(y) + 0.055) / (2.4 y) above
			// their linear segments; no further powers
			const float d_lin = x < 0.04045f
									? (1.0f / 12.92f)
									: 2.4f * lin * __builtin_amdgcn_rcpf(x + 0.055f);
			const float d_srgb = y < 0.0031308f
									 ? 12.92f
									 : (o.s + 0.055f) * __builtin_amdgcn_rcpf(2.4f * y);
			o.deriv = d_srgb * f * d_lin;
		}
	}
	return o;
}

// ------------------------------------------------------------------ albedo textures
//
// textureGrad (shading.glsl:145-150) on an RGBA8 atlas with the reference's sampler
// (src/lucid_base.h:30: linear magnification, minification and mip filter, repeat
// addressing), restated from the Vulkan sampling rules: the scale factors are the texture
// coordinate derivatives times the level-0 size, lambda = log2(max(|d/dx|, |d/dy|))
// (isotropic), the level is clamped to [0, levels - 1] and the two nearest levels are
// blended by its fraction; within a level, bilinear weights at u = s * size - 0.5 with the
// texel indices wrapped (repeat). s is reduced to [0, 1] first (repeat is periodic).

// First texel of `level`: the offsets are kernel arguments, picked with selects (no dynamic
// indexing of an argument array).
__device__ __forceinline__ uint32_t tex_level_offset(const TexDesc &t, int level) {
	uint32_t off = 0;
#pragma unroll
	for(int i = 1; i < kMaxTexLevels; i++)
		off = level == i ? t.offset[i] : off;
	return off;
}

__device__ __forceinline__ float4 tex_bilinear(const TexDesc &t, int level, float s, float v) {
	const int w = max(1, t.width >> level), h = max(1, t.height >> level);
	const uint32_t *texels = t.texels + tex_level_offset(t, level);
	const float su = s - floorf(s), sv = v - floorf(v);
	const float u = su * (float)w - 0.5f, uv = sv * (float)h - 0.5f;
	const float fu = floorf(u), fv = floorf(uv);
	const float au = u - fu, av = uv - fv;
	int x0 = (int)fu, y0 = (int)fv; // in [-1, size - 1]
	x0 = x0 < 0 ? w - 1 : x0;
	y0 = y0 < 0 ? h - 1 : y0;
	const int x1 = x0 + 1 == w ? 0 : x0 + 1, y1 = y0 + 1 == h ? 0 : y0 + 1;
	const float4 c00 = decode_rgba8(texels[(uint32_t)y0 * (uint32_t)w + (uint32_t)x0]);
	const float4 c10 = decode_rgba8(texels[(uint32_t)y0 * (uint32_t)w + (uint32_t)x1]);
	const float4 c01 = decode_rgba8(texels[(uint32_t)y1 * (uint32_t)w + (uint32_t)x0]);
	const float4 c11 = decode_rgba8(texels[(uint32_t)y1 * (uint32_t)w + (uint32_t)x1]);
	const float bu = 1.0f - au, bv = 1.0f - av;
	float4 r;
	r.x = (c00.x * bu + c10.x * au) * bv + (c01.x * bu + c11.x * au) * av;
	r.y = (c00.y * bu + c10.y * au) * bv + (c01.y * bu + c11.y * au) * av;
	r.z = (c00.z * bu + c10.z * au) * bv + (c01.z * bu + c11.z * au) * av;
	r.w = (c00.w * bu + c10.w * au) * bv + (c01.w * bu + c11.w * au) * av;
	return r;
}

__device__ float4 texture_grad(const TexDesc &t, float s, float v, float dsx, float dvx, float dsy,
							   float dvy) {
	const float fw = (float)t.width, fh = (float)t.height;
	const float ux = dsx * fw, vx = dvx * fh, uy = dsy * fw, vy = dvy * fh;
	const float rho = fmx(sqrtf(ux * ux + vx * vx), sqrtf(uy * uy + vy * vy));
	const float lambda = log2f(rho);
	if(!(lambda > 0.0f)) // magnification (or a degenerate footprint): level 0
		return tex_bilinear(t, 0, s, v);
	const float d = fmn(lambda, (float)(t.levels - 1));
	const float fl = floorf(d);
	const int lo = (int)fl;
	const float delta = d - fl;
	float4 r = tex_bilinear(t, lo, s, v);
	if(delta > 0.0f) {
		const float4 h = tex_bilinear(t, min(lo + 1, t.levels - 1), s, v);
		const float nd = 1.0f - delta;
		r.x = r.x * nd + h.x * delta;
		r.y = r.y * nd + h.y * delta;
		r.z = r.z * nd + h.z * delta;
		r.w = r.w * nd + h.w * delta;
	}
	return r;
}

// Albedo of a textured triangle at a pixel (shading.glsl:111-122,131-159): barycentric
// derivatives from the planes one pixel right and one pixel down, the texture coordinate
// and its derivatives from the triangle's corner coordinates, the instance's uv rectangle,
// then the opaque (alpha 1) or transparent atlas. bary0/1: before the offsets are removed.
__device__ __forceinline__ float4 albedo_sample(const FrameBufs &b, const uint4 *q, float inv_ray,
												float e0, float e1, float bary0, float bary1,
												float bp0, float bp1) {
	const uint32_t flags = q[0].w & 0xffffu, inst = q[0].w >> 16, t = q[6].z;
	const float dex = bits2f(q[0].x), dey = bits2f(q[0].y);
	const float ray_posx = __builtin_amdgcn_rcpf(inv_ray + dex);
	const float ray_posy = __builtin_amdgcn_rcpf(inv_ray + dey);
	const float bdx0 = (e0 + bits2f(q[1].x)) * ray_posx - bary0;
	const float bdx1 = (e1 + bits2f(q[2].x)) * ray_posx - bary1;
	const float bdy0 = (e0 + bits2f(q[1].y)) * ray_posy - bary0;
	const float bdy1 = (e1 + bits2f(q[2].y)) * ray_posy - bary1;
	const float b0 = bary0 - bp0, b1 = bary1 - bp1;
	const uint4 *tt = b.tri_tex + 2 * (size_t)tri_rec(b, t);
	const uint4 ta = tt[0], tb = tt[1];
	const float t0x = bits2f(ta.x), t0y = bits2f(ta.y), t1x = bits2f(ta.z), t1y = bits2f(ta.w);
	const float t2x = bits2f(tb.x), t2y = bits2f(tb.y);
	float s = b0 * t1x + (b1 * t2x + t0x), v = b0 * t1y + (b1 * t2y + t0y);
	float dsx = bdx0 * t1x + bdx1 * t2x, dvx = bdx0 * t1y + bdx1 * t2y;
	float dsy = bdy0 * t1x + bdy1 * t2x, dvy = bdy0 * t1y + bdy1 * t2y;
	if(flags & LUCID_INST_HAS_UV_RECT) {
		const float4 rc = b.inst_uv_rects[inst];
		s = rc.z * (s - floorf(s)) + rc.x;
		v = rc.w * (v - floorf(v)) + rc.y;
		dsx = dsx * rc.z, dsy = dsy * rc.z;
		dvx = dvx * rc.w, dvy = dvy * rc.w;
	}
	const bool opaque = (flags & LUCID_INST_TEX_OPAQUE) != 0;
	float4 col = texture_grad(opaque ? b.tex[0] : b.tex[1], s, v, dsx, dvx, dsy, dvy);
	col.w = opaque ? 1.0f : col.w;
	return col;
}

// The shading record of triangle t, built by setup (one 128-byte line per triangle).
__device__ __forceinline__ void load_shading_record(const FrameBufs &b, uint32_t t,
													uint4 (&q)[kRecQuads]) {
	const uint4 *src = b.tri_shade + (size_t)kRecStride * tri_rec(b, t);
#pragma unroll
	for(int j = 0; j < kRecQuads; j++)
		q[j] = src[j];
}

// Depth plane at a pixel: inverse distance along the pixel's ray (shading.glsl:108).
__device__ __forceinline__ float record_inv_depth(uint4 q0, float px, float py) {
	return bits2f(q0.x) * px + (bits2f(q0.y) * py + bits2f(q0.z));
}

// Everything shadeSample (shading.glsl:107-184, no textures) computes before lighting,
// from a shading record (LDS or registers): barycentrics, the vertex colour times the
// instance colour, and the per-channel lighting factors. `inv_ray` is record_inv_depth at
// the same pixel. Returns false when the sample's alpha is 0 (it contributes nothing).
struct ShadeState {
	float bary0, bary1;
	float4 color; // interpolated vertex colour (white without one) times the instance colour
	float4 inst;  // instance colour (white without one)
	float f[3];   // lighting factors
	float fp[3];  // f^(1/2.4)
};

template <bool TEX>
__device__ __forceinline__ bool shade_prepare(const FrameConst &c, const FrameBufs &b,
											  const uint4 *q, float px, float py, float inv_ray,
											  ShadeState &st) {
	const uint32_t flags = q[0].w & 0xffffu;
	const uint4 b0 = q[1], b1 = q[2];
	// GLSL's division is specified to 2.5 ulp; v_rcp_f32 (1 ulp) meets that
	float ray_pos = __builtin_amdgcn_rcpf(inv_ray);
	float e0 = bits2f(b0.x) * px + (bits2f(b0.y) * py + bits2f(b0.z));
	float e1 = bits2f(b1.x) * px + (bits2f(b1.y) * py + bits2f(b1.z));
	float bary0 = e0 * ray_pos, bary1 = e1 * ray_pos;
	float4 tex = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
	bool textured = false;
	if constexpr(TEX) {
		if(flags & LUCID_INST_HAS_ALBEDO_TEXTURE) {
			textured = true;
			tex = albedo_sample(b, q, inv_ray, e0, e1, bary0, bary1, bits2f(b0.w), bits2f(b1.w));
		}
	}
	bary0 = bary0 - bits2f(b0.w);
	bary1 = bary1 - bits2f(b1.w);
	st.bary0 = bary0, st.bary1 = bary1;

	// vertex colour (or white), then times the instance colour; products by an exact 1.0
	// are skipped, which leaves every result bit-identical
	float4 color = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
	const uint4 q6 = q[6];
	if(flags & LUCID_INST_HAS_VERTEX_COLORS) {
		float w0 = (1.0f - bary0) - bary1;
		const uint4 q3 = q[3];
		if(c.color_mode == kColorRGBA8) {
			float4 c0 = decode_rgba8(q3.x), c1 = decode_rgba8(q3.y), c2 = decode_rgba8(q3.z);
			color.x = w0 * c0.x + (bary0 * c1.x + bary1 * c2.x);
			color.y = w0 * c0.y + (bary0 * c1.y + bary1 * c2.y);
			color.z = w0 * c0.z + (bary0 * c1.z + bary1 * c2.z);
			color.w = w0 * c0.w + (bary0 * c1.w + bary1 * c2.w);
		} else if(c.color_mode == kColorF32) {
			const uint4 q4 = q[4], q5 = q[5];
			color.x = w0 * bits2f(q3.x) + (bary0 * bits2f(q4.x) + bary1 * bits2f(q5.x));
			color.y = w0 * bits2f(q3.y) + (bary0 * bits2f(q4.y) + bary1 * bits2f(q5.y));
			color.z = w0 * bits2f(q3.z) + (bary0 * bits2f(q4.z) + bary1 * bits2f(q5.z));
			color.w = bits2f(q3.w);
		}
	}
	st.inst = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
	if(textured) {
		// shading.glsl: colour = instance colour, times the texture, times the vertex colour
		float4 f = tex;
		if(flags & LUCID_INST_HAS_COLOR) {
			const float4 ic = decode_rgba8(q6.x);
			st.inst = ic;
			f = make_float4(ic.x * tex.x, ic.y * tex.y, ic.z * tex.z, ic.w * tex.w);
		}
		color.x = f.x * color.x;
		color.y = f.y * color.y;
		color.z = f.z * color.z;
		color.w = f.w * color.w;
	} else if(flags & LUCID_INST_HAS_COLOR) {
		const float4 ic = decode_rgba8(q6.x);
		st.inst = ic;
		color.x = ic.x * color.x;
		color.y = ic.y * color.y;
		color.z = ic.z * color.z;
		color.w = ic.w * color.w;
	}
	st.color = color;
	if(color.w == 0.0f)
		return false;
	if(c.identity_light) { // every lighting factor is 1 (no normal or light value is read)
		st.f[0] = st.f[1] = st.f[2] = 1.0f;
		st.fp[0] = st.fp[1] = st.fp[2] = 1.0f;
		return true;
	}
	if(flags & LUCID_INST_HAS_VERTEX_NORMALS) {
		V3 n0 = decode_normal(q[4].w), n1 = decode_normal(q[5].w), n2 = decode_normal(q6.y);
		n1 = sub3(n1, n0);
		n2 = sub3(n2, n0);
		V3 normal = mk3(bary0 * n1.x + (bary1 * n2.x + n0.x), bary0 * n1.y + (bary1 * n2.y + n0.y),
						bary0 * n1.z + (bary1 * n2.z + n0.z));
		const float light = light_value(c, normal);
#pragma unroll
		for(int k = 0; k < 3; k++) {
			st.f[k] = light_factor(c, k, light);
			st.fp[k] = glsl_pow(st.f[k], 1.0f / 2.4f);
		}
	} else {
		const float light = bits2f(q6.w);
#pragma unroll
		for(int k = 0; k < 3; k++)
			st.f[k] = light_factor(c, k, light);
		st.fp[0] = bits2f(q[4].w), st.fp[1] = bits2f(q[5].w), st.fp[2] = bits2f(q6.y);
	}
	return true;
}

// shadeSample (shading.glsl:107-184) without textures: lighting through the sRGB round trip
// of finalShading (funcs.glsl:261-271), then RGBA8 quantisation unless FP32.
template <bool FP32, bool TEX>
__device__ __forceinline__ SampleColor<FP32> shade_record(const FrameConst &c, const FrameBufs &b,
														  const uint4 *q, float px, float py,
														  float inv_ray) {
	ShadeState st;
	SampleColor<FP32> out;
	if(!shade_prepare<TEX>(c, b, q, px, py, inv_ray, st)) {
		out.zero();
		return out;
	}
	float rgb[3] = {st.color.x, st.color.y, st.color.z};
	if(c.identity_light) {
		// L2S(S2L(x) * 1) = x: the sRGB round trip of a unit lighting factor is skipped
#pragma unroll
		for(int k = 0; k < 3; k++)
			rgb[k] = sat(rgb[k]);
	} else {
#pragma unroll
		for(int k = 0; k < 3; k++)
			rgb[k] = sat(lit_channel<false>(rgb[k], st.f[k], st.fp[k]).s);
	}
	if constexpr(FP32)
		out.col = make_float4(rgb[0], rgb[1], rgb[2], st.color.w);
	else
		out.enc = encode_rgba8_unit_rgb(rgb[0], rgb[1], rgb[2], st.color.w);
	return out;
}

// Shades triangle t at a pixel straight from the setup records (exact fallback path).
template <bool FP32, bool TEX>
__device__ __forceinline__ SampleColor<FP32> shade_sample(const FrameConst &c, const FrameBufs &b,
														  uint32_t t, float px, float py,
														  float inv_ray) {
	uint4 q[kRecQuads];
	load_shading_record(b, t, q);
	return shade_record<FP32, TEX>(c, b, q, px, py, inv_ray);
}

__device__ __forceinline__ uint64_t pixel_key(float inv_ray, uint32_t t) {
	return ((uint64_t)inv_ord_f32(inv_ray) << 32) | t;
}

// ------------------------------------------------------------------ per-pixel state

struct Accum {
	float cr, cg, cb, tr;
	__device__ __forceinline__ void init() { cr = cg = cb = 0.0f, tr = 1.0f; }
	// shading.glsl:276-283 (front-to-back "over"; additive variant)
	__device__ __forceinline__ void blend(float4 col, bool additive) {
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
	// The same for an RGBA8 sample (bytes c and a): the byte scales fold into one weight per
	// sample, so each channel takes one fused multiply-add, cr += c (a tr / 255) with
	// a = byte / 255 (the sums differ from blend()'s by rounding only; the transmittance
	// step is blend()'s, operation for operation).
	__device__ __forceinline__ void blend(SampleColor<false> s, bool additive) {
		const float a = (float)(s.enc >> 24) * (1.0f / 255.0f);
		const float w = (additive ? a : a * tr) * (1.0f / 255.0f);
		cr = fmaf((float)(s.enc & 0xffu), w, cr);
		cg = fmaf((float)((s.enc >> 8) & 0xffu), w, cg);
		cb = fmaf((float)((s.enc >> 16) & 0xffu), w, cb);
		if(!additive)
			tr = tr * (1.0f - a);
	}
	__device__ __forceinline__ void blend(SampleColor<true> s, bool additive) { blend(s.col, additive); }
};

// Sample order key: nearer (larger inverse depth) first, equal depths by triangle id.
constexpr uint64_t kEmptyKey = ~0ull;

// Per-pixel ordering window of kWindow pending samples (payload P) in fixed slots. A
// sample that arrives when the window is full pushes out the nearest of the five, which is
// handed to the emit callback, so samples leave front to back. The pixel is exact iff the
// emitted keys strictly increase; otherwise `bad` marks it for the exact per-pixel sort.
// Slots are updated in place by selects: no register rotation and no dynamic indexing.
// P provides zero() and a component-wise static sel(bool, P, P).
template <class P>
struct OrderWindow {
	uint64_t k0, k1, k2, k3;
	P c0, c1, c2, c3;
	uint32_t n;           // filled slots
	uint64_t first, last; // first and last emitted key (kEmptyKey / 0 before the first)
	bool bad;

	__device__ __forceinline__ void init() {
		k0 = k1 = k2 = k3 = kEmptyKey;
		c0.zero(), c1.zero(), c2.zero(), c3.zero();
		n = 0, first = kEmptyKey, last = 0, bad = false;
	}

	template <class F> __device__ __forceinline__ void emit(uint64_t k, const P &col, F &&f) {
		bad |= k <= last;
		first = first == kEmptyKey ? k : first;
		last = k;
		f(col);
	}

	template <class F> __device__ __forceinline__ void insert(uint64_t key, const P &col, F &&f) {
		static_assert(kWindow == 4, "window slots are spelled out");
		if(n < (uint32_t)kWindow) {
			k0 = n == 0 ? key : k0, c0 = P::sel(n == 0, col, c0);
			k1 = n == 1 ? key : k1, c1 = P::sel(n == 1, col, c1);
			k2 = n == 2 ? key : k2, c2 = P::sel(n == 2, col, c2);
			k3 = n == 3 ? key : k3, c3 = P::sel(n == 3, col, c3);
			n++;
			return;
		}
		// nearest pending sample by a two-level tournament
		const bool a = k1 < k0, bq = k3 < k2;
		const uint64_t ka = a ? k1 : k0, kb = bq ? k3 : k2;
		const P ca = P::sel(a, c1, c0), cb = P::sel(bq, c3, c2);
		const bool m = kb < ka;
		const uint64_t km = m ? kb : ka;
		const P cm = P::sel(m, cb, ca);
		const bool take_new = key < km;
		emit(take_new ? key : km, P::sel(take_new, col, cm), f);
		// the new sample takes the emitted slot
		const bool s0 = !take_new && !m && !a, s1 = !take_new && !m && a;
		const bool s2 = !take_new && m && !bq, s3 = !take_new && m && bq;
		k0 = s0 ? key : k0, c0 = P::sel(s0, col, c0);
		k1 = s1 ? key : k1, c1 = P::sel(s1, col, c1);
		k2 = s2 ? key : k2, c2 = P::sel(s2, col, c2);
		k3 = s3 ? key : k3, c3 = P::sel(s3, col, c3);
	}

	template <class F> __device__ __forceinline__ void flush(F &&f) {
		for(uint32_t r = 0; r < n; r++) {
			const bool a = k1 < k0, bq = k3 < k2;
			const uint64_t ka = a ? k1 : k0, kb = bq ? k3 : k2;
			const P ca = P::sel(a, c1, c0), cb = P::sel(bq, c3, c2);
			const bool m = kb < ka;
			emit(m ? kb : ka, P::sel(m, cb, ca), f);
			const bool s0 = !m && !a, s1 = !m && a, s2 = m && !bq, s3 = m && bq;
			k0 = s0 ? kEmptyKey : k0;
			k1 = s1 ? kEmptyKey : k1;
			k2 = s2 ? kEmptyKey : k2;
			k3 = s3 ? kEmptyKey : k3;
		}
		n = 0;
	}
};

// Smaller ordering windows (N = 1 or 2 pending samples), same contract as OrderWindow: for
// lists whose block order is already (nearly) each pixel's depth order, where the window
// only has to catch neighbours in the wrong order; a pixel it cannot put in order is still
// detected (bad) and sorted exactly.
template <class P, int N, bool PH = (N == 1)>
struct SmallWindow {
	static_assert(N == 1 || N == 2, "one or two slots");
	static_assert(N == 2 || PH, "the one-slot window keeps a placeholder");
	// PH: the slots hold placeholders from the start (key 0, below every sample key; a zero
	// payload), which the first samples push out, so no fill count is kept; the emit
	// callback must treat the zero payload as nothing (the forward blend does; the
	// backward's order recording does not, so it keeps the count). Placeholders always
	// leave before any sample (they are the smallest keys), and the keys of one pixel are
	// distinct (one sample per triangle), so `<` finds what `<=` does. After flush such a
	// window takes no more samples.
	uint64_t k0, k1;
	P c0, c1;
	uint32_t n;
	uint64_t first, last;
	uint32_t bad; // out-of-order emits (a count: one add with carry per sample)

	__device__ __forceinline__ void init() {
		k0 = PH ? 0ull : kEmptyKey;
		k1 = PH ? 0ull : kEmptyKey;
		c0.zero(), c1.zero();
		n = 0, first = kEmptyKey, last = 0, bad = 0;
	}
	template <class F> __device__ __forceinline__ void emit(uint64_t k, const P &col, F &&f) {
		if constexpr(PH) {
			bad += k < last ? 1u : 0u; // (first is kept by insert)
		} else {
			bad += k <= last ? 1u : 0u;
			first = first == kEmptyKey ? k : first;
		}
		last = k;
		f(col);
	}
	template <class F> __device__ __forceinline__ void insert(uint64_t key, const P &col, F &&f) {
		// with placeholders, first is the smallest key inserted: while the emitted keys
		// increase (an exact pixel) that is the first sample emitted, and a pixel whose keys
		// do not increase is resolved by the exact fallback, whatever first holds
		if constexpr(PH)
			first = key < first ? key : first;
		if constexpr(N == 1) {
			const bool take_new = key < k0;
			emit(take_new ? key : k0, P::sel(take_new, col, c0), f);
			k0 = take_new ? k0 : key, c0 = P::sel(take_new, c0, col);
		} else {
			if constexpr(!PH) {
				if(n < (uint32_t)N) {
					k0 = n == 0 ? key : k0, c0 = P::sel(n == 0, col, c0);
					k1 = n == 1 ? key : k1, c1 = P::sel(n == 1, col, c1);
					n++;
					return;
				}
			}
			const bool a = k1 < k0;
			const uint64_t km = a ? k1 : k0;
			const P cm = P::sel(a, c1, c0);
			const bool take_new = key < km;
			emit(take_new ? key : km, P::sel(take_new, col, cm), f);
			const bool s0 = !take_new && !a, s1 = !take_new && a;
			k0 = s0 ? key : k0, c0 = P::sel(s0, col, c0);
			k1 = s1 ? key : k1, c1 = P::sel(s1, col, c1);
		}
	}
	template <class F> __device__ __forceinline__ void flush(F &&f) {
		if constexpr(N == 1) {
			emit(k0, c0, f); // the last sample, or the placeholder if there was none
		} else {
			for(uint32_t r = 0; r < (PH ? 2u : n); r++) { // (placeholders leave first)
				const bool a = k1 < k0;
				emit(a ? k1 : k0, P::sel(a, c1, c0), f);
				k0 = a ? k0 : kEmptyKey;
				k1 = a ? kEmptyKey : k1;
			}
			n = 0;
		}
	}
};

// Forward per-pixel state: the ordering window (N slots) feeding the front-to-back blend.
template <bool FP32, int N = kWindow>
struct PixelState {
	std::conditional_t<N == kWindow, OrderWindow<SampleColor<FP32>>, SmallWindow<SampleColor<FP32>, N, true>> w;
	Accum acc;

	__device__ __forceinline__ void init() {
		w.init();
		acc.init();
	}
	__device__ __forceinline__ void insert(uint64_t key, SampleColor<FP32> col, bool additive) {
		w.insert(key, col, [&](const SampleColor<FP32> &s) { acc.blend(s, additive); });
	}
	__device__ __forceinline__ void flush(bool additive) {
		w.flush([&](const SampleColor<FP32> &s) { acc.blend(s, additive); });
	}
};

__device__ __forceinline__ void write_pixel(const FrameConst &c, const FrameBufs &b, int px,
											int py, const Accum &a, uint32_t count) {
	if(px >= c.width || py >= c.height)
		return;
	float r = sat(a.cr + a.tr * c.background[0]);
	float g = sat(a.cg + a.tr * c.background[1]);
	float bl = sat(a.cb + a.tr * c.background[2]);
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

// ------------------------------------------------------------------ sorts

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
	uint32_t lo = __shfl((uint32_t)v, src, 64), hi = __shfl((uint32_t)(v >> 32), src, 64);
	return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
	uint32_t lo = __shfl_xor((uint32_t)v, m, 64), hi = __shfl_xor((uint32_t)(v >> 32), m, 64);
	return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl_xor_key(uint64_t v, int m) { return shfl_xor64(v, m); }
__device__ __forceinline__ uint32_t shfl_xor_key(uint32_t v, int m) { return __shfl_xor(v, m, 64); }

// Ascending bitonic sort of 64*E keys held in registers; element i = e * 64 + lane.
// Bitonic network steps j = j_start .. 1 of merge level k on 64 * E keys (K: 32- or 64-bit)
// held by one wave (key e * 64 + lane of the block in v[e]); `base` is the block's index in
// the whole sequence, which decides each pair's direction: ascending iff (index & k) == 0.
template <int E, class K>
__device__ __forceinline__ void bitonic_regs_steps(K (&v)[E], uint32_t base, uint32_t k,
												   int j_start) {
	const int lane = lane_id();
#pragma unroll
	for(int j = 64 * E / 2; j > 0; j >>= 1) {
		if(j > j_start)
			continue;
		if(j >= 64) {
			const int jr = j / 64;
#pragma unroll
			for(int e = 0; e < E; e++) {
				if((e & jr) == 0) {
					const int pe = e | jr;
					const bool asc = ((base + (uint32_t)(e * 64 + lane)) & k) == 0;
					K a = v[e], bb = v[pe];
					bool sw = (a > bb) == asc;
					v[e] = sw ? bb : a;
					v[pe] = sw ? a : bb;
				}
			}
		} else {
#pragma unroll
			for(int e = 0; e < E; e++) {
				const uint32_t i = base + (uint32_t)(e * 64 + lane);
				K p = shfl_xor_key(v[e], j);
				bool asc = (i & k) == 0;
				bool lower = (lane & j) == 0;
				K mn = v[e] < p ? v[e] : p, mx = v[e] < p ? p : v[e];
				v[e] = (asc == lower) ? mn : mx;
			}
		}
	}
}

// Bitonic sort levels k = 2 .. 64 * E of a 64 * E block starting at index `base` of a longer
// sequence (ascending for base = 0 alone; alternating directions per block otherwise, as the
// network's later levels expect).
template <int E, class K>
__device__ __forceinline__ void bitonic_regs_from(K (&v)[E], uint32_t base) {
#pragma unroll
	for(int k = 2; k <= 64 * E; k <<= 1)
		bitonic_regs_steps<E>(v, base, (uint32_t)k, k >> 1);
}

// Ascending bitonic sort of 64 * E keys held by one wave (key e * 64 + lane in v[e]).
template <int E, class K>
__device__ __forceinline__ void bitonic_regs(K (&v)[E]) {
	bitonic_regs_from<E>(v, 0u);
}

// Ascending bitonic sort of P (power of two) keys in LDS by one wave.
__device__ void bitonic_lds_wave(uint64_t *key, uint32_t P) {
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
__device__ void bitonic_global_wave(uint64_t *key, uint32_t P) {
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

// ------------------------------------------------------------------ records and keys

__device__ __forceinline__ uint4 make_record(uint32_t t, const uint32_t sp[8], uint32_t bxm) {
	return make_uint4(t, sp[0] | (sp[1] << 10) | (sp[2] << 20), sp[3] | (sp[4] << 10) | (sp[5] << 20),
					  sp[6] | (sp[7] << 10) | (bxm << 20));
}

// Inverse depth of triangle t inside a block at the centroid of its covered pixels, from
// its pixel mask m (bit 8 r + c: row r, column c of the block whose corner pixel is
// (px0, py0)): the block-level sort key (ties and centroid error are resolved exactly per
// pixel later). The column and row sums are integers, taken bit by bit with masked
// popcounts: the sum of the columns is sum_k 2^k popcount(m & pixels whose column has bit
// k), likewise for the rows.
__device__ __forceinline__ float mask_inv_depth(const FrameBufs &b, uint32_t t, uint64_t m,
												int px0, int py0) {
	auto pc = [](uint64_t x) { return (int)__popcll(x); };
	const int n = pc(m);
	const int sc = pc(m & 0xAAAAAAAAAAAAAAAAull) + 2 * pc(m & 0xCCCCCCCCCCCCCCCCull) +
				   4 * pc(m & 0xF0F0F0F0F0F0F0F0ull);
	const int sr = pc(m & 0xFF00FF00FF00FF00ull) + 2 * pc(m & 0xFFFF0000FFFF0000ull) +
				   4 * pc(m & 0xFFFFFFFF00000000ull);
	const float inv_n = 1.0f / (float)(n > 0 ? n : 1);
	const float cx = (float)px0 + (float)(2 * sc) * 0.5f * inv_n;
	const float cy = (float)py0 + (float)sr * inv_n;
	return sample_inv_depth(b, t, cx, cy);
}

// Depth rank of a block entry from its depth key ~ord(inverse depth) (ascending = nearest
// first): the key minus the block's smallest, shifted right just enough that the block's key
// range fits BITS bits. Monotone in the key, and float keys keep relative precision (about
// BITS - log2(binades spanned) mantissa bits), so ties are rare and are resolved per pixel.
struct DepthGrid {
	uint32_t lo;
	int shift;
	__device__ __forceinline__ DepthGrid(uint32_t lo_, uint32_t hi_, int bits) : lo(lo_) {
		const uint32_t span = hi_ - lo_;
		const int need = span ? 32 - __clz(span) : 0;
		shift = need > bits ? need - bits : 0;
	}
	__device__ __forceinline__ uint32_t rank(uint32_t key) const { return (key - lo) >> shift; }
};

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
	for(int m = 1; m < 64; m <<= 1)
		v = min(v, (uint32_t)__shfl_xor(v, m, 64));
	return v;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
	for(int m = 1; m < 64; m <<= 1)
		v = max(v, (uint32_t)__shfl_xor(v, m, 64));
	return v;
}

// Depth order of a block list of m <= 256 entries, sorted by one wave in registers: sorted
// entry 64 k + lane is list position pos[k]. Keys are (24-bit depth rank << 8 | position).
// With s_msk / s_tri (LDS, kRegSortCap entries each, or null), each entry's pixel mask and
// triangle id are also left there at its list position, for the caller to read in sorted
// order (after a wave_sync_lds) instead of gathering the records again.
__device__ __forceinline__ void sort_block_positions(const FrameBufs &b, const uint4 *recs,
													 const uint32_t *list, uint32_t m, int bx,
													 int px0, int py0, uint32_t (&pos)[4],
													 uint64_t *s_msk = nullptr,
													 uint32_t *s_tri = nullptr) {
	static_assert(kRegSortCap <= 256, "list positions fit 8 bits of the key");
	const int lane = lane_id();
	uint32_t dk[4];
	uint32_t lo = ~0u, hi = 0u;
#pragma unroll
	for(int k = 0; k < 4; k++) {
		const uint32_t j = k * 64 + lane;
		dk[k] = 0u;
		if(j < m) {
			const uint4 rec = recs[list[j]];
			const uint64_t msk = record_block_mask(rec, bx);
			if(s_msk) {
				s_msk[j] = msk;
				s_tri[j] = rec.x;
			}
			dk[k] = inv_ord_f32(mask_inv_depth(b, rec.x, msk, px0, py0));
			lo = min(lo, dk[k]);
			hi = max(hi, dk[k]);
		}
	}
	const DepthGrid grid(wave_min_u32(lo), wave_max_u32(hi), 24);
	// a real key never equals the padding: position 255 exists only when nothing is padded
	auto key = [&](int k) -> uint32_t {
		const uint32_t j = k * 64 + lane;
		return j < m ? (grid.rank(dk[k]) << 8) | j : ~0u;
	};
	// One key array for the three sort widths (copied in and out of the narrower sorts), so
	// that it stays in registers: positions of empty rows are never read.
	uint32_t w[4] = {key(0), key(1), key(2), key(3)};
	if(m <= 64) {
		uint32_t w1[1] = {w[0]};
		bitonic_regs<1>(w1);
		w[0] = w1[0];
	} else if(m <= 128) {
		uint32_t w2[2] = {w[0], w[1]};
		bitonic_regs<2>(w2);
		w[0] = w2[0], w[1] = w2[1];
	} else {
		bitonic_regs<4>(w);
	}
#pragma unroll
	for(int k = 0; k < 4; k++)
		pos[k] = w[k] & 0xffu;
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

struct BinInfo {
	uint32_t qc, tc, qo, to, nlevel;
	int bin_x, bin_y;
};

__device__ __forceinline__ BinInfo load_bin(const FrameConst &c, const FrameBufs &b, int bin) {
	const int NB = c.bin_count;
	const uint32_t *counts = b.info + LUCID_INFO_U32_SIZE;
	BinInfo bi;
	bi.qc = counts[LUCID_COUNTS_QUAD_COUNTS * NB + bin];
	bi.tc = counts[LUCID_COUNTS_TRI_COUNTS * NB + bin];
	bi.qo = counts[LUCID_COUNTS_QUAD_OFFSETS * NB + bin];
	bi.to = counts[LUCID_COUNTS_TRI_OFFSETS * NB + bin];
	bi.nlevel = bi.tc + 2 * bi.qc;
	bi.bin_x = (bin % c.bin_count_x) * kBinSize;
	bi.bin_y = (bin / c.bin_count_x) * kBinSize;
	return bi;
}

// Exact-fallback queue: one uint4 item per pixel, {block slot, px, py, whole}. Called by
// a whole wave; lanes with `mine` set queue their pixel (px, py).
__device__ __forceinline__ void enqueue_pixels(const FrameBufs &b, uint32_t slot, bool mine, int px,
											   int py, bool whole) {
	const int lane = lane_id();
	const uint64_t sel = ballot(mine);
	if(!sel)
		return;
	uint32_t base = 0;
	if(lane == 0)
		base = atomicAdd(&b.ctrl[kCtrlFallbackCount], (uint32_t)__popcll(sel));
	base = __shfl(base, 0, 64);
	if(mine) {
		const uint32_t i = base + (uint32_t)__popcll(sel & ((1ull << lane) - 1ull));
		if(i < b.fallback_cap)
			b.fallback_queue[i] = make_uint4(slot, (uint32_t)px, (uint32_t)py, whole ? 1u : 0u);
		else
			atomicOr(&b.ctrl[kCtrlOverflow], kOvfFallback);
	}
}

// ------------------------------------------------------------------ rowgen_low

__device__ __forceinline__ void rowgen_low_kernel_body(const FrameConst &c, const FrameBufs &b, WgPos wg) {
	if(b.ctrl[kCtrlOverflow] & kOvfBinLists)
		return;
	const int bin = wg.id;
	const BinInfo bi = load_bin(c, b, bin);
	uint4 *slots = b.blk_slot + (size_t)bin * 16;
	uint32_t *flags = b.blk_flags + (size_t)bin * 16;
	if(bi.nlevel == 0) {
		if(threadIdx.x < 16)
			flags[threadIdx.x] = LUCID_BIN_LEVEL_EMPTY;
		return;
	}
	if(bi.nlevel >= (uint32_t)kLowLevelLimit)
		return; // rowgen_high writes this bin's slots

	__shared__ uint32_t s_row_n[4];
	__shared__ uint32_t s_blk_n[16];
	if(threadIdx.x < 16)
		s_blk_n[threadIdx.x] = 0;
	if(threadIdx.x < 4)
		s_row_n[threadIdx.x] = 0;
	__syncthreads();

	uint4 *recs = b.low_records + (size_t)bin * 4 * kLowRowCap;
	uint32_t *lists = b.blk_lists + (size_t)bin * 16 * kRegSortCap;
	// The bin's live triangles (its list has fewer than kLowLevelLimit entries: one position
	// per thread and chunk) are counting-sorted by their number of 8-row bands, so that the
	// waves' band loops run about as long as each of their triangles needs.
	static_assert(kLowLevelLimit <= 4 * kRowgenThreads, "a low bin's list in one round");
	__shared__ uint32_t s_sorted[kLowLevelLimit];
	__shared__ uint32_t s_nb[4], s_nb_off[4], s_nlive;
	if(threadIdx.x < 4)
		s_nb[threadIdx.x] = 0;
	__syncthreads();
	uint32_t tk[4];
	int hk[4];
#pragma unroll
	for(int k = 0; k < 4; k++) {
		const uint32_t i = k * kRowgenThreads + threadIdx.x;
		const uint32_t t = i < bi.nlevel ? bin_tri_at(b, i, bi.qc, bi.qo, bi.to) : 0xffffffffu;
		int h = -1;
		if(t != 0xffffffffu) {
			const uint32_t yr = b.tri_scan[2 * (size_t)tri_rec(b, t)].w;
			const int lo_by = min(max((int)(yr & 0xffffu) - bi.bin_y, 0), kBinMask) >> 3;
			const int hi_by = min(max((int)(yr >> 16) - bi.bin_y, 0), kBinMask) >> 3;
			h = hi_by > lo_by ? hi_by - lo_by : 0;
			atomicAdd(&s_nb[h], 1u);
		}
		tk[k] = t, hk[k] = h;
	}
	__syncthreads();
	if(threadIdx.x == 0) {
		uint32_t run = 0;
		for(int h = 0; h < 4; h++)
			s_nb_off[h] = run, run += s_nb[h];
		s_nlive = run;
	}
	__syncthreads();
#pragma unroll
	for(int k = 0; k < 4; k++)
		if(hk[k] >= 0)
			s_sorted[atomicAdd(&s_nb_off[hk[k]], 1u)] = tk[k];
	__syncthreads();
	const uint32_t nlive = s_nlive;
	// block-row records, 8-row rule (raster_low.glsl:39-79), one live triangle per lane
	for(uint32_t q = threadIdx.x; q < nlive; q += kRowgenThreads) {
		const uint32_t t = s_sorted[q];
		uint4 s0 = b.tri_scan[2 * (size_t)tri_rec(b, t)], s1 = b.tri_scan[2 * (size_t)tri_rec(b, t) + 1];
		int min_by = min(max((int)(s0.w & 0xffffu) - bi.bin_y, 0), kBinMask) >> 3;
		int max_by = min(max((int)(s0.w >> 16) - bi.bin_y, 0), kBinMask) >> 3;
		RowScan rs;
		row_scan_init(rs, s0, s1, (float)bi.bin_x, (float)(bi.bin_y + min_by * 8));
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
					if(k < (uint32_t)kRegSortCap)
						lists[(by * 4 + bx) * kRegSortCap + k] = idx;
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
			b.info[LUCID_INFO_U32_SIZE + LUCID_COUNTS_HIGH_BINS * c.bin_count + id] = (uint32_t)bin;
		}
		if(threadIdx.x < 16)
			flags[threadIdx.x] = kSlotPromoted;
		return;
	}
	if(threadIdx.x < 16) {
		int blk = threadIdx.x, by = blk >> 2;
		slots[blk] = make_uint4((uint32_t)((size_t)bin * 16 * kRegSortCap + blk * kRegSortCap),
								s_blk_n[blk],
								(uint32_t)((size_t)bin * 4 * kLowRowCap + by * kLowRowCap),
								s_row_n[by]);
		flags[blk] = LUCID_BIN_LEVEL_LOW;
	}
}

__global__ __launch_bounds__(kRowgenThreads) void rowgen_low_kernel(FrameConst c, FrameBufs b) {
	rowgen_low_kernel_body(c, b, WgPos{blockIdx.x, gridDim.x});
}
__global__ __launch_bounds__(kRowgenThreads) void rowgen_low_kernel_batch(const FrameDesc *__restrict__ frames,
		const uint32_t *__restrict__ pre, uint32_t num_frames) {
	const StageBatch sb{frames, pre, num_frames};
	const uint32_t f = batch_frame(sb, blockIdx.x);
	const FrameDesc &d = sb.frames[f];
	rowgen_low_kernel_body(d.c, d.b, WgPos{blockIdx.x - sb.pre[f], sb.pre[f + 1] - sb.pre[f]});
}

// ------------------------------------------------------------------ high bins

// Allocates record and list space of every high bin (promotions included) with an exclusive
// scan over the high list, and clears the counters rowgen_high accumulates into.
__device__ __forceinline__ void high_prep_kernel_body(const FrameConst &c, const FrameBufs &b, WgPos wg) {
	if(b.ctrl[kCtrlOverflow] & kOvfBinLists)
		return;
	const lucid_info *info = (const lucid_info *)b.info;
	const uint32_t nh = (uint32_t)info->bin_level_counts[LUCID_BIN_LEVEL_HIGH];
	const uint32_t *high = b.info + LUCID_INFO_U32_SIZE + LUCID_COUNTS_HIGH_BINS * c.bin_count;
	__shared__ uint32_t s_wave[16];
	__shared__ uint32_t s_carry;
	__shared__ uint32_t s_ex[1024], s_n[1024]; // per bin of the chunk: earlier tris, its tris
	__shared__ uint32_t s_bad[1024 / 32];      // per bin of the chunk: allocation failed
	if(threadIdx.x == 0)
		s_carry = 0;
	__syncthreads();
	const uint64_t list0 = (uint64_t)c.bin_count * 16 * kRegSortCap;
	const int wave = threadIdx.x >> 6;
	for(uint32_t base = 0; base < nh; base += 1024) {
		uint32_t i = base + threadIdx.x;
		int bin = i < nh ? (int)high[i] : 0;
		uint32_t n = i < nh ? load_bin(c, b, bin).nlevel : 0u;
		uint32_t inc = wave_inclusive_add(n);
		if(lane_id() == 63)
			s_wave[wave] = inc;
		__syncthreads();
		if(threadIdx.x == 0) {
			uint32_t run = s_carry;
			for(int w = 0; w < 16; w++) {
				uint32_t v = s_wave[w];
				s_wave[w] = run;
				run += v;
			}
			s_carry = run;
		}
		__syncthreads();
		bool bad = false;
		if(i < nh) {
			uint64_t ex = (uint64_t)s_wave[wave] + inc - n; // tris of earlier high bins
			uint64_t rec = 4 * ex, lst = list0 + 16 * ex;
			bool ok = rec + 4ull * n <= b.high_records_cap && lst + 16ull * n <= b.blk_lists_cap;
			if(!ok)
				atomicOr(&b.ctrl[kCtrlOverflow], kOvfRecords | kOvfLists);
			bad = !ok; // (ex then need not fit 32 bits: the bin's slots are not used)
			if(i == nh - 1) { // sizes the host grows the pools to after an overflow
				b.ctrl[kCtrlRecordCursor] = (uint32_t)min(rec + 4ull * n, 0xffffffffull);
				b.ctrl[kCtrlListCursor] = (uint32_t)min(16ull * (ex + n), 0xffffffffull);
			}
			s_ex[threadIdx.x] = (uint32_t)ex;
			s_n[threadIdx.x] = n;
		}
		if((threadIdx.x & 31) == 0)
			s_bad[threadIdx.x >> 5] = 0;
		__syncthreads();
		if(bad)
			atomicOr(&s_bad[threadIdx.x >> 5], 1u << (threadIdx.x & 31));
		__syncthreads();
		// the chunk's block slots, 16 consecutive ones per bin, one per thread and step
		const uint32_t nb = min(nh - base, 1024u);
		for(uint32_t p = threadIdx.x; p < nb * 16; p += 1024) {
			const uint32_t k = p >> 4, blk = p & 15, band = blk >> 2, bx = blk & 3;
			const uint32_t bin = high[base + k], n = s_n[k];
			const uint64_t ex = s_ex[k], lst = list0 + 16 * ex;
			b.blk_slot[(size_t)bin * 16 + blk] =
				make_uint4((uint32_t)(lst + (uint64_t)(band * 4 + bx) * n), 0u,
						   (uint32_t)(4 * ex + (uint64_t)band * n), 0u);
			b.blk_count[(size_t)bin * 16 + blk] = 0;
			b.blk_flags[(size_t)bin * 16 + blk] = (s_bad[k >> 5] >> (k & 31)) & 1u ? kSlotPromoted : 0u;
		}
		for(uint32_t p = threadIdx.x; p < nb * 4; p += 1024)
			b.band_count[(size_t)high[base + (p >> 2)] * 4 + (p & 3)] = 0;
		__syncthreads();
	}
}

__global__ __launch_bounds__(1024) void high_prep_kernel(FrameConst c, FrameBufs b) {
	high_prep_kernel_body(c, b, WgPos{blockIdx.x, gridDim.x});
}
__global__ __launch_bounds__(1024) void high_prep_kernel_batch(const FrameDesc *__restrict__ frames,
		const uint32_t *__restrict__ pre, uint32_t num_frames) {
	const StageBatch sb{frames, pre, num_frames};
	const uint32_t f = batch_frame(sb, blockIdx.x);
	const FrameDesc &d = sb.frames[f];
	high_prep_kernel_body(d.c, d.b, WgPos{blockIdx.x - sb.pre[f], sb.pre[f + 1] - sb.pre[f]});
}

// Block-row records of a high bin with the 4-row rule (raster_high.glsl:54-90). Each lane
// takes one triangle of the bin and scans its rows once, from the first row of its 4-row
// group (the start row fixes the scanline arithmetic) to its last row, then emits one record
// per 8-row band it covers; kRowgenSplit work-groups share a bin. A wave's row loop runs as
// long as its tallest triangle, so each work-group first sorts its triangles by height (a
// counting sort over the 8 heights in 4-row groups, up to kRowgenRound list chunks at a
// time) and its waves scan them 64 at a time in that order. Record and list positions come
// from wave-aggregated atomics.
constexpr int kRowgenRound = 4; // list chunks of kRowgenThreads entries per sorting round

__device__ __forceinline__ void rowgen_high_kernel_body(const FrameConst &c, const FrameBufs &b, WgPos wg) {
	if(b.ctrl[kCtrlOverflow] & (kOvfBinLists | kOvfRecords | kOvfLists))
		return;
	const lucid_info *info = (const lucid_info *)b.info;
	const uint32_t slot = wg.id / kRowgenSplit;
	const int split = wg.id % kRowgenSplit;
	if(slot >= (uint32_t)info->bin_level_counts[LUCID_BIN_LEVEL_HIGH])
		return;
	const uint32_t *counts = b.info + LUCID_INFO_U32_SIZE;
	const int bin = (int)counts[LUCID_COUNTS_HIGH_BINS * c.bin_count + slot];
	const BinInfo bi = load_bin(c, b, bin);
	const uint4 *slot16 = b.blk_slot + (size_t)bin * 16;
	uint32_t *band_n = b.band_count + (size_t)bin * 4;
	uint32_t *blk_n = b.blk_count + (size_t)bin * 16;
	const int lane = lane_id(), wave = threadIdx.x >> 6;
	const uint32_t stride = kRowgenThreads * kRowgenSplit;
	constexpr int kHeights = kBinSize / 4;
	__shared__ uint32_t s_sorted[kRowgenRound * kRowgenThreads];
	__shared__ uint32_t s_cnt[kHeights], s_off[kHeights], s_n;
	// the row spans of each lane's triangle: rows[r][lane] of the wave
	__shared__ uint16_t s_rows[kRowgenThreads / 64][kBinSize][64];
	uint16_t *rows = &s_rows[wave][0][lane];
	// first row group and height in row groups of triangle t in this bin
	auto groups = [&](uint32_t yr, int &g0, int &g1) {
		g0 = min(max((int)(yr & 0xffffu) - bi.bin_y, 0), kBinMask) >> 2;
		g1 = min(max((int)(yr >> 16) - bi.bin_y, 0), kBinMask) >> 2;
	};
	auto scan_tri = [&](uint32_t t) { // called by the whole wave; t = ~0 on idle lanes
		int row_lo = 0, row_hi = -1;
		RowScan rs;
		if(t != 0xffffffffu) {
			const uint4 s0 = b.tri_scan[2 * (size_t)tri_rec(b, t)], s1 = b.tri_scan[2 * (size_t)tri_rec(b, t) + 1];
			int g0, g1;
			groups(s0.w, g0, g1);
			row_lo = g0 * 4, row_hi = g1 * 4 + 3;
			row_scan_init(rs, s0, s1, (float)bi.bin_x, (float)(bi.bin_y + row_lo));
		}
		uint32_t bxm_all = 0; // block-column masks of the four bands, 4 bits each
		for(int r = row_lo; r <= row_hi; r++) {
			const uint32_t v = row_scan_step(rs);
			rows[r * 64] = (uint16_t)v;
			bxm_all |= span_bx_mask(v) << (4 * (r >> 3));
		}
		// each band's record: rows 0-5 in lo (10 bits each), rows 6-7 in hi; rows the
		// triangle does not reach stay empty (31, 0)
		uint64_t lo_b[4];
		uint32_t hi_b[4], bxm_b[4];
#pragma unroll
		for(int by = 0; by < 4; by++) {
			const uint32_t bxm = (bxm_all >> (4 * by)) & 15u;
			lo_b[by] = 0, hi_b[by] = 0, bxm_b[by] = bxm;
			if(!ballot(bxm != 0))
				continue;
			if(bxm) {
				uint64_t lo = 0;
				uint32_t hi = 0;
#pragma unroll
				for(int k = 0; k < 8; k++) {
					const int r = by * 8 + k;
					const uint32_t v = r >= row_lo && r <= row_hi ? (uint32_t)rows[r * 64] : (uint32_t)kBinMask;
					if(k < 6)
						lo |= (uint64_t)v << (10 * k);
					else
						hi |= v << (10 * (k - 6));
				}
				lo_b[by] = lo, hi_b[by] = hi;
			}
		}
		// per band: its record counter, then its four block-list counters; lane 5 by + k
		// reserves counter k of band by (one atomic instruction for all of them)
		uint64_t emit[4], sel[4][4];
		uint32_t add = 0;
		uint32_t *ctr = band_n;
#pragma unroll
		for(int by = 0; by < 4; by++) {
			emit[by] = ballot(bxm_b[by] != 0);
			if(lane == 5 * by)
				add = (uint32_t)__popcll(emit[by]), ctr = band_n + by;
#pragma unroll
			for(int bx = 0; bx < 4; bx++) {
				sel[by][bx] = ballot((bxm_b[by] >> bx) & 1u);
				if(lane == 5 * by + 1 + bx)
					add = (uint32_t)__popcll(sel[by][bx]), ctr = blk_n + by * 4 + bx;
			}
		}
		uint32_t got = 0;
		if(lane < 20 && add)
			got = atomicAdd(ctr, add);
		const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
		for(int by = 0; by < 4; by++) {
			if(!emit[by])
				continue;
			const uint32_t bxm = bxm_b[by];
			const uint32_t rbase = __shfl(got, 5 * by, 64);
			const uint32_t idx = rbase + (uint32_t)__popcll(emit[by] & below);
			if(bxm)
				b.high_records[slot16[by * 4].z + idx] =
					make_uint4(t, (uint32_t)lo_b[by] & 0x3fffffffu,
							   (uint32_t)(lo_b[by] >> 30) & 0x3fffffffu, hi_b[by] | (bxm << 20));
#pragma unroll
			for(int bx = 0; bx < 4; bx++) {
				const uint32_t lb = __shfl(got, 5 * by + 1 + bx, 64);
				if((bxm >> bx) & 1u)
					b.blk_lists[slot16[by * 4 + bx].x + lb +
								(uint32_t)__popcll(sel[by][bx] & below)] = idx;
			}
		}
	};
	// the work-group's list positions: chunks split, split + kRowgenSplit, ... of
	// kRowgenThreads entries; culled triangles (every second one of a soup quad) are dropped
	const uint32_t round_len = stride * kRowgenRound;
	for(uint32_t base = split * kRowgenThreads; base < bi.nlevel; base += round_len) {
		if(threadIdx.x < kHeights)
			s_cnt[threadIdx.x] = 0;
		__syncthreads();
		uint32_t tk[kRowgenRound];
		int hk[kRowgenRound];
#pragma unroll
		for(int k = 0; k < kRowgenRound; k++) {
			const uint32_t i = base + k * stride + threadIdx.x;
			const uint32_t t = i < bi.nlevel ? bin_tri_at(b, i, bi.qc, bi.qo, bi.to) : 0xffffffffu;
			int h = -1;
			if(t != 0xffffffffu) {
				int g0, g1;
				groups(b.tri_scan[2 * (size_t)tri_rec(b, t)].w, g0, g1);
				h = g1 > g0 ? g1 - g0 : 0;
				atomicAdd(&s_cnt[h], 1u);
			}
			tk[k] = t, hk[k] = h;
		}
		__syncthreads();
		if(threadIdx.x == 0) {
			uint32_t run = 0;
			for(int h = 0; h < kHeights; h++)
				s_off[h] = run, run += s_cnt[h];
			s_n = run;
		}
		__syncthreads();
#pragma unroll
		for(int k = 0; k < kRowgenRound; k++)
			if(hk[k] >= 0)
				s_sorted[atomicAdd(&s_off[hk[k]], 1u)] = tk[k];
		__syncthreads();
		const uint32_t n = s_n;
		for(uint32_t q = wave * 64; q < n; q += kRowgenThreads)
			scan_tri(q + lane < n ? s_sorted[q + lane] : 0xffffffffu);
		if(base + round_len < bi.nlevel)
			__syncthreads(); // the next round reuses the sorted list
	}
}

__global__ __launch_bounds__(kRowgenThreads) void rowgen_high_kernel(FrameConst c, FrameBufs b) {
	rowgen_high_kernel_body(c, b, WgPos{blockIdx.x, gridDim.x});
}
__global__ __launch_bounds__(kRowgenThreads) void rowgen_high_kernel_batch(const FrameDesc *__restrict__ frames,
		const uint32_t *__restrict__ pre, uint32_t num_frames) {
	const StageBatch sb{frames, pre, num_frames};
	const uint32_t f = batch_frame(sb, blockIdx.x);
	const FrameDesc &d = sb.frames[f];
	rowgen_high_kernel_body(d.c, d.b, WgPos{blockIdx.x - sb.pre[f], sb.pre[f + 1] - sb.pre[f]});
}

// Entries (triangles of one block in depth order) are handed to the rasteriser in slices
// of up to kSliceEntries: entry lane + 64 k of the slice sits in register k of that lane.
// (The block rasterisers take the slice width W = 1 or 2 words as a template parameter,
// chosen per renderer by its tuning; the backward kernels use kSliceWords.)
constexpr int kSliceWords = 1;
constexpr int kSliceEntries = 64 * kSliceWords;
constexpr uint32_t kTieRun = 32; // longest run of equal block ranks put in exact order
constexpr int kBigWaves = 4; // waves of block_raster_big_kernel / the backward block kernels

// Pre-sorted block layout in the sorted pool (offsets in words from an even base):
//   [0] number of waves W, [1 .. W + 1] first entry of each wave's range (then m),
//   [kSortedHeader .. + m) triangle ids, [sorted_mask_offset(m) ..) 64-bit pixel masks.
constexpr uint32_t kSortedHeader = 10;
__device__ __forceinline__ uint32_t sorted_mask_offset(uint32_t m) {
	return (kSortedHeader + m + 1u) & ~1u;
}

// Finalises every block of the high bins: copies the list lengths into the slot table and
// pre-sorts long lists. A long list is sorted once by centroid depth (LDS bitonic by the
// whole work-group) and split into `parts` pixel-row groups, each a compacted list of the
// entries that cover it, so that several waves can rasterise a dense block.
// Two passes share the body: CAP = kMidSortCap does the slot bookkeeping of every block,
// sorts the lists of up to kMidSortCap entries (8 KB of keys in LDS, so twice as many
// work-groups per CU as with the full capacity) and queues the longer ones in big_queue;
// CAP = kBigSortCap then takes the queued blocks (a small persistent grid: the queue is
// usually empty). Both sort with the same keys, so the output does not depend on the pass.
constexpr uint32_t kMidSortCap = 2048;

template <uint32_t CAP>
__device__ __forceinline__ void high_finalize_kernel_body(const FrameConst &c, const FrameBufs &b, WgPos wg) {
	if(b.ctrl[kCtrlOverflow] & (kOvfBinLists | kOvfRecords | kOvfLists))
		return;
	const lucid_info *info = (const lucid_info *)b.info;
	const uint32_t hslot = wg.id >> 4;
	const int blk = wg.id & 15, by = blk >> 2, bx = blk & 3;
	if(hslot >= (uint32_t)info->bin_level_counts[LUCID_BIN_LEVEL_HIGH])
		return;
	const uint32_t *counts = b.info + LUCID_INFO_U32_SIZE;
	const int bin = (int)counts[LUCID_COUNTS_HIGH_BINS * c.bin_count + hslot];
	const uint32_t slot = (uint32_t)bin * 16 + blk;
	if(b.blk_flags[slot] & kSlotPromoted)
		return; // allocation failed: the frame reports an overflow
	const uint32_t m = b.blk_count[slot];
	const uint4 sl = b.blk_slot[slot];
	const int bin_x = (bin % c.bin_count_x) * kBinSize, bin_y = (bin / c.bin_count_x) * kBinSize;
	const int px0 = bin_x + bx * 8, py0 = bin_y + by * 8;
	if constexpr(CAP == kMidSortCap) {
		const uint32_t nrec = b.band_count[(size_t)bin * 4 + by];
		if(threadIdx.x == 0)
			b.blk_slot[slot] = make_uint4(sl.x, m, sl.z, nrec);
	}
	if(m <= c.presort_min) { // sorted in registers by the block rasteriser (<= kRegSortCap)
		if(CAP == kMidSortCap && threadIdx.x == 0)
			b.blk_flags[slot] = LUCID_BIN_LEVEL_HIGH;
		return;
	}
	static_assert(CAP == kMidSortCap || CAP == (uint32_t)kBigSortCap, "two sort passes");
	if(CAP == kMidSortCap && m > kMidSortCap) {
		if(threadIdx.x == 0) // the long-list pass takes it
			b.big_queue[atomicAdd(&b.ctrl[kCtrlBigSorts], 1u)] = wg.id;
		return;
	}
	if(m > (uint32_t)kBigSortCap) {
		if(threadIdx.x < 64) {
			const int px = px0 + (threadIdx.x & 7), py = py0 + (threadIdx.x >> 3);
			if(threadIdx.x == 0) {
				b.blk_flags[slot] = LUCID_BIN_LEVEL_HIGH | kSlotWhole;
				atomicAdd(&((lucid_info *)b.info)->stats[3], 1u); // blocks too long to pre-sort
			}
			enqueue_pixels(b, slot, px < c.width && py < c.height, px, py, true);
		}
		return;
	}
	__shared__ uint32_t s_key[CAP];
	__shared__ uint32_t s_lo[kSortThreads / 64], s_hi[kSortThreads / 64];
	__shared__ uint32_t s_wsum[2][kSortThreads / 64];
	__shared__ uint32_t s_tdk[2][kSortThreads]; // full depth keys of tied entries, per round
	__shared__ uint32_t s_bound[kBigWaves + 1];
	__shared__ uint32_t s_base, s_parts, s_total;
	__shared__ uint8_t s_smp[CAP]; // samples (pixel-mask bits) of each entry, in list order
	static_assert(kBigSortCap <= (1 << 13), "list positions fit 13 bits of the key");
	static_assert(kSortThreads == 256, "four waves share the chunks");
	const uint32_t *list = b.blk_lists + sl.x;
	const uint4 *recs = b.high_records + sl.z;
	const int wave_s = threadIdx.x >> 6, lane_s = lane_id();
	const uint32_t P = max(256u, next_pow2(m));
	const uint32_t nchunks = P / 256;
	// a key is (depth rank << pb | list position), with pb = log2(P) position bits, so
	// shorter lists keep more rank bits (24 at P = 256, 19 at 8192)
	const int pb = 31 - __clz(P);
	const uint32_t pmask = P - 1u;
	const bool unsorted = (c.opts & LUCID_OPT_UNSORTED) != 0; // the unsorted baseline
	// The block's pool space (allocated first): the sorted entries (header, triangle ids,
	// masks) and a scratch copy of the same in list order, which pass 1 fills so that the
	// sorted output is gathered from the block's own (cache-resident) scratch instead of
	// the scattered records.
	const uint32_t sorted_words = sorted_mask_offset(m) + 2 * m;
	const uint32_t scratch_mask = sorted_words + ((m + 1u) & ~1u);
	if(threadIdx.x == 0) {
		const uint32_t need = scratch_mask + 2 * m + 1; // + 1: even rounding of the base
		const uint32_t stripe = slot % kSortStripes;
		const uint64_t region = (b.sorted_cap / kSortStripes) & ~1ull;
		const uint32_t off = atomicAdd(sort_cursor(b.ctrl, stripe), need);
		uint32_t base = 0xffffffffu;
		if((uint64_t)off + need <= region) {
			base = (uint32_t)(stripe * region + off);
		} else {
			atomicOr(&b.ctrl[kCtrlOverflow], kOvfSorted);
			const uint64_t want = ((uint64_t)off + need + 2) * kSortStripes; // + even rounding
			atomicMax(&b.ctrl[kCtrlSortedCursor], (uint32_t)min(want, (uint64_t)0xffffffffu));
		}
		s_base = base == 0xffffffffu ? base : (base + 1u) & ~1u;
	}
	__syncthreads();
	if(s_base == 0xffffffffu) {
		if(threadIdx.x == 0)
			b.blk_flags[slot] = kSlotPromoted; // nothing rendered; the frame reports overflow
		return;
	}
	uint32_t *hdr = b.sorted_pool + s_base;
	uint32_t *scratch_tri = hdr + sorted_words;
	uint64_t *scratch_msk = (uint64_t *)(hdr + scratch_mask);
	// Keys are (depth rank << pb | list position). Pass 1 gathers the records once:
	// depth keys (in s_key), the scratch copy, the total of the sample counts and the
	// block's key range.
	uint32_t lo = ~0u, hi = 0u;
	uint32_t mine = 0;
	for(uint32_t ch = wave_s; ch < nchunks; ch += kSortThreads / 64) {
#pragma unroll
		for(int e = 0; e < 4; e++) {
			const uint32_t j = ch * 256 + e * 64 + lane_s;
			if(j < m) {
				const uint4 rec = recs[list[j]];
				const uint64_t msk = record_block_mask(rec, bx);
				scratch_tri[j] = rec.x;
				scratch_msk[j] = msk;
				s_smp[j] = (uint8_t)__popcll(msk);
				mine += (uint32_t)__popcll(msk);
				if(unsorted) {
					s_key[j] = j; // list order: rank 0, the position alone
				} else {
					const uint32_t dk = inv_ord_f32(mask_inv_depth(b, rec.x, msk, px0, py0));
					s_key[j] = dk;
					lo = min(lo, dk);
					hi = max(hi, dk);
				}
			}
		}
	}
	lo = wave_min_u32(lo);
	hi = wave_max_u32(hi);
	mine = __shfl(wave_inclusive_add(mine), 63, 64);
	if(lane_s == 0)
		s_lo[wave_s] = lo, s_hi[wave_s] = hi, s_wsum[1][wave_s] = mine;
	__syncthreads();
	uint32_t tot = 0;
#pragma unroll
	for(int w = 0; w < kSortThreads / 64; w++)
		lo = min(lo, s_lo[w]), hi = max(hi, s_hi[w]), tot += s_wsum[1][w];
	if(threadIdx.x == 0) {
		uint32_t nw = (tot + c.samples_per_wave - 1) / c.samples_per_wave;
		nw = nw < 1u ? 1u : nw > (uint32_t)kBigWaves ? (uint32_t)kBigWaves : nw;
		s_parts = nw;
		s_total = tot;
		for(int w = 0; w <= kBigWaves; w++)
			s_bound[w] = w == 0 ? 0u : m;
	}
	const DepthGrid grid(lo, hi, 32 - pb);
	// Bitonic sort: levels up to 256 in registers (one 256-key chunk per wave at a time; each
	// lane reads back only what it wrote in pass 1), then per level the strides >= 256 in LDS
	// and the rest in registers. A real key never equals the padding ~0u: position P - 1
	// exists only when nothing is padded.
	for(uint32_t ch = wave_s; ch < nchunks && !unsorted; ch += kSortThreads / 64) {
		uint32_t v[4];
#pragma unroll
		for(int e = 0; e < 4; e++) {
			const uint32_t j = ch * 256 + e * 64 + lane_s;
			v[e] = j < m ? (grid.rank(s_key[j]) << pb) | j : ~0u;
		}
		bitonic_regs_from<4>(v, ch * 256);
#pragma unroll
		for(int e = 0; e < 4; e++)
			s_key[ch * 256 + e * 64 + lane_s] = v[e];
	}
	__syncthreads();
	for(uint32_t k = 512; k <= P && !unsorted; k <<= 1) {
		for(uint32_t j = k >> 1; j >= 256; j >>= 1) {
			for(uint32_t i = threadIdx.x; i < P / 2; i += kSortThreads) {
				uint32_t lo_i = ((i & ~(j - 1)) << 1) | (i & (j - 1));
				uint32_t hi_i = lo_i | j;
				bool asc = (lo_i & k) == 0;
				uint32_t a = s_key[lo_i], bb = s_key[hi_i];
				if((a > bb) == asc) {
					s_key[lo_i] = bb;
					s_key[hi_i] = a;
				}
			}
			__syncthreads();
		}
		for(uint32_t ch = wave_s; ch < nchunks; ch += kSortThreads / 64) {
			uint32_t v[4];
#pragma unroll
			for(int e = 0; e < 4; e++)
				v[e] = s_key[ch * 256 + e * 64 + lane_s];
			bitonic_regs_steps<4>(v, ch * 256, k, 128);
#pragma unroll
			for(int e = 0; e < 4; e++)
				s_key[ch * 256 + e * 64 + lane_s] = v[e];
		}
		__syncthreads();
	}
	// s_key[i] & pmask is the list position of the i-th entry in depth order. Entries are
	// visited in rounds of kSortThreads consecutive positions (coalesced LDS and output).
	const uint32_t rounds = (m + kSortThreads - 1) / kSortThreads;
	const uint32_t nw = s_parts, total = s_total;
	uint32_t *tri_out = hdr + kSortedHeader;
	uint64_t *mask_out = (uint64_t *)(hdr + sorted_mask_offset(m));
	// wave w starts at the first entry with at least w * total / nw samples before it whose
	// depth rank differs from its predecessor's: the waves' ranges are chained per pixel
	// and must not split a group of equal ranks (whose order is only the list order). The
	// running prefix is a block scan per round (wave scans, then the wave totals).
	// Equal ranks are ordered by list position only. The forward's ordering window absorbs
	// that, but the backward's pass A flushes its window at every slice, so in frames with
	// fp32 samples (the differentiable ones) each run of equal ranks that lies within one
	// round (and is at most kTieRun long) is put in exact depth-key order here: its members
	// recompute their full depth key and take their place among the run.
	const bool exact_runs = (c.opts & LUCID_OPT_FP32_SAMPLES) != 0 && !unsorted;
	uint32_t carry = 0, found = 0;
	if(!exact_runs) {
		// Without the exact runs, the range starts need only the sample counts, which pass 1
		// left in LDS: each thread takes a contiguous run of sorted positions, one block scan
		// of the runs' totals gives the samples before each entry, and the output is a copy
		// with no barrier between its loads.
		const uint32_t per = (m + kSortThreads - 1) / kSortThreads;
		const uint32_t i0 = min(m, threadIdx.x * per), i1 = min(m, i0 + per);
		uint32_t sum = 0;
		for(uint32_t i = i0; i < i1; i++)
			sum += s_smp[s_key[i] & pmask];
		const uint32_t inc = wave_inclusive_add(sum);
		if(lane_s == 63)
			s_wsum[0][wave_s] = inc;
		__syncthreads();
		uint32_t run = inc - sum;
#pragma unroll
		for(int w = 0; w < kSortThreads / 64; w++)
			run += w < wave_s ? s_wsum[0][w] : 0u;
		for(uint32_t i = i0; i < i1; i++) {
			const uint32_t key = s_key[i];
			const bool cut = unsorted || i == 0 || (key >> pb) != (s_key[i - 1] >> pb);
			for(uint32_t w = 1; w < nw; w++) {
				const uint32_t target = (uint32_t)(((uint64_t)w * total) / nw);
				if(!((found >> w) & 1u) && run >= target && cut) {
					atomicMin(&s_bound[w], i);
					found |= 1u << w;
				}
			}
			run += s_smp[key & pmask];
		}
		for(uint32_t i = threadIdx.x; i < m; i += 4 * kSortThreads) {
			uint32_t tk[4];
			uint64_t mk[4];
#pragma unroll
			for(int k = 0; k < 4; k++) {
				const uint32_t ik = i + k * kSortThreads;
				const uint32_t pos = ik < m ? s_key[ik] & pmask : 0u;
				tk[k] = ik < m ? scratch_tri[pos] : 0u;
				mk[k] = ik < m ? scratch_msk[pos] : 0ull;
			}
#pragma unroll
			for(int k = 0; k < 4; k++) {
				const uint32_t ik = i + k * kSortThreads;
				if(ik < m)
					tri_out[ik] = tk[k], mask_out[ik] = mk[k];
			}
		}
	}
	for(uint32_t r = 0; r < rounds && exact_runs; r++) {
		const uint32_t i = r * kSortThreads + threadIdx.x;
		const uint32_t r0 = r * kSortThreads, r1 = min(m, r0 + kSortThreads);
		uint32_t tri = 0, key = 0, dk = 0;
		uint64_t msk = 0;
		bool tie = false;
		if(i < m) {
			key = s_key[i];
			tri = scratch_tri[key & pmask];
			msk = scratch_msk[key & pmask];
			tie = exact_runs && ((i > 0 && (s_key[i - 1] >> pb) == (key >> pb)) ||
								 (i + 1 < m && (s_key[i + 1] >> pb) == (key >> pb)));
			if(tie)
				dk = inv_ord_f32(mask_inv_depth(b, tri, msk, px0, py0));
		}
		s_tdk[r & 1][threadIdx.x] = dk;
		const uint32_t smp = (uint32_t)__popcll(msk);
		const uint32_t inc = wave_inclusive_add(smp);
		if(lane_s == 63)
			s_wsum[r & 1][wave_s] = inc;
		__syncthreads();
		uint32_t before = carry, round_tot = 0;
#pragma unroll
		for(int w = 0; w < kSortThreads / 64; w++) {
			const uint32_t t = s_wsum[r & 1][w];
			before += w < wave_s ? t : 0u;
			round_tot += t;
		}
		carry += round_tot;
		if(i < m) {
			const uint32_t run = before + inc - smp; // samples of the entries before i
			const bool cut = unsorted || i == 0 || (s_key[i] >> pb) != (s_key[i - 1] >> pb);
			for(uint32_t w = 1; w < nw; w++) {
				const uint32_t target = (uint32_t)(((uint64_t)w * total) / nw);
				if(!((found >> w) & 1u) && run >= target && cut) {
					atomicMin(&s_bound[w], i);
					found |= 1u << w;
				}
			}
			uint32_t out = i;
			if(tie) {
				const uint32_t rk = key >> pb;
				uint32_t gs = i, ge = i;
				while(gs > r0 && i - gs < kTieRun && (s_key[gs - 1] >> pb) == rk)
					gs--;
				while(ge + 1 < r1 && ge - i < kTieRun && (s_key[ge + 1] >> pb) == rk)
					ge++;
				// the whole run is [gs, ge]: it starts and ends within this round
				const bool inside = (gs == 0 || (s_key[gs - 1] >> pb) != rk) &&
									(ge + 1 >= m || (s_key[ge + 1] >> pb) != rk) && ge - gs < kTieRun;
				if(inside) {
					uint32_t ahead = 0;
					for(uint32_t k = gs; k <= ge; k++) {
						const uint32_t dk_k = s_tdk[r & 1][k - r0];
						const uint32_t j_k = s_key[k] & pmask;
						ahead += (dk_k < dk || (dk_k == dk && j_k < (key & pmask))) ? 1u : 0u;
					}
					out = gs + ahead;
				}
			}
			tri_out[out] = tri;
			mask_out[out] = msk;
		}
	}
	__syncthreads();
	if(threadIdx.x == 0) {
		hdr[0] = nw;
		for(uint32_t w = 0; w < nw; w++)
			hdr[1 + w] = s_bound[w];
		hdr[1 + nw] = m;
		b.blk_sorted[slot] = s_base;
		b.blk_flags[slot] = LUCID_BIN_LEVEL_HIGH | kSlotPresorted;
	}
}

__global__ __launch_bounds__(kSortThreads) void high_finalize_kernel(FrameConst c, FrameBufs b) {
	high_finalize_kernel_body<kMidSortCap>(c, b, WgPos{bin_run_id(blockIdx.x, gridDim.x), gridDim.x});
}

// The long-list pass over the blocks queued by the first: work-group w of `count` takes
// queue entries w, w + count, ... (the queue is complete: the first pass has finished).
__device__ __forceinline__ void high_finalize_long_body(const FrameConst &c, const FrameBufs &b,
														 WgPos wg) {
	if(b.ctrl[kCtrlOverflow] & (kOvfBinLists | kOvfRecords | kOvfLists))
		return;
	const uint32_t n = b.ctrl[kCtrlBigSorts];
	for(uint32_t q = wg.id; q < n; q += wg.count) {
		high_finalize_kernel_body<(uint32_t)kBigSortCap>(c, b, WgPos{b.big_queue[q], 0u});
		__syncthreads(); // the next block reuses the work-group's LDS
	}
}
__global__ __launch_bounds__(kSortThreads) void high_finalize_long_kernel(FrameConst c, FrameBufs b) {
	high_finalize_long_body(c, b, WgPos{blockIdx.x, gridDim.x});
}
__global__ __launch_bounds__(kSortThreads) void high_finalize_kernel_batch(const FrameDesc *__restrict__ frames,
		const uint32_t *__restrict__ pre, uint32_t num_frames) {
	const StageBatch sb{frames, pre, num_frames};
	const uint32_t wg = bin_run_id(blockIdx.x, gridDim.x);
	const uint32_t f = batch_frame(sb, wg);
	const FrameDesc &d = sb.frames[f];
	high_finalize_kernel_body<kMidSortCap>(d.c, d.b, WgPos{wg - sb.pre[f], sb.pre[f + 1] - sb.pre[f]});
}
__global__ __launch_bounds__(kSortThreads) void high_finalize_long_kernel_batch(const FrameDesc *__restrict__ frames,
		const uint32_t *__restrict__ pre, uint32_t num_frames) {
	const StageBatch sb{frames, pre, num_frames};
	const uint32_t f = batch_frame(sb, blockIdx.x);
	const FrameDesc &d = sb.frames[f];
	high_finalize_long_body(d.c, d.b, WgPos{blockIdx.x - sb.pre[f], sb.pre[f + 1] - sb.pre[f]});
}

// ------------------------------------------------------------------ block_raster

// 64 x 64 bit-matrix transpose across the wave: lane i holds row i (bit j = column j) and
// receives row i of the transpose. Six butterfly stages; at stage s lanes i and i ^ s swap
// the off-diagonal s x s blocks: the lower lane keeps its columns with (j & s) == 0 and
// takes the partner's as columns j + s, the upper lane keeps its columns with (j & s) != 0
// and takes the partner's, shifted down by s (Hacker's Delight, 7-3, one row per lane).
__device__ __forceinline__ uint64_t transpose64(uint64_t row) {
	const int lane = lane_id();
	{ // s = 32: whole 32-bit words
		uint32_t lo = (uint32_t)row, hi = (uint32_t)(row >> 32);
		const bool upper = (lane & 32) != 0;
		const uint32_t got = __shfl_xor(upper ? lo : hi, 32, 64);
		lo = upper ? got : lo;
		hi = upper ? hi : got;
		row = ((uint64_t)hi << 32) | lo;
	}
	constexpr uint64_t kM0[5] = {0x0000ffff0000ffffull, 0x00ff00ff00ff00ffull,
								 0x0f0f0f0f0f0f0f0full, 0x3333333333333333ull,
								 0x5555555555555555ull};
	// Selects only (the lane condition is a wave-constant lane mask; the column masks and
	// shifts are constants), so nothing lane-dependent stays live between calls.
#pragma unroll
	for(int st = 0; st < 5; st++) {
		const int sh = 16 >> st;
		const bool upper = (lane & sh) != 0;
		const uint64_t lo_cols = row & kM0[st], hi_cols = row & ~kM0[st];
		const uint64_t got = shfl_xor64(upper ? lo_cols : hi_cols, sh);
		const uint64_t moved = upper ? got >> sh : got << sh;
		row = (upper ? hi_cols : lo_cols) | moved;
	}
	return row;
}

// Transposes the coverage masks of a slice (64-bit pixel masks, one per entry) into
// per-pixel entry bitmaps: after the call bit e of bm[k] in lane p is set iff entry
// 64 k + e covers pixel p.
template <int W>
__device__ __forceinline__ void transpose_masks(const uint64_t (&msk)[W], int nwords,
												uint64_t (&bm)[W]) {
#pragma unroll
	for(int k = 0; k < W; k++)
		bm[k] = k < nwords ? transpose64(msk[k]) : 0ull;
}

// ------------------------------------------------------------------ flat colour records
//
// Shading modes of the block rasterisers (ShadeMode): the general record (shade_record),
// the same with albedo textures, and flat colour planes for frames with RGBA8 samples, unit
// lighting (identity_light), vertex colours absent or fp32 with a per-triangle alpha, and
// no textures. There a sample's colour is an affine function of the barycentrics, so per
// channel (with d1 = c1 - c0, d2 = c2 - c0 and the barycentric planes e0, e1 with offsets
// P0, P1 of shading.glsl:116-120)
//   c0 + b0 d1 + b1 d2 = K + ray_pos (Lx px + Ly py + Lz),
//   L = e0 d1 + e1 d2,  K = c0 - P0 d1 - P1 d2,
// which the staging computes once per (block, triangle) and the pixel walk evaluates with
// three fused multiply-adds per channel instead of the barycentric interpolation. The value
// is the same affine function (rounding differs by a few ulps, inside the 1/255 colour bar);
// the depth key keeps its unfused evaluation. The instance colour is folded into the planes.
constexpr int kFlatQuads = 5; // f0 depth plane | triangle id; f1-f3 r, g, b planes; f4 alpha
constexpr bool is_flat(int mode) { return mode == kShadeFlat || mode == kShadeFlatNarrow; }
constexpr int rec_quads(int mode) { return is_flat(mode) ? kFlatQuads : kRecQuads; }

__device__ __forceinline__ void make_flat_record(const uint4 *src, uint4 *dst) {
	const uint4 q0 = src[0], q1 = src[1], q2 = src[2], q3 = src[3], q4 = src[4], q5 = src[5],
				q6 = src[6];
	const uint32_t flags = q0.w & 0xffffu;
	float c0[3] = {1.0f, 1.0f, 1.0f}, c1[3] = {1.0f, 1.0f, 1.0f}, c2[3] = {1.0f, 1.0f, 1.0f};
	float a = 1.0f;
	if(flags & LUCID_INST_HAS_VERTEX_COLORS) {
		c0[0] = bits2f(q3.x), c0[1] = bits2f(q3.y), c0[2] = bits2f(q3.z), a = bits2f(q3.w);
		c1[0] = bits2f(q4.x), c1[1] = bits2f(q4.y), c1[2] = bits2f(q4.z);
		c2[0] = bits2f(q5.x), c2[1] = bits2f(q5.y), c2[2] = bits2f(q5.z);
	}
	if(flags & LUCID_INST_HAS_COLOR) {
		const float4 ic = decode_rgba8(q6.x);
		const float icc[3] = {ic.x, ic.y, ic.z};
#pragma unroll
		for(int k = 0; k < 3; k++)
			c0[k] = icc[k] * c0[k], c1[k] = icc[k] * c1[k], c2[k] = icc[k] * c2[k];
		a = ic.w * a;
	}
	const float e0[3] = {bits2f(q1.x), bits2f(q1.y), bits2f(q1.z)}, p0 = bits2f(q1.w);
	const float e1[3] = {bits2f(q2.x), bits2f(q2.y), bits2f(q2.z)}, p1 = bits2f(q2.w);
	dst[0] = make_uint4(q0.x, q0.y, q0.z, q6.z);
	// a zero alpha contributes nothing (shading.glsl:167-168): zero planes and a zero alpha
	// byte encode such a sample as 0, the RGBA8 word that contributes nothing
	const bool none = a == 0.0f;
#pragma unroll
	for(int k = 0; k < 3; k++) {
		const float d1 = c1[k] - c0[k], d2 = c2[k] - c0[k];
		const float lx = fmaf(e0[0], d1, e1[0] * d2), ly = fmaf(e0[1], d1, e1[1] * d2);
		const float lz = fmaf(e0[2], d1, e1[2] * d2);
		const float kk = fmaf(-p0, d1, fmaf(-p1, d2, c0[k]));
		dst[1 + k] = none ? make_uint4(0u, 0u, 0u, 0u)
						  : make_uint4(f2bits(lx), f2bits(ly), f2bits(lz), f2bits(kk));
	}
	dst[4] = make_uint4(f2u(a * 255.0f) << 24, 0u, 0u, 0u); // RGBA8 alpha byte
}

// RGBA8 sample of a flat record at pixel (px, py); inv_ray = record_inv_depth(f[0], ...).
__device__ __forceinline__ SampleColor<false> shade_flat(const uint4 *f, float px, float py,
														 float inv_ray) {
	const float rp = __builtin_amdgcn_rcpf(inv_ray);
	float rgb[3];
#pragma unroll
	for(int k = 0; k < 3; k++) {
		const uint4 pl = f[1 + k];
		const float l = fmaf(bits2f(pl.x), px, fmaf(bits2f(pl.y), py, bits2f(pl.z)));
		rgb[k] = sat(fmaf(rp, l, bits2f(pl.w)));
	}
	SampleColor<false> out;
	out.enc = (uint32_t)(rgb[0] * 255.0f) | ((uint32_t)(rgb[1] * 255.0f) << 8) |
			  ((uint32_t)(rgb[2] * 255.0f) << 16) | f[4].x;
	return out;
}

// Stages the shading records of a slice in LDS: lane l writes entries l + 64 k.
template <int W, int MODE = kShadeGeneral>
__device__ __forceinline__ void stage_slice(const FrameConst &c, const FrameBufs &b,
											const uint32_t (&tri)[W], uint32_t n,
											uint4 *s_rec) {
	const int lane = lane_id();
#pragma unroll
	for(int k = 0; k < W; k++) {
		const uint32_t i = k * 64 + lane;
		if(i < n) {
			// quad by quad (no staging array: with a record pointer read from memory, as in a
			// batched launch, an array between the load and the LDS store went to scratch)
			const uint4 *src = b.tri_shade + (size_t)kRecStride * tri_rec(b, tri[k]);
			if constexpr(is_flat(MODE)) {
				make_flat_record(src, s_rec + i * kFlatQuads);
			} else {
#pragma unroll
				for(int j = 0; j < kRecQuads; j++)
					s_rec[i * kRecQuads + j] = src[j];
			}
		}
	}
}

// LUCID_OPT_ALPHA_THRESHOLD (shading.glsl:62,282-284; raster_low.glsl:258-260): once every
// pixel of a wave has transmittance below 1/128, the wave's later samples are not shaded
// (their fragments are still counted). "Over" blending only.
constexpr float kAlphaThreshold = 1.0f / 128.0f;
__device__ __forceinline__ bool alpha_threshold_on(const FrameConst &c) {
	return (c.opts & LUCID_OPT_ALPHA_THRESHOLD) && !(c.opts & LUCID_OPT_ADDITIVE_BLENDING);
}
__device__ __forceinline__ bool all_below_alpha_threshold(float trans) {
	return ballot(!(trans < kAlphaThreshold)) == 0ull;
}

// Visits this lane's set bitmap bits in increasing entry order, one per step. The loop is
// divergent: a lane leaves when its bits run out (the wave steps until every lane is done),
// so the lane's state needs no copies for lanes that sit out a step.
template <int W, class Visit>
__device__ __forceinline__ void walk_bits(uint64_t (&bm)[W], Visit &&visit) {
	while(true) {
		bool nz[W], any = false;
#pragma unroll
		for(int k = 0; k < W; k++) {
			nz[k] = bm[k] != 0ull;
			any = any || nz[k];
		}
		if(!any)
			break;
		uint64_t cur = bm[W - 1];
		int word = W - 1;
#pragma unroll
		for(int k = W - 2; k >= 0; k--) {
			cur = nz[k] ? bm[k] : cur;
			word = nz[k] ? k : word;
		}
		const uint64_t rest = cur & (cur - 1ull);
#pragma unroll
		for(int k = 0; k < W; k++)
			bm[k] = word == k ? rest : bm[k];
		visit(word * 64 + (int)__builtin_ctzll(cur));
	}
}

// OVER: the common case compiled apart (sorted, front-to-back "over"), so that its loop
// carries no test of the blend options; otherwise they are read per sample.
template <bool FP32, int W, int MODE, bool OVER, class PS>
__device__ __forceinline__ void walk_pixel_t(const FrameConst &c, const FrameBufs &b,
											 const uint4 *s_rec, uint64_t (&bm)[W], float px,
											 float py, PS &ps, bool additive_opt) {
	const bool unsorted = !OVER && (c.opts & LUCID_OPT_UNSORTED) != 0;
	const bool additive = !OVER && additive_opt;
	if constexpr(is_flat(MODE)) {
		static_assert(!FP32, "flat records carry RGBA8 samples");
		walk_bits(bm, [&](int e) {
			const uint4 *rec = s_rec + e * kFlatQuads;
			const uint4 f0 = rec[0];
			const float inv = record_inv_depth(f0, px, py);
			const SampleColor<false> col = shade_flat(rec, px, py, inv);
			// a sample that contributes nothing (word 0) blends as nothing, so it is not
			// tested for: it takes its place in the window like any other
			if(unsorted)
				ps.acc.blend(col, additive);
			else
				ps.insert(pixel_key(inv, f0.w), col, additive);
		});
		return;
	}
	constexpr bool TEX = MODE == kShadeTextured;
	walk_bits(bm, [&](int e) {
		const uint4 *rec = s_rec + e * kRecQuads;
		const uint4 q0 = rec[0];
		const float inv = record_inv_depth(q0, px, py);
		const SampleColor<FP32> col = shade_record<FP32, TEX>(c, b, rec, px, py, inv);
		if(col.contributes()) {
			if(unsorted)
				ps.acc.blend(col, additive); // list order, no depth order
			else
				ps.insert(pixel_key(inv, rec[6].z), col, additive);
		}
	});
}

template <bool FP32, int W, int MODE, class PS>
__device__ __forceinline__ void walk_pixel(const FrameConst &c, const FrameBufs &b,
										   const uint4 *s_rec, uint64_t (&bm)[W], float px,
										   float py, PS &ps, bool additive) {
	if(!additive && !(c.opts & LUCID_OPT_UNSORTED))
		walk_pixel_t<FP32, W, MODE, true>(c, b, s_rec, bm, px, py, ps, false);
	else
		walk_pixel_t<FP32, W, MODE, false>(c, b, s_rec, bm, px, py, ps, additive);
}

template <int W>
__device__ __forceinline__ uint32_t bitmap_count(const uint64_t (&bm)[W]) {
	uint32_t n = 0;
#pragma unroll
	for(int k = 0; k < W; k++)
		n += (uint32_t)__popcll(bm[k]);
	return n;
}

// Writes the pixel, queues the block's mis-ordered pixels for the exact sort and adds the
// frame statistics (one lane per block pixel).
__device__ __forceinline__ void finish_pixels(const FrameConst &c, const FrameBufs &b,
											  const Accum &acc, bool bad, uint32_t count, int px,
											  int py, uint32_t slot, uint32_t m) {
	const int lane = lane_id();
	const bool on_screen = px < c.width && py < c.height;
	write_pixel(c, b, px, py, acc, count);
	enqueue_pixels(b, slot, bad && on_screen, px, py, false);
	const uint32_t frags = __shfl(wave_inclusive_add(on_screen ? count : 0u), 63, 64);
	if(lane == 0) {
		atomicAdd(stat_stripe(b.ctrl, 0, slot % kStatStripes), frags);
		if(m)
			atomicAdd(stat_stripe(b.ctrl, 1, slot % kStatStripes), m);
	}
}

// One wave per block whose list fits the register sort (every low bin, most high blocks):
// sorts the list by centroid depth in registers, then walks every pixel through it slice by
// slice (each slice's shading records staged in LDS).
template <bool FP32, int W, int MODE>
__device__ __forceinline__ void block_raster_kernel_body(const FrameConst &c, const FrameBufs &b, WgPos wg) {
	if(b.ctrl[kCtrlOverflow] & kOvfBinLists)
		return;
	__shared__ uint4 s_rec[(64 * W) * rec_quads(MODE)];
	const int lane = lane_id();
	const uint32_t slot = wg.id;
	const int bin = slot >> 4, blk = slot & 15, by = blk >> 2, bx = blk & 3;
	const int bin_x = (bin % c.bin_count_x) * kBinSize, bin_y = (bin / c.bin_count_x) * kBinSize;
	const int px0 = bin_x + bx * 8, py0 = bin_y + by * 8;
	const int px = px0 + (lane & 7), py = py0 + (lane >> 3);
	const uint32_t flags = b.blk_flags[slot];
	const uint32_t level = flags & kSlotLevelMask;
	if(flags & (kSlotWhole | kSlotPromoted | kSlotPresorted))
		return; // fallback kernel, overflowed frame, or block_raster_big
	Accum bg;
	bg.init();
	if(level == LUCID_BIN_LEVEL_EMPTY) {
		write_pixel(c, b, px, py, bg, 0u);
		return;
	}
	const uint4 sl = b.blk_slot[slot];
	const uint32_t m = sl.y;
	if(m == 0) {
		write_pixel(c, b, px, py, bg, 0u);
		return;
	}
	if(b.ctrl[kCtrlRecorded])
		return; // the recording pass A renders this block
	const uint4 *recs = (level == LUCID_BIN_LEVEL_LOW ? b.low_records : b.high_records) + sl.z;
	const uint32_t *list = b.blk_lists + sl.x;
	uint64_t msk_all[4];
	uint32_t tri_all[4];
	if(c.opts & LUCID_OPT_UNSORTED) { // list order
#pragma unroll
		for(int j = 0; j < 4; j++) {
			msk_all[j] = 0, tri_all[j] = 0;
			if(j * 64 + lane < (int)m) {
				const uint4 rec = recs[list[j * 64 + lane]];
				msk_all[j] = record_block_mask(rec, bx);
				tri_all[j] = rec.x;
			}
		}
	} else {
		// the sort leaves the masks and triangle ids by list position in the (not yet used)
		// slice staging area; the lanes read them back in depth order
		static_assert(sizeof(s_rec) >= kRegSortCap * (sizeof(uint64_t) + sizeof(uint32_t)),
					  "the sort's masks and ids fit the staging area");
		uint64_t *s_msk = reinterpret_cast<uint64_t *>(s_rec);
		uint32_t *s_tri = reinterpret_cast<uint32_t *>(s_msk + kRegSortCap);
		uint32_t pos[4];
		sort_block_positions(b, recs, list, m, bx, px0, py0, pos, s_msk, s_tri);
		wave_sync_lds();
#pragma unroll
		for(int j = 0; j < 4; j++) {
			msk_all[j] = 0, tri_all[j] = 0;
			if(j * 64 + lane < (int)m) {
				msk_all[j] = s_msk[pos[j]];
				tri_all[j] = s_tri[pos[j]];
			}
		}
		wave_sync_lds(); // read before the first slice is staged over them
	}
	const bool additive = (c.opts & (1u << 5)) != 0;
	const bool threshold = alpha_threshold_on(c);
	PixelState<FP32, MODE == kShadeFlatNarrow ? kNarrowWindow : kWindow> ps;
	ps.init();
	uint32_t count = 0;
	bool opaque = false; // every pixel of the block is nearly opaque (alpha threshold)
	static_assert(4 % W == 0, "register-sorted lists split into whole slices");
	for(uint32_t s0 = 0; s0 < m; s0 += (64 * W)) {
		const int j0 = (int)(s0 / 64);
		uint64_t msk[W];
		uint32_t tri[W];
#pragma unroll
		for(int k = 0; k < W; k++) {
			msk[k] = msk_all[k], tri[k] = tri_all[k];
#pragma unroll
			for(int j = W; j < 4; j++)
				if(j0 + k == j)
					msk[k] = msk_all[j], tri[k] = tri_all[j];
		}
		const uint32_t n = min(m - s0, (uint32_t)(64 * W));
		if(!opaque)
			stage_slice<W, MODE>(c, b, tri, n, s_rec); // loads in flight during the transpose
		uint64_t bm[W];
		transpose_masks(msk, (int)((n + 63) / 64), bm);
		count += bitmap_count(bm);
		if(opaque)
			continue; // counted, not shaded
		wave_sync_lds();
		walk_pixel<FP32, W, MODE>(c, b, s_rec, bm, (float)px, (float)py, ps, additive);
		wave_sync_lds(); // every lane is done with this slice's records
		opaque = threshold && all_below_alpha_threshold(ps.acc.tr);
	}
	ps.flush(additive);
	finish_pixels(c, b, ps.acc, ps.w.bad, count, px, py, slot, m);
}

template <bool FP32, int W, int MODE>
__global__ __launch_bounds__(64) void block_raster_kernel(FrameConst c, FrameBufs b) {
	block_raster_kernel_body<FP32, W, MODE>(c, b, WgPos{bin_run_id(blockIdx.x, gridDim.x), gridDim.x});
}
template <bool FP32, int W, int MODE>
__global__ __launch_bounds__(64) void block_raster_kernel_batch(const FrameDesc *__restrict__ frames,
		const uint32_t *__restrict__ pre, uint32_t num_frames) {
	const StageBatch sb{frames, pre, num_frames};
	const uint32_t wg = bin_run_id(blockIdx.x, gridDim.x);
	const uint32_t f = batch_frame(sb, wg);
	const FrameDesc &d = sb.frames[f];
	block_raster_kernel_body<FP32, W, MODE>(d.c, d.b, WgPos{wg - sb.pre[f], sb.pre[f + 1] - sb.pre[f]});
}

// Pre-sorted lists longer than one register sort: one work-group per block. Wave w takes
// the contiguous depth range [start[w], start[w + 1]) of the sorted list (ranges balanced
// by sample count), walks it slice by slice with one ordering window per pixel, and leaves
// a partial composite; wave 0 chains the partials front to back. A pixel is exact iff each
// range's blended keys increase and every range starts behind the previous one's last key.
template <bool FP32, int W, int MODE>
__device__ __forceinline__ void block_raster_big_kernel_body(const FrameConst &c, const FrameBufs &b, WgPos wg) {
	if(b.ctrl[kCtrlOverflow] & kOvfBinLists)
		return;
	const lucid_info *info = (const lucid_info *)b.info;
	const uint32_t hslot = wg.id >> 4;
	if(hslot >= (uint32_t)info->bin_level_counts[LUCID_BIN_LEVEL_HIGH])
		return;
	const int blk = wg.id & 15, by = blk >> 2, bx = blk & 3;
	const int bin = (int)b.info[LUCID_INFO_U32_SIZE + LUCID_COUNTS_HIGH_BINS * c.bin_count + hslot];
	const uint32_t slot = (uint32_t)bin * 16 + blk;
	const uint32_t flags = b.blk_flags[slot];
	if(!(flags & kSlotPresorted) || (flags & kSlotPromoted) || b.ctrl[kCtrlRecorded])
		return; // (a recording pass A renders the block)
	// Each wave's slice staging area holds, once its walk is done, what the chaining of the
	// ranges reads: 20 KB of LDS per work-group (flat records) instead of 29, so that more
	// work-groups fit on a CU.
	struct RangeEnd {
		float4 part[64]; // the range's partial composite
		uint64_t first[64], last[64];
		uint32_t cnt[64]; // fragments | 0x80000000 when the range's order is not exact
	};
	__shared__ uint4 s_rec[kBigWaves][(64 * W) * rec_quads(MODE)];
	static_assert(sizeof(RangeEnd) <= sizeof(s_rec[0]), "a range's end state fits its staging");
	auto range_end = [&](int w) -> RangeEnd & { return *reinterpret_cast<RangeEnd *>(&s_rec[w][0]); };
	const int wave = threadIdx.x >> 6, lane = lane_id();
	const uint4 sl = b.blk_slot[slot];
	const uint32_t m = sl.y;
	const uint32_t *hdr = b.sorted_pool + b.blk_sorted[slot];
	const uint32_t nw = hdr[0];
	const uint32_t *tri = hdr + kSortedHeader;
	const uint64_t *mask = (const uint64_t *)(hdr + sorted_mask_offset(m));
	const int bin_x = (bin % c.bin_count_x) * kBinSize, bin_y = (bin / c.bin_count_x) * kBinSize;
	const int px0 = bin_x + bx * 8, py0 = bin_y + by * 8;
	const int px = px0 + (lane & 7), py = py0 + (lane >> 3);
	const bool additive = (c.opts & (1u << 5)) != 0;
	if((uint32_t)wave < nw) {
		const uint32_t e0 = hdr[1 + wave], e1 = hdr[2 + wave];
		PixelState<FP32, MODE == kShadeFlatNarrow ? kNarrowBigWindow : kWindow> ps;
		ps.init();
		uint32_t count = 0;
		const bool threshold = alpha_threshold_on(c);
		bool opaque = false; // every pixel is nearly opaque within this range (alpha threshold)
		for(uint32_t s0 = e0; s0 < e1; s0 += (64 * W)) {
			const uint32_t n = min(e1 - s0, (uint32_t)(64 * W));
			uint64_t msk[W];
			uint32_t t[W];
#pragma unroll
			for(int k = 0; k < W; k++) {
				uint32_t i = k * 64 + lane;
				msk[k] = i < n ? mask[s0 + i] : 0ull;
				t[k] = i < n && !opaque ? tri[s0 + i] : 0u;
			}
			if(!opaque)
				stage_slice<W, MODE>(c, b, t, n, s_rec[wave]); // loads in flight during the transpose
			uint64_t bm[W];
			transpose_masks(msk, (int)((n + 63) / 64), bm);
			count += bitmap_count(bm);
			if(opaque)
				continue; // counted, not shaded
			wave_sync_lds();
			walk_pixel<FP32, W, MODE>(c, b, s_rec[wave], bm, (float)px, (float)py, ps, additive);
			wave_sync_lds(); // every lane is done with this slice's records
			// the range's own transmittance bounds the pixel's (earlier ranges only lower it)
			opaque = threshold && all_below_alpha_threshold(ps.acc.tr);
		}
		ps.flush(additive);
		wave_sync_lds(); // every lane is done with the staged records (also after an early end)
		RangeEnd &re = range_end(wave);
		re.part[lane] = make_float4(ps.acc.cr, ps.acc.cg, ps.acc.cb, ps.acc.tr);
		re.first[lane] = ps.w.first;
		re.last[lane] = ps.w.last;
		re.cnt[lane] = count | (ps.w.bad ? 0x80000000u : 0u);
	}
	__syncthreads();
	if(wave != 0)
		return;
	Accum acc;
	acc.init();
	uint64_t last = 0;
	bool bad = false;
	uint32_t count = 0;
	for(uint32_t w = 0; w < nw; w++) {
		const RangeEnd &re = range_end((int)w);
		const uint32_t cw = re.cnt[lane];
		count += cw & 0x7fffffffu;
		bad |= (cw >> 31) != 0;
		const uint64_t f = re.first[lane];
		if(f != kEmptyKey) {
			bad |= f <= last;
			last = re.last[lane];
			const float4 p = re.part[lane];
			acc.cr = acc.cr + acc.tr * p.x;
			acc.cg = acc.cg + acc.tr * p.y;
			acc.cb = acc.cb + acc.tr * p.z;
			acc.tr = acc.tr * p.w;
		}
	}
	finish_pixels(c, b, acc, bad, count, px, py, slot, m);
}

template <bool FP32, int W, int MODE>
__global__ __launch_bounds__(kBigWaves * 64) void block_raster_big_kernel(FrameConst c, FrameBufs b) {
	block_raster_big_kernel_body<FP32, W, MODE>(c, b, WgPos{bin_run_id(blockIdx.x, gridDim.x), gridDim.x});
}
template <bool FP32, int W, int MODE>
__global__ __launch_bounds__(kBigWaves * 64) void block_raster_big_kernel_batch(const FrameDesc *__restrict__ frames,
		const uint32_t *__restrict__ pre, uint32_t num_frames) {
	const StageBatch sb{frames, pre, num_frames};
	const uint32_t wg = bin_run_id(blockIdx.x, gridDim.x);
	const uint32_t f = batch_frame(sb, wg);
	const FrameDesc &d = sb.frames[f];
	block_raster_big_kernel_body<FP32, W, MODE>(d.c, d.b, WgPos{wg - sb.pre[f], sb.pre[f + 1] - sb.pre[f]});
}

// ------------------------------------------------------------------ exact fallback

// Collects the keys (depth key << 32 | triangle id) of one pixel's fragments into `dst`
// (LDS or global) from the block's list; returns the count. Entries beyond `cap` are
// counted but not stored.
__device__ uint32_t gather_pixel_keys(const FrameBufs &b, const uint4 *recs, const uint32_t *list,
									  uint32_t m, int row, int xb, float px, float py,
									  uint64_t *dst, uint32_t cap, bool global_dst) {
	constexpr int kU = 4; // entries per lane per step: their loads are issued together
	const int lane = lane_id();
	uint32_t n = 0;
	for(uint32_t base = 0; base < m; base += 64 * kU) {
		uint32_t idx[kU];
#pragma unroll
		for(int k = 0; k < kU; k++) {
			const uint32_t j = base + k * 64 + lane;
			idx[k] = j < m ? list[j] : 0xffffffffu;
		}
		uint4 rec[kU];
#pragma unroll
		for(int k = 0; k < kU; k++)
			rec[k] = idx[k] != 0xffffffffu ? recs[idx[k]] : make_uint4(0u, kBinMask, kBinMask, kBinMask);
		bool hit[kU];
		uint4 d[kU];
#pragma unroll
		for(int k = 0; k < kU; k++) {
			const uint32_t sp = record_span(rec[k], row);
			const int a = (int)(sp & 31u), e = (int)((sp >> 5) & 31u);
			hit[k] = idx[k] != 0xffffffffu && a <= xb && xb <= e;
			d[k] = hit[k] ? b.tri_depth[tri_rec(b, rec[k].x)] : make_uint4(0u, 0u, 0u, 0u);
		}
#pragma unroll
		for(int k = 0; k < kU; k++) {
			const uint64_t key = pixel_key(record_inv_depth(d[k], px, py), rec[k].x);
			const uint64_t bal = ballot(hit[k]);
			if(hit[k]) {
				const uint32_t pos = n + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
				if(pos < cap) {
					if(global_dst)
						__hip_atomic_store(&dst[pos], key, __ATOMIC_RELAXED,
										   __HIP_MEMORY_SCOPE_AGENT);
					else
						dst[pos] = key;
				}
			}
			n += (uint32_t)__popcll(bal);
		}
	}
	return n;
}

// Blends n sorted keys front to back: 64 samples are shaded in parallel, then composited
// in order.
template <bool FP32, bool TEX>
__device__ Accum blend_sorted(const FrameConst &c, const FrameBufs &b, const uint64_t *keys,
							  uint32_t n, float px, float py, bool global_src) {
	const int lane = lane_id();
	const bool additive = (c.opts & (1u << 5)) != 0;
	Accum acc;
	acc.init();
	for(uint32_t base = 0; base < n; base += 64) {
		const uint32_t i = base + lane;
		SampleColor<FP32> col;
		col.zero();
		if(i < n) {
			const uint64_t k = global_src ? __hip_atomic_load(&keys[i], __ATOMIC_RELAXED,
															  __HIP_MEMORY_SCOPE_AGENT)
										  : keys[i];
			const uint32_t t = (uint32_t)k;
			col = shade_sample<FP32, TEX>(c, b, t, px, py, sample_inv_depth(b, t, px, py));
		}
		const uint32_t cnt = min(64u, n - base);
		for(uint32_t j = 0; j < cnt; j++) {
			const SampleColor<FP32> c2 = col.shfl((int)j);
			if(c2.contributes())
				acc.blend(c2, additive);
		}
	}
	return acc;
}

// Exact path for one queued pixel: gathers its fragments from the block list, sorts the
// keys (LDS bitonic up to kFallbackCap, else a global spill), shades and blends in order.
template <bool FP32, bool TEX>
__device__ void fallback_pixel(const FrameConst &c, const FrameBufs &b, uint4 item,
							   uint64_t *s_keys) {
	const int lane = lane_id();
	const uint32_t slot = item.x;
	const int px = (int)item.y, py = (int)item.z;
	const bool whole = item.w != 0;
	const int blk = (int)(slot & 15u), bx = blk & 3;
	const uint32_t level = b.blk_flags[slot] & kSlotLevelMask;
	const uint4 sl = b.blk_slot[slot];
	const uint4 *recs = (level == LUCID_BIN_LEVEL_LOW ? b.low_records : b.high_records) + sl.z;
	const uint32_t *list = b.blk_lists + sl.x;
	const uint32_t m = sl.y;
	const int row = py & 7, xb = bx * 8 + (px & 7);
	const float fx = (float)px, fy = (float)py;
	const uint32_t n = gather_pixel_keys(b, recs, list, m, row, xb, fx, fy, s_keys, kFallbackCap, false);
	Accum acc;
	if(n <= (uint32_t)kFallbackCap) {
		const uint32_t P = next_pow2(n < 2 ? 2 : n);
		for(uint32_t i = n + lane; i < P; i += 64)
			s_keys[i] = ~0ull;
		wave_sync_lds();
		bitonic_lds_wave(s_keys, P);
		acc = blend_sorted<FP32, TEX>(c, b, s_keys, n, fx, fy, false);
	} else {
		// deep pixel: sort its keys in a global spill region
		const uint32_t P = next_pow2(n);
		uint32_t off = 0;
		if(lane == 0) {
			off = (uint32_t)atomicAdd(&b.ctrl[kCtrlSpillCursor], P);
			if((uint64_t)off + P > b.spill_cap) {
				atomicOr(&b.ctrl[kCtrlOverflow], kOvfSpill);
				off = 0xffffffffu;
			}
		}
		off = __shfl(off, 0, 64);
		if(off == 0xffffffffu)
			return; // the frame reports an overflow and is rendered again
		uint64_t *g = b.spill + off;
		gather_pixel_keys(b, recs, list, m, row, xb, fx, fy, g, P, true);
		for(uint32_t i = n + lane; i < P; i += 64)
			__hip_atomic_store(&g[i], ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		__threadfence();
		__builtin_amdgcn_wave_barrier();
		bitonic_global_wave(g, P);
		acc = blend_sorted<FP32, TEX>(c, b, g, n, fx, fy, true);
	}
	if(lane == 0) {
		write_pixel(c, b, px, py, acc, n);
		lucid_info *info = (lucid_info *)b.info;
		atomicAdd(&b.ctrl[kCtrlFallbackPixels], 1u);
		// stats[2]: pixels resolved by the exact per-pixel sort (the reference counts the
		// pixels its 3-deep window gets wrong there, shading.glsl:258-265)
		atomicAdd(&info->stats[2], 1u);
		if(whole)
			atomicAdd(&info->stats[0], n); // these pixels skipped the block rasteriser
	}
}

// Persistent over the per-pixel queue: wave w of the grid takes items w, w + waves, ...
template <bool FP32, bool TEX>
__device__ __forceinline__ void fallback_kernel_body(const FrameConst &c, const FrameBufs &b, WgPos wg) {
	__shared__ uint64_t s_keys[kFallbackThreads / 64][kFallbackCap];
	if(wg.id == 0 && threadIdx.x < 64) {
		// the block kernels' statistics stripes into LucidInfo (stats[0] fragments, [1]
		// block-list entries); the block kernels have finished
		static_assert(kStatStripes == 64, "one stripe per lane");
		uint32_t v[2] = {*stat_stripe(b.ctrl, 0, threadIdx.x), *stat_stripe(b.ctrl, 1, threadIdx.x)};
#pragma unroll
		for(int k = 0; k < 2; k++) {
			v[k] = __shfl(wave_inclusive_add(v[k]), 63, 64);
			if(threadIdx.x == 0 && v[k])
				atomicAdd(&((lucid_info *)b.info)->stats[k], v[k]);
		}
	}
	const uint32_t nq = min(b.ctrl[kCtrlFallbackCount], b.fallback_cap);
	const uint32_t waves = wg.count * (kFallbackThreads / 64);
	const uint32_t w = wg.id * (kFallbackThreads / 64) + (threadIdx.x >> 6);
	for(uint32_t qi = w; qi < nq; qi += waves)
		fallback_pixel<FP32, TEX>(c, b, b.fallback_queue[qi], s_keys[threadIdx.x >> 6]);
}

template <bool FP32, bool TEX>
__global__ __launch_bounds__(kFallbackThreads) void fallback_kernel(FrameConst c, FrameBufs b) {
	fallback_kernel_body<FP32, TEX>(c, b, WgPos{blockIdx.x, gridDim.x});
}
template <bool FP32, bool TEX>
__global__ __launch_bounds__(kFallbackThreads) void fallback_kernel_batch(const FrameDesc *__restrict__ frames,
		const uint32_t *__restrict__ pre, uint32_t num_frames) {
	const StageBatch sb{frames, pre, num_frames};
	const uint32_t f = batch_frame(sb, blockIdx.x);
	const FrameDesc &d = sb.frames[f];
	fallback_kernel_body<FP32, TEX>(d.c, d.b, WgPos{blockIdx.x - sb.pre[f], sb.pre[f + 1] - sb.pre[f]});
}

// ------------------------------------------------------------------ backward (adjoint)
//
// The adjoint of sorted front-to-back compositing for the fp32-sample forward. Fragment
// lists are recomputed (the forward stages run again), never stored across calls. Per block
// (its sorted list split into the depth ranges of the forward, one wave each):
//   pass A  (bwd_pass_a_kernel) walks every pixel through its range with the forward's
//           ordering window, records the order in which samples leave the window (one byte
//           per sample), and accumulates the range's partial composite in fp64; the
//           work-group then chains the ranges per pixel (exactness across ranges, the
//           saturation mask of the output, dL/d(pixel)) and stores each range's start state:
//           T before it and the remainder R = out - C_before;
//   pass B  (bwd_pass_b_kernel) replays the recorded order and per fragment forms
//             dout/dc_i = a_i T_i,   dout/da_i = T_i c_i - R_{i+1} / (1 - a_i)
//           (T and R in fp64, so pixels whose transmittance underflows fp32 stay exact),
//           chains them through the sRGB lighting stage and the barycentric weights, and
//           sums them per (block, entry) in LDS.
// The (block, entry) rows are then reduced per triangle and per vertex by sorted segmented
// sums (reduce.hip): no global atomics on gradient values. Pixels whose window order is not
// exact take the per-pixel exact sort, as in the forward.

// One sample with what its gradient needs.
struct BwdSample {
	float r, g, b, a;  // shaded sample (fp32, as the fp32-sample forward)
	float b0, b1;      // barycentrics: corner weights (1 - b0) - b1, b0, b1
	float dr, dg, db;  // d(sample channel) / d(vertex colour channel) per unit corner weight
	float da;          // d(sample alpha) / d(triangle alpha)
	__device__ __forceinline__ void zero() { r = g = b = a = b0 = b1 = dr = dg = db = da = 0.0f; }
};

// The fp32-sample shading of shade_record with the derivative of the lighting stage,
// s = sat(L2S(S2L(x) f)) (lit_channel: ds/dx = fp above both linear segments, f below both).
__device__ __forceinline__ bool shade_grad(const FrameConst &c, const FrameBufs &b, const uint4 *q,
										   float px, float py, float inv_ray, BwdSample &o) {
	ShadeState st;
	if(!shade_prepare<false>(c, b, q, px, py, inv_ray, st)) {
		o.zero();
		return false;
	}
	const float x[3] = {st.color.x, st.color.y, st.color.z};
	const float ic[3] = {st.inst.x, st.inst.y, st.inst.z};
	float sv[3], dv[3];
#pragma unroll
	for(int k = 0; k < 3; k++) {
		LitChannel l;
		if(c.identity_light) // s = x, ds/dx = 1 (the round trip of a unit factor)
			l.s = x[k], l.deriv = 1.0f;
		else
			l = lit_channel<true>(x[k], st.f[k], st.fp[k]);
		sv[k] = sat(l.s);
		// clamp passes the gradient on [0, 1] (as torch.clamp does)
		dv[k] = (l.s >= 0.0f && l.s <= 1.0f) ? l.deriv * ic[k] : 0.0f;
	}
	o.r = sv[0], o.g = sv[1], o.b = sv[2], o.a = st.color.w;
	o.b0 = st.bary0, o.b1 = st.bary1;
	o.dr = dv[0], o.dg = dv[1], o.db = dv[2];
	o.da = st.inst.w;
	return true;
}

// Pass A window payload: the fp32 sample and its entry within the slice.
struct BwdColor {
	float r, g, b, a;
	uint32_t e;
	__device__ __forceinline__ void zero() { r = g = b = a = 0.0f, e = 0; }
	static __device__ __forceinline__ BwdColor sel(bool p, const BwdColor &x, const BwdColor &y) {
		BwdColor o;
		o.r = p ? x.r : y.r, o.g = p ? x.g : y.g, o.b = p ? x.b : y.b, o.a = p ? x.a : y.a;
		o.e = p ? x.e : y.e;
		return o;
	}
};

// fp64 composite (C, T) plus the fp32 composite the forward computes (saturation mask).
struct BwdTotals {
	double t, cr, cg, cb;
	Accum acc;
	__device__ __forceinline__ void init() {
		t = 1.0, cr = cg = cb = 0.0;
		acc.init();
	}
	__device__ __forceinline__ void add(float r, float g, float b, float a) {
		const double da = a;
		cr += (double)r * da * t;
		cg += (double)g * da * t;
		cb += (double)b * da * t;
		t *= 1.0 - da;
		acc.blend(make_float4(r, g, b, a), false);
	}
};

// Adjoint of one fragment of front-to-back "over" compositing, out = sum_i s_i a_i T_i +
// T_end bg. State: t = T before the fragment and q = g . R, the gradient-weighted remainder
// R = out - (composite before the fragment), both fp64 (R cancels to tiny values behind
// deep fragments). dL/da = g . (s T - R_after / (1 - a)) = T (g . s) - q_after / (1 - a),
// dL/ds_c = g_c a T.
__device__ __forceinline__ void bwd_fragment(double &t, double &q, float gr, float gg, float gb,
											 const BwdSample &s, float (&out)[kGradWords]) {
	const double a = s.a;
	const double at = a * t;
	const double gs = (double)gr * (double)s.r + (double)gg * (double)s.g + (double)gb * (double)s.b;
	const double q_after = q - at * gs; // g . (remainder behind this fragment)
	const double om = 1.0 - a;          // exact: a is a float in [0, 1]
	// 1 / om: the v_rcp_f32 seed (relative error ~1e-7) refined by one Newton step
	const double r0 = (double)__builtin_amdgcn_rcpf((float)om);
	const double inv_om = om > 0.0 ? r0 * (2.0 - om * r0) : 0.0;
	const double ga = t * gs - q_after * inv_om;
	const float w = (float)at;
	const float xr = gr * w * s.dr, xg = gg * w * s.dg, xb = gb * w * s.db;
	const float w0 = (1.0f - s.b0) - s.b1;
	out[0] = xr * w0, out[1] = xg * w0, out[2] = xb * w0;
	out[3] = xr * s.b0, out[4] = xg * s.b0, out[5] = xb * s.b0;
	out[6] = xr * s.b1, out[7] = xg * s.b1, out[8] = xb * s.b1;
	out[9] = (float)ga * s.da;
	t *= om;
	q = q_after;
}

// Entries of one block in depth order (pre-sorted pool, or a register sort kept in LDS) and
// this wave's depth range [e0, e1) of them.
struct BlockEntries {
	const uint32_t *tri;
	const uint64_t *mask;
	uint32_t nw, e0, e1;
};

// Valid for the blocks the block kernels process (not empty, whole or promoted, m > 0).
// For lists sorted here only wave 0 may call it (it fills s_tri_all / s_mask_all).
__device__ __forceinline__ BlockEntries bwd_block_entries(const FrameConst &c, const FrameBufs &b,
														  uint32_t slot, uint32_t flags, uint4 sl,
														  int wave, uint32_t *s_tri_all,
														  uint64_t *s_mask_all) {
	const int lane = lane_id();
	const uint32_t m = sl.y;
	const int bin = slot >> 4, blk = slot & 15, by = blk >> 2, bx = blk & 3;
	const int px0 = (bin % c.bin_count_x) * kBinSize + bx * 8;
	const int py0 = (bin / c.bin_count_x) * kBinSize + by * 8;
	BlockEntries ent;
	if(flags & kSlotPresorted) {
		const uint32_t *hdr = b.sorted_pool + b.blk_sorted[slot];
		ent.tri = hdr + kSortedHeader;
		ent.mask = (const uint64_t *)(hdr + sorted_mask_offset(m));
		ent.nw = hdr[0];
		ent.e0 = (uint32_t)wave < ent.nw ? hdr[1 + wave] : m;
		ent.e1 = (uint32_t)wave < ent.nw ? hdr[2 + wave] : m;
		return ent;
	}
	const uint32_t level = flags & kSlotLevelMask;
	const uint4 *recs = (level == LUCID_BIN_LEVEL_LOW ? b.low_records : b.high_records) + sl.z;
	const uint32_t *list = b.blk_lists + sl.x;
	uint32_t pos[4];
	sort_block_positions(b, recs, list, m, bx, px0, py0, pos);
#pragma unroll
	for(int j = 0; j < 4; j++)
		if(j * 64 + lane < (int)m) {
			uint4 rec = recs[list[pos[j]]];
			s_mask_all[j * 64 + lane] = record_block_mask(rec, bx);
			s_tri_all[j * 64 + lane] = rec.x;
		}
	wave_sync_lds();
	ent.tri = s_tri_all;
	ent.mask = s_mask_all;
	ent.nw = 1, ent.e0 = 0, ent.e1 = m;
	return ent;
}

// Loads and stages slice [s0, s0 + n) of the block's entries; returns the coverage masks.
__device__ __forceinline__ void bwd_stage(const FrameConst &c, const FrameBufs &b,
										  const BlockEntries &ent, uint32_t s0, uint32_t n,
										  uint4 *s_rec, uint64_t (&msk)[kSliceWords]) {
	const int lane = lane_id();
	uint32_t t[kSliceWords];
#pragma unroll
	for(int k = 0; k < kSliceWords; k++) {
		const uint32_t i = k * 64 + lane;
		msk[k] = i < n ? ent.mask[s0 + i] : 0ull;
		t[k] = i < n ? ent.tri[s0 + i] : 0u;
	}
	stage_slice(c, b, t, n, s_rec);
	wave_sync_lds();
}

constexpr uint8_t kSeqEnd = 0xff; // end of a pixel's recorded order within a slice
static_assert(kSliceEntries <= 255, "recorded entries are bytes");

// Byte of the j-th sample that leaves pixel `lane`'s window in the slice starting at sorted
// position s0 of a block whose gradient rows start at row0.
__device__ __forceinline__ size_t seq_index(uint32_t row0, uint32_t s0, uint32_t j, int lane) {
	return ((size_t)row0 + s0 + j) * 64 + (size_t)lane;
}

// The recorded order, start states and rows are sized by the host from an earlier pass; a
// frame that needs more skips the block and per-pixel adjoints and reports kOvfGradRows
// (the host grows the buffers and the caller runs the backward pass again).
__device__ __forceinline__ bool bwd_fits(const FrameBufs &b, const BwdBufs &g) {
	const uint32_t rows = b.ctrl[kCtrlBwdRows], ranges = b.ctrl[kCtrlBwdRanges];
	return rows <= g.seq_rows_cap && rows <= g.rows_cap && ranges <= g.start_cap;
}

// RECORD: run by a forward frame (when the record gate admitted it) instead of the block
// rasterisers of the blocks it covers. It writes their pixels as the forward does (image,
// counts, statistics, exact-fallback queue) and, instead of the start states (which need
// dL/dpixel), the BwdRecord of each (range, pixel) and the pixel's clamp/exact flags;
// bwd_start_kernel turns them into start states when the backward runs.
// WIN: slots of the ordering window (4, or 2 with the narrow-window tuning).
template <bool RECORD, int WIN>
__global__ __launch_bounds__(kBigWaves * 64) void bwd_pass_a_kernel(FrameConst c, FrameBufs b,
																	 BwdBufs g) {
	if(b.ctrl[kCtrlOverflow] & kOvfBinLists)
		return;
	if constexpr(RECORD) {
		if(!b.ctrl[kCtrlRecorded])
			return;
	} else if(!bwd_fits(b, g)) {
		if(blockIdx.x == 0 && threadIdx.x == 0)
			atomicOr(&b.ctrl[kCtrlOverflow], kOvfGradRows);
		return;
	}
	// Each wave's slice staging area holds, once its walk is done, what the chaining of the
	// ranges reads (the range's composites and ordering keys per pixel): 31 KB of LDS per
	// work-group instead of 49, so that four work-groups fit on a CU.
	struct RangeEnd {
		double pc[4][64]; // C rgb, T (fp64)
		float4 p32[64];   // fp32 partial composite
		uint64_t first[64], last[64];
		uint32_t bad[64], count[64]; // count: RECORD, fragments per pixel
	};
	__shared__ uint4 s_rec[kBigWaves][kSliceEntries * kRecQuads];
	static_assert(sizeof(RangeEnd) <= sizeof(s_rec[0]), "a range's end state fits its staging");
	__shared__ uint32_t s_tri_all[kRegSortCap];
	__shared__ uint64_t s_mask_all[kRegSortCap];
	auto range_end = [&](int w) -> RangeEnd & { return *reinterpret_cast<RangeEnd *>(&s_rec[w][0]); };
	const int lane = lane_id(), wave = threadIdx.x >> 6;
	const uint32_t slot = blockIdx.x;
	const uint32_t flags = b.blk_flags[slot];
	const uint32_t level = flags & kSlotLevelMask;
	if(level == LUCID_BIN_LEVEL_EMPTY || (flags & (kSlotWhole | kSlotPromoted)))
		return;
	const uint4 sl = b.blk_slot[slot];
	if(sl.y == 0)
		return;
	const bool presorted = (flags & kSlotPresorted) != 0;
	if(!presorted && wave != 0)
		return; // one range: wave 0 alone (no work-group barrier below on this path)
	const int bin = slot >> 4, blk = slot & 15, by = blk >> 2, bx = blk & 3;
	const int px = (bin % c.bin_count_x) * kBinSize + bx * 8 + (lane & 7);
	const int py = (bin / c.bin_count_x) * kBinSize + by * 8 + (lane >> 3);
	const bool on_screen = px < c.width && py < c.height;
	const float fx = (float)px, fy = (float)py;
	const BlockEntries ent = bwd_block_entries(c, b, slot, flags, sl, wave, s_tri_all, s_mask_all);
	const uint32_t row0 = g.row_base[slot];

	if((uint32_t)wave < ent.nw) {
		std::conditional_t<WIN == kWindow, OrderWindow<BwdColor>, SmallWindow<BwdColor, WIN, false>> w;
		w.init();
		BwdTotals part;
		part.init();
		uint32_t count = 0;
		for(uint32_t s0 = ent.e0; s0 < ent.e1; s0 += kSliceEntries) {
			const uint32_t n = min(ent.e1 - s0, (uint32_t)kSliceEntries);
			uint64_t msk[kSliceWords];
			bwd_stage(c, b, ent, s0, n, s_rec[wave], msk);
			// the triangle of each gradient row (pass B fills the row itself)
			for(uint32_t e = lane; e < n; e += 64)
				g.row_tri[(size_t)row0 + s0 + e] = ent.tri[s0 + e];
			uint64_t bm[kSliceWords];
			transpose_masks(msk, (int)((n + 63) / 64), bm);
			if constexpr(RECORD)
				count += bitmap_count(bm);
			uint32_t j = 0;
			auto emit = [&](const BwdColor &s) {
				part.add(s.r, s.g, s.b, s.a);
				g.seq[seq_index(row0, s0, j, lane)] = (uint8_t)s.e;
				j++;
			};
			walk_bits(bm, [&](int e) {
				const uint4 *rec = s_rec[wave] + e * kRecQuads;
				const float inv = record_inv_depth(rec[0], fx, fy);
				const SampleColor<true> col = shade_record<true, false>(c, b, rec, fx, fy, inv);
				if(col.contributes()) {
					BwdColor bc;
					bc.r = col.col.x, bc.g = col.col.y, bc.b = col.col.z, bc.a = col.col.w;
					bc.e = (uint32_t)e;
					w.insert(pixel_key(inv, rec[6].z), bc, emit);
				}
			});
			// the window empties at every slice end, so pass B only needs this slice staged
			w.flush(emit);
			if(j < n)
				g.seq[seq_index(row0, s0, j, lane)] = kSeqEnd;
			wave_sync_lds();
		}
		RangeEnd &re = range_end(wave); // every lane is done with the last slice's records
		re.pc[0][lane] = part.cr, re.pc[1][lane] = part.cg;
		re.pc[2][lane] = part.cb, re.pc[3][lane] = part.t;
		re.p32[lane] = make_float4(part.acc.cr, part.acc.cg, part.acc.cb, part.acc.tr);
		re.first[lane] = w.first, re.last[lane] = w.last;
		re.bad[lane] = w.bad ? 1u : 0u;
		re.count[lane] = count;
	}
	if(presorted)
		__syncthreads();
	else
		wave_sync_lds();
	if((uint32_t)wave >= ent.nw)
		return;

	// chain the ranges front to back (each wave for its own start state)
	const float bg[3] = {c.background[0], c.background[1], c.background[2]};
	double ct[3] = {0.0, 0.0, 0.0}, tt = 1.0;   // composite of all ranges
	double cb4[3] = {0.0, 0.0, 0.0}, tb4 = 1.0; // composite of the ranges before this one
	Accum a32;
	a32.init();
	uint64_t last = 0;
	bool bad = false;
	for(uint32_t v = 0; v < ent.nw; v++) {
		if(v == (uint32_t)wave)
			cb4[0] = ct[0], cb4[1] = ct[1], cb4[2] = ct[2], tb4 = tt;
		const RangeEnd &re = range_end((int)v);
		bad |= re.bad[lane] != 0;
		const uint64_t f = re.first[lane];
		if(f != kEmptyKey) {
			bad |= f <= last;
			last = re.last[lane];
		}
		for(int k = 0; k < 3; k++)
			ct[k] = ct[k] + tt * re.pc[k][lane];
		tt = tt * re.pc[3][lane];
		const float4 p = re.p32[lane];
		a32.cr = a32.cr + a32.tr * p.x;
		a32.cg = a32.cg + a32.tr * p.y;
		a32.cb = a32.cb + a32.tr * p.z;
		a32.tr = a32.tr * p.w;
	}
	const float o32[3] = {a32.cr + a32.tr * bg[0], a32.cg + a32.tr * bg[1],
						  a32.cb + a32.tr * bg[2]};
	if constexpr(RECORD) {
		if(wave == 0) {
			uint32_t count = 0;
			for(uint32_t v = 0; v < ent.nw; v++)
				count += range_end((int)v).count[lane];
			finish_pixels(c, b, a32, bad, count, px, py, slot, sl.y);
			if(on_screen) {
				uint32_t pf = bad ? 0u : 8u;
#pragma unroll
				for(int k = 0; k < 3; k++)
					pf |= (o32[k] >= 0.0f && o32[k] <= 1.0f) ? 1u << k : 0u;
				g.pix_flags[(size_t)py * c.width + px] = (uint8_t)pf;
			}
		}
		BwdRecord rc;
		rc.t = tb4;
#pragma unroll
		for(int k = 0; k < 3; k++)
			rc.d[k] = (ct[k] + tt * (double)bg[k]) - cb4[k];
		g.rec[((size_t)g.range_base[slot] + wave) * 64 + lane] = rc;
		return;
	}
	if(wave == 0)
		enqueue_pixels(b, slot, on_screen && bad, px, py, false);
	float gsel[3] = {0.0f, 0.0f, 0.0f};
	if(on_screen && !bad) {
		const float4 gi = g.grad_image[(size_t)py * c.width + px];
		const float gv[3] = {gi.x, gi.y, gi.z};
#pragma unroll
		for(int k = 0; k < 3; k++)
			gsel[k] = (o32[k] >= 0.0f && o32[k] <= 1.0f) ? gv[k] : 0.0f;
	}
	BwdStart st;
	st.t = tb4;
	st.q = 0.0;
#pragma unroll
	for(int k = 0; k < 3; k++)
		st.q += (double)gsel[k] * ((ct[k] + tt * (double)bg[k]) - cb4[k]);
	st.gr = gsel[0], st.gg = gsel[1], st.gb = gsel[2];
	st.pad = 0.0f;
	g.start[((size_t)g.range_base[slot] + wave) * 64 + lane] = st;
}

// Start states of a recorded frame (what pass A computes at its end once dL/dpixel is
// known): one wave per (block, depth range), as pass B.
__global__ __launch_bounds__(kBigWaves * 64) void bwd_start_kernel(FrameConst c, FrameBufs b,
																	BwdBufs g) {
	if((b.ctrl[kCtrlOverflow] & kOvfBinLists) || !bwd_fits(b, g))
		return;
	const int lane = lane_id(), wave = threadIdx.x >> 6;
	const uint32_t slot = blockIdx.x;
	const uint32_t flags = b.blk_flags[slot];
	const uint32_t level = flags & kSlotLevelMask;
	if(level == LUCID_BIN_LEVEL_EMPTY || (flags & (kSlotWhole | kSlotPromoted)))
		return;
	if(b.blk_slot[slot].y == 0)
		return;
	const uint32_t nw = (flags & kSlotPresorted) ? b.sorted_pool[b.blk_sorted[slot]] : 1u;
	if((uint32_t)wave >= nw)
		return;
	const int bin = slot >> 4, blk = slot & 15, by = blk >> 2, bx = blk & 3;
	const int px = (bin % c.bin_count_x) * kBinSize + bx * 8 + (lane & 7);
	const int py = (bin / c.bin_count_x) * kBinSize + by * 8 + (lane >> 3);
	const size_t r = ((size_t)g.range_base[slot] + wave) * 64 + lane;
	const BwdRecord rc = g.rec[r];
	const uint32_t pf = (px < c.width && py < c.height) ? g.pix_flags[(size_t)py * c.width + px] : 0u;
	float gsel[3] = {0.0f, 0.0f, 0.0f};
	if(pf & 8u) {
		const float4 gi = g.grad_image[(size_t)py * c.width + px];
		const float gv[3] = {gi.x, gi.y, gi.z};
#pragma unroll
		for(int k = 0; k < 3; k++)
			gsel[k] = ((pf >> k) & 1u) ? gv[k] : 0.0f;
	}
	BwdStart st;
	st.t = rc.t;
	st.q = 0.0;
#pragma unroll
	for(int k = 0; k < 3; k++)
		st.q += (double)gsel[k] * rc.d[k];
	st.gr = gsel[0], st.gg = gsel[1], st.gb = gsel[2];
	st.pad = 0.0f;
	g.start[r] = st;
}

constexpr int kGradStride = kGradWords + 1; // odd stride against bank conflicts

// One wave per (block, depth range); single-wave work-groups (no LDS is shared between
// waves), so that more of them fit on a CU.
__global__ __launch_bounds__(64) void bwd_pass_b_kernel(FrameConst c, FrameBufs b, BwdBufs g) {
	if((b.ctrl[kCtrlOverflow] & kOvfBinLists) || !bwd_fits(b, g))
		return;
	__shared__ uint4 s_rec[kSliceEntries * kRecQuads];
	__shared__ unsigned long long s_grad[kSliceEntries * kGradStride];
	__shared__ uint32_t s_tri_all[kRegSortCap];
	__shared__ uint64_t s_mask_all[kRegSortCap];
	const int lane = lane_id();
	const uint32_t slot = blockIdx.x / kBigWaves;
	const int wave = (int)(blockIdx.x % kBigWaves);
	const uint32_t flags = b.blk_flags[slot];
	const uint32_t level = flags & kSlotLevelMask;
	if(level == LUCID_BIN_LEVEL_EMPTY || (flags & (kSlotWhole | kSlotPromoted)))
		return;
	const uint4 sl = b.blk_slot[slot];
	if(sl.y == 0)
		return;
	if(!(flags & kSlotPresorted) && wave != 0)
		return;
	if(flags & kSlotPresorted) {
		const uint32_t nw = b.sorted_pool[b.blk_sorted[slot]];
		if((uint32_t)wave >= nw)
			return;
	}
	const int bin = slot >> 4, blk = slot & 15, by = blk >> 2, bx = blk & 3;
	const int px = (bin % c.bin_count_x) * kBinSize + bx * 8 + (lane & 7);
	const int py = (bin / c.bin_count_x) * kBinSize + by * 8 + (lane >> 3);
	const float fx = (float)px, fy = (float)py;
	const BlockEntries ent = bwd_block_entries(c, b, slot, flags, sl, wave, s_tri_all, s_mask_all);
	const uint32_t row0 = g.row_base[slot];
	const BwdStart st = g.start[((size_t)g.range_base[slot] + wave) * 64 + lane];
	double t = st.t, q = st.q;
	const bool live = st.gr != 0.0f || st.gg != 0.0f || st.gb != 0.0f;
	const int fix_k = grad_fixed_exp(c, bits2f(b.ctrl[kCtrlBwdGradMax]));
	const double fix_scale = ldexp(1.0, fix_k);
	const bool rot_b0 = (lane & 1) != 0, rot_b1 = (lane & 2) != 0;
	const int rot = rot_b1 ? (rot_b0 ? 8 : 5) : (rot_b0 ? 3 : 0);
	for(uint32_t s0 = ent.e0; s0 < ent.e1; s0 += kSliceEntries) {
		const uint32_t n = min(ent.e1 - s0, (uint32_t)kSliceEntries);
		uint64_t msk[kSliceWords];
		bwd_stage(c, b, ent, s0, n, s_rec, msk);
		for(int i = lane; i < kSliceEntries * kGradStride; i += 64)
			s_grad[i] = 0ull;
		wave_sync_lds();
		// replay the order recorded by pass A (byte j of this lane at seq[64 j], read one
		// step ahead; a lane stops at its end marker and reads nothing beyond it). The loop
		// is divergent: each lane leaves at its own end, so its state (t, q) needs no copies.
		const uint8_t *seq = g.seq + seq_index(row0, s0, 0, lane);
		uint8_t next = seq[0];
		for(uint32_t j = 0; j < n && next != kSeqEnd; j++) {
			const uint8_t e = next;
			next = j + 1 < n ? seq[(size_t)(j + 1) * 64] : kSeqEnd;
			const uint4 *rec = s_rec + (uint32_t)e * kRecQuads;
			const float inv = record_inv_depth(rec[0], fx, fy);
			BwdSample smp;
			shade_grad(c, b, rec, fx, fy, inv, smp);
			float out[kGradWords];
			bwd_fragment(t, q, st.gr, st.gg, st.gb, smp, out);
			if(live) {
				// lanes add the ten components in four rotated orders (pixel column mod 4: by
				// 0, 3, 5 or 8), so that pixels at the same entry in the same step seldom hit
				// one address (integer sums: the order changes no bit). The order is picked with
				// selects on two lane bits: an equality chain on lane & 3 would compile to a
				// branchy switch per component.
#pragma unroll
				for(int k = 0; k < kGradWords; k++) {
					const int k1 = (k + 3) % kGradWords, k2 = (k + 5) % kGradWords;
					const int k3 = (k + 8) % kGradWords;
					const float v = rot_b1 ? (rot_b0 ? out[k3] : out[k2]) : (rot_b0 ? out[k1] : out[k]);
					const int kk0 = k + rot;
					const int kk = kk0 >= kGradWords ? kk0 - kGradWords : kk0;
					atomicAdd(&s_grad[(uint32_t)e * kGradStride + kk], to_fixed(v, fix_scale));
				}
			}
		}
		wave_sync_lds();
		for(uint32_t e = lane; e < n; e += 64) {
			const size_t row = (size_t)row0 + s0 + e;
			if(row < g.rows_cap) {
#pragma unroll
				for(int k = 0; k < kGradWords; k++)
					g.rows[row * kGradWords + k] =
						(float)ldexp((double)(long long)s_grad[e * kGradStride + k], -fix_k);
			}
		}
		wave_sync_lds();
	}
}

// Exclusive scans over the blocks the block kernels process: gradient rows (one per list
// entry) and depth ranges (start states of pass B).
// One thread queues every on-screen pixel of block `slot` for the exact per-pixel path (a
// list too long to pre-sort), as high_finalize does for a new frame.
__device__ void enqueue_block_whole(const FrameConst &c, const FrameBufs &b, uint32_t slot) {
	const int bin = (int)(slot >> 4), blk = (int)(slot & 15);
	const int x0 = (bin % c.bin_count_x) * kBinSize + (blk & 3) * 8;
	const int y0 = (bin / c.bin_count_x) * kBinSize + (blk >> 2) * 8;
	const int w = min(8, c.width - x0), h = min(8, c.height - y0);
	if(w <= 0 || h <= 0)
		return;
	const uint32_t base = atomicAdd(&b.ctrl[kCtrlFallbackCount], (uint32_t)(w * h));
	for(int i = 0; i < w * h; i++) {
		if(base + (uint32_t)i >= b.fallback_cap) {
			atomicOr(&b.ctrl[kCtrlOverflow], kOvfFallback);
			break;
		}
		b.fallback_queue[base + i] =
			make_uint4(slot, (uint32_t)(x0 + i % w), (uint32_t)(y0 + i / w), 1u);
	}
}

// REUSE (a finished frame's lists, reused by the backward): the control words start afresh
// and the exact-fallback queue is rebuilt: the blocks too long to pre-sort are queued here,
// pass A adds the mis-ordered pixels.
template <bool REUSE>
__global__ __launch_bounds__(1024) void grad_row_scan_kernel(FrameConst c, FrameBufs b, BwdBufs g) {
	const uint32_t n = (uint32_t)c.bin_count * 16;
	__shared__ uint32_t s_wave[2][16];
	__shared__ uint32_t s_carry[2];
	if constexpr(REUSE) {
		if(threadIdx.x < (uint32_t)kCtrlWords)
			b.ctrl[threadIdx.x] = 0u;
	}
	if(threadIdx.x < 2)
		s_carry[threadIdx.x] = 0;
	__syncthreads();
	const int wave = threadIdx.x >> 6;
	for(uint32_t base = 0; base < n; base += 1024) {
		const uint32_t slot = base + threadIdx.x;
		uint32_t v[2] = {0u, 0u};
		if(slot < n) {
			const uint32_t flags = b.blk_flags[slot];
			const uint32_t level = flags & kSlotLevelMask;
			const uint32_t m = b.blk_slot[slot].y;
			if(level != LUCID_BIN_LEVEL_EMPTY && !(flags & (kSlotWhole | kSlotPromoted)) && m) {
				v[0] = m;
				v[1] = (flags & kSlotPresorted) ? b.sorted_pool[b.blk_sorted[slot]] : 1u;
			}
			if(REUSE && (flags & kSlotWhole))
				enqueue_block_whole(c, b, slot);
		}
		uint32_t inc[2];
		for(int k = 0; k < 2; k++) {
			inc[k] = wave_inclusive_add(v[k]);
			if(lane_id() == 63)
				s_wave[k][wave] = inc[k];
		}
		__syncthreads();
		if(threadIdx.x < 2) {
			const int k = threadIdx.x;
			uint32_t run = s_carry[k];
			for(int w = 0; w < 16; w++) {
				uint32_t x = s_wave[k][w];
				s_wave[k][w] = run;
				run += x;
			}
			s_carry[k] = run;
		}
		__syncthreads();
		if(slot < n) {
			g.row_base[slot] = s_wave[0][wave] + inc[0] - v[0];
			g.range_base[slot] = s_wave[1][wave] + inc[1] - v[1];
		}
		__syncthreads();
	}
	if(threadIdx.x == 0) {
		b.ctrl[kCtrlBwdRows] = s_carry[0];
		b.ctrl[kCtrlBwdRanges] = s_carry[1];
		b.ctrl[kCtrlBwdExtra] = 0;
		b.ctrl[kCtrlBwdGradMax] = 0;
	}
}

// max |dL/dpixel| (rgb) over the image into kCtrlBwdGradMax (after the row scan cleared it):
// fixes the frame's gradient fixed-point scale.
__global__ __launch_bounds__(256) void grad_max_kernel(FrameConst c, BwdBufs g, uint32_t *ctrl) {
	const uint32_t n = (uint32_t)c.width * (uint32_t)c.height;
	float gm = 0.0f;
	for(uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
		const float4 v = g.grad_image[i];
		gm = fmaxf(gm, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fabsf(v.z)));
	}
#pragma unroll
	for(int d = 1; d < 64; d <<= 1)
		gm = fmaxf(gm, __shfl_xor(gm, d, 64));
	if(lane_id() == 0 && gm > 0.0f)
		atomicMax(&ctrl[kCtrlBwdGradMax], f2bits(gm)); // non-negative floats order as uints
}

// Exact adjoint of one queued pixel: its fragments sorted exactly (as fallback_pixel), then
// both passes over the sorted list; one gradient row per fragment, appended after the
// block rows.
__device__ void bwd_fallback_pixel(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
								   uint4 item, uint64_t *s_keys) {
	const int lane = lane_id();
	const uint32_t slot = item.x;
	const int px = (int)item.y, py = (int)item.z;
	const int blk = (int)(slot & 15u), bx = blk & 3;
	const uint32_t level = b.blk_flags[slot] & kSlotLevelMask;
	const uint4 sl = b.blk_slot[slot];
	const uint4 *recs = (level == LUCID_BIN_LEVEL_LOW ? b.low_records : b.high_records) + sl.z;
	const uint32_t *list = b.blk_lists + sl.x;
	const uint32_t m = sl.y;
	const int row = py & 7, xb = bx * 8 + (px & 7);
	const float fx = (float)px, fy = (float)py;
	const uint32_t n =
		gather_pixel_keys(b, recs, list, m, row, xb, fx, fy, s_keys, kFallbackCap, false);
	const uint64_t *keys = s_keys;
	bool global_keys = false;
	if(n <= (uint32_t)kFallbackCap) {
		const uint32_t P = next_pow2(n < 2 ? 2 : n);
		for(uint32_t i = n + lane; i < P; i += 64)
			s_keys[i] = ~0ull;
		wave_sync_lds();
		bitonic_lds_wave(s_keys, P);
	} else {
		const uint32_t P = next_pow2(n);
		uint32_t off = 0;
		if(lane == 0) {
			off = (uint32_t)atomicAdd(&b.ctrl[kCtrlSpillCursor], P);
			if((uint64_t)off + P > b.spill_cap) {
				atomicOr(&b.ctrl[kCtrlOverflow], kOvfSpill);
				off = 0xffffffffu;
			}
		}
		off = __shfl(off, 0, 64);
		if(off == 0xffffffffu)
			return;
		uint64_t *gk = b.spill + off;
		gather_pixel_keys(b, recs, list, m, row, xb, fx, fy, gk, P, true);
		for(uint32_t i = n + lane; i < P; i += 64)
			__hip_atomic_store(&gk[i], ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		__threadfence();
		__builtin_amdgcn_wave_barrier();
		bitonic_global_wave(gk, P);
		keys = gk;
		global_keys = true;
	}
	auto key_at = [&](uint32_t i) -> uint64_t {
		return global_keys ? __hip_atomic_load(&keys[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
						   : keys[i];
	};
	auto shade_at = [&](uint32_t i, BwdSample &smp) -> bool {
		const uint32_t t = (uint32_t)key_at(i);
		uint4 q[kRecQuads];
		load_shading_record(b, t, q);
		return shade_grad(c, b, q, fx, fy, record_inv_depth(q[0], fx, fy), smp);
	};
	// pass A (serial over the sorted list, 64 samples shaded at a time)
	BwdTotals tot;
	tot.init();
	for(uint32_t base = 0; base < n; base += 64) {
		BwdSample smp;
		smp.zero();
		const bool ok = base + lane < n && shade_at(base + lane, smp);
		const uint32_t cnt = min(64u, n - base);
		for(uint32_t j = 0; j < cnt; j++) {
			BwdSample s;
			s.r = __shfl(smp.r, (int)j, 64), s.g = __shfl(smp.g, (int)j, 64);
			s.b = __shfl(smp.b, (int)j, 64), s.a = __shfl(smp.a, (int)j, 64);
			if(__shfl((int)ok, (int)j, 64))
				tot.add(s.r, s.g, s.b, s.a);
		}
	}
	const float bg[3] = {c.background[0], c.background[1], c.background[2]};
	const float o32[3] = {tot.acc.cr + tot.acc.tr * bg[0], tot.acc.cg + tot.acc.tr * bg[1],
						  tot.acc.cb + tot.acc.tr * bg[2]};
	const float4 gi = g.grad_image[(size_t)py * c.width + px];
	const float gv[3] = {gi.x, gi.y, gi.z};
	float gsel[3];
#pragma unroll
	for(int k = 0; k < 3; k++)
		gsel[k] = (o32[k] >= 0.0f && o32[k] <= 1.0f) ? gv[k] : 0.0f;
	// pass B: rows [base_row, base_row + n)
	uint32_t base_row = 0;
	if(lane == 0) {
		base_row = atomicAdd(&b.ctrl[kCtrlBwdExtra], n);
		if((uint64_t)b.ctrl[kCtrlBwdRows] + base_row + n > g.rows_cap) {
			atomicOr(&b.ctrl[kCtrlOverflow], kOvfGradRows);
			base_row = 0xffffffffu;
		}
	}
	base_row = __shfl(base_row, 0, 64);
	if(base_row == 0xffffffffu)
		return;
	const size_t row0 = (size_t)b.ctrl[kCtrlBwdRows] + base_row;
	double t = 1.0;
	double q = (double)gsel[0] * (tot.cr + tot.t * (double)bg[0]) +
			   (double)gsel[1] * (tot.cg + tot.t * (double)bg[1]) +
			   (double)gsel[2] * (tot.cb + tot.t * (double)bg[2]);
	for(uint32_t base = 0; base < n; base += 64) {
		BwdSample smp;
		smp.zero();
		const bool ok = base + lane < n && shade_at(base + lane, smp);
		const uint32_t cnt = min(64u, n - base);
		float mine[kGradWords];
#pragma unroll
		for(int k = 0; k < kGradWords; k++)
			mine[k] = 0.0f;
		for(uint32_t j = 0; j < cnt; j++) {
			BwdSample s;
			s.r = __shfl(smp.r, (int)j, 64), s.g = __shfl(smp.g, (int)j, 64);
			s.b = __shfl(smp.b, (int)j, 64), s.a = __shfl(smp.a, (int)j, 64);
			s.b0 = __shfl(smp.b0, (int)j, 64), s.b1 = __shfl(smp.b1, (int)j, 64);
			s.dr = __shfl(smp.dr, (int)j, 64), s.dg = __shfl(smp.dg, (int)j, 64);
			s.db = __shfl(smp.db, (int)j, 64), s.da = __shfl(smp.da, (int)j, 64);
			if(__shfl((int)ok, (int)j, 64)) {
				float out[kGradWords];
				bwd_fragment(t, q, gsel[0], gsel[1], gsel[2], s, out);
				if((uint32_t)lane == j) {
#pragma unroll
					for(int k = 0; k < kGradWords; k++)
						mine[k] = out[k];
				}
			}
		}
		if(base + lane < n) {
			const size_t r = row0 + base + lane;
			g.row_tri[r] = (uint32_t)key_at(base + lane);
#pragma unroll
			for(int k = 0; k < kGradWords; k++)
				g.rows[r * kGradWords + k] = mine[k];
		}
	}
}

__global__ __launch_bounds__(kFallbackThreads) void bwd_fallback_kernel(FrameConst c, FrameBufs b,
																		 BwdBufs g) {
	__shared__ uint64_t s_keys[kFallbackThreads / 64][kFallbackCap];
	if((b.ctrl[kCtrlOverflow] & kOvfBinLists) || !bwd_fits(b, g))
		return;
	const uint32_t nq = min(b.ctrl[kCtrlFallbackCount], b.fallback_cap);
	const uint32_t waves = gridDim.x * (kFallbackThreads / 64);
	const uint32_t w = blockIdx.x * (kFallbackThreads / 64) + (threadIdx.x >> 6);
	for(uint32_t qi = w; qi < nq; qi += waves)
		bwd_fallback_pixel(c, b, g, b.fallback_queue[qi], s_keys[threadIdx.x >> 6]);
}

// Block bookkeeping and the pre-sort of long lists (up to kMidSortCap entries), then the
// queued lists longer than that (one work-group per CU).
static void launch_high_finalize(const FrameConst &c, const FrameBufs &b, hipStream_t stream) {
	hipLaunchKernelGGL(high_finalize_kernel, dim3(c.bin_count * 16), dim3(kSortThreads), 0, stream,
					   c, b);
	hipLaunchKernelGGL(high_finalize_long_kernel, dim3(256), dim3(kSortThreads), 0, stream, c, b);
}

template <bool FP32, int MODE>
static void launch_raster_t(const FrameConst &c, const FrameBufs &b, hipStream_t stream,
							hipEvent_t *ev) {
	constexpr bool TEX = MODE == kShadeTextured;
	hipLaunchKernelGGL(rowgen_low_kernel, dim3(c.bin_count), dim3(kRowgenThreads), 0, stream, c, b);
	if(ev)
		(void)hipEventRecord(ev[0], stream);
	hipLaunchKernelGGL(high_prep_kernel, dim3(1), dim3(1024), 0, stream, c, b);
	hipLaunchKernelGGL(rowgen_high_kernel, dim3(c.bin_count * kRowgenSplit), dim3(kRowgenThreads), 0,
					   stream, c, b);
	launch_high_finalize(c, b, stream);
	if(ev)
		(void)hipEventRecord(ev[1], stream);
	if(c.slice_words == 2) {
		hipLaunchKernelGGL((block_raster_kernel<FP32, 2, MODE>), dim3(c.bin_count * 16), dim3(64), 0,
						   stream, c, b);
		hipLaunchKernelGGL((block_raster_big_kernel<FP32, 2, MODE>), dim3(c.bin_count * 16),
						   dim3(kBigWaves * 64), 0, stream, c, b);
	} else {
		hipLaunchKernelGGL((block_raster_kernel<FP32, 1, MODE>), dim3(c.bin_count * 16), dim3(64), 0,
						   stream, c, b);
		hipLaunchKernelGGL((block_raster_big_kernel<FP32, 1, MODE>), dim3(c.bin_count * 16),
						   dim3(kBigWaves * 64), 0, stream, c, b);
	}
	if(ev)
		(void)hipEventRecord(ev[2], stream);
	hipLaunchKernelGGL((fallback_kernel<FP32, TEX>), dim3(1024), dim3(kFallbackThreads), 0, stream,
					   c, b);
	if(ev)
		(void)hipEventRecord(ev[3], stream);
}

template <bool FP32, int W, int MODE>
static void launch_raster_batch_t(const BatchPlan &p, hipStream_t stream) {
	constexpr bool TEX = MODE == kShadeTextured;
	hipLaunchKernelGGL(rowgen_low_kernel_batch, dim3(p.total[kBsRowgenLow]), dim3(kRowgenThreads), 0,
					   stream, p.frames, p.stage_pre(kBsRowgenLow), p.num_frames);
	hipLaunchKernelGGL(high_prep_kernel_batch, dim3(p.total[kBsHighPrep]), dim3(1024), 0, stream,
					   p.frames, p.stage_pre(kBsHighPrep), p.num_frames);
	hipLaunchKernelGGL(rowgen_high_kernel_batch, dim3(p.total[kBsRowgenHigh]),
					   dim3(kRowgenThreads), 0, stream, p.frames, p.stage_pre(kBsRowgenHigh), p.num_frames);
	hipLaunchKernelGGL(high_finalize_kernel_batch, dim3(p.total[kBsHighFinalize]),
					   dim3(kSortThreads), 0, stream, p.frames, p.stage_pre(kBsHighFinalize), p.num_frames);
	hipLaunchKernelGGL(high_finalize_long_kernel_batch, dim3(p.total[kBsHighFinalizeLong]),
					   dim3(kSortThreads), 0, stream, p.frames, p.stage_pre(kBsHighFinalizeLong),
					   p.num_frames);
	hipLaunchKernelGGL((block_raster_kernel_batch<FP32, W, MODE>), dim3(p.total[kBsBlockRaster]),
					   dim3(64), 0, stream, p.frames, p.stage_pre(kBsBlockRaster), p.num_frames);
	hipLaunchKernelGGL((block_raster_big_kernel_batch<FP32, W, MODE>),
					   dim3(p.total[kBsBlockRasterBig]), dim3(kBigWaves * 64), 0, stream,
					   p.frames, p.stage_pre(kBsBlockRasterBig), p.num_frames);
	hipLaunchKernelGGL((fallback_kernel_batch<FP32, TEX>), dim3(p.total[kBsFallback]),
					   dim3(kFallbackThreads), 0, stream, p.frames, p.stage_pre(kBsFallback), p.num_frames);
}

// Each frame's control words, for the host's overflow check (one copy per batch).
__global__ __launch_bounds__(64) void batch_status_kernel(const FrameDesc *frames, uint32_t n,
														 uint32_t *status) {
	const uint32_t f = blockIdx.x;
	if(f < n && threadIdx.x < (uint32_t)kCtrlWords)
		status[(size_t)f * kCtrlWords + threadIdx.x] = frames[f].b.ctrl[threadIdx.x];
}

// Row generation, sorting, block rasterisation and the exact fallback of every frame of a
// batch: one launch per stage, kernels chosen by the batch's sample format, slice width and
// shading mode; then the frames' control words into p.status.
template <int W>
static void launch_raster_batch_w(const BatchPlan &p, hipStream_t stream) {
	if(p.fp32)
		p.mode == kShadeTextured ? launch_raster_batch_t<true, W, kShadeTextured>(p, stream)
								 : launch_raster_batch_t<true, W, kShadeGeneral>(p, stream);
	else if(p.mode == kShadeTextured)
		launch_raster_batch_t<false, W, kShadeTextured>(p, stream);
	else if(p.mode == kShadeFlat)
		launch_raster_batch_t<false, W, kShadeFlat>(p, stream);
	else if(p.mode == kShadeFlatNarrow)
		launch_raster_batch_t<false, W, kShadeFlatNarrow>(p, stream);
	else
		launch_raster_batch_t<false, W, kShadeGeneral>(p, stream);
}

hipError_t launch_raster_batch(const BatchPlan &p, hipStream_t stream) {
	if(p.slice_words == 2)
		launch_raster_batch_w<2>(p, stream);
	else
		launch_raster_batch_w<1>(p, stream);
	hipLaunchKernelGGL(batch_status_kernel, dim3(p.num_frames), dim3(64), 0, stream, p.frames,
					   p.num_frames, p.status);
	return hipGetLastError();
}

// events (optional, 4): after rowgen_low, after the high-bin row generation and sorting,
// after block rasterisation, after the fallback kernel.
hipError_t launch_raster(const FrameConst &c, const FrameBufs &b, hipStream_t stream,
						 hipEvent_t *events) {
	// frames with albedo textures take kernels built with the texture path (the others carry
	// none of its code or registers); flat-colour frames take the colour-plane records
	const bool fp32 = (c.opts & LUCID_OPT_FP32_SAMPLES) != 0;
	const int mode = frame_shade_mode(c);
	if(fp32)
		mode == kShadeTextured ? launch_raster_t<true, kShadeTextured>(c, b, stream, events)
							   : launch_raster_t<true, kShadeGeneral>(c, b, stream, events);
	else if(mode == kShadeTextured)
		launch_raster_t<false, kShadeTextured>(c, b, stream, events);
	else if(mode == kShadeFlat)
		launch_raster_t<false, kShadeFlat>(c, b, stream, events);
	else if(mode == kShadeFlatNarrow)
		launch_raster_t<false, kShadeFlatNarrow>(c, b, stream, events);
	else
		launch_raster_t<false, kShadeGeneral>(c, b, stream, events);
	return hipGetLastError();
}

// Backward, first part: the forward's row generation and block sorting (fragment lists are
// recomputed, not stored), then the gradient row allocation.
hipError_t launch_backward_prepare(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
								   hipStream_t stream) {
	hipLaunchKernelGGL(rowgen_low_kernel, dim3(c.bin_count), dim3(kRowgenThreads), 0, stream, c, b);
	hipLaunchKernelGGL(high_prep_kernel, dim3(1), dim3(1024), 0, stream, c, b);
	hipLaunchKernelGGL(rowgen_high_kernel, dim3(c.bin_count * kRowgenSplit), dim3(kRowgenThreads), 0,
					   stream, c, b);
	launch_high_finalize(c, b, stream);
	hipLaunchKernelGGL(grad_row_scan_kernel<false>, dim3(1), dim3(1024), 0, stream, c, b, g);
	return hipGetLastError();
}

// max |dL/dpixel| for the gradient scale (pass B and the triangle sums read it); it may run
// on a helper stream once the row scan (which clears it) has run.
hipError_t launch_grad_max(const FrameConst &c, const BwdBufs &g, uint32_t *ctrl,
						   hipStream_t stream) {
	hipLaunchKernelGGL(grad_max_kernel, dim3(256), dim3(256), 0, stream, c, g, ctrl);
	return hipGetLastError();
}

// A frame recording the backward's pass A does so only if its gradient rows, recorded
// order and start states fit the buffers (sized by an earlier frame); otherwise the block
// rasterisers render it and the host grows the buffers from the row scan's counts.
__global__ __launch_bounds__(64) void record_gate_kernel(FrameBufs b, BwdBufs g) {
	if(threadIdx.x == 0)
		b.ctrl[kCtrlRecorded] = (!(b.ctrl[kCtrlOverflow] & kOvfBinLists) && bwd_fits(b, g)) ? 1u : 0u;
}

// Forward frame that also records pass A (fp32 samples): the raster stage of launch_raster
// with the gradient row scan, the gate and the recording pass before the block kernels
// (which then render only what the recording pass does not).
hipError_t launch_raster_recording(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
								   hipStream_t stream, hipEvent_t *ev) {
	hipLaunchKernelGGL(rowgen_low_kernel, dim3(c.bin_count), dim3(kRowgenThreads), 0, stream, c, b);
	if(ev)
		(void)hipEventRecord(ev[0], stream);
	hipLaunchKernelGGL(high_prep_kernel, dim3(1), dim3(1024), 0, stream, c, b);
	hipLaunchKernelGGL(rowgen_high_kernel, dim3(c.bin_count * kRowgenSplit), dim3(kRowgenThreads), 0,
					   stream, c, b);
	launch_high_finalize(c, b, stream);
	hipLaunchKernelGGL(grad_row_scan_kernel<false>, dim3(1), dim3(1024), 0, stream, c, b, g);
	hipLaunchKernelGGL(record_gate_kernel, dim3(1), dim3(64), 0, stream, b, g);
	if(ev)
		(void)hipEventRecord(ev[1], stream);
	if(c.window == 1)
		hipLaunchKernelGGL((bwd_pass_a_kernel<true, 2>), dim3(c.bin_count * 16), dim3(kBigWaves * 64),
						   0, stream, c, b, g);
	else
		hipLaunchKernelGGL((bwd_pass_a_kernel<true, kWindow>), dim3(c.bin_count * 16),
						   dim3(kBigWaves * 64), 0, stream, c, b, g);
	if(c.slice_words == 2) {
		hipLaunchKernelGGL((block_raster_kernel<true, 2, kShadeGeneral>), dim3(c.bin_count * 16), dim3(64), 0,
						   stream, c, b);
		hipLaunchKernelGGL((block_raster_big_kernel<true, 2, kShadeGeneral>), dim3(c.bin_count * 16),
						   dim3(kBigWaves * 64), 0, stream, c, b);
	} else {
		hipLaunchKernelGGL((block_raster_kernel<true, 1, kShadeGeneral>), dim3(c.bin_count * 16), dim3(64), 0,
						   stream, c, b);
		hipLaunchKernelGGL((block_raster_big_kernel<true, 1, kShadeGeneral>), dim3(c.bin_count * 16),
						   dim3(kBigWaves * 64), 0, stream, c, b);
	}
	if(ev)
		(void)hipEventRecord(ev[2], stream);
	hipLaunchKernelGGL((fallback_kernel<true, false>), dim3(1024), dim3(kFallbackThreads), 0, stream,
					   c, b);
	if(ev)
		(void)hipEventRecord(ev[3], stream);
	return hipGetLastError();
}

// Backward of a frame that recorded pass A: fresh backward control words (the frame's row
// scan results and exact-fallback queue stay), then the start states and the exact
// per-pixel adjoints.
__global__ __launch_bounds__(64) void recorded_begin_kernel(uint32_t *ctrl) {
	if(threadIdx.x == 0) {
		ctrl[kCtrlOverflow] = 0;
		ctrl[kCtrlSpillCursor] = 0;
		ctrl[kCtrlBwdExtra] = 0;
		ctrl[kCtrlBwdGradMax] = 0;
		ctrl[kCtrlTopoChanged] = 0;
	}
}

hipError_t launch_backward_recorded(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
									hipStream_t stream) {
	hipLaunchKernelGGL(recorded_begin_kernel, dim3(1), dim3(64), 0, stream, b.ctrl);
	return hipGetLastError();
}

hipError_t launch_backward_recorded_order(const FrameConst &c, const FrameBufs &b,
										  const BwdBufs &g, hipStream_t stream) {
	hipLaunchKernelGGL(bwd_start_kernel, dim3(c.bin_count * 16), dim3(kBigWaves * 64), 0, stream, c,
					   b, g);
	hipLaunchKernelGGL(bwd_fallback_kernel, dim3(1024), dim3(kFallbackThreads), 0, stream, c, b, g);
	return hipGetLastError();
}

// Backward, first part when the last finished frame's setup, bins and block lists are
// reused: fresh control words, the exact-pixel queue of the too-long blocks and the
// gradient row allocation, in one kernel.
hipError_t launch_backward_reuse(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
								 hipStream_t stream) {
	hipLaunchKernelGGL(grad_row_scan_kernel<true>, dim3(1), dim3(1024), 0, stream, c, b, g);
	return hipGetLastError();
}

// Backward, second part: the recorded blend order and start states (pass A, which also
// fills the triangle of every block gradient row), then the exact per-pixel adjoints (their
// rows follow the block rows). After these every row's triangle is known.
hipError_t launch_backward_order(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
								 hipStream_t stream) {
	if(c.window == 1) // narrow-window tuning: two slots
		hipLaunchKernelGGL((bwd_pass_a_kernel<false, 2>), dim3(c.bin_count * 16),
						   dim3(kBigWaves * 64), 0, stream, c, b, g);
	else
		hipLaunchKernelGGL((bwd_pass_a_kernel<false, kWindow>), dim3(c.bin_count * 16),
						   dim3(kBigWaves * 64), 0, stream, c, b, g);
	hipLaunchKernelGGL(bwd_fallback_kernel, dim3(1024), dim3(kFallbackThreads), 0, stream, c, b, g);
	return hipGetLastError();
}

// Backward, third part: the block gradient rows (pass B replays the recorded order).
hipError_t launch_backward_rows(const FrameConst &c, const FrameBufs &b, const BwdBufs &g,
								hipStream_t stream) {
	hipLaunchKernelGGL(bwd_pass_b_kernel, dim3(c.bin_count * 16 * kBigWaves), dim3(64), 0, stream, c,
					   b, g);
	return hipGetLastError();
}

} // namespace lucid
