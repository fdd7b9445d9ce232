// vv_device.hpp — device-side sampling and arithmetic shared by the LIC,
// ray-cast and adjoint kernels (gfx950, wave64).
//
// Texture semantics reproduce OpenGL linear filtering as the reference's
// shaders see it (VV/shader/inc_lic.glsl, raycast_lic3d_fragment.glsl):
// for a coordinate u over N texels, x = fma(u, N, -0.5), i0 = floor(x),
// f = x - i0, then texels i0 and i0+1 after the wrap.  Values blend as
// nested lerps (x, then y, then z) with lerp(a,b,t) = fma(t, b-a, a).
// UNORM8 textures blend raw byte values and scale the result by 1/255.
//
// Wrap modes are folded into padded layouts so no sample needs a branch:
//   CLAMP_TO_EDGE volumes are stored (n+2) per axis with replicated faces;
//     corner index p0 = clamp(floor(x), -1, n-1) + 1, p1 = p0 + 1.
//   REPEAT volumes are stored (n+1) per axis, plane n a copy of plane 0;
//     p0 = floor(x) mod n, p1 = p0 + 1.
// Both give exactly the texels GL would fetch, so results are bit-identical
// to the unpadded restatement in the oracle.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace vv {

constexpr float kInv255 = (float)(1.0 / 255.0);

// ---- scalar helpers ------------------------------------------------------
__device__ __forceinline__ float lerpf(float a, float b, float t) { return fmaf(t, b - a, a); }
__device__ __forceinline__ float clamp01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

// padded CLAMP_TO_EDGE corner index along one axis
__device__ __forceinline__ int tc_clamp(float u, int n, float& f)
{
    float x = fmaf(u, (float)n, -0.5f);
    float fl = floorf(x);
    f = x - fl;
    fl = fminf(fmaxf(fl, -1.0f), (float)(n - 1));
    return (int)fl + 1;
}

// padded REPEAT corner index along one axis
__device__ __forceinline__ int tc_repeat(float u, int n, float& f)
{
    float x = fmaf(u, (float)n, -0.5f);
    float fl = floorf(x);
    f = x - fl;
    fl = fminf(fmaxf(fl, -1.0e9f), 1.0e9f);
    int i = (int)fl;
    if ((unsigned)i >= (unsigned)n) {
        i %= n;
        if (i < 0) i += n;
    }
    return i;
}

// A padded volume: element pointer, logical size and element strides.
template <typename T>
struct Vol {
    const T* p;
    int n[3];
    int sy, sz;
};

struct Cell {
    int base;          // element index of corner (x0,y0,z0)
    float fx, fy, fz;
};

template <typename T>
__device__ __forceinline__ Cell cell_clamp(const Vol<T>& v, float x, float y, float z)
{
    Cell c;
    int i = tc_clamp(x, v.n[0], c.fx);
    int j = tc_clamp(y, v.n[1], c.fy);
    int k = tc_clamp(z, v.n[2], c.fz);
    c.base = k * v.sz + j * v.sy + i;
    return c;
}

template <typename T>
__device__ __forceinline__ Cell cell_repeat(const Vol<T>& v, float x, float y, float z)
{
    Cell c;
    int i = tc_repeat(x, v.n[0], c.fx);
    int j = tc_repeat(y, v.n[1], c.fy);
    int k = tc_repeat(z, v.n[2], c.fz);
    c.base = k * v.sz + j * v.sy + i;
    return c;
}

__device__ __forceinline__ float tri(float v000, float v100, float v010, float v110,
                                     float v001, float v101, float v011, float v111,
                                     const Cell& c)
{
    float c00 = lerpf(v000, v100, c.fx);
    float c10 = lerpf(v010, v110, c.fx);
    float c01 = lerpf(v001, v101, c.fx);
    float c11 = lerpf(v011, v111, c.fx);
    float c0 = lerpf(c00, c10, c.fy);
    float c1 = lerpf(c01, c11, c.fy);
    return lerpf(c0, c1, c.fz);
}

// ---- element decoders ----------------------------------------------------
// Field texel types: float4 (RGBA32F), half4_t (RGBA16F) and u8x4_t (RGBA8,
// the reference's UNORM8 field of fillTexDataChar[Interp]).  ld4/ld1 return
// raw channel values (byte values for u8x4_t); FieldScale turns a blend of
// them into the texture value (x 1/255 for UNORM8, as for every UNORM8
// texture here).
struct half4_t {
    __half x, y, z, w;
};

struct u8x4_t {
    uint8_t x, y, z, w;
};

template <typename FT>
struct FieldScale {
    __device__ static __forceinline__ float apply(float v) { return v; }
};
template <>
struct FieldScale<u8x4_t> {
    __device__ static __forceinline__ float apply(float v) { return v * kInv255; }
};

__device__ __forceinline__ float4 ld4(const float4* p, int i) { return p[i]; }
__device__ __forceinline__ float4 ld4(const half4_t* p, int i)
{
    // one 8-byte load, exact half -> float conversion
    uint2 u = reinterpret_cast<const uint2*>(p)[i];
    float4 r;
    r.x = __half2float(__ushort_as_half((unsigned short)(u.x & 0xffffu)));
    r.y = __half2float(__ushort_as_half((unsigned short)(u.x >> 16)));
    r.z = __half2float(__ushort_as_half((unsigned short)(u.y & 0xffffu)));
    r.w = __half2float(__ushort_as_half((unsigned short)(u.y >> 16)));
    return r;
}

__device__ __forceinline__ float4 ld4(const u8x4_t* p, int i)
{
    const uint32_t u = reinterpret_cast<const uint32_t*>(p)[i];
    return make_float4((float)(u & 0xffu), (float)((u >> 8) & 0xffu), (float)((u >> 16) & 0xffu),
                       (float)(u >> 24));
}

__device__ __forceinline__ float ld1(const float4* p, int i, int ch)
{
    return reinterpret_cast<const float*>(p)[4 * i + ch];
}
__device__ __forceinline__ float ld1(const half4_t* p, int i, int ch)
{
    unsigned short h = reinterpret_cast<const unsigned short*>(p)[4 * i + ch];
    return __half2float(__ushort_as_half(h));
}
__device__ __forceinline__ float ld1(const u8x4_t* p, int i, int ch)
{
    return (float)reinterpret_cast<const uint8_t*>(p)[4 * i + ch];
}

// rgb of a 4-channel field at a CLAMP cell
template <typename FT>
__device__ __forceinline__ float3 tri_rgb(const Vol<FT>& v, const Cell& c)
{
    const int b = c.base, sy = v.sy, sz = v.sz;
    float4 a000 = ld4(v.p, b);
    float4 a100 = ld4(v.p, b + 1);
    float4 a010 = ld4(v.p, b + sy);
    float4 a110 = ld4(v.p, b + sy + 1);
    float4 a001 = ld4(v.p, b + sz);
    float4 a101 = ld4(v.p, b + sz + 1);
    float4 a011 = ld4(v.p, b + sz + sy);
    float4 a111 = ld4(v.p, b + sz + sy + 1);
    float3 r;
    r.x = FieldScale<FT>::apply(tri(a000.x, a100.x, a010.x, a110.x, a001.x, a101.x, a011.x, a111.x, c));
    r.y = FieldScale<FT>::apply(tri(a000.y, a100.y, a010.y, a110.y, a001.y, a101.y, a011.y, a111.y, c));
    r.z = FieldScale<FT>::apply(tri(a000.z, a100.z, a010.z, a110.z, a001.z, a101.z, a011.z, a111.z, c));
    return r;
}

// all four channels of a field (direction and magnitude)
template <typename FT>
__device__ __forceinline__ float4 tri_rgba(const Vol<FT>& v, const Cell& c)
{
    const int b = c.base, sy = v.sy, sz = v.sz;
    float4 a000 = ld4(v.p, b);
    float4 a100 = ld4(v.p, b + 1);
    float4 a010 = ld4(v.p, b + sy);
    float4 a110 = ld4(v.p, b + sy + 1);
    float4 a001 = ld4(v.p, b + sz);
    float4 a101 = ld4(v.p, b + sz + 1);
    float4 a011 = ld4(v.p, b + sz + sy);
    float4 a111 = ld4(v.p, b + sz + sy + 1);
    float4 r;
    r.x = FieldScale<FT>::apply(tri(a000.x, a100.x, a010.x, a110.x, a001.x, a101.x, a011.x, a111.x, c));
    r.y = FieldScale<FT>::apply(tri(a000.y, a100.y, a010.y, a110.y, a001.y, a101.y, a011.y, a111.y, c));
    r.z = FieldScale<FT>::apply(tri(a000.z, a100.z, a010.z, a110.z, a001.z, a101.z, a011.z, a111.z, c));
    r.w = FieldScale<FT>::apply(tri(a000.w, a100.w, a010.w, a110.w, a001.w, a101.w, a011.w, a111.w, c));
    return r;
}

// one channel of a 4-channel field
template <typename FT>
__device__ __forceinline__ float tri_ch(const Vol<FT>& v, const Cell& c, int ch)
{
    const int b = c.base, sy = v.sy, sz = v.sz;
    return FieldScale<FT>::apply(tri(ld1(v.p, b, ch), ld1(v.p, b + 1, ch), ld1(v.p, b + sy, ch),
                                     ld1(v.p, b + sy + 1, ch), ld1(v.p, b + sz, ch), ld1(v.p, b + sz + 1, ch),
                                     ld1(v.p, b + sz + sy, ch), ld1(v.p, b + sz + sy + 1, ch), c));
}

// Host-side dispatch over the field texel type: fn(TypeTag<FT>{}) with FT
// float4 (type 0), half4_t (1) or u8x4_t (2)
template <typename T>
struct TypeTag {
    using type = T;
};

template <typename Fn>
inline hipError_t with_field_type(int type, Fn&& fn)
{
    switch (type) {
    case 0: fn(TypeTag<float4>{}); break;
    case 1: fn(TypeTag<half4_t>{}); break;
    case 2: fn(TypeTag<u8x4_t>{}); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// scalar float volume
__device__ __forceinline__ float tri_f(const Vol<float>& v, const Cell& c)
{
    const int b = c.base, sy = v.sy, sz = v.sz;
    const float* p = v.p;
    return tri(p[b], p[b + 1], p[b + sy], p[b + sy + 1], p[b + sz], p[b + sz + 1],
               p[b + sz + sy], p[b + sz + sy + 1], c);
}

// 4-channel float volume
__device__ __forceinline__ float4 tri_f4(const Vol<float4>& v, const Cell& c)
{
    const int b = c.base, sy = v.sy, sz = v.sz;
    const float4* p = v.p;
    float4 a000 = p[b], a100 = p[b + 1], a010 = p[b + sy], a110 = p[b + sy + 1];
    float4 a001 = p[b + sz], a101 = p[b + sz + 1], a011 = p[b + sz + sy], a111 = p[b + sz + sy + 1];
    float4 r;
    r.x = tri(a000.x, a100.x, a010.x, a110.x, a001.x, a101.x, a011.x, a111.x, c);
    r.y = tri(a000.y, a100.y, a010.y, a110.y, a001.y, a101.y, a011.y, a111.y, c);
    r.z = tri(a000.z, a100.z, a010.z, a110.z, a001.z, a101.z, a011.z, a111.z, c);
    r.w = tri(a000.w, a100.w, a010.w, a110.w, a001.w, a101.w, a011.w, a111.w, c);
    return r;
}

// UNORM8 scalar volume, returns the blended byte value (0..255)
__device__ __forceinline__ float tri_u8(const Vol<uint8_t>& v, const Cell& c)
{
    const int b = c.base, sy = v.sy, sz = v.sz;
    const uint8_t* p = v.p;
    return tri((float)p[b], (float)p[b + 1], (float)p[b + sy], (float)p[b + sy + 1],
               (float)p[b + sz], (float)p[b + sz + 1], (float)p[b + sz + sy],
               (float)p[b + sz + sy + 1], c);
}

// UNORM8 RGBA volume, returns blended byte values (0..255)
__device__ __forceinline__ float4 tri_u8x4(const Vol<uchar4>& v, const Cell& c)
{
    const int b = c.base, sy = v.sy, sz = v.sz;
    const uchar4* p = v.p;
    uchar4 a000 = p[b], a100 = p[b + 1], a010 = p[b + sy], a110 = p[b + sy + 1];
    uchar4 a001 = p[b + sz], a101 = p[b + sz + 1], a011 = p[b + sz + sy], a111 = p[b + sz + sy + 1];
    float4 r;
    r.x = tri(a000.x, a100.x, a010.x, a110.x, a001.x, a101.x, a011.x, a111.x, c);
    r.y = tri(a000.y, a100.y, a010.y, a110.y, a001.y, a101.y, a011.y, a111.y, c);
    r.z = tri(a000.z, a100.z, a010.z, a110.z, a001.z, a101.z, a011.z, a111.z, c);
    r.w = tri(a000.w, a100.w, a010.w, a110.w, a001.w, a101.w, a011.w, a111.w, c);
    return r;
}

// 1D CLAMP_TO_EDGE lookup in an LDS table of n entries with nc channels
__device__ __forceinline__ float lut1d(const float* t, int n, int nc, int ch, float u)
{
    float x = fmaf(u, (float)n, -0.5f);
    float fl = floorf(x);
    float f = x - fl;
    fl = fminf(fmaxf(fl, -1.0f), (float)(n - 1));
    int i = (int)fl;
    int i0 = i < 0 ? 0 : i;
    int i1 = i + 1 > n - 1 ? n - 1 : i + 1;
    return lerpf(t[i0 * nc + ch], t[i1 * nc + ch], f);
}

// 2D LINEAR + CLAMP_TO_EDGE lookup in a 256x256 table (row v, column u) with
// nc channels.  The coordinate is first limited to [0,1]: no change for
// finite values (the edge texels saturate there), and +-inf / NaN go to an
// edge deterministically (fmaxf(NaN, 0) = 0).
__device__ __forceinline__ float lut2d(const float* t, int nc, int ch, float u, float v)
{
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    v = fminf(fmaxf(v, 0.0f), 1.0f);
    const float x = fmaf(u, 256.0f, -0.5f), y = fmaf(v, 256.0f, -0.5f);
    const float flx = floorf(x), fly = floorf(y);
    const float fx = x - flx, fy = y - fly;
    const int i = (int)flx, j = (int)fly;  // -1 .. 255
    const int i0 = i < 0 ? 0 : i, i1 = i + 1 > 255 ? 255 : i + 1;
    const int j0 = j < 0 ? 0 : j, j1 = j + 1 > 255 ? 255 : j + 1;
    const float a = lerpf(t[(j0 * 256 + i0) * nc + ch], t[(j0 * 256 + i1) * nc + ch], fx);
    const float b = lerpf(t[(j1 * 256 + i0) * nc + ch], t[(j1 * 256 + i1) * nc + ch], fx);
    return lerpf(a, b, fy);
}

}  // namespace vv
