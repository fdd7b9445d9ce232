// vv_lic.hip — field encode, noise, noise-gradient and LIC-volume kernels.
//
//   field encode    VV/dataset.cpp:533-635 (fillTexDataFloatInterp)
//   noise gradients VV/gradient.cpp:190-532 (computeGradients SOBEL,
//                   filterGradients, quantize8)
//   LIC volume      VV/shader/lic3d_volume_fragment.glsl:2-20 with
//                   computeLIC / singleLICstep / freqSampling[Grad]
//                   (VV/shader/inc_lic.glsl:61-202)
//
// All floating-point expressions keep the operation order of the
// reference (left to right, no contraction: the library is built with
// -ffp-contract=off) and use fmaf only where the sampling rules say so.
#include "vv_device.hpp"
#include "vv_internal.hpp"

#include <cstring>
#include <type_traits>

namespace vv {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ float wave_max(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_i64(long long* p, long long v, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// |v| of the time-interpolated vector, 0 below EPS (VV/mmath.h:43)
__device__ __forceinline__ float field_len(const float* v0, const float* v1, float t, long long i,
                                           float tv[3])
{
    for (int c = 0; c < 3; ++c) {
        float a = v0[3 * i + c], b = v1[3 * i + c];
        tv[c] = a + t * (b - a);
    }
    float len = sqrtf(tv[0] * tv[0] + tv[1] * tv[1] + tv[2] * tv[2]);
    return len < 1e-5f ? 0.0f : len;
}

__global__ void __launch_bounds__(kThreads) k_field_maxlen(const float* v0, const float* v1,
                                                           float t, long long n, uint32_t* maxbits)
{
    __shared__ float red[kThreads / 64];
    float m = 0.0f;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        float tv[3];
        m = fmaxf(m, field_len(v0, v1, t, i, tv));
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kThreads / 64; ++w) m = fmaxf(m, red[w]);
        // non-negative floats order like their bit patterns
        atomicMax(maxbits, __float_as_uint(m));
    }
}

template <typename T>
__device__ __forceinline__ void store_enc(T* out, long long i, float r, float g, float b, float a);

template <>
__device__ __forceinline__ void store_enc<float4>(float4* out, long long i, float r, float g, float b,
                                                  float a)
{
    out[i] = make_float4(r, g, b, a);
}

template <>
__device__ __forceinline__ void store_enc<half4_t>(half4_t* out, long long i, float r, float g,
                                                   float b, float a)
{
    uint2 u;
    u.x = (uint32_t)__half_as_ushort(__float2half(r)) | ((uint32_t)__half_as_ushort(__float2half(g)) << 16);
    u.y = (uint32_t)__half_as_ushort(__float2half(b)) | ((uint32_t)__half_as_ushort(__float2half(a)) << 16);
    reinterpret_cast<uint2*>(out)[i] = u;
}

// UNORM8 direction byte of fillTexDataChar[Interp] (VV/dataset.cpp:760-762):
// (unsigned char)(q*127.0 + 128.0) in double.  127q is exact in double; the
// sum is exact too except for -2^-47 <= 127q < 0, where it rounds to 128.0
// (the double just below 128 is 128 - 2^-46; a tie goes to the even 128).
// So the byte is floor(127q + 128), but 128 for those q.  Without double
// arithmetic: k = floor of 127q + 128 rounded once to float; when the exact
// value lies just below k, fma(q, 127, 128 - k) (rounded once) is negative,
// and fma(q, 127, 2^-47) >= 0 tells the q that round to 128.
__device__ __forceinline__ uint32_t unorm8_dir_byte(float q)
{
    const float k = floorf(fmaf(q, 127.0f, 128.0f));
    const float r = fmaf(q, 127.0f, 128.0f - k);
    const bool to128 = k == 128.0f && fmaf(q, 127.0f, 0x1p-47f) >= 0.0f;
    return (uint32_t)(r < 0.0f && !to128 ? k - 1.0f : k);
}

// writes the CLAMP-padded encoded field; padded faces replicate the edge
template <typename T>
__global__ void __launch_bounds__(kThreads) k_field_encode(const float* v0, const float* v1, float t,
                                                           int nx, int ny, int nz,
                                                           const uint32_t* maxbits, int raw, T* out)
{
    const int px = nx + 2, py = ny + 2, pz = nz + 2;
    const long long np = (long long)px * py * pz;
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    int x = (int)(q % px), y = (int)((q / px) % py), z = (int)(q / ((long long)px * py));
    x = min(max(x - 1, 0), nx - 1);
    y = min(max(y - 1, 0), ny - 1);
    z = min(max(z - 1, 0), nz - 1);
    long long i = ((long long)z * ny + y) * nx + x;
    float tv[3];
    float len = field_len(v0, v1, t, i, tv);
    if constexpr (std::is_same<T, u8x4_t>::value) {
        // fillTexDataCharInterp (VV/dataset.cpp:704-823), float source branch
        uint32_t rgb = 128u | (128u << 8) | (128u << 16);
        if (len != 0.0f)
            rgb = unorm8_dir_byte(tv[0] / len) | (unorm8_dir_byte(tv[1] / len) << 8) |
                  (unorm8_dir_byte(tv[2] / len) << 16);
        const float maxlen = __uint_as_float(*maxbits);
        float m = maxlen > 0.0f ? len / maxlen * 255.0f : 0.0f;
        m = m > 255.0f ? 255.0f : (m < 0.0f ? 0.0f : m);
        reinterpret_cast<uint32_t*>(out)[q] = rgb | ((uint32_t)m << 24);
    } else {
        float r, g, b;
        if (len == 0.0f) {
            r = g = b = 0.5f;
        } else {
            r = 0.5f * tv[0] / len + 0.5f;
            g = 0.5f * tv[1] / len + 0.5f;
            b = 0.5f * tv[2] / len + 0.5f;
        }
        float a = len;  // raw: fillTexDataFloat keeps |v| (VV/dataset.cpp:512-513)
        if (!raw) {
            float maxlen = __uint_as_float(*maxbits);
            a = maxlen > 0.0f ? len / maxlen : 0.0f;
            a = a > 1.0f ? 1.0f : (a < 0.0f ? 0.0f : a);
        }
        store_enc<T>(out, q, r, g, b, a);
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_field_decode(const T* field, int nx, int ny, int nz,
                                                           float4* out)
{
    long long n = (long long)nx * ny * nz;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long long)nx * ny));
    long long q = ((long long)(z + 1) * (ny + 2) + (y + 1)) * (nx + 2) + (x + 1);
    const float4 v = ld4(field, (int)q);
    out[i] = make_float4(FieldScale<T>::apply(v.x), FieldScale<T>::apply(v.y), FieldScale<T>::apply(v.z),
                         FieldScale<T>::apply(v.w));
}

// LIC copy of the field for equal field and noise grids: every CLAMP-padded
// texel p keeps its rgb and carries in channel 3 the noise byte of REPEAT
// texel (p - 1) mod n per axis, a value each texel type holds exactly
template <typename T>
__global__ void __launch_bounds__(kThreads) k_pack_lic_field(const T* field, const uint8_t* noise, int nx, int ny,
                                                             int nz, T* out)
{
    const int px = nx + 2, py = ny + 2, pz = nz + 2;
    const long long np = (long long)px * py * pz;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const int x = (int)(q % px), y = (int)((q / px) % py), z = (int)(q / ((long long)px * py));
    const int ix = (x + nx - 1) % nx, iy = (y + ny - 1) % ny, iz = (z + nz - 1) % nz;
    const uint32_t b = noise[((long long)iz * (ny + 1) + iy) * (nx + 1) + ix];
    if constexpr (std::is_same<T, float4>::value) {
        float4 v = field[q];
        v.w = (float)b;
        out[q] = v;
    } else if constexpr (std::is_same<T, half4_t>::value) {
        uint2 u = reinterpret_cast<const uint2*>(field)[q];
        u.y = (u.y & 0xffffu) | ((uint32_t)__half_as_ushort(__float2half((float)b)) << 16);
        reinterpret_cast<uint2*>(out)[q] = u;
    } else {
        const uint32_t u = reinterpret_cast<const uint32_t*>(field)[q];
        reinterpret_cast<uint32_t*>(out)[q] = (u & 0x00ffffffu) | (b << 24);
    }
}

// the float4 LIC copy reordered into 8^3 bricks (brick_counts): texel (x, y,
// z) of the padded copy to element ((bz * nby + by) * nbx + bx) * 512 +
// (lz * 8 + ly) * 8 + lx; elements of bricks past the padded volume are not
// written (no fill reads them)
__global__ void __launch_bounds__(kThreads) k_pack_lic_field_brick(const float4* copy, int px, int py, int pz,
                                                                   int nbx, int nby, float4* out)
{
    const long long np = (long long)px * py * pz;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const int x = (int)(q % px), y = (int)((q / px) % py), z = (int)(q / ((long long)px * py));
    const long long e = (((long long)(z >> 3) * nby + (y >> 3)) * nbx + (x >> 3)) * 512 +
                        (((z & 7) * 8 + (y & 7)) * 8 + (x & 7));
    out[e] = copy[q];
}

// ---- noise -----------------------------------------------------------------
// noise[i] = (i+1)-th output of SplitMix64 seeded with `seed`
__device__ __forceinline__ uint64_t splitmix64_at(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + (i + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void __launch_bounds__(kThreads) k_noise_gen(uint64_t seed, int kind, uint64_t thr,
                                                        long long n, uint8_t* out)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t u = splitmix64_at(seed, (uint64_t)i);
    out[i] = kind == 0 ? (((u >> 32) < thr) ? 255 : 0) : (uint8_t)(u >> 56);
}

// computeGradients, SOBEL == 1 (VV/gradient.cpp:199-307)
__constant__ signed char c_sobel[3][3][3][3] = {
    {{{-1, -3, -1}, {-3, -6, -3}, {-1, -3, -1}},
     {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}},
     {{1, 3, 1}, {3, 6, 3}, {1, 3, 1}}},
    {{{-1, -3, -1}, {0, 0, 0}, {1, 3, 1}},
     {{-3, -6, -3}, {0, 0, 0}, {3, 6, 3}},
     {{-1, -3, -1}, {0, 0, 0}, {1, 3, 1}}},
    {{{-1, 0, 1}, {-3, 0, 3}, {-1, 0, 1}},
     {{-3, 0, 3}, {-6, 0, 6}, {-3, 0, 3}},
     {{-1, 0, 1}, {-3, 0, 3}, {-1, 0, 1}}}};

__global__ void __launch_bounds__(kThreads) k_noise_grad_raw(const uint8_t* v, int nx, int ny, int nz,
                                                             float sd0, float sd1, float sd2, float* g)
{
    long long n = (long long)nx * ny * nz;
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    int x = (int)(q % nx), y = (int)((q / nx) % ny), z = (int)(q / ((long long)nx * ny));
    auto vox = [&](int a, int b, int c) -> float {
        return (float)v[((long long)c * ny + b) * nx + a];
    };
    float sd[3] = {sd0, sd1, sd2};
    float* gp = g + 3 * q;
    if (x > 0 && x < nx - 1 && y > 0 && y < ny - 1 && z > 0 && z < nz - 1) {
        for (int d = 0; d < 3; ++d) {
            float acc = 0.0f;
            for (int i = -1; i < 2; ++i)
                for (int j = -1; j < 2; ++j)
                    for (int k = -1; k < 2; ++k)
                        acc += (float)c_sobel[d][i + 1][j + 1][k + 1] * vox(x + i, y + j, z + k);
            gp[d] = acc / (2.0f * sd[d]);
        }
    } else {
        gp[0] = (x < 1) ? (vox(x + 1, y, z) - vox(x, y, z)) / sd[0] : (vox(x, y, z) - vox(x - 1, y, z)) / sd[0];
        gp[1] = (y < 1) ? (vox(x, y + 1, z) - vox(x, y, z)) / sd[1] : (vox(x, y, z) - vox(x, y - 1, z)) / sd[1];
        gp[2] = (z < 1) ? (vox(x, y, z + 1) - vox(x, y, z)) / sd[2] : (vox(x, y, z) - vox(x, y, z - 1)) / sd[2];
    }
}

struct W27 {
    float w[27];
};

// quantize8 (VV/gradient.cpp:462-478): (unsigned char)(((double)f + 1) / 2 * 255)
// in float arithmetic.  The double expression equals floor(127.5 (f + 1))
// for every float f (no float is within its rounding of an integer
// boundary).  The float fma's rounding can cross a boundary, so the floor is
// corrected by exact sign tests: fma(f, 255, 255 - 2k) has the sign of
// 255 f + 255 - 2k (a nonzero exact value is a multiple of 2^-149, so it
// never rounds to zero).  Compared with the double expression for every
// float in [-1 - 4 ulp, 1 + 4 ulp] on the CPU: no mismatch.
__device__ __forceinline__ uint8_t quantize8(float f)
{
    float k = floorf(fmaf(f, 127.5f, 127.5f));
    if (fmaf(f, 255.0f, 255.0f - 2.0f * k) < 0.0f)
        k -= 1.0f;
    else if (fmaf(f, 255.0f, 255.0f - 2.0f * (k + 1.0f)) >= 0.0f)
        k += 1.0f;
    return (uint8_t)fminf(fmaxf(k, 0.0f), 255.0f);
}

// filterGradients (VV/gradient.cpp:377-459: window [-fw, fw-2], fw shrinking
// at the border) followed by quantize8 (:462-478)
__global__ void __launch_bounds__(kThreads) k_noise_grad_filter(const float* g, int nx, int ny, int nz,
                                                                W27 fw27, uint8_t* rgb)
{
    long long n = (long long)nx * ny * nz;
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    int x = (int)(q % nx), y = (int)((q / nx) % ny), z = (int)(q / ((long long)nx * ny));
    int bx = min(x, nx - x - 1), by = min(y, ny - y - 1), bz = min(z, nz - z - 1);
    int w = min(2, min(min(bx, by), bz));
    float f[3];
    for (int c = 0; c < 3; ++c) {
        float acc = 0.0f;
        for (int k = -w; k < w - 1; ++k)
            for (int j = -w; j < w - 1; ++j)
                for (int i = -w; i < w - 1; ++i) {
                    long long o = 3 * (((long long)(z + k) * ny + (y + j)) * nx + (x + i)) + c;
                    acc += fw27.w[((w + k) * 3 + (w + j)) * 3 + (w + i)] * g[o];
                }
        f[c] = acc;
    }
    float len = sqrtf(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (len < 1e-5f) {
        f[0] = f[1] = f[2] = 0.0f;
    } else {
        f[0] /= len;
        f[1] /= len;
        f[2] /= len;
    }
    for (int c = 0; c < 3; ++c) rgb[3 * q + c] = quantize8(f[c]);
}

// raw noise (+ quantised gradients) -> REPEAT-padded texture
__global__ void __launch_bounds__(kThreads) k_noise_tex(const uint8_t* noise, const uint8_t* rgb,
                                                        int nx, int ny, int nz, void* tex, int nc)
{
    const int px = nx + 1, py = ny + 1, pz = nz + 1;
    long long np = (long long)px * py * pz;
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    int x = (int)(q % px), y = (int)((q / px) % py), z = (int)(q / ((long long)px * py));
    if (x == nx) x = 0;
    if (y == ny) y = 0;
    if (z == nz) z = 0;
    long long i = ((long long)z * ny + y) * nx + x;
    if (nc == 1) {
        reinterpret_cast<uint8_t*>(tex)[q] = noise[i];
    } else {
        uchar4 u;
        u.x = rgb[3 * i];
        u.y = rgb[3 * i + 1];
        u.z = rgb[3 * i + 2];
        u.w = noise[i];
        reinterpret_cast<uchar4*>(tex)[q] = u;
    }
}

__global__ void __launch_bounds__(kThreads) k_noise_untex(const void* tex, int nc, int nx, int ny, int nz,
                                                          uint8_t* out)
{
    long long n = (long long)nx * ny * nz;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long long)nx * ny));
    long long q = ((long long)z * (ny + 1) + y) * (nx + 1) + x;
    if (nc == 1) {
        out[i] = reinterpret_cast<const uint8_t*>(tex)[q];
    } else {
        uchar4 u = reinterpret_cast<const uchar4*>(tex)[q];
        out[4 * i] = u.x;
        out[4 * i + 1] = u.y;
        out[4 * i + 2] = u.z;
        out[4 * i + 3] = u.w;
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_pad_clamp(const T* src, int nx, int ny, int nz, T* dst)
{
    const int px = nx + 2, py = ny + 2, pz = nz + 2;
    long long np = (long long)px * py * pz;
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    int x = (int)(q % px), y = (int)((q / px) % py), z = (int)(q / ((long long)px * py));
    x = min(max(x - 1, 0), nx - 1);
    y = min(max(y - 1, 0), ny - 1);
    z = min(max(z - 1, 0), nz - 1);
    dst[q] = src[((long long)z * ny + y) * nx + x];
}

__global__ void __launch_bounds__(kThreads) k_pad_repeat_f32(const float* src, int nc, int nx, int ny,
                                                             int nz, float* dst)
{
    const int px = nx + 1, py = ny + 1, pz = nz + 1;
    long long np = (long long)px * py * pz;
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    int x = (int)(q % px), y = (int)((q / px) % py), z = (int)(q / ((long long)px * py));
    if (x == nx) x = 0;
    if (y == ny) y = 0;
    if (z == nz) z = 0;
    long long i = ((long long)z * ny + y) * nx + x;
    for (int c = 0; c < nc; ++c) dst[q * nc + c] = src[i * nc + c];
}

__global__ void __launch_bounds__(kThreads) k_unpad_repeat_f32(const float* src, int nc, int nx, int ny,
                                                               int nz, float* dst)
{
    long long n = (long long)nx * ny * nz;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long long)nx * ny));
    long long q = ((long long)z * (ny + 1) + y) * (nx + 1) + x;
    for (int c = 0; c < nc; ++c) dst[i * nc + c] = src[q * nc + c];
}

// fills plane n of each axis of a REPEAT-padded volume from plane 0
__global__ void __launch_bounds__(kThreads) k_repeat_halo_f32(float* vol, int nc, int nx, int ny, int nz)
{
    const int px = nx + 1, py = ny + 1, pz = nz + 1;
    long long np = (long long)px * py * pz;
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    int x = (int)(q % px), y = (int)((q / px) % py), z = (int)(q / ((long long)px * py));
    if (x < nx && y < ny && z < nz) return;
    int sx = x == nx ? 0 : x, sy = y == ny ? 0 : y, sz = z == nz ? 0 : z;
    long long s = ((long long)sz * py + sy) * px + sx;
    for (int c = 0; c < nc; ++c) vol[q * nc + c] = vol[s * nc + c];
}

// ---- LIC ---------------------------------------------------------------------
#include "vv_lic_dev.hpp"

template <typename FT, bool SAME, bool BRICK = false>
__device__ __forceinline__ void lic_cached_tile(const LicArgs& a, int tile)
{
    int bx, by, bz;
    tile_coords(a, tile, bx, by, bz);
    const int t = threadIdx.x;
    const int x = bx * a.tx + t % a.tx;
    const int y = by * a.ty + (t / a.tx) % a.ty;
    const int z = bz * a.tz + t / (a.tx * a.ty);
    const int nx = a.field.n[0], ny = a.field.n[1], nz = a.field.n[2];
    if (x >= nx || y >= ny || z >= nz) return;

    // voxel-centre seed (VV/renderer.cpp:1353-1358, lic3d_volume_fragment.glsl:4-5)
    const float sx = ((float)x + 0.5f) / (float)nx * a.sc0;
    const float sy = ((float)y + 0.5f) / (float)ny * a.sc1;
    const float sz = ((float)z + 0.5f) / (float)nz * a.sc2;

    const CachedArgs c = cached_args(a);
    LicFieldCache<FT, SAME> fc;
    fc.reset();
    NoiseCache ncache;
    ncache.reset();
    const float acc = lic_point_cached<FT, SAME, BRICK>(a, c, fc, ncache, sx, sy, sz);
    store_repeat<float>(a.out, a.osy, a.osz, x, y, z, nx, ny, nz, acc * a.scl);
}

// SAME: at most 80 VGPRs (6 waves per SIMD instead of 5, 1-2 registers
// spilled), measured 1.5 % faster at 256^3 (FP32 5.71 vs 5.80 ms, FP16 4.79
// vs 4.86 ms); the separate-noise variants keep their 86 registers
// BRICK: the float4 LIC copy in 8^3 bricks (a.field.p, strides a.bsy/bsz)
template <typename FT, bool SAME, bool BRICK = false>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(SAME ? 6 : 1)))
k_lic_cached(LicArgs a)
{
    const int b = blockIdx.x;
    const int tile = (b & 7) * a.per_xcd + (b >> 3);
    if (tile >= a.nblocks) return;
    lic_cached_tile<FT, SAME, BRICK>(a, tile);
}

// LIC over the tiles of several volumes (one LicArgs per volume, same
// storage and noise grid relation), one block per tile of the concatenated
// list, so the hardware hands tiles to CUs as they free up.  Each volume holds
// a multiple of 8 tiles, so the XCD order of its tiles is the one of
// k_lic_cached and every XCD works on a contiguous part.
template <typename FT, bool SAME>
__global__ void __launch_bounds__(kThreads) k_lic_batch(const LicArgs* __restrict__ jobs,
                                                        const int* __restrict__ first, int njobs)
{
    const int item = blockIdx.x;
    int j = 0;
    while (j + 1 < njobs && item >= first[j + 1]) ++j;
    j = __builtin_amdgcn_readfirstlane(j);  // block-uniform: descriptor loads go to SGPRs
    const int i = item - first[j];
    const LicArgs a = jobs[j];  // by value: block-uniform, kept in SGPRs
    const int tile = (i & 7) * a.per_xcd + (i >> 3);
    if (tile < a.nblocks) lic_cached_tile<FT, SAME>(a, tile);
}

template <typename FT, int NC, bool GATE>
__global__ void __launch_bounds__(kThreads) k_lic(LicArgs a)
{
    // XCD-aware tile order: hardware block b runs on XCD b % 8; give each
    // XCD a contiguous run of tiles so streamlines of neighbouring tiles
    // share that XCD's L2.
    const int b = blockIdx.x;
    const int tile = (b & 7) * a.per_xcd + (b >> 3);
    if (tile >= a.nblocks) return;
    int bx, by, bz;
    tile_coords(a, tile, bx, by, bz);
    const int t = threadIdx.x;
    const int x = bx * a.tx + t % a.tx;
    const int y = by * a.ty + (t / a.tx) % a.ty;
    const int z = bz * a.tz + t / (a.tx * a.ty);
    const int nx = a.field.n[0], ny = a.field.n[1], nz = a.field.n[2];
    if (x >= nx || y >= ny || z >= nz) return;

    // voxel-centre seed (VV/renderer.cpp:1353-1358, lic3d_volume_fragment.glsl:4-5)
    const float sx = ((float)x + 0.5f) / (float)nx * a.sc0;
    const float sy = ((float)y + 0.5f) / (float)ny * a.sc1;
    const float sz = ((float)z + 0.5f) / (float)nz * a.sc2;

    const float4 acc = lic_point<FT, NC, GATE>(a, sx, sy, sz);
    if constexpr (NC == 1) {
        store_repeat<float>(a.out, a.osy, a.osz, x, y, z, nx, ny, nz, acc.x * a.scl);  // lic3d_volume_fragment.glsl:15
    } else {
        store_repeat<float4>(reinterpret_cast<float4*>(a.out), a.osy, a.osz, x, y, z, nx, ny, nz,
                             make_float4(acc.x, acc.y, acc.z, acc.w * a.scl));
    }
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n)), dim3(kThreads), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_fill_i64(long long* p, long long v, long long n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i64, dim3(blocks_for(n)), dim3(kThreads), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_field_maxlen(const float* v0, const float* v1, float t, long long n,
                               uint32_t* maxbits, hipStream_t s)
{
    unsigned g = blocks_for(n);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_field_maxlen, dim3(g), dim3(kThreads), 0, s, v0, v1, t, n, maxbits);
    return hipGetLastError();
}

hipError_t launch_field_encode(const float* v0, const float* v1, float t, const int d[3],
                               const uint32_t* maxbits, void* out, int field_type, hipStream_t s, int raw)
{
    long long np = (long long)(d[0] + 2) * (d[1] + 2) * (d[2] + 2);
    return with_field_type(field_type, [&](auto tag) {
        using FT = typename decltype(tag)::type;
        hipLaunchKernelGGL(k_field_encode<FT>, dim3(blocks_for(np)), dim3(kThreads), 0, s, v0, v1, t, d[0], d[1],
                           d[2], maxbits, raw, (FT*)out);
    });
}

hipError_t launch_field_decode(const void* field, int field_type, const int d[3], float* out4, hipStream_t s)
{
    long long n = (long long)d[0] * d[1] * d[2];
    return with_field_type(field_type, [&](auto tag) {
        using FT = typename decltype(tag)::type;
        hipLaunchKernelGGL(k_field_decode<FT>, dim3(blocks_for(n)), dim3(kThreads), 0, s, (const FT*)field, d[0],
                           d[1], d[2], (float4*)out4);
    });
}

hipError_t launch_pack_lic_field(const void* field, int field_type, const uint8_t* noise, const int d[3], void* out,
                                 hipStream_t s)
{
    const long long np = (long long)(d[0] + 2) * (d[1] + 2) * (d[2] + 2);
    return with_field_type(field_type, [&](auto tag) {
        using FT = typename decltype(tag)::type;
        hipLaunchKernelGGL(k_pack_lic_field<FT>, dim3(blocks_for(np)), dim3(kThreads), 0, s, (const FT*)field, noise,
                           d[0], d[1], d[2], (FT*)out);
    });
}

hipError_t launch_pack_lic_field_brick(const void* copy, const int d[3], void* brick, hipStream_t s)
{
    long long nb[3];
    brick_counts(d, nb);
    const long long np = (long long)(d[0] + 2) * (d[1] + 2) * (d[2] + 2);
    hipLaunchKernelGGL(k_pack_lic_field_brick, dim3(blocks_for(np)), dim3(kThreads), 0, s, (const float4*)copy,
                       d[0] + 2, d[1] + 2, d[2] + 2, (int)nb[0], (int)nb[1], (float4*)brick);
    return hipGetLastError();
}

hipError_t launch_noise_gen(uint64_t seed, int kind, uint64_t thr, const int d[3], uint8_t* out,
                            hipStream_t s)
{
    long long n = (long long)d[0] * d[1] * d[2];
    hipLaunchKernelGGL(k_noise_gen, dim3(blocks_for(n)), dim3(kThreads), 0, s, seed, kind, thr, n, out);
    return hipGetLastError();
}

hipError_t launch_noise_grad_raw(const uint8_t* noise, const int d[3], const float sd[3], float* g,
                                 hipStream_t s)
{
    long long n = (long long)d[0] * d[1] * d[2];
    hipLaunchKernelGGL(k_noise_grad_raw, dim3(blocks_for(n)), dim3(kThreads), 0, s, noise, d[0], d[1],
                       d[2], sd[0], sd[1], sd[2], g);
    return hipGetLastError();
}

hipError_t launch_noise_grad_filter(const float* g, const int d[3], const float w27[27], uint8_t* rgb,
                                    hipStream_t s)
{
    long long n = (long long)d[0] * d[1] * d[2];
    W27 w;
    for (int i = 0; i < 27; ++i) w.w[i] = w27[i];
    hipLaunchKernelGGL(k_noise_grad_filter, dim3(blocks_for(n)), dim3(kThreads), 0, s, g, d[0], d[1],
                       d[2], w, rgb);
    return hipGetLastError();
}

hipError_t launch_noise_tex(const uint8_t* noise, const uint8_t* rgb, const int d[3], void* tex, int nc,
                            hipStream_t s)
{
    long long np = (long long)(d[0] + 1) * (d[1] + 1) * (d[2] + 1);
    hipLaunchKernelGGL(k_noise_tex, dim3(blocks_for(np)), dim3(kThreads), 0, s, noise, rgb, d[0], d[1],
                       d[2], tex, nc);
    return hipGetLastError();
}

hipError_t launch_noise_untex(const void* tex, int nc, const int d[3], uint8_t* out, hipStream_t s)
{
    long long n = (long long)d[0] * d[1] * d[2];
    hipLaunchKernelGGL(k_noise_untex, dim3(blocks_for(n)), dim3(kThreads), 0, s, tex, nc, d[0], d[1],
                       d[2], out);
    return hipGetLastError();
}

hipError_t launch_pad_clamp_u8(const uint8_t* src, const int d[3], uint8_t* dst, hipStream_t s)
{
    long long np = (long long)(d[0] + 2) * (d[1] + 2) * (d[2] + 2);
    hipLaunchKernelGGL(k_pad_clamp<uint8_t>, dim3(blocks_for(np)), dim3(kThreads), 0, s, src, d[0], d[1],
                       d[2], dst);
    return hipGetLastError();
}

hipError_t launch_pad_clamp_f32(const float* src, const int d[3], float* dst, hipStream_t s)
{
    long long np = (long long)(d[0] + 2) * (d[1] + 2) * (d[2] + 2);
    hipLaunchKernelGGL(k_pad_clamp<float>, dim3(blocks_for(np)), dim3(kThreads), 0, s, src, d[0], d[1],
                       d[2], dst);
    return hipGetLastError();
}

hipError_t launch_pad_repeat_f32(const float* src, int nc, const int d[3], float* dst, hipStream_t s)
{
    long long np = (long long)(d[0] + 1) * (d[1] + 1) * (d[2] + 1);
    hipLaunchKernelGGL(k_pad_repeat_f32, dim3(blocks_for(np)), dim3(kThreads), 0, s, src, nc, d[0], d[1],
                       d[2], dst);
    return hipGetLastError();
}

hipError_t launch_unpad_repeat_f32(const float* src, int nc, const int d[3], float* dst, hipStream_t s)
{
    long long n = (long long)d[0] * d[1] * d[2];
    hipLaunchKernelGGL(k_unpad_repeat_f32, dim3(blocks_for(n)), dim3(kThreads), 0, s, src, nc, d[0], d[1],
                       d[2], dst);
    return hipGetLastError();
}

hipError_t launch_repeat_halo_f32(float* vol, int nc, const int d[3], hipStream_t s)
{
    long long np = (long long)(d[0] + 1) * (d[1] + 1) * (d[2] + 1);
    hipLaunchKernelGGL(k_repeat_halo_f32, dim3(blocks_for(np)), dim3(kThreads), 0, s, vol, nc, d[0], d[1],
                       d[2]);
    return hipGetLastError();
}

namespace {

// the kernel argument block of a LIC launch (volume, weights, tile order)
hipError_t make_lic_args(const LicLaunch& L, LicArgs& a)
{
    std::memset(&a, 0, sizeof(a));
    a.field = Vol<void>{L.field.p, {L.field.n[0], L.field.n[1], L.field.n[2]}, L.field.sy, L.field.sz};
    a.noise = Vol<void>{L.noise.p, {L.noise.n[0], L.noise.n[1], L.noise.n[2]}, L.noise.sy, L.noise.sz};
    a.scalar = Vol<uint8_t>{(const uint8_t*)L.scalar.p, {L.scalar.n[0], L.scalar.n[1], L.scalar.n[2]},
                            L.scalar.sy, L.scalar.sz};
    a.w = L.w;
    a.F = L.F;
    a.B = L.B;
    a.he = L.he;
    a.speed = L.speed;
    a.scl = L.scl;
    a.freq = L.freq;
    a.gate_lo = L.gate_lo;
    a.gate_hi = L.gate_hi;
    a.sc0 = L.scale[0];
    a.sc1 = L.scale[1];
    a.sc2 = L.scale[2];
    a.out = (float*)L.out.p;
    a.field_bytes = (unsigned)L.field_bytes;
    a.noise_bytes = (unsigned)L.noise_bytes;
    a.osy = L.out.sy;
    a.osz = L.out.sz;
    a.tx = L.tile[0];
    a.ty = L.tile[1];
    a.tz = L.tile[2];
    if (a.tx * a.ty * a.tz != kThreads) return hipErrorInvalidValue;
    a.sx = L.super_tile[0] > 0 ? L.super_tile[0] : 1;
    a.sy = L.super_tile[1] > 0 ? L.super_tile[1] : 1;
    a.sz = L.super_tile[2] > 0 ? L.super_tile[2] : 1;
    // the grid of tiles is rounded up to whole super-tiles; tiles past the
    // volume exit at the voxel bounds check
    a.sgx = (L.field.n[0] + a.tx * a.sx - 1) / (a.tx * a.sx);
    a.sgy = (L.field.n[1] + a.ty * a.sy - 1) / (a.ty * a.sy);
    const int sgz = (L.field.n[2] + a.tz * a.sz - 1) / (a.tz * a.sz);
    a.gx = a.sgx * a.sx;
    a.gy = a.sgy * a.sy;
    a.gz = sgz * a.sz;
    a.nblocks = a.gx * a.gy * a.gz;
    a.per_xcd = (a.nblocks + 7) / 8;
    return hipSuccess;
}

}  // namespace

hipError_t launch_lic(const LicLaunch& L, hipStream_t s)
{
    LicArgs a;
    const hipError_t e = make_lic_args(L, a);
    if (e != hipSuccess) return e;
    dim3 grid((unsigned)(a.per_xcd * 8));
    dim3 block(kThreads);
    const bool g = L.gate != 0;
    const bool cached = L.cache && L.nc == 1 && !g;
    // the SAME kernels read the noise from channel 3 of the LIC field copy
    const bool same = L.field_has_noise && L.freq == 1.0f && L.noise.n[0] == L.field.n[0] &&
                      L.noise.n[1] == L.field.n[1] && L.noise.n[2] == L.field.n[2];
    return with_field_type(L.field_type, [&](auto tag) {
        using FT = typename decltype(tag)::type;
        if constexpr (std::is_same<FT, float4>::value) {
            if (cached && same && L.field_brick) {
                // the LIC volume kernel alone reads the brick copy
                long long nb[3];
                const int d[3] = {L.field.n[0], L.field.n[1], L.field.n[2]};
                brick_counts(d, nb);
                LicArgs ab = a;
                ab.field.p = L.field_brick;
                ab.bsy = (int)(nb[0] * 512);
                ab.bsz = (int)(nb[0] * nb[1] * 512);
                hipLaunchKernelGGL((k_lic_cached<FT, true, true>), grid, block, 0, s, ab);
                return;
            }
        }
        if (cached && same)
            hipLaunchKernelGGL((k_lic_cached<FT, true>), grid, block, 0, s, a);
        else if (cached)
            hipLaunchKernelGGL((k_lic_cached<FT, false>), grid, block, 0, s, a);
        else if (L.nc == 4)
            hipLaunchKernelGGL((k_lic<FT, 4, false>), grid, block, 0, s, a);
        else if (g)
            hipLaunchKernelGGL((k_lic<FT, 1, true>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((k_lic<FT, 1, false>), grid, block, 0, s, a);
    });
}

}  // namespace vv

namespace vv {

size_t lic_args_size() { return sizeof(LicArgs); }

// per-volume kernel arguments of the cached kernel, as launch_lic builds them
hipError_t lic_args_fill(const LicLaunch& L, void* host_args, int* ntiles)
{
    LicArgs a;
    const hipError_t e = make_lic_args(L, a);
    if (e != hipSuccess) return e;
    std::memcpy(host_args, &a, sizeof(a));
    *ntiles = a.per_xcd * 8;
    return hipSuccess;
}

hipError_t launch_lic_batch(const LicBatchLaunch& B, hipStream_t s)
{
    if (B.njobs < 1 || B.total < 1) return hipErrorInvalidValue;
    const LicArgs* jobs = (const LicArgs*)B.jobs;
    return with_field_type(B.field_type, [&](auto tag) {
        using FT = typename decltype(tag)::type;
        if (B.same)
            hipLaunchKernelGGL((k_lic_batch<FT, true>), dim3(B.total), dim3(kThreads), 0, s, jobs, B.first, B.njobs);
        else
            hipLaunchKernelGGL((k_lic_batch<FT, false>), dim3(B.total), dim3(kThreads), 0, s, jobs, B.first, B.njobs);
    });
}

}  // namespace vv
