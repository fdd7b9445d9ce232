// vv_lic_dev.hpp — device code of the LIC integral shared by the LIC-volume
// kernels (vv_lic.hip) and the fused on-the-fly ray caster (vv_render.hip):
// the argument block, field and noise sampling, the per-lane cell caches and
// the LIC of one seed point.  Included inside an anonymous namespace.
#pragma once


struct LicArgs {
    Vol<void> field;
    Vol<void> noise;
    Vol<uint8_t> scalar;
    const float* w;
    int F, B;
    float he, scl, freq, gate_lo, gate_hi;
    float sc0, sc1, sc2;
    float* out;
    int osy, osz;
    unsigned field_bytes, noise_bytes;  // buffer sizes for the cached kernels
    int tx, ty, tz;      // block tile
    int gx, gy, gz;      // tiles per axis
    int nblocks;         // gx*gy*gz
    int per_xcd;         // tiles handled by each of the 8 XCDs
    int sx, sy, sz;      // super-tile size in tiles (3D-blocked tile order)
    int sgx, sgy;        // super-tiles per axis (x, y)
    int speed;           // SPEED_OF_FLOW (uncached path only)
    int bsy, bsz;        // BRICK kernels: field.p is the 8^3-brick copy with these brick-row and
                         // brick-plane strides (elements)
};

// logical tile t -> tile coordinates: super-tiles in raster order, tiles
// raster order inside each super-tile, so the tiles resident on one XCD at a
// time cover a compact 3D region whose streamline halo fits that XCD's L2
__device__ __forceinline__ void tile_coords(const LicArgs& a, int t, int& bx, int& by, int& bz)
{
    const int per_super = a.sx * a.sy * a.sz;
    const int st = t / per_super, r = t - st * per_super;
    const int ux = st % a.sgx, uy = (st / a.sgx) % a.sgy, uz = st / (a.sgx * a.sgy);
    bx = ux * a.sx + r % a.sx;
    by = uy * a.sy + (r / a.sx) % a.sy;
    bz = uz * a.sz + r / (a.sx * a.sy);
}

template <int NC>
struct NoiseT;
template <>
struct NoiseT<1> {
    using T = uint8_t;
};
template <>
struct NoiseT<4> {
    using T = uchar4;
};

// noise sample n(x): freqSampling (inc_lic.glsl:71-90) for scalar noise,
// freqSamplingGrad (:61-68, no frequency scale, no gate) for RGBA noise
template <int NC, bool GATE>
__device__ __forceinline__ float4 noise_at(const LicArgs& a, float x, float y, float z)
{
    float4 n = make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (NC == 4) {
        const Vol<uchar4> nv{(const uchar4*)a.noise.p, {a.noise.n[0], a.noise.n[1], a.noise.n[2]},
                             a.noise.sy, a.noise.sz};
        Cell c = cell_repeat(nv, x, y, z);
        float4 r = tri_u8x4(nv, c);
        n.x = r.x * kInv255;
        n.y = r.y * kInv255;
        n.z = r.z * kInv255;
        n.w = r.w * kInv255;
    } else {
        if constexpr (GATE) {
            Cell cs = cell_clamp(a.scalar, x, y, z);
            float s = tri_u8(a.scalar, cs) * kInv255;
            if (!(s > a.gate_lo && s < a.gate_hi)) return n;
        }
        const Vol<uint8_t> nv{(const uint8_t*)a.noise.p, {a.noise.n[0], a.noise.n[1], a.noise.n[2]},
                              a.noise.sy, a.noise.sz};
        Cell c = cell_repeat(nv, x * a.freq, y * a.freq, z * a.freq);
        n.x = tri_u8(nv, c) * kInv255;
    }
    return n;
}

template <typename FT>
__device__ __forceinline__ float3 field_rgb(const LicArgs& a, float x, float y, float z)
{
    const Vol<FT> fv{(const FT*)a.field.p, {a.field.n[0], a.field.n[1], a.field.n[2]}, a.field.sy,
                     a.field.sz};
    Cell c = cell_clamp(fv, x, y, z);
    return tri_rgb(fv, c);
}

// direction and magnitude (.a = |v|/max) of the field
template <typename FT>
__device__ __forceinline__ float4 field_rgba(const LicArgs& a, float x, float y, float z)
{
    const Vol<FT> fv{(const FT*)a.field.p, {a.field.n[0], a.field.n[1], a.field.n[2]}, a.field.sy,
                     a.field.sz};
    Cell c = cell_clamp(fv, x, y, z);
    return tri_rgba(fv, c);
}

// Stores one LIC voxel into the REPEAT-padded output and into the copies
// of plane 0 at index n along each axis (the wrap halo), so no separate
// halo pass is needed.
template <typename T>
__device__ __forceinline__ void store_repeat(T* out, int osy, int osz, int x, int y, int z, int nx, int ny,
                                             int nz, T v)
{
    const long long o = (long long)z * osz + (long long)y * osy + x;
    out[o] = v;
    const bool hx = x == 0, hy = y == 0, hz = z == 0;
    if (hx | hy | hz) {
        const long long ox = nx, oy = (long long)ny * osy, oz = (long long)nz * osz;
        if (hx) out[o + ox] = v;
        if (hy) out[o + oy] = v;
        if (hz) out[o + oz] = v;
        if (hx & hy) out[o + ox + oy] = v;
        if (hx & hz) out[o + ox + oz] = v;
        if (hy & hz) out[o + oy + oz] = v;
        if (hx & hy & hz) out[o + ox + oy + oz] = v;
    }
}

// Per-lane cache of the last trilinear cell.  Consecutive samples of a
// streamline are ~h_e*n voxels apart (0.38 at 256^3, h = 0.005), so the
// corrector sample almost always lands in the predictor's cell and the next
// predictor often in the same cell: a lane that hits skips its gathers (it
// is masked off the loads).  For each channel the cache keeps the four x0
// texels and the four differences (x1 - x0) of the cell, so the four x-lerps
// of a hit are single fmas; the values and the operation order are exactly
// those of tri(), so results do not change.
// NCH = 3: the field direction; NCH = 4: also the noise byte that the LIC
// copy of the field carries in channel 3 (equal field and noise grids)
// A cache is keyed by the three floors of its cell (x = u n - 0.5), so a hit
// costs three compares and no index arithmetic; NaN keys never hit.  (Cells
// outside the volume with different floors but the same clamped texels are
// reloaded, with the same values.)
// v_fma_mix_f32 with one half operand, which the instruction widens exactly
// (every half is a float), so each form is one rounding of the exact result:
//   fma_half<HI>(a, b, w) = fmaf(a, b, (float)h), h the low (HI 0) or high
//   (HI 1) half of w;
//   sub_half<HI>(w1, w0)  = (float)h1 - (float)h0.
// They let the FP16 cell cache keep its x0 texels packed as loaded and skip
// every half-to-float conversion of a cell fill.
template <int HI>
__device__ __forceinline__ float fma_half(float a, float b, unsigned w)
{
    float r;
    if constexpr (HI)
        asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "=v"(r) : "v"(a), "v"(b), "v"(w));
    else
        asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[0,0,1]" : "=v"(r) : "v"(a), "v"(b), "v"(w));
    return r;
}

template <int HI>
__device__ __forceinline__ float sub_half(unsigned w1, unsigned w0)
{
    float r;
    if constexpr (HI)
        asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel:[1,0,1] op_sel_hi:[1,0,1]" : "=v"(r) : "v"(w1), "v"(w0));
    else
        asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel_hi:[1,0,1]" : "=v"(r) : "v"(w1), "v"(w0));
    return r;
}

// FP32 and UNORM8 texels: x0 texels and differences as floats
template <typename FT, int NCH>
struct FieldCache {
    float kx, ky, kz;   // floors of the cached cell
    bool same;          // NCH 4: the cell's noise bytes are those of the REPEAT cell
    float v0[NCH][4];   // [channel][(y,z) corner]: texel at x0
    float dx[NCH][4];   // texel(x1) - texel(x0)
    __device__ __forceinline__ void reset() { kx = ky = kz = __builtin_nanf(""); same = false; }
    // x-lerp of channel CH on corner row k: fma(fx, x1 - x0, x0)
    template <int CH>
    __device__ __forceinline__ float xlerp(int k, float fx) const { return fmaf(fx, dx[CH][k], v0[CH][k]); }
    __device__ __forceinline__ void set_row(int k, const float* t0, const float* t1)
    {
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            v0[ch][k] = t0[ch];
            dx[ch][k] = t1[ch] - t0[ch];
        }
    }
};

// FP16 texels: the x0 texel of each corner row stays as the two words the
// row load returned (channels 0|1 and 2|3, low half first), the differences
// are floats; x-lerps and differences read the halves with fma_half and
// sub_half.  Values and operation order are those of the float cache on the
// widened halves, so results are the same bits.
template <int NCH>
struct FieldCache<half4_t, NCH> {
    float kx, ky, kz;
    bool same;
    unsigned v0h[4][2];  // [(y,z) corner][word]
    float dx[NCH][4];
    __device__ __forceinline__ void reset() { kx = ky = kz = __builtin_nanf(""); same = false; }
    template <int CH>
    __device__ __forceinline__ float xlerp(int k, float fx) const
    {
        return fma_half<CH & 1>(fx, dx[CH][k], v0h[k][CH >> 1]);
    }
    // a = the row's 16 bytes: texel x0 in .x/.y, texel x1 in .z/.w
    __device__ __forceinline__ void set_row(int k, const uint4& a)
    {
        v0h[k][0] = a.x;
        v0h[k][1] = a.y;
        dx[0][k] = sub_half<0>(a.z, a.x);
        dx[1][k] = sub_half<1>(a.z, a.x);
        dx[2][k] = sub_half<0>(a.w, a.y);
        if constexpr (NCH == 4) dx[3][k] = sub_half<1>(a.w, a.y);
    }
};

// trilinear value of channel CH of a cached cell: tri_cached's arithmetic
template <int CH, typename FC>
__device__ __forceinline__ float tri_fc(const FC& fc, float fx, float fy, float fz)
{
    const float c00 = fc.template xlerp<CH>(0, fx);
    const float c10 = fc.template xlerp<CH>(1, fx);
    const float c01 = fc.template xlerp<CH>(2, fx);
    const float c11 = fc.template xlerp<CH>(3, fx);
    const float c0 = lerpf(c00, c10, fy);
    const float c1 = lerpf(c01, c11, fy);
    return lerpf(c0, c1, fz);
}

struct NoiseCache {
    float kx, ky, kz;
    float v0[4], dx[4];
    __device__ __forceinline__ void reset() { kx = ky = kz = __builtin_nanf(""); }
};

__device__ __forceinline__ float tri_cached(const float* v0, const float* dx, float fx, float fy, float fz)
{
    const float c00 = fmaf(fx, dx[0], v0[0]);
    const float c10 = fmaf(fx, dx[1], v0[1]);
    const float c01 = fmaf(fx, dx[2], v0[2]);
    const float c11 = fmaf(fx, dx[3], v0[3]);
    const float c0 = lerpf(c00, c10, fy);
    const float c1 = lerpf(c01, c11, fy);
    return lerpf(c0, c1, fz);
}

// axis coordinate: floor and fraction of x = u*n - 0.5 (shared by the
// CLAMP field cell and, on the same grid, the REPEAT noise cell)
struct Ax {
    float fl, f;
};

__device__ __forceinline__ Ax axis(float u, int n)
{
    Ax r;
    const float x = fmaf(u, (float)n, -0.5f);
    r.fl = floorf(x);
    r.f = x - r.fl;
    return r;
}

// corner index of a CLAMP-padded axis: clamp(floor, -1, n-1) + 1
__device__ __forceinline__ unsigned clamp_idx(float fl, float nm1)
{
    return (unsigned)((int)__builtin_amdgcn_fmed3f(fl, -1.0f, nm1) + 1);
}

// corner index of a REPEAT-padded axis: floor mod n
__device__ __forceinline__ unsigned repeat_idx(float fl, int n)
{
    int i = (int)__builtin_amdgcn_fmed3f(fl, -1.0e9f, 1.0e9f);
    if ((unsigned)i >= (unsigned)n) {
        i %= n;
        if (i < 0) i += n;
    }
    return (unsigned)i;
}

// Texel loads at a 32-bit unsigned byte offset from a block-uniform base:
// base + zext(off) compiles to global_load ... voffset, s[base] (a VGPR
// offset of 32 bits, the 64-bit base in SGPRs), so arrays up to 4 GiB need
// no 64-bit address arithmetic per lane.  (Raw buffer loads, which also take
// a 32-bit VGPR offset, returned zeros for offsets of 2 GiB and more.)
// global address space: the batch kernel reads its array pointers from a
// descriptor in memory, where the compiler cannot infer it (flat loads)
typedef const __attribute__((address_space(1))) char* gchar;
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) unsigned gu32;

template <typename FT>
struct FieldLoad;

// The loaders fill NCH channels (3: rgb, 4: rgb + channel 3) of the texels
// x0 and x1 of a cell row.
// float4 texels: two 12- or 16-byte loads (x0 and x1)
template <>
struct FieldLoad<float4> {
    static constexpr unsigned kShift = 4;
    template <int NCH>
    __device__ static __forceinline__ void pair(const char* base, unsigned off, float* t0, float* t1)
    {
        const gfloat* q = (const gfloat*)((gchar)base + off);
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            t0[ch] = q[ch];
            t1[ch] = q[4 + ch];
        }
    }
};

// half4 texels: one 16-byte load holds x0 and x1, kept as the raw words for
// the FP16 cell cache
template <>
struct FieldLoad<half4_t> {
    static constexpr unsigned kShift = 3;
    __device__ static __forceinline__ uint4 row(const char* base, unsigned off)
    {
        const gu32* q = (const gu32*)((gchar)base + off);
        uint4 a;
        a.x = q[0];
        a.y = q[1];
        a.z = q[2];
        a.w = q[3];
        return a;
    }
};

// UNORM8 texels: one 8-byte load holds x0 and x1 (raw byte values)
template <>
struct FieldLoad<u8x4_t> {
    static constexpr unsigned kShift = 2;
    template <int NCH>
    __device__ static __forceinline__ void pair(const char* base, unsigned off, float* t0, float* t1)
    {
        const gu32* q = (const gu32*)((gchar)base + off);
        const unsigned a = q[0], b = q[1];
        t0[0] = (float)(a & 0xffu);
        t0[1] = (float)((a >> 8) & 0xffu);
        t0[2] = (float)((a >> 16) & 0xffu);
        t1[0] = (float)(b & 0xffu);
        t1[1] = (float)((b >> 8) & 0xffu);
        t1[2] = (float)((b >> 16) & 0xffu);
        if constexpr (NCH == 4) {
            t0[3] = (float)(a >> 24);
            t1[3] = (float)(b >> 24);
        }
    }
};

// One corner row of a cell into a cache, in two steps so a caller can issue
// several loads before it uses any: RowLoad<FT, NCH>::load, then store.
template <typename FT, int NCH>
struct RowLoad {
    float t0[NCH], t1[NCH];
    __device__ __forceinline__ void load(const char* base, unsigned off)
    {
        FieldLoad<FT>::template pair<NCH>(base, off, t0, t1);
    }
    __device__ __forceinline__ void store(FieldCache<FT, NCH>& fc, int k) const { fc.set_row(k, t0, t1); }
};

template <int NCH>
struct RowLoad<half4_t, NCH> {
    uint4 a;
    __device__ __forceinline__ void load(const char* base, unsigned off) { a = FieldLoad<half4_t>::row(base, off); }
    __device__ __forceinline__ void store(FieldCache<half4_t, NCH>& fc, int k) const { fc.set_row(k, a); }
};

// two adjacent noise bytes (x0, x1) in one 16-bit load; the address may be odd
typedef unsigned short __attribute__((aligned(1))) u16_unaligned;
__device__ __forceinline__ unsigned noise_pair(const unsigned char* base, unsigned off)
{
    gchar g = (gchar)(const char*)base;
    return *(const __attribute__((address_space(1))) u16_unaligned*)(g + off);
}

struct CachedArgs {
    const char* fr;           // field texels
    unsigned bsy, bsz;        // BRICK: brick-row and brick-plane strides (elements)
    const unsigned char* nr;  // noise bytes
    float fnm1[3];      // field n-1 per axis
    unsigned fsy, fsz;  // field element strides
    unsigned nsy, nsz;  // noise element strides
};

__device__ __forceinline__ bool key_hit(float kx, float ky, float kz, const Ax& ax, const Ax& ay, const Ax& az)
{
    return ax.fl == kx && ay.fl == ky && az.fl == kz;
}

// Equal field and noise grids with freq 1 (SAME): the LIC copy of the field
// holds, in channel 3 of the CLAMP-padded texel p, the noise byte of REPEAT
// texel (p - 1) mod n per axis.  While every floor lies in [-1, n-1] the
// CLAMP cell (clamp(floor) + 1, + 2) and the REPEAT cell (floor mod n, + 1)
// then hold the same noise bytes, so the field fill brings the noise cell
// with it: no noise loads.  Outside that range (a streamline that left the
// volume) the noise comes from the REPEAT array.  The noise value and its
// arithmetic are those of noise_from_axes.  field_from_axes sets fc.same
// for a cell whose three floors lie in that range.

template <typename FT, int NCH, bool BRICK = false>
__device__ __forceinline__ float3 field_from_axes(const CachedArgs& c, FieldCache<FT, NCH>& fc, const Ax& ax,
                                                  const Ax& ay, const Ax& az)
{
    if (!key_hit(fc.kx, fc.ky, fc.kz, ax, ay, az)) {
        fc.kx = ax.fl;
        fc.ky = ay.fl;
        fc.kz = az.fl;
        // clamped floors; a floor the clamp leaves unchanged lies in [-1, n-1]
        const float mx = __builtin_amdgcn_fmed3f(ax.fl, -1.0f, c.fnm1[0]);
        const float my = __builtin_amdgcn_fmed3f(ay.fl, -1.0f, c.fnm1[1]);
        const float mz = __builtin_amdgcn_fmed3f(az.fl, -1.0f, c.fnm1[2]);
        if constexpr (NCH == 4) fc.same = mx == ax.fl && my == ay.fl && mz == az.fl;
        if constexpr (BRICK) {
            // float4 texels in 8^3 bricks (c.fr): the element of padded
            // texel (x, y, z) is xb(x) + yb(y) + zb(z)
            const unsigned x0 = (unsigned)((int)mx + 1), y0 = (unsigned)((int)my + 1), z0 = (unsigned)((int)mz + 1);
            const unsigned xb[2] = {((x0 >> 3) << 9) | (x0 & 7u), (((x0 + 1) >> 3) << 9) | ((x0 + 1) & 7u)};
            const unsigned yb[2] = {__umul24(y0 >> 3, c.bsy) + ((y0 & 7u) << 3),
                                    __umul24((y0 + 1) >> 3, c.bsy) + (((y0 + 1) & 7u) << 3)};
            const unsigned zb[2] = {__umul24(z0 >> 3, c.bsz) + ((z0 & 7u) << 6),
                                    __umul24((z0 + 1) >> 3, c.bsz) + (((z0 + 1) & 7u) << 6)};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const unsigned row = yb[k & 1] + zb[k >> 1];
                const gfloat* q0 = (const gfloat*)((gchar)c.fr + ((row + xb[0]) << 4));
                const gfloat* q1 = (const gfloat*)((gchar)c.fr + ((row + xb[1]) << 4));
                float t0[NCH], t1[NCH];
#pragma unroll
                for (int ch = 0; ch < NCH; ++ch) {
                    t0[ch] = q0[ch];
                    t1[ch] = q1[ch];
                }
                fc.set_row(k, t0, t1);
            }
        } else {
            const unsigned base = __umul24((unsigned)((int)mz + 1), c.fsz) + __umul24((unsigned)((int)my + 1), c.fsy) +
                                  (unsigned)((int)mx + 1);
            // one byte offset for the four rows; each row's start is a
            // block-uniform pointer (scalar arithmetic)
            const unsigned boff = base << FieldLoad<FT>::kShift;
            const size_t off[4] = {0u, (size_t)c.fsy << FieldLoad<FT>::kShift, (size_t)c.fsz << FieldLoad<FT>::kShift,
                                   (size_t)(c.fsz + c.fsy) << FieldLoad<FT>::kShift};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                RowLoad<FT, NCH> row;
                row.load(c.fr + off[k], boff);
                row.store(fc, k);
            }
        }
    }
    float3 r;
    r.x = FieldScale<FT>::apply(tri_fc<0>(fc, ax.f, ay.f, az.f));
    r.y = FieldScale<FT>::apply(tri_fc<1>(fc, ax.f, ay.f, az.f));
    r.z = FieldScale<FT>::apply(tri_fc<2>(fc, ax.f, ay.f, az.f));
    return r;
}

// corrector sample: the field cell and the noise cell of the same position.
// Both fills are issued before either result is used, so the two gathers
// overlap (one memory round trip instead of two); values and arithmetic
// are those of field_from_axes and noise_from_axes.
template <typename FT>
__device__ __forceinline__ float3 field_and_noise(const LicArgs& a, const CachedArgs& c, FieldCache<FT, 3>& fc,
                                                  NoiseCache& nc, const Ax& ax, const Ax& ay,
                                                  const Ax& az, const Ax& nx_, const Ax& ny_,
                                                  const Ax& nz_, float& noise)
{
    const bool fmiss = !key_hit(fc.kx, fc.ky, fc.kz, ax, ay, az);
    const bool nmiss = !key_hit(nc.kx, nc.ky, nc.kz, nx_, ny_, nz_);
    unsigned short nraw[4];
    constexpr bool kHalf = std::is_same<FT, half4_t>::value;
    RowLoad<half4_t, 3> hrows[kHalf ? 4 : 1];  // FP16: the raw rows until both fills are in flight
    if (fmiss) {
        const unsigned fb = __umul24(clamp_idx(az.fl, c.fnm1[2]), c.fsz) +
                            __umul24(clamp_idx(ay.fl, c.fnm1[1]), c.fsy) + clamp_idx(ax.fl, c.fnm1[0]);
        const unsigned boff = fb << FieldLoad<FT>::kShift;
        const size_t off[4] = {0u, (size_t)c.fsy << FieldLoad<FT>::kShift, (size_t)c.fsz << FieldLoad<FT>::kShift,
                               (size_t)(c.fsz + c.fsy) << FieldLoad<FT>::kShift};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if constexpr (kHalf) {
                hrows[k].load(c.fr + off[k], boff);
            } else {
                // float caches hold the raw x1 texel in dx until both fills are in flight
                float t0[3], t1[3];
                FieldLoad<FT>::template pair<3>(c.fr + off[k], boff, t0, t1);
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) {
                    fc.v0[ch][k] = t0[ch];
                    fc.dx[ch][k] = t1[ch];
                }
            }
        }
    }
    if (nmiss) {
        const unsigned nb = __umul24(repeat_idx(nz_.fl, a.noise.n[2]), c.nsz) +
                            __umul24(repeat_idx(ny_.fl, a.noise.n[1]), c.nsy) + repeat_idx(nx_.fl, a.noise.n[0]);
        const unsigned off[4] = {0u, c.nsy, c.nsz, c.nsz + c.nsy};
#pragma unroll
        for (int k = 0; k < 4; ++k) nraw[k] = (unsigned short)noise_pair(c.nr, nb + off[k]);
    }
    if (fmiss) {
        fc.kx = ax.fl;
        fc.ky = ay.fl;
        fc.kz = az.fl;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if constexpr (kHalf) {
                hrows[k].store(fc, k);
            } else {
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) fc.dx[ch][k] = fc.dx[ch][k] - fc.v0[ch][k];
            }
        }
    }
    if (nmiss) {
        nc.kx = nx_.fl;
        nc.ky = ny_.fl;
        nc.kz = nz_.fl;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float t0 = (float)(nraw[k] & 0xffu), t1 = (float)(nraw[k] >> 8);
            nc.v0[k] = t0;
            nc.dx[k] = t1 - t0;
        }
    }
    noise = tri_cached(nc.v0, nc.dx, nx_.f, ny_.f, nz_.f);
    float3 r;
    r.x = FieldScale<FT>::apply(tri_fc<0>(fc, ax.f, ay.f, az.f));
    r.y = FieldScale<FT>::apply(tri_fc<1>(fc, ax.f, ay.f, az.f));
    r.z = FieldScale<FT>::apply(tri_fc<2>(fc, ax.f, ay.f, az.f));
    return r;
}

__device__ __forceinline__ float noise_from_axes(const LicArgs& a, const CachedArgs& c, NoiseCache& nc,
                                                 const Ax& ax, const Ax& ay, const Ax& az)
{
    if (!key_hit(nc.kx, nc.ky, nc.kz, ax, ay, az)) {
        nc.kx = ax.fl;
        nc.ky = ay.fl;
        nc.kz = az.fl;
        const unsigned base = __umul24(repeat_idx(az.fl, a.noise.n[2]), c.nsz) +
                              __umul24(repeat_idx(ay.fl, a.noise.n[1]), c.nsy) + repeat_idx(ax.fl, a.noise.n[0]);
        const unsigned off[4] = {0u, c.nsy, c.nsz, c.nsz + c.nsy};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            // two adjacent bytes (x0, x1) in one load
            const unsigned u = noise_pair(c.nr, base + off[k]);
            const float t0 = (float)(u & 0xffu), t1 = (float)(u >> 8);
            nc.v0[k] = t0;
            nc.dx[k] = t1 - t0;
        }
    }
    return tri_cached(nc.v0, nc.dx, ax.f, ay.f, az.f);
}

// REPEAT corner index (floor mod n) with float arithmetic instead of an
// integer remainder, which keeps the rarely taken noise fallback of the SAME
// kernels light on registers.  The floor is limited to +-2^20 first, so the
// quotient estimate fl * rcp(n) is off by at most one and one correction
// step lands in [0, n-1]; the final limit only guards that range.  Equal to
// repeat_idx for |floor| <= 2^20 (a million texels outside the volume).
__device__ __forceinline__ unsigned repeat_idx_f(float fl, float nm1)
{
    const float n = nm1 + 1.0f;
    fl = __builtin_amdgcn_fmed3f(fl, -1048576.0f, 1048576.0f);
    float w = fl - n * floorf(fl * __builtin_amdgcn_rcpf(n));
    w = w >= n ? w - n : (w < 0.0f ? w + n : w);
    return (unsigned)(int)__builtin_amdgcn_fmed3f(w, 0.0f, nm1);
}

// noise sample of a REPEAT cell on the field's grid, without the cache (a
// lane whose streamline left the volume)
__device__ __forceinline__ float noise_direct(const LicArgs& a, const CachedArgs& c, const Ax& ax, const Ax& ay,
                                              const Ax& az)
{
    const unsigned base = __umul24(repeat_idx_f(az.fl, c.fnm1[2]), c.nsz) +
                          __umul24(repeat_idx_f(ay.fl, c.fnm1[1]), c.nsy) + repeat_idx_f(ax.fl, c.fnm1[0]);
    const unsigned off[4] = {0u, c.nsy, c.nsz, c.nsz + c.nsy};
    float v0[4], dx[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const unsigned u = noise_pair(c.nr, base + off[k]);
        v0[k] = (float)(u & 0xffu);
        dx[k] = (float)(u >> 8) - v0[k];
    }
    return tri_cached(v0, dx, ax.f, ay.f, az.f);
}

// noise of the cell just fetched into fc (SAME): channel 3 of the cache, or
// the REPEAT array for a cell outside the volume
template <typename FT>
__device__ __forceinline__ float noise_same(const LicArgs& a, const CachedArgs& c, const FieldCache<FT, 4>& fc,
                                            const Ax& ax, const Ax& ay, const Ax& az)
{
    if (fc.same) return tri_fc<3>(fc, ax.f, ay.f, az.f);
    return noise_direct(a, c, ax, ay, az);
}

// LIC of one seed with the per-lane cell caches (scalar noise without gate):
// the shader's sum of (noise / 255) * weight, before the final scaling by
// scl.  The noise helpers of the cached path (noise_same, noise_from_axes,
// noise_direct, field_and_noise) return byte units (0..255).  SAME: the noise grid
// equals the field grid and freq == 1, so the noise cell of the corrector
// position reuses the corrector's floors and fractions.  The caches may hold
// cells of an earlier seed (the ray caster keeps them along a ray): they only
// decide whether a cell is reloaded, never its values.
template <typename FT, bool SAME>
using LicFieldCache = FieldCache<FT, SAME ? 4 : 3>;

// Step of one axis from an encoded field value s: (2s - 1) * h with the
// shader's two roundings.  2s is exact, so fma(2, s, -1) rounds 2s - 1
// once, as 2.0*s - 1.0 does.  (The one-rounding fma(s, 2h, -h) equals it
// only where 2s - 1 is exact, s >= 0.25; elsewhere a one-ulp step
// difference grows along a streamline through a rough field to several
// 1e-5 of the LIC range, so the literal form stays.)
__device__ __forceinline__ float decode_step(float s, float h)
{
    return fmaf(2.0f, s, -1.0f) * h;
}

template <typename FT, bool SAME, bool BRICK = false>
__device__ __forceinline__ float lic_point_cached(const LicArgs& a, const CachedArgs& c, LicFieldCache<FT, SAME>& fc,
                                                  NoiseCache& ncache, float sx, float sy, float sz)
{
    const int nx = a.field.n[0], ny = a.field.n[1], nz = a.field.n[2];
    auto noise = [&](float px, float py, float pz, const Ax& ax, const Ax& ay, const Ax& az) -> float {
        if constexpr (SAME) {
            return noise_same<FT>(a, c, fc, ax, ay, az);
        } else {
            const float q0 = px * a.freq, q1 = py * a.freq, q2 = pz * a.freq;
            return noise_from_axes(a, c, ncache, axis(q0, a.noise.n[0]), axis(q1, a.noise.n[1]),
                                   axis(q2, a.noise.n[2]));
        }
    };
    Ax ax = axis(sx, nx), ay = axis(sy, ny), az = axis(sz, nz);
    const float3 s0 = field_from_axes<FT, SAME ? 4 : 3, BRICK>(c, fc, ax, ay, az);
    // the weights are global memory; the cast keeps the batch kernel, which
    // reads this pointer from a descriptor, on scalar loads
    const gfloat* w = (const gfloat*)a.w;
    float acc = (noise(sx, sy, sz, ax, ay, az) * kInv255) * w[0];

#pragma unroll 1
    for (int dirn = 0; dirn < 2; ++dirn) {
        const int steps = dirn == 0 ? a.B : a.F;
        const float h = dirn == 0 ? -a.he : a.he;
        const gfloat* wk = dirn == 0 ? w + 1 : w + 1 + a.B;
        float px = sx, py = sy, pz = sz;
        float3 s = s0;
#pragma unroll 1
        for (int i = 0; i < steps; ++i) {
            // singleLICstep (inc_lic.glsl:93-145), Heun predictor/corrector.
            // 0.5*(d1 + d2) is exact, so fma(0.5, d1 + d2, p) rounds once,
            // as p + 0.5*(d1 + d2) does.
            const float d1x = decode_step(s.x, h);
            const float d1y = decode_step(s.y, h);
            const float d1z = decode_step(s.z, h);
            const float qx = px + d1x, qy = py + d1y, qz = pz + d1z;
            const float3 s2 =
                field_from_axes<FT, SAME ? 4 : 3, BRICK>(c, fc, axis(qx, nx), axis(qy, ny), axis(qz, nz));
            const float d2x = decode_step(s2.x, h);
            const float d2y = decode_step(s2.y, h);
            const float d2z = decode_step(s2.z, h);
            px = fmaf(0.5f, d1x + d2x, px);
            py = fmaf(0.5f, d1y + d2y, py);
            pz = fmaf(0.5f, d1z + d2z, pz);
            ax = axis(px, nx);
            ay = axis(py, ny);
            az = axis(pz, nz);
            float nval;
            if constexpr (SAME) {
                s = field_from_axes<FT, 4, BRICK>(c, fc, ax, ay, az);
                nval = noise_same<FT>(a, c, fc, ax, ay, az);
            } else {
                s = field_and_noise<FT>(a, c, fc, ncache, ax, ay, az, axis(px * a.freq, a.noise.n[0]),
                                        axis(py * a.freq, a.noise.n[1]), axis(pz * a.freq, a.noise.n[2]),
                                        nval);
            }
            acc = acc + (nval * kInv255) * wk[i];
        }
    }
    return acc;
}

__device__ __forceinline__ CachedArgs cached_args(const LicArgs& a)
{
    CachedArgs c;
    c.fr = (const char*)a.field.p;
    c.bsy = (unsigned)a.bsy;
    c.bsz = (unsigned)a.bsz;
    c.nr = (const unsigned char*)a.noise.p;
    c.fnm1[0] = (float)(a.field.n[0] - 1);
    c.fnm1[1] = (float)(a.field.n[1] - 1);
    c.fnm1[2] = (float)(a.field.n[2] - 1);
    c.fsy = (unsigned)a.field.sy;
    c.fsz = (unsigned)a.field.sz;
    c.nsy = (unsigned)a.noise.sy;
    c.nsz = (unsigned)a.noise.sz;
    return c;
}

// LIC of one seed, any noise (scalar or RGBA) and the optional scalar gate:
// the weighted noise sums, before the final scaling.  With a.speed
// (SPEED_OF_FLOW, inc_lic.glsl:107-110,120-122) both displacements of a step
// are scaled by the field magnitude .a at the step's start point (the shader
// scales licdir2 by step.a before step is re-read), in the shader's order:
// ((2s - 1) * a) * h.
// SPEED_OF_FLOW step in the shader's order: ((2s - 1) * a) * h
__device__ __forceinline__ float3 speed_step(const float3& s, float sa, float h)
{
    return make_float3(((2.0f * s.x - 1.0f) * sa) * h, ((2.0f * s.y - 1.0f) * sa) * h,
                       ((2.0f * s.z - 1.0f) * sa) * h);
}

template <typename FT, int NC, bool GATE>
__device__ __forceinline__ float4 lic_point(const LicArgs& a, float sx, float sy, float sz)
{
    const bool speed = a.speed != 0;
    const float4 s0 = speed ? field_rgba<FT>(a, sx, sy, sz) : make_float4(0.f, 0.f, 0.f, 1.f);
    float4 n = noise_at<NC, GATE>(a, sx, sy, sz);
    const float w0 = a.w[0];
    float4 acc;
    acc.x = n.x * w0;
    acc.y = n.y * w0;
    acc.z = n.z * w0;
    acc.w = n.w * w0;
    const float3 s0rgb = speed ? make_float3(s0.x, s0.y, s0.z) : field_rgb<FT>(a, sx, sy, sz);

#pragma unroll 1
    for (int dirn = 0; dirn < 2; ++dirn) {
        const int steps = dirn == 0 ? a.B : a.F;
        const float h = dirn == 0 ? -a.he : a.he;
        const float* wk = dirn == 0 ? a.w + 1 : a.w + 1 + a.B;
        float px = sx, py = sy, pz = sz;
        float3 s = s0rgb;
        float sa = s0.w;
#pragma unroll 1
        for (int i = 0; i < steps; ++i) {
            // singleLICstep (inc_lic.glsl:93-145), Heun predictor/corrector
            const float3 d1 = speed ? speed_step(s, sa, h) : make_float3(decode_step(s.x, h), decode_step(s.y, h),
                                                                         decode_step(s.z, h));
            const float d1x = d1.x, d1y = d1.y, d1z = d1.z;
            const float3 s2 = field_rgb<FT>(a, px + d1x, py + d1y, pz + d1z);
            const float3 d2 = speed ? speed_step(s2, sa, h) : make_float3(decode_step(s2.x, h), decode_step(s2.y, h),
                                                                          decode_step(s2.z, h));
            const float d2x = d2.x, d2y = d2.y, d2z = d2.z;
            px = px + 0.5f * (d1x + d2x);
            py = py + 0.5f * (d1y + d2y);
            pz = pz + 0.5f * (d1z + d2z);
            if (speed) {
                const float4 q = field_rgba<FT>(a, px, py, pz);
                s = make_float3(q.x, q.y, q.z);
                sa = q.w;
            } else {
                s = field_rgb<FT>(a, px, py, pz);
            }
            n = noise_at<NC, GATE>(a, px, py, pz);
            const float wi = wk[i];
            acc.x = acc.x + n.x * wi;
            if constexpr (NC == 4) {
                acc.y = acc.y + n.y * wi;
                acc.z = acc.z + n.z * wi;
                acc.w = acc.w + n.w * wi;
            }
        }
    }
    return acc;
}
