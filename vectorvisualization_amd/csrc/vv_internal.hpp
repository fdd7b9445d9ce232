// vv_internal.hpp — launch descriptors shared by the C-ABI layer (vv_api.cpp)
// and the HIP kernels (vv_lic.hip, vv_render.hip).  Plain structs only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vv {

// Padded volume geometry (see vv_device.hpp for the two padded layouts).
struct VolDesc {
    void* p;
    int n[3];      // logical voxels per axis
    int sy, sz;    // element strides of the padded layout
};

inline VolDesc make_clamp_desc(void* p, const int n[3])
{
    VolDesc d;
    d.p = p;
    for (int a = 0; a < 3; ++a) d.n[a] = n[a];
    d.sy = n[0] + 2;
    d.sz = (n[0] + 2) * (n[1] + 2);
    return d;
}

inline VolDesc make_repeat_desc(void* p, const int n[3])
{
    VolDesc d;
    d.p = p;
    for (int a = 0; a < 3; ++a) d.n[a] = n[a];
    d.sy = n[0] + 1;
    d.sz = (n[0] + 1) * (n[1] + 1);
    return d;
}

inline long long clamp_elems(const int n[3]) { return (long long)(n[0] + 2) * (n[1] + 2) * (n[2] + 2); }
inline long long repeat_elems(const int n[3]) { return (long long)(n[0] + 1) * (n[1] + 1) * (n[2] + 1); }

// bytes per field texel of a storage type (VV_STORE_FP32, _FP16, _UNORM8)
inline unsigned long long field_texel_bytes(int field_type) { return field_type == 1 ? 8ull : field_type == 2 ? 4ull : 16ull; }

struct LicLaunch {
    VolDesc field;     // CLAMP, float4, half4 or UNORM8 x4; with a noise volume on
                       // the same grid (nc 1) the LIC copy (launch_pack_lic_field)
    int field_type;    // 0 float4, 1 half4, 2 UNORM8 x4 (the VV_STORE_* value)
    VolDesc noise;     // REPEAT, uint8 (nc 1) or uchar4 (nc 4)
    int nc;
    VolDesc scalar;    // CLAMP uint8, used when gate
    int gate;
    float gate_lo, gate_hi;
    const float* w;    // 1 + B + F convolution weights
    int F, B;
    float he;          // step_size_lic * eye_factor
    float scl;         // (inv_area / (F+B)) * gradient_scale
    float freq;
    float scale[3];
    VolDesc out;       // REPEAT layout, float (nc 1) or float4 (nc 4)
    int tile[3];       // threads per block along x,y,z
    int super_tile[3]; // tiles per super-tile along x,y,z (tile order)
    int cache;         // 1: per-lane cell caches (scalar noise, no gate)
    unsigned long long field_bytes, noise_bytes;  // padded buffer sizes
    int field_has_noise; // field is the LIC copy: channel 3 = noise byte (same grid)
    int speed;         // SPEED_OF_FLOW: Heun displacements scaled by the field's channel 3
    const void* field_brick;  // float4 LIC copy in 8^3 bricks (launch_pack_lic_field_brick), or null
};

// 8^3-brick order of a CLAMP-padded volume of d + 2 texels per axis: texel
// (x, y, z) at element ((bz * nby + by) * nbx + bx) * 512 + (lz * 8 + ly) * 8
// + lx, b = coordinate / 8, l = coordinate % 8, nb = ceil((d + 2) / 8)
inline void brick_counts(const int d[3], long long nb[3])
{
    for (int a = 0; a < 3; ++a) nb[a] = (d[a] + 2 + 7) / 8;
}

struct ViewBlock {
    float cam[3], du[3], dv[3], dw[3];
    float inv_w2, inv_h2;
    float extent[3], scale[3], tex_max[3];
    int width, height;
};

struct RenderLaunch {
    VolDesc field;     // CLAMP
    int field_type;    // as LicLaunch::field_type
    int tf_ch;         // 2 (.b) or 3 (.a)
    VolDesc lic;       // REPEAT, float or float4
    int lc;            // 1 or 4
    VolDesc sigma;     // CLAMP float, p == nullptr when absent
    const float* tf;   // 256*4 rgba then 256 opacity (device)
    ViewBlock view;
    float step;
    float alpha_corr;
    int corr_mode;     // 0: c==1, 1: sqrt, 2: square, 3: powf
    long long max_samples;
    int illum;         // 0 illumLIC, 1 illumGradient, 2 illumMallo, 3 illumZoeckler
    const float* lut_zoeckler;    // 256x256 x 2 (device), illum 3
    const float* lut_mallo_diff;  // 256x256 x 3 (device), illum 2
    const float* lut_mallo_spec;  // 256x256 x 3 (device), illum 2
    float illum_scale;
    float light[3];    // normalised light direction
    float light_diffuse, light_specular, light_ambient;
    int spec_exp;
    float* image;
    unsigned long long* counter;  // total samples (may be null)
    int block[2];                 // pixel tile (threads), 0: 16x16
};

// Pixel tiles of the G-lane ray casters (G a power of two, 1..64): a block's
// 256/G rays form a TW x TH tile of 2 x 2 wave tiles, each wave's 64/G rays
// a WW x WH tile (WW = WH or 2 WH).
struct RgTile {
    int WW, WH, TW, TH;
};

__host__ __device__ constexpr RgTile rg_tile(int G)
{
    int lw = 0;  // log2 of the rays per wave
    while ((2 << lw) <= 64 / G) ++lw;
    const int ww = 1 << ((lw + 1) >> 1);
    return RgTile{ww, (64 / G) / ww, 2 * ww, 2 * ((64 / G) / ww)};
}

// tiles (blocks) of a width x height image
inline long long rg_blocks(int G, int width, int height)
{
    const RgTile t = rg_tile(G);
    return (long long)((width + t.TW - 1) / t.TW) * ((height + t.TH - 1) / t.TH);
}

struct RenderBatchLaunch {
    const RenderLaunch* views;    // device array, one per view
    const int* first;             // device: first work item of each view
    int nviews, total;            // views, work items (pixel tiles)
    unsigned* counter;            // device work counter, zero before launch
    unsigned long long* done;     // device: completion stamp per view
    unsigned long long* begun;    // device: start stamp per view (all ones before launch)
    int block[2];                 // pixel tile with one lane per pixel (group 1)
    int group;                    // lanes per ray: 1, or 2 / 4 over rg_tile(group) tiles of 256 threads
    int grid;                     // persistent blocks
    int field_type, lc;           // shared by all views
};

struct LicBatchLaunch {
    const void* jobs;             // device array of per-volume kernel arguments
    const int* first;             // device: first tile of each volume
    int njobs, total;             // volumes, tiles (one block each)
    int field_type, same;         // shared by all volumes
};

// Scratch of one adjoint call, carved from one device buffer (the caller of
// launch_render_adjoint owns it).  The parts a forward pass writes come
// first, so a recorded forward pass (vv_render_record) stays valid whatever
// size of accumulators a later adjoint call asks for:
//   header  2 x u64: [0] total samples of all rays (forward pass), [1] low
//           32 bits: the float bits of the largest bound on one sample's
//           gradient contribution (atomicMax; NaN bits when a bound is not
//           finite)
//   ckpt    one of two layouts, per ray:
//           compact (full_col 0): 2*max_ckpt float4 checkpoints (position,
//             D) every kAdjSeg samples, then one segment column (below)
//             that the forward pass leaves holding the last segment and the
//             backward walk refills from the checkpoints;
//           full (full_col 1): max_ckpt segment columns, one per segment,
//             all written by the forward pass (no checkpoints, no re-run)
//           A segment column: kAdjSeg floats (D.a before each sample),
//           kAdjSeg/16 u64 (the clamp masks of 16 samples, 4 bits each),
//           kAdjSeg/16 float4 (the position at the start of each 16-sample
//           part), each array [index][ray]
//   nsamp   per ray sample count (int)
//   bound   per ray bound on its contributions per unit of max |dL/dimage|
//           (float)
//   acc     1024 (+ sigma voxels) int64 fixed-point accumulators
constexpr int kAdjSeg = 64;          // checkpoint interval (samples)
constexpr int kAdjColumnBytes = kAdjSeg * 4 + (kAdjSeg / 16) * (8 + 16);  // per ray

struct AdjointLaunch {
    RenderLaunch r;               // r.image: the forward pass also writes the image (or null)
    const float* dimage;          // width*height*4 (device)
    float* ckpt;                  // per ray: checkpoints, then the segment column
    int max_ckpt;                 // checkpoints per ray
    int* nsamp;                   // per ray sample count
    float* ray_bound;             // per ray bound (per unit of max |dL/dimage|)
    unsigned long long* hdr;      // the two header words
    long long* dtf_fx;            // 256*4 fixed-point accumulators (zero before launch)
    long long* dsigma_fx;         // sigma voxel accumulators, or null
    int use_window;               // accumulate dsigma in an LDS window per segment
    int full_col;                 // the full layout: a column per segment, no checkpoints
    int group;                    // lanes per ray of the forward pass: 2 or 4, else one lane per pixel
};

// Scratch of vv_render_adjoint_noalloc: one static device array per GPU
// (vv_render.hip).  noalloc_scratch() returns the current device's address of
// it (hipGetSymbolAddress, looked up once per device and kept), to be called
// with that device current.
constexpr size_t kNoAllocScratchBytes = 128ull << 20;
hipError_t noalloc_scratch(int device, void** ptr);

// Bytes of adjoint scratch for nrays rays, max_ckpt checkpoints and nv sigma
// voxels (0 without sigma gradients), and the offsets of its parts.
struct AdjScratchLayout {
    size_t ckpt, nsamp, bound, acc, total;
};
inline AdjScratchLayout adj_scratch_layout(long long nrays, int max_ckpt, long long nv, bool full_col = false)
{
    AdjScratchLayout L;
    const size_t align = 256;
    auto up = [&](size_t x) { return (x + align - 1) / align * align; };
    L.ckpt = align;
    const size_t per_ray = full_col ? (size_t)kAdjColumnBytes * (size_t)max_ckpt
                                    : 16 * (size_t)(2 * max_ckpt) + kAdjColumnBytes;
    L.nsamp = up(L.ckpt + per_ray * (size_t)nrays);
    L.bound = up(L.nsamp + sizeof(int) * (size_t)nrays);
    L.acc = up(L.bound + sizeof(float) * (size_t)nrays);
    L.total = up(L.acc + sizeof(long long) * (size_t)(1024 + nv));
    return L;
}

// ---- launchers (vv_lic.hip / vv_render.hip) ----
// the float4 LIC copy (CLAMP-padded, d = field dims) reordered into 8^3 bricks
hipError_t launch_pack_lic_field_brick(const void* copy, const int d[3], void* brick, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, long long n, hipStream_t s);
hipError_t launch_field_maxlen(const float* v0, const float* v1, float t, long long n,
                               uint32_t* maxbits, hipStream_t s);
hipError_t launch_field_encode(const float* v0, const float* v1, float t, const int dims[3],
                               const uint32_t* maxbits, void* out, int field_type, hipStream_t s,
                               int raw = 0);  // raw: a = |v| (fillTexDataFloat), float types only
hipError_t launch_field_decode(const void* field, int field_type, const int dims[3], float* out4,
                               hipStream_t s);
// the LIC copy of a field for a noise volume on the same grid (REPEAT-padded
// scalar noise): rgb of the field, channel 3 the noise byte of the REPEAT
// texel (p - 1) mod n of CLAMP-padded texel p
hipError_t launch_pack_lic_field(const void* field, int field_type, const uint8_t* noise, const int dims[3],
                                 void* out, hipStream_t s);
hipError_t launch_noise_gen(uint64_t seed, int kind, uint64_t thr, const int dims[3],
                            uint8_t* out, hipStream_t s);
hipError_t launch_noise_grad_raw(const uint8_t* noise, const int dims[3], const float sd[3],
                                 float* g, hipStream_t s);
hipError_t launch_noise_grad_filter(const float* g, const int dims[3], const float w27[27],
                                    uint8_t* rgb, hipStream_t s);
hipError_t launch_noise_tex(const uint8_t* noise, const uint8_t* rgb, const int dims[3],
                            void* tex, int nc, hipStream_t s);
hipError_t launch_noise_untex(const void* tex, int nc, const int dims[3], uint8_t* out,
                              hipStream_t s);
hipError_t launch_pad_clamp_u8(const uint8_t* src, const int dims[3], uint8_t* dst, hipStream_t s);
hipError_t launch_pad_clamp_f32(const float* src, const int dims[3], float* dst, hipStream_t s);
hipError_t launch_pad_repeat_f32(const float* src, int nc, const int dims[3], float* dst,
                                 hipStream_t s);
hipError_t launch_unpad_repeat_f32(const float* src, int nc, const int dims[3], float* dst,
                                   hipStream_t s);
hipError_t launch_repeat_halo_f32(float* vol, int nc, const int dims[3], hipStream_t s);
hipError_t launch_lic(const LicLaunch& L, hipStream_t s);
hipError_t launch_render(const RenderLaunch& R, hipStream_t s);
// the same frame with G lanes per ray (G a power of two, 2..64; 256/G-ray pixel tiles)
hipError_t launch_render_rg(const RenderLaunch& R, int G, hipStream_t s);
hipError_t launch_render_batch(const RenderBatchLaunch& B, hipStream_t s);
hipError_t launch_stamp(unsigned long long* t0, hipStream_t s);
// fills the per-volume kernel arguments of a LIC batch (host memory, size
// returned by lic_args_size()), tiles per volume in *ntiles
size_t lic_args_size();
hipError_t lic_args_fill(const LicLaunch& L, void* host_args, int* ntiles);
hipError_t launch_lic_batch(const LicBatchLaunch& B, hipStream_t s);
int render_batch_blocks_per_cu(int field_type, int lc, int threads, int group);
// view-aligned slicing (ViewSlicing): planes v.(p - center) = d_i with
// d_i = -D/2 + (i + 1/2) D/n, i = 0..n-1, front to back
struct SliceSetup {
    float v[3];       // unit view vector in object space (components below 1e-8 zeroed)
    float center[3];  // volume centre (extent / 2)
    float D;          // depth range
    int n;            // number of slices
    int blend;        // 0: RGBA16F render target (lic3d_slicing_fragment.glsl);
                      // 1: blended into an 8-bit RGBA window framebuffer
                      //    (lic3d_slicingblend_fragment.glsl), image = uchar4
};

hipError_t launch_render_slices(const RenderLaunch& R, const LicLaunch& L, const SliceSetup& S, hipStream_t s);

// the fused caster (slices = 0) or slicing (slices = 1) with G lanes per ray
// (G a power of two, 1..64); identical images, more parallelism for small views
hipError_t launch_render_group(const RenderLaunch& R, const LicLaunch& L, const SliceSetup& S, int slices, int G,
                               hipStream_t s);

// fused on-the-fly LIC ray caster: R as for launch_render (its lic fields
// unused), L as for launch_lic (its out fields unused)
hipError_t launch_render_fused(const RenderLaunch& R, const LicLaunch& L, hipStream_t s);
// the forward pass of the adjoint (checkpoints, column, sample counts, ray
// bounds, header word 0; the image when A.r.image is set)
hipError_t launch_adjoint_forward(const AdjointLaunch& A, hipStream_t s);
// the contribution bound of dL/dimage (header word 1), then the backward walk
hipError_t launch_render_adjoint(const AdjointLaunch& A, hipStream_t s);
// dst[i] = src[i] / 2^k for the fixed-point scale 2^k the adjoint derived from
// its header words (NaN when the header holds a non-finite bound)
hipError_t launch_fixed_to_float(const unsigned long long* hdr, const long long* src, float* dst, long long n,
                                 hipStream_t s);
hipError_t launch_fill_i64(long long* p, long long v, long long n, hipStream_t s);

}  // namespace vv
