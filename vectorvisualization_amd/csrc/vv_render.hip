// vv_render.hip — fused front-to-back ray caster (TF lookup + illumination +
// compositing in one kernel) and its hand-written adjoint.
//
//   ray setup     VV/camera.cpp:56-68, VV/renderer.cpp:126-147, 682-736,
//                 VV/shader/raycast_lic3d_fragment.glsl:13-21
//   sample loop   VV/shader/raycast_lic3d_fragment.glsl:35-69
//   illumLIC      VV/shader/inc_illum.glsl:158-172
//   illumGradient VV/shader/inc_illum.glsl:1-41 (fed by the RGBA LIC volume)
//
// One thread per pixel; the 256x4 RGBA transfer function and the 256-entry
// opacity table live in LDS.  The build is -ffp-contract=off; fmaf marks
// every fused operation of the shared formulation.
#include <mutex>
#include <utility>

#include "vv_device.hpp"
#include "vv_internal.hpp"

namespace vv {

namespace {

#include "vv_lic_dev.hpp"

constexpr int kBX = 16, kBY = 16, kRThreads = kBX * kBY;
constexpr int kSeg = kAdjSeg;  // adjoint checkpoint interval (samples)

struct Sample {
    float tfu;     // transfer-function coordinate
    float l;       // LIC intensity
    float m[3];    // d rgb / d tf.rgb, per channel
    float ao;      // opacity-table value
    float sg;      // opacity scale at the sample
    float araw;    // ao * tf.a * sg
    float a;       // corrected alpha
    float rgb[3];  // colour before premultiplication
    float tfa;     // tf.a
    bool aclamp;   // opacity-scale clamp active
};

struct Ray {
    float pos[3], dir[3], V[3];
};

__device__ __forceinline__ void normalize3(const float v[3], float o[3])
{
    float d = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (d > 0.0f) {
        o[0] = v[0] / d;
        o[1] = v[1] / d;
        o[2] = v[2] / d;
    } else {
        o[0] = o[1] = o[2] = 0.0f;
    }
}

__device__ __forceinline__ float ipow(float x, int e)
{
    float r = 1.0f, b = x;
    while (e > 0) {
        if (e & 1) r = r * b;
        b = b * b;
        e >>= 1;
    }
    return r;
}

// b^c for b in [0, 1], c > 0, as exp2(c log2 b) in FP32 (within a few ulp of
// powf; the math library's powf extracts exponents through FP64 conversions)
__device__ __forceinline__ float pow01(float b, float c) { return exp2f(c * log2f(b)); }

__device__ __forceinline__ float pow_corr(float b, float c, int mode)
{
    switch (mode) {
    case 0: return b;
    case 1: return sqrtf(b);
    case 2: return b * b;
    default: return pow01(b, c);
    }
}

// per-pixel ray: analytic entry into [0,extent]^3 (the rasterised front
// faces of drawCubeFaces), dir = normalize(entry - camera) * scaleVol
__device__ __forceinline__ bool ray_setup(const ViewBlock& v, int i, int j, Ray& r)
{
    const float nxp = fmaf((float)i + 0.5f, v.inv_w2, -1.0f);
    const float nyp = fmaf((float)j + 0.5f, v.inv_h2, -1.0f);
    float d[3];
    float tn = -INFINITY, tf = INFINITY;
    for (int a = 0; a < 3; ++a) {
        d[a] = fmaf(nxp, v.du[a], fmaf(nyp, v.dv[a], v.dw[a]));
        const float c = v.cam[a], e = v.extent[a];
        if (d[a] != 0.0f) {
            const float inv = 1.0f / d[a];
            const float t0 = (0.0f - c) * inv, t1 = (e - c) * inv;
            const float lo = t0 < t1 ? t0 : t1, hi = t0 < t1 ? t1 : t0;
            if (lo > tn) tn = lo;
            if (hi < tf) tf = hi;
        } else if (c < 0.0f || c > e) {
            return false;
        }
    }
    if (!(tn <= tf) || !(tn > 0.0f)) return false;
    float p0[3], g[3], gn[3];
    for (int a = 0; a < 3; ++a) {
        const float q = fmaf(tn, d[a], v.cam[a]);
        const float e = v.extent[a];
        p0[a] = q < 0.0f ? 0.0f : (q > e ? e : q);
        g[a] = p0[a] - v.cam[a];
    }
    normalize3(g, gn);
    for (int a = 0; a < 3; ++a) {
        r.dir[a] = gn[a] * v.scale[a];
        r.pos[a] = p0[a] * v.scale[a];
    }
    const float vd[3] = {-r.dir[0], -r.dir[1], -r.dir[2]};
    normalize3(vd, r.V);
    return true;
}

__device__ __forceinline__ bool outside(const float p[3], const ViewBlock& v)
{
    return p[0] < 0.0f || p[0] > v.tex_max[0] || p[1] < 0.0f || p[1] > v.tex_max[1] ||
           p[2] < 0.0f || p[2] > v.tex_max[2];
}

// Shading and opacity of one sample from its LIC value(s) l4 (LC channels):
// TF lookup, illumination mode, opacity table, opacity scale and the alpha
// correction.  fv/cf: the field and its cell at pos.
template <typename FT, int LC>
__device__ __forceinline__ void shade_sample(const RenderLaunch& a, const float* tf, const float* op,
                                             const Ray& r, const float pos[3], const Vol<FT>& fv, const Cell& cf,
                                             const float l4[4], Sample& s)
{
    const float v = tri_ch(fv, cf, a.tf_ch);
    s.tfu = v;
    float t4[4];
    for (int ch = 0; ch < 4; ++ch) t4[ch] = lut1d(tf, 256, 4, ch, v);
    s.tfa = t4[3];
    const float g = a.illum_scale;
    const float l = LC == 1 ? l4[0] : l4[3];
    s.l = l;
    if (a.illum >= 2) {
        // line illumination with the field direction as tangent
        // (vectorData.xyz; tangent = normalize(2 t - 1))
        const float3 fr = tri_rgb(fv, cf);
        const float tg[3] = {fmaf(2.0f, fr.x, -1.0f), fmaf(2.0f, fr.y, -1.0f), fmaf(2.0f, fr.z, -1.0f)};
        float T[3];
        normalize3(tg, T);
        const float* L = a.light;
        if (a.illum == 3) {
            // illumZoeckler (inc_illum.glsl:117-154); the table is LUMINANCE_ALPHA,
            // so .rg reads (L, L): both shading terms are the luminance channel
            float V[3];
            normalize3(r.dir, V);
            const float u = fmaf(0.5f, L[0] * T[0] + L[1] * T[1] + L[2] * T[2], 0.5f);
            const float w = fmaf(0.5f, V[0] * T[0] + V[1] * T[1] + V[2] * T[2], 0.5f);
            const float lum = lut2d(a.lut_zoeckler, 2, 0, u, w);
            const float shr = lum * g;
            const float sp = 0.9f * lum;
            for (int c = 0; c < 3; ++c) {
                s.rgb[c] = (t4[c] * l) * shr + sp;
                s.m[c] = l * shr;
            }
        } else {
            // illumMallo (inc_illum.glsl:46-113)
            const float vd[3] = {-r.dir[0], -r.dir[1], -r.dir[2]};
            float V[3], bn[3], Bv[3], N[3], hv[3], H[3];
            normalize3(vd, V);
            bn[0] = T[1] * V[2] - T[2] * V[1];
            bn[1] = T[2] * V[0] - T[0] * V[2];
            bn[2] = T[0] * V[1] - T[1] * V[0];
            normalize3(bn, Bv);
            N[0] = Bv[1] * T[2] - Bv[2] * T[1];
            N[1] = Bv[2] * T[0] - Bv[0] * T[2];
            N[2] = Bv[0] * T[1] - Bv[1] * T[0];
            for (int c = 0; c < 3; ++c) hv[c] = V[c] + L[c];
            normalize3(hv, H);
            const float lx = L[0] * N[0] + L[1] * N[1] + L[2] * N[2];
            const float ly = L[0] * T[0] + L[1] * T[1] + L[2] * T[2];
            const float lz = H[0] * N[0] + H[1] * N[1] + H[2] * N[2];
            const float lw = H[0] * T[0] + H[1] * T[1] + H[2] * T[2];
            const float ql = sqrtf(fmaxf(1.0f - ly * ly, 0.0f)), qh = sqrtf(fmaxf(1.0f - lw * lw, 0.0f));
            const float tx = 1.0f / ql, ty = 1.0f / qh;
            const float ud = fmaf(0.5f, lx * tx, 0.5f);
            const float vdc = fmaf(0.5f, ly, 0.5f);
            const float vs = fmaf(0.5f, lz * ty, 0.5f);
            // pow(tmp.y, -spotExponent) = (1 - (H.T)^2)^(n/2)
            const float att = ipow(qh, a.spec_exp);
            for (int c = 0; c < 3; ++c) {
                const float dif = lut2d(a.lut_mallo_diff, 3, c, ud, vdc) * g;
                const float spc = clamp01(lut2d(a.lut_mallo_spec, 3, c, ud, vs) * att);
                s.rgb[c] = (t4[c] * l) * dif + spc;
                s.m[c] = l * dif;
            }
        }
    } else if (LC == 1 || a.illum == 0) {
        // illumLIC: rgb = illum * tf.rgb * gradient.g
        for (int c = 0; c < 3; ++c) s.rgb[c] = l * t4[c] * g;
        s.m[0] = s.m[1] = s.m[2] = l * g;
    } else {
        // illumGradient: N = normalize(-I.xyz), L = normalize(light) (scaleVolInv = 0)
        const float nn[3] = {-l4[0], -l4[1], -l4[2]};
        float N[3], rr[3], Rv[3];
        normalize3(nn, N);
        const float ldn = a.light[0] * N[0] + a.light[1] * N[1] + a.light[2] * N[2];
        for (int c = 0; c < 3; ++c) rr[c] = fmaf(2.0f * ldn, N[c], -a.light[c]);
        normalize3(rr, Rv);
        float sp = clamp01(Rv[0] * r.V[0] + Rv[1] * r.V[1] + Rv[2] * r.V[2]);
        sp = ipow(sp, a.spec_exp);
        sp = sp * l * a.light_specular;
        const float diff = clamp01(ldn) * a.light_diffuse * g;
        const float kk = (diff + 0.3f) + a.light_ambient;
        for (int c = 0; c < 3; ++c) {
            const float col = t4[c] * l;
            s.rgb[c] = fmaf(col, kk, sp) * g;
        }
        s.m[0] = s.m[1] = s.m[2] = l * kk * g;
    }
    s.ao = lut1d(op, 256, 1, 0, l * 1.3f);
    float al = s.ao * t4[3];
    s.sg = 1.0f;
    s.aclamp = false;
    if (a.sigma.p) {
        // per-voxel opacity scale: alpha * sigma, clamped to [0,1]
        const Vol<float> sv{(const float*)a.sigma.p, {a.sigma.n[0], a.sigma.n[1], a.sigma.n[2]},
                            a.sigma.sy, a.sigma.sz};
        const Cell cs = cell_clamp(sv, pos[0], pos[1], pos[2]);
        s.sg = tri_f(sv, cs);
        al = al * s.sg;
        if (al > 1.0f) {
            al = 1.0f;
            s.aclamp = true;
        }
        if (al < 0.0f) {
            al = 0.0f;
            s.aclamp = true;
        }
    }
    s.araw = al;
    s.a = 1.0f - pow_corr(1.0f - al, a.alpha_corr, a.corr_mode);
}

// One sample of the LIC-volume ray caster: LIC value(s) trilinear from the
// volume, then shade_sample.
template <typename FT, int LC>
__device__ __forceinline__ void eval_sample(const RenderLaunch& a, const float* tf, const float* op,
                                            const Ray& r, const float pos[3], Sample& s)
{
    const Vol<FT> fv{(const FT*)a.field.p, {a.field.n[0], a.field.n[1], a.field.n[2]}, a.field.sy,
                     a.field.sz};
    const Cell cf = cell_clamp(fv, pos[0], pos[1], pos[2]);
    float l4[4];
    if constexpr (LC == 1) {
        const Vol<float> lv{(const float*)a.lic.p, {a.lic.n[0], a.lic.n[1], a.lic.n[2]}, a.lic.sy,
                            a.lic.sz};
        const Cell cl = cell_repeat(lv, pos[0], pos[1], pos[2]);
        l4[0] = tri_f(lv, cl);
    } else {
        const Vol<float4> lv{(const float4*)a.lic.p, {a.lic.n[0], a.lic.n[1], a.lic.n[2]}, a.lic.sy,
                             a.lic.sz};
        const Cell cl = cell_repeat(lv, pos[0], pos[1], pos[2]);
        const float4 q = tri_f4(lv, cl);
        l4[0] = q.x;
        l4[1] = q.y;
        l4[2] = q.z;
        l4[3] = q.w;
    }
    shade_sample<FT, LC>(a, tf, op, r, pos, fv, cf, l4, s);
}

__device__ __forceinline__ void load_tables(const float* g, float* tf, float* op)
{
    const int nt = blockDim.x * blockDim.y;
    for (int i = threadIdx.y * blockDim.x + threadIdx.x; i < 1280; i += nt) {
        float v = g[i];
        if (i < 1024)
            tf[i] = v;
        else
            op[i - 1024] = v;
    }
    __syncthreads();
}

// composite one sample; returns the unclamped-out-of-range mask (bit c)
__device__ __forceinline__ int composite(float dest[4], const Sample& s)
{
    const float src[4] = {s.rgb[0] * s.a, s.rgb[1] * s.a, s.rgb[2] * s.a, s.a};
    const float t = 1.0f - dest[3];
    int mask = 0;
    for (int c = 0; c < 4; ++c) {
        const float u = fmaf(t, src[c], dest[c]);
        if (u < 0.0f || u > 1.0f) mask |= 1 << c;
        dest[c] = clamp01(u);
    }
    return mask;
}

// one pixel: ray setup, sample loop, image store; returns the sample count
template <typename FT, int LC>
__device__ __forceinline__ unsigned render_pixel(const RenderLaunch& a, const float* tf, const float* op,
                                                 int i, int j)
{
    unsigned ns = 0;
    float dest[4] = {0.f, 0.f, 0.f, 0.f};
    Ray r;
    if (ray_setup(a.view, i, j, r)) {
        float pos[3] = {r.pos[0], r.pos[1], r.pos[2]};
        for (long long k = 0; k < a.max_samples; ++k) {
            Sample s;
            eval_sample<FT, LC>(a, tf, op, r, pos, s);
            composite(dest, s);
            ++ns;
            for (int c = 0; c < 3; ++c) pos[c] = fmaf(r.dir[c], a.step, pos[c]);
            if (outside(pos, a.view) || dest[3] > 0.95f) break;
        }
    }
    const long long pix = (long long)j * a.view.width + i;
    reinterpret_cast<float4*>(a.image)[pix] = make_float4(dest[0], dest[1], dest[2], dest[3]);
    return ns;
}

// block-wide sum of per-thread sample counts into *counter
__device__ __forceinline__ void add_samples(unsigned long long* counter, unsigned long long ns,
                                            unsigned long long* red)
{
    for (int o = 32; o > 0; o >>= 1) ns += __shfl_xor(ns, o, 64);
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const int nw = (blockDim.x * blockDim.y + 63) / 64;
    if ((tid & 63) == 0) red[tid >> 6] = ns;
    __syncthreads();
    if (tid == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < nw; ++w) t += red[w];
        atomicAdd(counter, t);
    }
}

// blockDim = pixel tile (any shape of at most 256 threads, a multiple of 64)
template <typename FT, int LC>
__global__ void __launch_bounds__(kRThreads) k_render(RenderLaunch a)
{
    __shared__ float tf[1024];
    __shared__ float op[256];
    __shared__ unsigned long long red[kRThreads / 64];
    load_tables(a.tf, tf, op);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    unsigned long long ns = 0;
    if (i < a.view.width && j < a.view.height) ns = render_pixel<FT, LC>(a, tf, op, i, j);
    if (a.counter) add_samples(a.counter, ns, red);
}

// Lane g of each G-lane group to every lane of the group, G = 2 or 4: the
// groups are aligned to the quads of the wave, so one DPP quad permutation
// ((g, g, g, g) for 4 lanes, (g, g, 2 + g, 2 + g) for 2) moves the value.
template <int G, int g>
__device__ __forceinline__ float group_bcast(float v)
{
    static_assert(G == 2 || G == 4, "quad groups only");
    constexpr int ctrl = G == 4 ? g * 0x55 : (g | (g << 2) | ((2 + g) << 4) | ((2 + g) << 6));
    return __builtin_bit_cast(float,
                              __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), ctrl, 0xF, 0xF, true));
}

// sv[g] = src of lane g of the caller's group, for every g < G
template <int G, int... g>
__device__ __forceinline__ void group_gather(const float src[4], float (*sv)[4], int gl0,
                                             std::integer_sequence<int, g...>)
{
    if constexpr (G <= 4) {
        ((sv[g][0] = group_bcast<G, g>(src[0]), sv[g][1] = group_bcast<G, g>(src[1]),
          sv[g][2] = group_bcast<G, g>(src[2]), sv[g][3] = group_bcast<G, g>(src[3])),
         ...);
    } else {
        ((sv[g][0] = __shfl(src[0], gl0 + g), sv[g][1] = __shfl(src[1], gl0 + g),
          sv[g][2] = __shfl(src[2], gl0 + g), sv[g][3] = __shfl(src[3], gl0 + g)),
         ...);
    }
}

// Pixel tiles of the G-lane ray casters (rg_tile): blocks take the tiles of
// the image row by row.
constexpr bool rg_tiles_fit()
{
    for (int G = 1; G <= 64; G *= 2) {
        const RgTile t = rg_tile(G);
        if (t.WW * t.WH * G != 64 || t.TW * t.TH * G != kRThreads || t.WW < t.WH || t.WW > 2 * t.WH) return false;
    }
    return true;
}
static_assert(rg_tiles_fit(), "a wave tile holds 64/G rays and a block tile 256/G");

// pixel (i, j) of thread tid of block b
__device__ __forceinline__ void rg_pixel(int G, int width, int b, int tid, int& i, int& j)
{
    const RgTile t = rg_tile(G);
    const int tiles_x = (width + t.TW - 1) / t.TW;
    const int rw = 64 / G, r_in = tid / G;
    const int w = r_in / rw, lr = r_in % rw;
    i = (b % tiles_x) * t.TW + (w & 1) * t.WW + lr % t.WW;
    j = (b / tiles_x) * t.TH + (w >> 1) * t.WH + lr / t.WW;
}

// The LIC-volume ray caster with G lanes per ray (G a power of two, 2..64),
// for views whose longest rays would otherwise set the frame time: in each
// round lane g of a ray's group evaluates sample base + g, stepping from the
// round's first position with render_pixel's fma sequence, and the group
// composites the G samples in order with composite()'s arithmetic and
// render_pixel's stopping rules (a sample outside the box or past
// max_samples ends the ray, so does dest.a > 0.95 after a sample).  A lane
// tests only its own position: compositing stops at the first lane whose
// sample is not valid, and that lane's test is the one render_pixel makes
// at the same step.  Images and sample counts equal k_render's for every G.
// A block's 256/G rays form an rg_tile(G) pixel tile.  GT > 0 fixes G at
// compile time: the group's G samples then reach every lane through
// shuffles issued back to back (DPP quad permutations for 2 and 4 lanes)
// and one ballot before the ordered compositing, instead of a shuffle round
// trip per sample (GT = 0: G at run time, for 16 lanes and more).  The batch
// ray caster runs the same routine per ray.
//
// render_ray_rg: one ray of a G-lane group, the group of thread tid (tid's
// lane in its wave fixes the group): the rounds, the ordered compositing and
// the image store.  Returns the ray's sample count in the group's first lane
// and 0 in the others.  Every lane of the group must call it together.
template <typename FT, int LC, int GT>
__device__ __forceinline__ unsigned long long render_ray_rg(const RenderLaunch& a, const float* tf, const float* op,
                                                            int G, int tid, int i, int j)
{
    const int sub = tid & (G - 1);
    const int gl0 = (tid & 63) & ~(G - 1);  // first lane of the group in the wave
    const int W = a.view.width, H = a.view.height;
    const bool has_pix = i < W && j < H;
    float dest[4] = {0.f, 0.f, 0.f, 0.f};
    unsigned long long ns = 0;
    Ray r;
    // every lane of a group follows the same ray, so alive is uniform per group
    // and the shuffles below read lanes that run the loop with it
    bool alive = has_pix && ray_setup(a.view, i, j, r);
    float base[3] = {0.f, 0.f, 0.f};
    if (alive)
        for (int q = 0; q < 3; ++q) base[q] = r.pos[q];
    long long bk = 0;
    while (alive) {
        float pos[3] = {base[0], base[1], base[2]};
        for (int t = 0; t < sub; ++t)
            for (int q = 0; q < 3; ++q) pos[q] = fmaf(r.dir[q], a.step, pos[q]);
        const bool ok = bk + sub < a.max_samples && (sub == 0 || !outside(pos, a.view));
        float src[4] = {0.f, 0.f, 0.f, 0.f};
        if (ok) {
            Sample s;
            eval_sample<FT, LC>(a, tf, op, r, pos, s);
            src[0] = s.rgb[0] * s.a;
            src[1] = s.rgb[1] * s.a;
            src[2] = s.rgb[2] * s.a;
            src[3] = s.a;
        }
        // the position after this lane's sample: the next round's first one
        // when this is the group's last lane
        float nxt[3];
        for (int q = 0; q < 3; ++q) nxt[q] = fmaf(r.dir[q], a.step, pos[q]);
        const bool nxt_out = !ok || outside(nxt, a.view);
        if constexpr (GT > 0) {
            float sv[GT][4];
            group_gather<GT>(src, sv, gl0, std::make_integer_sequence<int, GT>{});
            // the group's ok bits; every lane of the group runs this code
            const unsigned okg = (unsigned)(__ballot((int)ok) >> gl0) & ((1u << GT) - 1u);
#pragma unroll
            for (int g = 0; g < GT; ++g) {
                if (!((okg >> g) & 1u)) {
                    alive = false;
                    break;
                }
                const float t = 1.0f - dest[3];
                for (int q = 0; q < 4; ++q) dest[q] = clamp01(fmaf(t, sv[g][q], dest[q]));
                ++ns;
                if (dest[3] > 0.95f) {
                    alive = false;
                    break;
                }
            }
        } else {
            for (int g = 0; g < G; ++g) {
                const int sl = gl0 + g;
                if (__shfl((int)ok, sl) == 0) {
                    alive = false;
                    break;
                }
                float sv[4];
                for (int q = 0; q < 4; ++q) sv[q] = __shfl(src[q], sl);
                const float t = 1.0f - dest[3];
                for (int q = 0; q < 4; ++q) dest[q] = clamp01(fmaf(t, sv[q], dest[q]));
                ++ns;
                if (dest[3] > 0.95f) {
                    alive = false;
                    break;
                }
            }
        }
        if (alive) {
            const int last = gl0 + G - 1;
            if constexpr (GT > 0 && GT <= 4) {
                for (int q = 0; q < 3; ++q) base[q] = group_bcast<GT, GT - 1>(nxt[q]);
            } else {
                for (int q = 0; q < 3; ++q) base[q] = __shfl(nxt[q], last);
            }
            if ((__ballot((int)nxt_out) >> last) & 1ull) alive = false;
        }
        bk += G;
    }
    if (has_pix && sub == 0) {
        const long long pix = (long long)j * W + i;
        reinterpret_cast<float4*>(a.image)[pix] = make_float4(dest[0], dest[1], dest[2], dest[3]);
    }
    return sub == 0 ? ns : 0ull;
}

template <typename FT, int LC, int GT>
__global__ void __launch_bounds__(kRThreads) k_render_rg(RenderLaunch a, int G_rt)
{
    const int G = GT > 0 ? GT : G_rt;
    __shared__ float tf[1024];
    __shared__ float op[256];
    __shared__ unsigned long long red[kRThreads / 64];
    load_tables(a.tf, tf, op);
    const int tid = threadIdx.x;
    int i, j;
    rg_pixel(G, a.view.width, (int)blockIdx.x, tid, i, j);
    const unsigned long long ns = render_ray_rg<FT, LC, GT>(a, tf, op, G, tid, i, j);
    if (a.counter) add_samples(a.counter, ns, red);
}

// Persistent ray caster over the pixel tiles of many views.  Blocks take
// work items (view, tile) from a counter (items of different views cost
// different amounts); the start of each view's first tile and the end of its
// last tile are recorded as the smallest and largest wall-clock stamps its
// tiles write.
// GT > 0: GT lanes per ray (2 or 4) in 256-thread blocks over rg_tile(GT)
// pixel tiles; GT = 0: one lane per pixel over blockDim pixel tiles.  Four
// waves per SIMD (at most 128 VGPRs): the scalar-LIC grouped kernels spill 2-4
// registers and still run the batch faster than at 143 VGPRs and 3 waves.
template <typename FT, int LC, int GT>
__global__ void __launch_bounds__(kRThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_render_batch(const RenderLaunch* views, const int* first,
                                                           int nviews, int total, unsigned* counter,
                                                           unsigned long long* done, unsigned long long* begun)
{
    __shared__ float tf[1024];
    __shared__ float op[256];
    __shared__ unsigned long long red[kRThreads / 64];
    __shared__ int item_s;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    int cur = -1;
    for (;;) {
        if (tid == 0) item_s = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int item = item_s;
        __syncthreads();
        if (item >= total) break;
        int v = 0;
        while (v + 1 < nviews && item >= first[v + 1]) ++v;
        const RenderLaunch& a = views[v];
        if (tid == 0) atomicMin(&begun[v], (unsigned long long)wall_clock64());
        if (v != cur) {
            load_tables(a.tf, tf, op);
            cur = v;
        }
        const int t = item - first[v];
        unsigned long long ns = 0;
        if constexpr (GT > 0) {
            int i, j;
            rg_pixel(GT, a.view.width, t, tid, i, j);
            ns = render_ray_rg<FT, LC, GT>(a, tf, op, GT, tid, i, j);
        } else {
            const int tx = (a.view.width + blockDim.x - 1) / blockDim.x;
            const int i = (t % tx) * blockDim.x + threadIdx.x;
            const int j = (t / tx) * blockDim.y + threadIdx.y;
            if (i < a.view.width && j < a.view.height) ns = render_pixel<FT, LC>(a, tf, op, i, j);
        }
        if (a.counter) add_samples(a.counter, ns, red);
        __syncthreads();
        if (tid == 0) atomicMax(&done[v], (unsigned long long)wall_clock64());
    }
}

__global__ void k_stamp(unsigned long long* t0) { *t0 = (unsigned long long)wall_clock64(); }

// ---- adjoint -----------------------------------------------------------------
//
// Forward  D' = clamp(D + (1 - D.a) s, 0, 1),  s = (rgb a, a).
// Backward G -> G masked where the clamp was active;
//          dL/ds = (1 - D.a) G;  dL/dD = G - (G . s) e_a.
// Phase 1 (k_adjoint_forward) runs the ray once and stores (pos, D) every
// kSeg samples.  Phase 2 (k_render_adjoint) walks the segments last to
// first: it re-runs a segment from its checkpoint, saving D per sample in a
// per-ray scratch column, then walks it backwards, re-evaluating each sample.
//
// Gradient contributions are accumulated as 64-bit integers, so the sums do
// not depend on thread order.  A contribution v becomes rint(v * 2^k).  k is
// chosen per call, on the device, so that no sum can leave the int64 range:
// phase 1 bounds every contribution a ray will make (below) and counts the
// samples of all rays, Ns.  One sample adds at most |v| + 4 quanta to any
// accumulator (its two TF corners, or its eight sigma corners, have weights
// summing to 1), so with C the largest bound and 2^k * C <= 2^61 / Ns, every
// sum stays within Ns (C 2^k + 4) <= 2^62.  The scale follows the gradient's
// own magnitude: multiplying dL/dimage by a power of two shifts k and leaves
// the integer sums unchanged.
//
// The bound of one ray (G = the pixel's dL/dimage, |G|max its largest entry).
// Along the backward walk the rgb entries of G keep their value or become 0,
// and G.a changes by -sum_c G.c src.c - G.a src.a with src.a in [0,1], so
// |G.a| <= |G|max * Gamma, Gamma = 1 + sum over the ray's samples of
// sum_c |src.c|.  Per sample, with t = 1 - D.a in [0,1]:
//   TF rgb       |t G.c a m.c|                      <= |G|max a |m.c|
//   TF alpha     |da dadraw ao sg|,  |da| <= |G|max (Gamma + sum_c |rgb.c|)
//   sigma        |da dadraw ao tf.a|
// so the ray's bound is |G|max * max(max a|m|, Gamma max q + max q R) with
// q = dadraw ao max(sg, tf.a) and R = sum_c |rgb.c|.  Rounding in these float
// sums is far inside the factor 2 between 2^62 and the int64 limit.

// d(corrected alpha) / d(raw alpha): 1 - (1 - a)^c differentiated
__device__ __forceinline__ float dalpha_draw(const RenderLaunch& a, float araw)
{
    const float b = 1.0f - araw;
    if (a.corr_mode == 0) return 1.0f;
    if (b > 0.0f) return a.alpha_corr * pow01(b, a.alpha_corr - 1.0f);
    return 0.0f;
}

// exponent k of the fixed-point scale 2^k from the header words; false when
// the bound is not finite (the gradients are then NaN) or zero (all
// contributions are zero)
__device__ __forceinline__ bool adj_scale_exp(const unsigned long long* hdr, int& k)
{
    const unsigned cb_bits = (unsigned)hdr[1];
    const unsigned long long ns = hdr[0];
    const float cb = __uint_as_float(cb_bits);
    if (!(cb > 0.0f) || cb_bits >= 0x7f800000u || ns == 0) return false;
    int ec;
    (void)frexpf(cb, &ec);  // cb < 2^ec
    const int en = ns <= 1 ? 0 : 64 - __clzll((long long)(ns - 1));  // ns <= 2^en
    k = 61 - en - ec;
    return true;
}

__device__ __forceinline__ long long to_fx(float v, int k) { return __float2ll_rn(ldexpf(v, k)); }

__device__ __forceinline__ void splat_tf(long long* acc, int ch, float u, float val, int k)
{
    const float x = fmaf(u, 256.0f, -0.5f);
    float fl = floorf(x);
    const float f = x - fl;
    fl = fminf(fmaxf(fl, -1.0f), 255.0f);
    const int i = (int)fl;
    const int i0 = i < 0 ? 0 : i;
    const int i1 = i + 1 > 255 ? 255 : i + 1;
    const long long a0 = to_fx((1.0f - f) * val, k);
    const long long a1 = to_fx(f * val, k);
    if (a0) atomicAdd((unsigned long long*)&acc[i0 * 4 + ch], (unsigned long long)a0);
    if (a1) atomicAdd((unsigned long long*)&acc[i1 * 4 + ch], (unsigned long long)a1);
}

// dsigma accumulation window in LDS: entries (int64 fixed point), one
// window per kSub samples of a segment.  Sized for occupancy: 24 KB of
// window (38 KB of LDS per block) lets four blocks share a CU, matching the
// kernel's 4 waves per SIMD (at most 128 VGPRs), and the voxel box of 16
// samples of a 16x16-ray tile fits it at 256^3 / 1024^2 (a box that does
// not fit goes to global atomics, with the same sums).
constexpr int kWin = 3072;
constexpr int kSub = 16;
constexpr int kNSub = kSeg / kSub;
static_assert(kSub == 16 && kSeg * 4 + kNSub * 24 == kAdjColumnBytes, "segment column layout");

// The segment column of a ray: [k][ray] D.a before sample k of the segment;
// [m][ray] the clamp masks of part m's samples (4 bits each, sample j of the
// part at bit 4j); [m][ray] the position at the start of part m.  Compact
// layout: one column, which phase 1 fills for every segment in turn, so it
// holds each ray's last segment when phase 2 starts, and phase 2 re-runs the
// earlier segments into it from their checkpoints.  Full layout (A.full_col):
// a column per segment, all filled by phase 1.
struct AdjColumn {
    float* da;
    unsigned long long* mask;
    float4* pos;
};

__device__ __forceinline__ AdjColumn adj_column(const AdjointLaunch& A, long long nrays, long long seg)
{
    AdjColumn c;
    char* base = A.full_col ? reinterpret_cast<char*>(A.ckpt) + seg * (long long)kAdjColumnBytes * nrays
                            : reinterpret_cast<char*>(reinterpret_cast<float4*>(A.ckpt) +
                                                      (long long)2 * A.max_ckpt * nrays);
    c.da = reinterpret_cast<float*>(base);
    c.mask = reinterpret_cast<unsigned long long*>(c.da + (long long)kSeg * nrays);
    c.pos = reinterpret_cast<float4*>(c.mask + (long long)kNSub * nrays);
    return c;
}

// Phase 1: the forward ray with a checkpoint (pos, D) every kSeg samples, the
// segment column, its sample count and its contribution bound per unit of
// max |dL/dimage|, the sample total of the header, and, when a.image is set,
// the image (the same samples and arithmetic as k_render).
template <typename FT, int LC>
__global__ void __launch_bounds__(kRThreads) k_adjoint_forward(AdjointLaunch A)
{
    __shared__ float tf[1024];
    __shared__ float op[256];
    __shared__ unsigned long long red[kRThreads / 64];
    const RenderLaunch& a = A.r;
    const int tid = threadIdx.y * kBX + threadIdx.x;
    load_tables(a.tf, tf, op);
    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const long long W = a.view.width, H = a.view.height;
    const long long nrays = W * H;
    const long long pix = (long long)j * W + i;
    float4* ck = reinterpret_cast<float4*>(A.ckpt);  // compact: [c][2][ray] checkpoint (pos, D)
    AdjColumn col = adj_column(A, nrays, 0);
    Ray r;
    const bool inside = i < W && j < H;
    const bool active = inside && ray_setup(a.view, i, j, r);
    long long n = 0;
    float bound = 0.0f;
    float dest[4] = {0.f, 0.f, 0.f, 0.f};
    if (active) {
        float pos[3] = {r.pos[0], r.pos[1], r.pos[2]};
        float max_am = 0.0f, max_q = 0.0f, max_qr = 0.0f, gamma = 1.0f;
        unsigned long long mbits = 0;
        for (long long k = 0; k < a.max_samples; ++k) {
            const int kk = (int)(k % kSeg);
            if (kk == 0) {
                const long long c = k / kSeg;
                if (A.full_col) {
                    col = adj_column(A, nrays, c);
                } else {
                    ck[(2 * c) * nrays + pix] = make_float4(pos[0], pos[1], pos[2], 0.f);
                    ck[(2 * c + 1) * nrays + pix] = make_float4(dest[0], dest[1], dest[2], dest[3]);
                }
            }
            if (kk % kSub == 0) col.pos[(long long)(kk / kSub) * nrays + pix] = make_float4(pos[0], pos[1], pos[2], 0.f);
            Sample s;
            eval_sample<FT, LC>(a, tf, op, r, pos, s);
            col.da[(long long)kk * nrays + pix] = dest[3];
            mbits |= (unsigned long long)composite(dest, s) << (4 * (kk % kSub));
            if (kk % kSub == kSub - 1) {
                col.mask[(long long)(kk / kSub) * nrays + pix] = mbits;
                mbits = 0;
            }
            ++n;
            const float rsum = fabsf(s.rgb[0]) + fabsf(s.rgb[1]) + fabsf(s.rgb[2]);
            const float mmax = fmaxf(fabsf(s.m[0]), fmaxf(fabsf(s.m[1]), fabsf(s.m[2])));
            const float q = s.aclamp ? 0.0f : dalpha_draw(a, s.araw) * s.ao * fmaxf(s.sg, s.tfa);
            max_am = fmaxf(max_am, s.a * mmax);
            max_q = fmaxf(max_q, q);
            max_qr = fmaxf(max_qr, q * rsum);
            gamma = fmaf(s.a, rsum, gamma);
            for (int c = 0; c < 3; ++c) pos[c] = fmaf(r.dir[c], a.step, pos[c]);
            if (outside(pos, a.view) || dest[3] > 0.95f) break;
        }
        // the masks of a last part the loop left unfinished
        const int kl = (int)((n - 1) % kSeg);
        if (kl % kSub != kSub - 1) col.mask[(long long)(kl / kSub) * nrays + pix] = mbits;
        bound = fmaxf(max_am, fmaf(gamma, max_q, max_qr));
        if (!isfinite(bound)) bound = __uint_as_float(0x7fc00000u);
    }
    if (inside) {
        A.nsamp[pix] = (int)n;
        A.ray_bound[pix] = bound;
        if (a.image) reinterpret_cast<float4*>(a.image)[pix] = make_float4(dest[0], dest[1], dest[2], dest[3]);
    }
    unsigned long long ns = (unsigned long long)n;
    for (int o = 32; o > 0; o >>= 1) ns += __shfl_xor(ns, o, 64);
    if ((tid & 63) == 0) red[tid >> 6] = ns;
    __syncthreads();
    if (tid == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kRThreads / 64; ++w) t += red[w];
        if (t) {
            atomicAdd(&A.hdr[0], t);
            if (a.counter) atomicAdd(a.counter, t);
        }
    }
}

// Phase 1 with GT lanes per ray (2 or 4, the groups are quads), for the
// views where one lane per pixel leaves the longest rays to set the time:
// lanes evaluate consecutive samples as k_render_rg does, the group
// composites them in order and sums gamma in order (every lane of the group
// runs the same compositing), and each lane then writes its own sample's
// column entries (D.a before it, the part's start position, the checkpoint)
// while the part's clamp masks, the same in every lane, are written by the
// group's first lane.  Columns, checkpoints, counts, bounds and image are
// k_adjoint_forward's bits.
template <typename FT, int LC, int GT>
__global__ void __launch_bounds__(kRThreads) k_adjoint_forward_rg(AdjointLaunch A)
{
    static_assert(GT == 2 || GT == 4, "quad groups only");
    static_assert(kSub % GT == 0, "a round lies in one part");
    __shared__ float tf[1024];
    __shared__ float op[256];
    __shared__ unsigned long long red[kRThreads / 64];
    const RenderLaunch& a = A.r;
    load_tables(a.tf, tf, op);
    const int tid = threadIdx.x;
    const int sub = tid & (GT - 1);
    const int gl0 = (tid & 63) & ~(GT - 1);
    const int W = a.view.width, H = a.view.height;
    int i, j;
    rg_pixel(GT, W, (int)blockIdx.x, tid, i, j);
    const long long nrays = (long long)W * H;
    const long long pix = (long long)j * W + i;
    float4* ck = reinterpret_cast<float4*>(A.ckpt);  // compact: [c][2][ray] checkpoint (pos, D)
    Ray r;
    const bool inside = i < W && j < H;
    bool alive = inside && ray_setup(a.view, i, j, r);
    const bool active = alive;
    long long n = 0;
    float bound = 0.0f;
    float dest[4] = {0.f, 0.f, 0.f, 0.f};
    float base[3] = {0.f, 0.f, 0.f};
    if (alive)
        for (int q = 0; q < 3; ++q) base[q] = r.pos[q];
    float max_am = 0.0f, max_q = 0.0f, max_qr = 0.0f, gamma = 1.0f;
    unsigned long long mbits = 0;
    long long bk = 0;
    while (alive) {
        float pos[3] = {base[0], base[1], base[2]};
        for (int t = 0; t < sub; ++t)
            for (int q = 0; q < 3; ++q) pos[q] = fmaf(r.dir[q], a.step, pos[q]);
        const long long k = bk + sub;
        const bool ok = k < a.max_samples && (sub == 0 || !outside(pos, a.view));
        float src[4] = {0.f, 0.f, 0.f, 0.f}, rsum = 0.f;
        float am = 0.f, qv = 0.f;
        if (ok) {
            Sample s;
            eval_sample<FT, LC>(a, tf, op, r, pos, s);
            src[0] = s.rgb[0] * s.a;
            src[1] = s.rgb[1] * s.a;
            src[2] = s.rgb[2] * s.a;
            src[3] = s.a;
            rsum = fabsf(s.rgb[0]) + fabsf(s.rgb[1]) + fabsf(s.rgb[2]);
            const float mmax = fmaxf(fabsf(s.m[0]), fmaxf(fabsf(s.m[1]), fabsf(s.m[2])));
            qv = s.aclamp ? 0.0f : dalpha_draw(a, s.araw) * s.ao * fmaxf(s.sg, s.tfa);
            am = s.a * mmax;
        }
        float nxt[3];
        for (int q = 0; q < 3; ++q) nxt[q] = fmaf(r.dir[q], a.step, pos[q]);
        const bool nxt_out = !ok || outside(nxt, a.view);
        float sv[GT][4], rs[GT];
        group_gather<GT>(src, sv, gl0, std::make_integer_sequence<int, GT>{});
        {
            // rsum of every sample of the round (only channel 0 is used)
            float one[4] = {rsum, 0.f, 0.f, 0.f}, got[GT][4];
            group_gather<GT>(one, got, gl0, std::make_integer_sequence<int, GT>{});
            for (int g = 0; g < GT; ++g) rs[g] = got[g][0];
        }
        const unsigned okg = (unsigned)(__ballot((int)ok) >> gl0) & ((1u << GT) - 1u);
        // compositing in sample order; this lane keeps D before its own sample
        float mine_d[4] = {0.f, 0.f, 0.f, 0.f};
        int ncomp = 0;
#pragma unroll
        for (int g = 0; g < GT; ++g) {
            if (!((okg >> g) & 1u)) {
                alive = false;
                break;
            }
            if (g == sub)
                for (int q = 0; q < 4; ++q) mine_d[q] = dest[q];
            const float t = 1.0f - dest[3];
            int mask = 0;
            for (int q = 0; q < 4; ++q) {
                const float u = fmaf(t, sv[g][q], dest[q]);
                if (u < 0.0f || u > 1.0f) mask |= 1 << q;
                dest[q] = clamp01(u);
            }
            const int kk = (int)((bk + g) % kSeg);
            mbits |= (unsigned long long)mask << (4 * (kk % kSub));
            gamma = fmaf(sv[g][3], rs[g], gamma);
            ++ncomp;
            if (dest[3] > 0.95f) {
                alive = false;
                break;
            }
        }
        // this lane's sample, when composited: its column entries
        if (sub < ncomp) {
            const int kk = (int)(k % kSeg);
            const long long c = k / kSeg;
            const AdjColumn col = adj_column(A, nrays, A.full_col ? c : 0);
            if (kk == 0 && !A.full_col) {
                ck[(2 * c) * nrays + pix] = make_float4(pos[0], pos[1], pos[2], 0.f);
                ck[(2 * c + 1) * nrays + pix] = make_float4(mine_d[0], mine_d[1], mine_d[2], mine_d[3]);
            }
            if (kk % kSub == 0)
                col.pos[(long long)(kk / kSub) * nrays + pix] = make_float4(pos[0], pos[1], pos[2], 0.f);
            col.da[(long long)kk * nrays + pix] = mine_d[3];
            max_am = fmaxf(max_am, am);
            max_q = fmaxf(max_q, qv);
            max_qr = fmaxf(max_qr, qv * rsum);
        }
        n += ncomp;
        // the part's masks once its last sample is composited (the loop's
        // tail writes an unfinished last part)
        const int kl = (int)((bk + GT - 1) % kSeg);
        if (ncomp == GT && kl % kSub == kSub - 1) {
            if (sub == 0) {
                const AdjColumn col = adj_column(A, nrays, A.full_col ? (bk + GT - 1) / kSeg : 0);
                col.mask[(long long)(kl / kSub) * nrays + pix] = mbits;
            }
            mbits = 0;
        }
        if (alive) {
            for (int q = 0; q < 3; ++q) base[q] = group_bcast<GT, GT - 1>(nxt[q]);
            if ((__ballot((int)nxt_out) >> (gl0 + GT - 1)) & 1ull) alive = false;
        }
        bk += GT;
    }
    if (active) {
        // the masks of a last part the loop left unfinished
        const int kl = (int)((n - 1) % kSeg);
        if (kl % kSub != kSub - 1 && sub == 0) {
            const AdjColumn col = adj_column(A, nrays, A.full_col ? (n - 1) / kSeg : 0);
            col.mask[(long long)(kl / kSub) * nrays + pix] = mbits;
        }
        // the bound terms are maxima over the ray's samples: reduce over the group
        for (int o = 1; o < GT; o <<= 1) {
            max_am = fmaxf(max_am, __shfl_xor(max_am, o, 64));
            max_q = fmaxf(max_q, __shfl_xor(max_q, o, 64));
            max_qr = fmaxf(max_qr, __shfl_xor(max_qr, o, 64));
        }
        bound = fmaxf(max_am, fmaf(gamma, max_q, max_qr));
        if (!isfinite(bound)) bound = __uint_as_float(0x7fc00000u);
    }
    if (inside && sub == 0) {
        A.nsamp[pix] = (int)n;
        A.ray_bound[pix] = bound;
        if (a.image) reinterpret_cast<float4*>(a.image)[pix] = make_float4(dest[0], dest[1], dest[2], dest[3]);
    }
    unsigned long long ns = sub == 0 ? (unsigned long long)n : 0ull;
    for (int o = 32; o > 0; o >>= 1) ns += __shfl_xor(ns, o, 64);
    if ((tid & 63) == 0) red[tid >> 6] = ns;
    __syncthreads();
    if (tid == 0) {
        unsigned long long t = 0;
        for (int q = 0; q < kRThreads / 64; ++q) t += red[q];
        if (t) {
            atomicAdd(&A.hdr[0], t);
            if (a.counter) atomicAdd(a.counter, t);
        }
    }
}

// The largest contribution bound of the call: each ray's bound per unit of
// max |dL/dimage| times its pixel's max |dL/dimage| (NaN when that pixel's
// gradient is not finite on a ray that has samples), into header word 1.
// The bound is >= 0 or NaN, whose bits exceed every finite one, so its bits
// order as unsigned integers.
__global__ void __launch_bounds__(256) k_adjoint_bound(AdjointLaunch A)
{
    __shared__ unsigned bmax;
    if (threadIdx.x == 0) bmax = 0u;
    __syncthreads();
    const long long nrays = (long long)A.r.view.width * A.r.view.height;
    const long long pix = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float bound = 0.0f;
    if (pix < nrays) {
        const float rb = A.ray_bound[pix];
        if (rb != 0.0f) {
            const float4 g = reinterpret_cast<const float4*>(A.dimage)[pix];
            const float gm = fmaxf(fmaxf(fabsf(g.x), fabsf(g.y)), fmaxf(fabsf(g.z), fabsf(g.w)));
            const bool finite_g = isfinite(g.x) && isfinite(g.y) && isfinite(g.z) && isfinite(g.w);
            bound = gm * rb;
            if (!finite_g || !isfinite(bound)) bound = __uint_as_float(0x7fc00000u);
        }
    }
    if (bound != 0.0f || isnan(bound)) atomicMax(&bmax, __float_as_uint(bound));
    __syncthreads();
    if (threadIdx.x == 0 && bmax) atomicMax(&A.hdr[1], (unsigned long long)bmax);
}


// Phase 2: the backward walk, segments last to first, all rays of the block
// in lockstep.
template <typename FT, int LC>
__global__ void __launch_bounds__(kRThreads) __attribute__((amdgpu_waves_per_eu(4))) k_render_adjoint(AdjointLaunch A)
{
    __shared__ float tf[1024];
    __shared__ float op[256];
    __shared__ long long dtf[1024];
    __shared__ long long win[kWin];
    __shared__ int box[kNSub][6];
    __shared__ int seg_max;
    int kexp;
    if (!adj_scale_exp(A.hdr, kexp)) return;  // zero or non-finite bound: nothing to add
    const RenderLaunch& a = A.r;
    const int tid = threadIdx.y * kBX + threadIdx.x;
    if (tid == 0) seg_max = 0;
    for (int q = tid; q < 1024; q += kRThreads) dtf[q] = 0;
    load_tables(a.tf, tf, op);

    const int i = blockIdx.x * kBX + threadIdx.x;
    const int j = blockIdx.y * kBY + threadIdx.y;
    const long long W = a.view.width, H = a.view.height;
    const long long nrays = W * H;
    const long long pix = (long long)j * W + i;
    const int* n3 = a.lic.n;  // the opacity scale lives on the LIC grid
    const float4* ck = reinterpret_cast<const float4*>(A.ckpt);  // compact: [c][2][ray] checkpoint (pos, D)

    Ray r;
    const bool active = i < W && j < H && ray_setup(a.view, i, j, r);
    const long long n = active ? A.nsamp[pix] : 0;
    const int nseg = (int)((n + kSeg - 1) / kSeg);
    if (nseg) atomicMax(&seg_max, nseg);
    __syncthreads();

    float G[4] = {0.f, 0.f, 0.f, 0.f};
    if (active) {
        const float4 gi = reinterpret_cast<const float4*>(A.dimage)[pix];
        G[0] = gi.x;
        G[1] = gi.y;
        G[2] = gi.z;
        G[3] = gi.w;
    }
    const int smax = seg_max;
    for (int c = smax - 1; c >= 0; --c) {
        const bool mine = active && c < nseg;
        const AdjColumn col = adj_column(A, nrays, c);
        const long long k0 = (long long)c * kSeg;
        const long long k1 = min(n, k0 + kSeg);
        if (tid < kNSub) {
            box[tid][0] = box[tid][1] = box[tid][2] = 0x7fffffff;
            box[tid][3] = box[tid][4] = box[tid][5] = -1;
        }
        __syncthreads();
        // (a) compact layout: re-run the segment from its checkpoint into the
        //     column (phase 1 left each ray's last segment there)
        if (mine && c < nseg - 1 && !A.full_col) {
            const float4 cp = ck[(2 * c) * nrays + pix];
            const float4 cd = ck[(2 * c + 1) * nrays + pix];
            float p[3] = {cp.x, cp.y, cp.z};
            float d[4] = {cd.x, cd.y, cd.z, cd.w};
            unsigned long long mbits = 0;
            for (long long k = k0; k < k1; ++k) {
                const int kk = (int)(k - k0);
                if (kk % kSub == 0) col.pos[(long long)(kk / kSub) * nrays + pix] = make_float4(p[0], p[1], p[2], 0.f);
                Sample s;
                eval_sample<FT, LC>(a, tf, op, r, p, s);
                col.da[(long long)kk * nrays + pix] = d[3];
                mbits |= (unsigned long long)composite(d, s) << (4 * (kk % kSub));
                if (kk % kSub == kSub - 1 || k == k1 - 1) {
                    col.mask[(long long)(kk / kSub) * nrays + pix] = mbits;
                    mbits = 0;
                }
                for (int q = 0; q < 3; ++q) p[q] = fmaf(r.dir[q], a.step, p[q]);
            }
        }
        // (b) the voxel box of each part's sigma splats: along a ray every
        //     axis coordinate is monotone (the same fma step each sample), so
        //     the part's first and last samples bound its cells
        if (mine && A.dsigma_fx) {
            for (int m = 0; m < kNSub; ++m) {
                const long long ka = k0 + (long long)m * kSub;
                const long long kb = min(k1, ka + kSub);
                if (ka >= kb) break;
                const float4 ps = col.pos[(long long)m * nrays + pix];
                float pe[3] = {ps.x, ps.y, ps.z};
                for (long long t = ka + 1; t < kb; ++t)
                    for (int q = 0; q < 3; ++q) pe[q] = fmaf(r.dir[q], a.step, pe[q]);
                const float pf[3] = {ps.x, ps.y, ps.z};
                for (int q = 0; q < 3; ++q) {
                    float f;
                    const int c0 = tc_clamp(pf[q], n3[q], f) - 1;
                    const int c1 = tc_clamp(pe[q], n3[q], f) - 1;
                    atomicMin(&box[m][q], max(min(c0, c1), 0));
                    atomicMax(&box[m][3 + q], min(max(c0, c1) + 1, n3[q] - 1));
                }
            }
        }
        __syncthreads();
        for (int m = kNSub - 1; m >= 0; --m) {
            const int bx0 = box[m][0], by0 = box[m][1], bz0 = box[m][2];
            const int wx = box[m][3] - bx0 + 1, wy = box[m][4] - by0 + 1, wz = box[m][5] - bz0 + 1;
            const long long wvol = (box[m][3] >= 0) ? (long long)wx * wy * wz : 0;
            const bool use_win = A.use_window && A.dsigma_fx && wvol > 0 && wvol <= kWin;
            if (use_win)
                for (int q = tid; q < (int)wvol; q += kRThreads) win[q] = 0;
            __syncthreads();
            // (b) walk this part of the segment backwards
            const long long ka = k0 + (long long)m * kSub;
            const long long kb = min(k1, ka + kSub);
            if (mine) {
                const float4 ps = col.pos[(long long)m * nrays + pix];
                const unsigned long long mbits = ka < kb ? col.mask[(long long)m * nrays + pix] : 0ull;
                for (long long k = kb - 1; k >= ka; --k) {
                    const int mask = (int)((mbits >> (4 * (int)(k - ka))) & 15ull);
                    const float dka = col.da[(k - k0) * nrays + pix];
                    // the sample position: the forward fma chain from the part's start
                    float pk[3] = {ps.x, ps.y, ps.z};
                    for (long long t = ka; t < k; ++t)
                        for (int q = 0; q < 3; ++q) pk[q] = fmaf(r.dir[q], a.step, pk[q]);
                    Sample s;
                    eval_sample<FT, LC>(a, tf, op, r, pk, s);
                    for (int q = 0; q < 4; ++q)
                        if (mask & (1 << q)) G[q] = 0.0f;
                    const float t = 1.0f - dka;
                    const float al = s.a;
                    const float src[4] = {s.rgb[0] * al, s.rgb[1] * al, s.rgb[2] * al, al};
                    float dsrc[4];
                    for (int q = 0; q < 4; ++q) dsrc[q] = t * G[q];
                    const float gs = G[0] * src[0] + G[1] * src[1] + G[2] * src[2] + G[3] * src[3];
                    G[3] = G[3] - gs;
                    float da = dsrc[3];
                    float drgb[3];
                    for (int q = 0; q < 3; ++q) {
                        drgb[q] = dsrc[q] * al;
                        da = da + dsrc[q] * s.rgb[q];
                    }
                    const float draw = s.aclamp ? 0.0f : da * dalpha_draw(a, s.araw);
                    for (int q = 0; q < 3; ++q) splat_tf(dtf, q, s.tfu, drgb[q] * s.m[q], kexp);
                    splat_tf(dtf, 3, s.tfu, draw * s.ao * s.sg, kexp);
                    if (A.dsigma_fx) {
                        const float dsg = draw * s.ao * s.tfa;
                        float fx, fy, fz;
                        const int px = tc_clamp(pk[0], n3[0], fx) - 1;
                        const int py = tc_clamp(pk[1], n3[1], fy) - 1;
                        const int pz = tc_clamp(pk[2], n3[2], fz) - 1;
                        const int xs[2] = {max(px, 0), min(px + 1, n3[0] - 1)};
                        const int ys[2] = {max(py, 0), min(py + 1, n3[1] - 1)};
                        const int zs[2] = {max(pz, 0), min(pz + 1, n3[2] - 1)};
                        const float wxs[2] = {1.0f - fx, fx}, wys[2] = {1.0f - fy, fy}, wzs[2] = {1.0f - fz, fz};
                        for (int cz = 0; cz < 2; ++cz)
                            for (int cy = 0; cy < 2; ++cy)
                                for (int cx = 0; cx < 2; ++cx) {
                                    const long long v = to_fx(wxs[cx] * wys[cy] * wzs[cz] * dsg, kexp);
                                    if (!v) continue;
                                    if (use_win) {
                                        const int o = ((zs[cz] - bz0) * wy + (ys[cy] - by0)) * wx + (xs[cx] - bx0);
                                        atomicAdd((unsigned long long*)&win[o], (unsigned long long)v);
                                    } else {
                                        const long long o = ((long long)zs[cz] * n3[1] + ys[cy]) * n3[0] + xs[cx];
                                        atomicAdd((unsigned long long*)&A.dsigma_fx[o], (unsigned long long)v);
                                    }
                                }
                    }
                }
            }
            __syncthreads();
            // (c) flush the window: one global integer atomic per touched voxel
            if (use_win) {
                for (int q = tid; q < (int)wvol; q += kRThreads) {
                    const long long v = win[q];
                    if (v) {
                        const int ox = q % wx, oy = (q / wx) % wy, oz = q / (wx * wy);
                        const long long o = ((long long)(bz0 + oz) * n3[1] + (by0 + oy)) * n3[0] + (bx0 + ox);
                        atomicAdd((unsigned long long*)&A.dsigma_fx[o], (unsigned long long)v);
                    }
                }
            }
            __syncthreads();
        }
    }
    __syncthreads();
    for (int q = tid; q < 1024; q += kRThreads) {
        const long long v = dtf[q];
        if (v) atomicAdd((unsigned long long*)&A.dtf_fx[q], (unsigned long long)v);
    }
}

// Phase 2 with GT lanes per ray (2 or 4), with k_adjoint_forward_rg's pixel
// tiles; COMPACT (the compact layout): the group re-runs each earlier segment
// from its checkpoint first (a variant of its own, so the full layout's
// kernel keeps its registers).  Within a part the lanes take
// consecutive samples from the end (lane g: sample top - g) and evaluate
// them at once; the group then runs the gradient recurrence of
// k_render_adjoint in its sample order (every lane of the group alike, on
// the samples' premultiplied colour and clamp masks), each lane keeping the
// state at its own sample, and each lane splats its own sample.  Integer
// sums do not depend on the order of the splats, so dTF and dsigma equal
// k_render_adjoint's bits.
template <typename FT, int LC, int GT, bool COMPACT>
__global__ void __launch_bounds__(kRThreads) __attribute__((amdgpu_waves_per_eu(4)))
k_render_adjoint_rg(AdjointLaunch A)
{
    static_assert(GT == 2 || GT == 4, "quad groups only");
    __shared__ float tf[1024];
    __shared__ float op[256];
    __shared__ long long dtf[1024];
    __shared__ long long win[kWin];
    __shared__ int box[kNSub][6];
    __shared__ int seg_max;
    int kexp;
    if (!adj_scale_exp(A.hdr, kexp)) return;  // zero or non-finite bound: nothing to add
    const RenderLaunch& a = A.r;
    const int tid = threadIdx.x;
    if (tid == 0) seg_max = 0;
    for (int q = tid; q < 1024; q += kRThreads) dtf[q] = 0;
    load_tables(a.tf, tf, op);

    const int sub = tid & (GT - 1);
    const int gl0 = (tid & 63) & ~(GT - 1);
    const int W = a.view.width, H = a.view.height;
    int i, j;
    rg_pixel(GT, W, (int)blockIdx.x, tid, i, j);
    const long long nrays = (long long)W * H;
    const long long pix = (long long)j * W + i;
    const int* n3 = a.lic.n;  // the opacity scale lives on the LIC grid
    const float4* ck = reinterpret_cast<const float4*>(A.ckpt);  // compact: [c][2][ray] checkpoint (pos, D)

    Ray r;
    const bool active = i < W && j < H && ray_setup(a.view, i, j, r);
    const long long n = active ? A.nsamp[pix] : 0;
    const int nseg = (int)((n + kSeg - 1) / kSeg);
    if (nseg && sub == 0) atomicMax(&seg_max, nseg);
    __syncthreads();

    float G[4] = {0.f, 0.f, 0.f, 0.f};
    if (active) {
        const float4 gi = reinterpret_cast<const float4*>(A.dimage)[pix];
        G[0] = gi.x;
        G[1] = gi.y;
        G[2] = gi.z;
        G[3] = gi.w;
    }
    const int smax = seg_max;
    for (int c = smax - 1; c >= 0; --c) {
        const bool mine = active && c < nseg;
        const AdjColumn col = adj_column(A, nrays, c);
        const long long k0 = (long long)c * kSeg;
        const long long k1 = min(n, k0 + kSeg);
        if (tid < kNSub) {
            box[tid][0] = box[tid][1] = box[tid][2] = 0x7fffffff;
            box[tid][3] = box[tid][4] = box[tid][5] = -1;
        }
        // compact layout: re-run the segment from its checkpoint into the
        // column (the forward pass left each ray's last segment there), the
        // group's lanes taking consecutive samples as in k_adjoint_forward_rg;
        // slot kk is written by lane kk mod GT, the lane that wrote it in
        // the forward pass, and read after the block barrier below
        if (COMPACT && mine && c < nseg - 1) {
            const float4 cp = ck[(2 * c) * nrays + pix];
            const float4 cd = ck[(2 * c + 1) * nrays + pix];
            float base[3] = {cp.x, cp.y, cp.z};
            float d[4] = {cd.x, cd.y, cd.z, cd.w};
            unsigned long long mbits = 0;
            for (long long bk = k0; bk < k1; bk += GT) {
                const long long k = bk + sub;
                const bool ok = k < k1;
                float pos[3] = {base[0], base[1], base[2]};
                for (int t = 0; t < sub; ++t)
                    for (int q = 0; q < 3; ++q) pos[q] = fmaf(r.dir[q], a.step, pos[q]);
                float src[4] = {0.f, 0.f, 0.f, 0.f};
                if (ok) {
                    Sample s;
                    eval_sample<FT, LC>(a, tf, op, r, pos, s);
                    src[0] = s.rgb[0] * s.a;
                    src[1] = s.rgb[1] * s.a;
                    src[2] = s.rgb[2] * s.a;
                    src[3] = s.a;
                }
                float sv[GT][4];
                group_gather<GT>(src, sv, gl0, std::make_integer_sequence<int, GT>{});
                float mine_da = 0.0f;
#pragma unroll
                for (int g = 0; g < GT; ++g) {
                    if (bk + g >= k1) break;
                    if (g == sub) mine_da = d[3];
                    const float t = 1.0f - d[3];
                    int mask = 0;
                    for (int q = 0; q < 4; ++q) {
                        const float u = fmaf(t, sv[g][q], d[q]);
                        if (u < 0.0f || u > 1.0f) mask |= 1 << q;
                        d[q] = clamp01(u);
                    }
                    mbits |= (unsigned long long)mask << (4 * (int)((bk + g - k0) % kSub));
                }
                if (ok) {
                    const int kk = (int)(k - k0);
                    if (kk % kSub == 0)
                        col.pos[(long long)(kk / kSub) * nrays + pix] = make_float4(pos[0], pos[1], pos[2], 0.f);
                    col.da[(long long)kk * nrays + pix] = mine_da;
                }
                // a part's masks once its last sample (or the segment's) is in
                const long long last = min(bk + GT, k1) - 1;
                const int kl = (int)(last - k0);
                if (kl % kSub == kSub - 1 || last == k1 - 1) {
                    if (sub == 0) col.mask[(long long)(kl / kSub) * nrays + pix] = mbits;
                    mbits = 0;
                }
                float nxt[3];
                for (int q = 0; q < 3; ++q) nxt[q] = fmaf(r.dir[q], a.step, pos[q]);
                for (int q = 0; q < 3; ++q) base[q] = group_bcast<GT, GT - 1>(nxt[q]);
            }
        }
        __syncthreads();
        // the voxel box of each part's sigma splats (lanes take parts in turn)
        if (mine && A.dsigma_fx) {
            for (int m = sub; m < kNSub; m += GT) {
                const long long ka = k0 + (long long)m * kSub;
                const long long kb = min(k1, ka + kSub);
                if (ka >= kb) continue;
                const float4 ps = col.pos[(long long)m * nrays + pix];
                float pe[3] = {ps.x, ps.y, ps.z};
                for (long long t = ka + 1; t < kb; ++t)
                    for (int q = 0; q < 3; ++q) pe[q] = fmaf(r.dir[q], a.step, pe[q]);
                const float pf[3] = {ps.x, ps.y, ps.z};
                for (int q = 0; q < 3; ++q) {
                    float f;
                    const int c0 = tc_clamp(pf[q], n3[q], f) - 1;
                    const int c1 = tc_clamp(pe[q], n3[q], f) - 1;
                    atomicMin(&box[m][q], max(min(c0, c1), 0));
                    atomicMax(&box[m][3 + q], min(max(c0, c1) + 1, n3[q] - 1));
                }
            }
        }
        __syncthreads();
        for (int m = kNSub - 1; m >= 0; --m) {
            const int bx0 = box[m][0], by0 = box[m][1], bz0 = box[m][2];
            const int wx = box[m][3] - bx0 + 1, wy = box[m][4] - by0 + 1, wz = box[m][5] - bz0 + 1;
            const long long wvol = (box[m][3] >= 0) ? (long long)wx * wy * wz : 0;
            const bool use_win = A.use_window && A.dsigma_fx && wvol > 0 && wvol <= kWin;
            if (use_win)
                for (int q = tid; q < (int)wvol; q += kRThreads) win[q] = 0;
            __syncthreads();
            const long long ka = k0 + (long long)m * kSub;
            const long long kb = min(k1, ka + kSub);
            if (mine && ka < kb) {
                const float4 ps = col.pos[(long long)m * nrays + pix];
                const unsigned long long mbits = col.mask[(long long)m * nrays + pix];
                for (long long top = kb - 1; top >= ka; top -= GT) {
                    const long long k = top - sub;
                    const bool valid = k >= ka;
                    int mask = 0;
                    float dka = 0.0f;
                    float pk[3] = {ps.x, ps.y, ps.z};
                    float src[4] = {0.f, 0.f, 0.f, 0.f};
                    Sample s;
                    if (valid) {
                        mask = (int)((mbits >> (4 * (int)(k - ka))) & 15ull);
                        dka = col.da[(k - k0) * nrays + pix];
                        // the sample position: the forward fma chain from the part's start
                        for (long long t = ka; t < k; ++t)
                            for (int q = 0; q < 3; ++q) pk[q] = fmaf(r.dir[q], a.step, pk[q]);
                        eval_sample<FT, LC>(a, tf, op, r, pk, s);
                        const float al = s.a;
                        src[0] = s.rgb[0] * al;
                        src[1] = s.rgb[1] * al;
                        src[2] = s.rgb[2] * al;
                        src[3] = al;
                    }
                    // the recurrence over the round's samples in k_render_adjoint's order
                    float sv[GT][4], mv[GT][4];
                    group_gather<GT>(src, sv, gl0, std::make_integer_sequence<int, GT>{});
                    {
                        // the 4-bit clamp masks travel as raw bits through the
                        // float broadcast (a register move, no arithmetic)
                        const float mk[4] = {__int_as_float(mask), 0.f, 0.f, 0.f};
                        group_gather<GT>(mk, mv, gl0, std::make_integer_sequence<int, GT>{});
                    }
                    float Gk[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int g = 0; g < GT; ++g) {
                        if (top - g < ka) break;
                        const int mg = __float_as_int(mv[g][0]);
                        for (int q = 0; q < 4; ++q)
                            if (mg & (1 << q)) G[q] = 0.0f;
                        if (g == sub)
                            for (int q = 0; q < 4; ++q) Gk[q] = G[q];
                        const float gs = G[0] * sv[g][0] + G[1] * sv[g][1] + G[2] * sv[g][2] + G[3] * sv[g][3];
                        G[3] = G[3] - gs;
                    }
                    if (valid) {
                        const float t = 1.0f - dka;
                        const float al = s.a;
                        float dsrc[4];
                        for (int q = 0; q < 4; ++q) dsrc[q] = t * Gk[q];
                        float da = dsrc[3];
                        float drgb[3];
                        for (int q = 0; q < 3; ++q) {
                            drgb[q] = dsrc[q] * al;
                            da = da + dsrc[q] * s.rgb[q];
                        }
                        const float draw = s.aclamp ? 0.0f : da * dalpha_draw(a, s.araw);
                        for (int q = 0; q < 3; ++q) splat_tf(dtf, q, s.tfu, drgb[q] * s.m[q], kexp);
                        splat_tf(dtf, 3, s.tfu, draw * s.ao * s.sg, kexp);
                        if (A.dsigma_fx) {
                            const float dsg = draw * s.ao * s.tfa;
                            float fx, fy, fz;
                            const int px = tc_clamp(pk[0], n3[0], fx) - 1;
                            const int py = tc_clamp(pk[1], n3[1], fy) - 1;
                            const int pz = tc_clamp(pk[2], n3[2], fz) - 1;
                            const int xs[2] = {max(px, 0), min(px + 1, n3[0] - 1)};
                            const int ys[2] = {max(py, 0), min(py + 1, n3[1] - 1)};
                            const int zs[2] = {max(pz, 0), min(pz + 1, n3[2] - 1)};
                            const float wxs[2] = {1.0f - fx, fx}, wys[2] = {1.0f - fy, fy}, wzs[2] = {1.0f - fz, fz};
                            for (int cz = 0; cz < 2; ++cz)
                                for (int cy = 0; cy < 2; ++cy)
                                    for (int cx = 0; cx < 2; ++cx) {
                                        const long long v = to_fx(wxs[cx] * wys[cy] * wzs[cz] * dsg, kexp);
                                        if (!v) continue;
                                        if (use_win) {
                                            const int o = ((zs[cz] - bz0) * wy + (ys[cy] - by0)) * wx + (xs[cx] - bx0);
                                            atomicAdd((unsigned long long*)&win[o], (unsigned long long)v);
                                        } else {
                                            const long long o = ((long long)zs[cz] * n3[1] + ys[cy]) * n3[0] + xs[cx];
                                            atomicAdd((unsigned long long*)&A.dsigma_fx[o], (unsigned long long)v);
                                        }
                                    }
                        }
                    }
                }
            }
            __syncthreads();
            // flush the window: one global integer atomic per touched voxel
            if (use_win) {
                for (int q = tid; q < (int)wvol; q += kRThreads) {
                    const long long v = win[q];
                    if (v) {
                        const int ox = q % wx, oy = (q / wx) % wy, oz = q / (wx * wy);
                        const long long o = ((long long)(bz0 + oz) * n3[1] + (by0 + oy)) * n3[0] + (bx0 + ox);
                        atomicAdd((unsigned long long*)&A.dsigma_fx[o], (unsigned long long)v);
                    }
                }
            }
            __syncthreads();
        }
    }
    __syncthreads();
    for (int q = tid; q < 1024; q += kRThreads) {
        const long long v = dtf[q];
        if (v) atomicAdd((unsigned long long*)&A.dtf_fx[q], (unsigned long long)v);
    }
}

// fixed point back to float: one rounding (int64 -> float), then the exact
// power-of-two scale
__global__ void k_fixed_to_float(const unsigned long long* hdr, const long long* src, float* dst, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k;
    if (adj_scale_exp(hdr, k))
        dst[i] = ldexpf(__ll2float_rn(src[i]), -k);
    else
        dst[i] = (unsigned)hdr[1] >= 0x7f800000u ? __uint_as_float(0x7fc00000u) : 0.0f;
}

// ---- fused on-the-fly LIC ray caster --------------------------------------
//
// lic3d_fragment.glsl:5-99 (the reference's default VOLIC_RAYCAST mode): no
// LIC volume; every ray sample computes the LIC integral of its own position
// (computeLIC, scaled by licKernel.b * gradient.r), shades it and composites
// it.  A ray stops when it leaves the box or when the sample's own alpha
// exceeds 0.95 (src.a, :91; the LIC-volume caster tests the accumulated
// dest.a).  MODE 0: any noise / gate; 1: per-lane cell caches (scalar noise,
// no gate), renewed for every sample; 2: the same with the noise grid equal
// to the field grid at freq 1.
template <typename FT, int NC, bool GATE, int MODE>
__global__ void __launch_bounds__(kRThreads) __attribute__((amdgpu_waves_per_eu(3)))
k_render_fused(RenderLaunch a, LicArgs L)
{
    __shared__ float tf[1024];
    __shared__ float op[256];
    __shared__ unsigned long long red[kRThreads / 64];
    load_tables(a.tf, tf, op);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    unsigned long long ns = 0;
    if (i < a.view.width && j < a.view.height) {
        float dest[4] = {0.f, 0.f, 0.f, 0.f};
        Ray r;
        if (ray_setup(a.view, i, j, r)) {
            const Vol<FT> fv{(const FT*)a.field.p, {a.field.n[0], a.field.n[1], a.field.n[2]}, a.field.sy,
                             a.field.sz};
            const CachedArgs c = cached_args(L);
            float pos[3] = {r.pos[0], r.pos[1], r.pos[2]};
            for (long long k = 0; k < a.max_samples; ++k) {
                float l4[4];
                if constexpr (MODE > 0) {
                    // cell caches per sample (their registers are free while shading)
                    LicFieldCache<FT, MODE == 2> fc;
                    fc.reset();
                    NoiseCache nc;
                    nc.reset();
                    const float acc = lic_point_cached<FT, MODE == 2>(L, c, fc, nc, pos[0], pos[1], pos[2]);
                    l4[0] = acc * L.scl;
                } else {
                    const float4 acc = lic_point<FT, NC, GATE>(L, pos[0], pos[1], pos[2]);
                    l4[0] = NC == 1 ? acc.x * L.scl : acc.x;
                    l4[1] = acc.y;
                    l4[2] = acc.z;
                    l4[3] = acc.w * L.scl;
                }
                const Cell cf = cell_clamp(fv, pos[0], pos[1], pos[2]);
                Sample s;
                shade_sample<FT, NC>(a, tf, op, r, pos, fv, cf, l4, s);
                composite(dest, s);
                ++ns;
                for (int q = 0; q < 3; ++q) pos[q] = fmaf(r.dir[q], a.step, pos[q]);
                if (outside(pos, a.view) || s.a > 0.95f) break;
            }
        }
        const long long pix = (long long)j * a.view.width + i;
        reinterpret_cast<float4*>(a.image)[pix] = make_float4(dest[0], dest[1], dest[2], dest[3]);
    }
    if (a.counter) add_samples(a.counter, ns, red);
}

// ---- view-aligned slicing -----------------------------------------------------
//
// Renderer::sliceVolume with lic3d_slicing_fragment.glsl (the render-target
// path): the view-aligned slices of ViewSlicing (VV/slicing.cpp:42-281) are
// drawn front to back; a pixel's fragment on slice i sits where the pixel's
// ray meets the slice plane inside the box.  A fragment is skipped when the
// accumulated alpha has reached 0.95 or when the sample's TF alpha is at most
// 0.05 (no LIC is computed there); otherwise it runs the LIC of its position,
// shades and composites it, and the result is stored in the RGBA16F render
// target, i.e. rounded to FP16, before the next slice reads it.
__device__ __forceinline__ float half_rn(float x) { return __half2float(__float2half_rn(x)); }

// The reference's default slicing output (Renderer::sliceVolume without
// render targets, VV/renderer.cpp:1149-1157, with
// lic3d_slicingblend_fragment.glsl): each slice's fragment src (rgb
// premultiplied, 0 where the TF alpha is at most 0.05) is blended into the
// 8-bit RGBA window framebuffer with glBlendFunc(GL_ONE_MINUS_DST_ALPHA,
// GL_ONE), dst = src (1 - dst.a) + dst: front to back "under", with no 0.95
// test.  The framebuffer is fixed point: the fragment colour is clamped to
// [0,1], the blend result clamped and stored as round(255 x).  After the
// slices a white screen quad is blended the same way (the background).
// dst holds byte values; blend_u8 blends one slice's src into them.
__device__ __forceinline__ float unorm8_rn(float x) { return floorf(fmaf(clamp01(x), 255.0f, 0.5f)); }

__device__ __forceinline__ void blend_u8(float dst[4], const float src[4])
{
    const float t = 1.0f - dst[3] * (1.0f / 255.0f);
    for (int c = 0; c < 4; ++c) dst[c] = unorm8_rn(fmaf(t, clamp01(src[c]), dst[c] * (1.0f / 255.0f)));
}

// the white background quad, then the pixel's bytes
__device__ __forceinline__ void store_blend(float* image, long long pix, float dst[4])
{
    const float white[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    blend_u8(dst, white);
    reinterpret_cast<uchar4*>(image)[pix] =
        make_uchar4((unsigned char)dst[0], (unsigned char)dst[1], (unsigned char)dst[2], (unsigned char)dst[3]);
}

template <typename FT, int NC, bool GATE, int MODE>
__global__ void __launch_bounds__(kRThreads) __attribute__((amdgpu_waves_per_eu(3)))
k_render_slices(RenderLaunch a, LicArgs L, SliceSetup S)
{
    __shared__ float tf[1024];
    __shared__ float op[256];
    __shared__ unsigned long long red[kRThreads / 64];
    load_tables(a.tf, tf, op);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    unsigned long long ns = 0;
    if (i < a.view.width && j < a.view.height) {
        const ViewBlock& vb = a.view;
        const float nxp = fmaf((float)i + 0.5f, vb.inv_w2, -1.0f);
        const float nyp = fmaf((float)j + 0.5f, vb.inv_h2, -1.0f);
        float d[3], rc[3];
        for (int q = 0; q < 3; ++q) {
            d[q] = fmaf(nxp, vb.du[q], fmaf(nyp, vb.dv[q], vb.dw[q]));
            rc[q] = vb.cam[q] - S.center[q];
        }
        const float vdir = S.v[0] * d[0] + S.v[1] * d[1] + S.v[2] * d[2];
        const float vcam = S.v[0] * rc[0] + S.v[1] * rc[1] + S.v[2] * rc[2];
        const Vol<FT> fv{(const FT*)a.field.p, {a.field.n[0], a.field.n[1], a.field.n[2]}, a.field.sy,
                         a.field.sz};
        const CachedArgs c = cached_args(L);
        float dest[4] = {0.f, 0.f, 0.f, 0.f};
        for (int k = 0; k < S.n && vdir != 0.0f; ++k) {
            // the blend adds nothing once dst.a = 1; the render-target path
            // stops at 0.95
            if (S.blend ? !(dest[3] < 255.0f) : !(dest[3] < 0.95f)) break;
            const float dk = -0.5f * S.D + ((float)k + 0.5f) * S.D / (float)S.n;
            const float sk = (dk - vcam) / vdir;
            if (!(sk > 0.0f)) continue;
            float p[3];
            bool in = true;
            for (int q = 0; q < 3; ++q) {
                p[q] = fmaf(sk, d[q], vb.cam[q]);
                in = in && p[q] >= 0.0f && p[q] <= vb.extent[q];
            }
            if (!in) continue;
            Ray r;
            float g[3], gn[3], pos[3];
            for (int q = 0; q < 3; ++q) g[q] = p[q] - vb.cam[q];
            normalize3(g, gn);
            for (int q = 0; q < 3; ++q) {
                r.dir[q] = gn[q] * vb.scale[q];
                pos[q] = p[q] * vb.scale[q];
            }
            const float vneg[3] = {-r.dir[0], -r.dir[1], -r.dir[2]};
            normalize3(vneg, r.V);
            const Cell cf = cell_clamp(fv, pos[0], pos[1], pos[2]);
            const float tfu = tri_ch(fv, cf, a.tf_ch);
            if (!(lut1d(tf, 256, 4, 3, tfu) > 0.05f)) continue;
            float l4[4];
            if constexpr (MODE > 0) {
                // cell caches per sample (their registers are free while shading)
                LicFieldCache<FT, MODE == 2> fc;
                fc.reset();
                NoiseCache nc;
                nc.reset();
                const float acc = lic_point_cached<FT, MODE == 2>(L, c, fc, nc, pos[0], pos[1], pos[2]);
                l4[0] = acc * L.scl;
            } else {
                const float4 acc = lic_point<FT, NC, GATE>(L, pos[0], pos[1], pos[2]);
                l4[0] = NC == 1 ? acc.x * L.scl : acc.x;
                l4[1] = acc.y;
                l4[2] = acc.z;
                l4[3] = acc.w * L.scl;
            }
            Sample s;
            shade_sample<FT, NC>(a, tf, op, r, pos, fv, cf, l4, s);
            if (S.blend) {
                const float src[4] = {s.rgb[0] * s.a, s.rgb[1] * s.a, s.rgb[2] * s.a, s.a};
                blend_u8(dest, src);
            } else {
                composite(dest, s);
                for (int q = 0; q < 4; ++q) dest[q] = half_rn(dest[q]);
            }
            ++ns;
        }
        const long long pix = (long long)j * a.view.width + i;
        if (S.blend)
            store_blend(a.image, pix, dest);
        else
            reinterpret_cast<float4*>(a.image)[pix] = make_float4(dest[0], dest[1], dest[2], dest[3]);
    }
    if (a.counter) add_samples(a.counter, ns, red);
}

// ---- ray groups: several lanes per ray for the on-the-fly casters ----------------
//
// With one lane per pixel a small view has too few lanes to hide the serial
// chain of samples (each a full LIC integral).  Here G lanes (a power of two,
// at most 64) share a ray: in every round lane g computes sample base + g,
// then the group composites the G results in sample order, each lane keeping
// the same running colour, so the compositing arithmetic and order are those
// of the one-lane casters and the image is identical.  Work past the end of a
// ray (termination inside a round) is at most G - 1 samples.
//   SLICES == false  fused caster: sample k at the k-th step along the ray
//                    (each lane steps from the round's first position to its
//                    own, with the same fma sequence), stop after a sample
//                    whose alpha exceeds 0.95 or when the next step leaves
//                    the box
//   SLICES == true   view-aligned slicing: sample k on slice k; skipped when
//                    the ray misses the slice in the box or the TF alpha is
//                    at most 0.05; stop once dest.a >= 0.95; dest rounded to
//                    FP16 after every composited sample
template <typename FT, int NC, bool GATE, int MODE, bool SLICES>
__global__ void __launch_bounds__(kRThreads) __attribute__((amdgpu_waves_per_eu(3)))
k_render_group(RenderLaunch a, LicArgs L, SliceSetup S, int G)
{
    __shared__ float tf[1024];
    __shared__ float op[256];
    __shared__ unsigned long long red[kRThreads / 64];
    load_tables(a.tf, tf, op);
    const int tid = threadIdx.x;
    const int sub = tid & (G - 1);
    const int gl0 = (tid & 63) & ~(G - 1);  // first lane of the group in the wave
    // the block's 256/G rays form a TW x TH pixel tile (TW = TH or 2 TH), so
    // each wave's rays are neighbours and their streamlines share cache lines
    const int W = a.view.width, H = a.view.height;
    const int nr = kRThreads / G;
    const int lg = 31 - __builtin_clz(nr);
    const int TW = 1 << ((lg + 1) >> 1), TH = nr / TW;
    const int tiles_x = (W + TW - 1) / TW;
    const int r_in = tid / G;
    const int i = (int)(blockIdx.x % tiles_x) * TW + (r_in & (TW - 1));
    const int j = (int)(blockIdx.x / tiles_x) * TH + r_in / TW;
    const bool has_pix = i < W && j < H;
    const ViewBlock& vb = a.view;
    const Vol<FT> fv{(const FT*)a.field.p, {a.field.n[0], a.field.n[1], a.field.n[2]}, a.field.sy, a.field.sz};
    const CachedArgs c = cached_args(L);
    float dest[4] = {0.f, 0.f, 0.f, 0.f};
    unsigned long long ns = 0;
    Ray r;
    bool alive;
    // ray policy state
    float base[3] = {0.f, 0.f, 0.f};
    // slice policy state
    float d[3] = {0.f, 0.f, 0.f}, vdir = 0.f, vcam = 0.f;
    if constexpr (SLICES) {
        const float nxp = fmaf((float)i + 0.5f, vb.inv_w2, -1.0f);
        const float nyp = fmaf((float)j + 0.5f, vb.inv_h2, -1.0f);
        float rc[3];
        for (int q = 0; q < 3; ++q) {
            d[q] = fmaf(nxp, vb.du[q], fmaf(nyp, vb.dv[q], vb.dw[q]));
            rc[q] = vb.cam[q] - S.center[q];
        }
        vdir = S.v[0] * d[0] + S.v[1] * d[1] + S.v[2] * d[2];
        vcam = S.v[0] * rc[0] + S.v[1] * rc[1] + S.v[2] * rc[2];
        alive = has_pix && vdir != 0.0f;
    } else {
        alive = has_pix && ray_setup(vb, i, j, r);
        for (int q = 0; q < 3; ++q) base[q] = r.pos[q];
    }
    long long bk = 0;
    while (alive) {
        const long long k = bk + sub;
        bool ok;
        float src[4] = {0.f, 0.f, 0.f, 0.f};
        float nxt[3] = {0.f, 0.f, 0.f};
        bool nxt_out = true;
        float pos[3];
        Ray rs;
        if constexpr (SLICES) {
            ok = k < S.n;
            if (ok) {
                const float dk = -0.5f * S.D + ((float)k + 0.5f) * S.D / (float)S.n;
                const float sk = (dk - vcam) / vdir;
                ok = sk > 0.0f;
                float p[3];
                for (int q = 0; q < 3; ++q) {
                    p[q] = fmaf(sk, d[q], vb.cam[q]);
                    ok = ok && p[q] >= 0.0f && p[q] <= vb.extent[q];
                }
                if (ok) {
                    float g[3], gn[3];
                    for (int q = 0; q < 3; ++q) g[q] = p[q] - vb.cam[q];
                    normalize3(g, gn);
                    for (int q = 0; q < 3; ++q) {
                        rs.dir[q] = gn[q] * vb.scale[q];
                        pos[q] = p[q] * vb.scale[q];
                    }
                    const float vneg[3] = {-rs.dir[0], -rs.dir[1], -rs.dir[2]};
                    normalize3(vneg, rs.V);
                }
            }
        } else {
            ok = k < a.max_samples;
            for (int q = 0; q < 3; ++q) pos[q] = base[q];
            for (int t = 0; t < sub; ++t) {
                for (int q = 0; q < 3; ++q) pos[q] = fmaf(r.dir[q], a.step, pos[q]);
                if (outside(pos, vb)) ok = false;
            }
            for (int q = 0; q < 3; ++q) {
                nxt[q] = fmaf(r.dir[q], a.step, pos[q]);
                rs.dir[q] = r.dir[q];
                rs.V[q] = r.V[q];
            }
            nxt_out = !ok || outside(nxt, vb);
        }
        if (ok) {
            const Cell cf = cell_clamp(fv, pos[0], pos[1], pos[2]);
            bool visible = true;
            if constexpr (SLICES) visible = lut1d(tf, 256, 4, 3, tri_ch(fv, cf, a.tf_ch)) > 0.05f;
            if (visible) {
                float l4[4];
                if constexpr (MODE > 0) {
                    // cell caches per sample: consecutive samples of a lane are G steps
                    // apart, so the previous streamline's last cell rarely serves the
                    // next one, and the caches' registers are free while shading
                    LicFieldCache<FT, MODE == 2> fc;
                    fc.reset();
                    NoiseCache nc;
                    nc.reset();
                    const float acc = lic_point_cached<FT, MODE == 2>(L, c, fc, nc, pos[0], pos[1], pos[2]);
                    l4[0] = acc * L.scl;
                } else {
                    const float4 acc = lic_point<FT, NC, GATE>(L, pos[0], pos[1], pos[2]);
                    l4[0] = NC == 1 ? acc.x * L.scl : acc.x;
                    l4[1] = acc.y;
                    l4[2] = acc.z;
                    l4[3] = acc.w * L.scl;
                }
                Sample sm;
                shade_sample<FT, NC>(a, tf, op, rs, pos, fv, cf, l4, sm);
                for (int q = 0; q < 3; ++q) src[q] = sm.rgb[q] * sm.a;
                src[3] = sm.a;
            } else {
                ok = false;  // slicing: nothing composited for this slice
            }
        }
        // composite the group's G samples in order (every lane of the group alike)
        for (int g = 0; g < G; ++g) {
            const int sl = gl0 + g;
            if constexpr (SLICES) {
                if (bk + g >= S.n || (S.blend ? !(dest[3] < 255.0f) : !(dest[3] < 0.95f))) {
                    alive = false;
                    break;
                }
                if (__shfl((int)ok, sl) == 0) continue;
            } else {
                if (__shfl((int)ok, sl) == 0) {
                    alive = false;
                    break;
                }
            }
            float sv[4];
            for (int q = 0; q < 4; ++q) sv[q] = __shfl(src[q], sl);
            if (SLICES && S.blend) {
                blend_u8(dest, sv);
            } else {
                const float t = 1.0f - dest[3];
                for (int q = 0; q < 4; ++q) {
                    dest[q] = clamp01(fmaf(t, sv[q], dest[q]));
                    if constexpr (SLICES) dest[q] = half_rn(dest[q]);
                }
            }
            ++ns;
            if constexpr (!SLICES) {
                if (sv[3] > 0.95f) {
                    alive = false;
                    break;
                }
            }
        }
        if constexpr (!SLICES) {
            if (alive) {
                const int last = gl0 + G - 1;
                for (int q = 0; q < 3; ++q) base[q] = __shfl(nxt[q], last);
                if (__shfl((int)nxt_out, last)) alive = false;
            }
        }
        bk += G;
    }
    if (has_pix && sub == 0) {
        const long long pix = (long long)j * W + i;
        if (SLICES && S.blend)
            store_blend(a.image, pix, dest);
        else
            reinterpret_cast<float4*>(a.image)[pix] = make_float4(dest[0], dest[1], dest[2], dest[3]);
    }
    if (a.counter) add_samples(a.counter, sub == 0 ? ns : 0ull, red);
}

}  // namespace

hipError_t launch_render(const RenderLaunch& R, hipStream_t s)
{
    const int bx = R.block[0] > 0 ? R.block[0] : kBX, by = R.block[1] > 0 ? R.block[1] : kBY;
    if (bx * by > kRThreads || (bx * by) % 64) return hipErrorInvalidValue;
    dim3 block(bx, by);
    dim3 grid((R.view.width + bx - 1) / bx, (R.view.height + by - 1) / by);
    return with_field_type(R.field_type, [&](auto tag) {
        using FT = typename decltype(tag)::type;
        if (R.lc == 4)
            hipLaunchKernelGGL((k_render<FT, 4>), grid, block, 0, s, R);
        else
            hipLaunchKernelGGL((k_render<FT, 1>), grid, block, 0, s, R);
    });
}

hipError_t launch_render_rg(const RenderLaunch& R, int G, hipStream_t s)
{
    if (G < 2 || G > 64 || (G & (G - 1))) return hipErrorInvalidValue;
    const long long nblocks = rg_blocks(G, R.view.width, R.view.height);
    if (nblocks > 0x7fffffffLL) return hipErrorInvalidValue;
    dim3 grid((unsigned)nblocks), block(kRThreads);
#define VV_RG_LAUNCH(LC)                                                                    \
    do {                                                                                    \
        if (G == 2)                                                                         \
            hipLaunchKernelGGL((k_render_rg<FT, LC, 2>), grid, block, 0, s, R, G);          \
        else if (G == 4)                                                                    \
            hipLaunchKernelGGL((k_render_rg<FT, LC, 4>), grid, block, 0, s, R, G);          \
        else if (G == 8)                                                                    \
            hipLaunchKernelGGL((k_render_rg<FT, LC, 8>), grid, block, 0, s, R, G);          \
        else                                                                                \
            hipLaunchKernelGGL((k_render_rg<FT, LC, 0>), grid, block, 0, s, R, G);          \
    } while (0)
    return with_field_type(R.field_type, [&](auto tag) {
        using FT = typename decltype(tag)::type;
        if (R.lc == 4)
            VV_RG_LAUNCH(4);
        else
            VV_RG_LAUNCH(1);
    });
#undef VV_RG_LAUNCH
}

namespace {

// LIC variant of the on-the-fly casters: 2 cached with the noise cell on
// the field grid, 1 cached, 0 uncached (RGBA noise or the scalar gate)
struct LicVariant {
    bool gate;
    int nc;
    int mode;
};

LicVariant lic_variant(const LicLaunch& L)
{
    LicVariant v;
    v.gate = L.gate != 0;
    v.nc = L.nc;
    const bool cached = L.cache && L.nc == 1 && !v.gate;
    // the SAME kernels read the noise from channel 3 of the LIC field copy
    const bool same = L.field_has_noise && L.freq == 1.0f && L.noise.n[0] == L.field.n[0] &&
                      L.noise.n[1] == L.field.n[1] && L.noise.n[2] == L.field.n[2];
    v.mode = cached ? (same ? 2 : 1) : 0;
    return v;
}

}  // namespace

hipError_t launch_render_fused(const RenderLaunch& R, const LicLaunch& L, hipStream_t s)
{
    const int bx = R.block[0] > 0 ? R.block[0] : kBX, by = R.block[1] > 0 ? R.block[1] : kBY;
    if (bx * by > kRThreads || (bx * by) % 64) return hipErrorInvalidValue;
    LicArgs A;
    int nt = 0;
    const hipError_t e = lic_args_fill(L, &A, &nt);
    if (e != hipSuccess) return e;
    dim3 block(bx, by);
    dim3 grid((R.view.width + bx - 1) / bx, (R.view.height + by - 1) / by);
    const LicVariant v = lic_variant(L);
    return with_field_type(R.field_type, [&](auto tag) {
        using FT = typename decltype(tag)::type;
        if (v.mode == 2)
            hipLaunchKernelGGL((k_render_fused<FT, 1, false, 2>), grid, block, 0, s, R, A);
        else if (v.mode == 1)
            hipLaunchKernelGGL((k_render_fused<FT, 1, false, 1>), grid, block, 0, s, R, A);
        else if (v.nc == 4)
            hipLaunchKernelGGL((k_render_fused<FT, 4, false, 0>), grid, block, 0, s, R, A);
        else if (v.gate)
            hipLaunchKernelGGL((k_render_fused<FT, 1, true, 0>), grid, block, 0, s, R, A);
        else
            hipLaunchKernelGGL((k_render_fused<FT, 1, false, 0>), grid, block, 0, s, R, A);
    });
}

hipError_t launch_render_slices(const RenderLaunch& R, const LicLaunch& L, const SliceSetup& S, hipStream_t s)
{
    const int bx = R.block[0] > 0 ? R.block[0] : kBX, by = R.block[1] > 0 ? R.block[1] : kBY;
    if (bx * by > kRThreads || (bx * by) % 64) return hipErrorInvalidValue;
    LicArgs A;
    int nt = 0;
    const hipError_t e = lic_args_fill(L, &A, &nt);
    if (e != hipSuccess) return e;
    dim3 block(bx, by);
    dim3 grid((R.view.width + bx - 1) / bx, (R.view.height + by - 1) / by);
    const LicVariant v = lic_variant(L);
    return with_field_type(R.field_type, [&](auto tag) {
        using FT = typename decltype(tag)::type;
        if (v.mode == 2)
            hipLaunchKernelGGL((k_render_slices<FT, 1, false, 2>), grid, block, 0, s, R, A, S);
        else if (v.mode == 1)
            hipLaunchKernelGGL((k_render_slices<FT, 1, false, 1>), grid, block, 0, s, R, A, S);
        else if (v.nc == 4)
            hipLaunchKernelGGL((k_render_slices<FT, 4, false, 0>), grid, block, 0, s, R, A, S);
        else if (v.gate)
            hipLaunchKernelGGL((k_render_slices<FT, 1, true, 0>), grid, block, 0, s, R, A, S);
        else
            hipLaunchKernelGGL((k_render_slices<FT, 1, false, 0>), grid, block, 0, s, R, A, S);
    });
}

hipError_t launch_render_group(const RenderLaunch& R, const LicLaunch& L, const SliceSetup& S, int slices, int G,
                               hipStream_t s)
{
    if (G < 1 || G > 64 || (G & (G - 1))) return hipErrorInvalidValue;
    LicArgs A;
    int nt = 0;
    const hipError_t e = lic_args_fill(L, &A, &nt);
    if (e != hipSuccess) return e;
    const int nr = kRThreads / G;
    int lg = 0;
    while ((1 << (lg + 1)) <= nr) ++lg;
    const int TW = 1 << ((lg + 1) >> 1), TH = nr / TW;
    const long long nblocks = (long long)((R.view.width + TW - 1) / TW) * ((R.view.height + TH - 1) / TH);
    if (nblocks > 0x7fffffffLL) return hipErrorInvalidValue;
    dim3 grid((unsigned)nblocks), block(kRThreads);
    const LicVariant v = lic_variant(L);
#define VV_GROUP_LAUNCH(FT, NC, GATE, MODE)                                                                 \
    do {                                                                                                   \
        if (slices)                                                                                        \
            hipLaunchKernelGGL((k_render_group<FT, NC, GATE, MODE, true>), grid, block, 0, s, R, A, S, G);  \
        else                                                                                               \
            hipLaunchKernelGGL((k_render_group<FT, NC, GATE, MODE, false>), grid, block, 0, s, R, A, S, G); \
    } while (0)
    return with_field_type(R.field_type, [&](auto tag) {
        using FT = typename decltype(tag)::type;
        if (v.mode == 2)
            VV_GROUP_LAUNCH(FT, 1, false, 2);
        else if (v.mode == 1)
            VV_GROUP_LAUNCH(FT, 1, false, 1);
        else if (v.nc == 4)
            VV_GROUP_LAUNCH(FT, 4, false, 0);
        else if (v.gate)
            VV_GROUP_LAUNCH(FT, 1, true, 0);
        else
            VV_GROUP_LAUNCH(FT, 1, false, 0);
    });
#undef VV_GROUP_LAUNCH
}

// ---- scratch of the allocation-free adjoint ---------------------------------
// vv_render_adjoint_noalloc takes its checkpoints, segment columns, sample
// counts and fixed-point accumulators from this array instead of device
// allocations.  Every GPU the library runs on holds its own copy until the
// process ends.  On the host the array's name is only a stand-in: host code
// uses the address noalloc_scratch() reports for the device.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];
}  // namespace

hipError_t noalloc_scratch(int device, void** ptr)
{
    *ptr = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}

hipError_t launch_adjoint_forward(const AdjointLaunch& A, hipStream_t s)
{
    if (A.group == 2 || A.group == 4) {
        const int G = A.group;
        const long long nblocks = rg_blocks(G, A.r.view.width, A.r.view.height);
        if (nblocks > 0x7fffffffLL) return hipErrorInvalidValue;
        dim3 grid((unsigned)nblocks), block(kRThreads);
        return with_field_type(A.r.field_type, [&](auto tag) {
            using FT = typename decltype(tag)::type;
            if (A.r.lc == 4) {
                if (G == 2)
                    hipLaunchKernelGGL((k_adjoint_forward_rg<FT, 4, 2>), grid, block, 0, s, A);
                else
                    hipLaunchKernelGGL((k_adjoint_forward_rg<FT, 4, 4>), grid, block, 0, s, A);
            } else {
                if (G == 2)
                    hipLaunchKernelGGL((k_adjoint_forward_rg<FT, 1, 2>), grid, block, 0, s, A);
                else
                    hipLaunchKernelGGL((k_adjoint_forward_rg<FT, 1, 4>), grid, block, 0, s, A);
            }
        });
    }
    dim3 block(kBX, kBY);
    dim3 grid((A.r.view.width + kBX - 1) / kBX, (A.r.view.height + kBY - 1) / kBY);
    return with_field_type(A.r.field_type, [&](auto tag) {
        using FT = typename decltype(tag)::type;
        if (A.r.lc == 4)
            hipLaunchKernelGGL((k_adjoint_forward<FT, 4>), grid, block, 0, s, A);
        else
            hipLaunchKernelGGL((k_adjoint_forward<FT, 1>), grid, block, 0, s, A);
    });
}

hipError_t launch_render_adjoint(const AdjointLaunch& A, hipStream_t s)
{
    const long long nrays = (long long)A.r.view.width * A.r.view.height;
    hipLaunchKernelGGL(k_adjoint_bound, dim3((unsigned)((nrays + 255) / 256)), dim3(256), 0, s, A);
    if (A.group == 2 || A.group == 4) {
        const int G = A.group;
        const long long nblocks = rg_blocks(G, A.r.view.width, A.r.view.height);
        if (nblocks > 0x7fffffffLL) return hipErrorInvalidValue;
        dim3 grid((unsigned)nblocks), block(kRThreads);
#define VV_ADJ_RG(LC, GT)                                                                        \
    do {                                                                                         \
        if (A.full_col)                                                                          \
            hipLaunchKernelGGL((k_render_adjoint_rg<FT, LC, GT, false>), grid, block, 0, s, A); \
        else                                                                                     \
            hipLaunchKernelGGL((k_render_adjoint_rg<FT, LC, GT, true>), grid, block, 0, s, A);  \
    } while (0)
        return with_field_type(A.r.field_type, [&](auto tag) {
            using FT = typename decltype(tag)::type;
            if (A.r.lc == 4) {
                if (G == 2)
                    VV_ADJ_RG(4, 2);
                else
                    VV_ADJ_RG(4, 4);
            } else {
                if (G == 2)
                    VV_ADJ_RG(1, 2);
                else
                    VV_ADJ_RG(1, 4);
            }
        });
#undef VV_ADJ_RG
    }
    dim3 block(kBX, kBY);
    dim3 grid((A.r.view.width + kBX - 1) / kBX, (A.r.view.height + kBY - 1) / kBY);
    return with_field_type(A.r.field_type, [&](auto tag) {
        using FT = typename decltype(tag)::type;
        if (A.r.lc == 4)
            hipLaunchKernelGGL((k_render_adjoint<FT, 4>), grid, block, 0, s, A);
        else
            hipLaunchKernelGGL((k_render_adjoint<FT, 1>), grid, block, 0, s, A);
    });
}

hipError_t launch_render_batch(const RenderBatchLaunch& B, hipStream_t s)
{
    const int bx = B.block[0], by = B.block[1];
    if (B.nviews < 1 || (B.group != 1 && B.group != 2 && B.group != 4)) return hipErrorInvalidValue;
    if (B.group == 1 && (bx * by > kRThreads || (bx * by) % 64)) return hipErrorInvalidValue;
    dim3 block = B.group > 1 ? dim3(kRThreads) : dim3(bx, by);
    dim3 grid((unsigned)B.grid);
#define VV_RB_LAUNCH(LC, GT)                                                                               \
    hipLaunchKernelGGL((k_render_batch<FT, LC, GT>), grid, block, 0, s, B.views, B.first, B.nviews, B.total, \
                       B.counter, B.done, B.begun)
    return with_field_type(B.field_type, [&](auto tag) {
        using FT = typename decltype(tag)::type;
        if (B.lc == 4) {
            if (B.group == 4) VV_RB_LAUNCH(4, 4);
            else if (B.group == 2) VV_RB_LAUNCH(4, 2);
            else VV_RB_LAUNCH(4, 0);
        } else {
            if (B.group == 4) VV_RB_LAUNCH(1, 4);
            else if (B.group == 2) VV_RB_LAUNCH(1, 2);
            else VV_RB_LAUNCH(1, 0);
        }
    });
#undef VV_RB_LAUNCH
}

hipError_t launch_stamp(unsigned long long* t0, hipStream_t s)
{
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(1), 0, s, t0);
    return hipGetLastError();
}

hipError_t launch_fixed_to_float(const unsigned long long* hdr, const long long* src, float* dst, long long n,
                                 hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fixed_to_float, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, hdr, src, dst, n);
    return hipGetLastError();
}

}  // namespace vv

namespace vv {

int render_batch_blocks_per_cu(int field_type, int lc, int threads, int group)
{
    int n = 0;
    hipError_t e = hipSuccess;
    const hipError_t d = with_field_type(field_type, [&](auto tag) {
        using FT = typename decltype(tag)::type;
        auto occ = [&](auto kernel, int nt) { e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, nt, 0); };
        if (group == 4)
            lc == 4 ? occ(k_render_batch<FT, 4, 4>, kRThreads) : occ(k_render_batch<FT, 1, 4>, kRThreads);
        else if (group == 2)
            lc == 4 ? occ(k_render_batch<FT, 4, 2>, kRThreads) : occ(k_render_batch<FT, 1, 2>, kRThreads);
        else
            lc == 4 ? occ(k_render_batch<FT, 4, 0>, threads) : occ(k_render_batch<FT, 1, 0>, threads);
    });
    return (d == hipSuccess && e == hipSuccess && n > 0) ? n : 1;
}

}  // namespace vv
