/*
 * vvoracle.c — CPU restatement of the reference's LIC + ray-cast path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Build: oracle/Makefile  (gcc -O2 -ffp-contract=off -fopenmp).
 *
 * Pinning (what fixes each function independently of its own formula;
 * the tests named here live in tests/test_oracle.py):
 *   orc_f32_to_f16/orc_f16_to_f32   numpy.float16 round trip over all 65536
 *                                    halves and random floats (RNE, denormals)
 *   orc_gen_noise                    SplitMix64 published vectors (seed
 *                                    1234567), binomial density bounds
 *   orc_noise_gradients              constant volume -> 127 bytes
 *                                    (VV/gradient.cpp:468-477); linear ramps ->
 *                                    axis normal; border layer zero; brute
 *                                    force 3x3x3 Sobel in numpy on 8^3
 *   orc_make_filter                  box inv-area 0.5 (VV/dataset.cpp:1412),
 *                                    closed-form triangle area, symmetry
 *   orc_encode_field_*               closed form, zero vector -> (.5,.5,.5,0)
 *   orc_encode_field_u8              axis directions -> bytes 255/1, zero
 *                                    vector -> 128, byte-boundary vectors
 *                                    (v/|v| = (k-128)/127 exactly), max |v|
 *                                    -> 255, all-zero field -> magnitude 0
 *   orc_lic_*                        constant noise closed form, zero noise,
 *                                    uniform field straight streamline with
 *                                    displacement n*h_e, numpy brute force on
 *                                    8^3, f32 vs f64 agreement
 *   orc_view_setup                   numpy gluPerspective/MV inverse, centre
 *                                    pixel hits the face centre
 *   orc_raycast_*                    constant volume closed form
 *                                    1-(1-a)^n and termination index, empty
 *                                    TF -> zero image, one-hot TF
 *   orc_raycast_adjoint_f64          central finite differences of
 *                                    orc_raycast_f64
 */
#include "vvoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* IEEE half <-> float, round to nearest even, denormals kept.        */
/* ------------------------------------------------------------------ */
uint16_t orc_f32_to_f16(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) {                 /* inf or nan */
        return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u | ((ax >> 13) & 0x3ffu) : 0u));
    }
    if (ax >= 0x477ff000u)                   /* rounds to >= 65520 -> inf */
        return (uint16_t)(sign | 0x7c00u);
    if (ax < 0x38800000u) {                  /* result is a half denormal (or zero) */
        if (ax < 0x33000000u)                /* < 2^-25: rounds to zero */
            return (uint16_t)sign;
        uint32_t e = ax >> 23;               /* 102..112 */
        uint32_t m = (ax & 0x7fffffu) | 0x800000u;
        uint32_t shift = 126u - e;           /* 14..24 */
        uint32_t q = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (q & 1u)))
            q += 1u;
        return (uint16_t)(sign | q);
    }
    uint32_t q = ax - 0x38000000u;           /* rebias exponent 127 -> 15 */
    uint32_t rem = q & 0x1fffu;
    q >>= 13;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u)))
        q += 1u;
    return (uint16_t)(sign | q);
}

float orc_f16_to_f32(uint16_t h)
{
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = ((uint32_t)h >> 10) & 0x1fu;
    uint32_t m = (uint32_t)h & 0x3ffu;
    uint32_t x;
    if (e == 0) {
        if (m == 0) {
            x = sign;
        } else {                             /* denormal: normalise */
            int k = 0;
            while (!(m & 0x400u)) { m <<= 1; ++k; }
            m &= 0x3ffu;
            x = sign | ((uint32_t)(113 - k) << 23) | (m << 13);
        }
    } else if (e == 31) {
        x = sign | 0x7f800000u | (m << 13);
    } else {
        x = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

void orc_f32_to_f16_array(const float* in, uint16_t* out, long n)
{
    for (long i = 0; i < n; ++i) out[i] = orc_f32_to_f16(in[i]);
}

void orc_f16_to_f32_array(const uint16_t* in, float* out, long n)
{
    for (long i = 0; i < n; ++i) out[i] = orc_f16_to_f32(in[i]);
}

/* ------------------------------------------------------------------ */
/* UNORM8 field encode: fillTexDataCharInterp, float source branch     */
/* (VV/dataset.cpp:704-823, :748-770 and :776-782).  The time lerp is  */
/* the float path's; direction byte = (unsigned char)(q*127.0 + 128.0) */
/* with q = v/|v| a float quotient and the rest in double, as written  */
/* (floor(127q + 128), except that the double sum rounds to 128 for    */
/* -2^-47 <= 127q < 0); 128 when |v| < EPS (1e-5f).  The               */
/* magnitude byte is (unsigned char)(|v|/max*255) in float, clamped to */
/* [0,255].  A field with max |v| = 0 (0/0 in the reference) gets      */
/* magnitude 0.  GL samples the bytes as UNORM8: value = byte/255.     */
/* ------------------------------------------------------------------ */
void orc_encode_field_u8(const float* v0, const float* v1, float t,
                         int nx, int ny, int nz, uint8_t* out)
{
    const long n = (long)nx * ny * nz;
    float* mag = (float*)malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
    float maxlen = -1.0f;
    if (!v1) v1 = v0;
    for (long i = 0; i < n; ++i) {
        float tv[3];
        for (int c = 0; c < 3; ++c) {
            const float a = v0[3 * i + c], b = v1[3 * i + c];
            tv[c] = a + t * (b - a);
        }
        float len = sqrtf(tv[0] * tv[0] + tv[1] * tv[1] + tv[2] * tv[2]);
        if (len < 1e-5f) {
            len = 0.0f;
            out[4 * i + 0] = out[4 * i + 1] = out[4 * i + 2] = 128;
        } else {
            for (int c = 0; c < 3; ++c) {
                const float q = tv[c] / len;
                out[4 * i + c] = (uint8_t)((double)q * 127.0 + 128.0);
            }
        }
        mag[i] = len;
        if (len > maxlen) maxlen = len;
    }
    for (long i = 0; i < n; ++i) {
        float m = maxlen > 0.0f ? mag[i] / maxlen * 255.0f : 0.0f;
        out[4 * i + 3] = (uint8_t)(m > 255.0f ? 255.0f : (m < 0.0f ? 0.0f : m));
    }
    free(mag);
}

/* ------------------------------------------------------------------ */
/* Seeded noise (replaces the MSVC rand() stream of                    */
/* VV/dataset.cpp:1158-1162, which cannot be reproduced).              */
/* noise[i] is the (i+1)-th output of a SplitMix64 generator seeded    */
/* with `seed`: state_i = seed + (i+1)*golden, out = finalize(state).  */
/* kind 0: binary {0,255}, 255 with probability `density`              */
/*         (test: (out>>32) < floor(density*2^32));                    */
/* kind 1: dense uniform byte (out>>56).                               */
/* ------------------------------------------------------------------ */
static inline uint64_t orc_splitmix64_at(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + (i + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t orc_splitmix64(uint64_t seed, uint64_t i) { return orc_splitmix64_at(seed, i); }

void orc_gen_noise(uint64_t seed, int kind, double density,
                   int nx, int ny, int nz, uint8_t* out)
{
    long n = (long)nx * ny * nz;
    double t = floor(density * 4294967296.0);
    if (t < 0.0) t = 0.0;
    if (t > 4294967296.0) t = 4294967296.0;
    uint64_t thr = (uint64_t)t;
    for (long i = 0; i < n; ++i) {
        uint64_t u = orc_splitmix64_at(seed, (uint64_t)i);
        if (kind == 0)
            out[i] = ((u >> 32) < thr) ? 255 : 0;
        else
            out[i] = (uint8_t)(u >> 56);
    }
}

/* ------------------------------------------------------------------ */
/* Noise gradients: VV/gradient.cpp:190-374 (computeGradients, SOBEL   */
/* == 1), :377-459 (filterGradients), :462-532 (quantize8).  All float */
/* arithmetic is the reference's own (float), quantize is double.      */
/* ------------------------------------------------------------------ */
static const int orc_sobel[3][3][3][3] = {
    {{{-1, -3, -1}, {-3, -6, -3}, {-1, -3, -1}},
     {{ 0,  0,  0}, { 0,  0,  0}, { 0,  0,  0}},
     {{ 1,  3,  1}, { 3,  6,  3}, { 1,  3,  1}}},
    {{{-1, -3, -1}, { 0,  0,  0}, { 1,  3,  1}},
     {{-3, -6, -3}, { 0,  0,  0}, { 3,  6,  3}},
     {{-1, -3, -1}, { 0,  0,  0}, { 1,  3,  1}}},
    {{{-1,  0,  1}, {-3,  0,  3}, {-1,  0,  1}},
     {{-3,  0,  3}, {-6,  0,  6}, {-3,  0,  3}},
     {{-1,  0,  1}, {-3,  0,  3}, {-1,  0,  1}}}};

void orc_noise_gradients_raw(const uint8_t* v, int nx, int ny, int nz,
                             const float sd[3], float* g)
{
#define VOX(x, y, z) ((float)v[((long)(z) * ny + (y)) * nx + (x)])
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                float* gp = g + 3 * (((long)z * ny + y) * nx + x);
                if (x > 0 && x < nx - 1 && y > 0 && y < ny - 1 && z > 0 && z < nz - 1) {
                    for (int d = 0; d < 3; ++d) {
                        gp[d] = 0.0f;
                        for (int i = -1; i < 2; ++i)
                            for (int j = -1; j < 2; ++j)
                                for (int k = -1; k < 2; ++k)
                                    gp[d] += (float)orc_sobel[d][i + 1][j + 1][k + 1] * VOX(x + i, y + j, z + k);
                        gp[d] /= 2.0f * sd[d];
                    }
                } else {
                    gp[0] = (x < 1) ? (VOX(x + 1, y, z) - VOX(x, y, z)) / sd[0]
                                    : (VOX(x, y, z) - VOX(x - 1, y, z)) / sd[0];
                    gp[1] = (y < 1) ? (VOX(x, y + 1, z) - VOX(x, y, z)) / sd[1]
                                    : (VOX(x, y, z) - VOX(x, y - 1, z)) / sd[1];
                    gp[2] = (z < 1) ? (VOX(x, y, z + 1) - VOX(x, y, z)) / sd[2]
                                    : (VOX(x, y, z) - VOX(x, y, z - 1)) / sd[2];
                }
            }
#undef VOX
}

void orc_filter_gradients(int nx, int ny, int nz, float* g)
{
    const int fs = 5;
    float filter[125];
    float sum = 0.0f;
    int fw = 2;
    memset(filter, 0, sizeof(filter));
    for (int k = -fw; k < fw - 1; ++k)
        for (int j = -fw; j < fw - 1; ++j)
            for (int i = -fw; i < fw - 1; ++i) {
                float w = expf(-(float)(i * i + j * j + k * k) / 5.0f);
                filter[((fw + k) * fs + fw + j) * fs + fw + i] = w;
                sum += w;
            }
    for (int k = -fw; k < fw - 1; ++k)
        for (int j = -fw; j < fw - 1; ++j)
            for (int i = -fw; i < fw - 1; ++i)
                filter[((fw + k) * fs + fw + j) * fs + fw + i] /= sum;

    long n = (long)nx * ny * nz;
    float* out = (float*)malloc(sizeof(float) * 3 * n);
    long gi = 0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                int bx = x < nx - x - 1 ? x : nx - x - 1;
                int by = y < ny - y - 1 ? y : ny - y - 1;
                int bz = z < nz - z - 1 ? z : nz - z - 1;
                int w = bx < by ? bx : by;
                w = w < bz ? w : bz;
                w = w < 2 ? w : 2;
                for (int c = 0; c < 3; ++c) {
                    float acc = 0.0f;
                    for (int k = -w; k < w - 1; ++k)
                        for (int j = -w; j < w - 1; ++j)
                            for (int i = -w; i < w - 1; ++i) {
                                long ogi = 3 * (((long)(z + k) * ny + (y + j)) * nx + (x + i)) + c;
                                acc += filter[((w + k) * fs + w + j) * fs + w + i] * g[ogi];
                            }
                    out[gi++] = acc;
                }
            }
    memcpy(g, out, sizeof(float) * 3 * n);
    free(out);
}

void orc_quantize_gradients(long n, float* g, uint8_t* rgb)
{
    for (long i = 0; i < n; ++i) {
        float* gp = g + 3 * i;
        float len = sqrtf(gp[0] * gp[0] + gp[1] * gp[1] + gp[2] * gp[2]);
        if (len < 1e-5f) {
            gp[0] = gp[1] = gp[2] = 0.0f;
        } else {
            gp[0] /= len;
            gp[1] /= len;
            gp[2] /= len;
        }
        for (int c = 0; c < 3; ++c)
            rgb[3 * i + c] = (uint8_t)(((double)gp[c] + 1.0) / 2.0 * 255);
    }
}

int orc_noise_gradients(const uint8_t* noise, int nx, int ny, int nz,
                        const float spacing[3], uint8_t* grad_rgb)
{
    long n = (long)nx * ny * nz;
    float* g = (float*)malloc(sizeof(float) * 3 * n);
    if (!g) return -1;
    orc_noise_gradients_raw(noise, nx, ny, nz, spacing, g);
    orc_filter_gradients(nx, ny, nz, g);
    orc_quantize_gradients(n, g, grad_rgb);
    free(g);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Filter LUT: W = 256 bytes (VV/dataset.cpp:1405-1413 box,            */
/* :1416-1467 PNG path).  The triangle/cos^2/Gaussian PNGs are not     */
/* shipped, so they are synthesised: texel i sits at t=(i+.5)/256,     */
/* s = (2t-1)*L steps, byte = round(255*w).                            */
/* inv_area per calcFilterKernelInvArea (:1503-1512), box = 0.5.       */
/* ------------------------------------------------------------------ */
float orc_filter_inv_area(const uint8_t* lut, int width)
{
    float area = 0.0f;
    for (int i = 0; i < width; ++i) area += lut[i];
    return 0.5f * width * 255.0f / area;
}

int orc_make_filter(int kind, int length, double sigma,
                    uint8_t lut[256], int* width, float* inv_area)
{
    const int W = 256;
    if (kind == 0) {
        memset(lut, 255, W);
        *width = W;
        *inv_area = 0.5f;
        return 0;
    }
    if (length <= 0) return -1;
    if (sigma <= 0.0) sigma = length / 4.0;
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < W; ++i) {
        double t = (i + 0.5) / W;
        double c = 2.0 * t - 1.0;
        double w;
        switch (kind) {
        case 1: w = 1.0 - fabs(c); break;
        case 2: { double q = cos(pi * (t - 0.5)); w = q * q; } break;
        case 3: { double s = c * length; w = exp(-s * s / (2.0 * sigma * sigma)); } break;
        default: return -1;
        }
        double b = floor(255.0 * w + 0.5);
        if (b < 0) b = 0;
        if (b > 255) b = 255;
        lut[i] = (uint8_t)b;
    }
    *width = W;
    *inv_area = orc_filter_inv_area(lut, W);
    return 0;
}

/* ------------------------------------------------------------------ */
/* View: gluPerspective(fovy, aspect, near, far) and                   */
/* MV = T(pos) T(0,0,-dist) R(q) T(-center)  (VV/camera.cpp:56-68,     */
/* VV/renderer.cpp:146).  Object-space camera = MV^-1 (0,0,0,1)        */
/* (gl_ModelViewMatrixInverse[3], raycast_lic3d_fragment.glsl:17).     */
/* Ray through NDC (x,y): eye dir (x*aspect*tan, y*tan, -1), rotated   */
/* by R^T.                                                             */
/* ------------------------------------------------------------------ */
static void orc_quat_matrix(const double q[4], double R[3][3])
{
    double x = q[0], y = q[1], z = q[2], w = q[3];
    double n = sqrt(x * x + y * y + z * z + w * w);
    if (n > 0) { x /= n; y /= n; z /= n; w /= n; }
    R[0][0] = 1 - 2 * (y * y + z * z); R[0][1] = 2 * (x * y - z * w);     R[0][2] = 2 * (x * z + y * w);
    R[1][0] = 2 * (x * y + z * w);     R[1][1] = 1 - 2 * (x * x + z * z); R[1][2] = 2 * (y * z - x * w);
    R[2][0] = 2 * (x * z - y * w);     R[2][1] = 2 * (y * z + x * w);     R[2][2] = 1 - 2 * (x * x + y * y);
}

void orc_view_setup(double fovy_deg, int width, int height, double dist,
                    const double pos[3], const double quat[4],
                    const double center[3], const double extent[3],
                    const double scale[3], orc_view* v)
{
    double R[3][3];
    orc_quat_matrix(quat, R);
    const double pi = 3.14159265358979323846;
    double th = tan(0.5 * fovy_deg * pi / 180.0);
    double aspect = (double)width / (double)height;
    double e[3] = {-pos[0], -pos[1], dist - pos[2]};
    for (int i = 0; i < 3; ++i) {
        /* R^T applied: column i of R */
        double rt_e = R[0][i] * e[0] + R[1][i] * e[1] + R[2][i] * e[2];
        v->cam[i] = (float)(center[i] + rt_e);
        v->du[i] = (float)(R[0][i] * aspect * th);
        v->dv[i] = (float)(R[1][i] * th);
        v->dw[i] = (float)(-R[2][i]);
        v->extent[i] = (float)extent[i];
        v->scale[i] = (float)scale[i];
        v->tex_max[i] = (float)(extent[i] * scale[i]);
    }
    v->inv_w2 = (float)(2.0 / width);
    v->inv_h2 = (float)(2.0 / height);
    v->width = width;
    v->height = height;
}

/* Renderer::updateLightPos (VV/renderer.cpp:431-466): the light's
 * Transform position (0,0,dist) rotated by its quaternion, mapped by
 * T(center) R(-camAngle) T(camPos) and loaded as an eye-space position. */
void orc_light_position(const double center[3], const double cam_pos[3],
                        const double cam_quat[4], double light_dist,
                        const double light_quat[4], double out[3])
{
    double RL[3][3], RC[3][3];
    orc_quat_matrix(light_quat, RL);
    orc_quat_matrix(cam_quat, RC);
    double lp[3];
    for (int i = 0; i < 3; ++i) lp[i] = RL[i][2] * light_dist + cam_pos[i];
    for (int i = 0; i < 3; ++i)
        out[i] = center[i] + RC[0][i] * lp[0] + RC[1][i] * lp[1] + RC[2][i] * lp[2];
}

/* ------------------------------------------------------------------ */
/* Precision-dependent part, compiled twice.                          */
/* ------------------------------------------------------------------ */
/* ------------------------------------------------------------------ */
/* Line-illumination tables (VV/illumination.cpp:96-390).  The          */
/* expressions keep the reference's types: material entries are float,  */
/* so ambient*lightColor is a float product; the rest is double.        */
/* ------------------------------------------------------------------ */
static double mallo_integrand(double beta, double n, double theta)
{
    double y = cos(theta - beta);
    if (y < 0.0) y = 0.0;
    return pow(y, n) * (cos(theta) / 2.0);
}

/* composite Simpson rule over [alpha - pi/2, pi/2] with 2m = 20 intervals,
 * summed in the reference's order: pairs 2 f(x_i) + 4 f(x_i+1) for even i,
 * then + f(b) - f(a) (f(a) entered the loop with weight 2), times h/3 */
double orc_mallo_spec_term(double alpha, double beta, double n)
{
    const double a = alpha - M_PI / 2.0, b = M_PI / 2.0;
    const int m = 10;
    const double h = (b - a) / (2.0 * m);
    double sum = 0.0;
    for (int i = 0; i < 2 * m; i += 2) {
        sum += 2.0 * mallo_integrand(beta, n, a + i * h);
        sum += 4.0 * mallo_integrand(beta, n, a + (i + 1) * h);
    }
    sum += mallo_integrand(beta, n, b);
    sum -= mallo_integrand(beta, n, a);
    return sum * (h / 3.0);
}

static float half_round(double v) { return orc_f16_to_f32(orc_f32_to_f16((float)v)); }
static double clamp01d(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

void orc_illum_luts(const orc_line_mat* m, float* zoeckler, float* mallo_diff, float* mallo_spec)
{
    const int W = 256, H = 256;
    const double irx = 1.0 / (double)(W - 1), iry = 1.0 / (double)(H - 1);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const long i = (long)y * W + x;
            /* Zoeckler: texel (x, y) <-> (L.T, V.T) = (2x/(W-1) - 1, 2y/(H-1) - 1) */
            const double lt = 2.0 * ((double)x * irx) - 1.0;
            const double vt = 2.0 * ((double)y * iry) - 1.0;
            double diffuse = pow(sqrt(1.0 - lt * lt), (double)m->diff_exp);
            double dp = lt * vt - sqrt(1.0 - lt * lt) * sqrt(1.0 - vt * vt);
            dp = dp < -1.0 ? -1.0 : (dp > 1.0 ? 1.0 : dp);
            const float amb = m->ambient[0] * m->light_color[0];
            double dterm = amb + diffuse * m->diffuse[0] * m->light_color[0];
            double sterm = pow(fabs(dp), (double)m->spec_exp) * m->specular[0] * m->light_color[0];
            dterm = clamp01d(dterm * 0.5);
            sterm = clamp01d(sterm * 0.5);
            zoeckler[2 * i] = half_round(dterm);
            zoeckler[2 * i + 1] = half_round((double)((float)sterm * 0.9f));
            /* Mallo: texel centres, alpha = acos(2s-1), beta = acos(2t-1), L.T = 2t-1 */
            const double s = ((double)x + 0.5) / W, t = ((double)y + 0.5) / H;
            const double alpha = acos(2.0 * s - 1.0), beta = acos(2.0 * t - 1.0);
            const double l2 = 2.0 * t - 1.0;
            const double fd = sqrt(1.0 - l2 * l2) * (sin(alpha) + (M_PI - alpha) * cos(alpha)) * 0.25;
            const double fs = 3.5 * orc_mallo_spec_term(alpha, beta, (double)m->spec_exp);
            for (int c = 0; c < 3; ++c) {
                mallo_diff[3 * i + c] = half_round(clamp01d(fd * m->diffuse[c] * m->light_color[c]));
                mallo_spec[3 * i + c] = half_round(clamp01d(fs * m->specular[c] * m->light_color[c]));
            }
        }
}

#define ORC_DOUBLE 0
#include "vvoracle_impl.h"
#undef ORC_DOUBLE
#define ORC_DOUBLE 1
#include "vvoracle_impl.h"
#undef ORC_DOUBLE
