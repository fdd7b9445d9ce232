// vv_illum.cpp — tables of the illuminated-streamline shading modes.
//
// Zoeckler et al. (VIS 1996): for the texel (x, y) of a 256x256 table,
// L.T = 2x/255 - 1 and V.T = 2y/255 - 1; the diffuse term is the ambient
// term plus sqrt(1 - (L.T)^2)^diffExp times the diffuse material, the
// specular term |L.T V.T - sqrt(1-(L.T)^2) sqrt(1-(V.T)^2)|^specExp times
// the specular material, each halved and clamped (VV/illumination.cpp:96-183).
// Mallo et al. (VIS 2005): at texel centres s, t, alpha = acos(2s-1),
// beta = acos(2t-1); diffuse sqrt(1-(2t-1)^2)(sin a + (pi-a) cos a)/4 and
// specular 3.5 times the integral over [a - pi/2, pi/2] of
// max(cos(theta - beta), 0)^n cos(theta)/2, by the composite Simpson rule on
// 20 intervals (VV/illumination.cpp:188-320, 348-390).  The reference builds
// float textures (VV/3DLIC.cpp:735), which GL stores as FP16: every value is
// rounded to the nearest half here.
#include "vv_ctx.hpp"

#include <cstdint>

using namespace vv_host;

namespace {

// round to the nearest FP16 value (ties to even), returned as float; the
// tables hold values in [0, 1]
float half_round(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = x & 0x80000000u, ax = x & 0x7fffffffu;
    float r;
    if (ax >= 0x7f800000u) return f;  // inf, NaN
    if (ax < 0x38800000u) {
        // below 2^-14: FP16 subnormals, a multiple of 2^-24 (exact scalings)
        r = std::nearbyint(std::fabs(f) * 16777216.0f) / 16777216.0f;
    } else {
        uint32_t y = ax + 0x0fffu + ((ax >> 13) & 1u);
        y &= 0xffffe000u;
        if (y > 0x477fe000u) y = 0x7f800000u;  // above 65504: inf
        std::memcpy(&r, &y, 4);
    }
    return sign ? -r : r;
}

double clamp01d(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

double spec_integrand(double beta, double n, double theta)
{
    double y = std::cos(theta - beta);
    if (y < 0.0) y = 0.0;
    return std::pow(y, n) * (std::cos(theta) / 2.0);
}

// Simpson's rule in the reference's summation order: 2 f(x_i) + 4 f(x_i+1)
// for even i, then + f(b) - f(a), times h/3
double mallo_spec_term(double alpha, double beta, double n)
{
    const double pi = 3.14159265358979323846;
    const double a = alpha - pi / 2.0, b = pi / 2.0;
    const int m = 10;
    const double h = (b - a) / (2.0 * m);
    double sum = 0.0;
    for (int i = 0; i < 2 * m; i += 2) {
        sum += 2.0 * spec_integrand(beta, n, a + i * h);
        sum += 4.0 * spec_integrand(beta, n, a + (i + 1) * h);
    }
    sum += spec_integrand(beta, n, b);
    sum -= spec_integrand(beta, n, a);
    return sum * (h / 3.0);
}

}  // namespace

namespace vv_host {

vv_status ensure_illum_luts(vv_ctx* c)
{
    if (!c->illum_dirty && c->illum_luts.p) return VV_OK;
    const size_t n = 256 * 256;
    std::vector<float> host(8 * n);
    vv_status s = vv_make_illum_tables(&c->line_mat, host.data(), host.data() + 2 * n, host.data() + 5 * n);
    if (s) return fail(c, s, "%s", g_create_error.c_str());
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, c->illum_luts.ensure(sizeof(float) * host.size()));
    HIPCHK(c, hipMemcpy(c->illum_luts.p, host.data(), sizeof(float) * host.size(), hipMemcpyHostToDevice));
    c->illum_dirty = false;
    return VV_OK;
}

}  // namespace vv_host

extern "C" {

void vv_default_line_material(vv_line_material* m)
{
    if (!m) return;
    for (int i = 0; i < 4; ++i) {
        m->light_color[i] = 1.0f;
        m->ambient[i] = 0.1f;
        m->diffuse[i] = 0.5f;
        m->specular[i] = 0.8f;
    }
    m->spec_exp = 40.0f;
    m->diff_exp = 2.0f;
}

vv_status vv_make_illum_tables(const vv_line_material* m, float* zoeckler, float* mallo_diff, float* mallo_spec)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!m || !zoeckler || !mallo_diff || !mallo_spec) return fail(nullptr, VV_EINVAL, "NULL argument");
        for (int i = 0; i < 4; ++i)
            if (!std::isfinite(m->light_color[i]) || !std::isfinite(m->ambient[i]) || !std::isfinite(m->diffuse[i]) ||
                !std::isfinite(m->specular[i]))
                return fail(nullptr, VV_EINVAL, "line material not finite");
        if (!(m->spec_exp >= 0.0f && m->spec_exp <= 1024.0f) || !(m->diff_exp >= 0.0f && m->diff_exp <= 1024.0f))
            return fail(nullptr, VV_EINVAL, "line material exponents must be in [0, 1024]");
        const double pi = 3.14159265358979323846;
        const double irx = 1.0 / 255.0, iry = 1.0 / 255.0;
        for (int y = 0; y < 256; ++y)
            for (int x = 0; x < 256; ++x) {
                const size_t i = (size_t)y * 256 + x;
                const double lt = 2.0 * ((double)x * irx) - 1.0;
                const double vt = 2.0 * ((double)y * iry) - 1.0;
                const double dif = std::pow(std::sqrt(1.0 - lt * lt), (double)m->diff_exp);
                double dot = lt * vt - std::sqrt(1.0 - lt * lt) * std::sqrt(1.0 - vt * vt);
                dot = dot < -1.0 ? -1.0 : (dot > 1.0 ? 1.0 : dot);
                // ambient * lightColor is a float product in the reference
                const float amb = m->ambient[0] * m->light_color[0];
                const double dz = clamp01d((amb + dif * m->diffuse[0] * m->light_color[0]) * 0.5);
                const double sz =
                    clamp01d(std::pow(std::fabs(dot), (double)m->spec_exp) * m->specular[0] * m->light_color[0] * 0.5);
                zoeckler[2 * i] = half_round((float)dz);
                zoeckler[2 * i + 1] = half_round((float)sz * 0.9f);

                const double s = ((double)x + 0.5) / 256.0, t = ((double)y + 0.5) / 256.0;
                const double alpha = std::acos(2.0 * s - 1.0), beta = std::acos(2.0 * t - 1.0);
                const double l2 = 2.0 * t - 1.0;
                const double fd = std::sqrt(1.0 - l2 * l2) * (std::sin(alpha) + (pi - alpha) * std::cos(alpha)) * 0.25;
                const double fs = 3.5 * mallo_spec_term(alpha, beta, (double)m->spec_exp);
                for (int ch = 0; ch < 3; ++ch) {
                    mallo_diff[3 * i + ch] = half_round((float)clamp01d(fd * m->diffuse[ch] * m->light_color[ch]));
                    mallo_spec[3 * i + ch] = half_round((float)clamp01d(fs * m->specular[ch] * m->light_color[ch]));
                }
            }
        return VV_OK;
    });
}

vv_status vv_set_line_material(vv_ctx* c, const vv_line_material* m)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!m) return fail(c, VV_EINVAL, "material is NULL");
        std::vector<float> probe(8 * 256 * 256);  // validates the material
        vv_status s = vv_make_illum_tables(m, probe.data(), probe.data() + 2 * 65536, probe.data() + 5 * 65536);
        if (s) return fail(c, s, "%s", g_create_error.c_str());
        c->line_mat = *m;
        c->illum_dirty = true;
        ++c->scene_version;
        return VV_OK;
    });
}

vv_status vv_get_line_material(const vv_ctx* c, vv_line_material* m)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!c || !m) return VV_EINVAL;
        *m = c->line_mat;
        return VV_OK;
    });
}

}  // extern "C"
