// vv_ctx.hpp — the context object behind the C ABI and the host helpers
// shared by vv_api.cpp (single volume / view) and vv_batch.cpp (batches).
#pragma once

#include "../../include/vvlic.h"
#include "vv_internal.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

using namespace vv;

namespace vv_host {

inline thread_local std::string g_create_error;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t ensure(size_t n)
    {
        if (n <= bytes && p) return hipSuccess;
        release();
        hipError_t e = hipMalloc(&p, n ? n : 1);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        bytes = n;
        return hipSuccess;
    }
};

constexpr int kMaxSteps = 4096;
constexpr long long kMaxVoxels = 1ll << 31;

}  // namespace vv_host

using vv_host::DevBuf;

struct vv_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // field (VolumeData)
    bool has_field = false;
    int fdims[3] = {0, 0, 0};
    float spacing[3] = {1, 1, 1};
    float extent[3] = {1, 1, 1}, scale[3] = {1, 1, 1}, center[3] = {0.5f, 0.5f, 0.5f};
    int storage = VV_STORE_FP32;
    DevBuf field;

    // noise
    bool has_noise = false;
    int ndims[3] = {0, 0, 0};
    int nch = 1;
    DevBuf noise;

    // LIC copy of the field with the noise byte in channel 3, for a scalar
    // noise volume on the field's grid (launch_pack_lic_field); rebuilt after
    // either input changes
    DevBuf lic_field;
    bool lic_field_ok = false;

    // scalar gate volume
    bool has_scalar = false;
    int sdims[3] = {0, 0, 0};
    DevBuf scalar;

    // filter
    std::vector<uint8_t> lut;
    float inv_area = 0.5f;

    // transfer function: 256*4 rgba + 256 opacity
    float tf[1280];
    DevBuf tf_dev;
    bool tf_dirty = true;

    vv_lic_params params;

    // LIC volume (REPEAT layout)
    bool has_lic = false;
    int ldims[3] = {0, 0, 0};
    int lch = 1;
    DevBuf lic;
    DevBuf weights;
    std::vector<float> weights_host;  // what `weights` holds

    // launch tuning (vv_set_tuning)
    int lic_tile[3] = {16, 4, 4};
    int lic_super[3] = {2, 16, 16};
    int lic_cache = 1;
    int lic_brick = 0;  // FP32: the LIC volume kernel reads an 8^3-brick copy (lic.brick)
    bool lic_field_brick = false;  // lic_field holds the brick copy after the linear one
    int adj_window = 1;
    // Version of what a frame depends on (field, LIC volume, opacity scale,
    // transfer function, LIC parameters, line material); every change bumps it.
    unsigned long long scene_version = 0;
    // The forward pass vv_render_record left in adj_ckpt for the adjoint of
    // the same camera and scene version (the compact layout's adjoint uses
    // it up; the full layout's keeps it).
    struct {
        bool valid = false;
        vv_camera cam;
        unsigned long long version = 0;
        int max_ckpt = 0;
        int full_col = 0;
        void* buffer = nullptr;
    } adj_rec;
    int render_block[2] = {16, 16};
    int ray_group = 0;  // lanes per ray of the on-the-fly casters, 0: automatic
    int volume_group = 0;  // lanes per ray of the LIC-volume caster, 0: automatic

    // line-illumination tables (built on first use of illum_mode 2 / 3)
    vv_line_material line_mat;
    bool illum_dirty = true;
    DevBuf illum_luts;  // zoeckler (2 ch), mallo diffuse (3 ch), mallo specular (3 ch)

    // opacity scale (CLAMP layout, LIC grid)
    bool has_sigma = false;
    DevBuf sigma;

    // staging / scratch
    DevBuf stage, stage2, image, counter, adj_ckpt, adj_dimg, adj_out;
    void* noalloc_scratch = nullptr;  // this device's scratch array (vv_render_adjoint_noalloc)
};

namespace vv_host {

inline vv_status fail(vv_ctx* c, vv_status s, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c)
        c->err = buf;
    else
        g_create_error = buf;
    return s;
}

#define HIPCHK(c, expr)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail((c), e_ == hipErrorOutOfMemory ? VV_ENOMEM : VV_EDEVICE, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                     \
    } while (0)

// The C ABI lets no C++ exception through: a host allocation sized by an
// input (a file's contents) that fails reports VV_ENOMEM instead.
template <typename F>
inline vv_status guarded(vv_ctx* c, F&& f)
{
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(c, VV_ENOMEM, "out of host memory");
    } catch (...) {
        return fail(c, VV_EDEVICE, "unexpected host error");
    }
}

inline bool storage_ok(int storage)
{
    return storage == VV_STORE_FP32 || storage == VV_STORE_FP16 || storage == VV_STORE_UNORM8;
}

// volume sizes dims_ok accepts (error messages)
constexpr const char* kDimsRule = "must be in [2,4096] per axis, under 2^29 voxels in total";

inline bool dims_ok(const int d[3])
{
    if (!d) return false;
    long long n = 1;
    for (int a = 0; a < 3; ++a) {
        if (d[a] < 2 || d[a] > 4096) return false;
        n *= d[a];
    }
    return n < kMaxVoxels / 4;
}

inline long long nvox(const int d[3]) { return (long long)d[0] * d[1] * d[2]; }

// copy n bytes from host or device memory into a device buffer
inline vv_status upload(vv_ctx* c, DevBuf& b, const void* src, size_t n, int on_device)
{
    HIPCHK(c, b.ensure(n));
    HIPCHK(c, hipMemcpyAsync(b.p, src, n, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                             c->stream));
    return VV_OK;
}

inline void default_params(vv_lic_params* p)
{
    p->step_size_vol = 1.0f / 128.0f;
    p->gradient_scale = 30.0f;
    p->illum_scale = 1.0f;
    p->freq_scale = 1.0f;
    p->num_iterations = 255;
    p->steps_forward = 32;
    p->steps_backward = 32;
    p->step_size_lic = 0.01f;
    p->eye_factor = 0.3f;
    p->illum_mode = VV_ILLUM_LIC;
    p->tf_index = VV_TF_INDEX_B;
    p->gate_enable = 0;
    p->gate_lo = 0.1f;
    p->gate_hi = 0.3f;
    p->speed_of_flow = 0;
}

inline void default_tf5(uint8_t* t)
{
    // TransferEdit::TransferEdit (VV/transferEdit.cpp:76-82)
    for (int i = 0; i < 256; ++i) {
        for (int ch = 0; ch < 3; ++ch) t[5 * i + ch] = (uint8_t)i;
        for (int ch = 3; ch < 5; ++ch) t[5 * i + ch] = (uint8_t)(i < 20 ? 0 : i - 20);
    }
}

inline void tf5_to_float(const uint8_t* t5, float* tf)
{
    // RGBA texture = channels 0..3, LA texture alpha = channel 4
    // (VV/transferEdit.cpp:507-531); UNORM8 -> c/255
    for (int i = 0; i < 256; ++i) {
        for (int ch = 0; ch < 4; ++ch) tf[4 * i + ch] = (float)t5[5 * i + ch] / 255.0f;
        tf[1024 + i] = (float)t5[5 * i + 4] / 255.0f;
    }
}

// k(u): 1D linear CLAMP_TO_EDGE over the byte LUT, scaled by 1/255
inline float lut_k(const std::vector<uint8_t>& lut, float u)
{
    const int n = (int)lut.size();
    float x = std::fmaf(u, (float)n, -0.5f);
    float fl = std::floor(x);
    float f = x - fl;
    if (!(fl > -1.0f)) fl = -1.0f;
    if (fl > (float)(n - 1)) fl = (float)(n - 1);
    int i = (int)fl;
    int i0 = i < 0 ? 0 : i, i1 = i + 1 > n - 1 ? n - 1 : i + 1;
    float a = (float)lut[i0], b = (float)lut[i1];
    return std::fmaf(f, b - a, a) * (float)(1.0 / 255.0);
}

// convolution weights in sample order (inc_lic.glsl:168-199): k(0.5), then
// B backward samples (u -= 0.5/B), then F forward samples (u += 0.5/F)
inline void lic_weights(const vv_ctx* c, std::vector<float>& w)
{
    const int F = c->params.steps_forward, B = c->params.steps_backward;
    const float kf = 0.5f / (float)F, kb = 0.5f / (float)B;
    w.resize(1 + F + B);
    float u = 0.5f;
    w[0] = lut_k(c->lut, u);
    for (int i = 0; i < B; ++i) {
        u = u - kb;
        w[1 + i] = lut_k(c->lut, u);
    }
    u = 0.5f;
    for (int i = 0; i < F; ++i) {
        u = u + kf;
        w[1 + B + i] = lut_k(c->lut, u);
    }
}

inline void quat_matrix(const float qf[4], double R[3][3])
{
    double x = qf[0], y = qf[1], z = qf[2], w = qf[3];
    double n = std::sqrt(x * x + y * y + z * z + w * w);
    if (n > 0) {
        x /= n;
        y /= n;
        z /= n;
        w /= n;
    }
    R[0][0] = 1 - 2 * (y * y + z * z);
    R[0][1] = 2 * (x * y - z * w);
    R[0][2] = 2 * (x * z + y * w);
    R[1][0] = 2 * (x * y + z * w);
    R[1][1] = 1 - 2 * (x * x + z * z);
    R[1][2] = 2 * (y * z - x * w);
    R[2][0] = 2 * (x * z - y * w);
    R[2][1] = 2 * (y * z + x * w);
    R[2][2] = 1 - 2 * (x * x + y * y);
}

// Camera::setCamera + glTranslatef(-center): MV = T(pos) T(0,0,-dist) R(q) T(-center).
// Object-space camera = MV^-1 (0,0,0,1); pixel ray through NDC (x,y) is
// R^T (x*aspect*tan(fovy/2), y*tan(fovy/2), -1).  Computed in double, rounded once.
inline ViewBlock view_block_geom(const float extent[3], const float scale[3], const float center[3],
                                 const vv_camera* cam);

inline ViewBlock view_block(const vv_ctx* c, const vv_camera* cam)
{
    return view_block_geom(c->extent, c->scale, c->center, cam);
}

inline ViewBlock view_block_geom(const float extent[3], const float scale[3], const float center[3],
                                 const vv_camera* cam)
{
    ViewBlock v;
    double R[3][3];
    quat_matrix(cam->quat, R);
    const double pi = 3.14159265358979323846;
    const double th = std::tan(0.5 * (double)cam->fovy_deg * pi / 180.0);
    const double aspect = (double)cam->width / (double)cam->height;
    const double e[3] = {-(double)cam->pos[0], -(double)cam->pos[1], (double)cam->dist - (double)cam->pos[2]};
    for (int i = 0; i < 3; ++i) {
        const double rt_e = R[0][i] * e[0] + R[1][i] * e[1] + R[2][i] * e[2];
        v.cam[i] = (float)((double)center[i] + rt_e);
        v.du[i] = (float)(R[0][i] * aspect * th);
        v.dv[i] = (float)(R[1][i] * th);
        v.dw[i] = (float)(-R[2][i]);
        v.extent[i] = extent[i];
        v.scale[i] = scale[i];
        v.tex_max[i] = (float)((double)extent[i] * (double)scale[i]);
    }
    v.inv_w2 = (float)(2.0 / cam->width);
    v.inv_h2 = (float)(2.0 / cam->height);
    v.width = cam->width;
    v.height = cam->height;
    return v;
}

inline vv_status check_camera(vv_ctx* c, const vv_camera* cam)
{
    if (!cam) return fail(c, VV_EINVAL, "camera is NULL");
    if (cam->width < 1 || cam->height < 1 || cam->width > 16384 || cam->height > 16384)
        return fail(c, VV_EINVAL, "image size %dx%d out of range", cam->width, cam->height);
    if (!(cam->fovy_deg > 0.0f && cam->fovy_deg < 180.0f))
        return fail(c, VV_EINVAL, "fovy %g out of range", (double)cam->fovy_deg);
    if (!std::isfinite(cam->dist)) return fail(c, VV_EINVAL, "camera distance not finite");
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(cam->pos[a]) || !std::isfinite(cam->light_pos[a]))
            return fail(c, VV_EINVAL, "camera/light position not finite");
    double qn = 0.0;
    for (int a = 0; a < 4; ++a) {
        if (!std::isfinite(cam->quat[a])) return fail(c, VV_EINVAL, "camera orientation not finite");
        qn += (double)cam->quat[a] * cam->quat[a];
    }
    if (!(qn > 0.0)) return fail(c, VV_EINVAL, "camera orientation is the zero quaternion");
    if (!std::isfinite(cam->light_diffuse) || !std::isfinite(cam->light_specular) ||
        !std::isfinite(cam->light_ambient))
        return fail(c, VV_EINVAL, "light intensities not finite");
    if (cam->spec_exponent < 0 || cam->spec_exponent > 1024)
        return fail(c, VV_EINVAL, "spec exponent out of range");
    return VV_OK;
}

inline vv_status check_params(vv_ctx* c, const vv_lic_params* p)
{
    if (!p) return fail(c, VV_EINVAL, "params NULL");
    if (p->steps_forward < 1 || p->steps_forward > kMaxSteps || p->steps_backward < 1 ||
        p->steps_backward > kMaxSteps)
        return fail(c, VV_EINVAL, "steps forward/backward must be in [1,%d]", kMaxSteps);
    if (!(p->step_size_vol > 0.0f) || !std::isfinite(p->step_size_vol))
        return fail(c, VV_EINVAL, "step_size_vol must be > 0");
    if (p->num_iterations < 1 || p->num_iterations > 4096)
        return fail(c, VV_EINVAL, "num_iterations out of range");
    const float fs[] = {p->gradient_scale, p->illum_scale, p->freq_scale, p->step_size_lic,
                        p->eye_factor, p->gate_lo, p->gate_hi};
    for (float f : fs)
        if (!std::isfinite(f)) return fail(c, VV_EINVAL, "non-finite parameter");
    if (p->illum_mode < VV_ILLUM_LIC || p->illum_mode > VV_ILLUM_ZOECKLER)
        return fail(c, VV_EINVAL, "illum_mode invalid");
    if (p->tf_index != VV_TF_INDEX_B && p->tf_index != VV_TF_INDEX_MAG)
        return fail(c, VV_EINVAL, "tf_index invalid");
    if (p->speed_of_flow != 0 && p->speed_of_flow != 1) return fail(c, VV_EINVAL, "speed_of_flow must be 0 or 1");
    return VV_OK;
}

inline vv_status ensure_tf(vv_ctx* c)
{
    if (!c->tf_dirty) return VV_OK;
    HIPCHK(c, c->tf_dev.ensure(sizeof(c->tf)));
    HIPCHK(c, hipMemcpyAsync(c->tf_dev.p, c->tf, sizeof(c->tf), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));  // host table may change after return
    c->tf_dirty = false;
    return VV_OK;
}

inline vv_status ensure_weights(vv_ctx* c)
{
    std::vector<float> w;
    lic_weights(c, w);
    if (w != c->weights_host) {
        HIPCHK(c, c->weights.ensure(sizeof(float) * w.size()));
        HIPCHK(c, hipMemcpyAsync(c->weights.p, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));  // w is a local
        c->weights_host = w;
    }
    return VV_OK;
}

// the line-illumination tables of c->line_mat, on the device
vv_status ensure_illum_luts(vv_ctx* c);

inline void illum_lut_ptrs(vv_ctx* c, RenderLaunch& R)
{
    const float* base = (const float*)c->illum_luts.p;
    R.lut_zoeckler = base;
    R.lut_mallo_diff = base + 256 * 256 * 2;
    R.lut_mallo_spec = base + 256 * 256 * 5;
}

// need_lic: the LIC-volume ray caster; without it (the fused caster) the
// shading channels are those of the noise
inline vv_status fill_render_launch(vv_ctx* c, const vv_camera* cam, RenderLaunch& R, bool need_lic = true)
{
    if (!c->has_field) return fail(c, VV_ESTATE, "no vector field uploaded");
    if (need_lic && !c->has_lic) return fail(c, VV_ESTATE, "no LIC volume (call vv_compute_lic or vv_upload_lic)");
    vv_status st = check_camera(c, cam);
    if (st) return st;
    st = ensure_tf(c);
    if (st) return st;
    const vv_lic_params& p = c->params;
    std::memset(&R, 0, sizeof(R));
    R.field = make_clamp_desc(c->field.p, c->fdims);
    R.field_type = c->storage;
    R.tf_ch = p.tf_index == VV_TF_INDEX_MAG ? 3 : 2;
    if (need_lic) {
        R.lic = make_repeat_desc(c->lic.p, c->ldims);
        R.lc = c->lch;
    } else {
        R.lc = c->nch;
    }
    if (c->has_sigma) R.sigma = make_clamp_desc(c->sigma.p, c->ldims);
    R.tf = (const float*)c->tf_dev.p;
    R.view = view_block(c, cam);
    R.step = p.step_size_vol;
    // alphaCorrection = stepSizeVol * 128 (VV/renderer.cpp:988-989)
    R.alpha_corr = p.step_size_vol * 128.0f;
    R.corr_mode = R.alpha_corr == 1.0f ? 0 : (R.alpha_corr == 0.5f ? 1 : (R.alpha_corr == 2.0f ? 2 : 3));
    R.max_samples = (long long)p.num_iterations * (long long)p.num_iterations;
    R.illum = p.illum_mode;
    if (p.illum_mode >= VV_ILLUM_MALLO) {
        st = ensure_illum_luts(c);
        if (st) return st;
        illum_lut_ptrs(c, R);
    }
    R.illum_scale = p.illum_scale;
    // L = normalize(gl_LightSource[0].position.xyz) (scaleVolInv is never uploaded)
    const float l0 = cam->light_pos[0], l1 = cam->light_pos[1], l2 = cam->light_pos[2];
    const float ld = std::sqrt(l0 * l0 + l1 * l1 + l2 * l2);
    if (ld > 0.0f) {
        R.light[0] = l0 / ld;
        R.light[1] = l1 / ld;
        R.light[2] = l2 / ld;
    }
    R.light_diffuse = cam->light_diffuse;
    R.light_specular = cam->light_specular;
    R.light_ambient = cam->light_ambient;
    R.spec_exp = cam->spec_exponent;
    R.block[0] = c->render_block[0];
    R.block[1] = c->render_block[1];
    const int nt = R.block[0] * R.block[1];
    if (nt > 256 || nt % 64)
        return fail(c, VV_EINVAL, "render block %dx%d must hold 64, 128, 192 or 256 threads", R.block[0],
                    R.block[1]);
    return VV_OK;
}

}  // namespace vv_host
