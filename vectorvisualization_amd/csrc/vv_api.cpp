// vv_api.cpp — the C ABI declared in include/vvlic.h.
//
// Host-side responsibilities only: argument validation, device-buffer
// ownership, the per-frame view block, the convolution-weight table and the
// uniform values of Renderer::setRenderVolParams (VV/renderer.cpp:925-996).
// Every volume operation runs in the HIP kernels of vv_lic.hip and
// vv_render.hip; there is no host compute path.
#include "vv_ctx.hpp"

using namespace vv_host;

extern "C" {

int vv_abi_version(void) { return VVLIC_ABI_VERSION; }

const char* vv_last_error(const vv_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

vv_status vv_create(vv_ctx** out, int device)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!out) return fail(nullptr, VV_EINVAL, "out is NULL");
        *out = nullptr;
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n == 0) return fail(nullptr, VV_EDEVICE, "no HIP device available");
        if (device < 0 || device >= n) return fail(nullptr, VV_EINVAL, "device %d out of range", device);
        e = hipSetDevice(device);
        if (e != hipSuccess) return fail(nullptr, VV_EDEVICE, "hipSetDevice: %s", hipGetErrorString(e));
        vv_ctx* c = new (std::nothrow) vv_ctx();
        if (!c) return fail(nullptr, VV_ENOMEM, "out of host memory");
        c->device = device;
        e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            return fail(nullptr, VV_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
        }
        // the allocation-free adjoint's scratch on this device, and the TF
        // table's device buffer, so that entry point never allocates
        e = noalloc_scratch(device, &c->noalloc_scratch);
        if (e == hipSuccess) e = c->tf_dev.ensure(sizeof(c->tf));
        if (e != hipSuccess) {
            c->tf_dev.release();
            (void)hipStreamDestroy(c->stream);
            delete c;
            return fail(nullptr, VV_EDEVICE, "device set-up: %s", hipGetErrorString(e));
        }
        default_params(&c->params);
        vv_default_line_material(&c->line_mat);
        uint8_t t5[256 * 5];
        default_tf5(t5);
        tf5_to_float(t5, c->tf);
        // default filter: box (VV/3DLIC.cpp:725-729 fallback)
        c->lut.assign(256, 255);
        c->inv_area = 0.5f;
        *out = c;
        return VV_OK;
    });
}

vv_status vv_destroy(vv_ctx* c)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_OK;
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        DevBuf* bufs[] = {&c->field, &c->noise, &c->scalar, &c->tf_dev, &c->lic, &c->weights,
                          &c->sigma, &c->stage, &c->stage2, &c->image, &c->counter, &c->adj_ckpt,
                          &c->adj_dimg, &c->adj_out, &c->lic_field, &c->illum_luts};
        for (DevBuf* b : bufs) b->release();
        (void)hipStreamDestroy(c->stream);
        delete c;
        return VV_OK;
    });
}

void* vv_stream(vv_ctx* c) { return c ? (void*)c->stream : nullptr; }

vv_status vv_synchronize(vv_ctx* c)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return VV_OK;
    });
}

vv_status vv_default_lic_params(vv_lic_params* p)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!p) return VV_EINVAL;
        default_params(p);
        return VV_OK;
    });
}

vv_status vv_default_camera(vv_camera* cam, int width, int height, const float center[3])
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!cam || width < 1 || height < 1) return VV_EINVAL;
        std::memset(cam, 0, sizeof(*cam));
        cam->width = width;
        cam->height = height;
        cam->fovy_deg = 35.0f;
        cam->dist = 4.0f;
        cam->quat[3] = 1.0f;
        const float c0 = center ? center[0] : 0.5f, c1 = center ? center[1] : 0.5f,
                    c2 = center ? center[2] : 0.5f;
        // updateLightPos with identity rotations: center + (0,0,light distance 1)
        cam->light_pos[0] = c0;
        cam->light_pos[1] = c1;
        cam->light_pos[2] = c2 + 1.0f;
        cam->light_diffuse = 1.0f;
        cam->light_specular = 1.0f;
        cam->light_ambient = 0.0f;
        cam->spec_exponent = 40;
        return VV_OK;
    });
}

// ---- field -------------------------------------------------------------------
vv_status vv_upload_field(vv_ctx* c, const float* v0, const float* v1, float lerp_t, const int dims[3],
                          const float spacing[3], int storage, int on_device)
{
    return guarded(c, [&]() -> vv_status {
        return vv_upload_field_ex(c, v0, v1, lerp_t, dims, spacing, storage, VV_ENCODE_INTERP, on_device);
    });
}

vv_status vv_upload_field_ex(vv_ctx* c, const float* v0, const float* v1, float lerp_t, const int dims[3],
                             const float spacing[3], int storage, int encoding, int on_device)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (encoding != VV_ENCODE_INTERP && encoding != VV_ENCODE_RAW_MAGNITUDE)
            return fail(c, VV_EINVAL, "unknown field encoding %d", encoding);
        if (encoding == VV_ENCODE_RAW_MAGNITUDE && storage == VV_STORE_UNORM8)
            return fail(c, VV_EINVAL, "the raw-magnitude encoding has float texels only (FP32, FP16)");
        if (!v0) return fail(c, VV_EINVAL, "field data is NULL");
        if (!dims_ok(dims)) return fail(c, VV_EINVAL, "field dims %s", kDimsRule);
        if (!storage_ok(storage)) return fail(c, VV_EINVAL, "unknown storage %d", storage);
        if (!std::isfinite(lerp_t)) return fail(c, VV_EINVAL, "lerp_t not finite");
        float sd[3] = {1, 1, 1};
        if (spacing)
            for (int a = 0; a < 3; ++a) {
                if (!(spacing[a] > 0.0f) || !std::isfinite(spacing[a]))
                    return fail(c, VV_EINVAL, "spacing must be > 0");
                sd[a] = spacing[a];
            }
        HIPCHK(c, hipSetDevice(c->device));
        const long long n = nvox(dims);
        const size_t raw = sizeof(float) * 3 * (size_t)n;
        const float* d0 = v0;
        const float* d1 = v1 ? v1 : v0;
        if (!on_device) {
            for (long long i = 0; i < 3 * n; ++i)
                if (!std::isfinite(v0[i]) || (v1 && !std::isfinite(v1[i])))
                    return fail(c, VV_EINVAL, "field contains a non-finite value at %lld", i / 3);
            vv_status st = upload(c, c->stage, v0, raw, 0);
            if (st) return st;
            d0 = (const float*)c->stage.p;
            d1 = d0;
            if (v1) {
                st = upload(c, c->stage2, v1, raw, 0);
                if (st) return st;
                d1 = (const float*)c->stage2.p;
            }
        }
        HIPCHK(c, c->field.ensure((size_t)field_texel_bytes(storage) * (size_t)clamp_elems(dims)));
        HIPCHK(c, c->counter.ensure(64));
        uint32_t* maxbits = (uint32_t*)c->counter.p;
        HIPCHK(c, launch_fill_u32(maxbits, 0u, 1, c->stream));
        HIPCHK(c, launch_field_maxlen(d0, d1, lerp_t, n, maxbits, c->stream));
        HIPCHK(c, launch_field_encode(d0, d1, lerp_t, dims, maxbits, c->field.p, storage, c->stream,
                                      encoding == VV_ENCODE_RAW_MAGNITUDE));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        // VectorDataSet::loadData geometry (VV/dataset.cpp:144-176)
        float vol[3], maxvol = 0.0f;
        int maxtex = 0;
        for (int a = 0; a < 3; ++a) {
            c->fdims[a] = dims[a];
            c->spacing[a] = sd[a];
            vol[a] = (float)dims[a] * sd[a];
            if (vol[a] > maxvol) maxvol = vol[a];
            if (dims[a] > maxtex) maxtex = dims[a];
        }
        for (int a = 0; a < 3; ++a) {
            c->scale[a] = (float)maxtex / ((float)dims[a] * sd[a]);
            c->extent[a] = (float)dims[a] * sd[a] / maxvol;
            c->center[a] = c->extent[a] / 2.0f;
        }
        c->storage = storage;
        c->has_field = true;
        c->lic_field_ok = false;
        ++c->scene_version;
        return VV_OK;
    });
}

vv_status vv_field_geometry(const vv_ctx* c, float extent[3], float scale[3], float center[3])
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!c || !c->has_field) return VV_ESTATE;
        for (int a = 0; a < 3; ++a) {
            if (extent) extent[a] = c->extent[a];
            if (scale) scale[a] = c->scale[a];
            if (center) center[a] = c->center[a];
        }
        return VV_OK;
    });
}

vv_status vv_download_field(vv_ctx* c, float* out4, int on_device)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!c->has_field) return fail(c, VV_ESTATE, "no field");
        if (!out4) return fail(c, VV_EINVAL, "out is NULL");
        const size_t bytes = sizeof(float) * 4 * (size_t)nvox(c->fdims);
        float* dst = out4;
        if (!on_device) {
            HIPCHK(c, c->stage.ensure(bytes));
            dst = (float*)c->stage.p;
        }
        HIPCHK(c, launch_field_decode(c->field.p, c->storage, c->fdims, dst, c->stream));
        if (!on_device) HIPCHK(c, hipMemcpyAsync(out4, dst, bytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return VV_OK;
    });
}

// ---- noise -------------------------------------------------------------------
namespace {

vv_status build_noise_tex(vv_ctx* c, const uint8_t* d_noise, const int dims[3], int with_gradients,
                          const uint8_t* grad_rgb, int grad_on_device)
{
    const long long n = nvox(dims);
    const int nc = with_gradients ? 4 : 1;
    const uint8_t* d_rgb = nullptr;
    if (with_gradients) {
        if (grad_rgb) {
            vv_status st = upload(c, c->stage2, grad_rgb, 3 * (size_t)n, grad_on_device);
            if (st) return st;
            d_rgb = (const uint8_t*)c->stage2.p;
        } else {
            // computeGradients -> filterGradients -> quantizeGradients
            // (VV/dataset.cpp:1241-1252) on the GPU
            HIPCHK(c, c->stage2.ensure(sizeof(float) * 3 * (size_t)n + 3 * (size_t)n));
            float* g = (float*)c->stage2.p;
            uint8_t* rgb = (uint8_t*)(g + 3 * n);
            const float sd[3] = {1.0f, 1.0f, 1.0f};  // noise VolumeData sliceDist default
            HIPCHK(c, launch_noise_grad_raw(d_noise, dims, sd, g, c->stream));
            // filterGradients kernel weights: exp(-(i^2+j^2+k^2)/SIGMA2) over
            // i,j,k in [-2,0], normalised (VV/gradient.cpp:394-409)
            float w[27], sum = 0.0f;
            for (int k = -2; k < 1; ++k)
                for (int j = -2; j < 1; ++j)
                    for (int i = -2; i < 1; ++i) {
                        float v = std::exp(-(float)(i * i + j * j + k * k) / 5.0f);
                        w[((k + 2) * 3 + (j + 2)) * 3 + (i + 2)] = v;
                        sum += v;
                    }
            for (float& v : w) v /= sum;
            HIPCHK(c, launch_noise_grad_filter(g, dims, w, rgb, c->stream));
            d_rgb = rgb;
        }
    }
    HIPCHK(c, c->noise.ensure((size_t)nc * (size_t)repeat_elems(dims)));
    HIPCHK(c, launch_noise_tex(d_noise, d_rgb, dims, c->noise.p, nc, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int a = 0; a < 3; ++a) c->ndims[a] = dims[a];
    c->nch = nc;
    c->has_noise = true;
    c->lic_field_ok = false;
    return VV_OK;
}

}  // namespace

vv_status vv_upload_noise(vv_ctx* c, const uint8_t* noise, const int dims[3], int with_gradients,
                          const uint8_t* grad_rgb, int on_device)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!noise) return fail(c, VV_EINVAL, "noise is NULL");
        if (!dims_ok(dims)) return fail(c, VV_EINVAL, "noise dims %s", kDimsRule);
        HIPCHK(c, hipSetDevice(c->device));
        vv_status st = upload(c, c->stage, noise, (size_t)nvox(dims), on_device);
        if (st) return st;
        return build_noise_tex(c, (const uint8_t*)c->stage.p, dims, with_gradients, grad_rgb, on_device);
    });
}

vv_status vv_generate_noise(vv_ctx* c, uint64_t seed, int kind, double density, const int dims[3],
                            int with_gradients)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!dims_ok(dims)) return fail(c, VV_EINVAL, "noise dims %s", kDimsRule);
        if (kind != VV_NOISE_BINARY && kind != VV_NOISE_UNIFORM) return fail(c, VV_EINVAL, "noise kind");
        if (kind == VV_NOISE_BINARY && !(density >= 0.0 && density <= 1.0))
            return fail(c, VV_EINVAL, "density must be in [0,1]");
        HIPCHK(c, hipSetDevice(c->device));
        double t = std::floor((kind == VV_NOISE_BINARY ? density : 0.0) * 4294967296.0);
        if (t < 0) t = 0;
        if (t > 4294967296.0) t = 4294967296.0;
        HIPCHK(c, c->stage.ensure((size_t)nvox(dims)));
        HIPCHK(c, launch_noise_gen(seed, kind, (uint64_t)t, dims, (uint8_t*)c->stage.p, c->stream));
        return build_noise_tex(c, (const uint8_t*)c->stage.p, dims, with_gradients, nullptr, 0);
    });
}

vv_status vv_download_noise(vv_ctx* c, uint8_t* out, int on_device)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!c->has_noise) return fail(c, VV_ESTATE, "no noise");
        if (!out) return fail(c, VV_EINVAL, "out is NULL");
        const size_t bytes = (size_t)c->nch * (size_t)nvox(c->ndims);
        uint8_t* dst = out;
        if (!on_device) {
            HIPCHK(c, c->stage.ensure(bytes));
            dst = (uint8_t*)c->stage.p;
        }
        HIPCHK(c, launch_noise_untex(c->noise.p, c->nch, c->ndims, dst, c->stream));
        if (!on_device) HIPCHK(c, hipMemcpyAsync(out, dst, bytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return VV_OK;
    });
}

vv_status vv_noise_info(const vv_ctx* c, int dims[3], int* channels)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!c || !c->has_noise) return VV_ESTATE;
        if (dims)
            for (int a = 0; a < 3; ++a) dims[a] = c->ndims[a];
        if (channels) *channels = c->nch;
        return VV_OK;
    });
}

// ---- filter ------------------------------------------------------------------
vv_status vv_make_filter(int kind, int length, float sigma, uint8_t lut[256], int* width, float* inv_area)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!lut || !width || !inv_area) return VV_EINVAL;
        const int W = 256;
        if (kind == VV_FILTER_BOX) {
            // LICFilter::createBoxFilter (VV/dataset.cpp:1405-1413)
            std::memset(lut, 255, W);
            *width = W;
            *inv_area = 0.5f;
            return VV_OK;
        }
        if (kind < VV_FILTER_TRIANGLE || kind > VV_FILTER_GAUSSIAN || length <= 0) return VV_EINVAL;
        double sg = sigma > 0.0f ? (double)sigma : length / 4.0;
        const double pi = 3.14159265358979323846;
        for (int i = 0; i < W; ++i) {
            const double t = (i + 0.5) / W;
            const double cc = 2.0 * t - 1.0;
            double w;
            if (kind == VV_FILTER_TRIANGLE) {
                w = 1.0 - std::fabs(cc);
            } else if (kind == VV_FILTER_COS2) {
                const double q = std::cos(pi * (t - 0.5));
                w = q * q;
            } else {
                const double s = cc * length;
                w = std::exp(-s * s / (2.0 * sg * sg));
            }
            double b = std::floor(255.0 * w + 0.5);
            lut[i] = (uint8_t)(b < 0 ? 0 : (b > 255 ? 255 : b));
        }
        *width = W;
        float area = 0.0f;
        for (int i = 0; i < W; ++i) area += lut[i];
        *inv_area = 0.5f * W * 255.0f / area;
        return VV_OK;
    });
}

vv_status vv_set_filter(vv_ctx* c, const uint8_t* lut, int width, float inv_area)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!lut) return fail(c, VV_EINVAL, "lut is NULL");
        if (width < 1 || width > 4096 || (width & (width - 1)))
            return fail(c, VV_EINVAL, "filter width must be a power of two <= 4096");
        float area = 0.0f;
        for (int i = 0; i < width; ++i) area += lut[i];
        if (!(area > 0.0f)) return fail(c, VV_EINVAL, "filter is all zero");
        if (!std::isfinite(inv_area)) return fail(c, VV_EINVAL, "inv_area not finite");
        c->lut.assign(lut, lut + width);
        // calcFilterKernelInvArea (VV/dataset.cpp:1503-1512)
        c->inv_area = inv_area > 0.0f ? inv_area : 0.5f * width * 255.0f / area;
        return VV_OK;
    });
}

// ---- transfer function -------------------------------------------------------
vv_status vv_set_tf(vv_ctx* c, const float* rgba, const float* opacity)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!rgba || !opacity) return fail(c, VV_EINVAL, "tf arrays NULL");
        for (int i = 0; i < 1024; ++i)
            if (!std::isfinite(rgba[i])) return fail(c, VV_EINVAL, "tf rgba not finite");
        for (int i = 0; i < 256; ++i)
            if (!std::isfinite(opacity[i])) return fail(c, VV_EINVAL, "tf opacity not finite");
        if (std::memcmp(c->tf, rgba, sizeof(float) * 1024) || std::memcmp(c->tf + 1024, opacity, sizeof(float) * 256))
            ++c->scene_version;
        std::memcpy(c->tf, rgba, sizeof(float) * 1024);
        std::memcpy(c->tf + 1024, opacity, sizeof(float) * 256);
        c->tf_dirty = true;
        return VV_OK;
    });
}

vv_status vv_set_tf_u8(vv_ctx* c, const uint8_t* tf5)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!tf5) return fail(c, VV_EINVAL, "tf NULL");
        float t[1280];
        tf5_to_float(tf5, t);
        if (std::memcmp(c->tf, t, sizeof(t))) ++c->scene_version;
        std::memcpy(c->tf, t, sizeof(t));
        c->tf_dirty = true;
        return VV_OK;
    });
}

vv_status vv_default_tf_u8(uint8_t tf5[256 * 5])
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!tf5) return VV_EINVAL;
        default_tf5(tf5);
        return VV_OK;
    });
}

// ---- scalar gate volume --------------------------------------------------------
vv_status vv_upload_scalar(vv_ctx* c, const uint8_t* scalar, const int dims[3], int on_device)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!scalar) return fail(c, VV_EINVAL, "scalar is NULL");
        if (!dims_ok(dims)) return fail(c, VV_EINVAL, "scalar dims %s", kDimsRule);
        HIPCHK(c, hipSetDevice(c->device));
        vv_status st = upload(c, c->stage, scalar, (size_t)nvox(dims), on_device);
        if (st) return st;
        HIPCHK(c, c->scalar.ensure((size_t)clamp_elems(dims)));
        HIPCHK(c, launch_pad_clamp_u8((const uint8_t*)c->stage.p, dims, (uint8_t*)c->scalar.p, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int a = 0; a < 3; ++a) c->sdims[a] = dims[a];
        c->has_scalar = true;
        return VV_OK;
    });
}

// ---- params ------------------------------------------------------------------
vv_status vv_set_lic_params(vv_ctx* c, const vv_lic_params* p)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        vv_status st = check_params(c, p);
        if (st) return st;
        c->params = *p;
        ++c->scene_version;
        return VV_OK;
    });
}

vv_status vv_get_lic_params(const vv_ctx* c, vv_lic_params* p)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!c || !p) return VV_EINVAL;
        *p = c->params;
        return VV_OK;
    });
}

vv_status vv_lic_weights(const vv_ctx* c, float* out)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!c || !out) return VV_EINVAL;
        std::vector<float> w;
        lic_weights(c, w);
        std::memcpy(out, w.data(), sizeof(float) * w.size());
        return VV_OK;
    });
}

// ---- LIC ---------------------------------------------------------------------
namespace {

// the LIC launch of the context's field, noise, filter and parameters
// (updateLICVolume + setRenderVolParams); out: the LIC volume, or none
vv_status fill_lic_launch(vv_ctx* c, LicLaunch& L, void* out)
{
    if (!c->has_field) return fail(c, VV_ESTATE, "no vector field uploaded");
    if (!c->has_noise) return fail(c, VV_ESTATE, "no noise uploaded");
    const vv_lic_params& p = c->params;
    if (p.gate_enable && c->nch == 1 && !c->has_scalar)
        return fail(c, VV_ESTATE, "gate enabled but no scalar volume uploaded");
    if (p.illum_mode == VV_ILLUM_GRADIENT && c->nch != 4)
        return fail(c, VV_ESTATE, "gradient illumination needs RGBA noise (with_gradients)");
    if (c->lic_tile[0] * c->lic_tile[1] * c->lic_tile[2] != 256)
        return fail(c, VV_EINVAL, "LIC tile %dx%dx%d must hold 256 threads", c->lic_tile[0],
                    c->lic_tile[1], c->lic_tile[2]);
    HIPCHK(c, hipSetDevice(c->device));
    vv_status wst = ensure_weights(c);
    if (wst) return wst;
    const int nc = c->nch;
    std::memset(&L, 0, sizeof(L));
    L.field = make_clamp_desc(c->field.p, c->fdims);
    L.field_type = c->storage;
    L.noise = make_repeat_desc(c->noise.p, c->ndims);
    L.nc = nc;
    if (c->has_scalar) L.scalar = make_clamp_desc(c->scalar.p, c->sdims);
    L.gate = p.gate_enable && nc == 1;
    L.gate_lo = p.gate_lo;
    L.gate_hi = p.gate_hi;
    L.w = (const float*)c->weights.p;
    L.F = p.steps_forward;
    L.B = p.steps_backward;
    // licParams.z * (logEyeDist*0.5 + 0.3), logEyeDist == 0 (inc_lic.glsl:114)
    L.he = p.step_size_lic * p.eye_factor;
    // licKernel.b * gradient.r = invArea/(F+B) * gradientScale (VV/renderer.cpp:982-984)
    L.scl = (c->inv_area / (float)(p.steps_forward + p.steps_backward)) * p.gradient_scale;
    L.freq = p.freq_scale;
    for (int a = 0; a < 3; ++a) L.scale[a] = c->scale[a];
    L.out = make_repeat_desc(out, c->fdims);
    for (int a = 0; a < 3; ++a) L.tile[a] = c->lic_tile[a];
    for (int a = 0; a < 3; ++a) L.super_tile[a] = c->lic_super[a];
    L.field_bytes = (unsigned long long)clamp_elems(c->fdims) * field_texel_bytes(L.field_type);
    L.noise_bytes = (unsigned long long)repeat_elems(c->ndims) * (unsigned long long)nc;
    // the cached kernels address through 32-bit byte offsets
    // and 24-bit index multiplies (plane strides below 2^24 elements); with
    // SPEED_OF_FLOW the LIC reads the field's magnitude channel, which the
    // cached kernels do not keep: the one-lane-per-voxel kernels run
    L.speed = p.speed_of_flow;
    L.cache = c->lic_cache && !L.speed && L.field_bytes < (1ull << 32) && L.noise_bytes < (1ull << 32) &&
              (long long)L.field.sz < (1ll << 24) && (long long)L.noise.sz < (1ll << 24);
    // scalar noise on the field's grid at freq 1 (the cached SAME kernels):
    // the LIC reads the field copy that carries the noise bytes, built once per
    // field/noise pair.  Without memory for the copy it reads the noise volume.
    if (L.cache && nc == 1 && !L.gate && L.freq == 1.0f && c->ndims[0] == c->fdims[0] &&
        c->ndims[1] == c->fdims[1] && c->ndims[2] == c->fdims[2]) {
        long long nb[3];
        brick_counts(c->fdims, nb);
        const size_t brick_bytes = (size_t)(nb[0] * nb[1] * nb[2] * 512) * 16;
        const bool brick = c->lic_brick && c->storage == VV_STORE_FP32 && brick_bytes < (1ull << 32) &&
                           nb[0] * nb[1] * 512 < (1ll << 24);
        if (c->lic_field_ok && c->lic_field_brick != brick) c->lic_field_ok = false;
        if (!c->lic_field_ok) {
            const hipError_t e = c->lic_field.ensure((size_t)L.field_bytes + (brick ? brick_bytes : 0));
            if (e == hipErrorOutOfMemory) {
                (void)hipGetLastError();
            } else {
                HIPCHK(c, e);
                HIPCHK(c, launch_pack_lic_field(c->field.p, c->storage, (const uint8_t*)c->noise.p, c->fdims,
                                                c->lic_field.p, c->stream));
                if (brick)
                    HIPCHK(c, launch_pack_lic_field_brick(c->lic_field.p, c->fdims,
                                                          (char*)c->lic_field.p + L.field_bytes, c->stream));
                c->lic_field_ok = true;
                c->lic_field_brick = brick;
            }
        }
        if (c->lic_field_ok) {
            L.field.p = c->lic_field.p;
            L.field_has_noise = 1;
            if (c->lic_field_brick) L.field_brick = (const char*)c->lic_field.p + L.field_bytes;
        }
    }
    return VV_OK;
}

}  // namespace

vv_status vv_compute_lic(vv_ctx* c)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!c->has_field) return fail(c, VV_ESTATE, "no vector field uploaded");
        LicLaunch L;
        vv_status st = fill_lic_launch(c, L, nullptr);
        if (st) return st;
        const int nc = c->nch;
        HIPCHK(c, c->lic.ensure(sizeof(float) * nc * (size_t)repeat_elems(c->fdims)));
        L.out = make_repeat_desc(c->lic.p, c->fdims);
        HIPCHK(c, launch_lic(L, c->stream));  // also writes the REPEAT halo planes
        // the opacity scale lives on the LIC grid: drop it when that grid changes
        if (c->ldims[0] != c->fdims[0] || c->ldims[1] != c->fdims[1] || c->ldims[2] != c->fdims[2])
            c->has_sigma = false;
        for (int a = 0; a < 3; ++a) c->ldims[a] = c->fdims[a];
        c->lch = nc;
        c->has_lic = true;
        ++c->scene_version;
        return VV_OK;
    });
}

vv_status vv_lic_info(const vv_ctx* c, int dims[3], int* channels)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!c || !c->has_lic) return VV_ESTATE;
        if (dims)
            for (int a = 0; a < 3; ++a) dims[a] = c->ldims[a];
        if (channels) *channels = c->lch;
        return VV_OK;
    });
}

vv_status vv_download_lic(vv_ctx* c, float* out, int on_device)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!c->has_lic) return fail(c, VV_ESTATE, "no LIC volume");
        if (!out) return fail(c, VV_EINVAL, "out is NULL");
        const size_t bytes = sizeof(float) * c->lch * (size_t)nvox(c->ldims);
        float* dst = out;
        if (!on_device) {
            HIPCHK(c, c->stage.ensure(bytes));
            dst = (float*)c->stage.p;
        }
        HIPCHK(c, launch_unpad_repeat_f32((const float*)c->lic.p, c->lch, c->ldims, dst, c->stream));
        if (!on_device) HIPCHK(c, hipMemcpyAsync(out, dst, bytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return VV_OK;
    });
}

vv_status vv_upload_lic(vv_ctx* c, const float* lic, int channels, const int dims[3], int on_device)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!lic) return fail(c, VV_EINVAL, "lic is NULL");
        if (channels != 1 && channels != 4) return fail(c, VV_EINVAL, "channels must be 1 or 4");
        if (!dims_ok(dims)) return fail(c, VV_EINVAL, "lic dims %s", kDimsRule);
        HIPCHK(c, hipSetDevice(c->device));
        const size_t bytes = sizeof(float) * channels * (size_t)nvox(dims);
        vv_status st = upload(c, c->stage, lic, bytes, on_device);
        if (st) return st;
        HIPCHK(c, c->lic.ensure(sizeof(float) * channels * (size_t)repeat_elems(dims)));
        HIPCHK(c, launch_pad_repeat_f32((const float*)c->stage.p, channels, dims, (float*)c->lic.p, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int a = 0; a < 3; ++a) c->ldims[a] = dims[a];
        c->lch = channels;
        c->has_lic = true;
        c->has_sigma = false;  // opacity scale lives on the LIC grid
        ++c->scene_version;
        return VV_OK;
    });
}

// ---- tuning ------------------------------------------------------------------
vv_status vv_set_tuning(vv_ctx* c, const char* key, int value)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!key) return fail(c, VV_EINVAL, "key is NULL");
        const std::string k(key);
        int* slot = nullptr;
        if (k == "lic.tile_x") slot = &c->lic_tile[0];
        else if (k == "lic.tile_y") slot = &c->lic_tile[1];
        else if (k == "lic.tile_z") slot = &c->lic_tile[2];
        else if (k == "lic.cache") slot = &c->lic_cache;
        else if (k == "lic.brick") slot = &c->lic_brick;
        else if (k == "lic.super_x") slot = &c->lic_super[0];
        else if (k == "lic.super_y") slot = &c->lic_super[1];
        else if (k == "lic.super_z") slot = &c->lic_super[2];
        else if (k == "adjoint.window") slot = &c->adj_window;
        else if (k == "render.block_x") slot = &c->render_block[0];
        else if (k == "render.block_y") slot = &c->render_block[1];
        else if (k == "render.ray_group") slot = &c->ray_group;
        else if (k == "render.volume_group") slot = &c->volume_group;
        else return fail(c, VV_EINVAL, "unknown tuning key '%s'", key);
        if (k == "lic.cache" || k == "lic.brick" || k == "adjoint.window") {
            if (value != 0 && value != 1) return fail(c, VV_EINVAL, "%s must be 0 or 1", key);
        } else if (k == "render.ray_group" || k == "render.volume_group") {
            if (value < 0 || value > 64 || (value & (value - 1)))
                return fail(c, VV_EINVAL, "%s must be 0 (automatic) or a power of two up to 64", key);
        } else if (value < 1 || value > 256 || (value & (value - 1))) {
            return fail(c, VV_EINVAL, "tile sizes must be powers of two in [1,256]");
        }
        *slot = value;  // render block pairs are checked when a frame is rendered
        return VV_OK;
    });
}

vv_status vv_get_tuning(const vv_ctx* c, const char* key, int* value)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!c || !key || !value) return VV_EINVAL;
        const std::string k(key);
        if (k == "lic.tile_x") *value = c->lic_tile[0];
        else if (k == "lic.tile_y") *value = c->lic_tile[1];
        else if (k == "lic.tile_z") *value = c->lic_tile[2];
        else if (k == "lic.cache") *value = c->lic_cache;
        else if (k == "lic.brick") *value = c->lic_brick;
        else if (k == "lic.super_x") *value = c->lic_super[0];
        else if (k == "lic.super_y") *value = c->lic_super[1];
        else if (k == "lic.super_z") *value = c->lic_super[2];
        else if (k == "adjoint.window") *value = c->adj_window;
        else if (k == "render.block_x") *value = c->render_block[0];
        else if (k == "render.block_y") *value = c->render_block[1];
        else if (k == "render.ray_group") *value = c->ray_group;
        else if (k == "render.volume_group") *value = c->volume_group;
        else return VV_EINVAL;
        return VV_OK;
    });
}

// ---- render ------------------------------------------------------------------
vv_status vv_view_block(const vv_ctx* c, const vv_camera* cam, float out24[24])
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!c || !cam || !out24) return VV_EINVAL;
        ViewBlock v = view_block(c, cam);
        float* o = out24;
        for (int a = 0; a < 3; ++a) o[a] = v.cam[a];
        for (int a = 0; a < 3; ++a) o[3 + a] = v.du[a];
        for (int a = 0; a < 3; ++a) o[6 + a] = v.dv[a];
        for (int a = 0; a < 3; ++a) o[9 + a] = v.dw[a];
        o[12] = v.inv_w2;
        o[13] = v.inv_h2;
        for (int a = 0; a < 3; ++a) o[14 + a] = v.extent[a];
        for (int a = 0; a < 3; ++a) o[17 + a] = v.scale[a];
        for (int a = 0; a < 3; ++a) o[20 + a] = v.tex_max[a];
        o[23] = 0.0f;
        return VV_OK;
    });
}

// Lanes per ray of the LIC-volume caster: the tuning value, else 4 for views
// of 2^18 rays or more and about 2^20 lanes in all below (at most 64 per
// ray).  With one lane per pixel the longest rays set the frame time while
// most waves have finished.  Measured on MI355X (256^3 LIC volume, step
// 1/256, mean of four views): 1024^2 0.70 -> 0.44 ms (G = 4), 512^2
// 0.50 -> 0.23 ms (4), 256^2 0.35 -> 0.14 ms (16).
static int volume_group_for(const vv_ctx* c, const vv_camera* cam)
{
    if (c->volume_group > 0) return c->volume_group;
    const long long rays = (long long)cam->width * cam->height;
    int G = 4;
    while (G < 64 && rays < (1ll << 18) && rays * G < (1ll << 20)) G *= 2;
    return G;
}

vv_status vv_render(vv_ctx* c, const vv_camera* cam, float* image, int on_device, uint64_t* samples_out)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!image) return fail(c, VV_EINVAL, "image is NULL");
        HIPCHK(c, hipSetDevice(c->device));
        RenderLaunch R;
        vv_status st = fill_render_launch(c, cam, R);
        if (st) return st;
        const size_t bytes = sizeof(float) * 4 * (size_t)cam->width * cam->height;
        float* dst = image;
        if (!on_device) {
            HIPCHK(c, c->image.ensure(bytes));
            dst = (float*)c->image.p;
        }
        R.image = dst;
        if (samples_out) {
            HIPCHK(c, c->counter.ensure(64));
            R.counter = (unsigned long long*)((char*)c->counter.p + 8);
            HIPCHK(c, launch_fill_u32((uint32_t*)R.counter, 0u, 2, c->stream));
        }
        const int G = volume_group_for(c, cam);
        if (G > 1)
            HIPCHK(c, launch_render_rg(R, G, c->stream));
        else
            HIPCHK(c, launch_render(R, c->stream));
        if (!on_device) HIPCHK(c, hipMemcpyAsync(image, dst, bytes, hipMemcpyDeviceToHost, c->stream));
        if (samples_out) {
            unsigned long long t = 0;
            HIPCHK(c, hipMemcpyAsync(&t, R.counter, 8, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            *samples_out = t;
        } else if (!on_device) {
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        return VV_OK;
    });
}

// Lanes per ray of the on-the-fly casters: enough lanes in flight to cover
// the serial chain of samples.  Measured at 256^3 L = 40 on 256^2..1024^2
// views: about 2^21 lanes for an FP32 field (its LIC is bound by L1 gathers,
// more lanes do not help) and 2^22 for FP16 / UNORM8 fields, at most 32
// lanes per ray.
static int ray_group_for(const vv_ctx* c, const vv_camera* cam)
{
    if (c->ray_group > 0) return c->ray_group;
    const long long rays = (long long)cam->width * cam->height;
    const long long lanes = c->storage == VV_STORE_FP32 ? (1ll << 21) : (1ll << 22);
    int G = 1;
    while (G < 32 && rays * G < lanes) G *= 2;
    return G;
}

vv_status vv_render_fused(vv_ctx* c, const vv_camera* cam, float* image, int on_device, uint64_t* samples_out)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!image) return fail(c, VV_EINVAL, "image is NULL");
        HIPCHK(c, hipSetDevice(c->device));
        LicLaunch L;
        vv_status st = fill_lic_launch(c, L, nullptr);
        if (st) return st;
        RenderLaunch R;
        st = fill_render_launch(c, cam, R, false);
        if (st) return st;
        const size_t bytes = sizeof(float) * 4 * (size_t)cam->width * cam->height;
        float* dst = image;
        if (!on_device) {
            HIPCHK(c, c->image.ensure(bytes));
            dst = (float*)c->image.p;
        }
        R.image = dst;
        if (samples_out) {
            HIPCHK(c, c->counter.ensure(64));
            R.counter = (unsigned long long*)((char*)c->counter.p + 8);
            HIPCHK(c, launch_fill_u32((uint32_t*)R.counter, 0u, 2, c->stream));
        }
        const int G = ray_group_for(c, cam);
        if (G > 1) {
            SliceSetup S;
            std::memset(&S, 0, sizeof(S));
            HIPCHK(c, launch_render_group(R, L, S, 0, G, c->stream));
        } else {
            HIPCHK(c, launch_render_fused(R, L, c->stream));
        }
        if (!on_device) HIPCHK(c, hipMemcpyAsync(image, dst, bytes, hipMemcpyDeviceToHost, c->stream));
        if (samples_out) {
            unsigned long long t = 0;
            HIPCHK(c, hipMemcpyAsync(&t, R.counter, 8, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            *samples_out = t;
        } else if (!on_device) {
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        return VV_OK;
    });
}

// View-aligned slicing into the RGBA16F render target (blend 0, float4
// pixels) or blended into the 8-bit window framebuffer (blend 1, 4 bytes per
// pixel).  Both slicing shaders index the TF with the field's .a
// (lic3d_slicing_fragment.glsl:43, lic3d_slicingblend_fragment.glsl:40).
static vv_status render_slices(vv_ctx* c, const vv_camera* cam, void* image, int on_device, uint64_t* samples_out,
                               int blend)
{
    if (!c) return VV_EINVAL;
    if (!image) return fail(c, VV_EINVAL, "image is NULL");
    HIPCHK(c, hipSetDevice(c->device));
    LicLaunch L;
    vv_status st = fill_lic_launch(c, L, nullptr);
    if (st) return st;
    RenderLaunch R;
    st = fill_render_launch(c, cam, R, false);
    if (st) return st;
    R.tf_ch = 3;
    // ViewSlicing::setupSlicing (VV/slicing.cpp:42-100): the view vector, the
    // depth range D of the box along it and n = D / stepSizeVol + 1 slices
    SliceSetup S;
    float inv = 1.0f / std::sqrt(R.view.dw[0] * R.view.dw[0] + R.view.dw[1] * R.view.dw[1] +
                                 R.view.dw[2] * R.view.dw[2]);
    float half[3], dv = 0.0f;
    for (int a = 0; a < 3; ++a) {
        S.v[a] = R.view.dw[a] * inv;
        if (std::fabs(S.v[a]) < 1.0e-8f) S.v[a] = 0.0f;  // VS_EPS (VV/slicing.cpp:129-134)
        S.center[a] = c->center[a];
        half[a] = 0.5f * c->extent[a];
    }
    dv = std::fabs(S.v[0]) * half[0] + std::fabs(S.v[1]) * half[1] + std::fabs(S.v[2]) * half[2];
    S.D = 2.0f * dv;
    const float nslices = S.D / c->params.step_size_vol;
    if (!(nslices < 1.0e6f)) return fail(c, VV_EINVAL, "too many slices (%g)", (double)nslices);
    S.n = (int)nslices + 1;
    S.blend = blend;
    const size_t bytes = (blend ? 4 : sizeof(float) * 4) * (size_t)cam->width * cam->height;
    void* dst = image;
    if (!on_device) {
        HIPCHK(c, c->image.ensure(bytes));
        dst = c->image.p;
    }
    R.image = static_cast<float*>(dst);
    if (samples_out) {
        HIPCHK(c, c->counter.ensure(64));
        R.counter = (unsigned long long*)((char*)c->counter.p + 8);
        HIPCHK(c, launch_fill_u32((uint32_t*)R.counter, 0u, 2, c->stream));
    }
    const int G = ray_group_for(c, cam);
    if (G > 1)
        HIPCHK(c, launch_render_group(R, L, S, 1, G, c->stream));
    else
        HIPCHK(c, launch_render_slices(R, L, S, c->stream));
    if (!on_device) HIPCHK(c, hipMemcpyAsync(image, dst, bytes, hipMemcpyDeviceToHost, c->stream));
    if (samples_out) {
        unsigned long long t = 0;
        HIPCHK(c, hipMemcpyAsync(&t, R.counter, 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        *samples_out = t;
    } else if (!on_device) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return VV_OK;
}

vv_status vv_render_slices(vv_ctx* c, const vv_camera* cam, float* image, int on_device, uint64_t* samples_out)
{
    return guarded(c, [&]() -> vv_status { return render_slices(c, cam, image, on_device, samples_out, 0); });
}

vv_status vv_render_slices_blend(vv_ctx* c, const vv_camera* cam, uint8_t* image, int on_device,
                                 uint64_t* samples_out)
{
    return guarded(c, [&]() -> vv_status { return render_slices(c, cam, image, on_device, samples_out, 1); });
}

// ---- differentiation -----------------------------------------------------------
vv_status vv_set_opacity_scale(vv_ctx* c, const float* sigma, int on_device)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        ++c->scene_version;
        if (!sigma) {
            c->has_sigma = false;
            return VV_OK;
        }
        if (!c->has_lic) return fail(c, VV_ESTATE, "opacity scale lives on the LIC grid: no LIC volume");
        HIPCHK(c, hipSetDevice(c->device));
        const long long n = nvox(c->ldims);
        vv_status st = upload(c, c->stage, sigma, sizeof(float) * (size_t)n, on_device);
        if (st) return st;
        HIPCHK(c, c->sigma.ensure(sizeof(float) * (size_t)clamp_elems(c->ldims)));
        HIPCHK(c, launch_pad_clamp_f32((const float*)c->stage.p, c->ldims, (float*)c->sigma.p, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->has_sigma = true;
        return VV_OK;
    });
}

// Launch descriptor and checkpoint count of an adjoint call.  Samples per ray
// are bounded by the box diagonal in texture space over the shortest
// per-step displacement (|dir| >= min scale).
static vv_status adjoint_setup(vv_ctx* c, const vv_camera* cam, AdjointLaunch& A)
{
    std::memset(&A, 0, sizeof(A));
    vv_status st = fill_render_launch(c, cam, A.r);
    if (st) return st;
    double diag = 0.0, smin = 1e30;
    for (int a = 0; a < 3; ++a) {
        diag += (double)A.r.view.tex_max[a] * A.r.view.tex_max[a];
        if (c->scale[a] < smin) smin = c->scale[a];
    }
    diag = std::sqrt(diag);
    const long long geo = (long long)(diag / (smin * (double)c->params.step_size_vol) * 1.01) + 4;
    A.r.max_samples = A.r.max_samples < geo ? A.r.max_samples : geo;
    A.max_ckpt = (int)((A.r.max_samples + kAdjSeg - 1) / kAdjSeg);
    A.use_window = c->adj_window;
    // the forward pass takes the frame's lanes per ray, at most 4
    const int G = volume_group_for(c, cam);
    A.group = G >= 4 ? 4 : G;
    return VV_OK;
}

// Pointers of A into an adjoint scratch buffer laid out by SL.
static void adjoint_bind(AdjointLaunch& A, char* scratch, const AdjScratchLayout& SL, bool with_sigma)
{
    A.hdr = reinterpret_cast<unsigned long long*>(scratch);
    A.ckpt = reinterpret_cast<float*>(scratch + SL.ckpt);
    A.nsamp = reinterpret_cast<int*>(scratch + SL.nsamp);
    A.ray_bound = reinterpret_cast<float*>(scratch + SL.bound);
    long long* acc = reinterpret_cast<long long*>(scratch + SL.acc);
    A.dtf_fx = acc;
    A.dsigma_fx = with_sigma ? acc + 1024 : nullptr;
}

// The adjoint's forward pass (header word 0 cleared first).
static vv_status adjoint_forward(vv_ctx* c, AdjointLaunch& A)
{
    HIPCHK(c, launch_fill_i64(reinterpret_cast<long long*>(A.hdr), 0, 1, c->stream));
    HIPCHK(c, launch_adjoint_forward(A, c->stream));
    return VV_OK;
}

// The backward walk and the fixed-point conversion on device buffers:
// dimage in, dtf (1024) and dsigma (nv, or null) out.
static vv_status adjoint_backward(vv_ctx* c, AdjointLaunch& A, const float* dimage, float* dtf, float* dsigma,
                                  long long nv)
{
    HIPCHK(c, launch_fill_i64(reinterpret_cast<long long*>(A.hdr) + 1, 0, 1, c->stream));
    HIPCHK(c, launch_fill_i64(A.dtf_fx, 0, 1024 + (dsigma ? nv : 0), c->stream));
    A.dimage = dimage;
    HIPCHK(c, launch_render_adjoint(A, c->stream));
    HIPCHK(c, launch_fixed_to_float(A.hdr, A.dtf_fx, dtf, 1024, c->stream));
    if (dsigma) HIPCHK(c, launch_fixed_to_float(A.hdr, A.dsigma_fx, dsigma, nv, c->stream));
    return VV_OK;
}

// The full layout (a column per segment, no re-run in the backward walk)
// when its columns take at most this many bytes; the compact one otherwise.
constexpr size_t kFullColumnCap = 8ull << 30;
static bool full_columns_fit(long long nrays, int max_ckpt, size_t cap)
{
    return (size_t)kAdjColumnBytes * (size_t)max_ckpt * (size_t)nrays <= cap;
}

// Both passes on a scratch buffer of adj_scratch_layout(nrays, A.max_ckpt,
// dsigma ? nv : 0, A.full_col) bytes.
static vv_status adjoint_enqueue(vv_ctx* c, AdjointLaunch& A, char* scratch, const float* dimage, float* dtf,
                                 float* dsigma, long long nv)
{
    const long long nrays = (long long)A.r.view.width * A.r.view.height;
    adjoint_bind(A, scratch, adj_scratch_layout(nrays, A.max_ckpt, dsigma ? nv : 0, A.full_col != 0),
                 dsigma != nullptr);
    vv_status st = adjoint_forward(c, A);
    if (st) return st;
    return adjoint_backward(c, A, dimage, dtf, dsigma, nv);
}

vv_status vv_render_record(vv_ctx* c, const vv_camera* cam, float* image, int on_device, uint64_t* samples_out)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!image) return fail(c, VV_EINVAL, "image is NULL");
        HIPCHK(c, hipSetDevice(c->device));
        AdjointLaunch A;
        vv_status st = adjoint_setup(c, cam, A);
        if (st) return st;
        c->adj_rec.valid = false;
        const long long nrays = (long long)cam->width * cam->height;
        A.full_col = full_columns_fit(nrays, A.max_ckpt, kFullColumnCap);
        // room for the opacity-scale gradient, so the adjoint can keep the buffer
        const AdjScratchLayout SL = adj_scratch_layout(nrays, A.max_ckpt, nvox(c->ldims), A.full_col != 0);
        HIPCHK(c, c->adj_ckpt.ensure(SL.total));
        const size_t bytes = sizeof(float) * 4 * (size_t)nrays;
        float* dst = image;
        if (!on_device) {
            HIPCHK(c, c->image.ensure(bytes));
            dst = (float*)c->image.p;
        }
        A.r.image = dst;
        if (samples_out) {
            HIPCHK(c, c->counter.ensure(64));
            A.r.counter = (unsigned long long*)((char*)c->counter.p + 8);
            HIPCHK(c, launch_fill_u32((uint32_t*)A.r.counter, 0u, 2, c->stream));
        }
        adjoint_bind(A, (char*)c->adj_ckpt.p, SL, false);
        st = adjoint_forward(c, A);
        if (st) return st;
        if (!on_device) HIPCHK(c, hipMemcpyAsync(image, dst, bytes, hipMemcpyDeviceToHost, c->stream));
        if (samples_out) {
            unsigned long long t = 0;
            HIPCHK(c, hipMemcpyAsync(&t, A.r.counter, 8, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            *samples_out = t;
        } else if (!on_device) {
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        c->adj_rec.valid = true;
        c->adj_rec.cam = *cam;
        c->adj_rec.version = c->scene_version;
        c->adj_rec.max_ckpt = A.max_ckpt;
        c->adj_rec.full_col = A.full_col;
        c->adj_rec.buffer = c->adj_ckpt.p;
        return VV_OK;
    });
}

vv_status vv_record_ray_samples(vv_ctx* c, const vv_camera* cam, int32_t* counts, int on_device)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!cam || !counts) return fail(c, VV_EINVAL, "cam/counts is NULL");
        if (!c->adj_rec.valid || c->adj_rec.buffer != c->adj_ckpt.p)
            return fail(c, VV_ESTATE, "no recorded forward pass (vv_render_record; an adjoint consumes it)");
        if (std::memcmp(&c->adj_rec.cam, cam, sizeof(vv_camera)) != 0)
            return fail(c, VV_ESTATE, "the recorded forward pass is of another camera");
        HIPCHK(c, hipSetDevice(c->device));
        const long long nrays = (long long)c->adj_rec.cam.width * c->adj_rec.cam.height;
        const AdjScratchLayout SL = adj_scratch_layout(nrays, c->adj_rec.max_ckpt, 0, c->adj_rec.full_col != 0);
        HIPCHK(c, hipMemcpyAsync(counts, (const char*)c->adj_ckpt.p + SL.nsamp, sizeof(int32_t) * (size_t)nrays,
                                 on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
        if (!on_device) HIPCHK(c, hipStreamSynchronize(c->stream));
        return VV_OK;
    });
}

vv_status vv_render_adjoint(vv_ctx* c, const vv_camera* cam, const float* dimage, float* dtf, float* dsigma,
                            int on_device)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!dimage || !dtf) return fail(c, VV_EINVAL, "dimage/dtf NULL");
        HIPCHK(c, hipSetDevice(c->device));
        AdjointLaunch A;
        vv_status st = adjoint_setup(c, cam, A);
        if (st) return st;
        const long long nrays = (long long)cam->width * cam->height;
        const long long nv = nvox(c->ldims);
        A.full_col = full_columns_fit(nrays, A.max_ckpt, kFullColumnCap);
        const AdjScratchLayout SL = adj_scratch_layout(nrays, A.max_ckpt, dsigma ? nv : 0, A.full_col != 0);
        HIPCHK(c, c->adj_ckpt.ensure(SL.total));
        // the forward pass vv_render_record left for this camera and scene
        const bool reuse = c->adj_rec.valid && c->adj_rec.version == c->scene_version &&
                           c->adj_rec.buffer == c->adj_ckpt.p && c->adj_rec.max_ckpt == A.max_ckpt &&
                           c->adj_rec.full_col == A.full_col &&
                           std::memcmp(&c->adj_rec.cam, cam, sizeof(vv_camera)) == 0;
        // a forward pass of this call overwrites the buffer, and the compact
        // layout's backward walk rewrites its segment column; only a reused
        // full-layout record stays as it was
        if (!reuse || !A.full_col) c->adj_rec.valid = false;
        const size_t img_bytes = sizeof(float) * 4 * (size_t)nrays;
        const float* d_dimg = dimage;
        float* d_dtf = dtf;
        float* d_dsigma = dsigma;
        if (!on_device) {
            st = upload(c, c->adj_dimg, dimage, img_bytes, 0);
            if (st) return st;
            d_dimg = (const float*)c->adj_dimg.p;
            HIPCHK(c, c->adj_out.ensure(sizeof(float) * (size_t)(1024 + (dsigma ? nv : 0))));
            d_dtf = (float*)c->adj_out.p;
            d_dsigma = dsigma ? d_dtf + 1024 : nullptr;
        }
        adjoint_bind(A, (char*)c->adj_ckpt.p, SL, dsigma != nullptr);
        if (!reuse) {
            st = adjoint_forward(c, A);
            if (st) return st;
        }
        st = adjoint_backward(c, A, d_dimg, d_dtf, d_dsigma, nv);
        if (st) return st;
        if (!on_device) {
            HIPCHK(c, hipMemcpyAsync(dtf, d_dtf, sizeof(float) * 1024, hipMemcpyDeviceToHost, c->stream));
            if (dsigma)
                HIPCHK(c, hipMemcpyAsync(dsigma, d_dsigma, sizeof(float) * (size_t)nv, hipMemcpyDeviceToHost,
                                         c->stream));
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return VV_OK;
    });
}

vv_status vv_render_adjoint_noalloc(vv_ctx* c, const vv_camera* cam, const float* dimage, float* dtf,
                                    float* dsigma)
{
    return guarded(c, [&]() -> vv_status {
        if (!c) return VV_EINVAL;
        if (!dimage || !dtf) return fail(c, VV_EINVAL, "dimage/dtf NULL");
        if (!c->noalloc_scratch) return fail(c, VV_ESTATE, "no scratch array on this device");
        if (c->params.illum_mode >= VV_ILLUM_MALLO && (c->illum_dirty || !c->illum_luts.p))
            return fail(c, VV_ESTATE,
                        "line-illumination tables are not on the device: render once with this material first");
        HIPCHK(c, hipSetDevice(c->device));
        AdjointLaunch A;
        vv_status st = adjoint_setup(c, cam, A);
        if (st) return st;
        const long long nrays = (long long)cam->width * cam->height;
        const long long nv = nvox(c->ldims);
        A.full_col = adj_scratch_layout(nrays, A.max_ckpt, dsigma ? nv : 0, true).total <= kNoAllocScratchBytes;
        const AdjScratchLayout SL = adj_scratch_layout(nrays, A.max_ckpt, dsigma ? nv : 0, A.full_col != 0);
        if (SL.total > kNoAllocScratchBytes)
            return fail(c, VV_EINVAL,
                        "the allocation-free adjoint needs %zu bytes of scratch for a %dx%d view (%d checkpoints per "
                        "ray%s); its array holds %zu",
                        SL.total, cam->width, cam->height, A.max_ckpt, dsigma ? ", opacity-scale gradient" : "",
                        kNoAllocScratchBytes);
        return adjoint_enqueue(c, A, static_cast<char*>(c->noalloc_scratch), dimage, dtf, dsigma, nv);
    });
}

}  // extern "C"
