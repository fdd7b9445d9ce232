// vv_anim.cpp — unsteady fields: the reference's animation of a time series.
//
// The reference keeps two raw time steps on the host (data, newData) and,
// every frame, re-encodes data + (interpIndex/InterpSize)(newData - data)
// into the field texture, recomputes the LIC volume, and after InterpSize
// frames moves to the next pair, wrapping at the end of the series
// (VV/3DLIC.cpp:700-708 init with InterpSize 10, :129-172 idle;
// VV/dataset.cpp:195-210 checkInterpolateStage, :533-635 the lerped encode).
// Here every time step stays resident on the GPU as raw float32 vectors and a
// frame is one device-side lerp + encode into the context's field
// (vv_upload_field with device inputs): no host work and no upload per frame.
// vv_batch_set_field re-encodes a batch time step the same way, so a captured
// batch graph replays over new data without being rebuilt.
#include "vv_ctx.hpp"

using namespace vv_host;

struct vv_anim {
    vv_ctx* ctx = nullptr;
    std::vector<DevBuf> steps;  // raw float32 x 3 per voxel, one per time step
    int dims[3] = {0, 0, 0};
    float spacing[3] = {1, 1, 1};
    int interp = 1;
    int storage = VV_STORE_FP32;
    long long frame = 0;
};

namespace {

vv_status anim_alloc(vv_ctx* c, int nsteps, const int dims[3], const float spacing[3], int interp, int storage,
                     vv_anim** out)
{
    if (!out) return fail(c, VV_EINVAL, "out is NULL");
    *out = nullptr;
    if (nsteps < 1) return fail(c, VV_EINVAL, "a series needs at least one time step");
    if (!dims_ok(dims)) return fail(c, VV_EINVAL, "field dims %s", kDimsRule);
    if (interp < 1) return fail(c, VV_EINVAL, "interp_size must be >= 1");
    if (!storage_ok(storage)) return fail(c, VV_EINVAL, "unknown storage");
    vv_anim* a = new (std::nothrow) vv_anim();
    if (!a) return fail(c, VV_ENOMEM, "out of host memory");
    a->ctx = c;
    a->steps.resize((size_t)nsteps);
    for (int k = 0; k < 3; ++k) {
        a->dims[k] = dims[k];
        a->spacing[k] = spacing ? spacing[k] : 1.0f;
    }
    a->interp = interp;
    a->storage = storage;
    const size_t bytes = sizeof(float) * 3 * (size_t)nvox(dims);
    for (DevBuf& d : a->steps) {
        hipError_t e = d.ensure(bytes);
        if (e) {
            for (DevBuf& x : a->steps) x.release();
            delete a;
            return fail(c, VV_ENOMEM, "time steps: %s", hipGetErrorString(e));
        }
    }
    *out = a;
    return VV_OK;
}

}  // namespace

extern "C" {

vv_status vv_anim_create(vv_ctx* ctx, const float* const* steps, int nsteps, const int dims[3],
                         const float spacing[3], int interp_size, int storage, int on_device, vv_anim** out)
{
    return guarded(ctx, [&]() -> vv_status {
        if (!ctx) return VV_EINVAL;
        if (!steps) return fail(ctx, VV_EINVAL, "steps is NULL");
        for (int t = 0; t < nsteps; ++t)
            if (!steps[t]) return fail(ctx, VV_EINVAL, "time step %d is NULL", t);
        HIPCHK(ctx, hipSetDevice(ctx->device));
        vv_anim* a = nullptr;
        vv_status s = anim_alloc(ctx, nsteps, dims, spacing, interp_size, storage, &a);
        if (s) return s;
        const size_t n = 3 * (size_t)nvox(dims);
        for (int t = 0; t < nsteps && !s; ++t) {
            if (!on_device)
                for (size_t i = 0; i < n; ++i)
                    if (!std::isfinite(steps[t][i])) {
                        s = fail(ctx, VV_EINVAL, "time step %d holds a non-finite value", t);
                        break;
                    }
            if (s) break;
            const hipError_t e = hipMemcpyAsync(a->steps[t].p, steps[t], sizeof(float) * n,
                                                on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                                ctx->stream);
            if (e) s = fail(ctx, VV_EDEVICE, "time step copy: %s", hipGetErrorString(e));
        }
        if (!s && hipStreamSynchronize(ctx->stream)) s = fail(ctx, VV_EDEVICE, "time step copy failed");
        if (s) {
            vv_anim_destroy(a);
            return s;
        }
        *out = a;
        return VV_OK;
    });
}

vv_status vv_anim_load_dat(vv_ctx* ctx, const char* dat_path, int interp_size, int storage, vv_anim** out)
{
    return guarded(ctx, [&]() -> vv_status {
        if (!ctx) return VV_EINVAL;
        vv_dat_info info;
        vv_status s = vv_dat_parse(dat_path, &info);
        if (s) return fail(ctx, s, "%s", g_create_error.c_str());
        if (!dims_ok(info.dims)) return fail(ctx, VV_EINVAL, "field dims %s", kDimsRule);
        HIPCHK(ctx, hipSetDevice(ctx->device));
        const int nsteps = info.t_end - info.t_begin + 1;
        vv_anim* a = nullptr;
        s = anim_alloc(ctx, nsteps, info.dims, info.spacing, interp_size, storage, &a);
        if (s) return s;
        std::vector<float> host(3 * (size_t)nvox(info.dims));
        for (int t = 0; t < nsteps && !s; ++t) {
            s = vv_dat_read_field(&info, info.t_begin + t, host.data());
            if (s) {
                s = fail(ctx, s, "%s", g_create_error.c_str());
                break;
            }
            for (float x : host)
                if (!std::isfinite(x)) {
                    s = fail(ctx, VV_EINVAL, "time step %d holds a non-finite value", info.t_begin + t);
                    break;
                }
            if (s) break;
            // synchronous: the host buffer is refilled for the next step
            if (hipMemcpy(a->steps[t].p, host.data(), sizeof(float) * host.size(), hipMemcpyHostToDevice))
                s = fail(ctx, VV_EDEVICE, "time step upload failed");
        }
        if (s) {
            vv_anim_destroy(a);
            return s;
        }
        *out = a;
        return VV_OK;
    });
}

vv_status vv_anim_destroy(vv_anim* a)
{
    return guarded((a ? a->ctx : nullptr), [&]() -> vv_status {
        if (!a) return VV_OK;
        (void)hipSetDevice(a->ctx->device);
        (void)hipStreamSynchronize(a->ctx->stream);
        for (DevBuf& d : a->steps) d.release();
        delete a;
        return VV_OK;
    });
}

vv_status vv_anim_seek(vv_anim* a, long long frame)
{
    return guarded((a ? a->ctx : nullptr), [&]() -> vv_status {
        if (!a) return VV_EINVAL;
        if (frame < 0) return fail(a->ctx, VV_EINVAL, "frame must be >= 0");
        a->frame = frame;
        return VV_OK;
    });
}

vv_status vv_anim_next(vv_anim* a, int* step, float* lerp_t)
{
    return guarded((a ? a->ctx : nullptr), [&]() -> vv_status {
        if (!a) return VV_EINVAL;
        const long long n = (long long)a->steps.size();
        const long long pair = a->frame / a->interp;
        const int index = (int)(a->frame % a->interp);
        const int cur = (int)(pair % n), next = (int)((pair + 1) % n);
        const float t = (float)index / (float)a->interp;  // (float)interpIndex / InterpSize
        vv_status s = vv_upload_field(a->ctx, (const float*)a->steps[cur].p, (const float*)a->steps[next].p, t, a->dims,
                                      a->spacing, a->storage, 1);
        if (s) return s;
        if (step) *step = cur;
        if (lerp_t) *lerp_t = t;
        ++a->frame;
        return VV_OK;
    });
}

}  // extern "C"
