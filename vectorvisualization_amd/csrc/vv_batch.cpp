// vv_batch.cpp — batched LIC + ray casting: several (unsteady) timesteps of
// different sizes and many views rendered from them, run as one captured
// hipGraph of persistent kernels:
//
//   k_stamp         start time of the batch (wall clock)
//   k_fill_u32      zero the render work counter, per-view sample counters
//                   and completion stamps (a kernel, not a memset, so the
//                   graph can be replayed)
//   k_lic_batch     every LIC tile of every timestep
//   k_render_batch  every pixel tile of every view, dynamically scheduled
//
// Everything runs on the context's stream: each timestep's LIC volume, then
// that timestep's views.  (Rendering a step's views on a second stream beside
// the next step's LIC volume paid while the ray caster ran one lane per
// pixel; with lanes per ray the two kernels only slow each other down.)
//
// The reference animates an unsteady field by re-encoding, re-uploading and
// recomputing the LIC volume frame by frame (VV/3DLIC.cpp:129-172,
// VV/dataset.cpp:202-210, 290-366); a batch holds all timesteps resident and
// produces all views in one launch.  The autotuner times the LIC tile /
// brick order per volume size and the render tile on the batch's own data.
#include "vv_ctx.hpp"

#include <algorithm>
#include <array>
#include <functional>
#include <map>
#include <mutex>
#include <tuple>

using namespace vv_host;

namespace {

struct BatchStep {
    int dims[3];
    float extent[3], scale[3], center[3];
    int storage;
    DevBuf field, noise, lic;
    DevBuf lic_field;  // field copy with the noise byte in channel 3 (launch_pack_lic_field)
    int tile[3] = {16, 4, 4};
    int super[3] = {2, 8, 8};
};

struct BatchView {
    int step;
    vv_camera cam;
    float* image;
};

// (dims, storage, F, B) -> tile x,y,z + brick (super-tile) x,y,z
using TuneKey = std::tuple<int, int, int, int, int, int>;
std::map<TuneKey, std::array<int, 6>> g_lic_tune;
std::mutex g_tune_mu;

}  // namespace

struct vv_batch {
    vv_ctx* ctx = nullptr;
    std::vector<BatchStep*> steps;
    std::vector<BatchView> views;
    DevBuf jobs, first, rviews, vfirst, ctr;
    std::vector<unsigned char> host_jobs;
    std::vector<int> host_first;   // per time step: 0, tiles
    std::vector<int> host_vfirst;  // per group: tile offsets of its views (0 first)
    std::vector<RenderLaunch> host_views;  // views grouped by time step, in run order
    std::vector<int> order;        // time steps in run order (fewest voxels first)
    std::vector<int> slot;         // user view -> position in host_views
    std::vector<int> gview, gfirst;  // per group: first view slot, first host_vfirst entry
    int ngroups = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<long long> graph_key;
    // descriptors as last uploaded to jobs / first / rviews / vfirst
    std::vector<unsigned char> up_jobs, up_views;
    std::vector<int> up_first, up_vfirst;
    int render_block[2] = {16, 16};  // pixel tile with one lane per pixel
    int render_group = 4;            // lanes per ray of the ray caster: 1, 2 or 4
    bool ran = false;
};

namespace {

void release_graph(vv_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_key.clear();
}

// counter block layout (u32 units): one render work counter per time step
// (rounded up to an even count), per-view u64 sample counters, per-view u64
// completion stamps, then u64 t0
size_t ctr_head(size_t nsteps) { return (nsteps + 1) & ~(size_t)1; }
size_t ctr_words(size_t nsteps, size_t nviews) { return ctr_head(nsteps) + 6 * nviews + 2; }

vv_status build_lic_launch(vv_ctx* c, BatchStep* st, LicLaunch& L)
{
    const vv_lic_params& p = c->params;
    std::memset(&L, 0, sizeof(L));
    L.field = make_clamp_desc(st->lic_field.p, st->dims);  // noise on the field's grid
    L.field_has_noise = 1;
    L.field_type = st->storage;
    L.noise = make_repeat_desc(st->noise.p, st->dims);
    L.nc = 1;
    L.w = (const float*)c->weights.p;
    L.F = p.steps_forward;
    L.B = p.steps_backward;
    L.he = p.step_size_lic * p.eye_factor;
    L.scl = (c->inv_area / (float)(p.steps_forward + p.steps_backward)) * p.gradient_scale;
    L.freq = p.freq_scale;
    for (int a = 0; a < 3; ++a) L.scale[a] = st->scale[a];
    L.out = make_repeat_desc(st->lic.p, st->dims);
    for (int a = 0; a < 3; ++a) {
        L.tile[a] = st->tile[a];
        L.super_tile[a] = st->super[a];
    }
    L.field_bytes = (unsigned long long)clamp_elems(st->dims) * field_texel_bytes(L.field_type);
    L.noise_bytes = (unsigned long long)repeat_elems(st->dims);
    L.cache = 1;
    if (!(L.field_bytes < (1ull << 32) && L.noise_bytes < (1ull << 32) && (long long)L.field.sz < (1ll << 24) &&
          (long long)L.noise.sz < (1ll << 24)))
        return fail(c, VV_EUNSUPPORTED, "batch timestep %dx%dx%d too large for 32-bit buffer offsets",
                    st->dims[0], st->dims[1], st->dims[2]);
    if (p.speed_of_flow) {
        // SPEED_OF_FLOW reads the magnitude channel of the encoded field: the
        // one-lane-per-voxel kernel on the field itself, not the LIC copy
        L.field = make_clamp_desc(st->field.p, st->dims);
        L.field_has_noise = 0;
        L.cache = 0;
        L.speed = 1;
    }
    return VV_OK;
}

vv_status build_view_launch(vv_batch* b, const BatchView& v, RenderLaunch& R)
{
    vv_ctx* c = b->ctx;
    BatchStep* st = b->steps[v.step];
    vv_status s = check_camera(c, &v.cam);
    if (s) return s;
    const vv_lic_params& p = c->params;
    std::memset(&R, 0, sizeof(R));
    R.field = make_clamp_desc(st->field.p, st->dims);
    R.field_type = st->storage;
    R.tf_ch = p.tf_index == VV_TF_INDEX_MAG ? 3 : 2;
    R.lic = make_repeat_desc(st->lic.p, st->dims);
    R.lc = 1;
    R.tf = (const float*)c->tf_dev.p;
    R.view = view_block_geom(st->extent, st->scale, st->center, &v.cam);
    R.step = p.step_size_vol;
    R.alpha_corr = p.step_size_vol * 128.0f;
    R.corr_mode = R.alpha_corr == 1.0f ? 0 : (R.alpha_corr == 0.5f ? 1 : (R.alpha_corr == 2.0f ? 2 : 3));
    R.max_samples = (long long)p.num_iterations * (long long)p.num_iterations;
    // batch LIC volumes are scalar: illumGradient falls back to illumLIC as
    // for a scalar volume in vv_render; the line modes need their tables
    R.illum = p.illum_mode;
    if (p.illum_mode >= VV_ILLUM_MALLO) {
        vv_status st = ensure_illum_luts(c);
        if (st) return st;
        illum_lut_ptrs(c, R);
    }
    R.illum_scale = p.illum_scale;
    const float l0 = v.cam.light_pos[0], l1 = v.cam.light_pos[1], l2 = v.cam.light_pos[2];
    const float ld = std::sqrt(l0 * l0 + l1 * l1 + l2 * l2);
    if (ld > 0.0f) {
        R.light[0] = l0 / ld;
        R.light[1] = l1 / ld;
        R.light[2] = l2 / ld;
    }
    R.light_diffuse = v.cam.light_diffuse;
    R.light_specular = v.cam.light_specular;
    R.light_ambient = v.cam.light_ambient;
    R.spec_exp = v.cam.spec_exponent;
    R.image = v.image;
    R.block[0] = b->render_block[0];
    R.block[1] = b->render_block[1];
    return VV_OK;
}

float time_ms(hipStream_t s, const std::function<hipError_t()>& fn, int reps, hipError_t* err)
{
    hipEvent_t a, e;
    *err = hipEventCreate(&a);
    if (*err) return 0;
    *err = hipEventCreate(&e);
    if (*err) {
        (void)hipEventDestroy(a);
        return 0;
    }
    float best = 1e30f;
    for (int r = 0; r < reps && *err == hipSuccess; ++r) {
        (void)hipEventRecord(a, s);
        *err = fn();
        (void)hipEventRecord(e, s);
        if (*err == hipSuccess) *err = hipEventSynchronize(e);
        float ms = 0;
        if (*err == hipSuccess) (void)hipEventElapsedTime(&ms, a, e);
        if (r > 0 || reps == 1) best = std::min(best, ms);  // first rep warms up
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(e);
    return best;
}

}  // namespace

extern "C" {

vv_status vv_batch_create(vv_ctx* ctx, vv_batch** out)
{
    return guarded(ctx, [&]() -> vv_status {
        if (!ctx || !out) return VV_EINVAL;
        *out = nullptr;
        vv_batch* b = new (std::nothrow) vv_batch();
        if (!b) return fail(ctx, VV_ENOMEM, "out of host memory");
        b->ctx = ctx;
        *out = b;
        return VV_OK;
    });
}

vv_status vv_batch_destroy(vv_batch* b)
{
    return guarded((b ? b->ctx : nullptr), [&]() -> vv_status {
        if (!b) return VV_OK;
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
        release_graph(b);
        for (BatchStep* st : b->steps) {
            st->field.release();
            st->noise.release();
            st->lic.release();
            st->lic_field.release();
            delete st;
        }
        DevBuf* bufs[] = {&b->jobs, &b->first, &b->rviews, &b->vfirst, &b->ctr};
        for (DevBuf* d : bufs) d->release();
        delete b;
        return VV_OK;
    });
}

vv_status vv_batch_info(const vv_batch* b, int* ntimesteps, int* nviews)
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!b) return VV_EINVAL;
        if (ntimesteps) *ntimesteps = (int)b->steps.size();
        if (nviews) *nviews = (int)b->views.size();
        return VV_OK;
    });
}

vv_status vv_batch_add_timestep(vv_batch* b, const float* v0, const int dims[3], const float spacing[3],
                                int storage, uint64_t noise_seed, int noise_kind, double density,
                                int on_device)
{
    return guarded((b ? b->ctx : nullptr), [&]() -> vv_status {
        if (!b) return VV_EINVAL;
        vv_ctx* c = b->ctx;
        if (!v0) return fail(c, VV_EINVAL, "field data is NULL");
        if (!dims_ok(dims)) return fail(c, VV_EINVAL, "field dims %s", kDimsRule);
        if (!storage_ok(storage)) return fail(c, VV_EINVAL, "unknown storage");
        if (!b->steps.empty() && b->steps[0]->storage != storage)
            return fail(c, VV_EINVAL, "all timesteps of a batch share one storage type");
        if (noise_kind != VV_NOISE_BINARY && noise_kind != VV_NOISE_UNIFORM) return fail(c, VV_EINVAL, "noise kind");
        if (noise_kind == VV_NOISE_BINARY && !(density >= 0.0 && density <= 1.0))
            return fail(c, VV_EINVAL, "density must be in [0,1]");
        float sd[3] = {1, 1, 1};
        if (spacing)
            for (int a = 0; a < 3; ++a) {
                if (!(spacing[a] > 0.0f) || !std::isfinite(spacing[a]))
                    return fail(c, VV_EINVAL, "spacing must be > 0");
                sd[a] = spacing[a];
            }
        HIPCHK(c, hipSetDevice(c->device));
        const long long n = nvox(dims);
        if (!on_device)
            for (long long i = 0; i < 3 * n; ++i)
                if (!std::isfinite(v0[i])) return fail(c, VV_EINVAL, "field contains a non-finite value");
        BatchStep* st = new (std::nothrow) BatchStep();
        if (!st) return fail(c, VV_ENOMEM, "out of host memory");
        auto bail = [&](vv_status s) {
            st->field.release();
            st->noise.release();
            st->lic.release();
            st->lic_field.release();
            delete st;
            return s;
        };
        const float* d0 = v0;
        if (!on_device) {
            vv_status s = upload(c, c->stage, v0, sizeof(float) * 3 * (size_t)n, 0);
            if (s) return bail(s);
            d0 = (const float*)c->stage.p;
        }
        hipError_t e = st->field.ensure((size_t)field_texel_bytes(storage) * (size_t)clamp_elems(dims));
        if (!e) e = st->noise.ensure((size_t)repeat_elems(dims));
        if (!e) e = st->lic.ensure(sizeof(float) * (size_t)repeat_elems(dims));
        if (!e) e = st->lic_field.ensure((size_t)field_texel_bytes(storage) * (size_t)clamp_elems(dims));
        if (!e) e = c->counter.ensure(64);
        if (!e) e = launch_fill_u32((uint32_t*)c->counter.p, 0u, 1, c->stream);
        if (!e) e = launch_field_maxlen(d0, d0, 0.0f, n, (uint32_t*)c->counter.p, c->stream);
        if (!e)
            e = launch_field_encode(d0, d0, 0.0f, dims, (const uint32_t*)c->counter.p, st->field.p, storage,
                                    c->stream);
        if (!e) e = hipStreamSynchronize(c->stream);  // the staging buffer is reused below
        if (e) return bail(fail(c, e == hipErrorOutOfMemory ? VV_ENOMEM : VV_EDEVICE, "timestep upload: %s",
                                hipGetErrorString(e)));
        double t = std::floor((noise_kind == VV_NOISE_BINARY ? density : 0.0) * 4294967296.0);
        t = std::min(std::max(t, 0.0), 4294967296.0);
        e = c->stage.ensure((size_t)n);
        if (!e) e = launch_noise_gen(noise_seed, noise_kind, (uint64_t)t, dims, (uint8_t*)c->stage.p, c->stream);
        if (!e) e = launch_noise_tex((const uint8_t*)c->stage.p, nullptr, dims, st->noise.p, 1, c->stream);
        if (!e)
            e = launch_pack_lic_field(st->field.p, storage, (const uint8_t*)st->noise.p, dims, st->lic_field.p,
                                      c->stream);
        if (!e) e = hipStreamSynchronize(c->stream);
        if (e) return bail(fail(c, VV_EDEVICE, "timestep noise: %s", hipGetErrorString(e)));
        // VectorDataSet::loadData geometry (VV/dataset.cpp:144-176)
        float vol[3], maxvol = 0.0f;
        int maxtex = 0;
        for (int a = 0; a < 3; ++a) {
            st->dims[a] = dims[a];
            vol[a] = (float)dims[a] * sd[a];
            maxvol = std::max(maxvol, vol[a]);
            maxtex = std::max(maxtex, dims[a]);
        }
        for (int a = 0; a < 3; ++a) {
            st->scale[a] = (float)maxtex / ((float)dims[a] * sd[a]);
            st->extent[a] = (float)dims[a] * sd[a] / maxvol;
            st->center[a] = st->extent[a] / 2.0f;
        }
        st->storage = storage;
        b->steps.push_back(st);
        release_graph(b);
        b->ran = false;  // the last run's counters no longer match the batch
        return VV_OK;
    });
}

vv_status vv_batch_set_field(vv_batch* b, int timestep, const float* v0, const float* v1, float lerp_t,
                             int on_device)
{
    return guarded((b ? b->ctx : nullptr), [&]() -> vv_status {
        if (!b) return VV_EINVAL;
        vv_ctx* c = b->ctx;
        if (timestep < 0 || timestep >= (int)b->steps.size())
            return fail(c, VV_EINVAL, "timestep %d out of range", timestep);
        if (!v0) return fail(c, VV_EINVAL, "field data is NULL");
        if (!std::isfinite(lerp_t)) return fail(c, VV_EINVAL, "lerp_t not finite");
        BatchStep* st = b->steps[timestep];
        const long long n = nvox(st->dims);
        HIPCHK(c, hipSetDevice(c->device));
        HIPCHK(c, hipStreamSynchronize(c->stream));  // a running batch may read this field
        const float* d0 = v0;
        const float* d1 = v1 ? v1 : v0;
        if (!on_device) {
            for (long long i = 0; i < 3 * n; ++i)
                if (!std::isfinite(v0[i]) || (v1 && !std::isfinite(v1[i])))
                    return fail(c, VV_EINVAL, "field contains a non-finite value");
            const size_t bytes = sizeof(float) * 3 * (size_t)n;
            vv_status s = upload(c, c->stage, v0, bytes, 0);
            if (!s && v1) s = upload(c, c->stage2, v1, bytes, 0);
            if (s) return s;
            d0 = (const float*)c->stage.p;
            d1 = v1 ? (const float*)c->stage2.p : d0;
        }
        const float t = v1 ? lerp_t : 0.0f;
        hipError_t e = c->counter.ensure(64);
        if (!e) e = launch_fill_u32((uint32_t*)c->counter.p, 0u, 1, c->stream);
        if (!e) e = launch_field_maxlen(d0, d1, t, n, (uint32_t*)c->counter.p, c->stream);
        if (!e) e = launch_field_encode(d0, d1, t, st->dims, (const uint32_t*)c->counter.p, st->field.p,
                                        st->storage, c->stream);
        if (!e) e = launch_pack_lic_field(st->field.p, st->storage, (const uint8_t*)st->noise.p, st->dims,
                                          st->lic_field.p, c->stream);
        if (!e) e = hipStreamSynchronize(c->stream);
        if (e) return fail(c, VV_EDEVICE, "timestep encode: %s", hipGetErrorString(e));
        return VV_OK;
    });
}

vv_status vv_batch_add_view(vv_batch* b, int timestep, const vv_camera* cam, float* image)
{
    return guarded((b ? b->ctx : nullptr), [&]() -> vv_status {
        if (!b) return VV_EINVAL;
        vv_ctx* c = b->ctx;
        if (timestep < 0 || timestep >= (int)b->steps.size())
            return fail(c, VV_EINVAL, "timestep %d out of range", timestep);
        vv_status s = check_camera(c, cam);
        if (s) return s;
        if (!image) return fail(c, VV_EINVAL, "image is NULL (a device buffer of width*height*4 floats)");
        BatchView v;
        v.step = timestep;
        v.cam = *cam;
        v.image = image;
        b->views.push_back(v);
        release_graph(b);
        b->ran = false;
        return VV_OK;
    });
}

// Runs the time steps one after the other, each LIC volume followed by the
// views that show it, fewest voxels first: a view is done as soon as its own
// time step is, so per-view latency grows with the work before it instead of
// all views waiting for every LIC volume.
vv_status vv_batch_run(vv_batch* b)
{
    return guarded((b ? b->ctx : nullptr), [&]() -> vv_status {
        if (!b) return VV_EINVAL;
        vv_ctx* c = b->ctx;
        if (b->steps.empty() || b->views.empty()) return fail(c, VV_ESTATE, "batch needs timesteps and views");
        HIPCHK(c, hipSetDevice(c->device));
        vv_status s = ensure_tf(c);
        if (!s) s = ensure_weights(c);
        if (s) return s;
        const int ftype = b->steps[0]->storage;
        const int same = c->params.freq_scale == 1.0f;
        const size_t nsteps = b->steps.size(), nv = b->views.size();
        // LIC jobs, one per time step
        const size_t asz = lic_args_size();
        b->host_jobs.resize(asz * nsteps);
        b->host_first.assign(2 * nsteps, 0);
        std::vector<LicLaunch> lics(nsteps);
        for (size_t i = 0; i < nsteps; ++i) {
            LicLaunch& L = lics[i];
            s = build_lic_launch(c, b->steps[i], L);
            if (s) return s;
            int nt = 0;
            if (lic_args_fill(L, b->host_jobs.data() + i * asz, &nt) != hipSuccess)
                return fail(c, VV_EINVAL, "LIC tile of timestep %zu must hold 256 threads", i);
            b->host_first[2 * i + 1] = nt;
        }
        b->order.resize(nsteps);
        for (size_t i = 0; i < nsteps; ++i) b->order[i] = (int)i;
        std::stable_sort(b->order.begin(), b->order.end(),
                         [&](int x, int y) { return nvox(b->steps[x]->dims) < nvox(b->steps[y]->dims); });
        // views grouped by time step in run order
        const int bx = b->render_block[0], by = b->render_block[1];
        const size_t head = ctr_head(nsteps);
        const size_t ctr_bytes = sizeof(uint32_t) * ctr_words(nsteps, nv);
        if (b->ctr.bytes < ctr_bytes) {
            HIPCHK(c, hipStreamSynchronize(c->stream));  // a previous run may still use the counters
            HIPCHK(c, b->ctr.ensure(ctr_bytes));
        }
        uint32_t* ctr = (uint32_t*)b->ctr.p;
        unsigned long long* samples = (unsigned long long*)(ctr + head);
        unsigned long long* done = samples + nv;
        unsigned long long* t0 = done + nv;
        unsigned long long* begun = t0 + 1;
        b->host_views.resize(nv);
        b->slot.assign(nv, 0);
        b->gview.assign(nsteps + 1, 0);
        b->gfirst.assign(nsteps, 0);
        b->host_vfirst.clear();
        std::vector<int> gtiles(nsteps, 0);
        size_t k = 0;
        for (size_t g = 0; g < nsteps; ++g) {
            const int t = b->order[g];
            b->gview[g] = (int)k;
            b->gfirst[g] = (int)b->host_vfirst.size();
            b->host_vfirst.push_back(0);
            for (size_t v = 0; v < nv; ++v) {
                if (b->views[v].step != t) continue;
                s = build_view_launch(b, b->views[v], b->host_views[k]);
                if (s) return s;
                b->host_views[k].counter = samples + k;
                b->slot[v] = (int)k;
                const int vw = b->views[v].cam.width, vh = b->views[v].cam.height;
                const int tiles = b->render_group > 1 ? (int)rg_blocks(b->render_group, vw, vh)
                                                      : ((vw + bx - 1) / bx) * ((vh + by - 1) / by);
                b->host_vfirst.push_back(b->host_vfirst.back() + tiles);
                ++k;
            }
            gtiles[g] = b->host_vfirst.back();
        }
        b->gview[nsteps] = (int)k;
        b->ngroups = (int)nsteps;
        int cus = 0;
        HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
        const int render_grid = cus * render_batch_blocks_per_cu(ftype, 1, bx * by, b->render_group);
        // Descriptors of this run (the graph's kernels read them from device
        // memory).  They are uploaded only when they differ from the last
        // upload, after the previous run has finished reading them; the
        // copies come from pageable host vectors, so the upload is waited
        // for.  A batch run again with unchanged views, parameters and
        // transfer function only enqueues its graph.
        const unsigned char* views_bytes = reinterpret_cast<const unsigned char*>(b->host_views.data());
        const size_t views_size = sizeof(RenderLaunch) * nv;
        const bool same_desc = b->exec && b->up_jobs == b->host_jobs && b->up_first == b->host_first &&
                               b->up_vfirst == b->host_vfirst && b->up_views.size() == views_size &&
                               std::memcmp(b->up_views.data(), views_bytes, views_size) == 0;
        if (!same_desc) {
            HIPCHK(c, hipStreamSynchronize(c->stream));
            HIPCHK(c, b->jobs.ensure(b->host_jobs.size()));
            HIPCHK(c, b->first.ensure(sizeof(int) * b->host_first.size()));
            HIPCHK(c, b->rviews.ensure(views_size));
            HIPCHK(c, b->vfirst.ensure(sizeof(int) * b->host_vfirst.size()));
            HIPCHK(c, hipMemcpyAsync(b->jobs.p, b->host_jobs.data(), b->host_jobs.size(), hipMemcpyHostToDevice,
                                     c->stream));
            HIPCHK(c, hipMemcpyAsync(b->first.p, b->host_first.data(), sizeof(int) * b->host_first.size(),
                                     hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(b->rviews.p, views_bytes, views_size, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(b->vfirst.p, b->host_vfirst.data(), sizeof(int) * b->host_vfirst.size(),
                                     hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            b->up_jobs = b->host_jobs;
            b->up_first = b->host_first;
            b->up_vfirst = b->host_vfirst;
            b->up_views.assign(views_bytes, views_bytes + views_size);
        }
        std::vector<long long> key = {(long long)nsteps, (long long)nv, render_grid, bx, by, b->render_group, ftype, same,
                                      (long long)(intptr_t)b->jobs.p, (long long)(intptr_t)b->first.p,
                                      (long long)(intptr_t)b->rviews.p, (long long)(intptr_t)b->vfirst.p,
                                      (long long)(intptr_t)b->ctr.p};
        for (size_t g = 0; g < nsteps; ++g) {
            key.push_back(b->order[g]);
            key.push_back(b->gview[g + 1] - b->gview[g]);
            key.push_back(gtiles[g]);
            key.push_back(b->host_first[2 * b->order[g] + 1]);
        }
        const bool speed = c->params.speed_of_flow != 0;
        key.push_back(speed);
        if (speed) {
            // the uncached LIC kernels take their arguments by value: a graph
            // holds the arguments it was captured with
            for (size_t i = 0; i + 8 <= b->host_jobs.size(); i += 8) {
                long long w;
                std::memcpy(&w, b->host_jobs.data() + i, 8);
                key.push_back(w);
            }
        }
        if (!b->exec || key != b->graph_key) {
            release_graph(b);
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = launch_stamp(t0, c->stream);
            if (!e) e = launch_fill_u32(ctr, 0u, (long long)(head + 4 * nv), c->stream);
            if (!e) e = launch_fill_u32((uint32_t*)begun, 0xffffffffu, (long long)(2 * nv), c->stream);
            for (size_t g = 0; g < nsteps && !e; ++g) {
                const int t = b->order[g];
                LicBatchLaunch LB;
                LB.jobs = (const char*)b->jobs.p + (size_t)t * asz;
                LB.first = (const int*)b->first.p + 2 * t;
                LB.njobs = 1;
                LB.total = b->host_first[2 * t + 1];
                LB.field_type = ftype;
                LB.same = same;
                e = speed ? launch_lic(lics[t], c->stream) : launch_lic_batch(LB, c->stream);
                const int nvg = b->gview[g + 1] - b->gview[g];
                if (e || nvg == 0) continue;
                RenderBatchLaunch RB;
                RB.views = (const RenderLaunch*)b->rviews.p + b->gview[g];
                RB.first = (const int*)b->vfirst.p + b->gfirst[g];
                RB.nviews = nvg;
                RB.total = gtiles[g];
                RB.counter = ctr + g;
                RB.done = done + b->gview[g];
                RB.begun = begun + b->gview[g];
                RB.block[0] = bx;
                RB.block[1] = by;
                RB.group = b->render_group;
                RB.grid = render_grid;
                RB.field_type = ftype;
                RB.lc = 1;
                // the views of this step, after its LIC volume
                e = launch_render_batch(RB, c->stream);
            }
            hipGraph_t gr = nullptr;
            hipError_t e2 = hipStreamEndCapture(c->stream, &gr);
            if (e || e2) {
                if (gr) (void)hipGraphDestroy(gr);
                return fail(c, VV_EDEVICE, "graph capture: %s", hipGetErrorString(e ? e : e2));
            }
            b->graph = gr;
            HIPCHK(c, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->graph_key = key;
        }
        HIPCHK(c, hipGraphLaunch(b->exec, c->stream));
        b->ran = true;
        return VV_OK;
    });
}

vv_status vv_batch_view_latency(vv_batch* b, double* ms)
{
    return guarded((b ? b->ctx : nullptr), [&]() -> vv_status {
        if (!b || !ms) return VV_EINVAL;
        vv_ctx* c = b->ctx;
        if (!b->ran) return fail(c, VV_ESTATE, "batch has not run");
        const size_t nv = b->views.size();
        std::vector<unsigned long long> h(nv + 1);
        const uint32_t* ctr = (const uint32_t*)b->ctr.p;
        HIPCHK(c, hipMemcpyAsync(h.data(), ctr + ctr_head(b->steps.size()) + 2 * nv,
                                 sizeof(unsigned long long) * (nv + 1), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        int khz = 0;
        HIPCHK(c, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
        if (khz <= 0) return fail(c, VV_EUNSUPPORTED, "no wall-clock rate");
        const unsigned long long t0 = h[nv];
        for (size_t v = 0; v < nv; ++v) ms[v] = (double)(h[b->slot[v]] - t0) / (double)khz;
        return VV_OK;
    });
}

vv_status vv_batch_view_span(vv_batch* b, double* ms)
{
    return guarded((b ? b->ctx : nullptr), [&]() -> vv_status {
        if (!b || !ms) return VV_EINVAL;
        vv_ctx* c = b->ctx;
        if (!b->ran) return fail(c, VV_ESTATE, "batch has not run");
        const size_t nv = b->views.size();
        // done[nv], t0, begun[nv]
        std::vector<unsigned long long> h(2 * nv + 1);
        const uint32_t* ctr = (const uint32_t*)b->ctr.p;
        HIPCHK(c, hipMemcpyAsync(h.data(), ctr + ctr_head(b->steps.size()) + 2 * nv,
                                 sizeof(unsigned long long) * (2 * nv + 1), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        int khz = 0;
        HIPCHK(c, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
        if (khz <= 0) return fail(c, VV_EUNSUPPORTED, "no wall-clock rate");
        for (size_t v = 0; v < nv; ++v) {
            const size_t k = b->slot[v];
            ms[v] = (double)(h[k] - h[nv + 1 + k]) / (double)khz;
        }
        return VV_OK;
    });
}

vv_status vv_batch_samples(vv_batch* b, uint64_t* total)
{
    return guarded((b ? b->ctx : nullptr), [&]() -> vv_status {
        if (!b || !total) return VV_EINVAL;
        vv_ctx* c = b->ctx;
        if (!b->ran) return fail(c, VV_ESTATE, "batch has not run");
        const size_t nv = b->views.size();
        std::vector<unsigned long long> h(nv);
        HIPCHK(c, hipMemcpyAsync(h.data(), (const uint32_t*)b->ctr.p + ctr_head(b->steps.size()),
                                 sizeof(unsigned long long) * nv, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        uint64_t t = 0;
        for (unsigned long long x : h) t += x;
        *total = t;
        return VV_OK;
    });
}

vv_status vv_batch_download_lic(vv_batch* b, int timestep, float* out, int on_device)
{
    return guarded((b ? b->ctx : nullptr), [&]() -> vv_status {
        if (!b) return VV_EINVAL;
        vv_ctx* c = b->ctx;
        if (timestep < 0 || timestep >= (int)b->steps.size()) return fail(c, VV_EINVAL, "timestep out of range");
        if (!out) return fail(c, VV_EINVAL, "out is NULL");
        BatchStep* st = b->steps[timestep];
        const size_t bytes = sizeof(float) * (size_t)nvox(st->dims);
        float* dst = out;
        if (!on_device) {
            HIPCHK(c, c->stage.ensure(bytes));
            dst = (float*)c->stage.p;
        }
        HIPCHK(c, launch_unpad_repeat_f32((const float*)st->lic.p, 1, st->dims, dst, c->stream));
        if (!on_device) HIPCHK(c, hipMemcpyAsync(out, dst, bytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return VV_OK;
    });
}

vv_status vv_batch_autotune(vv_batch* b)
{
    return guarded((b ? b->ctx : nullptr), [&]() -> vv_status {
        if (!b) return VV_EINVAL;
        vv_ctx* c = b->ctx;
        if (b->steps.empty()) return fail(c, VV_ESTATE, "batch has no timesteps");
        HIPCHK(c, hipSetDevice(c->device));
        vv_status s = ensure_tf(c);
        if (!s) s = ensure_weights(c);
        if (s) return s;
        // LIC: tile shape (256 threads) x brick edge in voxels (super-tile)
        static const int tiles[][3] = {{16, 4, 4}, {16, 8, 2}, {16, 16, 1}, {32, 4, 2}, {8, 8, 4}};
        static const int bricks[] = {16, 32, 64};
        for (BatchStep* st : b->steps) {
            const TuneKey key{st->dims[0], st->dims[1], st->dims[2], st->storage, c->params.steps_forward,
                              c->params.steps_backward};
            {
                std::lock_guard<std::mutex> lk(g_tune_mu);
                auto it = g_lic_tune.find(key);
                if (it != g_lic_tune.end()) {
                    for (int a = 0; a < 3; ++a) {
                        st->tile[a] = it->second[a];
                        st->super[a] = it->second[3 + a];
                    }
                    continue;
                }
            }
            float best = 1e30f;
            std::array<int, 6> pick{16, 4, 4, 2, 8, 8};
            for (const auto& t : tiles)
                for (int edge : bricks) {
                    for (int a = 0; a < 3; ++a) {
                        st->tile[a] = t[a];
                        st->super[a] = std::max(1, edge / t[a]);
                    }
                    LicLaunch L;
                    s = build_lic_launch(c, st, L);
                    if (s) return s;
                    hipError_t e = hipSuccess;
                    const float ms = time_ms(c->stream, [&] { return launch_lic(L, c->stream); }, 3, &e);
                    if (e) return fail(c, VV_EDEVICE, "autotune LIC: %s", hipGetErrorString(e));
                    if (ms < best) {
                        best = ms;
                        pick = {t[0], t[1], t[2], st->super[0], st->super[1], st->super[2]};
                    }
                }
            for (int a = 0; a < 3; ++a) {
                st->tile[a] = pick[a];
                st->super[a] = pick[3 + a];
            }
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_lic_tune[key] = pick;
        }
        // the ray caster, on the first view: one lane per pixel over five pixel
        // tiles, or 2 or 4 lanes per ray
        if (!b->views.empty()) {
            static const int blocks[][2] = {{16, 16}, {32, 8}, {64, 4}, {16, 4}, {32, 4}};
            float best = 1e30f;
            int pick[3] = {16, 16, 4};
            for (const auto& bl : blocks) {
                b->render_block[0] = bl[0];
                b->render_block[1] = bl[1];
                RenderLaunch R;
                s = build_view_launch(b, b->views[0], R);
                if (s) return s;
                hipError_t e = hipSuccess;
                const float ms = time_ms(c->stream, [&] { return launch_render(R, c->stream); }, 3, &e);
                if (e) return fail(c, VV_EDEVICE, "autotune render: %s", hipGetErrorString(e));
                if (ms < best) {
                    best = ms;
                    pick[0] = bl[0];
                    pick[1] = bl[1];
                    pick[2] = 1;
                }
            }
            for (int G : {2, 4}) {
                RenderLaunch R;
                s = build_view_launch(b, b->views[0], R);
                if (s) return s;
                hipError_t e = hipSuccess;
                const float ms = time_ms(c->stream, [&] { return launch_render_rg(R, G, c->stream); }, 3, &e);
                if (e) return fail(c, VV_EDEVICE, "autotune render: %s", hipGetErrorString(e));
                if (ms < best) {
                    best = ms;
                    pick[2] = G;
                }
            }
            b->render_block[0] = pick[0];
            b->render_block[1] = pick[1];
            b->render_group = pick[2];
        }
        release_graph(b);
        return VV_OK;
    });
}

vv_status vv_batch_get_tuning(const vv_batch* b, int timestep, int out8[8])
{
    return guarded(nullptr, [&]() -> vv_status {
        if (!b || !out8 || timestep < 0 || timestep >= (int)b->steps.size()) return VV_EINVAL;
        const BatchStep* st = b->steps[timestep];
        for (int a = 0; a < 3; ++a) {
            out8[a] = st->tile[a];
            out8[3 + a] = st->super[a];
        }
        const RgTile t = rg_tile(b->render_group);
        out8[6] = b->render_group > 1 ? t.TW : b->render_block[0];
        out8[7] = b->render_group > 1 ? t.TH : b->render_block[1];
        return VV_OK;
    });
}

}  // extern "C"
