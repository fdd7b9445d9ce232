/*
 * vvlic.h — C ABI of the MI355X-native 3D LIC + ray-cast engine.
 *
 * This ABI replaces the reference's C++ entry points on the LIC-volume and
 * ray-cast path (reference paths relative to VectorVisualization/, "VV/"):
 *   VectorDataSet::loadData + createTextureIterp   VV/dataset.cpp:97-180, 290-366, 533-635
 *   NoiseDataSet::loadData / createTexture          VV/dataset.cpp:1124-1344
 *   computeGradients / filterGradients / quantize   VV/gradient.cpp:190-532
 *   LICFilter::createBoxFilter / loadData           VV/dataset.cpp:1405-1512
 *   TransferEdit::updateTextures                    VV/transferEdit.cpp:480-543
 *   LICParams + Renderer::setRenderVolParams        VV/types.h:91-109, VV/renderer.cpp:925-996
 *   Renderer::updateLICVolume / renderLICVolume     VV/renderer.cpp:1311-1379
 *   Renderer::render / raycastLICVolume             VV/renderer.cpp:126-312, 1381-1405
 *   Camera::setCamera, Renderer::updateLightPos     VV/camera.cpp:42-68, VV/renderer.cpp:431-466
 *
 * Conventions
 *   - Every call returns vv_status; VV_OK is 0.  A failing call leaves the
 *     context usable and stores a message readable with vv_last_error().
 *     Nothing aborts or exits.
 *   - Volumes are x-fastest arrays of dims[0]*dims[1]*dims[2] voxels; vector
 *     fields are 3 interleaved float32 per voxel (VV/reader.cpp:290-295).
 *   - Images are width*height RGBA float32, row 0 at the bottom (OpenGL
 *     window convention).
 *   - Pointer arguments marked "host or device" take a memory-kind flag
 *     `on_device` (0 = host memory, 1 = device memory of the context's GPU).
 *   - All device work is ordered on the context's stream (vv_stream());
 *     calls that return data to host memory synchronise that stream.  Device
 *     inputs are read, and device outputs written, on that stream: a caller
 *     producing or consuming them on another stream orders the two streams
 *     itself (hipEventRecord + hipStreamWaitEvent), before and after the call.
 *   - A context is not thread-safe; use one context per host thread.
 */
#ifndef VVLIC_H
#define VVLIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VVLIC_ABI_VERSION 2

typedef enum vv_status {
    VV_OK = 0,
    VV_EINVAL = 1,        /* bad argument (size, pointer, NaN, range) */
    VV_ENOMEM = 2,        /* device or host allocation failed */
    VV_EDEVICE = 3,       /* HIP runtime error */
    VV_EUNSUPPORTED = 4,  /* valid request the library does not implement */
    VV_ESTATE = 5,        /* required input (field, noise, TF, ...) not loaded */
    VV_EIO = 6            /* file missing, unreadable or malformed */
} vv_status;

typedef enum vv_storage {
    VV_STORE_FP32 = 0,    /* 4 x float32 per voxel: encoded direction + |v|/max */
    VV_STORE_FP16 = 1,    /* 4 x float16 per voxel, FP32 arithmetic */
    VV_STORE_UNORM8 = 2   /* 4 x UNORM8 per voxel, the reference's GL_RGBA field
                           * (VectorDataSet::fillTexDataCharInterp,
                           * VV/dataset.cpp:704-823), FP32 arithmetic */
} vv_storage;

typedef enum vv_filter_kind {
    VV_FILTER_BOX = 0,
    VV_FILTER_TRIANGLE = 1,
    VV_FILTER_COS2 = 2,
    VV_FILTER_GAUSSIAN = 3
} vv_filter_kind;

typedef enum vv_noise_kind {
    VV_NOISE_BINARY = 0,  /* {0,255}, 255 with probability `density` */
    VV_NOISE_UNIFORM = 1  /* dense uniform bytes */
} vv_noise_kind;

typedef enum vv_illum {
    VV_ILLUM_LIC = 0,      /* illumLIC, VV/shader/inc_illum.glsl:158-172 */
    VV_ILLUM_GRADIENT = 1, /* illumGradient, VV/shader/inc_illum.glsl:1-41 */
    VV_ILLUM_MALLO = 2,    /* illumMallo, VV/shader/inc_illum.glsl:46-113 */
    VV_ILLUM_ZOECKLER = 3  /* illumZoeckler, VV/shader/inc_illum.glsl:117-154 */
} vv_illum;

/* Material of the line-illumination tables (LineMat, VV/illumination.h:40-62). */
typedef struct vv_line_material {
    float light_color[4], ambient[4], diffuse[4], specular[4];
    float spec_exp;   /* 40 */
    float diff_exp;   /* 2 */
} vv_line_material;

typedef enum vv_tf_index {
    VV_TF_INDEX_B = 0,    /* field .b, as both reference ray casters do */
    VV_TF_INDEX_MAG = 1   /* field .a = |v|/max|v| ("vector-length TF") */
} vv_tf_index;

/* Mirrors LICParams (VV/types.h:91-109) plus the shader switches the
 * reference selects with #defines.  vv_default_lic_params() fills the
 * reference defaults (VV/types.h:93-98). */
typedef struct vv_lic_params {
    float step_size_vol;   /* ray step, default 1/128; alphaCorrection = 128*step */
    float gradient_scale;  /* LIC intensity scale, default 30 */
    float illum_scale;     /* default 1 */
    float freq_scale;      /* noise frequency, default 1 */
    int   num_iterations;  /* ray loop bound n (n*n samples max), default 255 */
    int   steps_forward;   /* F, default 32 */
    int   steps_backward;  /* B, default 32 */
    float step_size_lic;   /* default 0.01 */
    float eye_factor;      /* (0.5*logEyeDist + 0.3); logEyeDist is never set: 0.3 */
    int   illum_mode;      /* vv_illum; VV_ILLUM_GRADIENT needs RGBA noise */
    int   tf_index;        /* vv_tf_index */
    int   gate_enable;     /* scalar gate of freqSampling (needs vv_upload_scalar) */
    float gate_lo;         /* 0.1 */
    float gate_hi;         /* 0.3 */
    int   speed_of_flow;   /* SPEED_OF_FLOW (inc_lic.glsl:107-110,120-122, key '6' in
                              VV/3DLIC.cpp:429-432): both Heun displacements of a
                              step scale with the field magnitude |v|/max at the
                              step's start point.  Default 0 */
} vv_lic_params;

/* Mirrors Camera (VV/camera.cpp:42-68) and the light set by
 * Renderer::updateLightPos (VV/renderer.cpp:431-466). */
typedef struct vv_camera {
    int   width, height;
    float fovy_deg;        /* 35 */
    float dist;            /* 4 */
    float pos[3];          /* camera translation, (0,0,0) */
    float quat[4];         /* orientation x,y,z,w, identity */
    float light_pos[3];    /* eye-space light position (gl_LightSource[0].position) */
    float light_diffuse;   /* 1 */
    float light_specular;  /* 1 */
    float light_ambient;   /* 0 */
    int   spec_exponent;   /* GL_SPOT_EXPONENT, 40 (VV/3DLIC.cpp:736) */
} vv_camera;

typedef struct vv_ctx vv_ctx;

/* ---- context ---------------------------------------------------- */
vv_status vv_create(vv_ctx** out, int device);
vv_status vv_destroy(vv_ctx* ctx);
/* Message of the last failing call on ctx (or of vv_create when ctx is NULL). */
const char* vv_last_error(const vv_ctx* ctx);
int vv_abi_version(void);
/* hipStream_t the context orders its work on. */
void* vv_stream(vv_ctx* ctx);
vv_status vv_synchronize(vv_ctx* ctx);
vv_status vv_default_lic_params(vv_lic_params* p);
/* camera defaults for a volume centred at center: fovy 35, dist 4, light at
 * center + (0,0,1) (light distance 1, VV/3DLIC.cpp:681) */
vv_status vv_default_camera(vv_camera* cam, int width, int height, const float center[3]);

/* ---- a1: vector field (VectorDataSet::loadData + createTextureIterp) ----
 * v0, v1: 3*N float32 (host or device); v1 may be NULL (no time lerp).
 * v = v0 + lerp_t*(v1 - v0) is encoded on the GPU as
 * rgb = 0.5 v/|v| + 0.5, a = |v|/max|v| and kept in `storage`
 * (fillTexDataFloatInterp, VV/dataset.cpp:533-635).  VV_STORE_UNORM8 keeps
 * the bytes of fillTexDataCharInterp instead: rgb = (unsigned char)
 * (v/|v|*127 + 128), 128 for |v| < 1e-5, a = (unsigned char)(|v|/max*255),
 * sampled as byte/255.
 * spacing = DAT slice distances; extent/scale/center follow
 * VV/dataset.cpp:144-176.
 * Volumes (field, noise, scalar, LIC) take 2..4096 samples per axis and fewer
 * than 2^29 in total, so padded element indices stay 32-bit. */
vv_status vv_upload_field(vv_ctx* ctx, const float* v0, const float* v1, float lerp_t,
                          const int dims[3], const float spacing[3], int storage,
                          int on_device);
/* Field encodings of vv_upload_field_ex. */
typedef enum vv_field_encoding {
    VV_ENCODE_INTERP = 0,        /* a = |v|/max|v| (fillTexDataFloatInterp / CharInterp), as vv_upload_field */
    VV_ENCODE_RAW_MAGNITUDE = 1  /* a = |v| unnormalised (fillTexDataFloat, VV/dataset.cpp:448-529, the
                                  * encoding of VectorDataSet::createTexture :213-288); FP32 / FP16 only */
} vv_field_encoding;
/* vv_upload_field with a choice of field encoding. */
vv_status vv_upload_field_ex(vv_ctx* ctx, const float* v0, const float* v1, float lerp_t,
                             const int dims[3], const float spacing[3], int storage, int encoding,
                             int on_device);
/* extent[3], scale[3], center[3] of the loaded field (VolumeData). */
vv_status vv_field_geometry(const vv_ctx* ctx, float extent[3], float scale[3],
                            float center[3]);
/* Encoded field, 4*N values (float32 for every storage; UNORM8 bytes b
 * come back as b * (1/255)). */
vv_status vv_download_field(vv_ctx* ctx, float* out4, int on_device);

/* ---- a2/a3: noise (NoiseDataSet) --------------------------------------
 * noise: N bytes (host or device).  with_gradients != 0 builds the RGBA
 * texture (grad.xyz, noise) of the reference's -g mode, computing the
 * gradients on the GPU (computeGradients + filterGradients + quantize8).
 * grad_rgb (3N bytes) may supply precomputed gradients instead (.grd cache). */
vv_status vv_upload_noise(vv_ctx* ctx, const uint8_t* noise, const int dims[3],
                          int with_gradients, const uint8_t* grad_rgb, int on_device);
/* Seeded noise generated on the GPU (SplitMix64 of seed and voxel index). */
vv_status vv_generate_noise(vv_ctx* ctx, uint64_t seed, int kind, double density,
                            const int dims[3], int with_gradients);
/* N bytes (1 channel) or 4N bytes (RGBA: grad.xyz, noise). */
vv_status vv_download_noise(vv_ctx* ctx, uint8_t* out, int on_device);
vv_status vv_noise_info(const vv_ctx* ctx, int dims[3], int* channels);

/* ---- a4: LIC filter kernel (LICFilter) -------------------------------- */
/* Synthesised 256-entry LUT: box (createBoxFilter), triangle, cos^2,
 * Gaussian exp(-s^2/(2 sigma^2)), s = (2t-1)*length (sigma <= 0: length/4). */
vv_status vv_make_filter(int kind, int length, float sigma, uint8_t lut[256],
                         int* width, float* inv_area);
/* Any LUT (e.g. a filter PNG row, VV/dataset.cpp:1416-1467): width bytes
 * (width a power of two <= 4096).  inv_area <= 0 computes
 * 0.5*W*255/sum (calcFilterKernelInvArea). */
vv_status vv_set_filter(vv_ctx* ctx, const uint8_t* lut, int width, float inv_area);

/* ---- a9: transfer function (TransferEdit::updateTextures) ------------- */
/* rgba: 256x4 float32, opacity: 256 float32 (the LA texture's alpha). */
vv_status vv_set_tf(vv_ctx* ctx, const float* rgba, const float* opacity);
/* 256x5 bytes R,G,B,A,LIC-opacity as TransferEdit::_tfData. */
vv_status vv_set_tf_u8(vv_ctx* ctx, const uint8_t* tf5);
/* default table of TransferEdit's constructor (VV/transferEdit.cpp:76-82) */
vv_status vv_default_tf_u8(uint8_t tf5[256 * 5]);

/* ---- scalar gate volume (VolumeDataSet, freqSampling) ----------------- */
vv_status vv_upload_scalar(vv_ctx* ctx, const uint8_t* scalar, const int dims[3],
                           int on_device);

/* ---- a11: parameters -------------------------------------------------- */
vv_status vv_set_lic_params(vv_ctx* ctx, const vv_lic_params* p);
vv_status vv_get_lic_params(const vv_ctx* ctx, vv_lic_params* p);

/* ---- a5/a6: LIC volume (Renderer::updateLICVolume) -------------------- */
/* Computes the LIC volume at the field's resolution into the context's
 * volume buffer: 1 float per voxel (scalar noise) or 4 (RGBA noise).
 * With scalar noise on the field's grid, freq_scale 1 and no scalar gate,
 * the first call after a field or noise upload also builds a LIC copy of the
 * field that carries the noise bytes (one more field-sized buffer, kept until
 * the next upload); the LIC then reads no separate noise volume.  Without
 * memory for it, the LIC reads the noise volume itself (same results). */
vv_status vv_compute_lic(vv_ctx* ctx);
vv_status vv_lic_info(const vv_ctx* ctx, int dims[3], int* channels);
vv_status vv_download_lic(vv_ctx* ctx, float* out, int on_device);
/* Replace the LIC volume with caller data (channels 1 or 4). */
vv_status vv_upload_lic(vv_ctx* ctx, const float* lic, int channels, const int dims[3],
                        int on_device);

/* ---- a7/a8/a10: ray cast (Renderer::render, raycastLICVolume) --------- */
/* image: width*height*4 float32 (host or device).  samples_out, if not
 * NULL, receives the total number of compositing samples (host). */
vv_status vv_render(vv_ctx* ctx, const vv_camera* cam, float* image, int on_device,
                    uint64_t* samples_out);

/* Renderer::raycastVolume with lic3d_fragment.glsl:5-99 (VOLIC_RAYCAST, the reference's
 * default mode, VV/renderer.cpp:34,184-185): no LIC volume; every ray sample computes the
 * LIC of its position on the fly, then shades and composites it as vv_render does.  A ray
 * stops when it leaves the box or when the sample's own alpha exceeds 0.95 (src.a, :91),
 * not the accumulated alpha.  Needs a field and noise; uses the LIC parameters, filter,
 * TF, illumination mode and (when set) the opacity scale. */
vv_status vv_render_fused(vv_ctx* ctx, const vv_camera* cam, float* image, int on_device,
                          uint64_t* samples);

/* Renderer::sliceVolume with lic3d_slicing_fragment.glsl (the render-target path): view-
 * aligned slices (ViewSlicing, VV/slicing.cpp:42-281) front to back, D / stepSizeVol + 1
 * of them over the box's depth range D.  A pixel's sample on a slice is where its ray meets
 * the slice plane inside the box.  It is skipped when the accumulated alpha has reached
 * 0.95 or its TF alpha is at most 0.05 (no LIC computed); otherwise the LIC of its position
 * is computed, shaded and composited, and the result rounded to FP16 (the RGBA16F render
 * target) before the next slice.  The TF is indexed by the field's .a (|v|/max), as the
 * slicing shaders do (lic3d_slicing_fragment.glsl:43), whatever tf_index says.
 * samples: the composited samples. */
vv_status vv_render_slices(vv_ctx* ctx, const vv_camera* cam, float* image, int on_device,
                           uint64_t* samples);

/* Renderer::sliceVolume without render targets, the reference's default
 * (_useFBO(false), VV/renderer.cpp:33,1149-1157,1233-1258), with
 * lic3d_slicingblend_fragment.glsl: the same slices, each fragment's colour
 * (rgb premultiplied; 0 where the TF alpha is at most 0.05, with no LIC
 * computed) blended into an 8-bit RGBA framebuffer cleared to 0 with
 * glBlendFunc(GL_ONE_MINUS_DST_ALPHA, GL_ONE), i.e. dst = src (1 - dst.a) +
 * dst, no 0.95 test; then a white quad blended the same way (the
 * background).  The framebuffer is fixed point: the fragment colour is
 * clamped to [0,1] and every blend result stored as round(255 x).  image:
 * width*height*4 bytes (RGBA).  samples: the blended fragments. */
vv_status vv_render_slices_blend(vv_ctx* ctx, const vv_camera* cam, uint8_t* image, int on_device,
                                 uint64_t* samples);

/* ---- differentiation ---------------------------------------------------
 * Per-voxel opacity scale sigma (N float32 at the LIC-volume resolution,
 * host or device) multiplies each sample's alpha before opacity
 * correction.  NULL removes it (sigma == 1). */
vv_status vv_set_opacity_scale(vv_ctx* ctx, const float* sigma, int on_device);
/* Gradient of a loss L w.r.t. the 256x4 RGBA transfer function and the
 * opacity scale, given dL/dimage (width*height*4).  dtf: 256*4 floats,
 * dsigma: N floats or NULL.  Checkpoints every 64 samples.  Reductions are
 * deterministic: contributions are rounded to a fixed point 2^-k and summed
 * as 64-bit integers.  k is chosen per call on the device from a bound on
 * every contribution (from max |dL/dimage| and the forward pass) and the
 * number of ray samples, so that no sum can overflow; scaling dL/dimage by
 * a power of two scales every gradient by exactly that power.  A non-finite
 * dL/dimage entry on a pixel the volume covers makes every gradient NaN.
 * The reference has no adjoint; this differentiates the compositing loop
 * of raycast_lic3d_fragment.glsl:35-69. */
vv_status vv_render_adjoint(vv_ctx* ctx, const vv_camera* cam, const float* dimage,
                            float* dtf, float* dsigma, int on_device);

/* vv_render that also keeps the adjoint's forward pass (per-segment state
 * of every ray, sample counts, per-ray bounds) in the context.  No
 * counterpart in the reference: a training step that renders and then
 * differentiates the same view calls this instead of vv_render, and the
 * following vv_render_adjoint with the same camera skips its own forward
 * pass.  The image and the sample count are vv_render's bit for bit.  Any
 * change to the scene (field, noise, LIC volume or parameters, transfer
 * function, opacity scale, line material) or another camera makes the next
 * vv_render_adjoint run the forward pass again; the gradients are the same
 * either way.  The pass keeps 352 bytes per ray and 64 samples of its
 * longest possible ray (about 2.5 GB at 1024^2 and sample distance 1/256);
 * above 8 GiB it keeps checkpoints every 64 samples instead (about 0.6 KB
 * per ray), the adjoint re-runs the segments, and a vv_render_adjoint uses
 * the recorded pass up. */
vv_status vv_render_record(vv_ctx* ctx, const vv_camera* cam, float* image, int on_device,
                           uint64_t* samples);

/* Per-ray sample counts of the forward pass vv_render_record keeps:
 * width*height int32 (row-major, as the image), the number of samples each
 * ray composited before it left the volume or its opacity passed 0.95 (its
 * early-termination index).  No counterpart in the reference: a diagnostic
 * (ray cost map) and the check of a finite difference, which is only
 * comparable with the adjoint on rays whose index does not change.
 * VV_ESTATE when there is no recorded pass (none yet, or an adjoint has
 * used it up) or it is not of this camera. */
vv_status vv_record_ray_samples(vv_ctx* ctx, const vv_camera* cam, int32_t* counts, int on_device);

/* vv_render_adjoint without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  dimage,
 * dtf and dsigma (or NULL) are device pointers.  The kernels' scratch
 * (checkpoints, segment columns, sample counts, fixed-point accumulators)
 * lives in a static device array of 128 MiB that every GPU the library runs
 * on holds until the process ends; a call whose view and opacity-scale grid
 * need more returns VV_EINVAL (about 0.6 KB per pixel for rays of up to
 * 512 samples, plus 8 B per opacity-scale voxel: a 384x384 view at sample
 * distance 1/256 with an opacity scale on a 128^3 grid fits; views small
 * enough for a column of 352 B per 64 samples of every ray keep those
 * instead of checkpoints, as vv_render_adjoint does).  The results
 * are those of vv_render_adjoint bit for bit.  The call is stream-ordered
 * and returns without waiting, except that a transfer function changed
 * since the last call is first uploaded with a blocking copy; line-
 * illumination tables must already be on the device (VV_ESTATE otherwise:
 * render once with the material).  Calls that overlap on one GPU, from any
 * context, race on its array. */
vv_status vv_render_adjoint_noalloc(vv_ctx* ctx, const vv_camera* cam, const float* dimage,
                                    float* dtf, float* dsigma);

/* ---- batching: many timesteps and views in one captured launch --------
 * A batch belongs to a context and uses its LIC parameters, filter and
 * transfer function at run time.  Timesteps are resident volumes (field +
 * seeded scalar noise + LIC volume) of any size; views render one timestep
 * each into a caller-owned device image (width*height*4 float32).
 * vv_batch_run computes every LIC volume (one block per tile) and every view
 * (a persistent kernel over pixel tiles), launches captured once into a
 * hipGraph and replayed on later runs.  When the views, LIC parameters,
 * filter and transfer function are those of its previous run it only
 * enqueues the graph and returns without waiting; otherwise it first waits
 * for the context's stream (the previous run reads the device descriptors),
 * uploads the new descriptors and waits for that upload.  Replaces the
 * reference's per-frame re-encode + LIC recompute of unsteady data
 * (VV/3DLIC.cpp:129-172). */
typedef struct vv_batch vv_batch;
vv_status vv_batch_create(vv_ctx* ctx, vv_batch** out);
vv_status vv_batch_destroy(vv_batch* b);
vv_status vv_batch_add_timestep(vv_batch* b, const float* v, const int dims[3],
                                const float spacing[3], int storage, uint64_t noise_seed,
                                int noise_kind, double density, int on_device);
vv_status vv_batch_add_view(vv_batch* b, int timestep, const vv_camera* cam, float* image);
/* re-encode time step `timestep` from new vectors (same dims; v1 != NULL lerps v0 + lerp_t
 * (v1 - v0)), so an unsteady series advances without rebuilding the batch or its graph */
vv_status vv_batch_set_field(vv_batch* b, int timestep, const float* v0, const float* v1,
                             float lerp_t, int on_device);
vv_status vv_batch_info(const vv_batch* b, int* ntimesteps, int* nviews);
/* times LIC tile shapes x brick (super-tile) edges per timestep size, and
 * the ray caster on the first view (one lane per pixel over five pixel
 * tiles, or 2 or 4 lanes per ray); picks the fastest (LIC choices cached per
 * size).  Before it runs, a batch renders with 4 lanes per ray. */
vv_status vv_batch_autotune(vv_batch* b);
/* tile x,y,z, brick tiles x,y,z of a timestep, then the ray caster's pixel
 * tile x,y: 16 x 8 with 2 lanes per ray, 8 x 8 with 4, any other shape with
 * one lane per pixel */
vv_status vv_batch_get_tuning(const vv_batch* b, int timestep, int out8[8]);
vv_status vv_batch_run(vv_batch* b);
/* per view: milliseconds from the start of the last run to the end of that
 * view's last pixel tile (device wall clock); waits for the run */
vv_status vv_batch_view_latency(vv_batch* b, double* ms);
/* per view: milliseconds from the start of that view's first pixel tile to
 * the end of its last (its own render span, without the time it waited in
 * the queue behind LIC volumes and earlier views); waits for the run */
vv_status vv_batch_view_span(vv_batch* b, double* ms);
/* total compositing samples of the last run (waits) */
vv_status vv_batch_samples(vv_batch* b, uint64_t* total);
vv_status vv_batch_download_lic(vv_batch* b, int timestep, float* out, int on_device);

/* ---- line illumination (illum_mode VV_ILLUM_MALLO / VV_ILLUM_ZOECKLER) -----
 * Both modes shade with 256x256 tables built from a line material: Zoeckler et al.
 * (diffuse, specular) over (L.T, V.T), Mallo et al. diffuse and specular (Simpson rule,
 * 20 intervals), computed in double and rounded to FP16 as the reference's float
 * textures hold them (Illumination::createIllumTextures, VV/illumination.cpp:46-320;
 * VV/3DLIC.cpp:735).  The tangent is the field direction at the sample. */
void vv_default_line_material(vv_line_material* m);
vv_status vv_set_line_material(vv_ctx* ctx, const vv_line_material* m);
vv_status vv_get_line_material(const vv_ctx* ctx, vv_line_material* m);
/* the tables of a material (host): zoeckler 256*256*2, mallo_diff and mallo_spec
 * 256*256*3 floats, row y column x at y*256 + x */
vv_status vv_make_illum_tables(const vv_line_material* m, float* zoeckler, float* mallo_diff,
                               float* mallo_spec);

/* ---- the reference's file formats -------------------------------------
 * Functions without a context report errors through vv_last_error(NULL). */
typedef enum vv_raw_type {
    VV_RAW_NONE = 0,
    VV_RAW_UCHAR = 1,
    VV_RAW_USHORT = 2,
    VV_RAW_FLOAT = 3
} vv_raw_type;

/* a parsed .dat descriptor (DatFile, VV/reader.h:38-77) */
typedef struct vv_dat_info {
    char raw_pattern[1024]; /* ObjectFileName: printf pattern of the time step, resolved
                               against the .dat file's directory when not found as given */
    int dims[3];            /* Resolution */
    float spacing[3];       /* SliceThickness (default 1 1 1) */
    int raw_type;           /* Format: vv_raw_type */
    int components;         /* Format's trailing count: FLOAT3 -> 3 (default 1) */
    int t_begin, t_end;     /* TimeDependent (default 0 0) */
} vv_dat_info;

/* DatFile::parseDatFile (VV/reader.cpp:81-258): also checks every time step's RAW file exists */
vv_status vv_dat_parse(const char* dat_path, vv_dat_info* info);
/* DatFile::readRawData (VV/reader.cpp:261-297): raw bytes of one time step */
vv_status vv_dat_read(const vv_dat_info* info, int timestep, void* out, size_t out_bytes);
/* one time step of a 3-component data set as float32 (UCHAR components c -> c - 128) */
vv_status vv_dat_read_field(const vv_dat_info* info, int timestep, float* out3);
/* VectorDataSet::loadData + loadTimeStep + createTextureIterp (VV/dataset.cpp:97-176,
 * 195-210, 290-366): reads `timestep` (and `next_timestep` when lerp_t != 0) and uploads
 * the field lerped by lerp_t, as vv_upload_field */
vv_status vv_load_field_dat(vv_ctx* ctx, const char* dat_path, int timestep, int next_timestep,
                            float lerp_t, int storage);
/* VolumeDataSet::loadData + createTexture (VV/dataset.cpp:836-1000) for the scalar gate:
 * UCHAR as is, FLOAT as UNORM8 round(clamp(v,0,1)*255) */
vv_status vv_load_scalar_dat(vv_ctx* ctx, const char* dat_path);
/* noise file: 3 x int32 sizes, then the bytes (NoiseDataSet::loadRawData,
 * VV/dataset.cpp:1347-1389).  out == NULL reads the sizes only. */
vv_status vv_noise_file_read(const char* path, int dims[3], uint8_t* out, size_t out_bytes);
vv_status vv_noise_file_write(const char* path, const uint8_t* noise, const int dims[3]);
/* NoiseDataSet::loadData + createTexture (VV/dataset.cpp:1124-1344): with gradients, the
 * 3-byte-per-voxel cache "<path>.grd" is read when present and valid, otherwise the
 * gradients are computed on the GPU and, with use_gradient_cache, written there
 * (loadGradients / saveGradients, VV/gradient.cpp:112-187) */
vv_status vv_load_noise_file(vv_ctx* ctx, const char* path, int with_gradients,
                             int use_gradient_cache);
/* NoiseDataSet::loadData with its fallback (VV/dataset.cpp:1124-1172): when the
 * noise file cannot be read (path NULL, missing, short or malformed), 256^3
 * binary white noise is generated instead, 255 with probability 1/6 as the
 * reference's floor(0.6 rand()/(RAND_MAX+1) + 0.5) * 255 (:1158-1162), from
 * `seed` (vv_generate_noise).  *generated (may be NULL) is 1 when it was. */
vv_status vv_load_noise_or_generate(vv_ctx* ctx, const char* path, int with_gradients,
                                    int use_gradient_cache, uint64_t seed, int* generated);
/* pngRead (VV/imageUtils.cpp:19-96): 8-bit gray, gray+alpha, RGB or RGBA, rows in file
 * order.  out == NULL reads the header only. */
vv_status vv_png_read(const char* path, uint8_t* out, size_t out_bytes, int* width, int* height,
                      int* channels);
/* pngWrite (VV/imageUtils.cpp:99-123); flip_rows writes the last row first */
vv_status vv_png_write(const char* path, const uint8_t* pixels, int width, int height, int channels,
                       int flip_rows);
/* a rendered image (RGBA float32, bottom row first) as an 8-bit RGBA PNG, the
 * screenshot path of the reference (UNORM8 conversion, rows flipped) */
vv_status vv_image_write_png(const char* path, const float* rgba, int width, int height);
/* LICFilter::loadData (VV/dataset.cpp:1420-1467): first row, first channel, centred in
 * the next power of two; invFilterArea computed (vv_set_filter with inv_area 0) */
vv_status vv_load_filter_png(vv_ctx* ctx, const char* path);
/* TransferEdit::loadTF / saveTF (VV/transferEdit.cpp:113-334) on the 256x5 byte table of
 * vv_set_tf_u8: "<stem>_rgba.png" and "<stem>_alpha.png".  Loading keeps the reference's
 * order: the alpha/opacity file is read only when the RGBA file cannot be; channels a
 * file does not hold keep their values in tf5. */
vv_status vv_tf_load_png(const char* name, uint8_t tf5[1280]);
vv_status vv_tf_save_png(const char* name, const uint8_t tf5[1280]);

/* ---- unsteady fields: the animation loop -------------------------------
 * All time steps of a series stay on the GPU as raw vectors.  Frame k (k = 0, 1, ...)
 * encodes into the context's field the lerp of steps (p mod n) and (p+1 mod n),
 * p = k / interp_size, at t = (k mod interp_size) / interp_size: the sequence of the
 * reference's init + idle loop (VV/3DLIC.cpp:700-708, :129-172; checkInterpolateStage,
 * VV/dataset.cpp:202-210; InterpSize 10 there).  The caller then runs vv_compute_lic and
 * vv_render as for a static field. */
typedef struct vv_anim vv_anim;
vv_status vv_anim_create(vv_ctx* ctx, const float* const* steps, int nsteps, const int dims[3],
                         const float spacing[3], int interp_size, int storage, int on_device,
                         vv_anim** out);
/* all time steps of a time-dependent DAT series (TimeDependent: begin end) */
vv_status vv_anim_load_dat(vv_ctx* ctx, const char* dat_path, int interp_size, int storage,
                           vv_anim** out);
/* encode the next frame; reports the series index of its first step and its lerp factor */
vv_status vv_anim_next(vv_anim* a, int* step, float* lerp_t);
vv_status vv_anim_seek(vv_anim* a, long long frame);
vv_status vv_anim_destroy(vv_anim* a);

/* ---- launch tuning (used by the autotuner) ----------------------------
 * keys: "lic.tile_x", "lic.tile_y", "lic.tile_z" (powers of two whose
 * product is 256 threads per block), "lic.super_x", "lic.super_y",
 * "lic.super_z" (tiles per super-tile: the 3D-blocked order in which tiles
 * are handed to the GPU), "lic.cache" (0/1: per-lane trilinear cell
 * caches), "lic.brick" (0/1: with an FP32 field the LIC volume kernel reads
 * a copy of the field in 8^3 bricks; measured no faster on the 256^3
 * workload, off by default), "adjoint.window" (0/1: accumulate the opacity-scale gradient in
 * an LDS window per ray segment before the global integer atomics),
 * "render.volume_group" (lanes per ray of vv_render, a power of two up to
 * 64; 1 = one lane per pixel; 0 = 4 for views of 2^18 rays or more, about
 * 2^20 lanes in total below; the forward pass of vv_render_record and of the
 * adjoints and the adjoints' backward walk take the same number up to 4;
 * gradients do not depend on it), "render.block_x", "render.block_y" (pixel
 * tile of vv_render with one lane per pixel, 64 to 256 threads in
 * multiples of 64), "render.ray_group" (lanes per ray of vv_render_fused /
 * vv_render_slices, a power of two up to 64; 0 = enough lanes for about
 * 2^21 in total with an FP32 field and 2^22 with FP16 or UNORM8, at most 32
 * per ray).  Images and sample counts do not depend on the lanes per ray.
 * Results do not depend on these settings. */
vv_status vv_set_tuning(vv_ctx* ctx, const char* key, int value);
vv_status vv_get_tuning(const vv_ctx* ctx, const char* key, int* value);

/* ---- helpers exposed for tests --------------------------------------- */
/* The 24 floats + 2 ints of the per-frame ray setup (see vv_view_block in
 * the implementation): cam[3] du[3] dv[3] dw[3] inv_w2 inv_h2 extent[3]
 * scale[3] tex_max[3], then width, height. */
vv_status vv_view_block(const vv_ctx* ctx, const vv_camera* cam, float out24[24]);
/* Convolution weights k(u_i) for the current filter and steps: 1+B+F floats. */
vv_status vv_lic_weights(const vv_ctx* ctx, float* out);

#ifdef __cplusplus
}
#endif
#endif
