/*
 * vvoracle — CPU restatement of the reference's 3D-LIC + ray-cast path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * vectorvisualization_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links, loads
 * or falls back to it.
 *
 * Every function exists in two precisions built from one source
 * (vvoracle_impl.h): *_f32 reproduces the operation order used by the GPU
 * kernels (compiled with -ffp-contract=off, explicit fmaf where the kernels
 * use fma), and *_f64 is the same algorithm in double, used as the error
 * reference for the FP16-storage variant and the adjoint.
 *
 * Reference citations (paths relative to the reference checkout,
 * VV = VectorVisualization/):
 *   field encode      VV/dataset.cpp:533-635   (fillTexDataFloatInterp)
 *   noise gradients   VV/gradient.cpp:190-532  (computeGradients, filterGradients, quantize8)
 *   filter LUT        VV/dataset.cpp:1405-1512 (createBoxFilter, loadData, calcFilterKernelInvArea)
 *   LIC               VV/shader/inc_lic.glsl:134-265, VV/shader/lic3d_volume_fragment.glsl:2-20
 *   ray cast          VV/shader/raycast_lic3d_fragment.glsl:5-73
 *   shading           VV/shader/inc_illum.glsl:1-41,158-172
 *   uniforms          VV/renderer.cpp:925-996
 *   camera            VV/camera.cpp:42-68, VV/renderer.cpp:126-147,682-736
 *
 * Pinning status of each function is listed in the header comment of
 * vvoracle.c.
 */
#ifndef VVORACLE_H
#define VVORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* LIC + render parameters (mirror of LICParams, VV/types.h:91-109, plus the
 * uniform values derived in VV/renderer.cpp:967-990). */
typedef struct orc_params {
    float step_size_vol;    /* stepSizeVol */
    float gradient_scale;   /* gradient.r */
    float illum_scale;      /* gradient.g */
    float freq_scale;       /* gradient.b */
    int   num_iterations;   /* numIterations */
    int   steps_forward;    /* licParams.x */
    int   steps_backward;   /* licParams.y */
    float step_size_lic;    /* licParams.z */
    float eye_factor;       /* (0.5*logEyeDist + 0.3) with logEyeDist == 0 */
    int   illum_mode;       /* 0 = illumLIC, 1 = illumGradient */
    int   tf_index;         /* 0 = field .b (reference), 1 = field .a (|v|/max) */
    int   gate_enable;      /* scalar gate of freqSampling */
    float gate_lo, gate_hi; /* 0.1, 0.3 */
    float inv_filter_area;  /* LICFilter::getInverseFilterArea() */
    float alpha_correction; /* stepSizeVol*128 */
    float light[3];         /* gl_LightSource[0].position.xyz (eye space) */
    float light_diffuse, light_specular, light_ambient;
    int   spec_exponent;    /* GL_SPOT_EXPONENT = 40 */
    /* line-illumination tables (orc_illum_luts), read by illum_mode 2 and 3 */
    const float* lut_zoeckler;   /* 256x256 x (diffuse, specular) */
    const float* lut_mallo_diff; /* 256x256 x rgb */
    const float* lut_mallo_spec; /* 256x256 x rgb */
    /* 1: the field array holds UNORM8 byte values (0..255, orc_encode_field_u8);
     * blended field values are scaled by 1/255 like every UNORM8 texture */
    int   field_unorm8;
    /* SPEED_OF_FLOW (inc_lic.glsl:107-110,120-122): both displacements of a
     * Heun step scaled by the field's .a at the step's start point */
    int   speed_of_flow;
} orc_params;

/* Line material of the illumination tables (LineMat, VV/illumination.h:40-62). */
typedef struct orc_line_mat {
    float light_color[4], ambient[4], diffuse[4], specular[4];
    float spec_exp, diff_exp;
} orc_line_mat;

/* Illumination textures: Zoeckler et al. (VV/illumination.cpp:96-183) and
 * Mallo et al. (:188-320, Simpson rule :348-390), computed in double and
 * rounded to FP16 as the reference's RGBA16F / LUMINANCE_ALPHA16F
 * textures store them.  Row y, column x at index y*256 + x. */
void orc_illum_luts(const orc_line_mat* m, float* zoeckler /* 256*256*2 */,
                    float* mallo_diff /* 256*256*3 */, float* mallo_spec /* 256*256*3 */);
double orc_mallo_spec_term(double alpha, double beta, double n);

/* Per-frame view block: every float the per-pixel ray setup consumes.
 * Computed in double from the camera and rounded once to float. */
typedef struct orc_view {
    float cam[3];      /* camera position in object space, gl_ModelViewMatrixInverse[3] */
    float du[3];       /* object-space ray-direction increment per unit NDC x */
    float dv[3];       /* per unit NDC y */
    float dw[3];       /* direction through NDC (0,0) */
    float inv_w2, inv_h2; /* 2/width, 2/height */
    float extent[3];   /* bounding box [0,extent] in object space */
    float scale[3];    /* scaleVol */
    float tex_max[3];  /* extent*scale */
    int   width, height;
} orc_view;

/* ---- integer / byte work (precision independent) ---- */
uint16_t orc_f32_to_f16(float f);
float    orc_f16_to_f32(uint16_t h);
void     orc_f32_to_f16_array(const float* in, uint16_t* out, long n);
void     orc_f16_to_f32_array(const uint16_t* in, float* out, long n);

void orc_gen_noise(uint64_t seed, int kind, double density,
                   int nx, int ny, int nz, uint8_t* out);

int  orc_noise_gradients(const uint8_t* noise, int nx, int ny, int nz,
                         const float spacing[3], uint8_t* grad_rgb);
void orc_noise_gradients_raw(const uint8_t* noise, int nx, int ny, int nz,
                             const float spacing[3], float* grad_xyz);
void orc_filter_gradients(int nx, int ny, int nz, float* grad_xyz);
void orc_quantize_gradients(long n, float* grad_xyz, uint8_t* grad_rgb);

int  orc_make_filter(int kind, int length, double sigma,
                     uint8_t lut[256], int* width, float* inv_area);
float orc_filter_inv_area(const uint8_t* lut, int width);

void orc_view_setup(double fovy_deg, int width, int height, double dist,
                    const double pos[3], const double quat[4],
                    const double center[3], const double extent[3],
                    const double scale[3], orc_view* out);
void orc_light_position(const double center[3], const double cam_pos[3],
                        const double cam_quat[4], double light_dist,
                        const double light_quat[4], double out[3]);

/* ---- floating point work, two precisions ---- */
void orc_encode_field_f32(const float* v0, const float* v1, float t,
                          int nx, int ny, int nz, float* out4);
/* fillTexDataFloat (VV/dataset.cpp:448-529): as orc_encode_field_* but
 * a = |v| unnormalised */
void orc_encode_field_raw_f32(const float* v0, const float* v1, float t,
                              int nx, int ny, int nz, float* out);
void orc_encode_field_raw_f64(const float* v0, const float* v1, double t,
                              int nx, int ny, int nz, double* out);
/* UNORM8 field (fillTexDataCharInterp, VV/dataset.cpp:704-823): 4 bytes per
 * voxel, direction bytes and |v|/max as a byte; see vvoracle.c */
void orc_encode_field_u8(const float* v0, const float* v1, float t,
                         int nx, int ny, int nz, uint8_t* out);
void orc_encode_field_f64(const float* v0, const float* v1, double t,
                          int nx, int ny, int nz, double* out4);

void orc_lic_weights_f32(const uint8_t* lut, int width, int steps_fwd, int steps_bwd,
                         float* w /* 1 + B + F */);
void orc_lic_weights_f64(const uint8_t* lut, int width, int steps_fwd, int steps_bwd,
                         double* w);

void orc_lic_f32(const float* field4, int nx, int ny, int nz, const float scale[3],
                 const uint8_t* noise, int noise_ch, const int ndims[3],
                 const uint8_t* scalar, const int sdims[3],
                 const uint8_t* lut, int lut_width, const orc_params* p,
                 int z0, int z1, int nthreads, float* out);
void orc_lic_f64(const double* field4, int nx, int ny, int nz, const float scale[3],
                 const uint8_t* noise, int noise_ch, const int ndims[3],
                 const uint8_t* scalar, const int sdims[3],
                 const uint8_t* lut, int lut_width, const orc_params* p,
                 int z0, int z1, int nthreads, double* out);

void orc_lic_box_f32(const float* field4, int nx, int ny, int nz, const float scale[3],
                     const uint8_t* noise, int noise_ch, const int ndims[3],
                     const uint8_t* scalar, const int sdims[3],
                     const uint8_t* lut, int lut_width, const orc_params* p,
                     const int box[6], int nthreads, float* out);
void orc_lic_box_f64(const double* field4, int nx, int ny, int nz, const float scale[3],
                     const uint8_t* noise, int noise_ch, const int ndims[3],
                     const uint8_t* scalar, const int sdims[3],
                     const uint8_t* lut, int lut_width, const orc_params* p,
                     const int box[6], int nthreads, double* out);

long orc_raycast_f32(const float* field4, const int fdims[3],
                     const float* lic, int lic_ch, const int ldims[3],
                     const float* tf_rgba, const float* tf_opac,
                     const float* sigma, const orc_params* p, const orc_view* v,
                     int nthreads, float* image, int* nsamples);
long orc_raycast_f64(const double* field4, const int fdims[3],
                     const double* lic, int lic_ch, const int ldims[3],
                     const double* tf_rgba, const double* tf_opac,
                     const double* sigma, const orc_params* p, const orc_view* v,
                     int nthreads, double* image, int* nsamples);

/* Fused ray caster (lic3d_fragment.glsl:5-99): the LIC of every sample's
 * position computed on the fly; sigma (optional) on a grid of sigma_dims. */
long orc_raycast_fused_f32(const float* field4, const int fdims[3],
                           const uint8_t* noise, int noise_ch, const int ndims[3],
                           const uint8_t* scalar, const int sdims[3],
                           const uint8_t* lut, int lut_width,
                           const float* tf_rgba, const float* tf_opac,
                           const float* sigma, const int sigma_dims[3],
                           const orc_params* p, const orc_view* v,
                           int nthreads, float* image, int* nsamples);
long orc_raycast_fused_f64(const double* field4, const int fdims[3],
                           const uint8_t* noise, int noise_ch, const int ndims[3],
                           const uint8_t* scalar, const int sdims[3],
                           const uint8_t* lut, int lut_width,
                           const double* tf_rgba, const double* tf_opac,
                           const double* sigma, const int sigma_dims[3],
                           const orc_params* p, const orc_view* v,
                           int nthreads, double* image, int* nsamples);

/* View-aligned slicing (Renderer::sliceVolume, lic3d_slicing_fragment.glsl):
 * the LIC of every slice sample on the fly, FP16 render target. */
long orc_render_slices_f32(const float* field4, const int fdims[3],
                           const uint8_t* noise, int noise_ch, const int ndims[3],
                           const uint8_t* scalar, const int sdims[3],
                           const uint8_t* lut, int lut_width,
                           const float* tf_rgba, const float* tf_opac,
                           const float* sigma, const int sigma_dims[3],
                           const orc_params* p, const orc_view* v,
                           int nthreads, float* image, int* nsamples);
long orc_render_slices_f64(const double* field4, const int fdims[3],
                           const uint8_t* noise, int noise_ch, const int ndims[3],
                           const uint8_t* scalar, const int sdims[3],
                           const uint8_t* lut, int lut_width,
                           const double* tf_rgba, const double* tf_opac,
                           const double* sigma, const int sigma_dims[3],
                           const orc_params* p, const orc_view* v,
                           int nthreads, double* image, int* nsamples);
/* The same slices blended into an 8-bit RGBA framebuffer (the reference's
 * default output, lic3d_slicingblend_fragment.glsl): image holds 4 bytes per
 * pixel. */
long orc_render_slices_blend_f32(const float* field4, const int fdims[3],
                                 const uint8_t* noise, int noise_ch, const int ndims[3],
                                 const uint8_t* scalar, const int sdims[3],
                                 const uint8_t* lut, int lut_width,
                                 const float* tf_rgba, const float* tf_opac,
                                 const float* sigma, const int sigma_dims[3],
                                 const orc_params* p, const orc_view* v,
                                 int nthreads, uint8_t* image, int* nsamples);
long orc_render_slices_blend_f64(const double* field4, const int fdims[3],
                                 const uint8_t* noise, int noise_ch, const int ndims[3],
                                 const uint8_t* scalar, const int sdims[3],
                                 const uint8_t* lut, int lut_width,
                                 const double* tf_rgba, const double* tf_opac,
                                 const double* sigma, const int sigma_dims[3],
                                 const orc_params* p, const orc_view* v,
                                 int nthreads, uint8_t* image, int* nsamples);

/* Reverse-mode derivative of orc_raycast_* with respect to tf_rgba (256x4)
 * and the per-voxel opacity scale sigma (ldims voxels), for a loss whose
 * image gradient is dimage (width*height*4).  Stores every ray's samples,
 * then walks each ray backwards. */
void orc_raycast_adjoint_f64(const double* field4, const int fdims[3],
                             const double* lic, int lic_ch, const int ldims[3],
                             const double* tf_rgba, const double* tf_opac,
                             const double* sigma, const orc_params* p, const orc_view* v,
                             const double* dimage, double* dtf, double* dsigma);

#ifdef __cplusplus
}
#endif
#endif
