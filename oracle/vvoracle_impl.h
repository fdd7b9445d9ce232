/*
 * vvoracle_impl.h — precision-dependent half of the oracle, included twice
 * by vvoracle.c (ORC_DOUBLE 0 -> *_f32, ORC_DOUBLE 1 -> *_f64).
 * TEST INFRASTRUCTURE ONLY (see vvoracle.c).
 *
 * Sampling follows OpenGL linear filtering: for a coordinate u over N texels
 * x = u*N - 0.5, i0 = floor(x), f = x - i0, and texels i0, i0+1 after the
 * wrap (CLAMP_TO_EDGE clamps to [0,N-1], REPEAT takes i mod N).  Values are
 * blended as nested lerps, x first, then y, then z, with
 * lerp(a,b,t) = fma(t, b-a, a).  UNORM8 textures blend the raw byte values
 * and scale the result by 1/255.
 */
#if ORC_DOUBLE
#define R double
#define FN(n) n##_f64
#define FMA fma
#define SQRT sqrt
#define FLOOR floor
#define POW pow
#define FMIN fmin
#define FMAX fmax
#else
#define R float
#define FN(n) n##_f32
#define FMA fmaf
#define SQRT sqrtf
#define FLOOR floorf
#define POW powf
#define FMIN fminf
#define FMAX fmaxf
#endif

#define K(x) ((R)(x))

static inline R FN(lerp)(R a, R b, R t) { return FMA(t, b - a, a); }
static inline R FN(clamp01)(R x) { return x < K(0) ? K(0) : (x > K(1) ? K(1) : x); }

static inline int FN(icl)(int i, int n) { return i < 0 ? 0 : (i > n - 1 ? n - 1 : i); }
static inline int FN(irp)(int i, int n) { int r = i % n; return r < 0 ? r + n : r; }

/* texel coordinate and fraction */
static inline int FN(tcoord)(R u, int n, R* f)
{
    R x = FMA(u, (R)n, K(-0.5));
    R fl = FLOOR(x);
    *f = x - fl;
    /* keep the int conversion defined for any float */
    if (!(fl > K(-1e9))) fl = K(-1e9);
    if (fl > K(1e9)) fl = K(1e9);
    return (int)fl;
}

typedef struct { int x0, x1, y0, y1, z0, z1; R fx, fy, fz; } FN(cell);

static inline void FN(cell_clamp)(const R p[3], const int d[3], FN(cell)* c)
{
    int i = FN(tcoord)(p[0], d[0], &c->fx);
    int j = FN(tcoord)(p[1], d[1], &c->fy);
    int k = FN(tcoord)(p[2], d[2], &c->fz);
    c->x0 = FN(icl)(i, d[0]); c->x1 = FN(icl)(i + 1, d[0]);
    c->y0 = FN(icl)(j, d[1]); c->y1 = FN(icl)(j + 1, d[1]);
    c->z0 = FN(icl)(k, d[2]); c->z1 = FN(icl)(k + 1, d[2]);
}

static inline void FN(cell_repeat)(const R p[3], const int d[3], FN(cell)* c)
{
    int i = FN(tcoord)(p[0], d[0], &c->fx);
    int j = FN(tcoord)(p[1], d[1], &c->fy);
    int k = FN(tcoord)(p[2], d[2], &c->fz);
    c->x0 = FN(irp)(i, d[0]); c->x1 = FN(irp)(i + 1, d[0]);
    c->y0 = FN(irp)(j, d[1]); c->y1 = FN(irp)(j + 1, d[1]);
    c->z0 = FN(irp)(k, d[2]); c->z1 = FN(irp)(k + 1, d[2]);
}

#define ORC_IDX(d, x, y, z) ((((long)(z)) * (d)[1] + (y)) * (d)[0] + (x))

/* trilinear blend of channel ch of an nc-channel array of type T */
#define ORC_TRILIN(T, arr, nc, ch, d, c, out)                                      \
    do {                                                                             \
        R v000 = (R)(arr)[ORC_IDX(d, (c).x0, (c).y0, (c).z0) * (nc) + (ch)];        \
        R v100 = (R)(arr)[ORC_IDX(d, (c).x1, (c).y0, (c).z0) * (nc) + (ch)];        \
        R v010 = (R)(arr)[ORC_IDX(d, (c).x0, (c).y1, (c).z0) * (nc) + (ch)];        \
        R v110 = (R)(arr)[ORC_IDX(d, (c).x1, (c).y1, (c).z0) * (nc) + (ch)];        \
        R v001 = (R)(arr)[ORC_IDX(d, (c).x0, (c).y0, (c).z1) * (nc) + (ch)];        \
        R v101 = (R)(arr)[ORC_IDX(d, (c).x1, (c).y0, (c).z1) * (nc) + (ch)];        \
        R v011 = (R)(arr)[ORC_IDX(d, (c).x0, (c).y1, (c).z1) * (nc) + (ch)];        \
        R v111 = (R)(arr)[ORC_IDX(d, (c).x1, (c).y1, (c).z1) * (nc) + (ch)];        \
        R c00 = FN(lerp)(v000, v100, (c).fx);                                        \
        R c10 = FN(lerp)(v010, v110, (c).fx);                                        \
        R c01 = FN(lerp)(v001, v101, (c).fx);                                        \
        R c11 = FN(lerp)(v011, v111, (c).fx);                                        \
        R c0 = FN(lerp)(c00, c10, (c).fy);                                           \
        R c1 = FN(lerp)(c01, c11, (c).fy);                                           \
        (out) = FN(lerp)(c0, c1, (c).fz);                                            \
    } while (0)

/* channel ch of the 4-channel field at cell c: a UNORM8 field holds byte
 * values and its blend is scaled by 1/255 */
#define ORC_FIELD_TRILIN(p, arr, ch, d, c, out)                                   \
    do {                                                                             \
        ORC_TRILIN(R, arr, 4, ch, d, c, out);                                        \
        if ((p)->field_unorm8) (out) = (out) * K(1.0 / 255.0);                       \
    } while (0)

/* 1D linear lookup with CLAMP_TO_EDGE over n entries of an nc-channel table */
static inline R FN(lut1d)(const R* t, int n, int nc, int ch, R u)
{
    R f;
    int i = FN(tcoord)(u, n, &f);
    int i0 = FN(icl)(i, n), i1 = FN(icl)(i + 1, n);
    return FN(lerp)(t[i0 * nc + ch], t[i1 * nc + ch], f);
}

/* 256x256 illumination table, LINEAR + CLAMP_TO_EDGE.  The coordinate is
 * first limited to [0,1]: no change for finite values (the edge texels
 * already saturate there), and +-inf / NaN map to an edge deterministically
 * (fmax(NaN, 0) = 0). */
static inline R FN(lut2d)(const float* t, int nc, int ch, R u, R v)
{
    u = FMIN(FMAX(u, K(0)), K(1));
    v = FMIN(FMAX(v, K(0)), K(1));
    R fx, fy;
    int i = FN(tcoord)(u, 256, &fx), j = FN(tcoord)(v, 256, &fy);
    int i0 = FN(icl)(i, 256), i1 = FN(icl)(i + 1, 256), j0 = FN(icl)(j, 256), j1 = FN(icl)(j + 1, 256);
    R a = FN(lerp)((R)t[(j0 * 256 + i0) * nc + ch], (R)t[(j0 * 256 + i1) * nc + ch], fx);
    R b = FN(lerp)((R)t[(j1 * 256 + i0) * nc + ch], (R)t[(j1 * 256 + i1) * nc + ch], fx);
    return FN(lerp)(a, b, fy);
}

/* byte LUT variant (filter kernel): blend bytes, then scale by 1/255 */
static inline R FN(lut1d_u8)(const uint8_t* t, int n, R u)
{
    R f;
    int i = FN(tcoord)(u, n, &f);
    int i0 = FN(icl)(i, n), i1 = FN(icl)(i + 1, n);
    return FN(lerp)((R)t[i0], (R)t[i1], f) * K(1.0 / 255.0);
}

/* ------------------------------------------------------------------ */
/* Field encode: VV/dataset.cpp:533-635 (fillTexDataFloatInterp).      */
/* v = v0 + t*(v1-v0); rgb = 0.5*v/|v| + 0.5 (|v| < 1e-5 -> 0.5, 0);   */
/* a = clamp(|v|/max|v|, 0, 1).  An all-zero field (max 0) gets a = 0 */
/* (the reference would store 0/0).                                   */
/* ------------------------------------------------------------------ */
/* raw = 0: fillTexDataFloatInterp (VV/dataset.cpp:533-635), a = |v|/max
 * clamped to [0,1]; raw = 1: fillTexDataFloat (:448-529), a = |v| as is
 * (its max normalisation is commented out, :517-525) */
static void FN(encode_field)(const float* v0, const float* v1, R t, int nx, int ny, int nz, int raw, R* out)
{
    long n = (long)nx * ny * nz;
    R maxlen = K(-1);
    if (!v1) v1 = v0;
    for (long i = 0; i < n; ++i) {
        R tv[3];
        for (int c = 0; c < 3; ++c) {
            R a = (R)v0[3 * i + c], b = (R)v1[3 * i + c];
            tv[c] = a + t * (b - a);
        }
        R len = SQRT(tv[0] * tv[0] + tv[1] * tv[1] + tv[2] * tv[2]);
        if (len < K(1e-5f)) {
            len = K(0);
            out[4 * i + 0] = out[4 * i + 1] = out[4 * i + 2] = K(0.5);
        } else {
            for (int c = 0; c < 3; ++c) out[4 * i + c] = K(0.5) * tv[c] / len + K(0.5);
        }
        out[4 * i + 3] = len;
        if (len > maxlen) maxlen = len;
    }
    if (raw) return;
    for (long i = 0; i < n; ++i) {
        R a = maxlen > K(0) ? out[4 * i + 3] / maxlen : K(0);
        out[4 * i + 3] = a > K(1) ? K(1) : (a < K(0) ? K(0) : a);
    }
}

#if ORC_DOUBLE
void orc_encode_field_f64(const float* v0, const float* v1, double t, int nx, int ny, int nz, double* out)
{
    encode_field_f64(v0, v1, t, nx, ny, nz, 0, out);
}
void orc_encode_field_raw_f64(const float* v0, const float* v1, double t, int nx, int ny, int nz, double* out)
{
    encode_field_f64(v0, v1, t, nx, ny, nz, 1, out);
}
#else
void orc_encode_field_f32(const float* v0, const float* v1, float t, int nx, int ny, int nz, float* out)
{
    encode_field_f32(v0, v1, t, nx, ny, nz, 0, out);
}
void orc_encode_field_raw_f32(const float* v0, const float* v1, float t, int nx, int ny, int nz, float* out)
{
    encode_field_f32(v0, v1, t, nx, ny, nz, 1, out);
}
#endif

/* ------------------------------------------------------------------ */
/* LIC kernel weights, one per convolution sample, in sample order:    */
/* w[0] = k(0.5) (inc_lic.glsl:168-172), w[1..B] backward with          */
/* u -= 0.5/B before each step (:178-185), w[B+1..B+F] forward with     */
/* u += 0.5/F (:192-199).  k(u) = LUT(u).r, GL_CLAMP taken as           */
/* CLAMP_TO_EDGE.  0.5/F, 0.5/B are float uniforms (VV/renderer.cpp:983)*/
/* ------------------------------------------------------------------ */
void FN(orc_lic_weights)(const uint8_t* lut, int width, int F, int B, R* w)
{
    const R kf = K(0.5) / (R)F, kb = K(0.5) / (R)B;
    R u = K(0.5);
    w[0] = FN(lut1d_u8)(lut, width, u);
    for (int i = 0; i < B; ++i) {
        u = u - kb;
        w[1 + i] = FN(lut1d_u8)(lut, width, u);
    }
    u = K(0.5);
    for (int i = 0; i < F; ++i) {
        u = u + kf;
        w[1 + B + i] = FN(lut1d_u8)(lut, width, u);
    }
}

/* ------------------------------------------------------------------ */
/* LIC volume: lic3d_volume_fragment.glsl:2-20 + computeLIC            */
/* (inc_lic.glsl:152-202), singleLICstep (inc_lic.glsl:93-145).        */
/* ------------------------------------------------------------------ */
typedef struct {
    const R* field; int fd[3];
    const uint8_t* noise; int nc; int nd[3];
    const uint8_t* scalar; int sd[3];
    const orc_params* p;
} FN(lic_ctx);

/* the field's direction (s[0..2]) and, with SPEED_OF_FLOW, its magnitude
 * .a (s[3]; 1 otherwise) */
static inline void FN(field_rgb)(const FN(lic_ctx)* c, const R x[3], R s[4])
{
    FN(cell) cl;
    FN(cell_clamp)(x, c->fd, &cl);
    ORC_FIELD_TRILIN(c->p, c->field, 0, c->fd, cl, s[0]);
    ORC_FIELD_TRILIN(c->p, c->field, 1, c->fd, cl, s[1]);
    ORC_FIELD_TRILIN(c->p, c->field, 2, c->fd, cl, s[2]);
    s[3] = K(1);
    if (c->p->speed_of_flow) ORC_FIELD_TRILIN(c->p, c->field, 3, c->fd, cl, s[3]);
}

/* freqSampling (inc_lic.glsl:71-90) / freqSamplingGrad (inc_lic.glsl:61-68) */
static inline void FN(noise_at)(const FN(lic_ctx)* c, const R x[3], R n[4])
{
    FN(cell) cl;
    if (c->nc == 4) {
        FN(cell_repeat)(x, c->nd, &cl);
        for (int ch = 0; ch < 4; ++ch) {
            R v;
            ORC_TRILIN(uint8_t, c->noise, 4, ch, c->nd, cl, v);
            n[ch] = v * K(1.0 / 255.0);
        }
        return;
    }
    if (c->p->gate_enable) {
        R s;
        FN(cell_clamp)(x, c->sd, &cl);
        ORC_TRILIN(uint8_t, c->scalar, 1, 0, c->sd, cl, s);
        s = s * K(1.0 / 255.0);
        if (!(s > (R)c->p->gate_lo && s < (R)c->p->gate_hi)) { n[0] = K(0); return; }
    }
    R fs = (R)c->p->freq_scale;
    R q[3] = {x[0] * fs, x[1] * fs, x[2] * fs};
    R v;
    FN(cell_repeat)(q, c->nd, &cl);
    ORC_TRILIN(uint8_t, c->noise, 1, 0, c->nd, cl, v);
    n[0] = v * K(1.0 / 255.0);
}

static void FN(lic_voxel)(const FN(lic_ctx)* c, const R* w, const R seed[3], R out[4])
{
    const orc_params* p = c->p;
    const int nc = c->nc;
    const R he = (R)p->step_size_lic * (R)p->eye_factor;
    const int speed = p->speed_of_flow;
    R s0[4], n[4];
    R acc[4] = {0, 0, 0, 0};
    FN(field_rgb)(c, seed, s0);
    FN(noise_at)(c, seed, n);
    for (int ch = 0; ch < nc; ++ch) acc[ch] = n[ch] * w[0];
    for (int dirn = 0; dirn < 2; ++dirn) {
        const int steps = dirn == 0 ? p->steps_backward : p->steps_forward;
        const R h = dirn == 0 ? -he : he;
        const R* wk = dirn == 0 ? w + 1 : w + 1 + p->steps_backward;
        R x[3] = {seed[0], seed[1], seed[2]};
        R s[4] = {s0[0], s0[1], s0[2], s0[3]};
        for (int i = 0; i < steps; ++i) {
            R d1[3], x2[3], s2[4], d2[3];
            for (int a = 0; a < 3; ++a) {
                R t = K(2) * s[a] - K(1);
                if (speed) t = t * s[3];                     /* :107-110 licdir *= step.a */
                d1[a] = t * h;                               /* :114 */
                x2[a] = x[a] + d1[a];                        /* :115 */
            }
            FN(field_rgb)(c, x2, s2);                        /* :116 */
            for (int a = 0; a < 3; ++a) {
                R t = K(2) * s2[a] - K(1);
                if (speed) t = t * s[3];                     /* :120-122 licdir2 *= step.a (old step) */
                d2[a] = t * h;                               /* :123 */
                x[a] = x[a] + K(0.5) * (d1[a] + d2[a]);      /* :125 */
            }
            FN(field_rgb)(c, x, s);                          /* :128 */
            FN(noise_at)(c, x, n);                           /* :134-139 */
            for (int ch = 0; ch < nc; ++ch) acc[ch] = acc[ch] + n[ch] * wk[i];  /* :142, :183/:197 */
        }
    }
    const R scl = ((R)p->inv_filter_area / (R)(p->steps_forward + p->steps_backward))
                  * (R)p->gradient_scale;
    if (nc == 1) {
        out[0] = acc[0] * scl;                               /* lic3d_volume_fragment.glsl:15 */
    } else {
        out[0] = acc[0]; out[1] = acc[1]; out[2] = acc[2];
        out[3] = acc[3] * scl;                               /* lic3d_fragment.glsl:67 */
    }
}

/* LIC of the voxels in [x0,x1) x [y0,y1) x [z0,z1); out holds the box,
 * x fastest */
void FN(orc_lic_box)(const R* field4, int nx, int ny, int nz, const float scale[3],
                     const uint8_t* noise, int noise_ch, const int ndims[3],
                     const uint8_t* scalar, const int sdims[3],
                     const uint8_t* lut, int lut_width, const orc_params* p,
                     const int box[6], int nthreads, R* out)
{
    FN(lic_ctx) c;
    c.field = field4; c.fd[0] = nx; c.fd[1] = ny; c.fd[2] = nz;
    c.noise = noise; c.nc = noise_ch;
    c.nd[0] = ndims[0]; c.nd[1] = ndims[1]; c.nd[2] = ndims[2];
    c.scalar = scalar;
    c.sd[0] = sdims ? sdims[0] : 1; c.sd[1] = sdims ? sdims[1] : 1; c.sd[2] = sdims ? sdims[2] : 1;
    c.p = p;
    int nw = 1 + p->steps_forward + p->steps_backward;
    R* w = (R*)malloc(sizeof(R) * nw);
    FN(orc_lic_weights)(lut, lut_width, p->steps_forward, p->steps_backward, w);
    int x0 = box[0] < 0 ? 0 : box[0], x1 = box[1] > nx ? nx : box[1];
    int y0 = box[2] < 0 ? 0 : box[2], y1 = box[3] > ny ? ny : box[3];
    int z0 = box[4] < 0 ? 0 : box[4], z1 = box[5] > nz ? nz : box[5];
    const int bw = x1 - x0, bh = y1 - y0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) collapse(2) num_threads(nthreads)
#endif
    for (int z = z0; z < z1; ++z)
        for (int y = y0; y < y1; ++y)
            for (int x = x0; x < x1; ++x) {
                R seed[3] = {((R)x + K(0.5)) / (R)nx * (R)scale[0],
                             ((R)y + K(0.5)) / (R)ny * (R)scale[1],
                             ((R)z + K(0.5)) / (R)nz * (R)scale[2]};
                R o[4];
                FN(lic_voxel)(&c, w, seed, o);
                long idx = (((long)(z - z0)) * bh + (y - y0)) * bw + (x - x0);
                for (int ch = 0; ch < noise_ch; ++ch) out[idx * noise_ch + ch] = o[ch];
            }
    free(w);
}

void FN(orc_lic)(const R* field4, int nx, int ny, int nz, const float scale[3],
                 const uint8_t* noise, int noise_ch, const int ndims[3],
                 const uint8_t* scalar, const int sdims[3],
                 const uint8_t* lut, int lut_width, const orc_params* p,
                 int z0, int z1, int nthreads, R* out)
{
    const int box[6] = {0, nx, 0, ny, z0, z1};
    FN(orc_lic_box)(field4, nx, ny, nz, scale, noise, noise_ch, ndims, scalar, sdims, lut,
                    lut_width, p, box, nthreads, out);
}

/* ------------------------------------------------------------------ */
/* Ray cast: raycast_lic3d_fragment.glsl:5-73 with illumLIC /          */
/* illumGradient (inc_illum.glsl:158-172 / :1-41).                     */
/* ------------------------------------------------------------------ */
static inline R FN(pow_corr)(R b, R c)
{
    if (c == K(1)) return b;
    if (c == K(0.5)) return SQRT(b);
    if (c == K(2)) return b * b;
    return POW(b, c);
}

static inline R FN(ipow)(R x, int e)
{
    R r = K(1), b = x;
    while (e > 0) {
        if (e & 1) r = r * b;
        b = b * b;
        e >>= 1;
    }
    return r;
}

static inline void FN(normalize3)(const R v[3], R o[3])
{
    R d = SQRT(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (d > K(0)) { o[0] = v[0] / d; o[1] = v[1] / d; o[2] = v[2] / d; }
    else { o[0] = o[1] = o[2] = K(0); }
}

/* per-pixel ray setup: returns 0 when the pixel misses the box */
static int FN(ray_setup)(const orc_view* v, int i, int j, R pos[3], R dir[3])
{
    R nxp = FMA((R)i + K(0.5), (R)v->inv_w2, K(-1));
    R nyp = FMA((R)j + K(0.5), (R)v->inv_h2, K(-1));
    R d[3], tn = -INFINITY, tf = INFINITY;
    for (int a = 0; a < 3; ++a) {
        d[a] = FMA(nxp, (R)v->du[a], FMA(nyp, (R)v->dv[a], (R)v->dw[a]));
        R c = (R)v->cam[a], e = (R)v->extent[a];
        if (d[a] != K(0)) {
            R inv = K(1) / d[a];
            R t0 = (K(0) - c) * inv, t1 = (e - c) * inv;
            R lo = t0 < t1 ? t0 : t1, hi = t0 < t1 ? t1 : t0;
            if (lo > tn) tn = lo;
            if (hi < tf) tf = hi;
        } else if (c < K(0) || c > e) {
            return 0;
        }
    }
    if (!(tn <= tf) || !(tn > K(0))) return 0;
    R p0[3], g[3], gn[3];
    for (int a = 0; a < 3; ++a) {
        R q = FMA(tn, d[a], (R)v->cam[a]);
        R e = (R)v->extent[a];
        p0[a] = q < K(0) ? K(0) : (q > e ? e : q);
        g[a] = p0[a] - (R)v->cam[a];
    }
    FN(normalize3)(g, gn);
    for (int a = 0; a < 3; ++a) {
        dir[a] = gn[a] * (R)v->scale[a];
        pos[a] = p0[a] * (R)v->scale[a];
    }
    return 1;
}

typedef struct {
    const R* field; int fd[3];
    const R* lic; int lc; int ld[3];
    const R* tf; const R* op;
    const R* sigma;
    const orc_params* p;
    R light[3];
} FN(rc_ctx);

/* one compositing sample; returns src (premultiplied) and the pieces the
 * adjoint needs */
typedef struct {
    R tfu;      /* TF coordinate */
    R opu;      /* opacity-table coordinate (1.3*l) */
    R m[3];     /* d rgb / d tf.rgb, per channel */
    R spec;     /* additive rgb term (illumGradient) */
    R ao;       /* opacity table value */
    R sg;       /* opacity scale at the sample (1 without sigma) */
    R araw;     /* ao*tf.a*sg */
    R a;        /* corrected alpha */
    R rgb[3];   /* unpremultiplied colour */
    R tf[4];
    int aclamp; /* the opacity-scale clamp was active */
} FN(rc_sample);

/* shading and opacity of one sample from its LIC value(s) l4 (c->lc channels) */
static void FN(rc_shade)(const FN(rc_ctx)* c, const R pos[3], const R dir[3], const R l4[4], FN(rc_sample)* s)
{
    const orc_params* p = c->p;
    FN(cell) cl;
    R v;
    FN(cell_clamp)(pos, c->fd, &cl);
    ORC_FIELD_TRILIN(p, c->field, p->tf_index ? 3 : 2, c->fd, cl, v);
    s->tfu = v;
    for (int ch = 0; ch < 4; ++ch) s->tf[ch] = FN(lut1d)(c->tf, 256, 4, ch, v);
    R g = (R)p->illum_scale;
    R l = c->lc == 1 ? l4[0] : l4[3];
    if (p->illum_mode == 2 || p->illum_mode == 3) {
        /* line illumination with the field direction as tangent
         * (vectorData.xyz; tangent = normalize(2 t - 1)) */
        FN(cell) cf;
        FN(cell_clamp)(pos, c->fd, &cf);
        R tg[3], T[3];
        for (int a = 0; a < 3; ++a) {
            R x;
            ORC_FIELD_TRILIN(p, c->field, a, c->fd, cf, x);
            tg[a] = FMA(K(2), x, K(-1));
        }
        FN(normalize3)(tg, T);
        const R* L = c->light;
        if (p->illum_mode == 3) {
            /* illumZoeckler (inc_illum.glsl:117-154).  The table is
             * LUMINANCE_ALPHA, so .rg reads (L, L): both shading terms are the
             * luminance (diffuse) channel. */
            R V[3];
            FN(normalize3)(dir, V);
            R u = FMA(K(0.5), L[0] * T[0] + L[1] * T[1] + L[2] * T[2], K(0.5));
            R w = FMA(K(0.5), V[0] * T[0] + V[1] * T[1] + V[2] * T[2], K(0.5));
            R lum = FN(lut2d)(p->lut_zoeckler, 2, 0, u, w);
            R shr = lum * g;
            R sp = K(0.9) * lum;
            for (int a = 0; a < 3; ++a) {
                s->rgb[a] = (s->tf[a] * l) * shr + sp;
                s->m[a] = l * shr;
            }
            s->spec = sp;
        } else {
            /* illumMallo (inc_illum.glsl:46-113) */
            R vd[3] = {-dir[0], -dir[1], -dir[2]}, V[3], bn[3], Bv[3], N[3], hv[3], Hh[3];
            FN(normalize3)(vd, V);
            bn[0] = T[1] * V[2] - T[2] * V[1];
            bn[1] = T[2] * V[0] - T[0] * V[2];
            bn[2] = T[0] * V[1] - T[1] * V[0];
            FN(normalize3)(bn, Bv);
            N[0] = Bv[1] * T[2] - Bv[2] * T[1];
            N[1] = Bv[2] * T[0] - Bv[0] * T[2];
            N[2] = Bv[0] * T[1] - Bv[1] * T[0];
            for (int a = 0; a < 3; ++a) hv[a] = V[a] + L[a];
            FN(normalize3)(hv, Hh);
            R lx = L[0] * N[0] + L[1] * N[1] + L[2] * N[2];
            R ly = L[0] * T[0] + L[1] * T[1] + L[2] * T[2];
            R lz = Hh[0] * N[0] + Hh[1] * N[1] + Hh[2] * N[2];
            R lw = Hh[0] * T[0] + Hh[1] * T[1] + Hh[2] * T[2];
            R ql = SQRT(FMAX(K(1) - ly * ly, K(0))), qh = SQRT(FMAX(K(1) - lw * lw, K(0)));
            R tx = K(1) / ql, ty = K(1) / qh;
            R ud = FMA(K(0.5), lx * tx, K(0.5));
            R vdc = FMA(K(0.5), ly, K(0.5));
            R vs = FMA(K(0.5), lz * ty, K(0.5));
            /* pow(tmp.y, -spotExponent) = (1 - (H.T)^2)^(n/2) */
            R att = FN(ipow)(qh, p->spec_exponent);
            for (int a = 0; a < 3; ++a) {
                R dif = FN(lut2d)(p->lut_mallo_diff, 3, a, ud, vdc) * g;
                R spc = FN(clamp01)(FN(lut2d)(p->lut_mallo_spec, 3, a, ud, vs) * att);
                s->rgb[a] = (s->tf[a] * l) * dif + spc;
                s->m[a] = l * dif;
            }
            s->spec = K(0);
        }
    } else if (p->illum_mode == 0 || c->lc == 1) {
        for (int a = 0; a < 3; ++a) s->rgb[a] = l * s->tf[a] * g;
        s->m[0] = s->m[1] = s->m[2] = l * g;
        s->spec = K(0);
    } else {
        R nn[3] = {-l4[0], -l4[1], -l4[2]}, N[3], V[3], Rv[3], rr[3];
        R vd[3] = {-dir[0], -dir[1], -dir[2]};
        FN(normalize3)(nn, N);
        FN(normalize3)(vd, V);
        const R* L = c->light;
        R ldn = L[0] * N[0] + L[1] * N[1] + L[2] * N[2];
        for (int a = 0; a < 3; ++a) rr[a] = FMA(K(2) * ldn, N[a], -L[a]);
        FN(normalize3)(rr, Rv);
        R sp = FN(clamp01)(Rv[0] * V[0] + Rv[1] * V[1] + Rv[2] * V[2]);
        sp = FN(ipow)(sp, p->spec_exponent);
        sp = sp * l * (R)p->light_specular;
        R diff = FN(clamp01)(ldn) * (R)p->light_diffuse * g;
        R kk = (diff + K(0.3)) + (R)p->light_ambient;
        for (int a = 0; a < 3; ++a) {
            R col = s->tf[a] * l;
            s->rgb[a] = FMA(col, kk, sp) * g;
        }
        s->m[0] = s->m[1] = s->m[2] = l * kk * g;
        s->spec = sp * g;
    }
    s->opu = l * K(1.3);
    s->ao = FN(lut1d)(c->op, 256, 1, 0, s->opu);
    R a = s->ao * s->tf[3];
    s->sg = K(1);
    s->aclamp = 0;
    if (c->sigma) {
        /* opacity scale (build addition): alpha * sigma, clamped to [0,1] */
        FN(cell_clamp)(pos, c->ld, &cl);
        ORC_TRILIN(R, c->sigma, 1, 0, c->ld, cl, s->sg);
        a = a * s->sg;
        if (a > K(1)) { a = K(1); s->aclamp = 1; }
        if (a < K(0)) { a = K(0); s->aclamp = 1; }
    }
    s->araw = a;
    s->a = K(1) - FN(pow_corr)(K(1) - a, (R)p->alpha_correction);
}

/* one sample of the LIC-volume ray caster: LIC value(s) from the volume */
static void FN(rc_eval)(const FN(rc_ctx)* c, const R pos[3], const R dir[3], FN(rc_sample)* s)
{
    FN(cell) cl;
    R l4[4] = {0, 0, 0, 0};
    FN(cell_repeat)(pos, c->ld, &cl);
    for (int ch = 0; ch < c->lc; ++ch) ORC_TRILIN(R, c->lic, c->lc, ch, c->ld, cl, l4[ch]);
    FN(rc_shade)(c, pos, dir, l4, s);
}

static inline int FN(outside)(const R pos[3], const orc_view* v)
{
    for (int a = 0; a < 3; ++a)
        if (pos[a] < K(0) || pos[a] > (R)v->tex_max[a]) return 1;
    return 0;
}

long FN(orc_raycast)(const R* field4, const int fdims[3],
                     const R* lic, int lic_ch, const int ldims[3],
                     const R* tf_rgba, const R* tf_opac,
                     const R* sigma, const orc_params* p, const orc_view* v,
                     int nthreads, R* image, int* nsamples)
{
    FN(rc_ctx) c;
    c.field = field4; c.lic = lic; c.lc = lic_ch;
    for (int a = 0; a < 3; ++a) { c.fd[a] = fdims[a]; c.ld[a] = ldims[a]; }
    c.tf = tf_rgba; c.op = tf_opac; c.sigma = sigma; c.p = p;
    R lraw[3] = {(R)p->light[0], (R)p->light[1], (R)p->light[2]};
    FN(normalize3)(lraw, c.light);
    const long maxs = (long)p->num_iterations * (long)p->num_iterations;
    const R step = (R)p->step_size_vol;
    long total = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total) num_threads(nthreads)
#endif
    for (int j = 0; j < v->height; ++j)
        for (int i = 0; i < v->width; ++i) {
            R pos[3], dir[3], dest[4] = {0, 0, 0, 0};
            long pix = (long)j * v->width + i;
            int ns = 0;
            if (FN(ray_setup)(v, i, j, pos, dir)) {
                for (long k = 0; k < maxs; ++k) {
                    FN(rc_sample) s;
                    FN(rc_eval)(&c, pos, dir, &s);
                    R src[4] = {s.rgb[0] * s.a, s.rgb[1] * s.a, s.rgb[2] * s.a, s.a};
                    R t = K(1) - dest[3];
                    for (int a = 0; a < 4; ++a) dest[a] = FN(clamp01)(FMA(t, src[a], dest[a]));
                    ++ns;
                    for (int a = 0; a < 3; ++a) pos[a] = FMA(dir[a], step, pos[a]);
                    if (FN(outside)(pos, v) || dest[3] > K(0.95)) break;
                }
            }
            for (int a = 0; a < 4; ++a) image[4 * pix + a] = dest[a];
            if (nsamples) nsamples[pix] = ns;
            total += ns;
        }
    return total;
}

/* ------------------------------------------------------------------ */
/* Fused ray caster: lic3d_fragment.glsl:5-99.  Every sample computes   */
/* the LIC of its position (computeLIC, scaled as the LIC volume is),   */
/* then shades it as rc_shade; a ray stops when it leaves the box or    */
/* when the sample's own alpha exceeds 0.95 (src.a, :91).               */
/* ------------------------------------------------------------------ */
long FN(orc_raycast_fused)(const R* field4, const int fdims[3],
                           const uint8_t* noise, int noise_ch, const int ndims[3],
                           const uint8_t* scalar, const int sdims[3],
                           const uint8_t* lut, int lut_width,
                           const R* tf_rgba, const R* tf_opac,
                           const R* sigma, const int sigma_dims[3],
                           const orc_params* p, const orc_view* v,
                           int nthreads, R* image, int* nsamples)
{
    FN(lic_ctx) lc;
    lc.field = field4;
    for (int a = 0; a < 3; ++a) { lc.fd[a] = fdims[a]; lc.nd[a] = ndims[a]; lc.sd[a] = sdims ? sdims[a] : 1; }
    lc.noise = noise; lc.nc = noise_ch; lc.scalar = scalar; lc.p = p;
    const int nw = 1 + p->steps_forward + p->steps_backward;
    R* w = (R*)malloc(sizeof(R) * nw);
    FN(orc_lic_weights)(lut, lut_width, p->steps_forward, p->steps_backward, w);
    FN(rc_ctx) c;
    c.field = field4; c.lic = NULL; c.lc = noise_ch;
    for (int a = 0; a < 3; ++a) { c.fd[a] = fdims[a]; c.ld[a] = sigma_dims ? sigma_dims[a] : fdims[a]; }
    c.tf = tf_rgba; c.op = tf_opac; c.sigma = sigma; c.p = p;
    R lraw[3] = {(R)p->light[0], (R)p->light[1], (R)p->light[2]};
    FN(normalize3)(lraw, c.light);
    const long maxs = (long)p->num_iterations * (long)p->num_iterations;
    const R step = (R)p->step_size_vol;
    long total = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total) num_threads(nthreads)
#endif
    for (int j = 0; j < v->height; ++j)
        for (int i = 0; i < v->width; ++i) {
            R pos[3], dir[3], dest[4] = {0, 0, 0, 0};
            long pix = (long)j * v->width + i;
            int ns = 0;
            if (FN(ray_setup)(v, i, j, pos, dir)) {
                for (long k = 0; k < maxs; ++k) {
                    R l4[4];
                    FN(lic_voxel)(&lc, w, pos, l4);
                    FN(rc_sample) s;
                    FN(rc_shade)(&c, pos, dir, l4, &s);
                    R src[4] = {s.rgb[0] * s.a, s.rgb[1] * s.a, s.rgb[2] * s.a, s.a};
                    R t = K(1) - dest[3];
                    for (int a = 0; a < 4; ++a) dest[a] = FN(clamp01)(FMA(t, src[a], dest[a]));
                    ++ns;
                    for (int a = 0; a < 3; ++a) pos[a] = FMA(dir[a], step, pos[a]);
                    if (FN(outside)(pos, v) || s.a > K(0.95)) break;
                }
            }
            for (int a = 0; a < 4; ++a) image[4 * pix + a] = dest[a];
            if (nsamples) nsamples[pix] = ns;
            total += ns;
        }
    free(w);
    return total;
}

/* ------------------------------------------------------------------ */
/* View-aligned slicing: Renderer::sliceVolume + ViewSlicing          */
/* (VV/slicing.cpp:42-281) with lic3d_slicing_fragment.glsl.  Slice k  */
/* is the plane v.(p - centre) = d_k, d_k = -D/2 + (k + 1/2) D/n; the   */
/* pixel's sample is where its ray meets the plane inside the box.      */
/* Skipped when dest.a >= 0.95 or tf.a <= 0.05; dest is stored in an    */
/* RGBA16F target, i.e. rounded to FP16 after every composited slice.  */
/* ------------------------------------------------------------------ */
static inline R FN(half_rn)(R x) { return (R)orc_f16_to_f32(orc_f32_to_f16((float)x)); }

/* The reference's default output (no render targets, VV/renderer.cpp:
 * 1149-1157,1233-1258; lic3d_slicingblend_fragment.glsl): the slices' colours
 * blended into an 8-bit RGBA framebuffer cleared to 0 with
 * glBlendFunc(GL_ONE_MINUS_DST_ALPHA, GL_ONE): dst = src (1 - dst.a) + dst,
 * the fragment colour clamped to [0,1] and each result stored as
 * round(255 x); no 0.95 test; then a white quad blended the same way.  The
 * blend arithmetic is float (the byte values read as b/255). */
static inline float FN(unorm8_rn)(float x)
{
    x = x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
    return floorf(fmaf(x, 255.0f, 0.5f));
}

static inline void FN(blend_u8)(float dst[4], const R src[4])
{
    const float t = 1.0f - dst[3] * (1.0f / 255.0f);
    for (int a = 0; a < 4; ++a) {
        float sc = (float)src[a];
        sc = sc < 0.0f ? 0.0f : (sc > 1.0f ? 1.0f : sc);
        dst[a] = FN(unorm8_rn)(fmaf(t, sc, dst[a] * (1.0f / 255.0f)));
    }
}

/* blend == 0: RGBA16F target into image (R); 1: framebuffer bytes into img8 */
static long FN(render_slices_impl)(const R* field4, const int fdims[3],
                                   const uint8_t* noise, int noise_ch, const int ndims[3],
                                   const uint8_t* scalar, const int sdims[3],
                                   const uint8_t* lut, int lut_width,
                                   const R* tf_rgba, const R* tf_opac,
                                   const R* sigma, const int sigma_dims[3],
                                   const orc_params* p, const orc_view* v,
                                   int nthreads, int blend, R* image, uint8_t* img8, int* nsamples)
{
    /* slicing setup (VV/slicing.cpp:42-100), in float as the reference */
    float sv[3], half[3], ctr[3];
    const float inv = 1.0f / sqrtf(v->dw[0] * v->dw[0] + v->dw[1] * v->dw[1] + v->dw[2] * v->dw[2]);
    for (int a = 0; a < 3; ++a) {
        sv[a] = v->dw[a] * inv;
        if (fabsf(sv[a]) < 1.0e-8f) sv[a] = 0.0f;
        half[a] = 0.5f * v->extent[a];
        ctr[a] = v->extent[a] / 2.0f;
    }
    const float D = 2.0f * (fabsf(sv[0]) * half[0] + fabsf(sv[1]) * half[1] + fabsf(sv[2]) * half[2]);
    const int nsl = (int)(D / p->step_size_vol) + 1;
    FN(lic_ctx) lc;
    lc.field = field4;
    for (int a = 0; a < 3; ++a) { lc.fd[a] = fdims[a]; lc.nd[a] = ndims[a]; lc.sd[a] = sdims ? sdims[a] : 1; }
    lc.noise = noise; lc.nc = noise_ch; lc.scalar = scalar; lc.p = p;
    const int nw = 1 + p->steps_forward + p->steps_backward;
    R* w = (R*)malloc(sizeof(R) * nw);
    FN(orc_lic_weights)(lut, lut_width, p->steps_forward, p->steps_backward, w);
    /* both slicing shaders index the TF with vectorData.a
     * (lic3d_slicing_fragment.glsl:43, lic3d_slicingblend_fragment.glsl:40) */
    orc_params ps = *p;
    ps.tf_index = 1;
    FN(rc_ctx) c;
    c.field = field4; c.lic = NULL; c.lc = noise_ch;
    for (int a = 0; a < 3; ++a) { c.fd[a] = fdims[a]; c.ld[a] = sigma_dims ? sigma_dims[a] : fdims[a]; }
    c.tf = tf_rgba; c.op = tf_opac; c.sigma = sigma; c.p = &ps;
    R lraw[3] = {(R)p->light[0], (R)p->light[1], (R)p->light[2]};
    FN(normalize3)(lraw, c.light);
    long total = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total) num_threads(nthreads)
#endif
    for (int j = 0; j < v->height; ++j)
        for (int i = 0; i < v->width; ++i) {
            const long pix = (long)j * v->width + i;
            const R nxp = FMA((R)i + K(0.5), (R)v->inv_w2, K(-1));
            const R nyp = FMA((R)j + K(0.5), (R)v->inv_h2, K(-1));
            R d[3], rc[3], dest[4] = {0, 0, 0, 0};
            for (int a = 0; a < 3; ++a) {
                d[a] = FMA(nxp, (R)v->du[a], FMA(nyp, (R)v->dv[a], (R)v->dw[a]));
                rc[a] = (R)v->cam[a] - (R)ctr[a];
            }
            const R vdir = (R)sv[0] * d[0] + (R)sv[1] * d[1] + (R)sv[2] * d[2];
            const R vcam = (R)sv[0] * rc[0] + (R)sv[1] * rc[1] + (R)sv[2] * rc[2];
            int ns = 0;
            float fb[4] = {0.0f, 0.0f, 0.0f, 0.0f};  /* framebuffer bytes (blend) */
            for (int k = 0; k < nsl && vdir != K(0); ++k) {
                if (!blend && !(dest[3] < K(0.95))) break;
                if (blend && fb[3] == 255.0f) break;  /* dst.a = 1: later slices add nothing */
                const float dkf = -0.5f * D + ((float)k + 0.5f) * D / (float)nsl;
                const R sk = ((R)dkf - vcam) / vdir;
                if (!(sk > K(0))) continue;
                R q[3], g[3], gn[3], pos[3], dir[3];
                int in = 1;
                for (int a = 0; a < 3; ++a) {
                    q[a] = FMA(sk, d[a], (R)v->cam[a]);
                    in = in && q[a] >= K(0) && q[a] <= (R)v->extent[a];
                }
                if (!in) continue;
                for (int a = 0; a < 3; ++a) g[a] = q[a] - (R)v->cam[a];
                FN(normalize3)(g, gn);
                for (int a = 0; a < 3; ++a) {
                    dir[a] = gn[a] * (R)v->scale[a];
                    pos[a] = q[a] * (R)v->scale[a];
                }
                FN(cell) cl;
                R tfu;
                FN(cell_clamp)(pos, fdims, &cl);
                ORC_FIELD_TRILIN(p, field4, 3, fdims, cl, tfu);
                if (!(FN(lut1d)(tf_rgba, 256, 4, 3, tfu) > K(0.05))) continue;
                R l4[4];
                FN(lic_voxel)(&lc, w, pos, l4);
                FN(rc_sample) s;
                FN(rc_shade)(&c, pos, dir, l4, &s);
                R src[4] = {s.rgb[0] * s.a, s.rgb[1] * s.a, s.rgb[2] * s.a, s.a};
                if (blend) {
                    FN(blend_u8)(fb, src);
                } else {
                    R t = K(1) - dest[3];
                    for (int a = 0; a < 4; ++a) dest[a] = FN(half_rn)(FN(clamp01)(FMA(t, src[a], dest[a])));
                }
                ++ns;
            }
            if (blend) {
                const R white[4] = {K(1), K(1), K(1), K(1)};
                FN(blend_u8)(fb, white);
                for (int a = 0; a < 4; ++a) img8[4 * pix + a] = (uint8_t)fb[a];
            } else {
                for (int a = 0; a < 4; ++a) image[4 * pix + a] = dest[a];
            }
            if (nsamples) nsamples[pix] = ns;
            total += ns;
        }
    free(w);
    return total;
}

long FN(orc_render_slices)(const R* field4, const int fdims[3],
                           const uint8_t* noise, int noise_ch, const int ndims[3],
                           const uint8_t* scalar, const int sdims[3],
                           const uint8_t* lut, int lut_width,
                           const R* tf_rgba, const R* tf_opac,
                           const R* sigma, const int sigma_dims[3],
                           const orc_params* p, const orc_view* v,
                           int nthreads, R* image, int* nsamples)
{
    return FN(render_slices_impl)(field4, fdims, noise, noise_ch, ndims, scalar, sdims, lut, lut_width, tf_rgba,
                                  tf_opac, sigma, sigma_dims, p, v, nthreads, 0, image, NULL, nsamples);
}

long FN(orc_render_slices_blend)(const R* field4, const int fdims[3],
                                 const uint8_t* noise, int noise_ch, const int ndims[3],
                                 const uint8_t* scalar, const int sdims[3],
                                 const uint8_t* lut, int lut_width,
                                 const R* tf_rgba, const R* tf_opac,
                                 const R* sigma, const int sigma_dims[3],
                                 const orc_params* p, const orc_view* v,
                                 int nthreads, uint8_t* image, int* nsamples)
{
    return FN(render_slices_impl)(field4, fdims, noise, noise_ch, ndims, scalar, sdims, lut, lut_width, tf_rgba,
                                  tf_opac, sigma, sigma_dims, p, v, nthreads, 1, NULL, image, nsamples);
}

#if ORC_DOUBLE
/* ------------------------------------------------------------------ */
/* Adjoint of orc_raycast_f64 w.r.t. tf_rgba and sigma.                */
/* Forward recurrence  D' = clamp(D + (1-D.a) s, 0, 1), s = (rgb a, a)  */
/* Backward            G  -> mask by the clamp, dL/ds = (1-D.a) G,      */
/*                     dL/dD = G - (G.s) e_a.                           */
/* The termination index is held fixed (piecewise-constant).           */
/* ------------------------------------------------------------------ */
static void orc_splat_lin(double* g, int n, int nc, int ch, double u, double val)
{
    double f;
    int i = tcoord_f64(u, n, &f);
    int i0 = icl_f64(i, n), i1 = icl_f64(i + 1, n);
    g[i0 * nc + ch] += (1.0 - f) * val;
    g[i1 * nc + ch] += f * val;
}

void orc_raycast_adjoint_f64(const double* field4, const int fdims[3],
                             const double* lic, int lic_ch, const int ldims[3],
                             const double* tf_rgba, const double* tf_opac,
                             const double* sigma, const orc_params* p, const orc_view* v,
                             const double* dimage, double* dtf, double* dsigma)
{
    rc_ctx_f64 c;
    c.field = field4; c.lic = lic; c.lc = lic_ch;
    for (int a = 0; a < 3; ++a) { c.fd[a] = fdims[a]; c.ld[a] = ldims[a]; }
    c.tf = tf_rgba; c.op = tf_opac; c.sigma = sigma; c.p = p;
    double lraw[3] = {p->light[0], p->light[1], p->light[2]};
    normalize3_f64(lraw, c.light);
    const long maxs = (long)p->num_iterations * (long)p->num_iterations;
    const double step = p->step_size_vol;
    const double cc = p->alpha_correction;
    memset(dtf, 0, sizeof(double) * 256 * 4);
    long nvox = (long)ldims[0] * ldims[1] * ldims[2];
    if (dsigma) memset(dsigma, 0, sizeof(double) * nvox);
    long cap = 1024;
    rc_sample_f64* ss = (rc_sample_f64*)malloc(sizeof(rc_sample_f64) * cap);
    double (*ps)[3] = malloc(sizeof(double[3]) * cap);
    double (*ds)[4] = malloc(sizeof(double[4]) * (cap + 1));
    double (*us)[4] = malloc(sizeof(double[4]) * cap);
    for (int j = 0; j < v->height; ++j)
        for (int i = 0; i < v->width; ++i) {
            double pos[3], dir[3];
            long pix = (long)j * v->width + i;
            if (!ray_setup_f64(v, i, j, pos, dir)) continue;
            long n = 0;
            double dest[4] = {0, 0, 0, 0};
            memcpy(ds[0], dest, sizeof(dest));
            for (long k = 0; k < maxs; ++k) {
                if (n == cap) {
                    cap *= 2;
                    ss = realloc(ss, sizeof(rc_sample_f64) * cap);
                    ps = realloc(ps, sizeof(double[3]) * cap);
                    ds = realloc(ds, sizeof(double[4]) * (cap + 1));
                    us = realloc(us, sizeof(double[4]) * cap);
                }
                rc_eval_f64(&c, pos, dir, &ss[n]);
                memcpy(ps[n], pos, sizeof(pos));
                double a = ss[n].a;
                double src[4] = {ss[n].rgb[0] * a, ss[n].rgb[1] * a, ss[n].rgb[2] * a, a};
                double t = 1.0 - dest[3];
                for (int q = 0; q < 4; ++q) {
                    us[n][q] = fma(t, src[q], dest[q]);
                    dest[q] = clamp01_f64(us[n][q]);
                }
                ++n;
                memcpy(ds[n], dest, sizeof(dest));
                for (int q = 0; q < 3; ++q) pos[q] = fma(dir[q], step, pos[q]);
                if (outside_f64(pos, v) || dest[3] > 0.95) break;
            }
            double G[4];
            for (int q = 0; q < 4; ++q) G[q] = dimage[4 * pix + q];
            for (long k = n - 1; k >= 0; --k) {
                rc_sample_f64* s = &ss[k];
                for (int q = 0; q < 4; ++q)
                    if (us[k][q] < 0.0 || us[k][q] > 1.0) G[q] = 0.0;
                double t = 1.0 - ds[k][3];
                double a = s->a;
                double src[4] = {s->rgb[0] * a, s->rgb[1] * a, s->rgb[2] * a, a};
                double dsrc[4];
                for (int q = 0; q < 4; ++q) dsrc[q] = t * G[q];
                double gs = G[0] * src[0] + G[1] * src[1] + G[2] * src[2] + G[3] * src[3];
                G[3] -= gs;
                double drgb[3], da = dsrc[3];
                for (int q = 0; q < 3; ++q) {
                    drgb[q] = dsrc[q] * a;
                    da += dsrc[q] * s->rgb[q];
                }
                double b = 1.0 - s->araw;
                double dadraw;
                if (cc == 1.0) dadraw = 1.0;
                else if (b > 0.0) dadraw = cc * pow(b, cc - 1.0);
                else dadraw = 0.0;
                double draw = s->aclamp ? 0.0 : da * dadraw;
                for (int q = 0; q < 3; ++q) orc_splat_lin(dtf, 256, 4, q, s->tfu, drgb[q] * s->m[q]);
                orc_splat_lin(dtf, 256, 4, 3, s->tfu, draw * s->ao * s->sg);
                if (dsigma) {
                    double dsg = draw * s->ao * s->tf[3];
                    cell_f64 cl;
                    cell_clamp_f64(ps[k], c.ld, &cl);
                    int xs[2] = {cl.x0, cl.x1}, ys[2] = {cl.y0, cl.y1}, zs[2] = {cl.z0, cl.z1};
                    double wx[2] = {1.0 - cl.fx, cl.fx}, wy[2] = {1.0 - cl.fy, cl.fy}, wz[2] = {1.0 - cl.fz, cl.fz};
                    for (int cz = 0; cz < 2; ++cz)
                        for (int cy = 0; cy < 2; ++cy)
                            for (int cx = 0; cx < 2; ++cx)
                                dsigma[ORC_IDX(c.ld, xs[cx], ys[cy], zs[cz])] += wx[cx] * wy[cy] * wz[cz] * dsg;
                }
            }
        }
    free(ss); free(ps); free(ds); free(us);
}
#endif

#undef R
#undef FN
#undef FMA
#undef SQRT
#undef FLOOR
#undef POW
#undef FMIN
#undef FMAX
#undef K
#undef ORC_IDX
#undef ORC_TRILIN
