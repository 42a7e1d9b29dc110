/*
 * Deterministic synthetic YUV clips for parity tests and benchmarks.
 *
 * Scene: a flat background plus `polygons` star-shaped polygons painted in
 * order (later ones on top).  Each region has a flat luma shade (uniform
 * 16..235, 8-bit scale) and flat U/V values (uniform 16..240).  Pixels whose
 * 3x3 neighbourhood touches another region are drawn as a dark outline
 * (luma 16), which gives a 2-pixel line across every boundary.  Gaussian
 * noise is added to luma and optionally chroma: each sample takes one of
 * 65,536 equiprobable normal quantiles (the inverse normal CDF at
 * (k + 0.5) / 65536, Acklam's rational approximation refined by one Halley
 * step), chosen by a counter hash of (seed, frame, plane, sample index), so
 * every sample's noise is independent of the others and of the thread
 * that generates the frame.  Chroma is
 * subsampled by a box average over columns {2i, 2i+1} and/or rows
 * {2j, 2j+1}; the frames are meant to be tagged left-sited (MPEG-2).
 *
 * Integer formats scale the 8-bit values by 2^(bits-8) before noise and
 * rounding (10-bit: 64..940).  Float formats map luma to Y/255 and chroma to
 * C/255 - 0.5.  With `motion`, every polygon translates by a fixed random
 * velocity of 0..4 px/frame.  All randomness comes from splitmix64.
 */
#include <math.h>
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct lgcr_synth_spec {
    uint64_t seed;
    int width, height;
    int ss_w, ss_h;
    int bits;            /* 8..16 for integer, 32 for float */
    int is_float;
    int polygons;
    int motion;          /* polygons translate 0..4 px/frame */
    int scene_per_frame; /* new scene for every frame (seed + frame) */
    double luma_noise;   /* sigma in output LSB (float: absolute) */
    double chroma_noise; /* sigma in output LSB (float: absolute) */
} lgcr_synth_spec;

static uint64_t sm64(uint64_t *s) {
    uint64_t z = (*s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
static double unif(uint64_t *s) { return (sm64(s) >> 11) * (1.0 / 9007199254740992.0); }
static int unif_int(uint64_t *s, int lo, int hi) { return lo + (int)(unif(s) * (hi - lo + 1)); }
/* stateless 64-bit mix (the splitmix64 finaliser) */
static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* Inverse standard normal CDF: P. J. Acklam's rational approximation
 * (relative error below 1.2e-9), refined by one Halley step on erfc. */
static double inv_norm(double p) {
    static const double a[6] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                                1.383577518672690e+02, -3.066479806614716e+01, 2.506628277459239e+00};
    static const double b[5] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                                6.680131188771972e+01, -1.328068155288572e+01};
    static const double c[6] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                                -2.549732539343734e+00, 4.374664141464968e+00, 2.938163982698783e+00};
    static const double d[4] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                                3.754408661907416e+00};
    const double lo = 0.02425, hi = 1.0 - lo;
    double x;
    if (p < lo) {
        const double q = sqrt(-2.0 * log(p));
        x = (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
            ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
    } else if (p <= hi) {
        const double q = p - 0.5, r = q * q;
        x = (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
            (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1.0);
    } else {
        const double q = sqrt(-2.0 * log(1.0 - p));
        x = -(((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
            ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
    }
    const double e = 0.5 * erfc(-x / sqrt(2.0)) - p;
    const double u = e * sqrt(2.0 * M_PI) * exp(x * x / 2.0);
    return x - u / (1.0 + x * u / 2.0);
}

#define NOISE_BITS 16
static double g_quantile[1 << NOISE_BITS];
static pthread_once_t g_quantile_once = PTHREAD_ONCE_INIT;
static void init_quantiles(void) {
    for (int k = 0; k < (1 << NOISE_BITS); ++k)
        g_quantile[k] = inv_norm((k + 0.5) / (double)(1 << NOISE_BITS));
}
/* N(0, 1) sample for counter `i` of stream `base` */
static double gauss_at(uint64_t base, uint64_t i) {
    return g_quantile[mix64(base + i * 0x9E3779B97F4A7C15ULL) >> (64 - NOISE_BITS)];
}

#define MAXV 8
typedef struct { double cx, cy, vx, vy; int n; double ang[MAXV], rad[MAXV]; int y8, u8, v8; } poly_t;

static int cmpd(const void *a, const void *b) {
    const double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* store one sample: float as is; integers rounded half up (values are
 * non-negative before the clamp wherever the clamp does not decide) */
static inline void put_sample(const lgcr_synth_spec *sp, uint8_t *row, int x, double val, int maxv) {
    if (sp->is_float) {
        ((float *)row)[x] = (float)val;
        return;
    }
    long q = (long)(val + 0.5);
    if (q < 0) q = 0;
    if (q > maxv) q = maxv;
    if (sp->bits <= 8) row[x] = (uint8_t)q;
    else ((uint16_t *)row)[x] = (uint16_t)q;
}

int lgcr_synth_frame(const lgcr_synth_spec *sp, int frame, void *const planes[3],
                     const ptrdiff_t strides[3]) {
    const int W = sp->width, H = sp->height;
    if (W <= 0 || H <= 0 || sp->polygons < 0 || sp->polygons > 32767)
        return -1;
    uint64_t rs = sp->seed * 0x2545F4914F6CDD1DULL + 0x1234567ULL;
    if (sp->scene_per_frame)
        rs += (uint64_t)frame * 0x9E3779B97F4A7C15ULL;
    const int np = sp->polygons;
    poly_t *P = (poly_t *)calloc((size_t)np + 1, sizeof(poly_t));
    const int bg_y = unif_int(&rs, 16, 235), bg_u = unif_int(&rs, 16, 240), bg_v = unif_int(&rs, 16, 240);
    const double mind = W < H ? W : H;
    for (int k = 0; k < np; ++k) {
        poly_t *p = &P[k];
        p->cx = unif(&rs) * W;
        p->cy = unif(&rs) * H;
        p->n = unif_int(&rs, 3, MAXV);
        const double r0 = (0.04 + 0.21 * unif(&rs)) * mind;
        for (int i = 0; i < p->n; ++i) p->ang[i] = unif(&rs) * 2.0 * M_PI;
        qsort(p->ang, (size_t)p->n, sizeof(double), cmpd);
        for (int i = 0; i < p->n; ++i) p->rad[i] = r0 * (0.5 + 0.5 * unif(&rs));
        p->y8 = unif_int(&rs, 16, 235);
        p->u8 = unif_int(&rs, 16, 240);
        p->v8 = unif_int(&rs, 16, 240);
        const double speed = 4.0 * unif(&rs), dir = unif(&rs) * 2.0 * M_PI;
        p->vx = speed * cos(dir);
        p->vy = speed * sin(dir);
    }
    int16_t *label = (int16_t *)malloc(sizeof(int16_t) * (size_t)W * H);
    for (size_t i = 0; i < (size_t)W * H; ++i) label[i] = -1;
    double xs[MAXV], ys[MAXV], cross[MAXV];
    for (int k = 0; k < np; ++k) {
        const poly_t *p = &P[k];
        const double t = sp->motion ? (double)frame : 0.0;
        const double cx = p->cx + t * p->vx, cy = p->cy + t * p->vy;
        double ymin = 1e300, ymax = -1e300;
        for (int i = 0; i < p->n; ++i) {
            xs[i] = cx + p->rad[i] * cos(p->ang[i]);
            ys[i] = cy + p->rad[i] * sin(p->ang[i]);
            if (ys[i] < ymin) ymin = ys[i];
            if (ys[i] > ymax) ymax = ys[i];
        }
        int y0 = (int)floor(ymin), y1 = (int)ceil(ymax);
        if (y0 < 0) y0 = 0;
        if (y1 > H - 1) y1 = H - 1;
        for (int y = y0; y <= y1; ++y) {
            const double py = y + 0.5;
            int nc = 0;
            for (int i = 0; i < p->n; ++i) {
                const int j = (i + 1) % p->n;
                const double ya = ys[i], yb = ys[j];
                if ((ya <= py && yb > py) || (yb <= py && ya > py))
                    cross[nc++] = xs[i] + (py - ya) / (yb - ya) * (xs[j] - xs[i]);
            }
            qsort(cross, (size_t)nc, sizeof(double), cmpd);
            for (int c = 0; c + 1 < nc; c += 2) {
                int xa = (int)ceil(cross[c] - 0.5), xb = (int)ceil(cross[c + 1] - 0.5);
                if (xa < 0) xa = 0;
                if (xb > W) xb = W;
                for (int x = xa; x < xb; ++x)
                    label[(size_t)y * W + x] = (int16_t)k;
            }
        }
    }
    /* full-resolution flat values (8-bit scale, integers), outlines on luma:
     * a pixel is an outline pixel when any pixel of its 3x3 neighbourhood
     * inside the frame has another label */
    uint8_t *fy = (uint8_t *)malloc((size_t)W * H);
    uint8_t *fu = (uint8_t *)malloc((size_t)W * H);
    uint8_t *fv = (uint8_t *)malloc((size_t)W * H);
    uint8_t ly8[32768 + 1], lu8[32768 + 1], lv8[32768 + 1];   /* by label + 1 */
    ly8[0] = (uint8_t)bg_y;
    lu8[0] = (uint8_t)bg_u;
    lv8[0] = (uint8_t)bg_v;
    for (int k = 0; k < np && k < 32768; ++k) {
        ly8[k + 1] = (uint8_t)P[k].y8;
        lu8[k + 1] = (uint8_t)P[k].u8;
        lv8[k + 1] = (uint8_t)P[k].v8;
    }
    for (int y = 0; y < H; ++y) {
        const int16_t *rm = label + (size_t)(y > 0 ? y - 1 : y) * W;
        const int16_t *rc = label + (size_t)y * W;
        const int16_t *rp = label + (size_t)(y < H - 1 ? y + 1 : y) * W;
        for (int x = 0; x < W; ++x) {
            const int xl = x > 0 ? x - 1 : x, xr = x < W - 1 ? x + 1 : x;
            const int16_t l = rc[x];
            /* clamped neighbours repeat the pixel's own row / column, which
             * cannot differ from it more than the in-frame ones do */
            const int edge = (rm[xl] != l) | (rm[x] != l) | (rm[xr] != l) | (rc[xl] != l) | (rc[xr] != l) |
                             (rp[xl] != l) | (rp[x] != l) | (rp[xr] != l);
            const size_t i = (size_t)y * W + x;
            fy[i] = edge ? 16 : ly8[l + 1];
            fu[i] = lu8[l + 1];
            fv[i] = lv8[l + 1];
        }
    }
    pthread_once(&g_quantile_once, init_quantiles);
    const uint64_t nbase = mix64(sp->seed * 0xD1B54A32D192ED03ULL + (uint64_t)frame * 0x8CB92BA72F3D8DD7ULL + 99);
    const int maxv = sp->is_float ? 0 : (1 << sp->bits) - 1;
    const double scale = sp->is_float ? 1.0 / 255.0 : (double)(1 << (sp->bits - 8));
    /* luma */
    for (int y = 0; y < H; ++y) {
        uint8_t *row = (uint8_t *)planes[0] + strides[0] * y;
        const uint8_t *fr = fy + (size_t)y * W;
        for (int x = 0; x < W; ++x) {
            double val = fr[x] * scale;
            if (sp->luma_noise > 0.0) val += sp->luma_noise * gauss_at(nbase, (size_t)y * W + x);
            put_sample(sp, row, x, val, maxv);
        }
    }
    /* chroma: box subsample (integer sums of the 8-bit scale values) */
    const int cw = W >> sp->ss_w, ch = H >> sp->ss_h;
    const int bx = 1 << sp->ss_w, by = 1 << sp->ss_h;
    const double inv_area = 1.0 / (double)(bx * by);
    for (int c = 0; c < 2; ++c) {
        const uint8_t *src = c == 0 ? fu : fv;
        const uint64_t cbase = mix64(nbase + 1 + (uint64_t)c);
        for (int y = 0; y < ch; ++y) {
            uint8_t *row = (uint8_t *)planes[1 + c] + strides[1 + c] * y;
            const uint8_t *r0 = src + (size_t)(y * by) * W;
            const uint8_t *r1 = src + (size_t)(y * by + by - 1) * W;
            for (int x = 0; x < cw; ++x) {
                const int x0 = x * bx, x1 = x0 + bx - 1;
                int sum = 0;
                if (bx <= 2 && by <= 2) {
                    sum = r0[x0];
                    if (bx == 2) sum += r0[x1];
                    if (by == 2) {
                        sum += r1[x0];
                        if (bx == 2) sum += r1[x1];
                    }
                } else {
                    for (int j = 0; j < by; ++j)
                        for (int i = 0; i < bx; ++i)
                            sum += src[(size_t)(y * by + j) * W + (x0 + i)];
                }
                const double m = sum * inv_area;
                double val = sp->is_float ? m / 255.0 - 0.5 : m * scale;
                if (sp->chroma_noise > 0.0) val += sp->chroma_noise * gauss_at(cbase, (size_t)y * cw + x);
                put_sample(sp, row, x, val, maxv);
            }
        }
    }
    free(label); free(fy); free(fu); free(fv); free(P);
    return 0;
}
