/*
 * LGCR oracle — TEST INFRASTRUCTURE ONLY (see lgcr_oracle.h).
 *
 * Scalar restatement of the reference's Recon path.  Every function names the
 * reference lines it follows; arithmetic order, float/double promotion and
 * clamping follow the reference's scalar build (Makefile:48-49: -O3, no
 * -mavx2/-mfma, so no FMA contraction).  Build with -ffp-contract=off.
 */
#define _GNU_SOURCE
#include "lgcr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline float clampf(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }
static inline double clampd(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }
static inline float minf(float a, float b) { return b < a ? b : a; }  /* std::min */
static inline float maxf(float a, float b) { return a < b ? b : a; }  /* std::max */
static inline int mini(int a, int b) { return b < a ? b : a; }
static inline int maxi(int a, int b) { return a < b ? b : a; }

/* Plane::at clamped read (lgcr.h:289-293) */
static inline float at(const float *p, int w, int h, int x, int y) {
    x = clampi(x, 0, w - 1);
    y = clampi(y, 0, h - 1);
    return p[(size_t)y * w + x];
}

/* ------------------------------------------------------------------------ */
/* closed-form kernels: lgcr.h:53-117                                        */
/* ------------------------------------------------------------------------ */

static double sinc_d(double x) {
    if (fabs(x) < 1e-8)
        return 1.0;
    x *= M_PI;
    return sin(x) / x;
}

double or_kernel_eval(int k, double x, double p1, double p2) {
    x = fabs(x);
    switch (k) {
    case OR_BILINEAR:
        return x < 1.0 ? 1.0 - x : 0.0;
    case OR_BICUBIC: {
        const double b = p1, c = p2;
        if (x < 1.0)
            return ((12 - 9 * b - 6 * c) * x * x * x + (-18 + 12 * b + 6 * c) * x * x + (6 - 2 * b)) / 6.0;
        if (x < 2.0)
            return ((-b - 6 * c) * x * x * x + (6 * b + 30 * c) * x * x + (-12 * b - 48 * c) * x + (8 * b + 24 * c)) / 6.0;
        return 0.0;
    }
    case OR_LANCZOS:
        return x < p1 ? sinc_d(x) * sinc_d(x / p1) : 0.0;
    case OR_SPLINE16:
        if (x < 1.0)
            return ((x - 9.0 / 5.0) * x - 1.0 / 5.0) * x + 1.0;
        if (x < 2.0)
            return ((-1.0 / 3.0 * (x - 1.0) + 4.0 / 5.0) * (x - 1.0) - 7.0 / 15.0) * (x - 1.0);
        return 0.0;
    case OR_SPLINE36:
        if (x < 1.0)
            return ((13.0 / 11.0 * x - 453.0 / 209.0) * x - 3.0 / 209.0) * x + 1.0;
        if (x < 2.0)
            return ((-6.0 / 11.0 * (x - 1.0) + 270.0 / 209.0) * (x - 1.0) - 156.0 / 209.0) * (x - 1.0);
        if (x < 3.0)
            return ((1.0 / 11.0 * (x - 2.0) - 45.0 / 209.0) * (x - 2.0) + 26.0 / 209.0) * (x - 2.0);
        return 0.0;
    default:
        return 0.0;
    }
}

/* jinc-lanczos profile (lgcr.h:98-105).  The reference calls libstdc++'s
 * std::cyl_bessel_j(1, .); this restatement uses glibc's j1(), an independent
 * implementation of the same function. */
double or_jinc_windowed(double x, double taps) {
    x = fabs(x);
    if (x >= taps)
        return 0.0;
    const double j = (x < 1e-8) ? 1.0 : 2.0 * j1(M_PI * x) / (M_PI * x);
    return j * sinc_d(x / taps);
}

double or_kernel_support(int k, double p1) {
    switch (k) {
    case OR_BILINEAR: return 1.0;
    case OR_BICUBIC: return 2.0;
    case OR_LANCZOS: return p1;
    case OR_SPLINE16: return 2.0;
    case OR_SPLINE36: return 3.0;
    case OR_JINC: return p1;
    }
    return 1.0;
}

/* LUT: 1024 entries per unit distance, support*1024+2 entries (plugin.cpp:1071-1076) */
void or_jinc_lut(double support, float *lut, int n) {
    for (int i = 0; i < n; ++i)
        lut[i] = (float)or_jinc_windowed((double)i / 1024.0, support);
}

/* lgcr.h:636-641 */
float or_lut_lookup(const float *lut, int n, float idx) {
    int i0 = (int)idx;
    i0 = clampi(i0, 0, n - 2);
    const float fr = minf(idx - (float)i0, 1.0f);
    return lut[i0] + (lut[i0 + 1] - lut[i0]) * fr;
}

/* ------------------------------------------------------------------------ */
/* weight tables: kernels.cpp:135-167                                        */
/* ------------------------------------------------------------------------ */

int or_build_weights(int src_n, int dst_n, int k, double kp1, double kp2,
                     double support, double shift, or_weights *t) {
    t->n = dst_n;
    const double scale = (double)src_n / dst_n;
    const double sup = (scale > 1.0) ? support * scale : support;
    const double inv_scale = (scale > 1.0) ? 1.0 / scale : 1.0;
    t->sup = maxi(1, 2 * (int)ceil(sup));
    t->start = (int *)malloc(sizeof(int) * (size_t)dst_n);
    t->w = (float *)calloc((size_t)dst_n * t->sup, sizeof(float));
    if (!t->start || !t->w)
        return -1;
    for (int i = 0; i < dst_n; ++i) {
        const double center = (i + 0.5) * scale - 0.5 + shift;
        int first = (int)ceil(center - sup);
        first = clampi(first, -(t->sup) + 1, src_n - 1);
        if (first + t->sup > src_n)
            first = maxi(-(t->sup) + 1, src_n - t->sup);
        t->start[i] = first;
        double sum = 0.0;
        for (int j = 0; j < t->sup; ++j) {
            const double d = (first + j) - center;
            const double w = or_kernel_eval(k, d * inv_scale, kp1, kp2);
            t->w[(size_t)i * t->sup + j] = (float)w;
            sum += w;
        }
        if (fabs(sum) > 1e-9)
            for (int j = 0; j < t->sup; ++j)
                t->w[(size_t)i * t->sup + j] /= (float)sum;
    }
    return 0;
}

void or_free_weights(or_weights *w) {
    free(w->start);
    free(w->w);
    memset(w, 0, sizeof(*w));
}

/* maps.cpp:190-209 (the AVX2 phase2x descriptor is not part of the result) */
void or_bilin_axis(const float *pos, int n, int src_n, int *i0, float *f) {
    if (src_n == 1) {
        for (int i = 0; i < n; ++i) { i0[i] = 0; f[i] = 0.0f; }
        return;
    }
    for (int i = 0; i < n; ++i) {
        const float x = pos[i];
        int j = (int)floorf(x);
        float fr = x - (float)j;
        if (j < 0) { j = 0; fr = 0.0f; }
        if (j > src_n - 2) { j = src_n - 2; fr = 1.0f; }
        i0[i] = j;
        f[i] = fr;
    }
}

/* kernels.cpp:337-404 */
int or_build_axis(int src_luma_n, int dst_luma_n, double r, double sit_shift,
                  const or_params *p, or_axis *a) {
    memset(a, 0, sizeof(*a));
    a->n = dst_luma_n;
    const double scale_c = ((double)src_luma_n / r) / dst_luma_n;
    const double widen_c = p->radial ? (scale_c > 1.0 ? scale_c : 1.0) : 1.0;
    a->sup = maxi(1, 2 * (int)ceil(p->support * widen_c));
    const size_t nn = (size_t)a->n, ns = nn * a->sup;
    a->start = (int *)malloc(sizeof(int) * nn);
    a->pos = (float *)malloc(sizeof(float) * nn);
    a->lpos = (float *)malloc(sizeof(float) * nn);
    a->w = (float *)calloc(ns, sizeof(float));
    a->am = (float *)calloc(ns, sizeof(float));
    a->wsum = (float *)calloc(nn, sizeof(float));
    a->wabs = (float *)calloc(nn, sizeof(float));
    a->active = (int *)calloc(nn, sizeof(int));
    a->cb_i0 = (int *)malloc(sizeof(int) * nn);
    a->cb_f = (float *)malloc(sizeof(float) * nn);
    a->lb_i0 = (int *)malloc(sizeof(int) * nn);
    a->lb_f = (float *)malloc(sizeof(float) * nn);
    if (!a->start || !a->pos || !a->lpos || !a->w || !a->am || !a->wsum || !a->wabs ||
        !a->active || !a->cb_i0 || !a->cb_f || !a->lb_i0 || !a->lb_f)
        return -1;
    const int ri = (int)r;
    const int src_chroma_n = (src_luma_n + ri - 1) / ri;
    const double scale = (double)src_luma_n / dst_luma_n;
    for (int i = 0; i < a->n; ++i) {
        const double lx = (i + 0.5) * scale - 0.5;
        const double sc = (lx + 0.5 - sit_shift) / r - 0.5;
        a->pos[i] = (float)sc;
        a->lpos[i] = (float)lx;
        int first = (int)ceil(sc - p->support * widen_c);
        first = clampi(first, 0, maxi(0, src_chroma_n - a->sup));
        a->start[i] = first;
        float *wi = a->w + (size_t)i * a->sup;
        float *mi = a->am + (size_t)i * a->sup;
        for (int j = 0; j < a->sup; ++j)
            if (fabs((first + j) - sc) < p->support * widen_c) {
                mi[j] = 1.0f;
                ++a->active[i];
            }
        if (p->radial)
            continue;
        double sum = 0.0;
        for (int j = 0; j < a->sup; ++j) {
            const double dist = (first + j) - sc;
            const double w = or_kernel_eval(p->kernel, dist, p->kp1, p->kp2);
            wi[j] = (float)w;
            sum += w;
        }
        if (fabs(sum) > 1e-9)
            for (int j = 0; j < a->sup; ++j)
                wi[j] /= (float)sum;
        for (int j = 0; j < a->sup; ++j) {
            if (mi[j] == 0.0f)
                continue;
            a->wsum[i] += wi[j];
            a->wabs[i] += fabsf(wi[j]);
        }
    }
    or_bilin_axis(a->pos, a->n, src_chroma_n, a->cb_i0, a->cb_f);
    or_bilin_axis(a->lpos, a->n, src_luma_n, a->lb_i0, a->lb_f);
    return 0;
}

void or_free_axis(or_axis *a) {
    free(a->start); free(a->pos); free(a->lpos); free(a->w); free(a->am);
    free(a->wsum); free(a->wabs); free(a->active);
    free(a->cb_i0); free(a->cb_f); free(a->lb_i0); free(a->lb_f);
    memset(a, 0, sizeof(*a));
}

/* ------------------------------------------------------------------------ */
/* conversions: lgcr.h:986-1013                                              */
/* ------------------------------------------------------------------------ */

void or_u8_to_float(const uint8_t *src, ptrdiff_t stride, float *dst, int w, int h,
                    double scale, double offset) {
    for (int y = 0; y < h; ++y) {
        const uint8_t *s = src + stride * y;
        for (int x = 0; x < w; ++x)
            dst[(size_t)y * w + x] = (float)(s[x] * scale + offset);
    }
}

void or_u16_to_float(const uint16_t *src, ptrdiff_t stride, float *dst, int w, int h,
                     double scale, double offset) {
    for (int y = 0; y < h; ++y) {
        const uint16_t *s = (const uint16_t *)((const uint8_t *)src + stride * y);
        for (int x = 0; x < w; ++x)
            dst[(size_t)y * w + x] = (float)(s[x] * scale + offset);
    }
}

void or_float_to_u8(const float *src, uint8_t *dst, ptrdiff_t stride, int w, int h,
                    double scale, double offset, int lo, int hi) {
    const float fs = (float)scale, fo = (float)offset;
    for (int y = 0; y < h; ++y) {
        uint8_t *d = dst + stride * y;
        for (int x = 0; x < w; ++x) {
            const int v = (int)(src[(size_t)y * w + x] * fs + fo + 0.5f);
            d[x] = (uint8_t)clampi(v, lo, hi);
        }
    }
}

void or_float_to_u16(const float *src, uint16_t *dst, ptrdiff_t stride, int w, int h,
                     double scale, double offset, int lo, int hi) {
    const float fs = (float)scale, fo = (float)offset;
    for (int y = 0; y < h; ++y) {
        uint16_t *d = (uint16_t *)((uint8_t *)dst + stride * y);
        for (int x = 0; x < w; ++x) {
            const int v = (int)(src[(size_t)y * w + x] * fs + fo + 0.5f);
            d[x] = (uint16_t)clampi(v, lo, hi);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* structure tensor + trust seed: maps.cpp:9-168                             */
/* ------------------------------------------------------------------------ */

/* sobelPixel (maps.cpp:9-20); the scalar interior loop (maps.cpp:36-59 with
 * one lane) evaluates the identical expression, so every pixel uses it. */
static void sobel_px(const float *p, int w, int h, int x, int y, float *gx, float *gy) {
    const float tl = at(p, w, h, x - 1, y - 1);
    const float t = at(p, w, h, x, y - 1);
    const float tr = at(p, w, h, x + 1, y - 1);
    const float l = at(p, w, h, x - 1, y);
    const float r = at(p, w, h, x + 1, y);
    const float bl = at(p, w, h, x - 1, y + 1);
    const float b = at(p, w, h, x, y + 1);
    const float br = at(p, w, h, x + 1, y + 1);
    *gx = (tr + 2.0f * r + br - tl - 2.0f * l - bl) * 0.125f;
    *gy = (bl + 2.0f * b + br - tl - 2.0f * t - tr) * 0.125f;
}

void or_sobel_tensor(const float *y, int w, int h, float *jxx, float *jxy, float *jyy,
                     uint8_t *seed, float threshold) {
    float *gx = (float *)malloc(sizeof(float) * (size_t)w * h);
    float *gy = (float *)malloc(sizeof(float) * (size_t)w * h);
    float *vxx = (float *)malloc(sizeof(float) * (size_t)w);
    float *vxy = (float *)malloc(sizeof(float) * (size_t)w);
    float *vyy = (float *)malloc(sizeof(float) * (size_t)w);
    for (int r = 0; r < h; ++r)
        for (int x = 0; x < w; ++x)
            sobel_px(y, w, h, x, r, &gx[(size_t)r * w + x], &gy[(size_t)r * w + x]);
    const float ninth = 1.0f / 9.0f;
    for (int r = 0; r < h; ++r) {
        const float *gx0 = gx + (size_t)maxi(0, r - 1) * w, *gx1 = gx + (size_t)r * w;
        const float *gx2 = gx + (size_t)mini(h - 1, r + 1) * w;
        const float *gy0 = gy + (size_t)maxi(0, r - 1) * w, *gy1 = gy + (size_t)r * w;
        const float *gy2 = gy + (size_t)mini(h - 1, r + 1) * w;
        for (int x = 0; x < w; ++x) {
            vxx[x] = gx0[x] * gx0[x] + gx1[x] * gx1[x] + gx2[x] * gx2[x];
            vxy[x] = gx0[x] * gy0[x] + gx1[x] * gy1[x] + gx2[x] * gy2[x];
            vyy[x] = gy0[x] * gy0[x] + gy1[x] * gy1[x] + gy2[x] * gy2[x];
        }
        float *oxx = jxx + (size_t)r * w, *oxy = jxy + (size_t)r * w, *oyy = jyy + (size_t)r * w;
        const int r1 = mini(1, w - 1);
        /* border columns divide by 9, the interior multiplies by float(1/9)
         * (maps.cpp:125-154) */
        oxx[0] = (2.0f * vxx[0] + vxx[r1]) / 9.0f;
        oxy[0] = (2.0f * vxy[0] + vxy[r1]) / 9.0f;
        oyy[0] = (2.0f * vyy[0] + vyy[r1]) / 9.0f;
        for (int x = 1; x + 1 <= w - 1; ++x) {
            oxx[x] = ninth * ((vxx[x - 1] + vxx[x]) + vxx[x + 1]);
            oxy[x] = ninth * ((vxy[x - 1] + vxy[x]) + vxy[x + 1]);
            oyy[x] = ninth * ((vyy[x - 1] + vyy[x]) + vyy[x + 1]);
        }
        if (w > 1) {
            oxx[w - 1] = (vxx[w - 2] + 2.0f * vxx[w - 1]) / 9.0f;
            oxy[w - 1] = (vxy[w - 2] + 2.0f * vxy[w - 1]) / 9.0f;
            oyy[w - 1] = (vyy[w - 2] + 2.0f * vyy[w - 1]) / 9.0f;
        }
        if (seed)
            for (int x = 0; x < w; ++x)
                seed[(size_t)r * w + x] = oxx[x] + oyy[x] > threshold;
    }
    free(gx); free(gy); free(vxx); free(vxy); free(vyy);
}

/* footprint luma map (maps.cpp:425-499, scalar tail 487-496) */
void or_lc_map(const float *y, int w, int h, int cw, int ch, double rw, double rh,
               double shift_x, double shift_y, float *lc) {
    int *xb = (int *)malloc(sizeof(int) * (size_t)cw), *xe = (int *)malloc(sizeof(int) * (size_t)cw);
    int *yb = (int *)malloc(sizeof(int) * (size_t)ch), *ye = (int *)malloc(sizeof(int) * (size_t)ch);
    for (int cx = 0; cx < cw; ++cx) {
        const double lx = (cx + 0.5) * rw - 0.5 + shift_x;
        int first = (int)ceil(lx - 0.5 * rw);
        int last = (int)ceil(lx + 0.5 * rw);
        first = clampi(first, 0, w - 1);
        last = clampi(last, first + 1, w);
        xb[cx] = first; xe[cx] = last;
    }
    for (int cy = 0; cy < ch; ++cy) {
        const double ly = (cy + 0.5) * rh - 0.5 + shift_y;
        int first = (int)ceil(ly - 0.5 * rh);
        int last = (int)ceil(ly + 0.5 * rh);
        first = clampi(first, 0, h - 1);
        last = clampi(last, first + 1, h);
        yb[cy] = first; ye[cy] = last;
    }
    for (int cy = 0; cy < ch; ++cy)
        for (int cx = 0; cx < cw; ++cx) {
            double sum = 0;
            for (int j = yb[cy]; j < ye[cy]; ++j)
                for (int i = xb[cx]; i < xe[cx]; ++i)
                    sum += y[(size_t)j * w + i];
            lc[(size_t)cy * cw + cx] = (float)(sum / ((double)(xe[cx] - xb[cx]) * (ye[cy] - yb[cy])));
        }
    free(xb); free(xe); free(yb); free(ye);
}

/* separable binary dilation by running counts (maps.cpp:295-339; same
 * result as dilateMask, maps.cpp:704-742): a pixel is set when any source
 * pixel within Chebyshev distance `radius` is set. */
void or_dilate(const uint8_t *src, int w, int h, int r, uint8_t *dst) {
    uint8_t *tmp = (uint8_t *)calloc((size_t)w * h, 1);
    for (int y = 0; y < h; ++y) {
        const uint8_t *s = src + (size_t)y * w;
        uint8_t *t = tmp + (size_t)y * w;
        int active = 0;
        for (int x = 0; x <= mini(w - 1, r); ++x)
            active += s[x] != 0;
        for (int x = 0; x < w; ++x) {
            t[x] = active != 0;
            if (x - r >= 0) active -= s[x - r] != 0;
            if (x + r + 1 < w) active += s[x + r + 1] != 0;
        }
    }
    int *cnt = (int *)calloc((size_t)w, sizeof(int));
    for (int y = 0; y <= mini(h - 1, r); ++y)
        for (int x = 0; x < w; ++x)
            cnt[x] += tmp[(size_t)y * w + x] != 0;
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x)
            dst[(size_t)y * w + x] = cnt[x] != 0;
        for (int x = 0; x < w; ++x) {
            if (y - r >= 0) cnt[x] -= tmp[(size_t)(y - r) * w + x] != 0;
            if (y + r + 1 < h) cnt[x] += tmp[(size_t)(y + r + 1) * w + x] != 0;
        }
    }
    free(cnt);
    free(tmp);
}

/* ------------------------------------------------------------------------ */
/* sparse workset: maps.cpp:342-422                                          */
/* ------------------------------------------------------------------------ */

int or_build_workset(const uint8_t *mask, int mask_w, int mask_h, int out_w, int out_h,
                     const or_axis *ax, const or_axis *ay, int cw, int ch,
                     or_workset *ws) {
    memset(ws, 0, sizeof(*ws));
    ws->mask_w = mask_w; ws->mask_h = mask_h;
    ws->out_w = out_w; ws->out_h = out_h;
    ws->chroma_w = cw; ws->chroma_h = ch;
    ws->mask = (uint8_t *)malloc((size_t)mask_w * mask_h);
    memcpy(ws->mask, mask, (size_t)mask_w * mask_h);
    ws->index_row_offsets = (size_t *)calloc((size_t)out_h + 1, sizeof(size_t));
    ws->span_row_offsets = (size_t *)calloc((size_t)out_h + 1, sizeof(size_t));
    ws->indices = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)out_w * out_h + 1));
    ws->spans = (int *)malloc(sizeof(int) * ((size_t)out_w * out_h + 2));
    ws->chroma_mask = (uint8_t *)calloc((size_t)cw * ch, 1);
    int *mx = (int *)malloc(sizeof(int) * (size_t)out_w);
    for (int x = 0; x < out_w; ++x)
        mx[x] = mini(mask_w - 1, (int)((int64_t)x * mask_w / out_w));
    for (int y = 0; y < out_h; ++y) {
        ws->span_row_offsets[y] = ws->n_spans;
        ws->index_row_offsets[y] = ws->n_indices;
        const int my = mini(mask_h - 1, (int)((int64_t)y * mask_h / out_h));
        const uint8_t *row = ws->mask + (size_t)my * mask_w;
        int x = 0;
        while (x < out_w) {
            while (x < out_w && row[mx[x]] == 0) ++x;
            const int begin = x;
            while (x < out_w && row[mx[x]] != 0) {
                ws->indices[ws->n_indices++] = (uint32_t)((size_t)y * out_w + x);
                ++x;
            }
            if (begin < x) {
                ws->spans[2 * ws->n_spans] = begin;
                ws->spans[2 * ws->n_spans + 1] = x;
                ws->n_spans++;
            }
        }
    }
    ws->span_row_offsets[out_h] = ws->n_spans;
    ws->index_row_offsets[out_h] = ws->n_indices;
    for (int y = 0; y < out_h; ++y) {
        const int cy0 = ay->cb_i0[y];
        const int cy1 = mini(cy0 + 1, ch - 1);
        uint8_t *a0 = ws->chroma_mask + (size_t)cy0 * cw;
        uint8_t *a1 = ws->chroma_mask + (size_t)cy1 * cw;
        for (size_t s = ws->span_row_offsets[y]; s < ws->span_row_offsets[y + 1]; ++s)
            for (int x = ws->spans[2 * s]; x < ws->spans[2 * s + 1]; ++x) {
                const int cx0 = ax->cb_i0[x];
                const int cx1 = mini(cx0 + 1, cw - 1);
                a0[cx0] = a0[cx1] = 1;
                a1[cx0] = a1[cx1] = 1;
            }
    }
    for (size_t i = 0; i < (size_t)cw * ch; ++i)
        ws->active_chroma += ws->chroma_mask[i] != 0;
    free(mx);
    return 0;
}

void or_free_workset(or_workset *ws) {
    free(ws->mask); free(ws->indices); free(ws->index_row_offsets);
    free(ws->spans); free(ws->span_row_offsets); free(ws->chroma_mask);
    memset(ws, 0, sizeof(*ws));
}

/* lgcr.h:780-785 */
int or_workset_output_dense(const or_workset *ws) {
    return ws->n_indices * 2 >= (size_t)ws->out_w * ws->out_h;
}
int or_workset_chroma_dense(const or_workset *ws) {
    return ws->active_chroma * 2 >= (size_t)ws->chroma_w * ws->chroma_h;
}

/* ------------------------------------------------------------------------ */
/* mutual-structure gate: maps.cpp:831-948                                   */
/* ------------------------------------------------------------------------ */

typedef struct { float coherence, nx, ny; } tdir;

/* principalTensorDirection (lgcr.h:516-538) */
static tdir principal_dir(float jxx, float jxy, float jyy) {
    tdir d = {0.0f, 1.0f, 0.0f};
    const float diff = jxx - jyy;
    const float sum = jxx + jyy;
    const float cross = 2.0f * jxy;
    const float disc = sqrtf(diff * diff + cross * cross);
    d.coherence = disc / (sum + 1e-12f);
    if (disc <= 1e-20f)
        return d;
    float vx, vy;
    if (diff >= 0.0f) { vx = disc + diff; vy = cross; }
    else { vx = cross; vy = disc - diff; }
    const float inv = 1.0f / sqrtf(vx * vx + vy * vy);
    d.nx = vx * inv;
    d.ny = vy * inv;
    return d;
}

/* SobelRowRing::sampleProjected (maps.cpp:782-791) at a bilinearPoint
 * (maps.cpp:175-186). */
static float sample_projected(const float *gx, const float *gy, int w, int h,
                              double px, double py, float nx, float ny) {
    const int rx = (int)floor(px), ry = (int)floor(py);
    const int x0 = clampi(rx, 0, w - 1), x1 = clampi(rx + 1, 0, w - 1);
    const int y0 = clampi(ry, 0, h - 1), y1 = clampi(ry + 1, 0, h - 1);
    const float fx = (float)(px - rx), fy = (float)(py - ry);
    const float *xt = gx + (size_t)y0 * w, *xb = gx + (size_t)y1 * w;
    const float *yt = gy + (size_t)y0 * w, *yb = gy + (size_t)y1 * w;
    const float a = xt[x0] * nx + yt[x0] * ny;
    const float b = xt[x1] * nx + yt[x1] * ny;
    const float c = xb[x0] * nx + yb[x0] * ny;
    const float d = xb[x1] * nx + yb[x1] * ny;
    return a + (b - a) * fx + (c - a) * fy + (a - b - c + d) * fx * fy;
}

void or_mutual_gate(const float *lc, const float *u, const float *v, int w, int h,
                    const uint8_t *chroma_mask, int chroma_dense, float *gate) {
    const int dense = !chroma_mask || chroma_dense;
    const size_t n = (size_t)w * h;
    float *g = (float *)malloc(sizeof(float) * n * 6);
    float *lx = g, *ly = g + n, *ux = g + 2 * n, *uy = g + 3 * n, *vx = g + 4 * n, *vy = g + 5 * n;
    /* Sobel gradients of the three planes.  The sparse reference computes
     * them only inside the 4-pixel-dilated mask and zero elsewhere; values
     * read by active pixels lie inside that halo, so computing them
     * everywhere gives identical gate values (see or_dilate for the halo). */
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            sobel_px(lc, w, h, x, y, &lx[i], &ly[i]);
            sobel_px(u, w, h, x, y, &ux[i], &uy[i]);
            sobel_px(v, w, h, x, y, &vx[i], &vy[i]);
        }
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            float *out = gate + (size_t)y * w + x;
            if (!dense && chroma_mask[(size_t)y * w + x] == 0) {
                *out = 0.0f;
                continue;
            }
            float sxx = 0, sxy = 0, syy = 0;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const float a = at(lx, w, h, x + dx, y + dy);
                    const float b = at(ly, w, h, x + dx, y + dy);
                    sxx += a * a; sxy += a * b; syy += b * b;
                }
            const float jsum = sxx + syy;
            if (jsum < 1e-10f) { *out = 0.0f; continue; }
            const tdir d = principal_dir(sxx, sxy, syy);
            const float nx = d.nx, ny = d.ny;
            float sumY = 0, sqY = 0, momY = 0, sumC = 0, sqC = 0, momC = 0;
            for (int k = -3; k <= 3; ++k) {
                /* int + float arithmetic happens in float, then widens */
                const double px = (float)x + (float)k * nx;
                const double py = (float)y + (float)k * ny;
                const float gY = fabsf(sample_projected(lx, ly, w, h, px, py, nx, ny));
                const float au = sample_projected(ux, uy, w, h, px, py, nx, ny);
                const float av = sample_projected(vx, vy, w, h, px, py, nx, ny);
                const float gC = sqrtf(au * au + av * av);
                sumY += gY; sqY += gY * gY; momY += (float)k * gY;
                sumC += gC; sqC += gC * gC; momC += (float)k * gC;
            }
            if (sumY < 1e-9f || sumC < 1e-9f) { *out = 0.0f; continue; }
            const float partY = sumY * sumY / (sqY + 1e-12f);
            const float partC = sumC * sumC / (sqC + 1e-12f);
            const float ratio = partC / (partY + 1e-6f);
            const float dcent = fabsf(momY / sumY - momC / sumC);
            const float tw = clampf((1.6f - ratio) / 0.4f, 0.0f, 1.0f);
            const float tp = clampf((1.0f - dcent) / 0.5f, 0.0f, 1.0f);
            *out = (tw * tw * (3.0f - 2.0f * tw)) * (tp * tp * (3.0f - 2.0f * tp));
        }
    free(g);
}

/* ------------------------------------------------------------------------ */
/* bilinear samplers: lgcr.h:467-508                                         */
/* ------------------------------------------------------------------------ */

static float bilinear(const float *p, int w, int h, double x, double y) {
    const int x0 = (int)floor(x), y0 = (int)floor(y);
    const float fx = (float)(x - x0), fy = (float)(y - y0);
    const float a = at(p, w, h, x0, y0), b = at(p, w, h, x0 + 1, y0);
    const float c = at(p, w, h, x0, y0 + 1), d = at(p, w, h, x0 + 1, y0 + 1);
    return a + (b - a) * fx + (c - a) * fy + (a - b - c + d) * fx * fy;
}

static float bilinear_fast(const float *p, int w, int h, int x0, float fx, int y0, float fy) {
    if (w == 1) {
        const float a = p[(size_t)y0 * w];
        if (h == 1)
            return a;
        const float c = p[(size_t)(y0 + 1) * w];
        return a + (c - a) * fy;
    }
    const float *r0 = p + (size_t)y0 * w + x0;
    if (h == 1)
        return r0[0] + (r0[1] - r0[0]) * fx;
    const float *r1 = r0 + w;
    const float a = r0[0], b = r0[1], c = r1[0], d = r1[1];
    return a + (b - a) * fx + (c - a) * fy + (a - b - c + d) * fx * fy;
}

/* ------------------------------------------------------------------------ */
/* plain chroma                                                              */
/* ------------------------------------------------------------------------ */

/* plainChromaSeparable, scalar path (recon.cpp:1234-1263 horizontal,
 * 1389-1437 vertical). */
static void plain_separable(const or_params *p, const float *u, const float *v, int cw, int ch,
                            const or_axis *ax, const or_axis *ay, int ow, int oh,
                            float *out_u, float *out_v) {
    const size_t hn = (size_t)ow * ch;
    float *hu = (float *)malloc(sizeof(float) * hn * 6);
    float *hv = hu + hn, *hmnu = hu + 2 * hn, *hmxu = hu + 3 * hn, *hmnv = hu + 4 * hn, *hmxv = hu + 5 * hn;
    for (int sy = 0; sy < ch; ++sy) {
        const float *su = u + (size_t)sy * cw, *sv = v + (size_t)sy * cw;
        for (int ox = 0; ox < ow; ++ox) {
            const int start = ax->start[ox];
            const float *wt = ax->w + (size_t)ox * ax->sup;
            const float *act = ax->am + (size_t)ox * ax->sup;
            double vu = 0.0, vv = 0.0;
            float mnu = 1e30f, mxu = -1e30f, mnv = 1e30f, mxv = -1e30f;
            for (int i = 0; i < ax->sup; ++i) {
                if (act[i] == 0.0f)
                    continue;
                const int x = clampi(start + i, 0, cw - 1);
                const float a = su[x], b = sv[x];
                vu += (double)wt[i] * a;
                mnu = minf(mnu, a); mxu = maxf(mxu, a);
                vv += (double)wt[i] * b;
                mnv = minf(mnv, b); mxv = maxf(mxv, b);
            }
            const size_t k = (size_t)sy * ow + ox;
            hu[k] = (float)vu; hv[k] = (float)vv;
            hmnu[k] = mnu; hmxu[k] = mxu; hmnv[k] = mnv; hmxv[k] = mxv;
        }
    }
    const int clamp_hull = p->ar_margin >= 0.0;
    const float margin = (float)(p->ar_margin > 0.0 ? p->ar_margin : 0.0);
    for (int oy = 0; oy < oh; ++oy) {
        const int start = ay->start[oy];
        const float *wt = ay->w + (size_t)oy * ay->sup;
        const float *act = ay->am + (size_t)oy * ay->sup;
        for (int ox = 0; ox < ow; ++ox) {
            double vu = 0.0, vv = 0.0;
            float mnu = 1e30f, mxu = -1e30f, mnv = 1e30f, mxv = -1e30f;
            for (int j = 0; j < ay->sup; ++j) {
                if (act[j] == 0.0f)
                    continue;
                const int y = clampi(start + j, 0, ch - 1);
                const size_t k = (size_t)y * ow + ox;
                vu += (double)wt[j] * hu[k];
                mnu = minf(mnu, hmnu[k]); mxu = maxf(mxu, hmxu[k]);
                vv += (double)wt[j] * hv[k];
                mnv = minf(mnv, hmnv[k]); mxv = maxf(mxv, hmxv[k]);
            }
            const double wsum = (double)ax->wsum[ox] * ay->wsum[oy];
            const double asum = (double)ax->wabs[ox] * ay->wabs[oy];
            if (wsum > 1e-6 && wsum >= 0.05 * asum) {
                vu /= wsum;
                vv /= wsum;
            } else {
                vu = bilinear(u, cw, ch, ax->pos[ox], ay->pos[oy]);
                vv = bilinear(v, cw, ch, ax->pos[ox], ay->pos[oy]);
            }
            if (clamp_hull) {
                vu = clampd(vu, (double)(mnu - margin), (double)(mxu + margin));
                vv = clampd(vv, (double)(mnv - margin), (double)(mxv + margin));
            }
            out_u[(size_t)oy * ow + ox] = (float)vu;
            out_v[(size_t)oy * ow + ox] = (float)vv;
        }
    }
    free(hu);
}

/* plainChromaBilinear (recon.cpp:1031-1149, scalar tail 1128-1132) */
static void plain_bilinear(const float *u, const float *v, int cw, int ch,
                           const or_axis *ax, const or_axis *ay, int ow, int oh,
                           float *out_u, float *out_v) {
    for (int oy = 0; oy < oh; ++oy)
        for (int ox = 0; ox < ow; ++ox) {
            out_u[(size_t)oy * ow + ox] = bilinear_fast(u, cw, ch, ax->cb_i0[ox], ax->cb_f[ox],
                                                        ay->cb_i0[oy], ay->cb_f[oy]);
            out_v[(size_t)oy * ow + ox] = bilinear_fast(v, cw, ch, ax->cb_i0[ox], ax->cb_f[ox],
                                                        ay->cb_i0[oy], ay->cb_f[oy]);
        }
}


/* ------------------------------------------------------------------------ */
/* radial (jinc) phase tables: kernels.cpp:449-526                           */
/* ------------------------------------------------------------------------ */

typedef struct {
    int valid, sup_x, sup_y, nx, ny;
    uint16_t *xphase, *yphase;
    float *w;                  /* nx*ny phases, sup_y*sup_x each */
} radial_table;

/* unique (offset bits, activity mask) phases in first-seen order */
static int collect_phases(const or_axis *a, uint16_t *index, float *offsets, uint64_t *activity) {
    if (a->sup > 64)
        return -1;
    int count = 0;
    uint32_t bits[65];
    for (int x = 0; x < a->n; ++x) {
        const float off = (float)a->start[x] - a->pos[x];
        uint32_t b;
        memcpy(&b, &off, 4);
        uint64_t m = 0;
        for (int t = 0; t < a->sup; ++t)
            if (a->am[(size_t)x * a->sup + t] != 0.0f)
                m |= (uint64_t)1 << t;
        int k = 0;
        while (k < count && !(bits[k] == b && activity[k] == m))
            ++k;
        if (k == count) {
            if (count == 64)
                return -1;
            bits[count] = b;
            offsets[count] = off;
            activity[count] = m;
            ++count;
        }
        index[x] = (uint16_t)k;
    }
    return count;
}

static void build_radial_table(const or_params *p, const or_axis *ax, const or_axis *ay,
                               int src_w, int src_h, radial_table *t) {
    memset(t, 0, sizeof(*t));
    t->sup_x = ax->sup;
    t->sup_y = ay->sup;
    float xo[64], yo[64];
    uint64_t xa[64], ya[64];
    t->xphase = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)ax->n);
    t->yphase = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)ay->n);
    const int nx = collect_phases(ax, t->xphase, xo, xa);
    const int ny = nx < 0 ? -1 : collect_phases(ay, t->yphase, yo, ya);
    if (nx < 0 || ny < 0)
        return;
    const size_t pairs = (size_t)nx * ny, per = (size_t)ax->sup * ay->sup;
    if (pairs > 1024 || pairs * per > 262144)
        return;
    t->nx = nx;
    t->ny = ny;
    t->w = (float *)calloc(pairs * per, sizeof(float));
    const float inv_x = (float)fmin(1.0, (double)ax->n / src_w);
    const float inv_y = (float)fmin(1.0, (double)ay->n / src_h);
    for (int py = 0; py < ny; ++py)
        for (int px = 0; px < nx; ++px) {
            float *w = t->w + ((size_t)py * nx + px) * per;
            for (int j = 0; j < ay->sup; ++j)
                for (int i = 0; i < ax->sup; ++i) {
                    if (!(xa[px] & ((uint64_t)1 << i)) || !(ya[py] & ((uint64_t)1 << j)))
                        continue;
                    const float dx = (xo[px] + (float)i) * inv_x;
                    const float dy = (yo[py] + (float)j) * inv_y;
                    w[(size_t)j * ax->sup + i] = or_lut_lookup(p->lut, p->lut_n, sqrtf(dx * dx + dy * dy) * p->lut_scale);
                }
        }
    t->valid = 1;
}

static const float *radial_at(const radial_table *t, int x, int y) {
    if (!t || !t->valid)
        return NULL;
    return t->w + ((size_t)t->yphase[y] * t->nx + t->xphase[x]) * (size_t)t->sup_x * t->sup_y;
}

static void free_radial_table(radial_table *t) {
    free(t->xphase); free(t->yphase); free(t->w);
    memset(t, 0, sizeof(*t));
}

/* plainChromaImpl, scalar path (recon.cpp:828-1029, 970-1005) */
static void plain_radial(const or_params *p, const float *u, const float *v, int cw, int ch,
                         const or_axis *ax, const or_axis *ay, const radial_table *rt,
                         int ow, int oh, float *out_u, float *out_v) {
    const float inv_wx = (float)fmin(1.0, (double)ow / cw);
    const float inv_wy = (float)fmin(1.0, (double)oh / ch);
    const int clamp_hull = p->ar_margin >= 0.0;
    const float margin = (float)(p->ar_margin > 0.0 ? p->ar_margin : 0.0);
    for (int oy = 0; oy < oh; ++oy) {
        const int ty0 = ay->start[oy];
        const float *amy = ay->am + (size_t)oy * ay->sup;
        const float scy = ay->pos[oy];
        for (int ox = 0; ox < ow; ++ox) {
            const int tx0 = ax->start[ox];
            const float *amx = ax->am + (size_t)ox * ax->sup;
            const float scx = ax->pos[ox];
            const float *cw8 = radial_at(rt, ox, oy);
            double au = 0.0, av = 0.0, ws = 0.0, sa = 0.0;
            float mnu = 1e30f, mxu = -1e30f, mnv = 1e30f, mxv = -1e30f;
            for (int j = 0; j < ay->sup; ++j) {
                if (amy[j] == 0.0f)
                    continue;
                const int ty = clampi(ty0 + j, 0, ch - 1);
                for (int i = 0; i < ax->sup; ++i) {
                    if (amx[i] == 0.0f)
                        continue;
                    const int tx = clampi(tx0 + i, 0, cw - 1);
                    float w;
                    if (cw8) {
                        w = cw8[(size_t)j * ax->sup + i];
                    } else {
                        const float dx = ((float)(tx0 + i) - scx) * inv_wx;
                        const float dy = ((float)(ty0 + j) - scy) * inv_wy;
                        w = or_lut_lookup(p->lut, p->lut_n, sqrtf(dx * dx + dy * dy) * p->lut_scale);
                    }
                    const float a = u[(size_t)ty * cw + tx], b = v[(size_t)ty * cw + tx];
                    au += (double)w * a;
                    ws += w;
                    sa += fabsf(w);
                    mnu = minf(mnu, a); mxu = maxf(mxu, a);
                    av += (double)w * b;
                    mnv = minf(mnv, b); mxv = maxf(mxv, b);
                }
            }
            double vu, vv;
            if (ws > 1e-6 && ws >= 0.05 * sa) {
                vu = au / ws;
                vv = av / ws;
            } else {
                vu = bilinear(u, cw, ch, scx, scy);
                vv = bilinear(v, cw, ch, scx, scy);
            }
            if (clamp_hull) {
                vu = clampd(vu, (double)(mnu - margin), (double)(mxu + margin));
                vv = clampd(vv, (double)(mnv - margin), (double)(mxv + margin));
            }
            out_u[(size_t)oy * ow + ox] = (float)vu;
            out_v[(size_t)oy * ow + ox] = (float)vv;
        }
    }
}

void or_plain_chroma(const or_params *p, const float *u, const float *v, int cw, int ch,
                     int sw, int sh, float *out_u, float *out_v) {
    const double rw = (double)sw / cw, rh = (double)sh / ch;
    or_axis ax, ay;
    or_build_axis(sw, p->out_w, rw, p->shift_x, p, &ax);
    or_build_axis(sh, p->out_h, rh, p->shift_y, p, &ay);
    if (p->radial) {
        radial_table rt;
        build_radial_table(p, &ax, &ay, cw, ch, &rt);
        plain_radial(p, u, v, cw, ch, &ax, &ay, &rt, p->out_w, p->out_h, out_u, out_v);
        free_radial_table(&rt);
    } else if (p->kernel == OR_BILINEAR)
        plain_bilinear(u, v, cw, ch, &ax, &ay, p->out_w, p->out_h, out_u, out_v);
    else
        plain_separable(p, u, v, cw, ch, &ax, &ay, p->out_w, p->out_h, out_u, out_v);
    or_free_axis(&ax);
    or_free_axis(&ay);
}

/* ------------------------------------------------------------------------ */
/* guided reconstruction: recon.cpp:14-819 (scalar branches)                 */
/* ------------------------------------------------------------------------ */

typedef struct {
    float luma, nx, ny, coherence, min_diff, ridge, phase, mfade;
    int has_dir;
} guide_px;

static void guide_meta(const or_params *p, const or_guide *g, const or_axis *ax,
                       const or_axis *ay, int cw, int ch, float sp_out, int ox, int oy,
                       guide_px *m) {
    const int bxi = ax->lb_i0[ox], byi = ay->lb_i0[oy];
    const float bxf = ax->lb_f[ox], byf = ay->lb_f[oy];
    const int sw = g->sw, sh = g->sh;
    m->luma = bilinear_fast(g->y, sw, sh, bxi, bxf, byi, byf);
    const float jxx = bilinear_fast(g->jxx, sw, sh, bxi, bxf, byi, byf);
    const float jxy = bilinear_fast(g->jxy, sw, sh, bxi, bxf, byi, byf);
    const float jyy = bilinear_fast(g->jyy, sw, sh, bxi, bxf, byi, byf);
    const float jsum = jxx + jyy;
    const tdir d = principal_dir(jxx, jxy, jyy);
    m->coherence = d.coherence;
    m->nx = d.nx;
    m->ny = d.ny;
    m->has_dir = jsum > 1e-6f;
    m->min_diff = 0.0f;
    m->ridge = 0.0f;
    const float lx = ax->lpos[ox], ly = ay->lpos[oy];
    if (m->has_dir) {
        const float ya = bilinear(g->y, sw, sh, lx - sp_out * m->nx, ly - sp_out * m->ny);
        const float yb = bilinear(g->y, sw, sh, lx + sp_out * m->nx, ly + sp_out * m->ny);
        m->min_diff = minf(fabsf(ya - m->luma), fabsf(yb - m->luma));
    }
    if (p->ridge && m->has_dir) {
        const float dl = 1.5f * sp_out;
        const float ym2 = bilinear(g->y, sw, sh, lx - 2 * dl * m->nx, ly - 2 * dl * m->ny);
        const float ym1 = bilinear(g->y, sw, sh, lx - dl * m->nx, ly - dl * m->ny);
        const float yp1 = bilinear(g->y, sw, sh, lx + dl * m->nx, ly + dl * m->ny);
        const float yp2 = bilinear(g->y, sw, sh, lx + 2 * dl * m->nx, ly + 2 * dl * m->ny);
        const float diff[4] = { ym1 - ym2, m->luma - ym1, yp1 - m->luma, yp2 - yp1 };
        for (int q = 0; q < 3; ++q) {
            const float a = diff[q], b = diff[q + 1];
            if (a > 0.0f && b < 0.0f)
                m->ridge = maxf(m->ridge, minf(a, -b));
            if (a < 0.0f && b > 0.0f)
                m->ridge = maxf(m->ridge, minf(-a, b));
        }
    }
    const float fx = ax->pos[ox] - floorf(ax->pos[ox]);
    const float fy = ay->pos[oy] - floorf(ay->pos[oy]);
    const float phx = minf(fx, 1.0f - fx);
    const float phy = minf(fy, 1.0f - fy);
    m->phase = (1.0f - 2.0f * phx) * (1.0f - 2.0f * phy);
    m->mfade = 1.0f;
    if (p->strength > 0.0 && p->ms > 0.0 && g->ms) {
        const float gate = bilinear_fast(g->ms, cw, ch, ax->cb_i0[ox], ax->cb_f[ox],
                                         ay->cb_i0[oy], ay->cb_f[oy]);
        m->mfade = 1.0f - (float)p->ms * (1.0f - gate);
    }
}

typedef struct {
    double vu, vv;
    float ss_ramp, guide_fade, nx, ny, L0;
    int has_dir;
} gres;

static void guided_pixel(const or_params *p, const float *U, const float *V, int cw, int ch,
                         const or_guide *g, const or_axis *ax, const or_axis *ay,
                         const radial_table *rt, float sp_out, double rw, double rh, int ox, int oy,
                         gres *res) {
    const float *radial_base = radial_at(rt, ox, oy);
    const float inv_wxc = (float)fmin(1.0, (double)p->out_w / cw);
    const float inv_wyc = (float)fmin(1.0, (double)p->out_h / ch);
    const int guided = p->strength > 0.0;
    guide_px meta;
    memset(&meta, 0, sizeof(meta));
    meta.mfade = 1.0f;
    if (guided)
        guide_meta(p, g, ax, ay, cw, ch, sp_out, ox, oy, &meta);
    const int tx0 = ax->start[ox], ty0 = ay->start[oy];
    const float *wxp = ax->w + (size_t)ox * ax->sup;
    const float *wyp = ay->w + (size_t)oy * ay->sup;
    const float *amx = ax->am + (size_t)ox * ax->sup;
    const float *amy = ay->am + (size_t)oy * ay->sup;
    const float scx = ax->pos[ox], scy = ay->pos[oy];
    const float L0 = meta.luma;
    const float nxv = meta.nx, nyv = meta.ny;
    const float *lc = g ? g->lc : NULL;

    /* pass 1: window luma range and the 2D chroma hull (recon.cpp:304-363) */
    float max_dl = 0.0f;
    float mnu = 1e30f, mxu = -1e30f, mnv = 1e30f, mxv = -1e30f;
    for (int j = 0; j < ay->sup; ++j) {
        if (amy[j] == 0.0f)
            continue;
        const int ty = clampi(ty0 + j, 0, ch - 1);
        for (int i = 0; i < ax->sup; ++i) {
            if (amx[i] == 0.0f)
                continue;
            const int tx = clampi(tx0 + i, 0, cw - 1);
            if (guided)
                max_dl = maxf(max_dl, fabsf(lc[(size_t)ty * cw + tx] - L0));
            const float a = U[(size_t)ty * cw + tx], b = V[(size_t)ty * cw + tx];
            mnu = minf(mnu, a); mxu = maxf(mxu, a);
            mnv = minf(mnv, b); mxv = maxf(mxv, b);
        }
    }
    /* adaptive sigma, ridge fade (recon.cpp:369-376) */
    const float t2 = clampf(meta.min_diff / (0.1f * max_dl + 1e-6f), 0.0f, 1.0f);
    const float ss_ramp = t2 * t2 * (3.0f - 2.0f * t2);
    const float sig_mult = (float)p->sratio + ((float)p->sdb - (float)p->sratio) * ss_ramp;
    const float sig_loc = maxf((float)p->sigma, sig_mult * max_dl);
    const float inv_sigma2 = 1.0f / (sig_loc * sig_loc);
    const float ridge_fade = 1.0f - clampf(2.0f * meta.ridge / (max_dl + 1e-6f), 0.0f, 1.0f);
    /* chroma-edge fade (recon.cpp:385-407) */
    float cedge_fade = 1.0f;
    if (p->cedge && guided && meta.has_dir) {
        const float rc = maxf(mxu - mnu, mxv - mnv);
        if (rc > 1e-6f) {
            float gmax = 0.0f;
            const float anx = fabsf(nxv), any = fabsf(nyv);
            for (int j = 0; j < ay->sup; ++j) {
                const int ty = clampi(ty0 + j, 0, ch - 1);
                for (int i = 0; i < ax->sup; ++i) {
                    const int tx = clampi(tx0 + i, 0, cw - 1);
                    const int tx1 = mini(tx + 1, cw - 1);
                    const int ty1 = mini(ty + 1, ch - 1);
                    const float gu = maxf(fabsf(at(U, cw, ch, tx1, ty) - at(U, cw, ch, tx, ty)) * anx,
                                          fabsf(at(U, cw, ch, tx, ty1) - at(U, cw, ch, tx, ty)) * any);
                    const float gv = maxf(fabsf(at(V, cw, ch, tx1, ty) - at(V, cw, ch, tx, ty)) * anx,
                                          fabsf(at(V, cw, ch, tx, ty1) - at(V, cw, ch, tx, ty)) * any);
                    gmax = maxf(gmax, maxf(gu, gv));
                }
            }
            const float wc = rc / (gmax + 1e-6f);
            cedge_fade = clampf((2.2f - wc) / 0.7f, 0.0f, 1.0f);
        }
    }
    const float guide_fade = ridge_fade * cedge_fade * meta.mfade * (float)p->strength;
    /* phase-0 rescue with kappa^2 (recon.cpp:425-475) */
    float rescue = 0.0f;
    if (guided) {
        const int ncx = clampi((int)lroundf(scx), 0, cw - 1);
        const int ncy = clampi((int)lroundf(scy), 0, ch - 1);
        const float dl0 = at(lc, cw, ch, ncx, ncy) - L0;
        const float sim0 = 1.0f / (1.0f + dl0 * dl0 * inv_sigma2);
        const float pg = meta.phase;
        rescue = (float)p->rescue * guide_fade * pg * pg * (1.0f - sim0);
        if (rescue > 0.0f) {
            float kappa;
            if (!p->radial) {
                float mxX = 0, ssX = 0, mxY = 0, ssY = 0;
                for (int i = 0; i < ax->sup; ++i)
                    if (amx[i] > 0.0f) {
                        const float w = wxp[i];
                        mxX = maxf(mxX, fabsf(w));
                        ssX += w * w;
                    }
                for (int j = 0; j < ay->sup; ++j)
                    if (amy[j] > 0.0f) {
                        const float w = wyp[j];
                        mxY = maxf(mxY, fabsf(w));
                        ssY += w * w;
                    }
                kappa = (mxX * mxX * mxY * mxY) / (ssX * ssY + 1e-12f);
            } else {
                float mx = 0, ss = 0;
                for (int j = 0; j < ay->sup; ++j) {
                    if (amy[j] == 0.0f)
                        continue;
                    const float dyc = ((float)(ty0 + j) - scy) * inv_wyc;
                    for (int i = 0; i < ax->sup; ++i) {
                        if (amx[i] == 0.0f)
                            continue;
                        const float dxc = ((float)(tx0 + i) - scx) * inv_wxc;
                        const float w = radial_base ? radial_base[(size_t)j * ax->sup + i]
                                                    : or_lut_lookup(p->lut, p->lut_n,
                                                                    sqrtf(dxc * dxc + dyc * dyc) * p->lut_scale);
                        mx = maxf(mx, fabsf(w));
                        ss += w * w;
                    }
                }
                kappa = (mx * mx) / (ss + 1e-12f);
            }
            rescue *= kappa * kappa;
        }
    }
    const float aniso = 1.0f + (float)p->stretch * meta.coherence;
    const float inv_g2 = 1.0f / (float)(p->gsigma * p->gsigma);
    const float inv_g2_par = inv_g2 / (aniso * aniso);
    const float frw = (float)rw, frh = (float)rh;

    /* tap loop (recon.cpp:581-627) */
    double acc_u = 0.0, acc_v = 0.0, wsum = 0.0, sum_abs = 0.0;
    for (int j = 0; j < ay->sup; ++j) {
        if (amy[j] == 0.0f)
            continue;
        const int ty = clampi(ty0 + j, 0, ch - 1);
        const float wy = wyp[j];
        const float dy = ((float)(ty0 + j) - scy) * frh;
        for (int i = 0; i < ax->sup; ++i) {
            if (amx[i] == 0.0f)
                continue;
            const int tx = clampi(tx0 + i, 0, cw - 1);
            float wx;
            if (radial_base) {
                wx = radial_base[(size_t)j * ax->sup + i];
            } else if (p->radial) {
                const float dxc = ((float)(tx0 + i) - scx) * inv_wxc;
                const float dyc = ((float)(ty0 + j) - scy) * inv_wyc;
                wx = or_lut_lookup(p->lut, p->lut_n, sqrtf(dxc * dxc + dyc * dyc) * p->lut_scale);
            } else {
                wx = wxp[i] * wy;
            }
            float w;
            if (guided) {
                const float dx = ((float)(tx0 + i) - scx) * frw;
                const float dL = lc[(size_t)ty * cw + tx] - L0;
                const float dp = dx * nxv + dy * nyv;
                const float dt = -dx * nyv + dy * nxv;
                const float sim_raw = 1.0f / (1.0f + dL * dL * inv_sigma2);
                const float sim = 1.0f - guide_fade * (1.0f - sim_raw);
                const float bump = 1.0f / (1.0f + dp * dp * inv_g2 + dt * dt * inv_g2_par);
                w = sim * (wx + rescue * bump);
            } else {
                w = wx;
            }
            acc_u += (double)w * U[(size_t)ty * cw + tx];
            acc_v += (double)w * V[(size_t)ty * cw + tx];
            wsum += w;
            sum_abs += fabsf(w);
        }
    }
    /* temporal taps of TRecon (recon.cpp:642-745) */
    if (guided && g->n_nbrs > 0) {
        const int lxi = clampi((int)ax->lpos[ox], 0, g->sw - 1);
        const int lyi = clampi((int)ay->lpos[oy], 0, g->sh - 1);
        for (int q = 0; q < g->n_nbrs; ++q) {
            const or_nbr *nb = &g->nbrs[q];
            const int bxx = mini(nb->bw - 1, lxi / nb->block);
            const int byy = mini(nb->bh - 1, lyi / nb->block);
            const size_t bi = (size_t)byy * nb->bw + bxx;
            const float tconf = nb->conf[bi];
            if (tconf < 0.05f)
                continue;
            const float mx = (float)nb->mvx[bi], my = (float)nb->mvy[bi];
            const int tnx = clampi((int)lroundf(scx + mx / frw), 0, cw - 1);
            const int tny = clampi((int)lroundf(scy + my / frh), 0, ch - 1);
            const float ddx = ((float)tnx - scx) * frw - mx;
            const float ddy = ((float)tny - scy) * frh - my;
            const int stat = (fabsf(mx) + fabsf(my) <= 1.0f) && tconf > 0.5f;
            float dL;
            if (stat) {
                dL = at(nb->lc, cw, ch, tnx, tny) - L0;
            } else {
                const double lx = (tnx + 0.5) * rw - 0.5 + p->shift_x - mx;
                const double ly = (tny + 0.5) * rh - 0.5 + p->shift_y - my;
                const int x0 = (int)ceil(lx - 0.5 * rw), x1 = (int)ceil(lx + 0.5 * rw);
                const int y0 = (int)ceil(ly - 0.5 * rh), y1 = (int)ceil(ly + 0.5 * rh);
                float md = 0.0f;
                for (int jj = y0; jj < y1; ++jj)
                    for (int ii = x0; ii < x1; ++ii)
                        md = maxf(md, fabsf(at(g->y, g->sw, g->sh, ii, jj) - L0));
                dL = md;
            }
            const float sig_t = stat ? sig_loc : 0.5f * sig_loc;
            const float sim_raw = 1.0f / (1.0f + dL * dL / (sig_t * sig_t));
            const float sim = stat ? 1.0f - guide_fade * (1.0f - sim_raw) : sim_raw;
            float kw;
            if (p->radial) {
                const float dxc = ddx / frw * inv_wxc, dyc = ddy / frh * inv_wyc;
                kw = or_lut_lookup(p->lut, p->lut_n, sqrtf(dxc * dxc + dyc * dyc) * p->lut_scale);
            } else {
                kw = (float)(or_kernel_eval(p->kernel, fabsf(ddx) / rw * inv_wxc, p->kp1, p->kp2) *
                             or_kernel_eval(p->kernel, fabsf(ddy) / rh * inv_wyc, p->kp1, p->kp2));
            }
            float nw = 1.0f;
            if (!stat) {
                const float fx = mx / frw, fy = my / frh;
                const float px = fx - floorf(fx), py = fy - floorf(fy);
                nw = maxf(fabsf(sinf((float)M_PI * px)), fabsf(sinf((float)M_PI * py)));
                nw *= 1.0f - ss_ramp;
            }
            const float w = tconf * sim * kw * nw;
            const float nu = at(nb->u, cw, ch, tnx, tny), nv = at(nb->v, cw, ch, tnx, tny);
            acc_u += (double)w * nu;
            acc_v += (double)w * nv;
            wsum += w;
            sum_abs += fabsf(w);
            mnu = minf(mnu, nu); mxu = maxf(mxu, nu);
            mnv = minf(mnv, nv); mxv = maxf(mxv, nv);
        }
    }
    double vu, vv;
    if (wsum > 1e-6 && wsum >= 0.05 * sum_abs) {
        vu = acc_u / wsum;
        vv = acc_v / wsum;
    } else {
        vu = bilinear(U, cw, ch, scx, scy);
        vv = bilinear(V, cw, ch, scx, scy);
    }
    if (p->ar_margin >= 0.0) {
        const float margin = (float)(p->ar_margin > 0.0 ? p->ar_margin : 0.0);
        vu = clampd(vu, (double)(mnu - margin), (double)(mxu + margin));
        vv = clampd(vv, (double)(mnv - margin), (double)(mxv + margin));
    }
    res->vu = vu;
    res->vv = vv;
    res->ss_ramp = ss_ramp;
    res->guide_fade = guide_fade;
    res->nx = nxv;
    res->ny = nyv;
    res->L0 = L0;
    res->has_dir = meta.has_dir;
}

/* algo-4 selector modes of reconstructChromaImpl (recon.cpp:762-811) */
enum { SEL_NONE = 0, SEL_MAPS, SEL_COMPRESSED, SEL_PLANES };
typedef struct {
    int mode;
    float strength;                       /* selectorStrength */
    const float *plain_u, *plain_v;       /* base (output grid) */
    const float *au, *bu, *cu, *av, *bv, *cv;  /* LGF maps (chroma grid), SEL_MAPS */
    float *du, *dv, *w3c;                 /* SEL_COMPRESSED, indexed by workset order */
    float *w2p, *w3p;                     /* SEL_PLANES, output grid */
} selector_cfg;

static void guided_run(const or_params *p, const float *u, const float *v, int cw, int ch,
                       const or_guide *g, const or_workset *ws, const selector_cfg *sel,
                       float *out_u, float *out_v) {
    const int sw = g->sw, sh = g->sh;
    const double rw = (double)sw / cw, rh = (double)sh / ch;
    or_axis ax, ay;
    or_build_axis(sw, p->out_w, rw, p->shift_x, p, &ax);
    or_build_axis(sh, p->out_h, rh, p->shift_y, p, &ay);
    const float sp_out = (float)(rw > rh ? rw : rh);
    const int ow = p->out_w, oh = p->out_h;
    radial_table rt;
    memset(&rt, 0, sizeof(rt));
    if (p->radial)
        build_radial_table(p, &ax, &ay, cw, ch, &rt);
    const int sparse = ws && ws->out_w == ow && ws->out_h == oh;
    const size_t total = sparse ? ws->n_indices : (size_t)ow * oh;
    for (size_t k = 0; k < total; ++k) {
        const uint32_t idx = sparse ? ws->indices[k] : (uint32_t)k;
        const int ox = (int)(idx % (uint32_t)ow), oy = (int)(idx / (uint32_t)ow);
        const size_t o = (size_t)oy * ow + ox;
        gres r;
        guided_pixel(p, u, v, cw, ch, g, &ax, &ay, &rt, sp_out, rw, rh, ox, oy, &r);
        double vu = r.vu, vv = r.vv;
        if (sel && sel->mode != SEL_NONE) {
            const float diag = r.has_dir ? minf(1.0f, 2.0f * fabsf(r.nx * r.ny)) : 0.0f;
            const float hedgy = (1.0f - r.ss_ramp) * r.guide_fade;
            const float w2 = hedgy * (1.0f - diag) * sel->strength;
            const float w3 = hedgy * diag * sel->strength;
            if (sel->mode == SEL_COMPRESSED) {
                const float bu = sel->plain_u[o], bv = sel->plain_v[o];
                if (w2 < 1e-4f && w3 < 1e-4f) {
                    sel->du[k] = 0.0f; sel->dv[k] = 0.0f; sel->w3c[k] = 0.0f;
                } else {
                    sel->du[k] = hedgy * (1.0f - diag) * ((float)vu - bu);
                    sel->dv[k] = hedgy * (1.0f - diag) * ((float)vv - bv);
                    sel->w3c[k] = hedgy * diag;
                }
                continue;  /* the compressed pass leaves the output alone */
            } else if (sel->mode == SEL_PLANES) {
                sel->w2p[o] = hedgy * (1.0f - diag);
                sel->w3p[o] = hedgy * diag;
            } else if (w2 >= 1e-4f || w3 >= 1e-4f) {
                const int cx = ax.cb_i0[ox], cy = ay.cb_i0[oy];
                const float cxf = ax.cb_f[ox], cyf = ay.cb_f[oy];
                const float bu = out_u[o], bv = out_v[o];
                const float au = bilinear_fast(sel->au, cw, ch, cx, cxf, cy, cyf);
                const float bbu = bilinear_fast(sel->bu, cw, ch, cx, cxf, cy, cyf);
                const float av = bilinear_fast(sel->av, cw, ch, cx, cxf, cy, cyf);
                const float bbv = bilinear_fast(sel->bv, cw, ch, cx, cxf, cy, cyf);
                const float cu = bilinear_fast(sel->cu, cw, ch, cx, cxf, cy, cyf);
                const float cv = bilinear_fast(sel->cv, cw, ch, cx, cxf, cy, cyf);
                vu = bu + w2 * ((float)vu - bu) + w3 * cu * (au * r.L0 + bbu - bu);
                vv = bv + w2 * ((float)vv - bv) + w3 * cv * (av * r.L0 + bbv - bv);
            } else {
                vu = out_u[o];
                vv = out_v[o];
            }
        }
        out_u[o] = (float)vu;
        out_v[o] = (float)vv;
    }
    free_radial_table(&rt);
    or_free_axis(&ax);
    or_free_axis(&ay);
}

void or_guided(const or_params *p, const float *u, const float *v, int cw, int ch,
               const or_guide *g, const or_workset *ws, float *out_u, float *out_v) {
    guided_run(p, u, v, cw, ch, g, ws, NULL, out_u, out_v);
}


/* ------------------------------------------------------------------------ */
/* algo 4: LGF regression maps (algos.cpp:12-68, 241-315, 366-448)           */
/* ------------------------------------------------------------------------ */

/* pointSampledLuma (algos.cpp:300-315) */
static void point_luma(const float *y, int w, int h, int cw, int ch, double rw, double rh,
                       double sx, double sy, float *ls) {
    float *xp = (float *)malloc(sizeof(float) * (size_t)cw), *yp = (float *)malloc(sizeof(float) * (size_t)ch);
    int *xi = (int *)malloc(sizeof(int) * (size_t)cw), *yi = (int *)malloc(sizeof(int) * (size_t)ch);
    float *xf = (float *)malloc(sizeof(float) * (size_t)cw), *yf = (float *)malloc(sizeof(float) * (size_t)ch);
    for (int x = 0; x < cw; ++x) xp[x] = (float)((x + 0.5) * rw - 0.5 + sx);
    for (int j = 0; j < ch; ++j) yp[j] = (float)((j + 0.5) * rh - 0.5 + sy);
    or_bilin_axis(xp, cw, w, xi, xf);
    or_bilin_axis(yp, ch, h, yi, yf);
    for (int j = 0; j < ch; ++j)
        for (int x = 0; x < cw; ++x)
            ls[(size_t)j * cw + x] = bilinear_fast(y, w, h, xi[x], xf[x], yi[j], yf[j]);
    free(xp); free(yp); free(xi); free(yi); free(xf); free(yf);
}

/* slidingMinMax / minMax5 (algos.cpp:70-139, 241-298): min and max over
 * the in-bounds (2r+1)^2 window */
static void window_minmax(const float *src, int w, int h, int r, float *mn, float *mx) {
    float *hmn = (float *)malloc(sizeof(float) * (size_t)w * h), *hmx = (float *)malloc(sizeof(float) * (size_t)w * h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            float lo = src[(size_t)y * w + maxi(0, x - r)], hi = lo;
            for (int i = maxi(0, x - r); i <= mini(w - 1, x + r); ++i) {
                lo = minf(lo, src[(size_t)y * w + i]);
                hi = maxf(hi, src[(size_t)y * w + i]);
            }
            hmn[(size_t)y * w + x] = lo;
            hmx[(size_t)y * w + x] = hi;
        }
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            float lo = hmn[(size_t)maxi(0, y - r) * w + x], hi = hmx[(size_t)maxi(0, y - r) * w + x];
            for (int j = maxi(0, y - r); j <= mini(h - 1, y + r); ++j) {
                lo = minf(lo, hmn[(size_t)j * w + x]);
                hi = maxf(hi, hmx[(size_t)j * w + x]);
            }
            mn[(size_t)y * w + x] = lo;
            mx[(size_t)y * w + x] = hi;
        }
    free(hmn); free(hmx);
}

typedef struct { float *au, *bu, *cu, *av, *bv, *cv; } lgf_maps;

/* buildLGFPair (algos.cpp:366-448) with the rolling double window sums of
 * forEachWindowSum<6> (algos.cpp:12-68); roi == NULL or a dense ROI
 * finalises every sample, else only ROI samples (the rest stay 0). */
static void lgf_pair(const float *y, int w, int h, int cw, int ch, double rw, double rh,
                     double sx, double sy, const float *U, const float *V, int r, double eps,
                     int cedge, const uint8_t *roi, lgf_maps *m) {
    const size_t nc = (size_t)cw * ch;
    float *ls = (float *)malloc(sizeof(float) * nc);
    point_luma(y, w, h, cw, ch, rw, rh, sx, sy, ls);
    float *mnu = NULL, *mxu = NULL, *mnv = NULL, *mxv = NULL;
    if (cedge) {
        mnu = (float *)malloc(sizeof(float) * nc); mxu = (float *)malloc(sizeof(float) * nc);
        mnv = (float *)malloc(sizeof(float) * nc); mxv = (float *)malloc(sizeof(float) * nc);
        window_minmax(U, cw, ch, r, mnu, mxu);
        window_minmax(V, cw, ch, r, mnv, mxv);
    }
    int dense = !roi;
    if (roi) {
        size_t a = 0;
        for (size_t i = 0; i < nc; ++i) a += roi[i] != 0;
        dense = a * 2 >= nc;
    }
    if (roi)
        for (size_t i = 0; i < nc; ++i)
            m->au[i] = m->bu[i] = m->cu[i] = m->av[i] = m->bv[i] = m->cv[i] = 0.0f;
    enum { N = 6 };
    double *cols = (double *)calloc((size_t)cw * N, sizeof(double));
    const int rows = maxi(1, 2 * r + 2);
    double *cache = (double *)calloc((size_t)rows * cw * N, sizeof(double));
    int top_cur = 0, bot_cur = -1;
    for (int yy = 0; yy < ch; ++yy) {
        const int top = maxi(0, yy - r), bottom = mini(ch - 1, yy + r);
        while (bot_cur < bottom) {
            ++bot_cur;
            double *c = cache + (size_t)(bot_cur % rows) * cw * N;
            for (int x = 0; x < cw; ++x) {
                const double l = ls[(size_t)bot_cur * cw + x];
                const double u = U[(size_t)bot_cur * cw + x], v = V[(size_t)bot_cur * cw + x];
                double *val = c + (size_t)x * N;
                val[0] = l; val[1] = l * l; val[2] = u; val[3] = l * u; val[4] = v; val[5] = l * v;
                for (int k = 0; k < N; ++k) cols[(size_t)x * N + k] += val[k];
            }
        }
        while (top_cur < top) {
            const double *c = cache + (size_t)(top_cur % rows) * cw * N;
            for (int x = 0; x < cw; ++x)
                for (int k = 0; k < N; ++k) cols[(size_t)x * N + k] -= c[(size_t)x * N + k];
            ++top_cur;
        }
        double tot[N] = {0, 0, 0, 0, 0, 0};
        for (int x = 0; x <= mini(cw - 1, r); ++x)
            for (int k = 0; k < N; ++k) tot[k] += cols[(size_t)x * N + k];
        for (int x = 0; x < cw; ++x) {
            const int left = maxi(0, x - r), right = mini(cw - 1, x + r);
            const int n = (right - left + 1) * (bottom - top + 1);
            const size_t i = (size_t)yy * cw + x;
            if (dense || roi[i]) {
                const double mean_y = tot[0] / n;
                const double var = tot[1] / n - mean_y * mean_y;
                const double mean_u = tot[2] / n, mean_v = tot[4] / n;
                const double su = (tot[3] / n - mean_y * mean_u) / (var + eps);
                const double sv = (tot[5] / n - mean_y * mean_v) / (var + eps);
                m->au[i] = (float)su;
                m->bu[i] = (float)(mean_u - su * mean_y);
                m->av[i] = (float)sv;
                m->bv[i] = (float)(mean_v - sv * mean_y);
                const float base = (float)(var / (var + eps));
                float fu = 1.0f, fv = 1.0f;
                if (cedge) {
                    const float lgx = (at(ls, cw, ch, x + 1, yy) - at(ls, cw, ch, x - 1, yy)) * 0.5f;
                    const float lgy = (at(ls, cw, ch, x, yy + 1) - at(ls, cw, ch, x, yy - 1)) * 0.5f;
                    const float gl = hypotf(lgx, lgy);
                    for (int c = 0; c < 2; ++c) {
                        const float *P = c ? V : U;
                        const float range = (c ? mxv : mxu)[i] - (c ? mnv : mnu)[i];
                        float f = 1.0f;
                        if (!(gl <= 1e-9f || range <= 1e-6f)) {
                            const float cgx = (at(P, cw, ch, x + 1, yy) - at(P, cw, ch, x - 1, yy)) * 0.5f;
                            const float cgy = (at(P, cw, ch, x, yy + 1) - at(P, cw, ch, x, yy - 1)) * 0.5f;
                            const float gn = fabsf(cgx * lgx / gl + cgy * lgy / gl);
                            const float width = range / (gn + 1e-6f);
                            f = clampf((2.2f - width) / 0.7f, 0.0f, 1.0f);
                        }
                        if (c) fv = f; else fu = f;
                    }
                }
                m->cu[i] = base * fu;
                m->cv[i] = base * fv;
            }
            if (x - r >= 0)
                for (int k = 0; k < N; ++k) tot[k] -= cols[(size_t)(x - r) * N + k];
            if (x + r + 1 < cw)
                for (int k = 0; k < N; ++k) tot[k] += cols[(size_t)(x + r + 1) * N + k];
        }
    }
    free(cols); free(cache); free(ls);
    free(mnu); free(mxu); free(mnv); free(mxv);
}

/* ------------------------------------------------------------------------ */
/* back-projection: algos.cpp:663-677                                        */
/* ------------------------------------------------------------------------ */

void or_back_project(float *out, int ow, int oh, const float *src, int cw, int ch, float bp) {
    for (int cy = 0; cy < ch; ++cy)
        for (int cx = 0; cx < cw; ++cx) {
            const int x0 = 2 * cx, y0 = 2 * cy;
            const int xa = clampi(x0, 0, ow - 1), xb = clampi(x0 + 1, 0, ow - 1);
            const int ya = clampi(y0, 0, oh - 1), yb = clampi(y0 + 1, 0, oh - 1);
            const float p00 = out[(size_t)ya * ow + xa], p10 = out[(size_t)ya * ow + xb];
            const float p01 = out[(size_t)yb * ow + xa], p11 = out[(size_t)yb * ow + xb];
            const float r = src[(size_t)cy * cw + cx] - 0.25f * (p00 + p10 + p01 + p11);
            const float d = bp * r;
            out[(size_t)ya * ow + xa] += d;
            out[(size_t)ya * ow + xb] += d;
            out[(size_t)yb * ow + xa] += d;
            out[(size_t)yb * ow + xb] += d;
        }
}


void or_lgf_maps(const float *y, int w, int h, int cw, int ch, double rw, double rh, double sx,
                 double sy, const float *u, const float *v, double eps, int cedge,
                 const uint8_t *roi, float *out6) {
    const size_t nc = (size_t)cw * ch;
    lgf_maps m = { out6, out6 + nc, out6 + 2 * nc, out6 + 3 * nc, out6 + 4 * nc, out6 + 5 * nc };
    lgf_pair(y, w, h, cw, ch, rw, rh, sx, sy, u, v, 2, eps, cedge, roi, &m);
}

/* algo 4 orchestration (plugin.cpp:396-572) on a plain-filled output */
static void recon_algo4(const or_params *p, or_debug *D, const or_axis *ax, const or_axis *ay,
                        int is_float, float *out_u, float *out_v) {
    const int cw = D->cw, ch = D->ch, ow = D->ow, oh = D->oh;
    const size_t nc = (size_t)cw * ch, no = (size_t)ow * oh;
    const double rw = (double)D->sw / cw, rh = (double)D->sh / ch;
    or_params p2 = *p;
    p2.algo = 2;
    p2.strength = 1.0;
    lgf_maps m;
    float *lgf = (float *)malloc(sizeof(float) * nc * 6);
    m.au = lgf; m.bu = lgf + nc; m.cu = lgf + 2 * nc; m.av = lgf + 3 * nc; m.bv = lgf + 4 * nc; m.cv = lgf + 5 * nc;
    or_guide g = { D->y, D->sw, D->sh, D->jxx, D->jxy, D->jyy, D->lc, D->ms, NULL, 0 };
    const or_workset *ws = D->has_ws ? &D->ws : NULL;
    const float lam = (float)p->strength;
    if (!is_float) {
        lgf_pair(D->y, D->sw, D->sh, cw, ch, rw, rh, p->shift_x, p->shift_y, D->u, D->v, 2,
                 p->reg * p->reg, p->cedge, NULL, &m);
        if (D->ms)
            for (size_t i = 0; i < nc; ++i) {
                const float f = 1.0f - (float)p->ms * (1.0f - D->ms[i]);
                m.cu[i] *= f;
                m.cv[i] *= f;
            }
        selector_cfg sel;
        memset(&sel, 0, sizeof(sel));
        sel.mode = SEL_MAPS;
        sel.strength = lam;
        sel.au = m.au; sel.bu = m.bu; sel.cu = m.cu; sel.av = m.av; sel.bv = m.bv; sel.cv = m.cv;
        guided_run(&p2, D->u, D->v, cw, ch, &g, ws, &sel, out_u, out_v);
        free(lgf);
        return;
    }
    const int compressed = ws && !or_workset_output_dense(ws);
    selector_cfg sel;
    memset(&sel, 0, sizeof(sel));
    sel.strength = lam;
    float *gu = NULL, *gv = NULL;
    if (compressed) {
        sel.mode = SEL_COMPRESSED;
        sel.plain_u = out_u; sel.plain_v = out_v;
        sel.du = (float *)malloc(sizeof(float) * (ws->n_indices + 1));
        sel.dv = (float *)malloc(sizeof(float) * (ws->n_indices + 1));
        sel.w3c = (float *)malloc(sizeof(float) * (ws->n_indices + 1));
        guided_run(&p2, D->u, D->v, cw, ch, &g, ws, &sel, out_u, out_v);
    } else {
        sel.mode = SEL_PLANES;
        gu = (float *)calloc(no, sizeof(float)); gv = (float *)calloc(no, sizeof(float));
        sel.w2p = (float *)calloc(no, sizeof(float)); sel.w3p = (float *)calloc(no, sizeof(float));
        guided_run(&p2, D->u, D->v, cw, ch, &g, ws, &sel, gu, gv);
    }
    /* ROI of the diagonal route (plugin.cpp:496-531) */
    uint8_t *roi = (uint8_t *)calloc(nc, 1);
    const size_t total = ws ? ws->n_indices : no;
    for (size_t k = 0; k < total; ++k) {
        const uint32_t idx = ws ? ws->indices[k] : (uint32_t)k;
        const int ox = (int)(idx % (uint32_t)ow), oy = (int)(idx / (uint32_t)ow);
        const float w3 = compressed ? sel.w3c[k] : sel.w3p[idx];
        if (w3 < 1e-4f)
            continue;
        const int cx0 = ax->cb_i0[ox], cx1 = mini(cx0 + 1, cw - 1);
        const int cy0 = ay->cb_i0[oy], cy1 = mini(cy0 + 1, ch - 1);
        roi[(size_t)cy0 * cw + cx0] = roi[(size_t)cy0 * cw + cx1] = 1;
        roi[(size_t)cy1 * cw + cx0] = roi[(size_t)cy1 * cw + cx1] = 1;
    }
    lgf_pair(D->y, D->sw, D->sh, cw, ch, rw, rh, p->shift_x, p->shift_y, D->u, D->v, 2,
             p->reg * p->reg, p->cedge, roi, &m);
    if (D->ms)
        for (size_t i = 0; i < nc; ++i) {
            if (!roi[i])
                continue;
            const float f = 1.0f - (float)p->ms * (1.0f - D->ms[i]);
            m.cu[i] *= f;
            m.cv[i] *= f;
        }
    /* blend (blendSelectorCompressed / blendSelector, algos.cpp:453-593) */
    for (size_t k = 0; k < total; ++k) {
        const uint32_t idx = ws ? ws->indices[k] : (uint32_t)k;
        const int ox = (int)(idx % (uint32_t)ow), oy = (int)(idx / (uint32_t)ow);
        const int cx = ax->cb_i0[ox], cy = ay->cb_i0[oy];
        const float cxf = ax->cb_f[ox], cyf = ay->cb_f[oy];
        if (compressed) {
            const float pu = out_u[idx], pv = out_v[idx];
            const float w3 = lam * sel.w3c[k];
            float vu = pu + lam * sel.du[k], vv = pv + lam * sel.dv[k];
            if (w3 != 0.0f) {
                const float luma = bilinear_fast(D->y, D->sw, D->sh, ax->lb_i0[ox], ax->lb_f[ox], ay->lb_i0[oy], ay->lb_f[oy]);
                const float au = bilinear_fast(m.au, cw, ch, cx, cxf, cy, cyf);
                const float bu = bilinear_fast(m.bu, cw, ch, cx, cxf, cy, cyf);
                const float av = bilinear_fast(m.av, cw, ch, cx, cxf, cy, cyf);
                const float bv = bilinear_fast(m.bv, cw, ch, cx, cxf, cy, cyf);
                const float cu = bilinear_fast(m.cu, cw, ch, cx, cxf, cy, cyf);
                const float cv = bilinear_fast(m.cv, cw, ch, cx, cxf, cy, cyf);
                vu += w3 * cu * (au * luma + bu - pu);
                vv += w3 * cv * (av * luma + bv - pv);
            }
            out_u[idx] = vu;
            out_v[idx] = vv;
        } else {
            const float w2v = sel.w2p[idx] * lam, w3v = sel.w3p[idx] * lam;
            if (w2v < 1e-4f && w3v < 1e-4f)
                continue;
            const float L0 = bilinear_fast(D->y, D->sw, D->sh, ax->lb_i0[ox], ax->lb_f[ox], ay->lb_i0[oy], ay->lb_f[oy]);
            const float au = bilinear_fast(m.au, cw, ch, cx, cxf, cy, cyf);
            const float bu = bilinear_fast(m.bu, cw, ch, cx, cxf, cy, cyf);
            const float av = bilinear_fast(m.av, cw, ch, cx, cxf, cy, cyf);
            const float bv = bilinear_fast(m.bv, cw, ch, cx, cxf, cy, cyf);
            const float cu = bilinear_fast(m.cu, cw, ch, cx, cxf, cy, cyf);
            const float cv = bilinear_fast(m.cv, cw, ch, cx, cxf, cy, cyf);
            out_u[idx] += w2v * (gu[idx] - out_u[idx]) + w3v * cu * (au * L0 + bu - out_u[idx]);
            out_v[idx] += w2v * (gv[idx] - out_v[idx]) + w3v * cv * (av * L0 + bv - out_v[idx]);
        }
    }
    free(roi); free(lgf); free(gu); free(gv);
    free(sel.du); free(sel.dv); free(sel.w3c); free(sel.w2p); free(sel.w3p);
}

/* ------------------------------------------------------------------------ */
/* luma resampling: kernels.cpp:171-271 (scalar), 276-334 (radial, scalar)   */
/* ------------------------------------------------------------------------ */

static void resample_h(const float *src, int sw, int sh, float *dst, const or_weights *t) {
    for (int y = 0; y < sh; ++y)
        for (int x = 0; x < t->n; ++x) {
            const float *wp = t->w + (size_t)x * t->sup;
            const int s = t->start[x];
            double acc = 0.0;
            for (int j = 0; j < t->sup; ++j)
                acc += (double)wp[j] * at(src, sw, sh, s + j, y);
            dst[(size_t)y * t->n + x] = (float)acc;
        }
}

static void resample_v(const float *src, int sw, int sh, float *dst, const or_weights *t) {
    for (int y = 0; y < t->n; ++y) {
        const float *wp = t->w + (size_t)y * t->sup;
        const int s = t->start[y];
        for (int x = 0; x < sw; ++x) {
            double acc = 0.0;
            for (int j = 0; j < t->sup; ++j)
                acc += (double)wp[j] * at(src, sw, sh, x, s + j);
            dst[(size_t)y * sw + x] = (float)acc;
        }
    }
}

static void resample_radial(const or_params *p, const float *src, int sw, int sh,
                            float *dst, int dw, int dh) {
    const double sx = (double)sw / dw, sy = (double)sh / dh;
    const double widen = fmax(1.0, fmax(sx, sy));
    const float inv_w = (float)(1.0 / widen);
    const int radius = (int)ceil(p->support * widen);
    const int sup = 2 * radius;
    for (int oy = 0; oy < dh; ++oy) {
        const double cy = (oy + 0.5) * sy - 0.5;
        const int y0 = (int)ceil(cy - radius);
        for (int ox = 0; ox < dw; ++ox) {
            const double cx = (ox + 0.5) * sx - 0.5;
            const int x0 = (int)ceil(cx - radius);
            double acc = 0.0, wsum = 0.0;
            for (int j = 0; j < sup; ++j) {
                const int ty = clampi(y0 + j, 0, sh - 1);
                const float dy = (float)((y0 + j - cy) * inv_w);
                for (int i = 0; i < sup; ++i) {
                    const int tx = clampi(x0 + i, 0, sw - 1);
                    const float dx = (float)((x0 + i - cx) * inv_w);
                    const float w = or_lut_lookup(p->lut, p->lut_n,
                                                  sqrtf(dx * dx + dy * dy) * p->lut_scale);
                    acc += (double)w * src[(size_t)ty * sw + tx];
                    wsum += w;
                }
            }
            dst[(size_t)oy * dw + ox] = (float)(wsum > 1e-6 ? acc / wsum : 0.0);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* algo 6: affine credibility maps, detail map, detail transfer             */
/* (maps.cpp:501-680, algos.cpp:809-1698)                                    */
/* ------------------------------------------------------------------------ */

static float median3f(float a, float b, float c) {
    const float mn = minf(minf(a, b), c) , mx = maxf(maxf(a, b), c);
    return a + b + c - mn - mx;
}

typedef struct {
    float *au, *av, *g, *mnu, *mxu, *mnv, *mxv, *rng, *deg;  /* chroma grid */
} affine_maps;

/* candidate degradations D_k(Y) on the chroma grid: box (= lc), bilinear and
 * bicubic(0, 0.6) polyphase downsamples (maps.cpp:512-680) */
static void yc_filtered(const float *y, int w, int h, int cw, int ch, double sx, double sy,
                        float *yc1, float *yc2) {
    for (int c = 0; c < 2; ++c) {
        const int k = c ? OR_BICUBIC : OR_BILINEAR;
        const double p1 = 0.0, p2 = c ? 0.6 : 0.0, sup = c ? 2.0 : 1.0;
        or_weights th, tv;
        or_build_weights(w, cw, k, p1, p2, sup, sx, &th);
        or_build_weights(h, ch, k, p1, p2, sup, sy, &tv);
        float *tmp = (float *)malloc(sizeof(float) * (size_t)cw * h);
        resample_h(y, w, h, tmp, &th);
        resample_v(tmp, cw, h, c ? yc2 : yc1, &tv);
        free(tmp);
        or_free_weights(&th);
        or_free_weights(&tv);
    }
}

/* the per-sample finalise shared by both affine paths; `inv` selects the
 * sparse path's reciprocal multiply (algos.cpp:1082-1120) over the dense
 * path's divisions (algos.cpp:1483-1512) */
static void affine_finalize(const double *sum, int n, int sparse_form, double eps, double eps_sig,
                            float *au, float *av, float *g) {
    double meanU, meanV, varU, varV, qk[3], aUk[3], aVk[3];
    const double inv = 1.0 / n;
    if (sparse_form) {
        meanU = sum[0] * inv; meanV = sum[2] * inv;
        varU = fmax(0.0, sum[1] * inv - meanU * meanU);
        varV = fmax(0.0, sum[3] * inv - meanV * meanV);
    } else {
        meanU = sum[0] / n; meanV = sum[2] / n;
        varU = fmax(0.0, sum[1] / n - meanU * meanU);
        varV = fmax(0.0, sum[3] / n - meanV * meanV);
    }
    for (int k = 0; k < 3; ++k) {
        const int b = 4 + 4 * k;
        double meanY, varY, covU, covV;
        if (sparse_form) {
            meanY = sum[b] * inv;
            varY = fmax(0.0, sum[b + 1] * inv - meanY * meanY);
            covU = sum[b + 2] * inv - meanY * meanU;
            covV = sum[b + 3] * inv - meanY * meanV;
        } else {
            meanY = sum[b] / n;
            varY = fmax(0.0, sum[b + 1] / n - meanY * meanY);
            covU = sum[b + 2] / n - meanY * meanU;
            covV = sum[b + 3] / n - meanY * meanV;
        }
        aUk[k] = clampd(covU / (varY + eps), -16.0, 16.0);
        aVk[k] = clampd(covV / (varY + eps), -16.0, 16.0);
        qk[k] = clampd((covU * covU + covV * covV) / ((varY + eps) * (varU + varV + eps)), 0.0, 1.0);
    }
    *au = median3f((float)aUk[0], (float)aUk[1], (float)aUk[2]);
    *av = median3f((float)aVk[0], (float)aVk[1], (float)aVk[2]);
    const double qmax = fmax(fmax(qk[0], qk[1]), qk[2]);
    const double qmin = fmin(fmin(qk[0], qk[1]), qk[2]);
    const double stab = qmax > 1e-9 ? qmin / qmax : 0.0;
    const double sig = (varU + varV) / (varU + varV + eps_sig);
    if (sparse_form)
        *g = (float)(((qk[0] + qk[1] + qk[2]) / 3.0) * stab * sig);
    else
        *g = (float)(((qk[0] + qk[1] + qk[2]) / 3.0) * stab * sig);
}

/* buildAffineMaps (algos.cpp:1382-1526): sparse ROI path = direct window sums
 * over chroma-mask samples (affineSparseFixed5, algos.cpp:1026-1137); dense
 * path = rolling sums (forEachWindowSum<16>) + sliding min/max. */
static void affine_maps_build(const or_params *p, const float *U, const float *V, int cw, int ch,
                              const float *yc[3], const uint8_t *cmask, int sparse_roi, affine_maps *m) {
    const size_t nc = (size_t)cw * ch;
    const double eps = p->reg * p->reg;
    const double eps_sig = (0.25 * p->sigma) * (0.25 * p->sigma);
    for (size_t i = 0; i < nc; ++i)
        m->deg[i] = median3f(yc[0][i], yc[1][i], yc[2][i]);
    if (sparse_roi) {
        for (size_t i = 0; i < nc; ++i)
            m->au[i] = m->av[i] = m->g[i] = m->mnu[i] = m->mxu[i] = m->mnv[i] = m->mxv[i] = m->rng[i] = 0.0f;
        for (int y = 0; y < ch; ++y)
            for (int x = 0; x < cw; ++x) {
                const size_t i = (size_t)y * cw + x;
                if (!cmask[i])
                    continue;
                double sum[16] = {0};
                const int top = maxi(0, y - 2), bot = mini(ch - 1, y + 2);
                const int lef = maxi(0, x - 2), rig = mini(cw - 1, x + 2);
                float hmnu = U[(size_t)top * cw + lef], hmxu = hmnu;
                float hmnv = V[(size_t)top * cw + lef], hmxv = hmnv;
                for (int py = top; py <= bot; ++py)
                    for (int px = lef; px <= rig; ++px) {
                        const size_t q = (size_t)py * cw + px;
                        const double u = U[q], v = V[q];
                        sum[0] += u; sum[1] += u * u; sum[2] += v; sum[3] += v * v;
                        hmnu = minf(hmnu, (float)u); hmxu = maxf(hmxu, (float)u);
                        hmnv = minf(hmnv, (float)v); hmxv = maxf(hmxv, (float)v);
                        for (int k = 0; k < 3; ++k) {
                            const double l = yc[k][q];
                            const int b = 4 + 4 * k;
                            sum[b] += l; sum[b + 1] += l * l; sum[b + 2] += l * u; sum[b + 3] += l * v;
                        }
                    }
                const int n = (rig - lef + 1) * (bot - top + 1);
                m->mnu[i] = hmnu; m->mxu[i] = hmxu; m->mnv[i] = hmnv; m->mxv[i] = hmxv;
                m->rng[i] = maxf(hmxu - hmnu, hmxv - hmnv);
                affine_finalize(sum, n, 1, eps, eps_sig, &m->au[i], &m->av[i], &m->g[i]);
            }
        return;
    }
    window_minmax(U, cw, ch, 2, m->mnu, m->mxu);
    window_minmax(V, cw, ch, 2, m->mnv, m->mxv);
    enum { N = 16 };
    const int r = 2;
    double *cols = (double *)calloc((size_t)cw * N, sizeof(double));
    const int rows = 2 * r + 2;
    double *cache = (double *)calloc((size_t)rows * cw * N, sizeof(double));
    int top_cur = 0, bot_cur = -1;
    for (int yy = 0; yy < ch; ++yy) {
        const int top = maxi(0, yy - r), bottom = mini(ch - 1, yy + r);
        while (bot_cur < bottom) {
            ++bot_cur;
            double *c = cache + (size_t)(bot_cur % rows) * cw * N;
            for (int x = 0; x < cw; ++x) {
                const size_t q = (size_t)bot_cur * cw + x;
                const double u = U[q], v = V[q];
                double *val = c + (size_t)x * N;
                val[0] = u; val[1] = u * u; val[2] = v; val[3] = v * v;
                for (int k = 0; k < 3; ++k) {
                    const double l = yc[k][q];
                    const int b = 4 + 4 * k;
                    val[b] = l; val[b + 1] = l * l; val[b + 2] = l * u; val[b + 3] = l * v;
                }
                for (int k = 0; k < N; ++k) cols[(size_t)x * N + k] += val[k];
            }
        }
        while (top_cur < top) {
            const double *c = cache + (size_t)(top_cur % rows) * cw * N;
            for (int x = 0; x < cw; ++x)
                for (int k = 0; k < N; ++k) cols[(size_t)x * N + k] -= c[(size_t)x * N + k];
            ++top_cur;
        }
        double tot[N];
        for (int k = 0; k < N; ++k) tot[k] = 0.0;
        for (int x = 0; x <= mini(cw - 1, r); ++x)
            for (int k = 0; k < N; ++k) tot[k] += cols[(size_t)x * N + k];
        for (int x = 0; x < cw; ++x) {
            const int left = maxi(0, x - r), right = mini(cw - 1, x + r);
            const int n = (right - left + 1) * (bottom - top + 1);
            const size_t i = (size_t)yy * cw + x;
            m->rng[i] = maxf(m->mxu[i] - m->mnu[i], m->mxv[i] - m->mnv[i]);
            affine_finalize(tot, n, 0, eps, eps_sig, &m->au[i], &m->av[i], &m->g[i]);
            if (x - r >= 0)
                for (int k = 0; k < N; ++k) tot[k] -= cols[(size_t)(x - r) * N + k];
            if (x + r + 1 < cw)
                for (int k = 0; k < N; ++k) tot[k] += cols[(size_t)(x + r + 1) * N + k];
        }
    }
    free(cols);
    free(cache);
}

static int nyquist_levels(double ratio) {
    int levels = 0;
    while (ratio > 1.5 && levels < 4) {
        ratio *= 0.5;
        ++levels;
    }
    return levels;
}

/* dyadic [1/4, 1/2, 1/4] passes with step 2^pass along x then y, edge
 * clamped (DetailRowPipeline / suppressSourceNyquist, algos.cpp:1148-1364) */
static void nyquist_passes(float *cur, int w, int h, double rw, double rh) {
    const int xl = nyquist_levels(rw), yl = nyquist_levels(rh);
    float *nx = (float *)malloc(sizeof(float) * (size_t)w * h);
    for (int pass = 0; pass < xl; ++pass) {
        const int step = 1 << pass;
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                const float *r = cur + (size_t)y * w;
                nx[(size_t)y * w + x] = 0.25f * r[maxi(0, x - step)] + 0.5f * r[x] + 0.25f * r[mini(w - 1, x + step)];
            }
        memcpy(cur, nx, sizeof(float) * (size_t)w * h);
    }
    for (int pass = 0; pass < yl; ++pass) {
        const int step = 1 << pass;
        for (int y = 0; y < h; ++y) {
            const float *t = cur + (size_t)maxi(0, y - step) * w, *m = cur + (size_t)y * w;
            const float *b = cur + (size_t)mini(h - 1, y + step) * w;
            for (int x = 0; x < w; ++x)
                nx[(size_t)y * w + x] = 0.25f * t[x] + 0.5f * m[x] + 0.25f * b[x];
        }
        memcpy(cur, nx, sizeof(float) * (size_t)w * h);
    }
    free(nx);
}

/* buildDetailMap (algos.cpp:1528-1585): detail = Y - P(deg) on the source
 * grid, Nyquist-suppressed, then resampled to the output grid */
static void detail_map(const or_params *p, const float *y, int sw, int sh, const float *deg, int cw,
                       int ch, float *detail_out) {
    const double rw = (double)sw / cw, rh = (double)sh / ch;
    or_params ps = *p;
    ps.out_w = sw;
    ps.out_h = sh;
    const size_t ns = (size_t)sw * sh;
    float *rec = (float *)malloc(sizeof(float) * ns), *tmp = (float *)malloc(sizeof(float) * ns);
    or_plain_chroma(&ps, deg, deg, cw, ch, sw, sh, rec, tmp);
    for (size_t i = 0; i < ns; ++i)
        rec[i] = y[i] - rec[i];
    nyquist_passes(rec, sw, sh, rw, rh);
    if (p->out_w == sw && p->out_h == sh) {
        memcpy(detail_out, rec, sizeof(float) * ns);
    } else if (p->radial) {
        resample_radial(p, rec, sw, sh, detail_out, p->out_w, p->out_h);
    } else {
        or_weights th, tv;
        or_build_weights(sw, p->out_w, p->kernel, p->kp1, p->kp2, p->support, 0.0, &th);
        or_build_weights(sh, p->out_h, p->kernel, p->kp1, p->kp2, p->support, 0.0, &tv);
        float *t2 = (float *)malloc(sizeof(float) * (size_t)p->out_w * sh);
        resample_h(rec, sw, sh, t2, &th);
        resample_v(t2, p->out_w, sh, detail_out, &tv);
        free(t2);
        or_free_weights(&th);
        or_free_weights(&tv);
    }
    free(rec);
    free(tmp);
}

/* detailTransferImpl (algos.cpp:1587-1689) */
static void detail_transfer(const or_params *p, const or_debug *D, const affine_maps *m,
                            const float *detail, const or_axis *ax, const or_axis *ay,
                            float *out_u, float *out_v) {
    const int ow = D->ow, oh = D->oh, cw = D->cw, ch = D->ch, sw = D->sw, sh = D->sh;
    const int use_ms = p->ms > 0.0 && D->ms;
    const float mss = (float)p->ms, qs = (float)p->qgate, st = (float)p->strength;
    const float ar = (float)(p->ar_margin > 0.0 ? p->ar_margin : 0.0);
    const int clamp_hull = p->ar_margin >= 0.0;
    const or_workset *ws = D->has_ws ? &D->ws : NULL;
    const size_t total = ws ? ws->n_indices : (size_t)ow * oh;
    for (size_t k = 0; k < total; ++k) {
        const uint32_t idx = ws ? ws->indices[k] : (uint32_t)k;
        const int ox = (int)(idx % (uint32_t)ow), oy = (int)(idx / (uint32_t)ow);
        const int lxi = ax->lb_i0[ox], lyi = ay->lb_i0[oy];
        const float lxf = ax->lb_f[ox], lyf = ay->lb_f[oy];
        const float jxx = bilinear_fast(D->jxx, sw, sh, lxi, lxf, lyi, lyf);
        const float jxy = bilinear_fast(D->jxy, sw, sh, lxi, lxf, lyi, lyf);
        const float jyy = bilinear_fast(D->jyy, sw, sh, lxi, lxf, lyi, lyf);
        const tdir d = principal_dir(jxx, jxy, jyy);
        const int cxi = ax->cb_i0[ox], cyi = ay->cb_i0[oy];
        const float cxf = ax->cb_f[ox], cyf = ay->cb_f[oy];
        const float q = bilinear_fast(m->g, cw, ch, cxi, cxf, cyi, cyf);
        float g = (1.0f - qs * (1.0f - q)) * st;
        if (use_ms) {
            const float msv = bilinear_fast(D->ms, cw, ch, cxi, cxf, cyi, cyf);
            g *= 1.0f - mss * (1.0f - msv);
        }
        const float au = bilinear_fast(m->au, cw, ch, cxi, cxf, cyi, cyf);
        const float av = bilinear_fast(m->av, cw, ch, cxi, cxf, cyi, cyf);
        const float coh = d.coherence;
        const float tx = -d.ny, ty = d.nx;
        const float dc = detail[idx];
        const float s = 1.0f + (float)p->stretch * coh;
        const float dtp = bilinear(detail, ow, oh, (float)ox + tx * s, (float)oy + ty * s);
        const float dtm = bilinear(detail, ow, oh, (float)ox - tx * s, (float)oy - ty * s);
        const float dt = 0.25f * dtp + 0.5f * dc + 0.25f * dtm;
        const float det = dc + (dt - dc) * coh;
        const float cap = 0.5f * bilinear_fast(m->rng, cw, ch, cxi, cxf, cyi, cyf);
        const float cu = clampf(g * au * det, -cap, cap);
        const float cv = clampf(g * av * det, -cap, cap);
        float ou = out_u[idx] + cu, ov = out_v[idx] + cv;
        if (clamp_hull) {
            const float bu = out_u[idx], bv = out_v[idx];
            const float lou = minf(bu, bilinear_fast(m->mnu, cw, ch, cxi, cxf, cyi, cyf) - ar);
            const float hiu = maxf(bu, bilinear_fast(m->mxu, cw, ch, cxi, cxf, cyi, cyf) + ar);
            const float lov = minf(bv, bilinear_fast(m->mnv, cw, ch, cxi, cxf, cyi, cyf) - ar);
            const float hiv = maxf(bv, bilinear_fast(m->mxv, cw, ch, cxi, cxf, cyi, cyf) + ar);
            ou = clampf(ou, lou, hiu);
            ov = clampf(ov, lov, hiv);
        }
        out_u[idx] = ou;
        out_v[idx] = ov;
    }
}

/* test access: the credibility g of the next algo-6 frame's affine maps
 * (chroma grid) */
static float *g_affine_debug;
void or_affine_debug(float *g) { g_affine_debug = g; }

/* algo 6 orchestration (plugin.cpp:573-626) on a plain-filled output */
static void recon_algo6(const or_params *p, or_debug *D, const or_axis *ax, const or_axis *ay,
                        float *out_u, float *out_v) {
    const int cw = D->cw, ch = D->ch;
    const size_t nc = (size_t)cw * ch;
    float *buf = (float *)malloc(sizeof(float) * nc * 11);
    affine_maps m = { buf, buf + nc, buf + 2 * nc, buf + 3 * nc, buf + 4 * nc, buf + 5 * nc,
                      buf + 6 * nc, buf + 7 * nc, buf + 8 * nc };
    float *yc1 = buf + 9 * nc, *yc2 = buf + 10 * nc;
    yc_filtered(D->y, D->sw, D->sh, cw, ch, p->shift_x, p->shift_y, yc1, yc2);
    const float *yc[3] = { D->lc, yc1, yc2 };
    const int sparse_roi = D->has_ws && !or_workset_chroma_dense(&D->ws);
    affine_maps_build(p, D->u, D->v, cw, ch, yc, D->has_ws ? D->ws.chroma_mask : NULL, sparse_roi, &m);
    if (g_affine_debug)
        memcpy(g_affine_debug, m.g, sizeof(float) * nc);
    float *detail = (float *)malloc(sizeof(float) * (size_t)D->ow * D->oh);
    detail_map(p, D->y, D->sw, D->sh, m.deg, cw, ch, detail);
    detail_transfer(p, D, &m, detail, ax, ay, out_u, out_v);
    free(detail);
    free(buf);
}

/* ------------------------------------------------------------------------ */
/* collaborative (BM3D-style) chroma filter: bm.cpp:15-442                   */
/* ------------------------------------------------------------------------ */

#define BM_P 8
#define BM_G 4
static const float bm_inv_sqrt2 = 0.7071067811865475f;

typedef struct { float cost; int x, y; } bm_match;

static void bm_insert(bm_match *m, int *count, bm_match c) {
    int pos = mini(*count, BM_G - 1);
    if (*count == BM_G && c.cost >= m[BM_G - 1].cost)
        return;
    if (*count < BM_G)
        ++*count;
    while (pos > 0 && c.cost < m[pos - 1].cost) {
        if (pos < BM_G)
            m[pos] = m[pos - 1];
        --pos;
    }
    m[pos] = c;
}

static void haar_fwd(float *v, int n) {
    float t[BM_P];
    for (int len = n; len > 1; len /= 2) {
        const int half = len / 2;
        for (int i = 0; i < half; ++i) {
            const float a = v[2 * i], b = v[2 * i + 1];
            t[i] = (a + b) * bm_inv_sqrt2;
            t[half + i] = (a - b) * bm_inv_sqrt2;
        }
        memcpy(v, t, sizeof(float) * (size_t)len);
    }
}

static void haar_inv(float *v, int n) {
    float t[BM_P];
    for (int len = 2; len <= n; len *= 2) {
        const int half = len / 2;
        for (int i = 0; i < half; ++i) {
            const float a = v[i], b = v[half + i];
            t[2 * i] = (a + b) * bm_inv_sqrt2;
            t[2 * i + 1] = (a - b) * bm_inv_sqrt2;
        }
        memcpy(v, t, sizeof(float) * (size_t)len);
    }
}

static void haar2d_fwd(float *p) {
    float col[BM_P];
    for (int y = 0; y < BM_P; ++y)
        haar_fwd(p + y * BM_P, BM_P);
    for (int x = 0; x < BM_P; ++x) {
        for (int y = 0; y < BM_P; ++y) col[y] = p[y * BM_P + x];
        haar_fwd(col, BM_P);
        for (int y = 0; y < BM_P; ++y) p[y * BM_P + x] = col[y];
    }
}

static void haar2d_inv(float *p) {
    float col[BM_P];
    for (int x = 0; x < BM_P; ++x) {
        for (int y = 0; y < BM_P; ++y) col[y] = p[y * BM_P + x];
        haar_inv(col, BM_P);
        for (int y = 0; y < BM_P; ++y) p[y * BM_P + x] = col[y];
    }
    for (int y = 0; y < BM_P; ++y)
        haar_inv(p + y * BM_P, BM_P);
}

static int block_starts(int extent, int *starts) {
    int n = 0;
    if (extent < BM_P)
        return 0;
    const int last = extent - BM_P;
    for (int pos = 0; pos <= last; pos += 8)
        starts[n++] = pos;
    if (n == 0 || starts[n - 1] != last)
        starts[n++] = last;
    return n;
}

static float patch_cost(const float *g, const float *u, const float *v, int w, int ax, int ay, int bx, int by) {
    float cost = 0.0f;
    int samples = 0;
    for (int py = 0; py < BM_P; py += 2)
        for (int px = 0; px < BM_P; px += 2) {
            const size_t a = (size_t)(ay + py) * w + ax + px, b = (size_t)(by + py) * w + bx + px;
            const float dg = g[a] - g[b], du = u[a] - u[b], dv = v[a] - v[b];
            cost += dg * dg + 0.25f * (du * du + dv * dv);
            ++samples;
        }
    return cost / samples;
}

/* collaborativeChromaFilter (bm.cpp:250-442); guide is the source luma,
 * u/v the reconstructed output chroma (modified in place). */
void or_collaborative(const float *guide, int gw, int gh, float *u, float *v, int w, int h, float guide_sigma) {
    if (w < BM_P || h < BM_P)
        return;
    const size_t n = (size_t)w * h;
    float *rg = NULL;
    const float *mg = guide;
    if (gw != w || gh != h) {
        rg = (float *)malloc(sizeof(float) * n);
        const double sx = (double)gw / w, sy = (double)gh / h;
        for (int y = 0; y < h; ++y) {
            const double gy = (y + 0.5) * sy - 0.5;
            for (int x = 0; x < w; ++x)
                rg[(size_t)y * w + x] = bilinear(guide, gw, gh, (x + 0.5) * sx - 0.5, gy);
        }
        mg = rg;
    }
    float *su = (float *)malloc(sizeof(float) * n), *sv = (float *)malloc(sizeof(float) * n);
    float *wt = (float *)calloc(n, sizeof(float));
    memcpy(su, u, sizeof(float) * n);
    memcpy(sv, v, sizeof(float) * n);
    memset(u, 0, sizeof(float) * n);
    memset(v, 0, sizeof(float) * n);
    int *sxs = (int *)malloc(sizeof(int) * (size_t)(w / 8 + 2)), *sys = (int *)malloc(sizeof(int) * (size_t)(h / 8 + 2));
    const int nx = block_starts(w, sxs), ny = block_starts(h, sys);
    const float sigma = clampf(0.6f * guide_sigma, 0.002f, 0.012f);
    const float thr = 2.7f * sigma;
    const float limit = maxf(4e-5f, 4.0f * guide_sigma * guide_sigma);
    static const float win[BM_P] = {0.25f, 0.50f, 0.75f, 1.0f, 1.0f, 0.75f, 0.50f, 0.25f};
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            const int ax = sxs[ix], ay = sys[iy];
            bm_match m[BM_G];
            for (int k = 0; k < BM_G; ++k) { m[k].cost = INFINITY; m[k].x = m[k].y = 0; }
            int mc = 0, cx[25], cy[25], cc = 0;
            for (int dy = -8; dy <= 8; dy += 4)
                for (int dx = -8; dx <= 8; dx += 4) {
                    const int bx = clampi(ax + dx, 0, w - BM_P), by = clampi(ay + dy, 0, h - BM_P);
                    int dup = 0;
                    for (int i = 0; i < cc; ++i)
                        if (cx[i] == bx && cy[i] == by) { dup = 1; break; }
                    if (dup)
                        continue;
                    cx[cc] = bx; cy[cc] = by; ++cc;
                    bm_match c = { bx == ax && by == ay ? 0.0f : patch_cost(mg, su, sv, w, ax, ay, bx, by), bx, by };
                    bm_insert(m, &mc, c);
                }
            int elig = 0;
            while (elig < mc && m[elig].cost <= limit)
                ++elig;
            const int gc = elig >= 4 ? 4 : (elig >= 2 ? 2 : 1);
            if (gc < 2)
                continue;
            float fu[BM_G * 64], fv[BM_G * 64], mnu[64], mxu[64], mnv[64], mxv[64];
            memset(fu, 0, sizeof(fu)); memset(fv, 0, sizeof(fv));
            for (int o = 0; o < 64; ++o) { mnu[o] = mnv[o] = INFINITY; mxu[o] = mxv[o] = -INFINITY; }
            for (int k = 0; k < gc; ++k)
                for (int py = 0; py < BM_P; ++py)
                    for (int px = 0; px < BM_P; ++px) {
                        const int o = py * BM_P + px;
                        const size_t q = (size_t)(m[k].y + py) * w + m[k].x + px;
                        const float a = su[q], b = sv[q];
                        fu[o * BM_G + k] = a; fv[o * BM_G + k] = b;
                        mnu[o] = minf(mnu[o], a); mxu[o] = maxf(mxu[o], a);
                        mnv[o] = minf(mnv[o], b); mxv[o] = maxf(mxv[o], b);
                    }
            float patch[64];
            for (int c = 0; c < 2; ++c) {
                float *pl = c ? fv : fu;
                for (int k = 0; k < gc; ++k) {
                    for (int o = 0; o < 64; ++o) patch[o] = pl[o * BM_G + k];
                    haar2d_fwd(patch);
                    for (int o = 0; o < 64; ++o) pl[o * BM_G + k] = patch[o];
                }
            }
            int retained = 0;
            for (int co = 0; co < 64; ++co) {
                float *vu = fu + co * BM_G, *vv = fv + co * BM_G;
                haar_fwd(vu, gc);
                haar_fwd(vv, gc);
                for (int k = 0; k < gc; ++k) {
                    const int dc = co == 0 && k == 0;
                    if (!dc && fabsf(vu[k]) < thr) vu[k] = 0.0f; else ++retained;
                    if (!dc && fabsf(vv[k]) < thr) vv[k] = 0.0f; else ++retained;
                }
                haar_inv(vu, gc);
                haar_inv(vv, gc);
            }
            for (int c = 0; c < 2; ++c) {
                float *pl = c ? fv : fu;
                for (int k = 0; k < gc; ++k) {
                    for (int o = 0; o < 64; ++o) patch[o] = pl[o * BM_G + k];
                    haar2d_inv(patch);
                    for (int o = 0; o < 64; ++o) pl[o * BM_G + k] = patch[o];
                }
            }
            for (int k = 0; k < gc; ++k)
                for (int o = 0; o < 64; ++o) {
                    fu[o * BM_G + k] = clampf(fu[o * BM_G + k], mnu[o], mxu[o]);
                    fv[o * BM_G + k] = clampf(fv[o * BM_G + k], mnv[o], mxv[o]);
                }
            const float gwt = 1.0f / (1.0f + retained);
            for (int k = 0; k < gc; ++k)
                for (int py = 0; py < BM_P; ++py)
                    for (int px = 0; px < BM_P; ++px) {
                        const int o = py * BM_P + px;
                        const float wgt = gwt * win[px] * win[py];
                        const size_t q = (size_t)(m[k].y + py) * w + m[k].x + px;
                        u[q] += wgt * fu[o * BM_G + k];
                        v[q] += wgt * fv[o * BM_G + k];
                        wt[q] += wgt;
                    }
        }
    for (size_t i = 0; i < n; ++i) {
        if (wt[i] > 0.0f) {
            const float fu = u[i] / wt[i], fv = v[i] / wt[i];
            u[i] = su[i] + 0.65f * (fu - su[i]);
            v[i] = sv[i] + 0.65f * (fv - sv[i]);
        } else {
            u[i] = su[i];
            v[i] = sv[i];
        }
    }
    free(sxs); free(sys); free(su); free(sv); free(wt); free(rg);
}

/* ------------------------------------------------------------------------ */
/* whole-frame orchestration: plugin.cpp:177-721                             */
/* ------------------------------------------------------------------------ */

static __thread or_debug g_dbg;

const or_debug *or_last_debug(void) { return &g_dbg; }

void or_clear_debug(void) {
    free(g_dbg.y); free(g_dbg.u); free(g_dbg.v);
    free(g_dbg.jxx); free(g_dbg.jxy); free(g_dbg.jyy); free(g_dbg.lc); free(g_dbg.ms);
    free(g_dbg.seed); free(g_dbg.mask);
    if (g_dbg.has_ws)
        or_free_workset(&g_dbg.ws);
    free(g_dbg.plain_u); free(g_dbg.plain_v);
    free(g_dbg.out_u); free(g_dbg.out_v);
    memset(&g_dbg, 0, sizeof(g_dbg));
}

static void to_float(const or_frame_format *f, const void *src, ptrdiff_t stride, float *dst,
                     int w, int h, double scale, double offset) {
    if (f->sample_type == 1) {
        for (int y = 0; y < h; ++y)
            memcpy(dst + (size_t)y * w, (const uint8_t *)src + stride * y, sizeof(float) * (size_t)w);
    } else if (f->bytes_per_sample == 1) {
        or_u8_to_float((const uint8_t *)src, stride, dst, w, h, scale, offset);
    } else {
        or_u16_to_float((const uint16_t *)src, stride, dst, w, h, scale, offset);
    }
}

static void from_float(const or_frame_format *f, const float *src, void *dst, ptrdiff_t stride,
                       int w, int h, double scale, double offset) {
    if (f->sample_type == 1) {
        for (int y = 0; y < h; ++y)
            memcpy((uint8_t *)dst + stride * y, src + (size_t)y * w, sizeof(float) * (size_t)w);
    } else if (f->bytes_per_sample == 1) {
        or_float_to_u8(src, (uint8_t *)dst, stride, w, h, scale, offset, 0, 255);
    } else {
        or_float_to_u16(src, (uint16_t *)dst, stride, w, h, scale, offset, 0, (1 << f->bits) - 1);
    }
}

/* ------------------------------------------------------------------------ */
/* Sharpen: sharpenPlane (algos.cpp:595-659) and sharpenGetFrame            */
/* (plugin.cpp:14-90)                                                       */
/* ------------------------------------------------------------------------ */

static void sort4(float *v) {
    for (int i = 1; i < 4; ++i) {
        const float t = v[i];
        int j = i;
        while (j > 0 && t < v[j - 1]) { v[j] = v[j - 1]; --j; }
        v[j] = t;
    }
}

void or_sharpen_plane(float *p, int w, int h, const float *lc, const float *gy, int gw, int gh,
                      double rw, double rh, double sx, double sy, const or_sharpen_params *d) {
    float *src = (float *)malloc(sizeof(float) * (size_t)w * h);
    memcpy(src, p, sizeof(float) * (size_t)w * h);
    const float inv_g2 = 1.0f / (float)(d->gspatial * d->gspatial);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const float c0 = at(src, w, h, x, y);
            const float l0 = lc ? bilinear(gy, gw, gh, (x + 0.5) * rw - 0.5 + sx, (y + 0.5) * rh - 0.5 + sy) : c0;
            float max_dl = 0.0f, lo = 1e30f, hi = -1e30f;
            for (int j = -2; j <= 2; ++j)
                for (int i = -2; i <= 2; ++i) {
                    const float g = lc ? at(lc, w, h, x + i, y + j) : at(src, w, h, x + i, y + j);
                    max_dl = maxf(max_dl, fabsf(g - l0));
                    lo = minf(lo, at(src, w, h, x + i, y + j));
                    hi = maxf(hi, at(src, w, h, x + i, y + j));
                }
            float nd[4] = {
                fabsf(at(src, w, h, x + 1, y) - c0), fabsf(at(src, w, h, x - 1, y) - c0),
                fabsf(at(src, w, h, x, y + 1) - c0), fabsf(at(src, w, h, x, y - 1) - c0),
            };
            sort4(nd);
            const float med = (nd[1] + nd[2]) * 0.5f;
            const float sig = maxf(maxf((float)d->sigma, (float)d->sratio * max_dl), 3.0f * med);
            const float inv_s2 = 1.0f / (sig * sig);
            float hp = 0.0f, ws = 0.0f;
            for (int j = -2; j <= 2; ++j)
                for (int i = -2; i <= 2; ++i) {
                    if (i == 0 && j == 0)
                        continue;
                    const float dl = (lc ? at(lc, w, h, x + i, y + j) : at(src, w, h, x + i, y + j)) - l0;
                    const float sim = 1.0f / (1.0f + dl * dl * inv_s2);
                    const float sp = 1.0f / (1.0f + (float)(i * i + j * j) * inv_g2);
                    const float n = sp * sim;
                    hp += n * (c0 - at(src, w, h, x + i, y + j));
                    ws += n;
                }
            float out = c0;
            if (ws > 1e-6f)
                out = c0 + (float)d->alpha * hp / ws;
            if (d->ar_margin >= 0.0)
                out = clampf(out, lo - (float)d->ar_margin, hi + (float)d->ar_margin);
            p[(size_t)y * w + x] = out;
        }
    free(src);
}

int or_sharpen_frame(const or_sharpen_params *d, const or_frame_format *f, const void *const in[3],
                     const ptrdiff_t in_stride[3], void *const out[3], const ptrdiff_t out_stride[3]) {
    const int sw = f->width, sh = f->height;
    const int cw = sw >> f->ss_w, ch = sh >> f->ss_h;
    const int is_float = f->sample_type == 1;
    const double ysc = is_float ? 1.0 : 1.0 / (double)((1 << f->bits) - 1);
    float *pl[3];
    const int pw[3] = {sw, cw, cw}, ph[3] = {sh, ch, ch};
    for (int i = 0; i < 3; ++i) {
        pl[i] = (float *)malloc(sizeof(float) * (size_t)pw[i] * ph[i]);
        to_float(f, in[i], in_stride[i], pl[i], pw[i], ph[i], ysc, (i == 0 || is_float) ? 0.0 : -0.5);
    }
    or_sharpen_plane(pl[0], sw, sh, NULL, pl[0], sw, sh, 1.0, 1.0, 0.0, 0.0, d);
    if (cw != sw || ch != sh) {
        const double rw = (double)sw / cw, rh = (double)sh / ch;
        float *lc = (float *)malloc(sizeof(float) * (size_t)cw * ch);
        or_lc_map(pl[0], sw, sh, cw, ch, rw, rh, -0.5, 0.0, lc);
        or_sharpen_plane(pl[1], cw, ch, lc, pl[0], sw, sh, rw, rh, -0.5, 0.0, d);
        or_sharpen_plane(pl[2], cw, ch, lc, pl[0], sw, sh, rw, rh, -0.5, 0.0, d);
        free(lc);
    } else {
        or_sharpen_plane(pl[1], cw, ch, NULL, pl[1], cw, ch, 1.0, 1.0, 0.0, 0.0, d);
        or_sharpen_plane(pl[2], cw, ch, NULL, pl[2], cw, ch, 1.0, 1.0, 0.0, 0.0, d);
    }
    const double osc = is_float ? 1.0 : (double)((1 << f->bits) - 1);
    for (int i = 0; i < 3; ++i) {
        from_float(f, pl[i], out[i], out_stride[i], pw[i], ph[i], osc, (i == 0 || is_float) ? 0.0 : 0.5 * osc);
        free(pl[i]);
    }
    return 0;
}

int or_recon_frame(const or_params *pin, const or_frame_format *f,
                   const void *const in[3], const ptrdiff_t in_stride[3],
                   void *const out[3], const ptrdiff_t out_stride[3]) {
    or_clear_debug();
    /* siting only applies on subsampled axes (plugin.cpp:159-162, 1130-1133) */
    or_params pl = *pin;
    if (f->ss_w == 0)
        pl.shift_x = 0.0;
    if (f->ss_h == 0)
        pl.shift_y = 0.0;
    const or_params *p = &pl;
    const int sw = f->width, sh = f->height;
    const int cw = sw >> f->ss_w, ch = sh >> f->ss_h;
    const int ow = p->out_w, oh = p->out_h;
    const int is_float = f->sample_type == 1;
    const double yscale = is_float ? 1.0 : 1.0 / (double)((1 << f->bits) - 1);
    const double coff = is_float ? 0.0 : -0.5;
    or_debug *D = &g_dbg;
    D->sw = sw; D->sh = sh; D->cw = cw; D->ch = ch; D->ow = ow; D->oh = oh;
    const size_t ns = (size_t)sw * sh, nc = (size_t)cw * ch, no = (size_t)ow * oh;
    D->y = (float *)malloc(sizeof(float) * ns);
    D->u = (float *)malloc(sizeof(float) * nc);
    D->v = (float *)malloc(sizeof(float) * nc);
    to_float(f, in[0], in_stride[0], D->y, sw, sh, yscale, 0.0);
    to_float(f, in[1], in_stride[1], D->u, cw, ch, yscale, coff);
    to_float(f, in[2], in_stride[2], D->v, cw, ch, yscale, coff);
    const double rw = (double)sw / cw, rh = (double)sh / ch;
    const int same = ow == sw && oh == sh;
    const double oscale = is_float ? 1.0 : (double)((1 << f->bits) - 1);

    /* luma (plugin.cpp:273-316) */
    if (same) {
        for (int y = 0; y < sh; ++y)
            memcpy((uint8_t *)out[0] + out_stride[0] * y, (const uint8_t *)in[0] + in_stride[0] * y,
                   (size_t)sw * (is_float ? 4 : f->bytes_per_sample));
    } else {
        float *oy = (float *)malloc(sizeof(float) * no);
        if (p->radial) {
            resample_radial(p, D->y, sw, sh, oy, ow, oh);
        } else {
            or_weights th, tv;
            or_build_weights(sw, ow, p->kernel, p->kp1, p->kp2, p->support, 0.0, &th);
            or_build_weights(sh, oh, p->kernel, p->kp1, p->kp2, p->support, 0.0, &tv);
            float *tmp = (float *)malloc(sizeof(float) * (size_t)ow * sh);
            resample_h(D->y, sw, sh, tmp, &th);
            resample_v(tmp, ow, sh, oy, &tv);
            free(tmp);
            or_free_weights(&th);
            or_free_weights(&tv);
        }
        from_float(f, oy, out[0], out_stride[0], ow, oh, oscale, 0.0);
        free(oy);
    }

    /* guide maps, trust mask, workset, mutual gate (plugin.cpp:318-369) */
    or_axis ax, ay;
    or_build_axis(sw, ow, rw, p->shift_x, p, &ax);
    or_build_axis(sh, oh, rh, p->shift_y, p, &ay);
    if (p->strength > 0.0) {
        const double thr = p->algo == 4 ? 0.25 * p->sigma : p->sigma;
        D->jxx = (float *)malloc(sizeof(float) * ns);
        D->jxy = (float *)malloc(sizeof(float) * ns);
        D->jyy = (float *)malloc(sizeof(float) * ns);
        D->lc = (float *)malloc(sizeof(float) * nc);
        if (p->sparse)
            D->seed = (uint8_t *)malloc(ns);
        or_sobel_tensor(D->y, sw, sh, D->jxx, D->jxy, D->jyy, D->seed, (float)(thr * thr));
        or_lc_map(D->y, sw, sh, cw, ch, rw, rh, p->shift_x, p->shift_y, D->lc);
        if (p->sparse) {
            const double rmax = rw > rh ? rw : rh;
            const int dil = (int)ceil(p->support * rmax) + 8;
            D->mask = (uint8_t *)malloc(ns);
            or_dilate(D->seed, sw, sh, dil, D->mask);
            or_build_workset(D->mask, sw, sh, ow, oh, &ax, &ay, cw, ch, &D->ws);
            D->has_ws = 1;
        }
        if (p->ms > 0.0) {
            D->ms = (float *)malloc(sizeof(float) * nc);
            or_mutual_gate(D->lc, D->u, D->v, cw, ch, D->has_ws ? D->ws.chroma_mask : NULL,
                           D->has_ws ? or_workset_chroma_dense(&D->ws) : 1, D->ms);
        }
    }

    D->out_u = (float *)malloc(sizeof(float) * no);
    D->out_v = (float *)malloc(sizeof(float) * no);
    if (p->algo == 4) {
        or_plain_chroma(p, D->u, D->v, cw, ch, sw, sh, D->out_u, D->out_v);
        if (p->strength > 0.0)
            recon_algo4(p, D, &ax, &ay, is_float, D->out_u, D->out_v);
    } else if (p->algo == 6) {
        or_plain_chroma(p, D->u, D->v, cw, ch, sw, sh, D->out_u, D->out_v);
        if (p->strength > 0.0)
            recon_algo6(p, D, &ax, &ay, D->out_u, D->out_v);
    } else if (p->strength == 0.0) {
        or_plain_chroma(p, D->u, D->v, cw, ch, sw, sh, D->out_u, D->out_v);
    } else {
        if (D->has_ws) {
            or_plain_chroma(p, D->u, D->v, cw, ch, sw, sh, D->out_u, D->out_v);
            D->plain_u = (float *)malloc(sizeof(float) * no);
            D->plain_v = (float *)malloc(sizeof(float) * no);
            memcpy(D->plain_u, D->out_u, sizeof(float) * no);
            memcpy(D->plain_v, D->out_v, sizeof(float) * no);
        }
        or_guide g = { D->y, sw, sh, D->jxx, D->jxy, D->jyy, D->lc, D->ms, NULL, 0 };
        or_guided(p, D->u, D->v, cw, ch, &g, D->has_ws ? &D->ws : NULL, D->out_u, D->out_v);
    }
    or_free_axis(&ax);
    or_free_axis(&ay);
    if (p->bm)
        or_collaborative(D->y, sw, sh, D->out_u, D->out_v, ow, oh, (float)p->sigma);

    /* back-projection, same-size 4:2:0 only (plugin.cpp:374-375, 683-689) */
    if (p->bp > 0.0 && same && cw * 2 == sw && ch * 2 == sh) {
        or_back_project(D->out_u, ow, oh, D->u, cw, ch, (float)p->bp);
        or_back_project(D->out_v, ow, oh, D->v, cw, ch, (float)p->bp);
    }

    /* output conversion (plugin.cpp:691-711) */
    from_float(f, D->out_u, out[1], out_stride[1], ow, oh, oscale, is_float ? 0.0 : 0.5 * oscale);
    from_float(f, D->out_v, out[2], out_stride[2], ow, oh, oscale, is_float ? 0.0 : 0.5 * oscale);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* TRecon: block ME, chroma consistency, temporal taps                      */
/* (algos.cpp:683-807, plugin.cpp:740-919)                                   */
/* ------------------------------------------------------------------------ */

void or_block_match(const float *cur, const float *nbr, int w, int h, int block, int search,
                    float tsad, int16_t *mvx, int16_t *mvy, float *conf, int *pbw, int *pbh) {
    const int bw = (w + block - 1) / block, bh = (h + block - 1) / block;
    *pbw = bw;
    *pbh = bh;
    const float inv_t2 = 1.0f / (tsad * tsad);
    for (int by = 0; by < bh; ++by)
        for (int bx = 0; bx < bw; ++bx) {
            const int x0 = bx * block, y0 = by * block;
            const int bw_ = mini(block, w - x0), bh_ = mini(block, h - y0);
            float best = 1e30f, sad0 = 1e30f;
            double sum = 0.0, sum2 = 0.0;
            int bdx = 0, bdy = 0;
            for (int j = 0; j < bh_; j += 2)
                for (int i = 0; i < bw_; i += 2) {
                    const float v = at(cur, w, h, x0 + i, y0 + j);
                    sum += v;
                    sum2 += (double)v * v;
                }
            const float npix = (float)((bw_ + 1) / 2 * ((bh_ + 1) / 2));
            for (int dy = -search; dy <= search; ++dy)
                for (int dx = -search; dx <= search; ++dx) {
                    if (x0 + dx < 0 || y0 + dy < 0 || x0 + dx + bw_ > w || y0 + dy + bh_ > h)
                        continue;
                    float sad = 0;
                    for (int j = 0; j < bh_; j += 2)
                        for (int i = 0; i < bw_; i += 2)
                            sad += fabsf(at(cur, w, h, x0 + i, y0 + j) - at(nbr, w, h, x0 + dx + i, y0 + dy + j));
                    if (dx == 0 && dy == 0)
                        sad0 = sad;
                    const int mn = abs(dx) + abs(dy), bn = abs(bdx) + abs(bdy);
                    if (sad < best || (sad == best && mn < bn)) {
                        best = sad;
                        bdx = dx;
                        bdy = dy;
                    }
                }
            if ((bdx != 0 || bdy != 0) && sad0 - best <= 0.10f * tsad * npix) {
                bdx = 0;
                bdy = 0;
                best = sad0;
            }
            const size_t idx = (size_t)by * bw + bx;
            mvx[idx] = (int16_t)bdx;
            mvy[idx] = (int16_t)bdy;
            const float sad_px = best / npix;
            const float var = (float)fmax(0.0, sum2 / npix - (sum / npix) * (sum / npix));
            const float match_q = 1.0f / (1.0f + sad_px * sad_px * inv_t2);
            const float uniq = var / (var + 0.25f * tsad * tsad);
            conf[idx] = match_q * uniq;
        }
}

static int cmp_float(const void *a, const void *b) {
    const float x = *(const float *)a, y = *(const float *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

void or_chroma_consistency(const float *cu, const float *cv, const float *nu, const float *nv,
                           int cw, int ch, int luma_w, int luma_h, int block, const int16_t *mvx,
                           const int16_t *mvy, int bw, int bh, float *conf) {
    const double rw = (double)luma_w / cw, rh = (double)luma_h / ch;
    float *delta = (float *)malloc(sizeof(float) * (size_t)(block + 2) * (block + 2));
    for (int by = 0; by < bh; ++by)
        for (int bx = 0; bx < bw; ++bx) {
            const size_t bi = (size_t)by * bw + bx;
            if (conf[bi] <= 0.0f)
                continue;
            const int lx0 = bx * block, ly0 = by * block;
            const int lx1 = mini(luma_w, lx0 + block), ly1 = mini(luma_h, ly0 + block);
            const int cx0 = clampi((int)floor(lx0 / rw), 0, cw - 1);
            const int cy0 = clampi((int)floor(ly0 / rh), 0, ch - 1);
            const int cx1 = clampi((int)ceil(lx1 / rw), cx0 + 1, cw);
            const int cy1 = clampi((int)ceil(ly1 / rh), cy0 + 1, ch);
            const double mx = mvx[bi] / rw, my = mvy[bi] / rh;
            size_t n = 0;
            for (int y = cy0; y < cy1; ++y)
                for (int x = cx0; x < cx1; ++x) {
                    const int nx = clampi((int)lround(x + mx), 0, cw - 1);
                    const int ny = clampi((int)lround(y + my), 0, ch - 1);
                    delta[n++] = maxf(fabsf(at(cu, cw, ch, x, y) - at(nu, cw, ch, nx, ny)),
                                      fabsf(at(cv, cw, ch, x, y) - at(nv, cw, ch, nx, ny)));
                }
            /* nth_element median; even counts average with the largest of
             * the lower half (algos.cpp:794-800) */
            qsort(delta, n, sizeof(float), cmp_float);
            const size_t mid = n / 2;
            float median = delta[mid];
            if ((n & 1) == 0)
                median = 0.5f * (delta[mid - 1] + median);
            const float ratio = median / 0.04f;
            const float r2 = ratio * ratio;
            conf[bi] *= 1.0f / (1.0f + r2 * r2);
        }
    free(delta);
}

int or_trecon_frame(const or_params *pin, const or_frame_format *f,
                    const void *const in[3], const ptrdiff_t in_stride[3],
                    int n_nbrs, const void *const *nbr_planes, const ptrdiff_t *nbr_strides,
                    void *const out[3], const ptrdiff_t out_stride[3]) {
    or_clear_debug();
    or_params pl = *pin;
    if (f->ss_w == 0) pl.shift_x = 0.0;
    if (f->ss_h == 0) pl.shift_y = 0.0;
    pl.out_w = f->width;
    pl.out_h = f->height;
    const or_params *p = &pl;
    const int sw = f->width, sh = f->height, cw = sw >> f->ss_w, ch = sh >> f->ss_h;
    const int is_float = f->sample_type == 1;
    const double ys = is_float ? 1.0 : 1.0 / (double)((1 << f->bits) - 1);
    const double co = is_float ? 0.0 : -0.5;
    const double rw = (double)sw / cw, rh = (double)sh / ch;
    const size_t ns = (size_t)sw * sh, nc = (size_t)cw * ch;
    or_debug *D = &g_dbg;
    D->sw = D->ow = sw; D->sh = D->oh = sh; D->cw = cw; D->ch = ch;
    D->y = (float *)malloc(sizeof(float) * ns);
    D->u = (float *)malloc(sizeof(float) * nc);
    D->v = (float *)malloc(sizeof(float) * nc);
    to_float(f, in[0], in_stride[0], D->y, sw, sh, ys, 0.0);
    to_float(f, in[1], in_stride[1], D->u, cw, ch, ys, co);
    to_float(f, in[2], in_stride[2], D->v, cw, ch, ys, co);
    const int strength = p->strength > 0.0;
    const int nn = strength ? n_nbrs : 0;
    or_nbr *nbrs = (or_nbr *)calloc((size_t)(nn > 0 ? nn : 1), sizeof(or_nbr));
    float **keep = (float **)calloc((size_t)(nn > 0 ? nn : 1) * 5, sizeof(float *));
    for (int k = 0; k < nn; ++k) {
        float *ny = (float *)malloc(sizeof(float) * ns), *nu = (float *)malloc(sizeof(float) * nc);
        float *nv = (float *)malloc(sizeof(float) * nc), *nl = (float *)malloc(sizeof(float) * nc);
        to_float(f, nbr_planes[3 * k], nbr_strides[3 * k], ny, sw, sh, ys, 0.0);
        to_float(f, nbr_planes[3 * k + 1], nbr_strides[3 * k + 1], nu, cw, ch, ys, co);
        to_float(f, nbr_planes[3 * k + 2], nbr_strides[3 * k + 2], nv, cw, ch, ys, co);
        const int bw = (sw + 15) / 16, bh = (sh + 15) / 16;
        int16_t *mv = (int16_t *)malloc(sizeof(int16_t) * (size_t)bw * bh * 2);
        float *cf = (float *)malloc(sizeof(float) * (size_t)bw * bh);
        int obw, obh;
        or_block_match(D->y, ny, sw, sh, 16, p->tsearch, (float)p->tsad, mv, mv + (size_t)bw * bh, cf, &obw, &obh);
        or_chroma_consistency(D->u, D->v, nu, nv, cw, ch, sw, sh, 16, mv, mv + (size_t)bw * bh, bw, bh, cf);
        or_lc_map(ny, sw, sh, cw, ch, rw, rh, p->shift_x, p->shift_y, nl);
        nbrs[k].u = nu; nbrs[k].v = nv; nbrs[k].lc = nl;
        nbrs[k].mvx = mv; nbrs[k].mvy = mv + (size_t)bw * bh; nbrs[k].conf = cf;
        nbrs[k].bw = bw; nbrs[k].bh = bh; nbrs[k].block = 16;
        keep[5 * k] = ny; keep[5 * k + 1] = nu; keep[5 * k + 2] = nv; keep[5 * k + 3] = nl;
        keep[5 * k + 4] = (float *)(void *)mv;
    }
    if (strength) {
        D->jxx = (float *)malloc(sizeof(float) * ns);
        D->jxy = (float *)malloc(sizeof(float) * ns);
        D->jyy = (float *)malloc(sizeof(float) * ns);
        D->lc = (float *)malloc(sizeof(float) * nc);
        or_sobel_tensor(D->y, sw, sh, D->jxx, D->jxy, D->jyy, NULL, 0.0f);
        or_lc_map(D->y, sw, sh, cw, ch, rw, rh, p->shift_x, p->shift_y, D->lc);
        if (p->ms > 0.0) {
            D->ms = (float *)malloc(sizeof(float) * nc);
            or_mutual_gate(D->lc, D->u, D->v, cw, ch, NULL, 1, D->ms);
        }
    }
    /* Y is rewritten through the output conversion (plugin.cpp:850-860) */
    const double os = is_float ? 1.0 : (double)((1 << f->bits) - 1);
    from_float(f, D->y, out[0], out_stride[0], sw, sh, os, 0.0);
    D->out_u = (float *)calloc(ns, sizeof(float));
    D->out_v = (float *)calloc(ns, sizeof(float));
    or_workset ws;
    int has_mask = 0;
    if (p->sparse && strength) {
        uint8_t *seed = (uint8_t *)malloc(ns);
        const float eth = (float)(p->sigma * p->sigma);
        for (size_t i = 0; i < ns; ++i)
            seed[i] = D->jxx[i] + D->jyy[i] > eth;
        D->mask = (uint8_t *)malloc(ns);
        const int dil = (int)ceil(p->support * (rw > rh ? rw : rh)) + 8;
        or_dilate(seed, sw, sh, dil, D->mask);
        free(seed);
        or_axis ax, ay;
        or_build_axis(sw, sw, rw, p->shift_x, p, &ax);
        or_build_axis(sh, sh, rh, p->shift_y, p, &ay);
        or_build_workset(D->mask, sw, sh, sw, sh, &ax, &ay, cw, ch, &ws);
        or_free_axis(&ax);
        or_free_axis(&ay);
        has_mask = 1;
        or_plain_chroma(p, D->u, D->v, cw, ch, sw, sh, D->out_u, D->out_v);
    }
    or_guide g = { D->y, sw, sh, D->jxx, D->jxy, D->jyy, D->lc, D->ms, nbrs, nn };
    guided_run(p, D->u, D->v, cw, ch, &g, has_mask ? &ws : NULL, NULL, D->out_u, D->out_v);
    if (has_mask)
        or_free_workset(&ws);
    if (p->bm)
        or_collaborative(D->y, sw, sh, D->out_u, D->out_v, sw, sh, (float)p->sigma);
    from_float(f, D->out_u, out[1], out_stride[1], sw, sh, os, is_float ? 0.0 : 0.5 * os);
    from_float(f, D->out_v, out[2], out_stride[2], sw, sh, os, is_float ? 0.0 : 0.5 * os);
    for (int k = 0; k < nn; ++k) {
        for (int j = 0; j < 4; ++j) free(keep[5 * k + j]);
        free(keep[5 * k + 4]);
        free((void *)nbrs[k].conf);
    }
    free(keep);
    free(nbrs);
    return 0;
}

/* ------------------------------------------------------------------------
 * Downsample 4:4:4 -> 4:2:0 (downsample.cpp), scalar build: every sum is a
 * float accumulation in the order written there.
 * ---------------------------------------------------------------------- */

typedef struct ds_plane {
    const uint8_t *data;
    ptrdiff_t stride;
    int w, h, type; /* type: 0 u8, 1 u16, 2 f32 */
    float scale, offset;
} ds_plane;

/* SourcePlane::at (downsample.cpp:101-106): clamped sample * scale + offset */
static float ds_at(const ds_plane *p, int x, int y) {
    x = clampi(x, 0, p->w - 1);
    y = clampi(y, 0, p->h - 1);
    const uint8_t *row = p->data + (ptrdiff_t)y * p->stride;
    float raw;
    if (p->type == 0)
        raw = (float)row[x];
    else if (p->type == 1)
        raw = (float)((const uint16_t *)row)[x];
    else
        raw = ((const float *)row)[x];
    return raw * p->scale + p->offset;
}

/* buildBinomialWeights (downsample.cpp:52-84) and buildDownsampleWeights
 * (downsample.cpp:86-92): spline36 / lanczos3 go through buildWeights at
 * half size (support 3, widened to 6 by the 2x decimation) */
static int ds_weights(int src_n, int kernel, float shift, or_weights *t) {
    if (kernel != 2)
        return or_build_weights(src_n, src_n / 2, kernel == 0 ? 4 : 2, 3.0, 0.0, 3.0, (double)shift, t);
    const int n = src_n / 2;
    const int integer_phase = fabsf(fabsf(shift) - 0.5f) < 1e-6f;
    t->n = n;
    t->sup = integer_phase ? 5 : 4;
    t->start = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    t->w = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1) * (size_t)t->sup);
    if (!t->start || !t->w)
        return -1;
    static const float iw[5] = {1.0f / 16.0f, 4.0f / 16.0f, 6.0f / 16.0f, 4.0f / 16.0f, 1.0f / 16.0f};
    static const float hw[4] = {1.0f / 8.0f, 3.0f / 8.0f, 3.0f / 8.0f, 1.0f / 8.0f};
    for (int i = 0; i < n; ++i) {
        const float center = 2.0f * (float)i + 0.5f + shift;
        if (integer_phase) {
            t->start[i] = (int)lroundf(center) - 2;
            memcpy(t->w + (size_t)i * 5, iw, sizeof(iw));
        } else {
            t->start[i] = (int)floorf(center) - 1;
            memcpy(t->w + (size_t)i * 4, hw, sizeof(hw));
        }
    }
    return 0;
}

/* BaselineRing (downsample.cpp:129-261), scalar branch: horizontal float
 * sums of clamped samples per source row, then vertical float sums over the
 * clamped rows of each output row */
static int ds_baseline(const ds_plane *u, const ds_plane *v, const or_weights *xw, const or_weights *yw,
                       float *bu, float *bv) {
    const int cw = xw->n, ch = yw->n, h = u->h;
    float *ru = (float *)malloc(sizeof(float) * (size_t)cw * (size_t)h);
    float *rv = (float *)malloc(sizeof(float) * (size_t)cw * (size_t)h);
    if (!ru || !rv) {
        free(ru);
        free(rv);
        return -1;
    }
    for (int sy = 0; sy < h; ++sy)
        for (int x = 0; x < cw; ++x) {
            const float *wt = xw->w + (size_t)x * xw->sup;
            float su = 0.0f, sv = 0.0f;
            for (int t = 0; t < xw->sup; ++t) {
                const int sx = xw->start[x] + t;
                su += wt[t] * ds_at(u, sx, sy);
                sv += wt[t] * ds_at(v, sx, sy);
            }
            ru[(size_t)sy * cw + x] = su;
            rv[(size_t)sy * cw + x] = sv;
        }
    for (int oy = 0; oy < ch; ++oy) {
        float *ou = bu + (size_t)oy * cw, *ov = bv + (size_t)oy * cw;
        for (int x = 0; x < cw; ++x) {
            ou[x] = 0.0f;
            ov[x] = 0.0f;
        }
        const float *wy = yw->w + (size_t)oy * yw->sup;
        for (int t = 0; t < yw->sup; ++t) {
            const int sy = clampi(yw->start[oy] + t, 0, h - 1);
            const float w = wy[t];
            for (int x = 0; x < cw; ++x) {
                ou[x] += w * ru[(size_t)sy * cw + x];
                ov[x] += w * rv[(size_t)sy * cw + x];
            }
        }
    }
    free(ru);
    free(rv);
    return 0;
}

typedef struct ds_dir { float x, y; } ds_dir;
static const ds_dir ds_dirs[8] = {
    {1.0f, 0.0f}, {0.9238795f, 0.3826834f}, {0.7071068f, 0.7071068f}, {0.3826834f, 0.9238795f},
    {0.0f, 1.0f}, {-0.3826834f, 0.9238795f}, {-0.7071068f, 0.7071068f}, {-0.9238795f, 0.3826834f},
};

typedef struct ds_stats { int primary, secondary; float mix, gate; int isotropic; } ds_stats;

static float ds_clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* classify (downsample.cpp:281-385) */
static ds_stats ds_classify(const ds_plane *yp, const ds_plane *up, const ds_plane *vp, float cx, float cy,
                            int quality) {
    ds_stats st = {0, 1, 0.0f, 0.0f, 1};
    const int ix = (int)floorf(cx + 0.5f), iy = (int)floorf(cy + 0.5f);
    float jxx = 0.0f, jxy = 0.0f, jyy = 0.0f, cgx = 0.0f, cgy = 0.0f;
    if (quality == 0) {
        const float tl = ds_at(yp, ix - 1, iy - 1), tc = ds_at(yp, ix, iy - 1), tr = ds_at(yp, ix + 1, iy - 1);
        const float ml = ds_at(yp, ix - 1, iy), mr = ds_at(yp, ix + 1, iy);
        const float bl = ds_at(yp, ix - 1, iy + 1), bc = ds_at(yp, ix, iy + 1), br = ds_at(yp, ix + 1, iy + 1);
        cgx = (tr + 2.0f * mr + br - tl - 2.0f * ml - bl) * 0.125f;
        cgy = (bl + 2.0f * bc + br - tl - 2.0f * tc - tr) * 0.125f;
        jxx = cgx * cgx;
        jxy = cgx * cgy;
        jyy = cgy * cgy;
    } else {
        for (int k = -1; k <= 1; ++k) {
            const float gx = 0.5f * (ds_at(yp, ix + 1, iy + k) - ds_at(yp, ix - 1, iy + k));
            const float gy = 0.5f * (ds_at(yp, ix, iy + k + 1) - ds_at(yp, ix, iy + k - 1));
            jxx += gx * gx;
            jxy += gx * gy;
            jyy += gy * gy;
            if (k == 0) {
                cgx = gx;
                cgy = gy;
            }
        }
    }
    const float trace = jxx + jyy;
    if (trace < 1.0e-6f)
        return st;
    const float diff = jxx - jyy;
    const float disc2 = diff * diff + 4.0f * jxy * jxy;
    const float coh2 = disc2 / (trace * trace + 1e-12f);
    float best = -1.0f, second = -1.0f;
    const int step = quality == 0 ? 2 : 1;
    for (int d = 0; d < 8; d += step) {
        const ds_dir n = ds_dirs[d];
        const float resp = n.x * n.x * jxx + 2.0f * n.x * n.y * jxy + n.y * n.y * jyy;
        if (resp > best) {
            second = best;
            st.secondary = st.primary;
            best = resp;
            st.primary = d;
        } else if (resp > second) {
            second = resp;
            st.secondary = d;
        }
    }
    st.isotropic = coh2 < (quality == 0 ? 0.20f : 0.10f);
    if (quality > 0 && best > 1e-12f) {
        const float adj = 0.8535534f;
        const float ratio = ds_clampf(second / best, 0.0f, 1.0f);
        st.mix = 0.5f * ds_clampf((ratio - adj) / (1.0f - adj), 0.0f, 1.0f);
    }
    const float ugx = 0.5f * (ds_at(up, ix + 1, iy) - ds_at(up, ix - 1, iy));
    const float ugy = 0.5f * (ds_at(up, ix, iy + 1) - ds_at(up, ix, iy - 1));
    const float vgx = 0.5f * (ds_at(vp, ix + 1, iy) - ds_at(vp, ix - 1, iy));
    const float vgy = 0.5f * (ds_at(vp, ix, iy + 1) - ds_at(vp, ix, iy - 1));
    const float cen = ugx * ugx + ugy * ugy + vgx * vgx + vgy * vgy;
    const ds_dir nrm = ds_dirs[st.primary];
    const float un = ugx * nrm.x + ugy * nrm.y, vn = vgx * nrm.x + vgy * nrm.y;
    const float aligned = un * un + vn * vn;
    const int nsx = (int)lroundf(nrm.x), nsy = (int)lroundf(nrm.y);
    const float cu = ds_at(up, ix, iy), cv = ds_at(vp, ix, iy);
    const float curu = ds_at(up, ix + nsx, iy + nsy) - 2.0f * cu + ds_at(up, ix - nsx, iy - nsy);
    const float curv = ds_at(vp, ix + nsx, iy + nsy) - 2.0f * cv + ds_at(vp, ix - nsx, iy - nsy);
    const float curE = curu * curu + curv * curv;
    const float sharp = curE / (curE + cen + 1.0e-12f);
    const float luma_g = trace / (trace + 3.0e-5f);
    const float conf_g = coh2 / (coh2 + 0.08f);
    const float chroma_g = cen / (cen + 2.0e-5f);
    const float dir_g = aligned / (cen + 1.0e-12f);
    const float centerE = cgx * cgx + cgy * cgy;
    const float pos_g = centerE / (centerE + 1.0e-5f);
    st.gate = ds_clampf(luma_g * conf_g * chroma_g * dir_g * pos_g * sharp, 0.0f, 1.0f);
    if (quality == 0)
        st.gate *= 0.985f;
    else if (quality == 2)
        st.gate = fminf(1.0f, st.gate * 1.01f);
    if (sharp < (quality == 0 ? 0.20f : 0.12f))
        st.gate = 0.0f;
    return st;
}

typedef struct ds_env { ds_dir normal; int radius; float aniso; int isotropic; } ds_env;
typedef struct ds_acc { float w, u, v; } ds_acc;

/* spatialWeight (downsample.cpp:402-411) */
static float ds_spatial(const ds_env *e, float dx, float dy) {
    const float r2 = dx * dx + dy * dy;
    if (e->isotropic)
        return 1.0f / (1.0f + 0.42f * r2);
    const float nd = dx * e->normal.x + dy * e->normal.y;
    const float td = -dx * e->normal.y + dy * e->normal.x;
    return 1.0f / (1.0f + (0.9f + e->aniso) * nd * nd + 0.16f * td * td);
}

/* accumulateGuided (downsample.cpp:413-458) */
static void ds_accumulate(const ds_plane *yp, const ds_plane *up, const ds_plane *vp, float cx, float cy,
                          const ds_env *env, int n, ds_acc *acc) {
    int maxr = 0;
    for (int i = 0; i < n; ++i) {
        acc[i].w = acc[i].u = acc[i].v = 0.0f;
        if (env[i].radius > maxr)
            maxr = env[i].radius;
    }
    const int x0 = (int)ceilf(cx - (float)maxr), x1 = (int)floorf(cx + (float)maxr);
    const int y0 = (int)ceilf(cy - (float)maxr), y1 = (int)floorf(cy + (float)maxr);
    const int cx0 = (int)floorf(cx), cy0 = (int)floorf(cy);
    const float fx = cx - (float)cx0, fy = cy - (float)cy0;
    const float ya = ds_at(yp, cx0, cy0), yb = ds_at(yp, cx0 + 1, cy0);
    const float yc = ds_at(yp, cx0, cy0 + 1), yd = ds_at(yp, cx0 + 1, cy0 + 1);
    const float cl = ya + (yb - ya) * fx + (yc - ya) * fy + (ya - yb - yc + yd) * fx * fy;
    const float rs2 = 2.25e-4f;
    for (int sy = y0; sy <= y1; ++sy)
        for (int sx = x0; sx <= x1; ++sx) {
            const float dx = (float)sx - cx, dy = (float)sy - cy;
            const float delta = ds_at(yp, sx, sy) - cl;
            const float rw = rs2 / (rs2 + delta * delta);
            const float su = ds_at(up, sx, sy), sv = ds_at(vp, sx, sy);
            for (int i = 0; i < n; ++i) {
                if (fabsf(dx) > (float)env[i].radius || fabsf(dy) > (float)env[i].radius)
                    continue;
                const float w = rw * ds_spatial(&env[i], dx, dy);
                acc[i].w += w;
                acc[i].u += w * su;
                acc[i].v += w * sv;
            }
        }
}

static float ds_norm(float sum, float w, float fallback) { return w > 1e-12f ? sum / w : fallback; }

/* writeRow (downsample.cpp:117-126): integers round with the chroma offset */
static void ds_store(const or_frame_format *f, void *out, ptrdiff_t stride, int x, int y, float v) {
    uint8_t *row = (uint8_t *)out + (ptrdiff_t)y * stride;
    if (f->sample_type == 1) {
        ((float *)row)[x] = v;
        return;
    }
    const int maxv = (1 << f->bits) - 1;
    const float scale = (float)maxv, offset = 0.5f * scale;
    const int q = clampi((int)(v * scale + offset + 0.5f), 0, maxv);
    if (f->bytes_per_sample == 1)
        row[x] = (uint8_t)q;
    else
        ((uint16_t *)row)[x] = (uint16_t)q;
}

static void ds_planes(const or_frame_format *f, const void *const in[3], const ptrdiff_t in_stride[3],
                      ds_plane p[3]) {
    const int is_float = f->sample_type == 1;
    const int maxv = is_float ? 0 : (1 << f->bits) - 1;
    const float scale = is_float ? 1.0f : 1.0f / (float)maxv;
    const int type = is_float ? 2 : (f->bytes_per_sample == 1 ? 0 : 1);
    for (int i = 0; i < 3; ++i) {
        p[i].data = (const uint8_t *)in[i];
        p[i].stride = in_stride[i];
        p[i].w = f->width;
        p[i].h = f->height;
        p[i].type = type;
        p[i].scale = scale;
        p[i].offset = (i > 0 && !is_float) ? -0.5f : 0.0f;
    }
}

/* generateCandidateRow (downsample.cpp:526-568) for every output row:
 * gate and four candidates (planes c * nc + i) per chroma sample */
static void ds_candidates(const or_downsample_params *d, const ds_plane *p, int cw, int ch, const float *bu,
                          const float *bv, float *gate, float *cu, float *cv) {
    const size_t nc = (size_t)cw * (size_t)ch;
    for (int y = 0; y < ch; ++y) {
        const float cy = 2.0f * (float)y + 0.5f + d->shift_y;
        for (int x = 0; x < cw; ++x) {
            const size_t i = (size_t)y * cw + x;
            const float cx = 2.0f * (float)x + 0.5f + d->shift_x;
            const ds_stats st = ds_classify(&p[0], &p[1], &p[2], cx, cy, 2);
            gate[i] = st.gate > 1.0e-4f ? st.gate : 0.0f;
            if (gate[i] == 0.0f) {
                for (int c = 0; c < 4; ++c) {
                    cu[(size_t)c * nc + i] = bu[i];
                    cv[(size_t)c * nc + i] = bv[i];
                }
                continue;
            }
            const ds_env env[5] = {{ds_dirs[st.primary], 3, 2.2f, st.isotropic},
                                   {ds_dirs[st.secondary], 3, 2.2f, st.isotropic},
                                   {ds_dirs[st.primary], 4, 2.7f, st.isotropic},
                                   {ds_dirs[st.secondary], 4, 2.7f, st.isotropic},
                                   {ds_dirs[st.primary], 4, 0.0f, 1}};
            ds_acc a[5];
            ds_accumulate(&p[0], &p[1], &p[2], cx, cy, env, 5, a);
            const float u0 = ds_norm(a[0].u, a[0].w, bu[i]), v0 = ds_norm(a[0].v, a[0].w, bv[i]);
            const float u1 = ds_norm(a[1].u, a[1].w, bu[i]), v1 = ds_norm(a[1].v, a[1].w, bv[i]);
            cu[i] = u0 + st.mix * (u1 - u0);
            cv[i] = v0 + st.mix * (v1 - v0);
            for (int c = 1; c < 4; ++c) {
                cu[(size_t)c * nc + i] = ds_norm(a[c + 1].u, a[c + 1].w, bu[i]);
                cv[(size_t)c * nc + i] = ds_norm(a[c + 1].v, a[c + 1].w, bv[i]);
            }
        }
    }
}

/* test access: the quality-2 gates and candidates of the next frame */
static struct { float *gate, *cu, *cv; } g_ds_debug;
void or_downsample_debug(float *gate, float *cu, float *cv) {
    g_ds_debug.gate = gate;
    g_ds_debug.cu = cu;
    g_ds_debug.cv = cv;
}

int or_downsample_base(const or_downsample_params *d, const or_frame_format *f, const void *const in[3],
                       const ptrdiff_t in_stride[3], float *base_u, float *base_v) {
    ds_plane p[3];
    ds_planes(f, in, in_stride, p);
    or_weights xw = {0}, yw = {0};
    int rc = ds_weights(f->width, d->kernel, d->shift_x, &xw);
    if (!rc)
        rc = ds_weights(f->height, d->kernel, d->shift_y, &yw);
    if (!rc)
        rc = ds_baseline(&p[1], &p[2], &xw, &yw, base_u, base_v);
    or_free_weights(&xw);
    or_free_weights(&yw);
    return rc;
}

/* processFrame (downsample.cpp:628-722) with makeGuidedPixel
 * (downsample.cpp:466-507), generateCandidateRow (downsample.cpp:526-568)
 * and scoreCandidateRow (downsample.cpp:576-626); the three-row candidate
 * ring is whole planes here (rows outside the frame clamp, as the ring's
 * lookup does) */
int or_downsample_frame(const or_downsample_params *d, const or_frame_format *f, const void *const in[3],
                        const ptrdiff_t in_stride[3], void *const out[3], const ptrdiff_t out_stride[3]) {
    const int w = f->width, h = f->height, cw = w / 2, ch = h / 2;
    const size_t nc = (size_t)cw * (size_t)ch;
    ds_plane p[3];
    ds_planes(f, in, in_stride, p);
    float *bu = (float *)malloc(sizeof(float) * (nc ? nc : 1));
    float *bv = (float *)malloc(sizeof(float) * (nc ? nc : 1));
    if (!bu || !bv || or_downsample_base(d, f, in, in_stride, bu, bv)) {
        free(bu);
        free(bv);
        return -1;
    }
    const float s = d->strength;
    if (s == 0.0f || d->quality < 2) {
        for (int y = 0; y < ch; ++y) {
            const float cy = 2.0f * (float)y + 0.5f + d->shift_y;
            for (int x = 0; x < cw; ++x) {
                const size_t i = (size_t)y * cw + x;
                float ru = bu[i], rv = bv[i];
                if (s != 0.0f) {
                    const float cx = 2.0f * (float)x + 0.5f + d->shift_x;
                    const ds_stats st = ds_classify(&p[0], &p[1], &p[2], cx, cy, d->quality);
                    float gu = bu[i], gv = bv[i], gate = st.gate;
                    if (gate <= 1.0e-4f) {
                        gate = 0.0f;
                    } else if (d->quality == 0) {
                        const ds_env env = {ds_dirs[st.primary], 2, 1.8f, st.isotropic};
                        ds_acc a;
                        ds_accumulate(&p[0], &p[1], &p[2], cx, cy, &env, 1, &a);
                        gu = ds_norm(a.u, a.w, bu[i]);
                        gv = ds_norm(a.v, a.w, bv[i]);
                    } else {
                        const ds_env env[2] = {{ds_dirs[st.primary], 3, 2.2f, st.isotropic},
                                               {ds_dirs[st.secondary], 3, 2.2f, st.isotropic}};
                        ds_acc a[2];
                        ds_accumulate(&p[0], &p[1], &p[2], cx, cy, env, 2, a);
                        const float u0 = ds_norm(a[0].u, a[0].w, bu[i]), v0 = ds_norm(a[0].v, a[0].w, bv[i]);
                        const float u1 = ds_norm(a[1].u, a[1].w, bu[i]), v1 = ds_norm(a[1].v, a[1].w, bv[i]);
                        gu = u0 + st.mix * (u1 - u0);
                        gv = v0 + st.mix * (v1 - v0);
                    }
                    ru = bu[i] + s * gate * (gu - bu[i]);
                    rv = bv[i] + s * gate * (gv - bv[i]);
                }
                ds_store(f, out[1], out_stride[1], x, y, ru);
                ds_store(f, out[2], out_stride[2], x, y, rv);
            }
        }
        free(bu);
        free(bv);
        return 0;
    }
    /* quality 2: gate and four candidates per chroma sample, then the
     * loopback score over the sample's 2x2 source block */
    float *gate = (float *)malloc(sizeof(float) * nc);
    float *cu = (float *)malloc(sizeof(float) * nc * 4), *cv = (float *)malloc(sizeof(float) * nc * 4);
    if (!gate || !cu || !cv) {
        free(bu); free(bv); free(gate); free(cu); free(cv);
        return -1;
    }
    ds_candidates(d, p, cw, ch, bu, bv, gate, cu, cv);
    if (g_ds_debug.gate) {
        memcpy(g_ds_debug.gate, gate, sizeof(float) * nc);
        memcpy(g_ds_debug.cu, cu, sizeof(float) * nc * 4);
        memcpy(g_ds_debug.cv, cv, sizeof(float) * nc * 4);
    }
    for (int y = 0; y < ch; ++y)
        for (int x = 0; x < cw; ++x) {
            const size_t i = (size_t)y * cw + x;
            float ru = bu[i], rv = bv[i];
            if (gate[i] != 0.0f) {
                float score[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                for (int hy = 2 * y; hy <= 2 * y + 1; ++hy)
                    for (int hx = 2 * x; hx <= 2 * x + 1; ++hx) {
                        const float gx = 0.5f * (ds_at(&p[0], hx + 1, hy) - ds_at(&p[0], hx - 1, hy));
                        const float gy = 0.5f * (ds_at(&p[0], hx, hy + 1) - ds_at(&p[0], hx, hy - 1));
                        const float e2 = gx * gx + gy * gy;
                        const float ew = 0.125f + e2 / (e2 + 1.0e-5f);
                        const float lx = ((float)hx - 0.5f - d->shift_x) * 0.5f;
                        const float ly = ((float)hy - 0.5f - d->shift_y) * 0.5f;
                        const float su = ds_at(&p[1], hx, hy), sv = ds_at(&p[2], hx, hy);
                        const int rx0 = (int)floorf(lx), ry0 = (int)floorf(ly);
                        const int x0 = clampi(rx0, 0, cw - 1), x1 = clampi(rx0 + 1, 0, cw - 1);
                        const int r0 = clampi(ry0, 0, ch - 1), r1 = clampi(ry0 + 1, 0, ch - 1);
                        const float fx = lx - (float)rx0, fy = ly - (float)ry0;
                        const float wts[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy, fx * fy};
                        const size_t idx[4] = {(size_t)r0 * cw + x0, (size_t)r0 * cw + x1, (size_t)r1 * cw + x0,
                                               (size_t)r1 * cw + x1};
                        for (int c = 0; c < 4; ++c) {
                            float recu = 0.0f, recv = 0.0f;
                            for (int k = 0; k < 4; ++k) {
                                const size_t j = idx[k];
                                float vu = bu[j], vv = bv[j];
                                vu += gate[j] * (cu[(size_t)c * nc + j] - vu);
                                vv += gate[j] * (cv[(size_t)c * nc + j] - vv);
                                recu += wts[k] * vu;
                                recv += wts[k] * vv;
                            }
                            const float du = recu - su, dv = recv - sv;
                            score[c] += ew * (du * du + dv * dv);
                        }
                    }
                int best = 0;
                for (int c = 1; c < 4; ++c)
                    if (score[c] < score[best])
                        best = c;
                ru = bu[i] + s * gate[i] * (cu[(size_t)best * nc + i] - bu[i]);
                rv = bv[i] + s * gate[i] * (cv[(size_t)best * nc + i] - bv[i]);
            }
            ds_store(f, out[1], out_stride[1], x, y, ru);
            ds_store(f, out[2], out_stride[2], x, y, rv);
        }
    free(bu); free(bv); free(gate); free(cu); free(cv);
    return 0;
}
