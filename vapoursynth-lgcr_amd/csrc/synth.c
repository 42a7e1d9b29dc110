/*
 * Deterministic synthetic YUV clips for parity tests and benchmarks.
 *
 * Scene: a flat background plus `polygons` star-shaped polygons painted in
 * order (later ones on top).  Each region has a flat luma shade (uniform
 * 16..235, 8-bit scale) and flat U/V values (uniform 16..240).  Pixels whose
 * 3x3 neighbourhood touches another region are drawn as a dark outline
 * (luma 16), which gives a 2-pixel line across every boundary.  Gaussian
 * noise (Box-Muller) is added to luma and optionally chroma.  Chroma is
 * subsampled by a box average over columns {2i, 2i+1} and/or rows
 * {2j, 2j+1}; the frames are meant to be tagged left-sited (MPEG-2).
 *
 * Integer formats scale the 8-bit values by 2^(bits-8) before noise and
 * rounding (10-bit: 64..940).  Float formats map luma to Y/255 and chroma to
 * C/255 - 0.5.  With `motion`, every polygon translates by a fixed random
 * velocity of 0..4 px/frame.  All randomness comes from splitmix64.
 */
#include <math.h>
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
static double gauss(uint64_t *s) {
    double u1 = unif(s);
    if (u1 < 1e-300) u1 = 1e-300;
    const double u2 = unif(s);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

#define MAXV 8
typedef struct { double cx, cy, vx, vy; int n; double ang[MAXV], rad[MAXV]; int y8, u8, v8; } poly_t;

static int cmpd(const void *a, const void *b) {
    const double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

int lgcr_synth_frame(const lgcr_synth_spec *sp, int frame, void *const planes[3],
                     const ptrdiff_t strides[3]) {
    const int W = sp->width, H = sp->height;
    if (W <= 0 || H <= 0 || sp->polygons < 0 || sp->polygons > 100000)
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
    /* full-resolution flat values (8-bit scale), outlines on luma */
    double *fy = (double *)malloc(sizeof(double) * (size_t)W * H);
    double *fu = (double *)malloc(sizeof(double) * (size_t)W * H);
    double *fv = (double *)malloc(sizeof(double) * (size_t)W * H);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            const int l = label[i];
            int edge = 0;
            for (int dy = -1; dy <= 1 && !edge; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int xx = x + dx, yy = y + dy;
                    if (xx < 0 || yy < 0 || xx >= W || yy >= H) continue;
                    if (label[(size_t)yy * W + xx] != l) { edge = 1; break; }
                }
            fy[i] = edge ? 16.0 : (l < 0 ? bg_y : P[l].y8);
            fu[i] = l < 0 ? bg_u : P[l].u8;
            fv[i] = l < 0 ? bg_v : P[l].v8;
        }
    uint64_t ns = sp->seed * 0xD1B54A32D192ED03ULL + (uint64_t)frame * 0x8CB92BA72F3D8DD7ULL + 99;
    const int maxv = sp->is_float ? 0 : (1 << sp->bits) - 1;
    const double scale = sp->is_float ? 1.0 / 255.0 : (double)(1 << (sp->bits - 8));
    /* luma */
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            double val = fy[(size_t)y * W + x] * scale;
            if (sp->luma_noise > 0.0) val += sp->luma_noise * gauss(&ns);
            if (sp->is_float) {
                ((float *)((uint8_t *)planes[0] + strides[0] * y))[x] = (float)val;
            } else {
                long q = lround(val);
                if (q < 0) q = 0;
                if (q > maxv) q = maxv;
                if (sp->bits <= 8) ((uint8_t *)planes[0] + strides[0] * y)[x] = (uint8_t)q;
                else ((uint16_t *)((uint8_t *)planes[0] + strides[0] * y))[x] = (uint16_t)q;
            }
        }
    }
    /* chroma: box subsample */
    const int cw = W >> sp->ss_w, ch = H >> sp->ss_h;
    const int bx = 1 << sp->ss_w, by = 1 << sp->ss_h;
    for (int c = 0; c < 2; ++c) {
        const double *src = c == 0 ? fu : fv;
        for (int y = 0; y < ch; ++y)
            for (int x = 0; x < cw; ++x) {
                double s = 0.0;
                for (int j = 0; j < by; ++j)
                    for (int i = 0; i < bx; ++i)
                        s += src[(size_t)(y * by + j) * W + (x * bx + i)];
                s /= (double)(bx * by);
                double val = sp->is_float ? s / 255.0 - 0.5 : s * scale;
                if (sp->chroma_noise > 0.0) val += sp->chroma_noise * gauss(&ns);
                uint8_t *row = (uint8_t *)planes[1 + c] + strides[1 + c] * y;
                if (sp->is_float) {
                    ((float *)row)[x] = (float)val;
                } else {
                    long q = lround(val);
                    if (q < 0) q = 0;
                    if (q > maxv) q = maxv;
                    if (sp->bits <= 8) row[x] = (uint8_t)q;
                    else ((uint16_t *)row)[x] = (uint16_t)q;
                }
            }
    }
    free(label); free(fy); free(fu); free(fv); free(P);
    return 0;
}
