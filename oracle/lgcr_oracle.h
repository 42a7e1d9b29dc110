/*
 * LGCR oracle — TEST INFRASTRUCTURE ONLY.
 *
 * A scalar CPU restatement, in plain C, of the reference's chroma
 * reconstruction path as its scalar build computes it (ScalarBackend, every
 * AVX2 block compiled out, no FMA contraction).  Only tests/, the smoke()
 * entry and bench.py's CPU-baseline leg may load it, and only as the checker
 * or the timed CPU baseline.  Nothing in the product library links it.
 *
 * Planes are contiguous row-major float arrays (stride == width) unless a
 * stride argument says otherwise.
 *
 * Pinning: closed forms, tables, the Sobel tensor, footprint luma, dilation,
 * workset, plain base, back-projection, LGF regression, the collaborative
 * filter and Sharpen are checked against the reference's known-answer tests,
 * brute force, closed forms or independent models in tests/.  Parity
 * unpinned (only invariants such as sparse == dense, continuity, constants
 * and identities constrain them; no golden vector exists and the reference
 * cannot be built without VapourSynth headers): the heuristic internals of
 * or_guided (metadata probes, knee, ridge and cedge fades, rescue, fallback;
 * recon.cpp:155-760), or_mutual_gate (maps.cpp:853-908), the LGF cedge fade
 * and selector weights (algos.cpp:366-593), the algo-6 affine q/stability
 * maps, Nyquist detail and detail transfer (algos.cpp:1382-1689), the block
 * match confidence and chroma consistency (algos.cpp:683-807) and the
 * TRecon temporal taps (recon.cpp:642-745).
 */
#ifndef LGCR_ORACLE_H
#define LGCR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_BILINEAR = 0, OR_BICUBIC, OR_LANCZOS, OR_SPLINE16, OR_SPLINE36, OR_JINC };

typedef struct or_params {
    int kernel;
    double kp1, kp2, support;
    int radial;
    const float *lut;      /* jinc profile LUT (radial only) */
    int lut_n;
    float lut_scale;
    double strength, sigma, sratio, sdb, reg, stretch, gsigma, rescue;
    int algo, ridge, cedge;
    double ms, qgate;
    int sparse;
    double bp;
    int bm;
    double ar_margin;
    double shift_x, shift_y;
    int out_w, out_h;
    /* TRecon */
    int trad, tsearch;
    double tsad;
} or_params;

typedef struct or_weights {
    int n, sup;
    int *start;
    float *w;
} or_weights;

typedef struct or_axis {
    int n, sup;
    int *start;
    float *w, *am;
    float *wsum, *wabs;
    int *active;
    float *pos, *lpos;
    int *cb_i0; float *cb_f;   /* pos on the source chroma grid */
    int *lb_i0; float *lb_f;   /* lpos on the source luma grid */
} or_axis;

typedef struct or_workset {
    int mask_w, mask_h, out_w, out_h, chroma_w, chroma_h;
    uint8_t *mask;               /* mask_w * mask_h (owned copy) */
    uint32_t *indices;           /* active output pixel indices, row-major */
    size_t n_indices;
    size_t *index_row_offsets;   /* out_h + 1 */
    int *spans;                  /* (begin, end) pairs */
    size_t n_spans;
    size_t *span_row_offsets;    /* out_h + 1 */
    uint8_t *chroma_mask;        /* chroma_w * chroma_h */
    size_t active_chroma;
} or_workset;

/* closed forms (lgcr.h:53-117) */
double or_kernel_eval(int kernel, double x, double p1, double p2);
double or_jinc_windowed(double x, double taps);
double or_kernel_support(int kernel, double p1);
void or_jinc_lut(double support, float *lut, int n);
float or_lut_lookup(const float *lut, int n, float idx);

/* tables (kernels.cpp:135-167, 337-404; maps.cpp:190-249) */
int or_build_weights(int src_n, int dst_n, int kernel, double kp1, double kp2,
                     double support, double shift, or_weights *out);
void or_free_weights(or_weights *w);
int or_build_axis(int src_luma_n, int dst_luma_n, double r, double sit_shift,
                  const or_params *p, or_axis *out);
void or_free_axis(or_axis *a);
void or_bilin_axis(const float *pos, int n, int src_n, int *i0, float *f);

/* conversions (lgcr.h:986-1013) */
void or_u8_to_float(const uint8_t *src, ptrdiff_t stride, float *dst, int w, int h,
                    double scale, double offset);
void or_u16_to_float(const uint16_t *src, ptrdiff_t stride, float *dst, int w, int h,
                     double scale, double offset);
void or_float_to_u8(const float *src, uint8_t *dst, ptrdiff_t stride, int w, int h,
                    double scale, double offset, int lo, int hi);
void or_float_to_u16(const float *src, uint16_t *dst, ptrdiff_t stride, int w, int h,
                     double scale, double offset, int lo, int hi);

/* guide maps (maps.cpp:9-168, 252-340, 425-499) */
void or_sobel_tensor(const float *y, int w, int h, float *jxx, float *jxy, float *jyy,
                     uint8_t *seed, float threshold);
void or_lc_map(const float *y, int w, int h, int cw, int ch, double rw, double rh,
               double shift_x, double shift_y, float *lc);
void or_dilate(const uint8_t *src, int w, int h, int radius, uint8_t *dst);
int or_build_workset(const uint8_t *mask, int mask_w, int mask_h, int out_w, int out_h,
                     const or_axis *ax, const or_axis *ay, int cw, int ch,
                     or_workset *out);
void or_free_workset(or_workset *ws);
int or_workset_output_dense(const or_workset *ws);
int or_workset_chroma_dense(const or_workset *ws);

/* mutual-structure gate (maps.cpp:831-948); chroma_mask may be NULL (dense) */
void or_mutual_gate(const float *lc, const float *u, const float *v, int w, int h,
                    const uint8_t *chroma_mask, int chroma_dense, float *gate);

/* plain chroma (recon.cpp:1151-1461, 828-1029, 1031-1149) */
void or_plain_chroma(const or_params *p, const float *u, const float *v, int cw, int ch,
                     int sw, int sh, float *out_u, float *out_v);

/* guided reconstruction, algo 2 (recon.cpp:14-819) without temporal taps.
 * workset may be NULL (dense visit of every output pixel). */
/* one motion-compensated neighbour frame of TRecon (TemporalNbr,
 * lgcr.h:729-736) */
typedef struct or_nbr {
    const float *u, *v, *lc;             /* neighbour chroma + footprint luma (cw x ch) */
    const int16_t *mvx, *mvy;            /* block motion (luma px) */
    const float *conf;                   /* block confidence */
    int bw, bh, block;
} or_nbr;

typedef struct or_guide {
    const float *y; int sw, sh;          /* source luma */
    const float *jxx, *jxy, *jyy;        /* sw x sh */
    const float *lc;                     /* cw x ch */
    const float *ms;                     /* cw x ch or NULL */
    const or_nbr *nbrs; int n_nbrs;      /* temporal taps (TRecon) */
} or_guide;
void or_guided(const or_params *p, const float *u, const float *v, int cw, int ch,
               const or_guide *g, const or_workset *ws, float *out_u, float *out_v);

/* LGF regression maps, radius 2 (buildLGFPair, algos.cpp:366-448); out6 =
 * aU, bU, confU, aV, bV, confV planes of cw*ch.  roi may be NULL. */
void or_lgf_maps(const float *y, int w, int h, int cw, int ch, double rw, double rh, double sx,
                 double sy, const float *u, const float *v, double eps, int cedge,
                 const uint8_t *roi, float *out6);

/* collaborative chroma filter (bm.cpp:250-442), in place on u/v (w x h)
 * guided by the source luma (gw x gh) */
void or_collaborative(const float *guide, int gw, int gh, float *u, float *v, int w, int h,
                      float guide_sigma);

/* back-projection (algos.cpp:663-677) */
void or_back_project(float *out, int ow, int oh, const float *src, int cw, int ch, float bp);

/* whole-frame Recon (plugin.cpp:177-721).  sample_type 0 = integer, 1 = float.
 * in[] / out[] are host planes with byte strides.  Intermediates of the last
 * call are kept for inspection (or_last_*). */
typedef struct or_frame_format {
    int sample_type, bits, bytes_per_sample, ss_w, ss_h, width, height;
} or_frame_format;
int or_recon_frame(const or_params *p, const or_frame_format *fmt,
                   const void *const in[3], const ptrdiff_t in_stride[3],
                   void *const out[3], const ptrdiff_t out_stride[3]);

/* TRecon frame (plugin.cpp:740-919): center planes plus the neighbour
 * frames already chosen by temporalFrameNumbers (plugin.cpp:729-738);
 * output is 4:4:4 at the input size with Y rewritten. */
int or_trecon_frame(const or_params *p, const or_frame_format *fmt,
                    const void *const in[3], const ptrdiff_t in_stride[3],
                    int n_nbrs, const void *const *nbr_planes, const ptrdiff_t *nbr_strides,
                    void *const out[3], const ptrdiff_t out_stride[3]);

/* Sharpen (SharpenData, lgcr.h; sharpenPlane algos.cpp:595-659;
 * sharpenGetFrame plugin.cpp:14-90).  lc == NULL: self-guided plane. */
typedef struct or_sharpen_params { double alpha, sigma, sratio, gspatial, ar_margin; } or_sharpen_params;
void or_sharpen_plane(float *p, int w, int h, const float *lc, const float *gy, int gw, int gh,
                      double rw, double rh, double sx, double sy, const or_sharpen_params *d);
int or_sharpen_frame(const or_sharpen_params *d, const or_frame_format *f, const void *const in[3],
                     const ptrdiff_t in_stride[3], void *const out[3], const ptrdiff_t out_stride[3]);

/* block matching and chroma consistency (algos.cpp:683-807) */
void or_block_match(const float *cur, const float *nbr, int w, int h, int block, int search,
                    float tsad, int16_t *mvx, int16_t *mvy, float *conf, int *bw, int *bh);
void or_chroma_consistency(const float *cu, const float *cv, const float *nu, const float *nv,
                           int cw, int ch, int luma_w, int luma_h, int block, const int16_t *mvx,
                           const int16_t *mvy, int bw, int bh, float *conf);

/* Downsample 4:4:4 -> 4:2:0 (downsample.cpp).  kernel: 0 spline36,
 * 1 lanczos3, 2 binomial; shifts are the output siting in source units.
 * out[1], out[2] receive U and V at (width/2, height/2); out[0] is unused
 * (the plugin shares the source luma). */
typedef struct or_downsample_params { int quality, kernel; float strength, shift_x, shift_y; } or_downsample_params;
int or_downsample_frame(const or_downsample_params *d, const or_frame_format *f, const void *const in[3],
                        const ptrdiff_t in_stride[3], void *const out[3], const ptrdiff_t out_stride[3]);
/* test access: the next quality-2 frame also copies its gates (nc) and
 * candidates (4 * nc, candidate-major) here; NULL turns it off */
void or_downsample_debug(float *gate, float *cand_u, float *cand_v);
/* test access: copy the affine credibility g (chroma grid) of the next
 * algo-6 Recon frame into g; NULL stops */
void or_affine_debug(float *g);
/* the plain separable baseline alone (float U, V at the output size) */
int or_downsample_base(const or_downsample_params *d, const or_frame_format *f, const void *const in[3],
                       const ptrdiff_t in_stride[3], float *base_u, float *base_v);

/* intermediates of the last or_recon_frame on this thread */
typedef struct or_debug {
    int sw, sh, cw, ch, ow, oh;
    float *y, *u, *v;            /* converted source planes */
    float *jxx, *jxy, *jyy, *lc, *ms;
    uint8_t *seed, *mask;
    or_workset ws; int has_ws;
    float *plain_u, *plain_v;    /* plain base (algo 2 sparse) */
    float *out_u, *out_v;        /* float result before output conversion */
} or_debug;
const or_debug *or_last_debug(void);
void or_clear_debug(void);

#ifdef __cplusplus
}
#endif
#endif
