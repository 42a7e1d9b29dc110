// Internal declarations shared by the HIP kernels and the host runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace lgcr_hip {

// Per-axis geometry of the chroma reconstruction (ChromaAxis,
// src/lgcr.h:739-755), resident in device memory.  Arrays have n entries;
// w and am have n * sup.
struct AxisDev {
    int n = 0, sup = 0;
    const int *start = nullptr;
    const float *w = nullptr;
    const float *am = nullptr;
    const float *wsum = nullptr;
    const float *wabs = nullptr;
    const float *pos = nullptr;
    const float *lpos = nullptr;
    const int *cb_i0 = nullptr;   // pos on the source chroma grid
    const float *cb_f = nullptr;
    const int *lb_i0 = nullptr;   // lpos on the source luma grid
    const float *lb_f = nullptr;
    // sup <= 6: one 64-byte record per output, 4 float4:
    // {start, pos, lpos, active-tap bits}, {lb_i0, lb_f, cb_i0, cb_f},
    // {w0, w1, w2, w3}, {w4, w5, wsum, wabs} (ints stored by bit pattern)
    const float4 *rec = nullptr;
    // largest number of source rows/columns read by 16 consecutive outputs
    // (start[o + 15] + sup - start[o] over o = 0, 16, 32, ...)
    int tile_rows16 = 0;
};

// Separable polyphase weight table (WeightTable, src/lgcr.h:601-606).
struct WeightsDev {
    int n = 0, sup = 0;
    const int *start = nullptr;
    const float *w = nullptr;
};

// Output columns per block of the class-table radial kernels, output rows
// per thread, and the largest source window (rows x columns) a block of the
// radial plain base stages in LDS.
constexpr int kSegCols = 256;
constexpr int kRadialRows = 4;
constexpr int kRadTileRows = 10, kRadTileCols = 264;

// Radial (jinc) base-weight phase table (RadialWeightTable,
// src/lgcr.h:793-811).  valid == 0 means the kernels evaluate the LUT per tap.
struct RadialDev {
    int valid = 0, sup_x = 0, sup_y = 0, nx = 0, ny = 0;
    const uint16_t *xphase = nullptr, *yphase = nullptr;
    const float *w = nullptr;
    const double *wd = nullptr;  // w as doubles
    // output columns of each x phase (xoff[p] .. xoff[p + 1]) and, per phase
    // pair, the sum and absolute sum of the active weights in (j, i) order
    const int *xs = nullptr, *xoff = nullptr;
    const double *wsa = nullptr;
    int max_count = 0;
    // waves per segment of kSegCols output columns: segw[segoff[s] ..
    // segoff[s + 1]) = {phase, first xs entry, end xs entry, 0}
    const int *segoff = nullptr;
    const int4 *segw = nullptr;
    int nseg = 0;
    int tile_fit = 0;  // every block's source window fits kRadTileRows x kRadTileCols
    const float *lut = nullptr;
    int lut_n = 0;
    float lut_scale = 1024.0f;
    float inv_wx = 1.0f, inv_wy = 1.0f;  // per-axis downscale AA factors
};

// Tables of the radial (jinc) luma resample (resampleRadial,
// src/kernels.cpp:276-334) when output coordinates fall into few classes of
// identical tap offsets per axis (2x upscaling: two).  w holds the tap
// weights of every (y class, x class) pair, sup x sup in (j, i) order as
// doubles, and wsum their sums in that order; xs lists the output columns
// of each x class (xoff[c] .. xoff[c + 1]).  valid == 0: no tables.
struct RadialResampleDev {
    int valid = 0, ncx = 0, ncy = 0, sup = 0, max_count = 0;
    const int *x0 = nullptr, *y0 = nullptr, *xs = nullptr, *xoff = nullptr;
    const uint8_t *ycls = nullptr;
    const double *w = nullptr, *wsum = nullptr;
    // waves per segment of kSegCols output columns (as in RadialDev)
    const int *segoff = nullptr;
    const int4 *segw = nullptr;
    int nseg = 0;
};

// Scalar parameters the kernels read, already narrowed exactly as the
// reference narrows them (float(d.sigma) etc.).
struct GuideParams {
    float sigma, sratio, sdb, strength, rescue, stretch, inv_gsigma2, ms, margin, sp_out;
    float frw, frh;
    int ridge, cedge, clamp_hull, ms_on, radial;
    // the planes come from an integer format, so every sample is finite
    // (the exact tap-loop shortcuts rely on finite weights)
    int finite = 0;
};

// One motion-compensated neighbour of TRecon (TemporalNbr, src/lgcr.h:729-736).
struct NbrDev {
    const float *u, *v, *lc;
    const int16_t *mvx, *mvy;
    const float *conf;
    int bw, bh, block;
};

// Temporal-tap context of the guided pass (recon.cpp:642-745).
// Temporal tap of one output column (row) for one integer block vector:
// the neighbour chroma sample it lands on, the residual offset in luma
// units and the separable kernel weight of that offset (recon.cpp:662-707).
struct TempTap {
    int tn;
    float dd;
    double k;
};

struct TemporalDev {
    const NbrDev *nbrs = nullptr;   // device array
    int n = 0;
    // per-clip tables over (output position, vector + search): TempTap for
    // x and y, and |sin(pi * frac(v / r))| for x then y
    const TempTap *tx = nullptr, *ty = nullptr;
    const float *nsin = nullptr;
    int search = 0;
    double sx = 0.0, sy = 0.0;      // frame siting (luma units)
    double rw = 1.0, rh = 1.0;
    int kernel = 2;
    double kp1 = 3.0, kp2 = 0.6;
    float inv_wx = 1.0f, inv_wy = 1.0f;
};

// algo-4 selector state of the guided pass (recon.cpp:762-811).
enum SelectorMode : int { SEL_NONE = 0, SEL_MAPS = 1, SEL_FLOAT = 2 };
struct SelectorDev {
    int mode = SEL_NONE;
    float strength = 0.0f;          // selectorStrength
    const float *lgf = nullptr;     // aU, bU, confU, aV, bV, confV planes (chroma grid)
    float *du = nullptr, *dv = nullptr, *w3c = nullptr;   // compressed, workset order
    float *gu = nullptr, *gv = nullptr, *w2p = nullptr, *w3p = nullptr;  // planes, output grid
};

// Sample formats of raw planes.
enum RawType : int { RAW_U8 = 0, RAW_U16 = 1, RAW_F32 = 2 };

// Frame planes of a captured frame graph, read on the device: the graph's
// conversion kernels take plane `slot` of entry `counter` and the graph's
// last node advances `counter`, so one graph replays over a table of frames
// queued in stream order on one workspace.
struct IoSlot {
    const void *in[3];
    ptrdiff_t in_stride[3];
    void *out[3];
    ptrdiff_t out_stride[3];
};
constexpr int kIoSlots = 1024;
// Upper bound on the frame workspaces of one clip (frames in flight).
constexpr int kMaxInflight = 64;
struct IoTable {
    uint32_t counter;
    uint32_t pad[63];
    IoSlot slot[kIoSlots];
};

struct RawPlane {
    const void *data;
    ptrdiff_t stride;   // bytes
    int w, h;
    float *dst;         // converted f32, stride w
    double scale, offset;
    const IoTable *io;  // non-null: data/stride are in[slot] of io's current entry
    int slot;
};

struct RawOut {
    const float *src;   // f32, stride w
    void *data;
    ptrdiff_t stride;   // bytes
    int w, h;
    float scale, offset;
    int lo, hi;
    const IoTable *io;  // non-null: data/stride are out[slot] of io's current entry
    int slot;
};

// Downsample 4:4:4 -> 4:2:0 (downsample.cpp).  Source planes are read raw
// (type RAW_*) and converted per sample as SourcePlane::at does.
struct DsPlane {
    const void *data;
    ptrdiff_t stride;  // bytes
    float scale, offset;
};
struct DsArgs {
    DsPlane y, u, v;
    int type;          // RAW_U8, RAW_U16 or RAW_F32
    int w, h, cw, ch;  // source (4:4:4) and output chroma sizes
    WeightsDev xw, yw; // baseline taps (start may be negative: loads clamp)
    float *base_u, *base_v;  // baseline, cw x ch
    float *gate, *cand_u, *cand_v;  // quality 2: gate (cw x ch), 4 candidate planes each
    int quality;
    float strength, shift_x, shift_y;
    RawOut out_u, out_v;     // packed U, V (src unused)
};

// baseline kernel: kDsBaseTileW x kDsBaseTileH outputs per block; the
// tile's source columns (start of its first output to start of its last +
// support) and source rows must span at most kDsBaseSegW and kDsBaseSegH
// Group launches of the stages before the tile pass: one launch serves up
// to kGroupMax frames (blockIdx.z = frame).  The frames' workspaces share
// one layout, so a buffer of frame k lies d[k] bytes from the same buffer of
// frame 0 (d[0] = 0).
constexpr int kGroupMax = 8;
struct GroupOff {
    long long d[kGroupMax];
};

constexpr int kDsBaseTileW = 32, kDsBaseTileH = 8, kDsBaseSegW = 80, kDsBaseSegH = 32;

// Kernel launchers (lgcr_kernels.hip).  All are asynchronous on `s`.
hipError_t launch_check_rcp(uint32_t lo, uint32_t hi, unsigned long long *out, hipStream_t s);
hipError_t launch_to_float(const RawPlane *planes, int count, int type, hipStream_t s);
// the conversion of up to kGroupMax frames (planes 3k .. 3k + 2: frame k)
hipError_t launch_to_float_group(const RawPlane *planes, int count, int type, hipStream_t s);
hipError_t launch_from_float(const RawOut *planes, int count, int type, hipStream_t s);
// structure tensor written interleaved with the luma: yt = {Y, jxx, jxy, jyy}
// seed: one bit per pixel, rows of (w + 63) / 64 words
// tile_rows: 4, 8 or 16 (tile 64 x tile_rows)
hipError_t launch_sobel_tensor(const float *y, int w, int h, float4 *yt, uint64_t *seed, float threshold,
                               hipStream_t s, int tile_rows = 8, const GroupOff *go = nullptr, int nf = 1);
// trust mask (dilation of the seed bits by radius) and the sparse workset:
// per-output-row counts, row offsets, row-major indices, chroma mask and
// its population in counters[0]
hipError_t launch_trust_workset(const uint64_t *seed, uint64_t *tmp, uint64_t *mask, int w, int h, int radius,
                                uint32_t *row_counts, uint32_t *row_offsets, uint32_t *indices, uint8_t *chroma_mask,
                                int cw, int ch, const AxisDev &ax, const AxisDev &ay, uint32_t *counters,
                                hipStream_t s);
// footprint luma; with luv also the interleaved chroma guide {lc, U, V, 0}
hipError_t launch_lc_map(const float *y, int w, int h, float *lc, int cw, int ch,
                         const int *xb, const int *xe, const int *yb, const int *ye,
                         hipStream_t s, const float *u = nullptr, const float *v = nullptr,
                         float4 *luv = nullptr, const GroupOff *go = nullptr, int nf = 1);
// count / ncount: the marked chroma samples are count[0] + ... + count[ncount - 1]
// tile_cols: 16, 32 or 64 (tiles 16x16, 32x8, 64x4)
hipError_t launch_mutual_gate(const float *lc, const float *u, const float *v, int w, int h,
                              const uint8_t *chroma_mask, const uint32_t *count, int ncount, float *gate,
                              hipStream_t s, int tile_cols = 32, const GroupOff *go = nullptr, int nf = 1);
hipError_t launch_plain_separable(const float *u, const float *v, int cw, int ch,
                                  const AxisDev &ax, const AxisDev &ay, float *hbuf,
                                  float *out_u, float *out_v, int clamp_hull, float margin,
                                  hipStream_t s);
hipError_t launch_plain_bilinear(const float *u, const float *v, int cw, int ch,
                                 const AxisDev &ax, const AxisDev &ay, float *out_u,
                                 float *out_v, hipStream_t s);
hipError_t launch_plain_radial(const float *u, const float *v, int cw, int ch,
                               const AxisDev &ax, const AxisDev &ay, const RadialDev &rd,
                               float *out_u, float *out_v, int clamp_hull, float margin,
                               hipStream_t s);
// Downsample stages: the separable baseline; then either the packed
// baseline (strength 0), the guided pass of quality 0/1 with packing, or the
// quality-2 candidates followed by the scored, packed result.
hipError_t launch_ds_base(const DsArgs &a, hipStream_t s);
hipError_t launch_ds_output(const DsArgs &a, hipStream_t s);
hipError_t launch_ds_candidates(const DsArgs &a, hipStream_t s);
hipError_t launch_ds_score(const DsArgs &a, hipStream_t s);
hipError_t launch_guided(const float *u, const float *v, int cw, int ch, const float *y,
                         int sw, int sh, const float4 *yt, const float4 *luv,
                         const float *lc, const float *ms, const AxisDev &ax, const AxisDev &ay,
                         const RadialDev &rd, const GuideParams &gp, const uint32_t *row_offsets,
                         const uint32_t *indices, float *out_u, float *out_v,
                         const SelectorDev &sel, const TemporalDev &tmp, hipStream_t s,
                         const uint64_t *mask_bits = nullptr, int mask_w = 0, int mask_h = 0,
                         int tiles_fit = 0, int plain_in_tiles = 0,
                         // non-null: the tiled pass also back-projects
                         // (bp_on) and writes these two output planes of
                         // out_type itself (out_u/out_v are not written)
                         const RawOut *fused_out = nullptr, int out_type = 0, float bp = 0.0f, int bp_on = 0,
                         // non-null and the pass is the fused tile pass:
                         // nothing is launched, the frame's arguments go here
                         // for launch_guided_batch (tiles > 0 when deferred)
                         struct GuidedDeferred *defer = nullptr);
// A frame's fused tile pass held back for a multi-frame launch: the kernel
// arguments (opaque here), the kernel variant and the frame's tile count.
constexpr size_t kGuidedDeferBytes = 1024;
struct GuidedDeferred {
    alignas(16) unsigned char args[kGuidedDeferBytes];
    int variant = -1;
    int tiles = 0;
};
// Device and pinned host tables of one multi-frame launch.
constexpr int kBatchMaxFrames = 64;
// lanes (streams) of a clip's multi-frame launches: the hardware queues a
// process gets by default
constexpr int kBatchLanes = 4;
size_t guided_batch_table_bytes();
// One launch of the fused tile pass over the deferred frames items[0..n)
// (all of one variant, n <= kBatchMaxFrames): the tables go through the
// pinned host staging `host` into `dev` (both guided_batch_table_bytes()
// long, the staging not rewritten until this launch's work has finished),
// then either one wave per tile over a (tiles, frames) grid, or (persistent)
// a grid of the device's resident workgroups strides over the tiles of all
// frames.
hipError_t launch_guided_batch(const GuidedDeferred *items, int n, void *dev, void *host, bool persistent,
                               hipStream_t s);
// whether every 8x8 output tile's chroma footprint (6 taps per axis) fits
// the LDS window of the tiled guided pass
int guided_tiles_fit(const int *xs, int ow, const int *ys, int oh);
// true when launch_guided takes the tiled pass, which reads the trust mask
// bits instead of the compacted index list
bool guided_tiled(int tiles_fit, int radial, int cedge, const AxisDev &ax, const AxisDev &ay, int sel_mode);
// Trust mask for the tiled pass at output size == luma size: dilated mask
// bits and chroma mask, no index list, one block per band of cband chroma
// rows.  band[b] = {first owned mask row, end of owned rows, first and last
// mask row the band reads}; the owned ranges partition the mask rows.
// Counters: counters[kCntChromaWords + b] marked chroma samples of band b,
// counters[kCntChromaWords + kMaxBands + b] active pixels of its mask rows.
// xr / yr: per chroma column / row, the range of output columns / rows whose
// bilinear footprint contains it.
constexpr int kCntChromaWords = 64, kMaxBands = 4096, kMaxCband = 64;
constexpr int kCounterWords = kCntChromaWords + 2 * kMaxBands;
size_t trust_fused_lds(int w, int radius, int rows, int cw);
hipError_t launch_trust_mask_tiles(const uint64_t *seed, uint64_t *mask, int w, int h, int radius, const int4 *band,
                                   int nbands, int cband, int max_rows, uint8_t *chroma_mask, int cw, int ch,
                                   const int2 *xr, const int2 *yr, uint32_t *counters, hipStream_t s,
                                   const GroupOff *go = nullptr, int nf = 1);
// TRecon motion search and chroma consistency (algos.cpp:683-807)
hipError_t launch_block_match(const float *cur, const float *nbr, int w, int h, int block,
                              int search, float tsad, int16_t *mvx, int16_t *mvy, float *conf,
                              hipStream_t s, const GroupOff *go = nullptr, int nf = 1);
hipError_t launch_chroma_consistency(const float *cu, const float *cv, const float *nu,
                                     const float *nv, int cw, int ch, int luma_w, int luma_h,
                                     int block, const int16_t *mvx, const int16_t *mvy, int bw,
                                     int bh, float *conf, hipStream_t s, const GroupOff *go = nullptr,
                                     int nf = 1);
// algo 4 (algos.cpp:241-593)
hipError_t launch_point_luma(const float *y, int w, int h, float *ls, int cw, int ch,
                             const int *xi, const float *xf, const int *yi, const float *yf,
                             hipStream_t s);
hipError_t launch_minmax5(const float *u, const float *v, int cw, int ch, float *mm, hipStream_t s);
hipError_t launch_lgf(const float *ls, const float *u, const float *v, int cw, int ch, double eps,
                      int cedge, const float *mm, const uint8_t *roi, const uint32_t *roi_count,
                      const float *gate, float ms, float *lgf, hipStream_t s);
hipError_t launch_roi_mark(const uint32_t *row_offsets, const uint32_t *indices, int ow, int oh,
                           const SelectorDev &sel, const AxisDev &ax, const AxisDev &ay,
                           uint8_t *roi, int cw, int ch, uint32_t *roi_count, hipStream_t s);
hipError_t launch_blend4(const uint32_t *row_offsets, const uint32_t *indices, const float *y,
                         int sw, int sh, const AxisDev &ax, const AxisDev &ay, int cw, int ch,
                         const SelectorDev &sel, float *out_u, float *out_v, hipStream_t s);
// algo 6 (algos.cpp:809-1698)
struct AffineDev {
    float *au, *av, *g, *mnu, *mxu, *mnv, *mxv, *rng, *deg;  // chroma grid
};
hipError_t launch_affine(const float *u, const float *v, int cw, int ch, const float *yc0,
                         const float *yc1, const float *yc2, const uint8_t *chroma_mask,
                         const uint32_t *chroma_count, double eps, double eps_sig,
                         const AffineDev &m, hipStream_t s);
hipError_t launch_detail_sub(const float *y, float *rec, size_t n, hipStream_t s);
hipError_t launch_nyquist(float *a, float *b, int w, int h, int xlevels, int ylevels,
                          float **result, hipStream_t s);
hipError_t launch_detail_transfer(const uint32_t *row_offsets, const uint32_t *indices,
                                  const float4 *yt, int sw,
                                  int sh, const AxisDev &ax, const AxisDev &ay, int cw, int ch,
                                  const AffineDev &m, const float *ms, const float *detail,
                                  float ms_strength, float qgate, float strength, float stretch,
                                  float ar, int clamp_hull, float *out_u, float *out_v,
                                  hipStream_t s);
// Sharpen (sharpenPlane, algos.cpp:595-659); scalars narrowed as the
// reference narrows them (float(d.alpha) etc.)
struct SharpenParams {
    float inv_g2, sigma, sratio, alpha, margin;
    int clamp_hull;
};
hipError_t launch_sharpen(const float *src, float *dst, int w, int h, const float *lc, const float *gy, int gw,
                          int gh, double rw, double rh, double sx, double sy, const SharpenParams &p,
                          hipStream_t s);
// collaborative chroma filter (bm.cpp:250-442), in place on u/v (w x h);
// guide_tmp (w x h floats) is used when the guide differs in size; scratch
// holds bm_scratch_bytes(w, h) bytes of per-anchor groups.
size_t bm_scratch_bytes(int w, int h);
hipError_t launch_collaborative(const float *guide, int gw, int gh, float *u, float *v, int w, int h,
                                float guide_sigma, float *guide_tmp, void *scratch, hipStream_t s);
// back-projection fused with packing of the two chroma planes (same-size
// 4:2:0); keep_float also stores the corrected float planes.
hipError_t launch_bp_pack(float *out_u, float *out_v, int ow, int oh, const float *u, const float *v, int cw,
                          int ch, float bp, const RawOut *planes, int type, int keep_float, hipStream_t s);
// io->counter += 1 (the last node of a frame graph)
hipError_t launch_io_advance(IoTable *io, hipStream_t s);
hipError_t launch_back_project(float *out_u, float *out_v, int ow, int oh, const float *u,
                               const float *v, int cw, int ch, float bp, hipStream_t s);
hipError_t launch_resample_h(const float *src, int sw, int sh, float *dst,
                             const WeightsDev &t, hipStream_t s);
hipError_t launch_resample_v(const float *src, int sw, int sh, float *dst,
                             const WeightsDev &t, hipStream_t s);
hipError_t launch_resample_radial_tab(const float *src, int sw, int sh, float *dst, int dw, int dh,
                                      const RadialResampleDev &t, hipStream_t s);
hipError_t launch_resample_radial(const float *src, int sw, int sh, float *dst, int dw, int dh,
                                  double support, const float *lut, int lut_n, float lut_scale,
                                  hipStream_t s);

}  // namespace lgcr_hip
