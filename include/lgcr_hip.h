/*
 * lgcr_hip.h — C ABI of the MI355X (gfx950) LGCR chroma-reconstruction
 * backend.  Plain C types only: pointers, sizes and POD structs.
 *
 * The reference has no device boundary: its VapourSynth glue calls the C++
 * stage functions directly (src/plugin.cpp:177-721) through a compile-time
 * SIMD vocabulary (src/backend.h:13-96), and ships an unused CUDA executor
 * (src/cuda_backend.h:36-65).  This header replaces both: the plugin glue
 * (or any other host, e.g. a serving loop) hands whole frames to one device
 * backend.  Each declaration below cites the reference interface it stands
 * in for.
 *
 * Threading: every entry point is thread-safe.  Concurrent lgcr_recon_frame
 * calls on one clip (VapourSynth fmParallel, plugin.cpp:1157) run on
 * separate per-call workspaces and HIP streams.
 */
#ifndef LGCR_HIP_H
#define LGCR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LGCR_ABI_VERSION 1

/* Status codes: the reference's StageStatus (src/pipeline.h:390-395) plus
 * allocation and admission failures. */
typedef enum lgcr_status {
    LGCR_OK = 0,
    LGCR_UNSUPPORTED = 1,       /* StageStatus::Unsupported */
    LGCR_INVALID_CONTEXT = 2,   /* StageStatus::InvalidContext */
    LGCR_BACKEND_FAILURE = 3,   /* StageStatus::BackendFailure */
    LGCR_OUT_OF_MEMORY = 4,
    LGCR_REJECTED = 5,          /* admission refused (HBM budget) */
    LGCR_INVALID_ARGUMENT = 6,
    LGCR_BUSY = 7               /* no free workspace (submit) / not finished (query) */
} lgcr_status;

/* Resampling kernels, same order as lgcr::Kernel (src/lgcr.h:46). */
typedef enum lgcr_kernel {
    LGCR_KERNEL_BILINEAR = 0,
    LGCR_KERNEL_BICUBIC = 1,
    LGCR_KERNEL_LANCZOS = 2,
    LGCR_KERNEL_SPLINE16 = 3,
    LGCR_KERNEL_SPLINE36 = 4,
    LGCR_KERNEL_JINC = 5
} lgcr_kernel;

/* Filter parameters: the fields of LGCRData (src/lgcr.h:128-159) after
 * argument parsing (src/plugin.cpp:1025-1133 for Recon, 939-976 for TRecon).
 * Fill with lgcr_params_default() and override. */
typedef struct lgcr_params {
    int kernel;               /* lgcr_kernel */
    double kp1, kp2;          /* lanczos/jinc taps, or bicubic b, c */
    double support;
    int radial;               /* 1 for jinc */
    double strength, sigma, sratio, sdb, reg, stretch, gsigma, rescue;
    int algo;                 /* 2, 4 or 6 */
    int ridge, cedge;
    double ms, qgate;
    int sparse;
    double bp;
    int bm;
    int trad, tsearch;        /* TRecon */
    double tsad;
    double ar_margin;         /* <0 disables the hull clamp */
    double shift_x, shift_y;  /* default chroma siting, luma units */
    int loc_set;              /* 1: ignore per-frame _ChromaLocation */
    int out_w, out_h;         /* output luma size */
} lgcr_params;

/* Input clip format (VSVideoFormat + constant size, src/plugin.cpp:1012-1023). */
typedef struct lgcr_format {
    int sample_type;          /* 0 integer, 1 float (32-bit) */
    int bits_per_sample;      /* 8..16, or 32 for float */
    int bytes_per_sample;     /* 1, 2 or 4 */
    int sub_sampling_w;       /* 0 or 1 (Sharpen: 0..2) */
    int sub_sampling_h;       /* 0 or 1 (Sharpen: 0..2) */
    int width, height;        /* luma size */
} lgcr_format;

typedef enum lgcr_memory { LGCR_MEMORY_HOST = 0, LGCR_MEMORY_DEVICE = 1 } lgcr_memory;

/* One frame: three planes with byte strides, all in host memory (the
 * VapourSynth frame buffers, borrowed for the call) or all in device memory
 * (resident frames of a serving loop).  Mirrors PlaneView
 * (src/pipeline.h:46-62). */
typedef struct lgcr_frame {
    void *data[3];
    ptrdiff_t stride[3];
    int memory;               /* lgcr_memory */
} lgcr_frame;

/* Stage ids, same order as lgcr::Stage (src/pipeline.h:98-118). */
enum {
    LGCR_STAGE_CONVERT_INPUT = 0, LGCR_STAGE_RESAMPLE_LUMA, LGCR_STAGE_BUILD_GUIDE_MAPS,
    LGCR_STAGE_BUILD_TRUST_MASK, LGCR_STAGE_BUILD_MUTUAL_GATE, LGCR_STAGE_BUILD_BASE_CHROMA,
    LGCR_STAGE_BUILD_LGF, LGCR_STAGE_BUILD_AFFINE_MAPS, LGCR_STAGE_BUILD_DETAIL,
    LGCR_STAGE_APPLY_GUIDED_CORRECTION, LGCR_STAGE_APPLY_SELECTOR, LGCR_STAGE_APPLY_DETAIL_TRANSFER,
    LGCR_STAGE_APPLY_COLLABORATIVE_FILTER, LGCR_STAGE_DOWNSAMPLE_BASE, LGCR_STAGE_DOWNSAMPLE_GUIDE,
    LGCR_STAGE_DOWNSAMPLE_CANDIDATE_SCORE, LGCR_STAGE_DOWNSAMPLE_OUTPUT, LGCR_STAGE_BACK_PROJECT,
    LGCR_STAGE_CONVERT_OUTPUT, LGCR_STAGE_COUNT
};

/* Per-frame metrics: PipelineMetrics (src/pipeline.h:315-336) as exported by
 * writeProfile (src/plugin.cpp:126-155).  Stage times are measured with HIP
 * events on the frame's stream. */
typedef struct lgcr_stats {
    double stage_us[LGCR_STAGE_COUNT];
    uint64_t stage_pixels[LGCR_STAGE_COUNT];
    double total_us;
    uint64_t output_pixels;
    uint64_t active_output_pixels;   /* sparse workset size */
    uint64_t active_chroma_pixels;
    uint64_t sparse_total_pixels;
    /* taps the reference's stages visit for this frame (PipelineMetrics::taps
     * and tapsVisited, src/pipeline.h:315-336): the plain base
     * (recon.cpp:1144-1147, 1453-1459) and the guided pass (recon.cpp:253-257) */
    uint64_t stage_taps[LGCR_STAGE_COUNT];
    uint64_t taps_visited;
} lgcr_stats;

typedef struct lgcr_device lgcr_device;
typedef struct lgcr_clip lgcr_clip;

/* Library / ABI identification. */
int lgcr_abi_version(void);

/* Thread-local message of the last failing call on this thread
 * (CudaPipeline::lastError, src/cuda_backend.h:40; cuda_backend.cpp:76-87). */
const char *lgcr_last_error(void);

/* Defaults of lgcrCreate (src/plugin.cpp:1025-1133): lanczos 3, algo 2,
 * strength 0.8, sparse 1 ...; out_w/out_h are left 0 (same size). */
void lgcr_params_default(lgcr_params *p);

/* Open a device context; hbm_budget_bytes caps the clip workspaces admitted
 * on it (0 = 90 % of free HBM).  Replaces CudaPipeline construction and
 * available() (src/cuda_backend.h:36-39, cuda_backend.cpp:60-74). */
lgcr_status lgcr_device_open(int ordinal, uint64_t hbm_budget_bytes, lgcr_device **out);
void lgcr_device_close(lgcr_device *dev);
/* Bytes currently reserved by clips on the device. */
uint64_t lgcr_device_reserved(const lgcr_device *dev);

/* Create a clip: validates params like lgcrCreate (src/plugin.cpp:1012-1142),
 * builds and uploads the geometry tables (GeometryCache,
 * src/kernels.cpp:94-131, 411-428) and jinc LUT (plugin.cpp:1071-1076), and
 * reserves the per-frame device workspace (FrameWorkspacePool,
 * src/lgcr.h:431-464; CudaPipeline::allocate, cuda_backend.cpp:108-131).
 * max_inflight bounds the workspaces (0 = 4, at most 64); admission fails with
 * LGCR_REJECTED when one workspace does not fit the device budget. */
lgcr_status lgcr_clip_create(lgcr_device *dev, const lgcr_params *p, const lgcr_format *fmt,
                             int max_inflight, lgcr_clip **out);
void lgcr_clip_destroy(lgcr_clip *clip);
/* Device bytes of one frame workspace of this clip. */
uint64_t lgcr_clip_workspace_bytes(const lgcr_clip *clip);

/* Reconstruct one frame (lgcrGetFrame's arAllFramesReady branch,
 * src/plugin.cpp:188-721).  shift_x/shift_y are the frame's chroma siting
 * after applyFrameChromaSiting (plugin.cpp:157-175); pass NAN to use the
 * clip defaults.  out->data[0] may be NULL when the output has the input
 * size: the luma plane is then shared with the source (plugin.cpp:265-269).
 * Blocks until the output is written.  stats may be NULL. */
lgcr_status lgcr_recon_frame(lgcr_clip *clip, const lgcr_frame *in, double shift_x,
                             double shift_y, lgcr_frame *out, lgcr_stats *stats);

/* Reconstruct n device-resident frames back to back on the clip's
 * workspaces without host synchronisation between frames; returns when all
 * are written.  The batched form of lgcr_recon_frame for serving loops. */
lgcr_status lgcr_recon_batch(lgcr_clip *clip, int n, const lgcr_frame *in,
                             lgcr_frame *out);

/* TRecon (tReconCreate / tReconGetFrame, src/plugin.cpp:740-990).  Params
 * from lgcr_params_default_trecon (lanczos 3, algo 2, sparse 0; strength,
 * sigma, sratio, sdb, gsigma, stretch, ar, ridge, sparse, ms, trad, tsearch,
 * tsad and bm are the TRecon arguments).  The output is 4:4:4 at the input
 * size with the luma plane rewritten (plugin.cpp:850-860).  The caller
 * supplies the neighbour frames chosen by temporalFrameNumbers
 * (plugin.cpp:729-738): n-trad..n+trad clamped to the clip, without the
 * centre and duplicates; at most 2*trad of them. */
void lgcr_params_default_trecon(lgcr_params *p);
lgcr_status lgcr_trecon_create(lgcr_device *dev, const lgcr_params *p, const lgcr_format *fmt,
                               int max_inflight, lgcr_clip **out);
lgcr_status lgcr_trecon_frame(lgcr_clip *clip, const lgcr_frame *in, const lgcr_frame *nbrs,
                              int n_nbrs, double shift_x, double shift_y, lgcr_frame *out,
                              lgcr_stats *stats);

/* Asynchronous submission for serving loops: the frame is queued on a free
 * workspace of the clip and the call returns at once with a ticket.  When
 * every workspace of the clip is busy (and no new one fits the device
 * budget) the call returns LGCR_BUSY without queuing anything, so a
 * scheduler can submit another clip's request instead.  Frames must be
 * device resident.  The workspace stays owned by the ticket until
 * lgcr_ticket_wait.  This is the non-blocking split of lgcr_recon_frame /
 * lgcr_trecon_frame (getFrame's arAllFramesReady branch,
 * src/plugin.cpp:188-721, 740-919), which VapourSynth's own frame
 * scheduler drives with blocking calls. */
typedef struct lgcr_ticket lgcr_ticket;
lgcr_status lgcr_recon_submit(lgcr_clip *clip, const lgcr_frame *in, double shift_x, double shift_y,
                              const lgcr_frame *out, lgcr_ticket **ticket);
lgcr_status lgcr_trecon_submit(lgcr_clip *clip, const lgcr_frame *in, const lgcr_frame *nbrs, int n_nbrs,
                               double shift_x, double shift_y, const lgcr_frame *out, lgcr_ticket **ticket);
/* LGCR_OK when the ticket's frame is complete, LGCR_BUSY while it runs. */
lgcr_status lgcr_ticket_query(const lgcr_ticket *ticket);
/* Block until the frame is complete, return its workspace to the clip and
 * free the ticket. */
lgcr_status lgcr_ticket_wait(lgcr_ticket *ticket);
/* No reference counterpart (serving loops): wait up to timeout_us
 * microseconds for any of n tickets to complete, spinning on their
 * completion events.  LGCR_OK with *index set to a completed ticket,
 * LGCR_BUSY on timeout (timeout_us <= 0 checks once).  Tickets stay owned
 * by the caller (retire them with lgcr_ticket_wait). */
lgcr_status lgcr_ticket_wait_any(const lgcr_ticket *const *tickets, int n, double timeout_us, int *index);

/* Sharpen (sharpenCreate / sharpenGetFrame, src/plugin.cpp:14-111): the
 * fields of SharpenData (src/lgcr.h:968-976).  The clip keeps its input
 * format (any planar YUV; chroma subsampling up to 2 per axis here).  Luma
 * is sharpened self-guided; subsampled chroma is gated by the footprint luma
 * of the sharpened luma with MPEG-2 siting (plugin.cpp:41-53).  Destroy
 * with lgcr_clip_destroy. */
typedef struct lgcr_sharpen_params {
    double alpha, sigma, sratio, gspatial;
    double ar_margin;         /* <0 disables the hull clamp */
} lgcr_sharpen_params;
void lgcr_sharpen_params_default(lgcr_sharpen_params *p);
lgcr_status lgcr_sharpen_create(lgcr_device *dev, const lgcr_sharpen_params *p, const lgcr_format *fmt,
                                int max_inflight, lgcr_clip **out);
lgcr_status lgcr_sharpen_frame(lgcr_clip *clip, const lgcr_frame *in, lgcr_frame *out);

/* Downsample (downsampleCreate / downsampleGetFrame, src/downsample.cpp:
 * 725-906): the fields of DownsampleData (src/lgcr.h:611-623).  Input is
 * planar YUV 4:4:4 (8-16 bit integer or 32-bit float) with even width and
 * height; the output is 4:2:0 of the same sample type.  out->data[0] may
 * be NULL: the plugin shares the source luma plane (downsample.cpp:737-742);
 * otherwise Y is copied.  U and V of `out` are (width/2) x (height/2).
 * Destroy with lgcr_clip_destroy. */
enum {
    LGCR_DS_KERNEL_SPLINE36 = 0, LGCR_DS_KERNEL_LANCZOS3 = 1, LGCR_DS_KERNEL_BINOMIAL = 2
};
typedef struct lgcr_downsample_params {
    int quality;              /* 0 fast, 1 balanced, 2 high */
    int kernel;               /* LGCR_DS_KERNEL_* */
    double strength;          /* guided blend, 0..1 (float in the filter) */
    int chroma_location;      /* H.273 _ChromaLocation written on the output:
                                 0 left, 1 center, 2 topleft, 3 top,
                                 4 bottomleft, 5 bottom (downsample.cpp:875-893) */
} lgcr_downsample_params;
void lgcr_downsample_params_default(lgcr_downsample_params *p);
lgcr_status lgcr_downsample_create(lgcr_device *dev, const lgcr_downsample_params *p, const lgcr_format *fmt,
                                   int max_inflight, lgcr_clip **out);
/* stats (may be NULL): DownsampleBase / Guide / CandidateScore / Output
 * device times, as writeDownsampleProfile reports them (downsample.cpp:31-49) */
lgcr_status lgcr_downsample_frame(lgcr_clip *clip, const lgcr_frame *in, lgcr_frame *out, lgcr_stats *stats);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, hosts that
 * forbid hipMalloc): lgcr_downsample_frame on device-resident frames whose
 * float intermediates live in a fixed 32 MiB array per GPU, set up when the
 * device is opened, instead of a clip workspace.  The call allocates
 * nothing; it returns LGCR_UNSUPPORTED when a frame needs more than 32 MiB
 * (44 bytes per chroma sample at quality 2 with strength above 0, else 8:
 * 1920x1080 fits at every quality, 3840x2160 only below quality 2) and
 * LGCR_INVALID_ARGUMENT for host frames.  Every call on one GPU shares that
 * GPU's array; calls from one process are serialised, and the call returns
 * after its work has finished.  Results equal lgcr_downsample_frame's. */
lgcr_status lgcr_downsample_frame_noalloc(lgcr_clip *clip, const lgcr_frame *in, lgcr_frame *out);

/* Per-stage device time accumulated while timing is on: the nanosecond
 * counters of PipelineMetrics (src/pipeline.h:315-336), here measured with
 * HIP events recorded around each stage on the workspace stream that runs
 * it.  lgcr_clip_set_timing resets the totals; lgcr_clip_stage_times copies
 * LGCR_STAGE_COUNT totals (microseconds) and launch counts. */
lgcr_status lgcr_clip_set_timing(lgcr_clip *clip, int enable);
lgcr_status lgcr_clip_stage_times(lgcr_clip *clip, double *us, uint64_t *launches);

/* Steady-state frame graphs (off by default).  When enabled, device-resident
 * Recon frames (lgcr_recon_frame, lgcr_recon_batch, lgcr_recon_submit)
 * replay a hipGraph of the frame's kernels captured once per workspace and
 * geometry; the frames' plane pointers reach it through a per-workspace
 * table in device memory.  Stage timing and stage capture use the direct
 * launches.  No counterpart in the reference (its frames run stage by stage
 * on the host). */
lgcr_status lgcr_clip_set_graphs(lgcr_clip *clip, int enable);

/* Multi-frame launches for lgcr_recon_batch: with frames_per_launch > 0 the
 * frames of a batch call go in groups of up to that many (at most the
 * clip's workspace count, at most 64), spread over up to four streams; each
 * frame runs its stages up to the mutual gate on its own workspace, then ONE
 * launch computes the fused tile pass (plain base, guided correction,
 * back-projection, output packing) over the tiles of the whole group.
 * Results are identical to frame-by-frame launches.  Pays for frames too
 * small to fill the device on their own (640x360); at 1080p frame-by-frame
 * launches on several streams are as fast or faster.  0 (the default) turns
 * it off.  Applies to device-resident frames without stage timing, stage
 * capture or frame graphs; frames whose guided pass is not the fused tile
 * pass complete on their group's stream.  No counterpart in the reference
 * (its frames run one at a time on the host). */
lgcr_status lgcr_clip_set_batching(lgcr_clip *clip, int frames_per_launch);
int lgcr_clip_get_batching(const lgcr_clip *clip);

/* Tile shapes of the kernels whose best shape depends on the frame size.
 * No counterpart in the reference (its CPU stages have no tiles). */
typedef struct lgcr_tuning {
    int tensor_tile_rows; /* Sobel tensor strips of 64 x {8, 16, 32} rows */
    int gate_tile_cols;   /* mutual-gate tiles of {16x16, 32x8, 64x4} (by columns) */
} lgcr_tuning;
lgcr_status lgcr_clip_get_tuning(const lgcr_clip *clip, lgcr_tuning *tuning);
/* LGCR_INVALID_ARGUMENT for a shape outside the candidates. */
lgcr_status lgcr_clip_set_tuning(lgcr_clip *clip, const lgcr_tuning *tuning);
/* Autotuner: runs `frame` (host or device planes of this Recon clip) once,
 * times every candidate shape of each tunable kernel on that frame's
 * intermediates with HIP events, and keeps the fastest.  The choice is
 * remembered in this process for (width, height, subsampling, sample type,
 * bits, algo): clips created later with the same key start with it.  Every
 * candidate computes identical results.  Call it (and lgcr_clip_set_tuning)
 * while no frame of the clip is in flight. */
lgcr_status lgcr_clip_autotune(lgcr_clip *clip, const lgcr_frame *frame);

/* Intermediates of the last frame processed on workspace 0 of a clip, for
 * parity checks against the stage functions of src/lgcr.h:555-962.
 * Copies up to `bytes` into host memory `dst` and stores the full size in
 * *needed.  Returns LGCR_INVALID_ARGUMENT for an unknown id. */
enum {
    LGCR_FETCH_Y = 0, LGCR_FETCH_U, LGCR_FETCH_V,           /* converted source, f32 */
    LGCR_FETCH_JXX, LGCR_FETCH_JXY, LGCR_FETCH_JYY,         /* structure tensor, f32 */
    LGCR_FETCH_LC, LGCR_FETCH_MS,                           /* chroma res, f32 */
    LGCR_FETCH_SEED, LGCR_FETCH_MASK,                       /* u8, luma res */
    LGCR_FETCH_CHROMA_MASK,                                 /* u8, chroma res */
    LGCR_FETCH_INDICES,                                     /* u32 active output indices */
    LGCR_FETCH_PLAIN_U, LGCR_FETCH_PLAIN_V,                 /* plain base, f32 */
    LGCR_FETCH_OUT_U, LGCR_FETCH_OUT_V,                     /* result before packing, f32 */
    LGCR_FETCH_MV,      /* TRecon, first neighbour: int16 mvx[nb] then mvy[nb], nb = 16x16 luma blocks */
    LGCR_FETCH_CONF,    /* TRecon, first neighbour: f32 block confidence after the chroma check */
    LGCR_FETCH_COUNT
};
lgcr_status lgcr_clip_fetch(lgcr_clip *clip, int what, void *dst, size_t bytes, size_t *needed);

/* Arithmetic self-check, no reference counterpart.  The guided tap weights
 * and the Downsample spatial weights take 1/y for y >= 1 as the hardware
 * reciprocal refined by one FMA Newton step, which the kernels rely on to
 * equal the IEEE quotient 1.0f / y.  This compares the two, bit for bit, for
 * every float bit pattern in [lo_bits, hi_bits) on device `ordinal`, and
 * returns the number of mismatches and the smallest mismatching pattern
 * (0xffffffff if none). */
lgcr_status lgcr_check_rcp(int ordinal, uint32_t lo_bits, uint32_t hi_bits, uint64_t *mismatches,
                           uint32_t *first_bad);

#ifdef __cplusplus
}
#endif
#endif
