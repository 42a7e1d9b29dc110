// VapourSynth 4 plugin glue of the MI355X backend: core.lgcr.Recon, TRecon,
// Sharpen and Downsample with the reference's registration
// (VapourSynthPluginInit2, src/plugin.cpp:1160-1184, src/downsample.cpp:913-918),
// argument parsing, defaults, ranges and error strings (lgcrCreate
// src/plugin.cpp:998-1158, tReconCreate 921-990, sharpenCreate 84-111,
// downsampleCreate src/downsample.cpp:816-906), frame requests and frame
// properties.  Every pixel is computed through include/lgcr_hip.h; this file
// holds no compute of its own.
#include <VapourSynth4.h>

#include "lgcr_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#ifndef LGCR_SUFFIX
#define LGCR_SUFFIX ""
#endif

namespace {

// ---------------------------------------------------------------------------
// shared helpers
// ---------------------------------------------------------------------------

// One backend device per process, opened on first use and kept until exit.
struct DeviceSlot {
    std::mutex m;
    bool tried = false;
    lgcr_device *dev = nullptr;
    std::string error;
};
DeviceSlot g_device;

lgcr_device *backend_device(std::string &error) {
    std::lock_guard<std::mutex> lock(g_device.m);
    if (!g_device.tried) {
        g_device.tried = true;
        if (lgcr_device_open(0, 0, &g_device.dev) != LGCR_OK) {
            g_device.dev = nullptr;
            g_device.error = lgcr_last_error();
        }
    }
    if (!g_device.dev)
        error = g_device.error;
    return g_device.dev;
}

// LGCR_PROFILE set and not "0" (plugin.cpp:114-124).  Read per frame; the
// reference reads it once per process.
bool profiling_enabled() {
    const char *value = std::getenv("LGCR_PROFILE");
    return value && value[0] != '\0' && value[0] != '0';
}

// Stage property names in the backend's LGCR_STAGE_* order (stageName,
// src/pipeline.h:222-244).
const char *const kStageNames[LGCR_STAGE_COUNT] = {
    "input_conversion", "luma_resample", "buildGuideMaps", "buildTrustMask", "buildMutualGate",
    "plainChroma", "buildLGF", "buildAffineMaps", "buildDetail", "reconstructChroma",
    "selectorBlend", "detailTransfer", "collaborative_chroma", "downsample_base", "downsample_guide",
    "downsample_candidate_score", "downsample_output", "backProject", "output_conversion"};

// CPU sub-stage timers of the reference (cpuProfileSlotName,
// src/pipeline.h:282-312).  The backend computes on the device, so their CPU
// time is 0; the keys keep profile readers working.
const char *const kCpuSlots[] = {
    "guide_sobel", "guide_tensor", "guide_lcMap", "affine_candidate_box", "affine_candidate_bilinear",
    "affine_candidate_bicubic", "affine_minmax_u", "affine_minmax_v", "affine_window_moments",
    "affine_finalize_median", "guided_active_rows", "guided_metadata", "guided_tap_accumulation",
    "guided_normalization_selector", "trust_seed", "trust_dilate", "mutual_gradients", "mutual_gate",
    "plain_horizontal", "plain_vertical", "lgf_moments", "lgf_finalize", "affine_rolling", "affine_consume",
    "detail_reconstruct", "detail_nyquist", "detail_transfer"};

// writeProfile (plugin.cpp:126-155) from the backend's per-frame stats.
void write_profile(VSFrame *frame, const lgcr_stats &st, const uint64_t pixels[LGCR_STAGE_COUNT],
                   const VSAPI *vsapi) {
    VSMap *props = vsapi->getFramePropertiesRW(frame);
    for (int i = 0; i < LGCR_STAGE_COUNT; ++i) {
        const std::string base = std::string("_LGCR_") + kStageNames[i];
        vsapi->mapSetFloat(props, (base + "_us").c_str(), st.stage_us[i], maReplace);
        vsapi->mapSetInt(props, (base + "_pixels").c_str(), int64_t(pixels[i]), maReplace);
        vsapi->mapSetInt(props, (base + "_taps").c_str(), int64_t(st.stage_taps[i]), maReplace);
    }
    for (const char *slot : kCpuSlots)
        vsapi->mapSetFloat(props, (std::string("_LGCR_cpu_") + slot + "_us").c_str(), 0.0, maReplace);
    vsapi->mapSetFloat(props, "_LGCR_total_us", st.total_us, maReplace);
    vsapi->mapSetInt(props, "_LGCR_output_pixels", int64_t(st.output_pixels), maReplace);
    vsapi->mapSetInt(props, "_LGCR_taps_visited", int64_t(st.taps_visited), maReplace);
    const double active = st.sparse_total_pixels ? double(st.active_output_pixels) / double(st.sparse_total_pixels)
                                                 : 1.0;
    vsapi->mapSetFloat(props, "_LGCR_sparse_active_ratio", active, maReplace);
}

// applyFrameChromaSiting (plugin.cpp:157-175): H.273 _ChromaLocation of the
// frame sets the siting on subsampled axes unless loc was given.
void frame_siting(const VSVideoFormat *fmt, const VSFrame *frame, bool loc_set, double &sx, double &sy,
                  const VSAPI *vsapi) {
    if (fmt->subSamplingW == 0)
        sx = 0.0;
    if (fmt->subSamplingH == 0)
        sy = 0.0;
    if (loc_set)
        return;
    int err = 0;
    const int64_t cl = vsapi->mapGetInt(vsapi->getFramePropertiesRO(frame), "_ChromaLocation", 0, &err);
    if (err)
        return;
    if (fmt->subSamplingW > 0)
        sx = (cl == 0 || cl == 2 || cl == 4) ? -0.5 : 0.0;
    if (fmt->subSamplingH > 0)
        sy = (cl == 2 || cl == 3) ? -0.5 : ((cl == 4 || cl == 5) ? 0.5 : 0.0);
}

lgcr_format backend_format(const VSVideoFormat *f, int width, int height) {
    lgcr_format fmt;
    fmt.sample_type = f->sampleType == stFloat ? 1 : 0;
    fmt.bits_per_sample = f->bitsPerSample;
    fmt.bytes_per_sample = f->bytesPerSample;
    fmt.sub_sampling_w = f->subSamplingW;
    fmt.sub_sampling_h = f->subSamplingH;
    fmt.width = width;
    fmt.height = height;
    return fmt;
}

lgcr_frame read_frame(const VSFrame *f, const VSAPI *vsapi) {
    lgcr_frame in;
    for (int p = 0; p < 3; ++p) {
        in.data[p] = const_cast<uint8_t *>(vsapi->getReadPtr(f, p));
        in.stride[p] = vsapi->getStride(f, p);
    }
    in.memory = LGCR_MEMORY_HOST;
    return in;
}

// float argument read the way the reference reads it: narrowed to float
// (mapGetFloatSaturated) when given, the double default otherwise
double float_arg(const VSMap *in, const char *key, double dflt, const VSAPI *vsapi) {
    int err = 0;
    const float v = vsapi->mapGetFloatSaturated(in, key, 0, &err);
    return err ? dflt : double(v);
}

int64_t int_arg(const VSMap *in, const char *key, int64_t dflt, const VSAPI *vsapi) {
    int err = 0;
    const int v = vsapi->mapGetIntSaturated(in, key, 0, &err);
    return err ? dflt : int64_t(v);
}

// ---------------------------------------------------------------------------
// Recon
// ---------------------------------------------------------------------------

struct ReconData {
    VSNode *node = nullptr;
    const VSVideoInfo *vi_in = nullptr;
    VSVideoInfo vi_out{};
    lgcr_params p{};
    lgcr_clip *hip = nullptr;
    bool same = true;
};

const VSFrame *VS_CC recon_get_frame(int n, int reason, void *instance, void **, VSFrameContext *ctx,
                                     VSCore *core, const VSAPI *vsapi) {
    ReconData *d = static_cast<ReconData *>(instance);
    if (reason == arInitial) {
        vsapi->requestFrameFilter(n, d->node, ctx);
        return nullptr;
    }
    if (reason != arAllFramesReady)
        return nullptr;
    const VSFrame *src = vsapi->getFrameFilter(n, d->node, ctx);
    const VSVideoFormat *fmt = &d->vi_in->format;
    double sx = d->p.shift_x, sy = d->p.shift_y;
    frame_siting(fmt, src, d->p.loc_set != 0, sx, sy, vsapi);
    // same-size Recon shares the source luma plane (plugin.cpp:265-269)
    VSFrame *dst;
    if (d->same) {
        const VSFrame *plane_src[3] = {src, nullptr, nullptr};
        const int planes[3] = {0, 0, 0};
        dst = vsapi->newVideoFrame2(&d->vi_out.format, d->vi_out.width, d->vi_out.height, plane_src, planes, src,
                                    core);
    } else {
        dst = vsapi->newVideoFrame(&d->vi_out.format, d->vi_out.width, d->vi_out.height, src, core);
    }
    const lgcr_frame in = read_frame(src, vsapi);
    lgcr_frame out;
    for (int p = 0; p < 3; ++p) {
        const bool shared = d->same && p == 0;
        out.data[p] = shared ? nullptr : vsapi->getWritePtr(dst, p);
        out.stride[p] = shared ? 0 : vsapi->getStride(dst, p);
    }
    out.memory = LGCR_MEMORY_HOST;
    const bool profile = profiling_enabled();
    lgcr_stats st;
    if (lgcr_recon_frame(d->hip, &in, sx, sy, &out, profile ? &st : nullptr) != LGCR_OK) {
        vsapi->setFilterError((std::string("LGCR: ") + lgcr_last_error()).c_str(), ctx);
        vsapi->freeFrame(dst);
        vsapi->freeFrame(src);
        return nullptr;
    }
    // 4:4:4 output: the siting no longer applies (plugin.cpp:268-271)
    vsapi->mapDeleteKey(vsapi->getFramePropertiesRW(dst), "_ChromaLocation");
    if (profile) {
        // conversion work as the reference counts it (plugin.cpp:250-253,
        // 708-710): float input is viewed, float output converted only with bm
        uint64_t pixels[LGCR_STAGE_COUNT];
        for (int i = 0; i < LGCR_STAGE_COUNT; ++i)
            pixels[i] = st.stage_pixels[i];
        const bool is_float = fmt->sampleType == stFloat;
        const uint64_t sw = uint64_t(d->vi_in->width), sh = uint64_t(d->vi_in->height);
        const uint64_t cw = sw >> fmt->subSamplingW, ch = sh >> fmt->subSamplingH;
        pixels[LGCR_STAGE_CONVERT_INPUT] = is_float ? 0 : sw * sh + 2 * cw * ch;
        pixels[LGCR_STAGE_CONVERT_OUTPUT] =
            (!is_float || d->p.bm) ? 2 * uint64_t(d->vi_out.width) * uint64_t(d->vi_out.height) : 0;
        write_profile(dst, st, pixels, vsapi);
    }
    vsapi->freeFrame(src);
    return dst;
}

void VS_CC recon_free(void *instance, VSCore *, const VSAPI *vsapi) {
    ReconData *d = static_cast<ReconData *>(instance);
    lgcr_clip_destroy(d->hip);
    vsapi->freeNode(d->node);
    delete d;
}

void VS_CC recon_create(const VSMap *in, VSMap *out, void *, VSCore *core, const VSAPI *vsapi) {
    auto d = std::make_unique<ReconData>();
    d->node = vsapi->mapGetNode(in, "clip", 0, nullptr);
    d->vi_in = vsapi->getVideoInfo(d->node);
    const VSVideoFormat *fmt = &d->vi_in->format;
    auto fail = [&](const std::string &msg) {
        vsapi->mapSetError(out, msg.c_str());
        vsapi->freeNode(d->node);
    };
    if (fmt->colorFamily != cfYUV || (fmt->sampleType == stFloat && fmt->bitsPerSample != 32))
        return fail("LGCR: input must be planar YUV (8-16 bit int or 32 bit float)");
    if (fmt->subSamplingW > 1 || fmt->subSamplingH > 1)
        return fail("LGCR: only 4:4:4 / 4:2:2 / 4:2:0 input is supported");
    if (d->vi_in->width == 0 || d->vi_in->height == 0)
        return fail("LGCR: constant input dimensions required");
    lgcr_params &p = d->p;
    lgcr_params_default(&p);
    p.out_w = int(int_arg(in, "width", d->vi_in->width, vsapi));
    p.out_h = int(int_arg(in, "height", d->vi_in->height, vsapi));
    if (p.out_w <= 0 || p.out_h <= 0)
        return fail("LGCR: invalid width/height");
    int err = 0;
    const char *kname = vsapi->mapGetData(in, "kernel", 0, &err);
    const std::string k = err ? "lanczos" : kname;
    p.radial = 0;
    if (k == "bilinear") {
        p.kernel = LGCR_KERNEL_BILINEAR;
        p.support = 1.0;
    } else if (k == "bicubic") {
        p.kernel = LGCR_KERNEL_BICUBIC;
        p.support = 2.0;
        p.kp1 = float_arg(in, "b", 0.0, vsapi);
        p.kp2 = float_arg(in, "c", 0.6, vsapi);
    } else if (k == "lanczos" || k == "jinc") {
        const int64_t taps = int_arg(in, "taps", 3, vsapi);
        if (taps < 1 || taps > 64)
            return fail("LGCR: " + k + " taps must be in the range 1..64");
        p.kernel = k == "jinc" ? LGCR_KERNEL_JINC : LGCR_KERNEL_LANCZOS;
        p.kp1 = double(taps);
        p.support = p.kp1;
        p.radial = k == "jinc" ? 1 : 0;
    } else if (k == "spline16") {
        p.kernel = LGCR_KERNEL_SPLINE16;
        p.support = 2.0;
    } else if (k == "spline36") {
        p.kernel = LGCR_KERNEL_SPLINE36;
        p.support = 3.0;
    } else {
        return fail("LGCR: kernel must be \"bilinear\", \"bicubic\", \"lanczos\", \"spline16\", \"spline36\" or "
                    "\"jinc\"");
    }
    p.strength = float_arg(in, "strength", 0.8, vsapi);
    p.sigma = float_arg(in, "sigma", 0.01, vsapi);
    p.sratio = float_arg(in, "sratio", 0.15, vsapi);
    p.algo = int(int_arg(in, "algo", 2, vsapi));
    if (p.algo != 2 && p.algo != 4 && p.algo != 6)
        return fail("LGCR: algo must be 2 (guided), 4 (selector) or 6 (detail transfer)");
    p.sdb = float_arg(in, "sdb", 3.0, vsapi);
    p.stretch = float_arg(in, "stretch", 1.0, vsapi);
    p.gsigma = float_arg(in, "gsigma", 2.5, vsapi);
    p.rescue = float_arg(in, "rescue", 1.0, vsapi);
    p.ridge = int_arg(in, "ridge", 1, vsapi) != 0 ? 1 : 0;
    p.cedge = int_arg(in, "cedge", 0, vsapi) != 0 ? 1 : 0;
    p.ms = float_arg(in, "ms", 1.0, vsapi);
    p.qgate = float_arg(in, "qgate", 1.0, vsapi);
    p.ar_margin = float_arg(in, "ar", 0.0, vsapi);
    p.reg = float_arg(in, "reg", 0.005, vsapi);
    const char *loc = vsapi->mapGetData(in, "loc", 0, &err);
    p.shift_x = -0.5;   // MPEG-2 horizontal siting unless the frame says otherwise
    p.shift_y = 0.0;
    p.loc_set = 0;
    if (!err) {
        const std::string l = loc;
        if (l == "left")
            p.shift_x = -0.5;
        else if (l == "center")
            p.shift_x = 0.0;
        else
            return fail("LGCR: loc must be \"left\" or \"center\"");
        p.loc_set = 1;
    }
    p.sparse = int_arg(in, "sparse", 1, vsapi) != 0 ? 1 : 0;
    p.bm = int_arg(in, "bm", 0, vsapi) != 0 ? 1 : 0;
    p.bp = float_arg(in, "bp", 0.0, vsapi);
    if (fmt->subSamplingW == 0)
        p.shift_x = 0.0;
    if (fmt->subSamplingH == 0)
        p.shift_y = 0.0;
    if (p.strength < 0.0 || p.strength > 1.0 || p.sigma <= 0.0 || p.sratio <= 0.0 || p.sdb <= 0.0 ||
        p.stretch < 0.0 || p.gsigma <= 0.0 || p.reg <= 0.0 || p.bp < 0.0 || p.bp > 1.0 || p.ms < 0.0 ||
        p.ms > 1.0 || p.qgate < 0.0 || p.qgate > 1.0 || p.rescue < 0.0 || p.rescue > 1.0)
        return fail("LGCR: invalid parameter range (need 0<=strength/rescue<=1, 0<=bp/ms/qgate<=1, "
                    "sigma/sratio/sdb/gsigma/reg>0, stretch>=0)");
    d->vi_out = *d->vi_in;
    d->vi_out.width = p.out_w;
    d->vi_out.height = p.out_h;
    if (!vsapi->queryVideoFormat(&d->vi_out.format, cfYUV, fmt->sampleType, fmt->bitsPerSample, 0, 0, core))
        return fail("LGCR: failed to query 4:4:4 output format");
    d->same = p.out_w == d->vi_in->width && p.out_h == d->vi_in->height;
    std::string derr;
    lgcr_device *dev = backend_device(derr);
    if (!dev)
        return fail(derr);
    const lgcr_format bf = backend_format(fmt, d->vi_in->width, d->vi_in->height);
    if (lgcr_clip_create(dev, &p, &bf, 0, &d->hip) != LGCR_OK)
        return fail(std::string("LGCR: ") + lgcr_last_error());
    ReconData *data = d.release();
    const VSFilterDependency deps[] = {{data->node, rpGeneral}};
    vsapi->createVideoFilter(out, "LGCR", &data->vi_out, recon_get_frame, recon_free, fmParallel, deps, 1, data,
                             core);
}

// ---------------------------------------------------------------------------
// TRecon
// ---------------------------------------------------------------------------

struct TReconData {
    VSNode *node = nullptr;
    const VSVideoInfo *vi_in = nullptr;
    VSVideoInfo vi_out{};
    lgcr_params p{};
    lgcr_clip *hip = nullptr;
};

// temporalFrameNumbers (plugin.cpp:729-738): n-trad..n+trad clamped to the
// clip, the centre and duplicates removed
std::vector<int> temporal_frame_numbers(int n, int n_last, int trad) {
    std::vector<int> out;
    for (int k = -trad; k <= trad; ++k) {
        const int fn = std::min(std::max(n + k, 0), n_last);
        if (fn != n && std::find(out.begin(), out.end(), fn) == out.end())
            out.push_back(fn);
    }
    return out;
}

const VSFrame *VS_CC trecon_get_frame(int n, int reason, void *instance, void **, VSFrameContext *ctx,
                                      VSCore *core, const VSAPI *vsapi) {
    TReconData *d = static_cast<TReconData *>(instance);
    const std::vector<int> nbr_ids = temporal_frame_numbers(n, d->vi_in->numFrames - 1, d->p.trad);
    if (reason == arInitial) {
        vsapi->requestFrameFilter(n, d->node, ctx);
        for (int k : nbr_ids)
            vsapi->requestFrameFilter(k, d->node, ctx);
        return nullptr;
    }
    if (reason != arAllFramesReady)
        return nullptr;
    const VSFrame *src = vsapi->getFrameFilter(n, d->node, ctx);
    std::vector<const VSFrame *> nbr_frames;
    std::vector<lgcr_frame> nbrs;
    for (int k : nbr_ids) {
        nbr_frames.push_back(vsapi->getFrameFilter(k, d->node, ctx));
        nbrs.push_back(read_frame(nbr_frames.back(), vsapi));
    }
    double sx = -0.5, sy = 0.0;
    frame_siting(&d->vi_in->format, src, false, sx, sy, vsapi);
    // TRecon rewrites luma too (plugin.cpp:850-860)
    VSFrame *dst = vsapi->newVideoFrame(&d->vi_out.format, d->vi_out.width, d->vi_out.height, src, core);
    const lgcr_frame in = read_frame(src, vsapi);
    lgcr_frame out;
    for (int p = 0; p < 3; ++p) {
        out.data[p] = vsapi->getWritePtr(dst, p);
        out.stride[p] = vsapi->getStride(dst, p);
    }
    out.memory = LGCR_MEMORY_HOST;
    const lgcr_status st = lgcr_trecon_frame(d->hip, &in, nbrs.data(), int(nbrs.size()), sx, sy, &out, nullptr);
    for (const VSFrame *f : nbr_frames)
        vsapi->freeFrame(f);
    vsapi->freeFrame(src);
    if (st != LGCR_OK) {
        vsapi->setFilterError((std::string("TRecon: ") + lgcr_last_error()).c_str(), ctx);
        vsapi->freeFrame(dst);
        return nullptr;
    }
    vsapi->mapDeleteKey(vsapi->getFramePropertiesRW(dst), "_ChromaLocation");
    return dst;
}

void VS_CC trecon_free(void *instance, VSCore *, const VSAPI *vsapi) {
    TReconData *d = static_cast<TReconData *>(instance);
    lgcr_clip_destroy(d->hip);
    vsapi->freeNode(d->node);
    delete d;
}

void VS_CC trecon_create(const VSMap *in, VSMap *out, void *, VSCore *core, const VSAPI *vsapi) {
    auto d = std::make_unique<TReconData>();
    d->node = vsapi->mapGetNode(in, "clip", 0, nullptr);
    d->vi_in = vsapi->getVideoInfo(d->node);
    const VSVideoFormat *fmt = &d->vi_in->format;
    auto fail = [&](const std::string &msg) {
        vsapi->mapSetError(out, msg.c_str());
        vsapi->freeNode(d->node);
    };
    if (fmt->colorFamily != cfYUV || (fmt->sampleType == stFloat && fmt->bitsPerSample != 32) ||
        fmt->subSamplingW > 1 || fmt->subSamplingH > 1)
        return fail("TRecon: planar YUV 444/422/420, 8-16 bit int or 32 bit float");
    if (d->vi_in->width == 0 || d->vi_in->height == 0 || d->vi_in->numFrames == 0)
        return fail("TRecon: constant dimensions and frame count required");
    lgcr_params &p = d->p;
    lgcr_params_default_trecon(&p);
    p.strength = float_arg(in, "strength", 0.8, vsapi);
    p.sigma = float_arg(in, "sigma", 0.01, vsapi);
    p.sratio = float_arg(in, "sratio", 0.15, vsapi);
    p.sdb = float_arg(in, "sdb", 3.0, vsapi);
    p.gsigma = float_arg(in, "gsigma", 2.5, vsapi);
    p.stretch = float_arg(in, "stretch", 1.0, vsapi);
    p.ar_margin = float_arg(in, "ar", 0.0, vsapi);
    p.ms = float_arg(in, "ms", 1.0, vsapi);
    p.trad = int(int_arg(in, "trad", 1, vsapi));
    p.tsearch = int(int_arg(in, "tsearch", 6, vsapi));
    p.tsad = float_arg(in, "tsad", 0.02, vsapi);
    if (p.strength < 0.0 || p.strength > 1.0 || p.sigma <= 0.0 || p.sratio <= 0.0 || p.sdb <= 0.0 ||
        p.stretch < 0.0 || p.gsigma <= 0.0 || p.ms < 0.0 || p.ms > 1.0 || p.trad < 0 || p.trad > 8 ||
        p.tsearch < 0 || p.tsearch > 64 || p.tsad <= 0.0)
        return fail("TRecon: invalid parameter range (need 0<=strength<=1, "
                    "sigma/sratio/sdb/gsigma/tsad>0, stretch>=0, 0<=trad<=8, 0<=tsearch<=64)");
    p.ridge = int_arg(in, "ridge", 1, vsapi) != 0 ? 1 : 0;
    p.bm = int_arg(in, "bm", 0, vsapi) != 0 ? 1 : 0;
    p.sparse = int_arg(in, "sparse", 0, vsapi) != 0 ? 1 : 0;
    d->vi_out = *d->vi_in;
    if (!vsapi->queryVideoFormat(&d->vi_out.format, cfYUV, fmt->sampleType, fmt->bitsPerSample, 0, 0, core))
        return fail("TRecon: failed to query 4:4:4 output format");
    std::string derr;
    lgcr_device *dev = backend_device(derr);
    if (!dev)
        return fail(derr);
    const lgcr_format bf = backend_format(fmt, d->vi_in->width, d->vi_in->height);
    if (lgcr_trecon_create(dev, &p, &bf, 0, &d->hip) != LGCR_OK)
        return fail(std::string("TRecon: ") + lgcr_last_error());
    TReconData *data = d.release();
    const VSFilterDependency deps[] = {{data->node, rpGeneral}};
    vsapi->createVideoFilter(out, "TRecon", &data->vi_out, trecon_get_frame, trecon_free, fmParallelRequests,
                             deps, 1, data, core);
}

// ---------------------------------------------------------------------------
// Sharpen
// ---------------------------------------------------------------------------

struct SharpenData {
    VSNode *node = nullptr;
    const VSVideoInfo *vi = nullptr;
    lgcr_clip *hip = nullptr;
};

const VSFrame *VS_CC sharpen_get_frame(int n, int reason, void *instance, void **, VSFrameContext *ctx,
                                       VSCore *core, const VSAPI *vsapi) {
    SharpenData *d = static_cast<SharpenData *>(instance);
    if (reason == arInitial) {
        vsapi->requestFrameFilter(n, d->node, ctx);
        return nullptr;
    }
    if (reason != arAllFramesReady)
        return nullptr;
    const VSFrame *src = vsapi->getFrameFilter(n, d->node, ctx);
    // same format; the source props (and _ChromaLocation) stay (plugin.cpp:60)
    VSFrame *dst = vsapi->newVideoFrame(&d->vi->format, d->vi->width, d->vi->height, src, core);
    const lgcr_frame in = read_frame(src, vsapi);
    lgcr_frame out;
    for (int p = 0; p < 3; ++p) {
        out.data[p] = vsapi->getWritePtr(dst, p);
        out.stride[p] = vsapi->getStride(dst, p);
    }
    out.memory = LGCR_MEMORY_HOST;
    const lgcr_status st = lgcr_sharpen_frame(d->hip, &in, &out);
    vsapi->freeFrame(src);
    if (st != LGCR_OK) {
        vsapi->setFilterError((std::string("LGCR Sharpen: ") + lgcr_last_error()).c_str(), ctx);
        vsapi->freeFrame(dst);
        return nullptr;
    }
    return dst;
}

void VS_CC sharpen_free(void *instance, VSCore *, const VSAPI *vsapi) {
    SharpenData *d = static_cast<SharpenData *>(instance);
    lgcr_clip_destroy(d->hip);
    vsapi->freeNode(d->node);
    delete d;
}

void VS_CC sharpen_create(const VSMap *in, VSMap *out, void *, VSCore *core, const VSAPI *vsapi) {
    auto d = std::make_unique<SharpenData>();
    d->node = vsapi->mapGetNode(in, "clip", 0, nullptr);
    d->vi = vsapi->getVideoInfo(d->node);
    const VSVideoFormat *fmt = &d->vi->format;
    auto fail = [&](const std::string &msg) {
        vsapi->mapSetError(out, msg.c_str());
        vsapi->freeNode(d->node);
    };
    if (fmt->colorFamily != cfYUV || (fmt->sampleType == stFloat && fmt->bitsPerSample != 32))
        return fail("LGCR Sharpen: planar YUV only");
    lgcr_sharpen_params sp;
    lgcr_sharpen_params_default(&sp);
    sp.alpha = float_arg(in, "alpha", 0.3, vsapi);
    sp.sigma = float_arg(in, "sigma", 0.01, vsapi);
    sp.sratio = float_arg(in, "sratio", 0.15, vsapi);
    sp.gspatial = float_arg(in, "gspatial", 1.2, vsapi);
    sp.ar_margin = float_arg(in, "ar", 0.0, vsapi);
    if (sp.sigma <= 0.0 || sp.sratio <= 0.0 || sp.gspatial <= 0.0)
        return fail("LGCR Sharpen: sigma/sratio/gspatial must be > 0");
    std::string derr;
    lgcr_device *dev = backend_device(derr);
    if (!dev)
        return fail(derr);
    const lgcr_format bf = backend_format(fmt, d->vi->width, d->vi->height);
    if (lgcr_sharpen_create(dev, &sp, &bf, 0, &d->hip) != LGCR_OK)
        return fail(std::string("LGCR Sharpen: ") + lgcr_last_error());
    SharpenData *data = d.release();
    const VSFilterDependency deps[] = {{data->node, rpGeneral}};
    vsapi->createVideoFilter(out, "Sharpen", data->vi, sharpen_get_frame, sharpen_free, fmParallelRequests, deps, 1,
                             data, core);
}

// ---------------------------------------------------------------------------
// Downsample
// ---------------------------------------------------------------------------

struct DownsampleData {
    VSNode *node = nullptr;
    const VSVideoInfo *vi_in = nullptr;
    VSVideoInfo vi_out{};
    lgcr_clip *hip = nullptr;
    int chroma_location = 0;
    int quality = 0, kernel = 0;
    float strength = 1.0f;
};

// work counts of processFrame (downsample.cpp:767-789) for the profile
void downsample_profile(VSFrame *dst, const DownsampleData *d, const lgcr_stats &st, const VSAPI *vsapi) {
    const uint64_t w = uint64_t(d->vi_in->width), h = uint64_t(d->vi_in->height);
    const uint64_t cw = w / 2, ch = h / 2, pixels = 2 * cw * ch;
    uint64_t xsup = 12, ysup = 12;
    if (d->kernel == LGCR_DS_KERNEL_BINOMIAL) {
        // 5 taps at integer phase (shift -0.5 or +0.5), else 4 (downsample.cpp:52-58)
        const int loc = d->chroma_location;
        xsup = (loc == 0 || loc == 2 || loc == 4) ? 5 : 4;
        ysup = (loc == 0 || loc == 1) ? 4 : 5;
    }
    const uint64_t base_taps = 2 * (h * cw * xsup + ch * cw * ysup);
    uint64_t px[LGCR_STAGE_COUNT] = {}, taps[LGCR_STAGE_COUNT] = {};
    px[LGCR_STAGE_DOWNSAMPLE_BASE] = pixels;
    taps[LGCR_STAGE_DOWNSAMPLE_BASE] = base_taps;
    px[LGCR_STAGE_DOWNSAMPLE_OUTPUT] = pixels;
    if (d->strength > 0.0f) {
        const uint64_t diameter = 2 * uint64_t(d->quality + 2) + 1;
        const uint64_t candidates = d->quality == 0 ? 1 : (d->quality == 1 ? 2 : 5);
        px[LGCR_STAGE_DOWNSAMPLE_GUIDE] = pixels;
        taps[LGCR_STAGE_DOWNSAMPLE_GUIDE] = cw * ch * diameter * diameter * candidates;
        if (d->quality == 2) {
            px[LGCR_STAGE_DOWNSAMPLE_CANDIDATE_SCORE] = pixels;
            taps[LGCR_STAGE_DOWNSAMPLE_CANDIDATE_SCORE] = cw * ch * 16;
        }
    }
    // writeDownsampleProfile (downsample.cpp:31-49)
    VSMap *props = vsapi->getFramePropertiesRW(dst);
    for (int i = 0; i < LGCR_STAGE_COUNT; ++i) {
        const std::string base = std::string("_LGCR_") + kStageNames[i];
        vsapi->mapSetFloat(props, (base + "_us").c_str(), st.stage_us[i], maReplace);
        vsapi->mapSetInt(props, (base + "_pixels").c_str(), int64_t(px[i]), maReplace);
        vsapi->mapSetInt(props, (base + "_taps").c_str(), int64_t(taps[i]), maReplace);
    }
    vsapi->mapSetFloat(props, "_LGCR_total_us", st.total_us, maReplace);
    vsapi->mapSetInt(props, "_LGCR_output_pixels", int64_t(w * h + pixels), maReplace);
    vsapi->mapSetInt(props, "_LGCR_taps_visited", int64_t(base_taps), maReplace);
}

const VSFrame *VS_CC downsample_get_frame(int n, int reason, void *instance, void **, VSFrameContext *ctx,
                                          VSCore *core, const VSAPI *vsapi) {
    DownsampleData *d = static_cast<DownsampleData *>(instance);
    if (reason == arInitial) {
        vsapi->requestFrameFilter(n, d->node, ctx);
        return nullptr;
    }
    if (reason != arAllFramesReady)
        return nullptr;
    const VSFrame *src = vsapi->getFrameFilter(n, d->node, ctx);
    // 4:2:0 output with the source luma plane attached (plane sharing)
    const VSFrame *plane_src[3] = {src, nullptr, nullptr};
    const int planes[3] = {0, 0, 0};
    VSFrame *dst = vsapi->newVideoFrame2(&d->vi_out.format, d->vi_out.width, d->vi_out.height, plane_src, planes,
                                         src, core);
    const lgcr_frame in = read_frame(src, vsapi);
    lgcr_frame out;
    out.data[0] = nullptr;
    out.stride[0] = 0;
    for (int p = 1; p < 3; ++p) {
        out.data[p] = vsapi->getWritePtr(dst, p);
        out.stride[p] = vsapi->getStride(dst, p);
    }
    out.memory = LGCR_MEMORY_HOST;
    const bool profile = profiling_enabled();
    lgcr_stats st;
    const lgcr_status rc = lgcr_downsample_frame(d->hip, &in, &out, profile ? &st : nullptr);
    vsapi->freeFrame(src);
    if (rc != LGCR_OK) {
        vsapi->setFilterError((std::string("Downsample: ") + lgcr_last_error()).c_str(), ctx);
        vsapi->freeFrame(dst);
        return nullptr;
    }
    vsapi->mapSetInt(vsapi->getFramePropertiesRW(dst), "_ChromaLocation", d->chroma_location, maReplace);
    if (profile)
        downsample_profile(dst, d, st, vsapi);
    return dst;
}

void VS_CC downsample_free(void *instance, VSCore *, const VSAPI *vsapi) {
    DownsampleData *d = static_cast<DownsampleData *>(instance);
    lgcr_clip_destroy(d->hip);
    vsapi->freeNode(d->node);
    delete d;
}

void VS_CC downsample_create(const VSMap *in, VSMap *out, void *, VSCore *core, const VSAPI *vsapi) {
    auto d = std::make_unique<DownsampleData>();
    d->node = vsapi->mapGetNode(in, "clip", 0, nullptr);
    d->vi_in = vsapi->getVideoInfo(d->node);
    const VSVideoFormat *fmt = &d->vi_in->format;
    auto fail = [&](const std::string &msg) {
        vsapi->mapSetError(out, msg.c_str());
        vsapi->freeNode(d->node);
    };
    const bool integer = fmt->sampleType == stInteger && fmt->bitsPerSample >= 8 && fmt->bitsPerSample <= 16;
    const bool floating = fmt->sampleType == stFloat && fmt->bitsPerSample == 32;
    if (fmt->colorFamily != cfYUV || fmt->subSamplingW != 0 || fmt->subSamplingH != 0 || !(integer || floating))
        return fail("Downsample: input must be planar YUV444 (8-16 bit integer or float32)");
    if (d->vi_in->width <= 0 || d->vi_in->height <= 0 || d->vi_in->width % 2 || d->vi_in->height % 2)
        return fail("Downsample: constant even input width and height are required");
    lgcr_downsample_params dp;
    lgcr_downsample_params_default(&dp);
    dp.quality = int(int_arg(in, "quality", 0, vsapi));
    if (dp.quality < 0 || dp.quality > 2)
        return fail("Downsample: quality must be 0, 1, or 2");
    int err = 0;
    const char *kname = vsapi->mapGetData(in, "kernel", 0, &err);
    const std::string k = err ? "spline36" : kname;
    if (k == "spline36")
        dp.kernel = LGCR_DS_KERNEL_SPLINE36;
    else if (k == "lanczos3")
        dp.kernel = LGCR_DS_KERNEL_LANCZOS3;
    else if (k == "binomial")
        dp.kernel = LGCR_DS_KERNEL_BINOMIAL;
    else
        return fail("Downsample: kernel must be \"spline36\", \"lanczos3\", or \"binomial\"");
    const float strength = float(float_arg(in, "strength", 1.0, vsapi));
    if (!std::isfinite(strength) || strength < 0.0f || strength > 1.0f)
        return fail("Downsample: strength must be in the range 0..1");
    dp.strength = strength;
    const char *loc = vsapi->mapGetData(in, "loc", 0, &err);
    const std::string l = err ? "left" : loc;
    static const char *const kLocs[6] = {"left", "center", "topleft", "top", "bottomleft", "bottom"};
    int cl = -1;
    for (int i = 0; i < 6; ++i)
        if (l == kLocs[i])
            cl = i;
    if (cl < 0)
        return fail("Downsample: loc must be left, center, topleft, top, bottomleft, or bottom");
    dp.chroma_location = cl;
    d->chroma_location = cl;
    d->quality = dp.quality;
    d->kernel = dp.kernel;
    d->strength = strength;
    d->vi_out = *d->vi_in;
    if (!vsapi->queryVideoFormat(&d->vi_out.format, cfYUV, fmt->sampleType, fmt->bitsPerSample, 1, 1, core))
        return fail("Downsample: failed to query the 4:2:0 output format");
    std::string derr;
    lgcr_device *dev = backend_device(derr);
    if (!dev)
        return fail(derr);
    const lgcr_format bf = backend_format(fmt, d->vi_in->width, d->vi_in->height);
    if (lgcr_downsample_create(dev, &dp, &bf, 0, &d->hip) != LGCR_OK)
        return fail(std::string("Downsample: ") + lgcr_last_error());
    DownsampleData *data = d.release();
    const VSFilterDependency deps[] = {{data->node, rpGeneral}};
    vsapi->createVideoFilter(out, "Downsample", &data->vi_out, downsample_get_frame, downsample_free,
                             fmParallelRequests, deps, 1, data, core);
}

}  // namespace

VS_EXTERNAL_API(void) VapourSynthPluginInit2(VSPlugin *plugin, const VSPLUGINAPI *vspapi) {
    vspapi->configPlugin("dev.bsflab.lgcr" LGCR_SUFFIX, "lgcr" LGCR_SUFFIX,
                         "Luma-guided chroma reconstruction with direction-aware kernels", VS_MAKE_VERSION(2, 1),
                         VAPOURSYNTH_API_VERSION, 0, plugin);
    vspapi->registerFunction("Recon",
                             "clip:vnode;width:int:opt;height:int:opt;kernel:data:opt;taps:int:opt;algo:int:opt;"
                             "b:float:opt;c:float:opt;strength:float:opt;sigma:float:opt;sratio:float:opt;"
                             "sdb:float:opt;stretch:float:opt;gsigma:float:opt;rescue:float:opt;ridge:int:opt;"
                             "cedge:int:opt;ar:float:opt;reg:float:opt;loc:data:opt;sparse:int:opt;bp:float:opt;"
                             "ms:float:opt;qgate:float:opt;bm:int:opt",
                             "clip:vnode", recon_create, nullptr, plugin);
    vspapi->registerFunction("Sharpen",
                             "clip:vnode;alpha:float:opt;sigma:float:opt;sratio:float:opt;gspatial:float:opt;"
                             "ar:float:opt",
                             "clip:vnode", sharpen_create, nullptr, plugin);
    vspapi->registerFunction("TRecon",
                             "clip:vnode;strength:float:opt;sigma:float:opt;sratio:float:opt;sdb:float:opt;"
                             "gsigma:float:opt;stretch:float:opt;ar:float:opt;ridge:int:opt;sparse:int:opt;"
                             "ms:float:opt;trad:int:opt;tsearch:int:opt;tsad:float:opt;bm:int:opt",
                             "clip:vnode", trecon_create, nullptr, plugin);
    vspapi->registerFunction("Downsample", "clip:vnode;quality:int:opt;kernel:data:opt;strength:float:opt;loc:data:opt",
                             "clip:vnode", downsample_create, nullptr, plugin);
}
