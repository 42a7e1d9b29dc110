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
};

// Separable polyphase weight table (WeightTable, src/lgcr.h:601-606).
struct WeightsDev {
    int n = 0, sup = 0;
    const int *start = nullptr;
    const float *w = nullptr;
};

// Scalar parameters the kernels read, already narrowed exactly as the
// reference narrows them (float(d.sigma) etc.).
struct GuideParams {
    float sigma, sratio, sdb, strength, rescue, stretch, inv_gsigma2, ms, margin, sp_out;
    float frw, frh;
    int ridge, cedge, clamp_hull, ms_on, radial;
};

// Sample formats of raw planes.
enum RawType : int { RAW_U8 = 0, RAW_U16 = 1, RAW_F32 = 2 };

struct RawPlane {
    const void *data;
    ptrdiff_t stride;   // bytes
    int w, h;
    float *dst;         // converted f32, stride w
    double scale, offset;
};

struct RawOut {
    const float *src;   // f32, stride w
    void *data;
    ptrdiff_t stride;   // bytes
    int w, h;
    float scale, offset;
    int lo, hi;
};

// Kernel launchers (lgcr_kernels.hip).  All are asynchronous on `s`.
hipError_t launch_to_float(const RawPlane *planes, int count, int type, hipStream_t s);
hipError_t launch_from_float(const RawOut *planes, int count, int type, hipStream_t s);
hipError_t launch_sobel_tensor(const float *y, int w, int h, float *jxx, float *jxy, float *jyy,
                               uint8_t *seed, float threshold, hipStream_t s);
hipError_t launch_lc_map(const float *y, int w, int h, float *lc, int cw, int ch,
                         const int *xb, const int *xe, const int *yb, const int *ye,
                         hipStream_t s);
hipError_t launch_dilate(const uint8_t *src, uint8_t *tmp, uint8_t *dst, int w, int h,
                         int radius, hipStream_t s);
hipError_t launch_workset(const uint8_t *mask, int w, int h, uint32_t *row_offsets,
                          uint32_t *indices, uint8_t *chroma_mask, int cw, int ch,
                          const AxisDev &ax, const AxisDev &ay, uint32_t *counters,
                          hipStream_t s);
hipError_t launch_mutual_gate(const float *lc, const float *u, const float *v, int w, int h,
                              float *grad, const uint8_t *chroma_mask,
                              const uint32_t *counters, float *gate, hipStream_t s);
hipError_t launch_plain_separable(const float *u, const float *v, int cw, int ch,
                                  const AxisDev &ax, const AxisDev &ay, float *hbuf,
                                  float *out_u, float *out_v, int clamp_hull, float margin,
                                  hipStream_t s);
hipError_t launch_plain_bilinear(const float *u, const float *v, int cw, int ch,
                                 const AxisDev &ax, const AxisDev &ay, float *out_u,
                                 float *out_v, hipStream_t s);
hipError_t launch_guided(const float *u, const float *v, int cw, int ch, const float *y,
                         int sw, int sh, const float *jxx, const float *jxy, const float *jyy,
                         const float *lc, const float *ms, const AxisDev &ax, const AxisDev &ay,
                         const GuideParams &gp, const uint32_t *row_offsets,
                         const uint32_t *indices, float *out_u, float *out_v,
                         hipStream_t s);
hipError_t launch_back_project(float *out_u, float *out_v, int ow, int oh, const float *u,
                               const float *v, int cw, int ch, float bp, hipStream_t s);
hipError_t launch_resample_h(const float *src, int sw, int sh, float *dst,
                             const WeightsDev &t, hipStream_t s);
hipError_t launch_resample_v(const float *src, int sw, int sh, float *dst,
                             const WeightsDev &t, hipStream_t s);
hipError_t launch_resample_radial(const float *src, int sw, int sh, float *dst, int dw, int dh,
                                  double support, const float *lut, int lut_n, float lut_scale,
                                  hipStream_t s);

}  // namespace lgcr_hip
