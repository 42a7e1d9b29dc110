// gfx950 kernels of the LGCR Recon path.
//
// Everything is FP32 with FP64 accumulators wherever the reference's scalar
// build accumulates in double; the library is compiled with
// -ffp-contract=off so every product and sum rounds as written.  Comments
// cite the reference stage each kernel computes.
#include "lgcr_internal.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <type_traits>

namespace lgcr_hip {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
// std::min / std::max / std::clamp semantics (NaN propagation included).
__device__ __forceinline__ float fmn(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float fmx(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float fclamp(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }
__device__ __forceinline__ double dclamp(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }
// max(m, |a|, |b|) for m >= +0 and finite a, b: |a| and |b| are finite and
// non-negative, so the VALU maximum equals the ordered compares' result.
__device__ __forceinline__ float max3abs(float m, float a, float b) {
    float r;
    __asm__("v_max3_f32 %0, %1, |%2|, |%3|" : "=v"(r) : "v"(m), "v"(a), "v"(b));
    return r;
}

// First and last active tap of a 6-tap axis record (the active taps form
// one run [lo, hi]) and the offsets of taps clamp(i, lo, hi) * scale.
__device__ __forceinline__ void tap_run(uint32_t bits, int off[6], int scale) {
    const int lo = bits ? __builtin_ctz(bits) : 0;
    const int hi = bits ? 31 - __builtin_clz(bits) : 5;
    _Pragma("unroll") for (int i = 0; i < 6; ++i)
        off[i] = min(max(i, lo), hi) * scale;
}

// Maximum of finite non-negative floats (no -0.0): such floats order like
// their bit patterns read as unsigned integers.
__device__ __forceinline__ float nn_max(float a, float b) {
    const uint32_t x = __float_as_uint(a), y = __float_as_uint(b);
    return __uint_as_float(x < y ? y : x);
}

// Planes hold fewer than 2^31 samples (checked at clip creation), so sample
// offsets are 32-bit.
__device__ __forceinline__ float at(const float *p, int w, int h, int x, int y) {
    return p[(uint32_t)(clampi(y, 0, h - 1) * w + clampi(x, 0, w - 1))];
}

// Correctly rounded 1/y for y >= 1.  rcp_nr is the hardware reciprocal
// refined by one FMA Newton step; on gfx950 it equals the IEEE quotient
// 1.0f / y for every float y in [1, 2^126) (checked exhaustively on the
// device).  recip_ge1 sends larger, inf and NaN arguments to the division.
__device__ __forceinline__ float rcp_nr(float y) {
    const float r = __builtin_amdgcn_rcpf(y);
    const float e = __builtin_fmaf(-y, r, 1.0f);
    return __builtin_fmaf(e, r, r);
}
__device__ __forceinline__ float recip_ge1(float y) {
    if (y < 0x1p126f)
        return rcp_nr(y);
    return 1.0f / y;
}

// acc + (double)a * b.  A product of two floats is exact in double, so the
// fused form rounds once, exactly like the separate multiply and add.
__device__ __forceinline__ double dacc(double acc, float a, float b) {
    return __builtin_fma((double)a, (double)b, acc);
}

// XCD-aware block order.  Workgroups are dispatched round-robin over the 8
// XCDs, each with its own L2; remapping the linear block id gives every XCD
// a contiguous range of tiles, so halos shared by neighbouring tiles hit the
// same L2.  Bijective for any block count.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_linear() {
    const unsigned nb = gridDim.x * gridDim.y;
    const unsigned lin = blockIdx.y * gridDim.x + blockIdx.x;
    const unsigned xcd = lin % kXcds, slot = lin / kXcds;
    const unsigned q = nb / kXcds, r = nb % kXcds;
    return xcd * q + min(xcd, r) + slot;
}
// (x, y) block coordinates of the remapped order for a 2D grid
__device__ __forceinline__ void xcd_block2(unsigned &bx, unsigned &by) {
    const unsigned l = xcd_linear();
    bx = l % gridDim.x;
    by = l / gridDim.x;
}

// Bilinear sample at double coordinates with clamped taps (lgcr.h:467-475).
__device__ float bilinear(const float *p, int w, int h, double x, double y) {
    const int x0 = (int)floor(x), y0 = (int)floor(y);
    const float fx = (float)(x - x0), fy = (float)(y - y0);
    const float a = at(p, w, h, x0, y0), b = at(p, w, h, x0 + 1, y0);
    const float c = at(p, w, h, x0, y0 + 1), d = at(p, w, h, x0 + 1, y0 + 1);
    return a + (b - a) * fx + (c - a) * fy + (a - b - c + d) * fx * fy;
}

// The same sample for a coordinate that is a float value: floorf and the
// float subtraction give exactly the integer part and the rounded fraction
// the double path computes.
__device__ __forceinline__ float bilinear_f(const float *p, int w, int h, float x, float y) {
    const float flx = floorf(x), fly = floorf(y);
    const int x0 = (int)flx, y0 = (int)fly;
    const float fx = x - flx, fy = y - fly;
    const float a = at(p, w, h, x0, y0), b = at(p, w, h, x0 + 1, y0);
    const float c = at(p, w, h, x0, y0 + 1), d = at(p, w, h, x0 + 1, y0 + 1);
    return a + (b - a) * fx + (c - a) * fy + (a - b - c + d) * fx * fy;
}

// Bilinear sample with pre-clamped index/fraction (lgcr.h:494-508).
__device__ float bilinear_fast(const float *p, int w, int h, int x0, float fx, int y0, float fy) {
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

// bilinear_fast on every component of an interleaved float4 plane.
__device__ float4 bilinear_fast4(const float4 *p, int w, int h, int x0, float fx, int y0, float fy) {
    float4 r;
    if (w == 1) {
        const float4 a = p[(size_t)y0 * w];
        if (h == 1)
            return a;
        const float4 c = p[(size_t)(y0 + 1) * w];
        r.x = a.x + (c.x - a.x) * fy; r.y = a.y + (c.y - a.y) * fy;
        r.z = a.z + (c.z - a.z) * fy; r.w = a.w + (c.w - a.w) * fy;
        return r;
    }
    const float4 *r0 = p + (size_t)y0 * w + x0;
    const float4 a = r0[0], b = r0[1];
    if (h == 1) {
        r.x = a.x + (b.x - a.x) * fx; r.y = a.y + (b.y - a.y) * fx;
        r.z = a.z + (b.z - a.z) * fx; r.w = a.w + (b.w - a.w) * fx;
        return r;
    }
    const float4 c = r0[w], d = r0[w + 1];
    r.x = a.x + (b.x - a.x) * fx + (c.x - a.x) * fy + (a.x - b.x - c.x + d.x) * fx * fy;
    r.y = a.y + (b.y - a.y) * fx + (c.y - a.y) * fy + (a.y - b.y - c.y + d.y) * fx * fy;
    r.z = a.z + (b.z - a.z) * fx + (c.z - a.z) * fy + (a.z - b.z - c.z + d.z) * fx * fy;
    r.w = a.w + (b.w - a.w) * fx + (c.w - a.w) * fy + (a.w - b.w - c.w + d.w) * fx * fy;
    return r;
}

struct Dir { float coherence, nx, ny; };

// Principal eigenvector of a 2x2 structure tensor (lgcr.h:516-538).
__device__ Dir principal_dir(float jxx, float jxy, float jyy) {
    Dir d{0.0f, 1.0f, 0.0f};
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


// ---------------------------------------------------------------------------
// format conversion (planeToFloat / floatToPlane, lgcr.h:986-1013)
// ---------------------------------------------------------------------------

struct InBatch { RawPlane p[3]; };
struct OutBatch { RawOut p[3]; };

// A device pointer read from memory (a table entry) as a global-memory
// pointer, so its loads are global rather than flat ones.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gmem(const T *p) {
    return (const __attribute__((address_space(1))) T *)p;
}

// plane address and stride of a conversion: by value, or from the frame
// table of a captured frame graph
__device__ __forceinline__ const char *plane_in(const RawPlane &P, ptrdiff_t &stride) {
    if (P.io) {
        const IoSlot &e = P.io->slot[P.io->counter];
        stride = e.in_stride[P.slot];
        return static_cast<const char *>(e.in[P.slot]);
    }
    stride = P.stride;
    return static_cast<const char *>(P.data);
}
__device__ __forceinline__ char *plane_out(const RawOut &P, ptrdiff_t &stride) {
    if (P.io) {
        const IoSlot &e = P.io->slot[P.io->counter];
        stride = e.out_stride[P.slot];
        return static_cast<char *>(e.out[P.slot]);
    }
    stride = P.stride;
    return static_cast<char *>(P.data);
}

__global__ void k_io_advance(IoTable *io) {
    if (threadIdx.x == 0)
        io->counter += 1;
}

// One block per row and plane; each thread converts columns x, x+256, ...
// four at a time so several loads are in flight per lane.
constexpr int kConvUnroll = 4;

// Four samples of one type, loaded together.
template <typename T>
struct alignas(4 * sizeof(T)) Quad {
    T v[4];
};

// One row of one plane.  COPY: float input, copied bit for bit.  Rows whose
// source and destination are aligned for it move four samples per lane and
// load (the tail and unaligned rows go one sample at a time); the value of
// every sample is the same expression either way.
template <typename T, bool COPY>
__device__ __forceinline__ void to_float_row(const RawPlane &P, int y) {
    ptrdiff_t stride;
    const char *base = plane_in(P, stride);
    const T *row = reinterpret_cast<const T *>(base + stride * y);
    float *dst = P.dst + (size_t)y * P.w;
    const double scale = P.scale, offset = P.offset;
    auto cv = [&](T v) { return COPY ? (float)v : (float)((double)v * scale + offset); };
    int x_tail = 0;
    if ((reinterpret_cast<uintptr_t>(row) % sizeof(Quad<T>)) == 0 && (reinterpret_cast<uintptr_t>(dst) % 16) == 0) {
        const int n4 = P.w / 4;
        const Quad<T> *src4 = reinterpret_cast<const Quad<T> *>(row);
        float4 *dst4 = reinterpret_cast<float4 *>(dst);
        for (int i0 = threadIdx.x; i0 < n4; i0 += kBlock * 2) {
            Quad<T> q[2];
#pragma unroll
            for (int k = 0; k < 2; ++k)
                if (i0 + k * kBlock < n4)
                    q[k] = src4[i0 + k * kBlock];
#pragma unroll
            for (int k = 0; k < 2; ++k)
                if (i0 + k * kBlock < n4)
                    dst4[i0 + k * kBlock] = make_float4(cv(q[k].v[0]), cv(q[k].v[1]), cv(q[k].v[2]), cv(q[k].v[3]));
        }
        x_tail = 4 * n4;
    }
    for (int x0 = x_tail + (int)threadIdx.x; x0 < P.w; x0 += kBlock * kConvUnroll) {
        T v[kConvUnroll];
#pragma unroll
        for (int i = 0; i < kConvUnroll; ++i) {
            const int x = x0 + i * kBlock;
            v[i] = x < P.w ? gmem(row)[x] : T(0);
        }
#pragma unroll
        for (int i = 0; i < kConvUnroll; ++i) {
            const int x = x0 + i * kBlock;
            if (x < P.w)
                dst[x] = cv(v[i]);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_to_float(InBatch b) {
    const RawPlane &P = b.p[blockIdx.z];
    if ((int)blockIdx.y < P.h)
        to_float_row<T, false>(P, blockIdx.y);
}

__global__ __launch_bounds__(kBlock) void k_copy_f32(InBatch b) {
    const RawPlane &P = b.p[blockIdx.z];
    if ((int)blockIdx.y < P.h)
        to_float_row<float, true>(P, blockIdx.y);
}

// The conversion of a group of frames in one launch: planes 3k .. 3k + 2
// are frame k's Y, U and V.
struct InBatchG { RawPlane p[3 * kGroupMax]; };

template <typename T, bool COPY>
__global__ __launch_bounds__(kBlock) void k_to_float_group(InBatchG b) {
    const RawPlane &P = b.p[blockIdx.z];
    if ((int)blockIdx.y < P.h)
        to_float_row<T, COPY>(P, blockIdx.y);
}

// Group launches of the per-frame stages: blockIdx.z is the frame of the
// group, and every frame's workspace has the same layout, so each buffer of
// frame z lies go.d[z] bytes from the same buffer of frame 0 (d[0] = 0; a
// frame-by-frame launch passes zeros and one grid layer).  Null pointers
// stay null.
__device__ __forceinline__ long long group_off(const GroupOff &go) {
    const unsigned z = blockIdx.z;
    long long d = 0;
    _Pragma("unroll") for (int i = 1; i < kGroupMax; ++i)
        d = z == (unsigned)i ? go.d[i] : d;
    return d;
}
// byte arithmetic on the pointer itself (no round trip through an integer,
// which would hide that it addresses global memory and turn its loads and
// stores into flat ones)
template <class P>
__device__ __forceinline__ const P *gshift(const P *p, long long d) {
    return p ? reinterpret_cast<const P *>(reinterpret_cast<const char *>(p) + d) : p;
}
template <class P>
__device__ __forceinline__ P *gshift(P *p, long long d) {
    return p ? reinterpret_cast<P *>(reinterpret_cast<char *>(p) + d) : p;
}

template <typename T> __device__ __forceinline__ T pack_sample(float s, const RawOut &P) {
    const int v = (int)(s * P.scale + P.offset + 0.5f);
    return (T)clampi(v, P.lo, P.hi);
}

// One output row of one plane.  COPY: float output, copied bit for bit;
// otherwise packed (floatToPlane).  Aligned rows move four samples per lane
// load and store, the tail and unaligned rows one sample at a time.
template <typename T, bool COPY>
__device__ __forceinline__ void from_float_row(const RawOut &P, int y) {
    ptrdiff_t stride;
    char *base = plane_out(P, stride);
    T *row = reinterpret_cast<T *>(base + stride * y);
    const float *src = P.src + (size_t)y * P.w;
    auto pk = [&](float v) -> T {
        if constexpr (COPY)
            return v;
        else
            return pack_sample<T>(v, P);
    };
    int x_tail = 0;
    if ((reinterpret_cast<uintptr_t>(row) % sizeof(Quad<T>)) == 0 && (reinterpret_cast<uintptr_t>(src) % 16) == 0) {
        const int n4 = P.w / 4;
        const float4 *src4 = reinterpret_cast<const float4 *>(src);
        Quad<T> *dst4 = reinterpret_cast<Quad<T> *>(row);
        for (int i0 = threadIdx.x; i0 < n4; i0 += kBlock * 2) {
            float4 q[2];
#pragma unroll
            for (int k = 0; k < 2; ++k)
                if (i0 + k * kBlock < n4)
                    q[k] = src4[i0 + k * kBlock];
#pragma unroll
            for (int k = 0; k < 2; ++k)
                if (i0 + k * kBlock < n4) {
                    Quad<T> o;
                    o.v[0] = pk(q[k].x);
                    o.v[1] = pk(q[k].y);
                    o.v[2] = pk(q[k].z);
                    o.v[3] = pk(q[k].w);
                    dst4[i0 + k * kBlock] = o;
                }
        }
        x_tail = 4 * n4;
    }
    for (int x0 = x_tail + (int)threadIdx.x; x0 < P.w; x0 += kBlock * kConvUnroll) {
        float v[kConvUnroll];
#pragma unroll
        for (int i = 0; i < kConvUnroll; ++i) {
            const int x = x0 + i * kBlock;
            v[i] = x < P.w ? src[x] : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < kConvUnroll; ++i) {
            const int x = x0 + i * kBlock;
            if (x < P.w)
                row[x] = pk(v[i]);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_from_float(OutBatch b) {
    const RawOut &P = b.p[blockIdx.z];
    if ((int)blockIdx.y < P.h)
        from_float_row<T, false>(P, blockIdx.y);
}

__global__ __launch_bounds__(kBlock) void k_from_float_f32(OutBatch b) {
    const RawOut &P = b.p[blockIdx.z];
    if ((int)blockIdx.y < P.h)
        from_float_row<float, true>(P, blockIdx.y);
}

// Back-projection fused with output packing (backProject, algos.cpp:663-677,
// then floatToPlane): same-size 4:2:0 only, where every output pixel lies in
// exactly one 2x2 block of one chroma sample.  keep_float also writes the
// corrected float planes back (intermediate capture).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bp_pack(float *ou, float *ov, int ow, int oh, const float *u,
                                                    const float *v, int cw, int ch, float bp, OutBatch b,
                                                    int keep_float) {
    const int cx = blockIdx.x * kBlock + threadIdx.x, cy = blockIdx.y;
    if (cx >= cw)
        return;
    const int xa = min(2 * cx, ow - 1), xb = min(2 * cx + 1, ow - 1);
    const int ya = min(2 * cy, oh - 1), yb = min(2 * cy + 1, oh - 1);
    for (int c = 0; c < 2; ++c) {
        float *o = c ? ov : ou;
        const float *s = c ? v : u;
        const RawOut &P = b.p[c];
        const float p00 = o[(size_t)ya * ow + xa], p10 = o[(size_t)ya * ow + xb];
        const float p01 = o[(size_t)yb * ow + xa], p11 = o[(size_t)yb * ow + xb];
        const float r = s[(size_t)cy * cw + cx] - 0.25f * (p00 + p10 + p01 + p11);
        const float d = bp * r;
        const float q00 = p00 + d, q10 = p10 + d, q01 = p01 + d, q11 = p11 + d;
        if (keep_float) {
            o[(size_t)ya * ow + xa] = q00;
            o[(size_t)ya * ow + xb] = q10;
            o[(size_t)yb * ow + xa] = q01;
            o[(size_t)yb * ow + xb] = q11;
        }
        ptrdiff_t stride;
        char *base = plane_out(P, stride);
        T *ra = reinterpret_cast<T *>(base + stride * ya);
        T *rb = reinterpret_cast<T *>(base + stride * yb);
        if constexpr (sizeof(T) == 4) {
            ra[xa] = q00; ra[xb] = q10; rb[xa] = q01; rb[xb] = q11;
        } else {
            ra[xa] = pack_sample<T>(q00, P); ra[xb] = pack_sample<T>(q10, P);
            rb[xa] = pack_sample<T>(q01, P); rb[xb] = pack_sample<T>(q11, P);
        }
    }
}

// ---------------------------------------------------------------------------
// structure tensor + fused trust seed (sobelTensor3x3, maps.cpp:64-168)
//
// One wave per 64 x TH strip (TH tuned per clip, 16 by default).  Lane L
// owns column x0 + L and walks down the strip: the three luma rows of the
// Sobel stencil (at clamped centres and taps, the reference's border rule)
// and the last three rows of gradient products stay in registers, so each
// luma row is read once and each product formed once; the column sums of
// the 3x3 box come from the neighbouring lanes (DPP wave shifts), and lanes
// 0 and 63 also form the strip's halo columns x0 - 1 and x0 + 64.  Every
// sum keeps the reference's operand order.
// ---------------------------------------------------------------------------

constexpr int kSobelW = 64;

// v of the lane to the left (wave_shr:1) / right (wave_shl:1); lanes 0 and
// 63 get `edge` instead
__device__ __forceinline__ float from_left(float v, float edge, int lane) {
    const float l = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
    return lane == 0 ? edge : l;
}
__device__ __forceinline__ float from_right(float v, float edge, int lane) {
    const float r = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
    return lane == kSobelW - 1 ? edge : r;
}

// One strip.  EDGE: the strip touches the image border, where the stencil
// rows clamp and the box sums of column 0 / w - 1 repeat the edge column;
// otherwise every stencil row is inside and the strip simply rolls down one
// luma row per step.  Its own instantiation keeps the border logic out of
// the interior strips.  The row loop is unrolled, so the rolling rows and
// products rotate through registers without copies.  Luma is read through
// a buffer resource: a wave-uniform row offset plus the per-lane column
// offsets, which are fixed for the strip.
template <int TH, bool EDGE>
__device__ __forceinline__ void sobel_strip(const float *__restrict__ y, int w, int h, float4 *yt, uint64_t *seed,
                                            int w64, float thr, int x0, int y0, int word) {
    const int lane = threadIdx.x;
    const int x = x0 + lane;
    // clamped stencil columns (byte offsets) of the own column and of the
    // halo column (x0 - 1 for lane 0, x0 + 64 for lane 63; unused elsewhere)
    const int cx = clampi(x, 0, w - 1), hc = clampi(lane == 0 ? x0 - 1 : x0 + kSobelW, 0, w - 1);
    const uint32_t oxl = 4u * clampi(cx - 1, 0, w - 1), oxc = 4u * cx, oxr = 4u * clampi(cx + 1, 0, w - 1);
    const uint32_t ohl = 4u * clampi(hc - 1, 0, w - 1), ohc = 4u * hc, ohr = 4u * clampi(hc + 1, 0, w - 1);
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(y), 0, (int)(4u * (uint32_t)w * (uint32_t)h), 0x00020000);
    struct Row {
        float l, c, r, hl, hc, hr;
    };
    auto load_row = [&](int r) {
        const int ro = (int)(4u * (uint32_t)r * (uint32_t)w);
        auto ld = [&](uint32_t off) { return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc, off, ro, 0)); };
        return Row{ld(oxl), ld(oxc), ld(oxr), ld(ohl), ld(ohc), ld(ohr)};
    };
    // rows (iu, ic, id) = (clamp(cy - 1), cy, clamp(cy + 1)) of the current
    // gradient centre cy
    int iu = -1, ic = -1, id = -1;
    Row ru{}, rc{}, rd{};
    if constexpr (!EDGE) {
        rc = load_row(y0 - 2);
        rd = load_row(y0 - 1);
    }
    // gradient products of the last three gradient rows ([0] the oldest)
    float pxx[3] = {}, pxy[3] = {}, pyy[3] = {}, hxx[3] = {}, hxy[3] = {}, hyy[3] = {};
    float yc_prev = 0.0f;  // luma at (x, previous gradient centre)
    const float ninth = 1.0f / 9.0f;
    _Pragma("unroll") for (int k = y0 - 1, step = 0; step < TH + 2; ++k, ++step) {
        const float y_out = yc_prev;  // luma of output row k - 1
        if constexpr (EDGE) {
            const int cy = clampi(k, 0, h - 1);
            const int yu = clampi(cy - 1, 0, h - 1), yd = clampi(cy + 1, 0, h - 1);
            if (yu == ic && cy == id) {
                ru = rc;
                rc = rd;
                rd = load_row(yd);
            } else if (!(yu == iu && cy == ic && yd == id)) {
                ru = load_row(yu);
                rc = load_row(cy);
                rd = load_row(yd);
            }
            iu = yu;
            ic = cy;
            id = yd;
        } else {
            ru = rc;
            rc = rd;
            rd = load_row(k + 1);
        }
        yc_prev = rc.c;
        // Sobel / 8 at the clamped centre
        const float gx = (ru.r + 2.0f * rc.r + rd.r - ru.l - 2.0f * rc.l - rd.l) * 0.125f;
        const float gy = (rd.l + 2.0f * rd.c + rd.r - ru.l - 2.0f * ru.c - ru.r) * 0.125f;
        const float hgx = (ru.hr + 2.0f * rc.hr + rd.hr - ru.hl - 2.0f * rc.hl - rd.hl) * 0.125f;
        const float hgy = (rd.hl + 2.0f * rd.hc + rd.hr - ru.hl - 2.0f * ru.hc - ru.hr) * 0.125f;
        pxx[0] = pxx[1]; pxx[1] = pxx[2]; pxx[2] = gx * gx;
        pxy[0] = pxy[1]; pxy[1] = pxy[2]; pxy[2] = gx * gy;
        pyy[0] = pyy[1]; pyy[1] = pyy[2]; pyy[2] = gy * gy;
        hxx[0] = hxx[1]; hxx[1] = hxx[2]; hxx[2] = hgx * hgx;
        hxy[0] = hxy[1]; hxy[1] = hxy[2]; hxy[2] = hgx * hgy;
        hyy[0] = hyy[1]; hyy[1] = hyy[2]; hyy[2] = hgy * hgy;
        if (step < 2)
            continue;
        const int r = k - 1;  // output row: gradient rows r - 1, r, r + 1
        if (EDGE && r >= h)
            break;  // uniform
        // vertical 3-sums of the products at this column and the halo
        const float vxx = pxx[0] + pxx[1] + pxx[2], vxy = pxy[0] + pxy[1] + pxy[2], vyy = pyy[0] + pyy[1] + pyy[2];
        const float wxx = hxx[0] + hxx[1] + hxx[2], wxy = hxy[0] + hxy[1] + hxy[2], wyy = hyy[0] + hyy[1] + hyy[2];
        const float lxx = from_left(vxx, wxx, lane), lxy = from_left(vxy, wxy, lane), lyy = from_left(vyy, wyy, lane);
        const float rxx = from_right(vxx, wxx, lane), rxy = from_right(vxy, wxy, lane),
                    ryy = from_right(vyy, wyy, lane);
        float oxx = ninth * ((lxx + vxx) + rxx);
        float oxy = ninth * ((lxy + vxy) + rxy);
        float oyy = ninth * ((lyy + vyy) + ryy);
        if constexpr (EDGE) {
            if (x == 0) {
                // the box repeats the edge column
                const float nxx = w > 1 ? rxx : vxx, nxy = w > 1 ? rxy : vxy, nyy = w > 1 ? ryy : vyy;
                oxx = (2.0f * vxx + nxx) / 9.0f;
                oxy = (2.0f * vxy + nxy) / 9.0f;
                oyy = (2.0f * vyy + nyy) / 9.0f;
            } else if (x == w - 1) {
                oxx = (lxx + 2.0f * vxx) / 9.0f;
                oxy = (lxy + 2.0f * vxy) / 9.0f;
                oyy = (lyy + 2.0f * vyy) / 9.0f;
            }
        }
        const bool inside = x < w;
        // trust seed (fused in buildGuideMaps, maps.cpp:252-276) as one bit
        // per pixel: the wave's ballot is the row's 64-pixel word
        const uint64_t bits = __ballot(inside && oxx + oyy > thr);
        if (seed && lane == 0)
            seed[(size_t)r * w64 + word] = bits;
        if (inside)
            yt[(uint32_t)r * (uint32_t)w + (uint32_t)x] = make_float4(y_out, oxx, oxy, oyy);
    }
}

template <int TH, bool GROUP>
__global__ __launch_bounds__(kSobelW) void k_sobel_tensor(const float *__restrict__ y, int w, int h, float4 *yt,
                                                         uint64_t *seed, int w64, float thr, GroupOff go) {
    if constexpr (GROUP) {
        const long long gd = group_off(go);
        y = gshift(y, gd);
        yt = gshift(yt, gd);
        seed = gshift(seed, gd);
    }
    unsigned bx, by;
    xcd_block2(bx, by);
    const int x0 = (int)bx * kSobelW, y0 = (int)by * TH;
    // interior: stencil rows y0 - 2 .. y0 + TH + 1 inside, neither border
    // column in the strip
    if (x0 == 0 || x0 + kSobelW >= w || y0 < 2 || y0 + TH + 1 > h - 1)
        sobel_strip<TH, true>(y, w, h, yt, seed, w64, thr, x0, y0, (int)bx);
    else
        sobel_strip<TH, false>(y, w, h, yt, seed, w64, thr, x0, y0, (int)bx);
}

// ---------------------------------------------------------------------------
// footprint luma map (buildLcMap, maps.cpp:425-499); with luv it also writes
// the interleaved chroma guide {lc, U, V, 0} the guided pass gathers from
// ---------------------------------------------------------------------------

template <bool GROUP>
__global__ __launch_bounds__(kBlock) void k_lc_map(const float *__restrict__ y, int w, float *lc, int cw,
                                                   int ch, const int *xb, const int *xe, const int *yb,
                                                   const int *ye, const float *u, const float *v, float4 *luv,
                                                   GroupOff go) {
    if constexpr (GROUP) {
        const long long gd = group_off(go);
        y = gshift(y, gd);
        lc = gshift(lc, gd);
        u = gshift(u, gd);
        v = gshift(v, gd);
        luv = gshift(luv, gd);
    }
    const int cx = blockIdx.x * kBlock + threadIdx.x, cy = blockIdx.y;
    if (cx >= cw || cy >= ch)
        return;
    const int x0 = xb[cx], x1 = xe[cx], y0 = yb[cy], y1 = ye[cy];
    double sum = 0.0;
    for (int j = y0; j < y1; ++j)
        for (int i = x0; i < x1; ++i)
            sum += y[(size_t)j * w + i];
    const size_t i = (size_t)cy * cw + cx;
    const float l = (float)(sum / ((double)(x1 - x0) * (y1 - y0)));
    lc[i] = l;
    if (luv)
        luv[i] = make_float4(l, u[i], v[i], 0.0f);
}

// ---------------------------------------------------------------------------
// trust mask and sparse workset (buildTrustMask, maps.cpp:280-340;
// buildSparseWorkset, maps.cpp:342-422) on bit planes: row-major list of
// active output pixels via per-row counts, a per-row prefix and an ordered
// per-row compaction (wave ballots), then the chroma mask of the bilinear
// footprints and its population.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int block_sum(int v, int *scratch) {
    // wave64 reduction then across the 4 waves of the block
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        scratch[wave] = v;
    __syncthreads();
    int t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; ++i)
            t += scratch[i];
    __syncthreads();
    return t;  // valid in thread 0
}

// ---------------------------------------------------------------------------
// trust mask on bit planes: one bit per luma pixel, 64 per word, rows of
// w64 words.  Horizontal pass: each output word ORs the row shifted by every
// offset in [-r, r]; vertical pass: OR over rows y-r..y+r, with the row
// population when the mask has the output size.
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t row_bits(const uint64_t *row, int w64, long pos) {
    // 64 bits starting at bit `pos` of the row; zeros outside it
    const long q = pos >> 6;
    const int sft = (int)(pos & 63);
    const uint64_t lo = (q >= 0 && q < w64) ? row[q] : 0ull;
    if (sft == 0)
        return lo;
    const uint64_t hi = (q + 1 >= 0 && q + 1 < w64) ? row[q + 1] : 0ull;
    return (lo >> sft) | (hi << (64 - sft));
}

__global__ __launch_bounds__(kBlock) void k_dilate_hbits(const uint64_t *in, uint64_t *out, int w, int w64, int h,
                                                         int r, uint8_t *zero8, size_t n_zero8, uint32_t *counters) {
    const int wx = blockIdx.x * kBlock + threadIdx.x, y = blockIdx.y;
    // clear the chroma mask and the counters the workset fills next
    const size_t tid = ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * kBlock + threadIdx.x;
    const size_t nthreads = (size_t)gridDim.x * gridDim.y * kBlock;
    for (size_t i = tid; i < n_zero8; i += nthreads)
        zero8[i] = 0;
    if (tid < 8)
        counters[tid] = 0;
    if (wx >= w64)
        return;
    const uint64_t *row = in + (size_t)y * w64;
    uint64_t acc = 0;
    for (int d = -r; d <= r; ++d)
        acc |= row_bits(row, w64, (long)wx * 64 + d);
    const int valid = w - wx * 64;   // pixels of this word inside the row
    if (valid < 64)
        acc &= (1ull << valid) - 1ull;
    out[(size_t)y * w64 + wx] = acc;
}

__global__ __launch_bounds__(64) void k_dilate_vbits(const uint64_t *in, uint64_t *out, int w64, int h, int r,
                                                     uint32_t *row_count) {
    const int y = blockIdx.x;
    const int j0 = max(0, y - r), j1 = min(h - 1, y + r);
    int cnt = 0;
    for (int wx = threadIdx.x; wx < w64; wx += 64) {
        uint64_t acc = 0;
        for (int j = j0; j <= j1; ++j)
            acc |= in[(size_t)j * w64 + wx];
        out[(size_t)y * w64 + wx] = acc;
        cnt += __popcll(acc);
    }
    if (row_count) {
        for (int o = 32; o > 0; o >>= 1)
            cnt += __shfl_xor(cnt, o, 64);
        if (threadIdx.x == 0)
            row_count[y] = (uint32_t)cnt;
    }
}

// Trust mask without the index list, for the tiled guided pass (it reads
// the mask bits) at output size == mask size, in one kernel.  Block b owns
// a band of chroma rows and the mask rows band[b].x .. band[b].y - 1 (the
// host partitions the mask rows); it stages the seed rows its mask rows
// band[b].z .. band[b].w read (r above and below, zero outside the image),
// dilates them in LDS, stores and counts its own mask rows, and marks its
// chroma rows from the LDS rows.  Each block writes
// its two counts to its own partial slot: a shared counter would serialise
// the atomics of every block across the XCDs.
//
// Horizontal dilation for radius < 64: a word and its two neighbours hold
// every bit within the radius, so the OR of the 2r+1 shifted rows is formed
// in registers.
template <bool GROUP>
__global__ __launch_bounds__(kBlock) void k_trust_fused(const uint64_t *seed, uint64_t *mask, int w, int w64, int h,
                                                        int r, const int4 *band, int cband, int cw, int ch,
                                                        const int2 *xr, const int2 *yr, uint8_t *cmask,
                                                        uint32_t *counters, GroupOff go) {
    if constexpr (GROUP) {
        const long long gd = group_off(go);
        seed = gshift(seed, gd);
        mask = gshift(mask, gd);
        cmask = gshift(cmask, gd);
        counters = gshift(counters, gd);
    }
    extern __shared__ uint64_t tf_lds[];
    __shared__ int scratch[kBlock / 64];
    __shared__ int2 ys_band[kMaxCband];
    const int4 bd = band[blockIdx.x];
    const int rows = bd.w - bd.z + 1, srows = rows + 2 * r, ws = w64 + 2;
    uint64_t *raw = tf_lds;                 // srows x ws seed rows, zero outside the image
    uint64_t *vd = raw + srows * ws;        // rows x ws vertical ORs, one zero word each side
    uint64_t *mk = vd + rows * ws;          // rows x w64 (mask rows bd.z ..)
    int2 *xs_l = reinterpret_cast<int2 *>(mk + rows * w64);  // cw column ranges
    const int s0 = bd.z - r;
    const int cy0 = blockIdx.x * cband, ncy = min(cband, ch - cy0);
    for (int i = threadIdx.x; i < cw; i += kBlock)
        xs_l[i] = xr[i];
    if (threadIdx.x < ncy)
        ys_band[threadIdx.x] = yr[cy0 + threadIdx.x];
    if (blockIdx.x == 0 && threadIdx.x < kCntChromaWords)
        counters[threadIdx.x] = 0;  // per-frame counters of the other passes
    for (int i = threadIdx.x; i < srows * ws; i += kBlock) {
        const int k = i / ws, c = i - k * ws, sy = s0 + k, wx = c - 1;
        raw[i] = (sy >= 0 && sy < h && wx >= 0 && wx < w64) ? seed[(uint32_t)(sy * w64 + wx)] : 0ull;
    }
    __syncthreads();
    // the square dilation is separable in either order: vertical OR of the
    // 2r+1 seed rows first (only the band's rows), then the horizontal pass
    for (int i = threadIdx.x; i < rows * ws; i += kBlock) {
        const int k = i / ws;
        uint64_t acc = 0;
        for (int j = 0; j <= 2 * r; ++j)
            acc |= raw[(k + j) * ws + (i - k * ws)];
        vd[i] = acc;
    }
    __syncthreads();
    int act = 0;
    for (int i = threadIdx.x; i < rows * w64; i += kBlock) {
        const int k = i / w64, wx = i - k * w64, y = bd.z + k;
        const uint64_t *q = vd + k * ws + wx;
        const uint64_t lo = q[0], mid = q[1], hi = q[2];
        uint64_t acc = mid;
        for (int d = 1; d <= r; ++d)
            acc |= (mid << d) | (lo >> (64 - d)) | (mid >> d) | (hi << (64 - d));
        const int valid = w - wx * 64;  // pixels of this word inside the row
        if (valid < 64)
            acc &= (1ull << valid) - 1ull;
        mk[i] = acc;
        if (y >= bd.x && y < bd.y) {
            mask[(uint32_t)(y * w64 + wx)] = acc;
            act += __popcll(acc);
        }
    }
    __syncthreads();
    // chroma marks (buildSparseWorkset, maps.cpp:342-360): sample (cx, cy)
    // is marked when an active output pixel has it in its 2x2 bilinear
    // footprint; xr / yr give, per chroma column / row, the contiguous range
    // of output columns / rows whose footprint contains it
    int marked = 0;
    for (int i = threadIdx.x; i < ncy * cw; i += kBlock) {
        const int k = i / cw, cx = i - k * cw, cy = cy0 + k;
        const int2 xs = xs_l[cx], ys = ys_band[k];
        uint64_t any = 0;
        if (xs.x <= xs.y && ys.x <= ys.y) {
            const int wa = xs.x >> 6, wb = xs.y >> 6;
            const uint64_t ma = ~0ull << (xs.x & 63);
            const uint64_t mb = (xs.y & 63) == 63 ? ~0ull : (1ull << ((xs.y & 63) + 1)) - 1ull;
            for (int y = ys.x; y <= ys.y; ++y) {
                const uint64_t *row = mk + (y - bd.z) * w64;
                if (wa == wb) {
                    any |= row[wa] & ma & mb;
                } else {
                    any |= row[wa] & ma;
                    for (int wd = wa + 1; wd < wb; ++wd)
                        any |= row[wd];
                    any |= row[wb] & mb;
                }
            }
        }
        const int hit = any != 0ull;
        cmask[(uint32_t)(cy * cw + cx)] = (uint8_t)hit;
        marked += hit;
    }
    const int ta = block_sum(act, scratch);
    const int tm = block_sum(marked, scratch);
    if (threadIdx.x == 0) {
        counters[kCntChromaWords + blockIdx.x] = (uint32_t)tm;
        counters[kCntChromaWords + kMaxBands + blockIdx.x] = (uint32_t)ta;
    }
}

// Output-pixel view of a bit mask at mask resolution (nearest mapping of
// buildSparseWorkset, maps.cpp:342-360).
struct MaskBits { const uint64_t *bits; int mw, mh, w64, ow, oh; };

__device__ __forceinline__ bool out_active_bits(const MaskBits &m, int x, int y) {
    const int mx = min(m.mw - 1, (int)((int64_t)x * m.mw / m.ow));
    const int my = min(m.mh - 1, (int)((int64_t)y * m.mh / m.oh));
    return (m.bits[(size_t)my * m.w64 + (mx >> 6)] >> (mx & 63)) & 1ull;
}

__global__ __launch_bounds__(kBlock) void k_row_count_bits(MaskBits m, uint32_t *row_count) {
    __shared__ int scratch[kBlock / 64];
    const int y = blockIdx.x;
    int c = 0;
    for (int x = threadIdx.x; x < m.ow; x += kBlock)
        c += out_active_bits(m, x, y);
    const int t = block_sum(c, scratch);
    if (threadIdx.x == 0)
        row_count[y] = (uint32_t)t;
}

// One block per output row: exclusive prefix of the row counts (summed
// here, so no separate scan pass), ordered ballot compaction of the row's
// active pixels, and the 2x2 bilinear chroma footprints of each.
__global__ __launch_bounds__(kBlock) void k_compact_bits(MaskBits m, const uint32_t *row_count, uint32_t *off,
                                                         uint32_t *indices, const int *cx_i0, const int *cy_i0,
                                                         uint8_t *cmask, int cw, int ch) {
    __shared__ int scratch[kBlock / 64];
    __shared__ int wave_cnt[kBlock / 64];
    const int y = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int part = 0;
    for (int i = threadIdx.x; i < y; i += kBlock)
        part += (int)row_count[i];
    const int prefix = block_sum(part, scratch);
    __shared__ uint32_t base_s;
    if (threadIdx.x == 0) {
        base_s = (uint32_t)prefix;
        off[y] = (uint32_t)prefix;
        if (y == m.oh - 1)
            off[m.oh] = (uint32_t)prefix + row_count[y];
    }
    __syncthreads();
    uint32_t base = base_s;
    const int cy0 = cy_i0[y], cy1 = min(cy0 + 1, ch - 1);
    for (int x0 = 0; x0 < m.ow; x0 += kBlock) {
        const int x = x0 + threadIdx.x;
        const bool a = x < m.ow && out_active_bits(m, x, y);
        const uint64_t bal = __ballot(a);
        const int before = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0)
            wave_cnt[wave] = __popcll(bal);
        __syncthreads();
        int prior = 0, total = 0;
        for (int i = 0; i < kBlock / 64; ++i) {
            prior += i < wave ? wave_cnt[i] : 0;
            total += wave_cnt[i];
        }
        if (a) {
            indices[base + prior + before] = (uint32_t)((size_t)y * m.ow + x);
            const int cx0 = cx_i0[x], cx1 = min(cx0 + 1, cw - 1);
            cmask[(size_t)cy0 * cw + cx0] = 1;
            cmask[(size_t)cy0 * cw + cx1] = 1;
            cmask[(size_t)cy1 * cw + cx0] = 1;
            cmask[(size_t)cy1 * cw + cx1] = 1;
        }
        base += total;
        __syncthreads();
    }
}

// population of a 0/1 byte plane, 16 bytes per thread step
__global__ __launch_bounds__(kBlock) void k_count_u8x16(const uint8_t *p, size_t n, uint32_t *counter) {
    __shared__ int scratch[kBlock / 64];
    int c = 0;
    const size_t n16 = n / 16;
    const uint4 *q = reinterpret_cast<const uint4 *>(p);
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n16; i += (size_t)gridDim.x * kBlock) {
        const uint4 v = q[i];
        // bytes are 0 or 1: the byte sums fit in each byte lane
        const uint32_t s4 = v.x + v.y + v.z + v.w;
        c += (int)((s4 & 0xff) + ((s4 >> 8) & 0xff) + ((s4 >> 16) & 0xff) + (s4 >> 24));
    }
    if (blockIdx.x == 0)
        for (size_t i = n16 * 16 + threadIdx.x; i < n; i += kBlock)
            c += p[i] != 0;
    const int t = block_sum(c, scratch);
    if (threadIdx.x == 0 && t)
        atomicAdd(counter, (uint32_t)t);
}

__global__ __launch_bounds__(kBlock) void k_fill_u8(uint8_t *p, size_t n, uint8_t v) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        p[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_count_u8(const uint8_t *p, size_t n, uint32_t *counter) {
    __shared__ int scratch[kBlock / 64];
    int c = 0;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        c += p[i] != 0;
    const int t = block_sum(c, scratch);
    if (threadIdx.x == 0 && t)
        atomicAdd(counter, (uint32_t)t);
}

__global__ void k_zero_u32(uint32_t *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0;
}

// ---------------------------------------------------------------------------
// mutual-structure gate (buildMutualGate, maps.cpp:831-948)
// ---------------------------------------------------------------------------

// Mutual-structure gate with the gradient pass fused in: a block owns a
// 32x8 tile of chroma samples, stages lc/U/V for the tile plus a 5-sample
// apron in LDS, computes the three Sobel gradients for the tile plus a
// 4-sample apron (every sample the 3x3 tensor and the seven projected
// bilinear samples can touch, as |k n| <= 3 + rounding) and evaluates the
// gate from there.  LDS entries are addressed by clamped image coordinates,
// so each lookup returns exactly the value k_grad3 + k_gate read.
constexpr int kGateHalo = 4;

// 3x3 Sobel / 8 (maps.cpp:9-20) from an LDS tile; the caller passes clamped
// neighbour indices.
template <int SW>
__device__ __forceinline__ void sobel_lds(const float (*p)[SW], int xm, int xc, int xp, int ym, int yc, int yp,
                                          float &gx, float &gy) {
    const float tl = p[ym][xm], t = p[ym][xc], tr = p[ym][xp];
    const float l = p[yc][xm], r = p[yc][xp];
    const float bl = p[yp][xm], b = p[yp][xc], br = p[yp][xp];
    gx = (tr + 2.0f * r + br - tl - 2.0f * l - bl) * 0.125f;
    gy = (bl + 2.0f * b + br - tl - 2.0f * t - tr) * 0.125f;
}

// The mutual gate of one chroma sample (x, y) from the block's LDS gradients
// (buildMutualGate's processPixel, maps.cpp:853-905).
template <int kGateGW, int kGateGH>
__device__ __forceinline__ void gate_sample(int x, int y, int w, int h, int gx0, int gy0,
                                            const float4 (*gyu)[kGateGW], const float2 (*gvv)[kGateGW],
                                            float *gate) {
    float *out = gate + (size_t)y * w + x;
    // LDS column / row of a clamped image coordinate (bounded for any input)
    auto cx = [&](int X) { return clampi(clampi(X, 0, w - 1) - gx0, 0, kGateGW - 1); };
    auto cy = [&](int Y) { return clampi(clampi(Y, 0, h - 1) - gy0, 0, kGateGH - 1); };
    float sxx = 0, sxy = 0, syy = 0;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const int r = cy(y + dy), c = cx(x + dx);
            const float4 g = gyu[r][c];
            const float a = g.x;
            const float b = g.y;
            sxx += a * a;
            sxy += a * b;
            syy += b * b;
        }
    if (sxx + syy < 1e-10f) {
        *out = 0.0f;
        return;
    }
    const Dir d = principal_dir(sxx, sxy, syy);
    const float nx = d.nx, ny = d.ny;
    float sumY = 0, sqY = 0, momY = 0, sumC = 0, sqC = 0, momC = 0;
    for (int k = -3; k <= 3; ++k) {
        const double px = (float)x + (float)k * nx;
        const double py = (float)y + (float)k * ny;
        const int rx = (int)floor(px), ry = (int)floor(py);
        const int c0 = cx(rx), c1 = cx(rx + 1), r0 = cy(ry), r1 = cy(ry + 1);
        const float fx = (float)(px - rx), fy = (float)(py - ry);
        const float4 p00 = gyu[r0][c0], p01 = gyu[r0][c1], p10 = gyu[r1][c0], p11 = gyu[r1][c1];
        const float2 q00 = gvv[r0][c0], q01 = gvv[r0][c1], q10 = gvv[r1][c0], q11 = gvv[r1][c1];
        const float gxs[3][4] = {{p00.x, p01.x, p10.x, p11.x}, {p00.z, p01.z, p10.z, p11.z},
                                 {q00.x, q01.x, q10.x, q11.x}};
        const float gys[3][4] = {{p00.y, p01.y, p10.y, p11.y}, {p00.w, p01.w, p10.w, p11.w},
                                 {q00.y, q01.y, q10.y, q11.y}};
        float res[3];
        _Pragma("unroll") for (int q = 0; q < 3; ++q) {
            const float a = gxs[q][0] * nx + gys[q][0] * ny;
            const float b = gxs[q][1] * nx + gys[q][1] * ny;
            const float c = gxs[q][2] * nx + gys[q][2] * ny;
            const float dd = gxs[q][3] * nx + gys[q][3] * ny;
            res[q] = a + (b - a) * fx + (c - a) * fy + (a - b - c + dd) * fx * fy;
        }
        const float gY = fabsf(res[0]);
        const float au = res[1];
        const float av = res[2];
        const float gC = sqrtf(au * au + av * av);
        sumY += gY; sqY += gY * gY; momY += (float)k * gY;
        sumC += gC; sqC += gC * gC; momC += (float)k * gC;
    }
    if (sumY < 1e-9f || sumC < 1e-9f) {
        *out = 0.0f;
        return;
    }
    const float partY = sumY * sumY / (sqY + 1e-12f);
    const float partC = sumC * sumC / (sqC + 1e-12f);
    const float ratio = partC / (partY + 1e-6f);
    const float dcent = fabsf(momY / sumY - momC / sumC);
    const float tw = fclamp((1.6f - ratio) / 0.4f, 0.0f, 1.0f);
    const float tp = fclamp((1.0f - dcent) / 0.5f, 0.0f, 1.0f);
    *out = (tw * tw * (3.0f - 2.0f * tw)) * (tp * tp * (3.0f - 2.0f * tp));
}

template <int kGateTW, int kGateTH, bool GROUP>
__global__ __launch_bounds__(kBlock) void k_gate_fused(const float *__restrict__ lc, const float *__restrict__ u,
                                                       const float *__restrict__ v, int w, int h,
                                                       const uint8_t *cmask, const uint32_t *count, int ncount,
                                                       float *gate, GroupOff go) {
    static_assert(kGateTW * kGateTH == kBlock, "one thread per tile sample");
    if constexpr (GROUP) {
        const long long gd = group_off(go);
        lc = gshift(lc, gd);
        u = gshift(u, gd);
        v = gshift(v, gd);
        cmask = gshift(cmask, gd);
        count = gshift(count, gd);
        gate = gshift(gate, gd);
    }
    constexpr int kGateGW = kGateTW + 2 * kGateHalo, kGateGH = kGateTH + 2 * kGateHalo;
    constexpr int kGateSW = kGateGW + 2, kGateSH = kGateGH + 2;
    __shared__ float src[3][kGateSH][kGateSW];
    // gradients per sample, interleaved so a bilinear corner is one 16-byte
    // and one 8-byte read: {gx, gy} of lc and U, then of V
    __shared__ float4 gyu[kGateGH][kGateGW];
    __shared__ float2 gvv[kGateGH][kGateGW];
    __shared__ int s_dense;
    unsigned bx, by;
    // hardware (round-robin) block order: the work per block follows the
    // chroma mask, and round-robin spreads it evenly over the XCDs
    bx = blockIdx.x;
    by = blockIdx.y;
    const int tx0 = (int)bx * kGateTW, ty0 = (int)by * kGateTH;
    const int gx0 = tx0 - kGateHalo, gy0 = ty0 - kGateHalo, sx0 = gx0 - 1, sy0 = gy0 - 1;
    const int x = tx0 + (int)(threadIdx.x % kGateTW), y = ty0 + (int)(threadIdx.x / kGateTW);
    const size_t n = (size_t)w * h;
    // dense unless at most half the chroma samples are marked (the marked
    // count is the sum of count[0..ncount))
    if (threadIdx.x < 64) {
        uint32_t c = 0u;
        if (cmask) {
            _Pragma("unroll 8") for (int i = threadIdx.x; i < ncount; i += 64)
                c += count[i];
        }
        for (int o = 32; o > 0; o >>= 1)
            c += __shfl_xor(c, o, 64);
        if (threadIdx.x == 0)
            s_dense = !cmask || (uint64_t)c * 2 >= (uint64_t)n;
    }
    __syncthreads();
    const bool dense = s_dense != 0;
    const bool inside = x < w && y < h;
    const bool need = inside && (dense || cmask[(size_t)y * w + x] != 0);
    if (!__syncthreads_or(need)) {
        if (inside)
            gate[(size_t)y * w + x] = 0.0f;
        return;
    }
    for (int e = threadIdx.x; e < kGateSH * kGateSW; e += kBlock) {
        const int r = e / kGateSW, c = e - r * kGateSW;
        const size_t k = (size_t)clampi(sy0 + r, 0, h - 1) * w + clampi(sx0 + c, 0, w - 1);
        src[0][r][c] = lc[k];
        src[1][r][c] = u[k];
        src[2][r][c] = v[k];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kGateGH * kGateGW; e += kBlock) {
        const int r = e / kGateGW, c = e - r * kGateGW;
        const int P = clampi(gx0 + c, 0, w - 1), Q = clampi(gy0 + r, 0, h - 1);
        const int xm = clampi(P - 1, 0, w - 1) - sx0, xc = P - sx0, xp = clampi(P + 1, 0, w - 1) - sx0;
        const int ym = clampi(Q - 1, 0, h - 1) - sy0, yc = Q - sy0, yp = clampi(Q + 1, 0, h - 1) - sy0;
        float g[6];
        _Pragma("unroll") for (int q = 0; q < 3; ++q)
            sobel_lds<kGateSW>(src[q], xm, xc, xp, ym, yc, yp, g[2 * q], g[2 * q + 1]);
        gyu[r][c] = make_float4(g[0], g[1], g[2], g[3]);
        gvv[r][c] = make_float2(g[4], g[5]);
    }
    // the block's needed samples, gathered in row-major order into the
    // first lanes, so the gate below runs on full waves
    __shared__ uint16_t s_list[kBlock];
    __shared__ int s_wcount[kBlock / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t nb = __ballot(need);
    if (lane == 0)
        s_wcount[wv] = __popcll(nb);
    if (inside && !need)
        gate[(size_t)y * w + x] = 0.0f;
    __syncthreads();
    int base = 0, n_need = 0;
    _Pragma("unroll") for (int k = 0; k < kBlock / 64; ++k) {
        base += k < wv ? s_wcount[k] : 0;
        n_need += s_wcount[k];
    }
    if (need)
        s_list[base + __popcll(nb & ((1ull << lane) - 1ull))] = (uint16_t)threadIdx.x;
    __syncthreads();
    if ((int)threadIdx.x >= n_need)
        return;
    {
        const int p = s_list[threadIdx.x];
        gate_sample<kGateGW, kGateGH>(tx0 + p % kGateTW, ty0 + p / kGateTW, w, h, gx0, gy0, gyu, gvv, gate);
    }
}


// Per-output-pixel axis data: first tap, positions, bilinear indices, the
// tap weights and active-tap flags.  6-tap axes read everything from one
// 64-byte record (AxisDev::rec) and keep it in registers; other supports
// read the weights from the tables.
template <int SUP> struct PixAxis {
    int start, lb_i0, cb_i0;
    float pos, lpos, lb_f, cb_f;
    const float *w, *am;
    __device__ __forceinline__ bool on(int i) const { return am[i] != 0.0f; }
    __device__ __forceinline__ float wt(int i) const { return w[i]; }
};
template <> struct PixAxis<6> {
    int start, lb_i0, cb_i0;
    float pos, lpos, lb_f, cb_f;
    uint32_t bits;
    float w[6];
    float wsum, wabs;  // sum and absolute sum of the active weights
    __device__ __forceinline__ bool on(int i) const { return (bits >> i) & 1u; }
    __device__ __forceinline__ float wt(int i) const { return w[i]; }
};

template <int SUP> __device__ __forceinline__ PixAxis<SUP> pix_axis(const AxisDev &a, int o) {
    PixAxis<SUP> p;
    if constexpr (SUP == 6) {
        const float4 *r = a.rec + 4 * (size_t)o;
        const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
        p.start = __float_as_int(r0.x);
        p.pos = r0.y;
        p.lpos = r0.z;
        p.bits = __float_as_uint(r0.w);
        p.lb_i0 = __float_as_int(r1.x);
        p.lb_f = r1.y;
        p.cb_i0 = __float_as_int(r1.z);
        p.cb_f = r1.w;
        p.w[0] = r2.x; p.w[1] = r2.y; p.w[2] = r2.z; p.w[3] = r2.w;
        p.w[4] = r3.x; p.w[5] = r3.y;
        p.wsum = r3.z;
        p.wabs = r3.w;
    } else {
        p.start = a.start[o];
        p.pos = a.pos[o];
        p.lpos = a.lpos[o];
        p.lb_i0 = a.lb_i0[o];
        p.lb_f = a.lb_f[o];
        p.cb_i0 = a.cb_i0[o];
        p.cb_f = a.cb_f[o];
        p.w = a.w + (size_t)o * a.sup;
        p.am = a.am + (size_t)o * a.sup;
    }
    return p;
}

// ---------------------------------------------------------------------------
// plain chroma base (plainChromaSeparable, recon.cpp:1151-1461): horizontal
// pass per source chroma row with 1D hulls, vertical pass per output pixel
// with the separable normalisation, cancellation fallback and 2D hull clamp.
// ---------------------------------------------------------------------------

// SUP == 6: the 6 taps are straight-line code with every load issued up
// front (clamped indices are always in bounds); an inactive tap adds an
// exact zero and leaves the hull alone, so the sums equal the skipping loop.
template <int SUP>
__global__ __launch_bounds__(kBlock) void k_plain_h(const float *__restrict__ u, const float *__restrict__ v,
                                                    int cw, int ch, AxisDev ax, float *hb) {
    unsigned bx, by;
    xcd_block2(bx, by);
    const int ox = bx * kBlock + threadIdx.x, sy = by;
    if (ox >= ax.n)
        return;
    const float *su = u + (size_t)sy * cw, *sv = v + (size_t)sy * cw;
    double vu = 0.0, vv = 0.0;
    float mnu = 1e30f, mxu = -1e30f, mnv = 1e30f, mxv = -1e30f;
    if constexpr (SUP == 6) {
        const PixAxis<6> px = pix_axis<6>(ax, ox);
        float a[6], b[6];
        _Pragma("unroll") for (int i = 0; i < 6; ++i) {
            const int x = clampi(px.start + i, 0, cw - 1);
            a[i] = su[x];
            b[i] = sv[x];
        }
        _Pragma("unroll") for (int i = 0; i < 6; ++i) {
            const bool on = px.on(i);
            const float w = on ? px.w[i] : 0.0f;
            vu = dacc(vu, w, a[i]);
            vv = dacc(vv, w, b[i]);
            const float nmu = fmn(mnu, a[i]), nxu = fmx(mxu, a[i]), nmv = fmn(mnv, b[i]), nxv = fmx(mxv, b[i]);
            mnu = on ? nmu : mnu;
            mxu = on ? nxu : mxu;
            mnv = on ? nmv : mnv;
            mxv = on ? nxv : mxv;
        }
    } else if constexpr (SUP > 6) {
        // fixed tap count from the plain tables: every load issued up
        // front, inactive taps leave the sums and hull untouched (selects)
        const int start = ax.start[ox];
        const float *wt = ax.w + (size_t)ox * SUP;
        const float *am = ax.am + (size_t)ox * SUP;
        float a[SUP], b[SUP], w[SUP];
        bool on[SUP];
        _Pragma("unroll") for (int i = 0; i < SUP; ++i) {
            const int x = clampi(start + i, 0, cw - 1);
            a[i] = su[x];
            b[i] = sv[x];
            w[i] = wt[i];
            on[i] = am[i] != 0.0f;
        }
        _Pragma("unroll") for (int i = 0; i < SUP; ++i) {
            const double nu = dacc(vu, w[i], a[i]), nv = dacc(vv, w[i], b[i]);
            vu = on[i] ? nu : vu;
            vv = on[i] ? nv : vv;
            const float nmu = fmn(mnu, a[i]), nxu = fmx(mxu, a[i]), nmv = fmn(mnv, b[i]), nxv = fmx(mxv, b[i]);
            mnu = on[i] ? nmu : mnu;
            mxu = on[i] ? nxu : mxu;
            mnv = on[i] ? nmv : mnv;
            mxv = on[i] ? nxv : mxv;
        }
    } else {
        const int start = ax.start[ox];
        const float *wt = ax.w + (size_t)ox * ax.sup;
        const float *am = ax.am + (size_t)ox * ax.sup;
        for (int i = 0; i < ax.sup; ++i) {
            if (am[i] == 0.0f)
                continue;
            const int x = clampi(start + i, 0, cw - 1);
            const float a = su[x], b = sv[x];
            vu = dacc(vu, wt[i], a);
            vv = dacc(vv, wt[i], b);
            mnu = fmn(mnu, a); mxu = fmx(mxu, a);
            mnv = fmn(mnv, b); mxv = fmx(mxv, b);
        }
    }
    // interleaved: {vu, vv} then {min u, max u, min v, max v}
    const size_t n = (size_t)ax.n * ch, k = (size_t)sy * ax.n + ox;
    reinterpret_cast<float2 *>(hb)[k] = make_float2((float)vu, (float)vv);
    reinterpret_cast<float4 *>(hb + ((2 * n + 3) & ~(size_t)3))[k] = make_float4(mnu, mxu, mnv, mxv);
}

template <int SUP>
__global__ __launch_bounds__(kBlock) void k_plain_v(const float *__restrict__ u, const float *__restrict__ v,
                                                    int cw, int ch, AxisDev ax, AxisDev ay,
                                                    const float *__restrict__ hb, float *out_u,
                                                    float *out_v, int clamp_hull, float margin) {
    unsigned bx, by;
    xcd_block2(bx, by);
    const int ox = bx * kBlock + threadIdx.x, oy = by;
    if (ox >= ax.n)
        return;
    const int ow = ax.n;
    const size_t n = (size_t)ow * ch;
    const float2 *hv = reinterpret_cast<const float2 *>(hb);
    const float4 *hh = reinterpret_cast<const float4 *>(hb + ((2 * n + 3) & ~(size_t)3));
    const int start = ay.start[oy];
    const float *wt = ay.w + (size_t)oy * ay.sup;
    const float *am = ay.am + (size_t)oy * ay.sup;
    double vu = 0.0, vv = 0.0;
    float mnu = 1e30f, mxu = -1e30f, mnv = 1e30f, mxv = -1e30f;
    if constexpr (SUP == 6) {
        float2 val[6];
        float4 hull[6];
        _Pragma("unroll") for (int j = 0; j < 6; ++j) {
            const size_t k = (size_t)clampi(start + j, 0, ch - 1) * ow + ox;
            val[j] = hv[k];
            hull[j] = hh[k];
        }
        _Pragma("unroll") for (int j = 0; j < 6; ++j) {
            const bool on = am[j] != 0.0f;   // uniform per block row
            const float w = on ? wt[j] : 0.0f;
            vu = dacc(vu, w, val[j].x);
            vv = dacc(vv, w, val[j].y);
            const float nmu = fmn(mnu, hull[j].x), nxu = fmx(mxu, hull[j].y);
            const float nmv = fmn(mnv, hull[j].z), nxv = fmx(mxv, hull[j].w);
            mnu = on ? nmu : mnu;
            mxu = on ? nxu : mxu;
            mnv = on ? nmv : mnv;
            mxv = on ? nxv : mxv;
        }
    } else if constexpr (SUP > 6) {
        float2 val[SUP];
        float4 hull[SUP];
        _Pragma("unroll") for (int j = 0; j < SUP; ++j) {
            const size_t k = (size_t)clampi(start + j, 0, ch - 1) * ow + ox;
            val[j] = hv[k];
            hull[j] = hh[k];
        }
        _Pragma("unroll") for (int j = 0; j < SUP; ++j) {
            if (am[j] == 0.0f)   // uniform per block row
                continue;
            vu = dacc(vu, wt[j], val[j].x);
            vv = dacc(vv, wt[j], val[j].y);
            mnu = fmn(mnu, hull[j].x); mxu = fmx(mxu, hull[j].y);
            mnv = fmn(mnv, hull[j].z); mxv = fmx(mxv, hull[j].w);
        }
    } else {
        for (int j = 0; j < ay.sup; ++j) {
            if (am[j] == 0.0f)
                continue;
            const size_t k = (size_t)clampi(start + j, 0, ch - 1) * ow + ox;
            const float2 val = hv[k];
            const float4 hull = hh[k];
            vu = dacc(vu, wt[j], val.x);
            vv = dacc(vv, wt[j], val.y);
            mnu = fmn(mnu, hull.x); mxu = fmx(mxu, hull.y);
            mnv = fmn(mnv, hull.z); mxv = fmx(mxv, hull.w);
        }
    }
    const double wsum = (double)ax.wsum[ox] * ay.wsum[oy];
    const double asum = (double)ax.wabs[ox] * ay.wabs[oy];
    if (wsum > 1e-6 && wsum >= 0.05 * asum) {
        vu /= wsum;
        vv /= wsum;
    } else {
        vu = bilinear(u, cw, ch, ax.pos[ox], ay.pos[oy]);
        vv = bilinear(v, cw, ch, ax.pos[ox], ay.pos[oy]);
    }
    if (clamp_hull) {
        vu = dclamp(vu, (double)(mnu - margin), (double)(mxu + margin));
        vv = dclamp(vv, (double)(mnv - margin), (double)(mxv + margin));
    }
    const size_t o = (size_t)oy * ow + ox;
    out_u[o] = (float)vu;
    out_v[o] = (float)vv;
}

// Both passes in one kernel for 6-tap axes: a block owns 64 output columns
// x 16 output rows, computes the horizontal pass of the source rows those
// outputs read into LDS (each element exactly as k_plain_h<6> computes it)
// and runs the vertical pass from there.  The host enables it when every
// tile reads at most kPlainRows source rows.
constexpr int kPlainCols = 64, kPlainTileRows = 16, kPlainRows = 24;

__global__ __launch_bounds__(kBlock) void k_plain_fused(const float *__restrict__ u, const float *__restrict__ v,
                                                        int cw, int ch, AxisDev ax, AxisDev ay, float *out_u,
                                                        float *out_v, int clamp_hull, float margin) {
    __shared__ float2 sval[kPlainRows][kPlainCols];
    __shared__ float4 shull[kPlainRows][kPlainCols];
    unsigned bx, by;
    xcd_block2(bx, by);
    const int c = threadIdx.x & (kPlainCols - 1), g = threadIdx.x / kPlainCols;
    const int ow = ax.n, oh = ay.n;
    const int ox = (int)bx * kPlainCols + c, oy0 = (int)by * kPlainTileRows;
    const int r0 = ay.start[oy0];
    const int nr = ay.start[min(oy0 + kPlainTileRows, oh) - 1] + 6 - r0;
    // columns past the right edge repeat the last one in the first pass
    const PixAxis<6> px = pix_axis<6>(ax, min(ox, ow - 1));
    int xs[6];
    _Pragma("unroll") for (int i = 0; i < 6; ++i) xs[i] = clampi(px.start + i, 0, cw - 1);
    for (int t = g; t < nr; t += kBlock / kPlainCols) {
        const size_t row = (size_t)clampi(r0 + t, 0, ch - 1) * cw;
        float a[6], b[6];
        _Pragma("unroll") for (int i = 0; i < 6; ++i) {
            a[i] = u[row + xs[i]];
            b[i] = v[row + xs[i]];
        }
        double vu = 0.0, vv = 0.0;
        float mnu = 1e30f, mxu = -1e30f, mnv = 1e30f, mxv = -1e30f;
        _Pragma("unroll") for (int i = 0; i < 6; ++i) {
            const bool on = px.on(i);
            const float w = on ? px.w[i] : 0.0f;
            vu = dacc(vu, w, a[i]);
            vv = dacc(vv, w, b[i]);
            const float nmu = fmn(mnu, a[i]), nxu = fmx(mxu, a[i]), nmv = fmn(mnv, b[i]), nxv = fmx(mxv, b[i]);
            mnu = on ? nmu : mnu;
            mxu = on ? nxu : mxu;
            mnv = on ? nmv : mnv;
            mxv = on ? nxv : mxv;
        }
        sval[t][c] = make_float2((float)vu, (float)vv);
        shull[t][c] = make_float4(mnu, mxu, mnv, mxv);
    }
    __syncthreads();
    if (ox >= ow)
        return;
    const float axw = ax.wsum[ox], axa = ax.wabs[ox], axp = ax.pos[ox];
    constexpr int kPer = kPlainTileRows / (kBlock / kPlainCols);
    for (int k = 0; k < kPer; ++k) {
        const int oy = oy0 + g * kPer + k;
        if (oy >= oh)
            break;
        const int t0 = ay.start[oy] - r0;
        const float *wt = ay.w + (size_t)oy * 6;
        const float *am = ay.am + (size_t)oy * 6;
        double vu = 0.0, vv = 0.0;
        float mnu = 1e30f, mxu = -1e30f, mnv = 1e30f, mxv = -1e30f;
        _Pragma("unroll") for (int j = 0; j < 6; ++j) {
            const bool on = am[j] != 0.0f;
            const float w = on ? wt[j] : 0.0f;
            const float2 val = sval[t0 + j][c];
            const float4 hull = shull[t0 + j][c];
            vu = dacc(vu, w, val.x);
            vv = dacc(vv, w, val.y);
            const float nmu = fmn(mnu, hull.x), nxu = fmx(mxu, hull.y);
            const float nmv = fmn(mnv, hull.z), nxv = fmx(mxv, hull.w);
            mnu = on ? nmu : mnu;
            mxu = on ? nxu : mxu;
            mnv = on ? nmv : mnv;
            mxv = on ? nxv : mxv;
        }
        const double wsum = (double)axw * ay.wsum[oy];
        const double asum = (double)axa * ay.wabs[oy];
        if (wsum > 1e-6 && wsum >= 0.05 * asum) {
            vu /= wsum;
            vv /= wsum;
        } else {
            vu = bilinear(u, cw, ch, axp, ay.pos[oy]);
            vv = bilinear(v, cw, ch, axp, ay.pos[oy]);
        }
        if (clamp_hull) {
            vu = dclamp(vu, (double)(mnu - margin), (double)(mxu + margin));
            vv = dclamp(vv, (double)(mnv - margin), (double)(mxv + margin));
        }
        const size_t o = (size_t)oy * ow + ox;
        out_u[o] = (float)vu;
        out_v[o] = (float)vv;
    }
}

// plainChromaBilinear (recon.cpp:1031-1149)
__global__ __launch_bounds__(kBlock) void k_plain_bilinear(const float *u, const float *v, int cw, int ch,
                                                           AxisDev ax, AxisDev ay, float *out_u,
                                                           float *out_v) {
    const int ox = blockIdx.x * kBlock + threadIdx.x, oy = blockIdx.y;
    if (ox >= ax.n)
        return;
    const size_t o = (size_t)oy * ax.n + ox;
    out_u[o] = bilinear_fast(u, cw, ch, ax.cb_i0[ox], ax.cb_f[ox], ay.cb_i0[oy], ay.cb_f[oy]);
    out_v[o] = bilinear_fast(v, cw, ch, ax.cb_i0[ox], ax.cb_f[ox], ay.cb_i0[oy], ay.cb_f[oy]);
}

__device__ __forceinline__ float lut_lookup(const float *lut, int n, float idx) {
    int i0 = (int)idx;
    i0 = clampi(i0, 0, n - 2);
    const float fr = fmn(idx - (float)i0, 1.0f);
    return lut[i0] + (lut[i0 + 1] - lut[i0]) * fr;
}

__device__ __forceinline__ const float *radial_at(const RadialDev &r, int x, int y) {
    if (!r.valid)
        return nullptr;
    return r.w + ((size_t)r.yphase[y] * r.nx + r.xphase[x]) * (size_t)r.sup_x * r.sup_y;
}

// Radial (jinc) base weight of tap (i, j) of output (ox, oy): the phase
// table entry when the table exists, else the LUT at the AA-scaled distance
// (recon.cpp:598-606, 984-991).
__device__ __forceinline__ float radial_weight(const RadialDev &r, const float *base, int sup_x, int i, int j,
                                               int tx0, int ty0, float scx, float scy) {
    if (base)
        return base[(size_t)j * sup_x + i];
    const float dx = ((float)(tx0 + i) - scx) * r.inv_wx;
    const float dy = ((float)(ty0 + j) - scy) * r.inv_wy;
    return lut_lookup(r.lut, r.lut_n, sqrtf(dx * dx + dy * dy) * r.lut_scale);
}

// plainChromaImpl for the radial kernel (recon.cpp:828-1029)
__global__ __launch_bounds__(kBlock) void k_plain_radial(const float *__restrict__ u, const float *__restrict__ v,
                                                         int cw, int ch, AxisDev ax, AxisDev ay, RadialDev rd,
                                                         float *out_u, float *out_v, int clamp_hull, float margin) {
    const int ox = blockIdx.x * kBlock + threadIdx.x, oy = blockIdx.y;
    if (ox >= ax.n)
        return;
    const int tx0 = ax.start[ox], ty0 = ay.start[oy];
    const float *amx = ax.am + (size_t)ox * ax.sup, *amy = ay.am + (size_t)oy * ay.sup;
    const float scx = ax.pos[ox], scy = ay.pos[oy];
    const float *base = radial_at(rd, ox, oy);
    double au = 0.0, av = 0.0, ws = 0.0, sa = 0.0;
    float mnu = 1e30f, mxu = -1e30f, mnv = 1e30f, mxv = -1e30f;
    for (int j = 0; j < ay.sup; ++j) {
        if (amy[j] == 0.0f)
            continue;
        const size_t row = (size_t)clampi(ty0 + j, 0, ch - 1) * cw;
        for (int i = 0; i < ax.sup; ++i) {
            if (amx[i] == 0.0f)
                continue;
            const float w = radial_weight(rd, base, ax.sup, i, j, tx0, ty0, scx, scy);
            const size_t k = row + clampi(tx0 + i, 0, cw - 1);
            const float a = u[k], b = v[k];
            au = dacc(au, w, a);
            ws += w;
            sa += fabsf(w);
            mnu = fmn(mnu, a); mxu = fmx(mxu, a);
            av = dacc(av, w, b);
            mnv = fmn(mnv, b); mxv = fmx(mxv, b);
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
        vu = dclamp(vu, (double)(mnu - margin), (double)(mxu + margin));
        vv = dclamp(vv, (double)(mnv - margin), (double)(mxv + margin));
    }
    const size_t o = (size_t)oy * ax.n + ox;
    out_u[o] = (float)vu;
    out_v[o] = (float)vv;
}

// The radial plain base with the phase tables, one x phase per block row:
// the phase pair's weights, tap activity and weight sums are uniform across
// the wave (scalar loads), so a tap costs two source loads, two FP64 FMAs
// and the hull updates.  Same arithmetic and order as k_plain_radial.
template <int SUP>
__global__ __launch_bounds__(kBlock) void k_plain_radial_tab(const float *__restrict__ u,
                                                             const float *__restrict__ v, int cw, int ch,
                                                             AxisDev ax, AxisDev ay, RadialDev rd, float *out_u,
                                                             float *out_v, int clamp_hull, float margin) {
    const int cls = blockIdx.z, oy = blockIdx.y;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    const int beg = rd.xoff[cls], cnt = rd.xoff[cls + 1] - beg;
    if (k >= cnt)
        return;
    const int SX = SUP ? SUP : ax.sup, SY = SUP ? SUP : ay.sup;
    const int ox = rd.xs[beg + k];
    const int pair = (int)rd.yphase[oy] * rd.nx + cls;
    const float *__restrict__ base = rd.w + (size_t)pair * SX * SY;
    // tap activity is part of the phase: the class's first column has it
    const float *__restrict__ amx = ax.am + (size_t)rd.xs[beg] * SX;
    const float *__restrict__ amy = ay.am + (size_t)oy * SY;
    const int tx0 = ax.start[ox], ty0 = ay.start[oy];
    double au = 0.0, av = 0.0;
    float mnu = 1e30f, mxu = -1e30f, mnv = 1e30f, mxv = -1e30f;
    for (int j = 0; j < SY; ++j) {
        if (amy[j] == 0.0f)
            continue;
        const size_t row = (size_t)clampi(ty0 + j, 0, ch - 1) * cw;
        _Pragma("unroll") for (int i = 0; i < SX; ++i) {
            if (amx[i] == 0.0f)
                continue;
            const float w = base[j * SX + i];
            const size_t q = row + clampi(tx0 + i, 0, cw - 1);
            const float a = u[q], b = v[q];
            au = dacc(au, w, a);
            av = dacc(av, w, b);
            mnu = fmn(mnu, a); mxu = fmx(mxu, a);
            mnv = fmn(mnv, b); mxv = fmx(mxv, b);
        }
    }
    const double ws = rd.wsa[2 * pair], sa = rd.wsa[2 * pair + 1];
    const float scx = ax.pos[ox], scy = ay.pos[oy];
    double vu, vv;
    if (ws > 1e-6 && ws >= 0.05 * sa) {
        vu = au / ws;
        vv = av / ws;
    } else {
        vu = bilinear(u, cw, ch, scx, scy);
        vv = bilinear(v, cw, ch, scx, scy);
    }
    if (clamp_hull) {
        vu = dclamp(vu, (double)(mnu - margin), (double)(mxu + margin));
        vv = dclamp(vv, (double)(mnv - margin), (double)(mxv + margin));
    }
    const size_t o = (size_t)oy * ax.n + ox;
    out_u[o] = (float)vu;
    out_v[o] = (float)vv;
}

// The radial plain base with 6x6 taps.  A block computes R consecutive
// output rows of one segment of kSegCols output columns: each of its waves
// takes the columns of one x phase inside the segment (rd.segw), so the
// phase pair's weights and tap activity are wave-uniform, and each thread
// computes its column in the R rows by walking the union of their source
// rows once.  A source row's six samples are loaded once and feed every
// output row whose taps include it (at 2x or more upscaling R rows share
// most of their source rows), and the row's hull over the active columns is
// formed once for all of them.  The results go through LDS, so the block
// writes whole rows of the segment (a phase's columns alone are strided,
// and partial-line writes cost a read of the rest of the line).  Each output
// accumulates its active taps in the reference's (j, i) order; its hull is
// the running minimum/maximum over the rows' minima/maxima in j order, which
// picks the same sample as the running hull over the taps in (j, i) order
// (ties keep the earlier sample, NaN never replaces a bound), so the results
// are those of the per-tap order.
// TILE (rd.tile_fit): the block's source rows and columns are first staged
// in LDS at clamped coordinates (every read then equals the clamped global
// read), so the taps read LDS instead of global memory.
template <int R, bool TILE>
__global__ __launch_bounds__(kBlock) void k_plain_radial_rows(const float *__restrict__ u,
                                                              const float *__restrict__ v, int cw, int ch,
                                                              AxisDev ax, AxisDev ay, RadialDev rd, float *out_u,
                                                              float *out_v, int clamp_hull, float margin) {
    static_assert(R == kRadialRows, "the host checks the source window for kRadialRows rows");
    __shared__ float2 tile[R][kSegCols];
    __shared__ float2 srct[TILE ? kRadTileRows : 1][TILE ? kRadTileCols : 1];
    const int sg = blockIdx.x, oy0 = blockIdx.y * R;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int nrow = min(R, ay.n - oy0);
    int ty[R], ypair[R];
    uint32_t my[R];
    int rlo = INT_MAX, rhi = INT_MIN;
    _Pragma("unroll") for (int s = 0; s < R; ++s) {
        const int oy = min(oy0 + s, ay.n - 1);
        ty[s] = ay.start[oy];
        ypair[s] = (int)rd.yphase[oy] * rd.nx;
        const float *__restrict__ amy = ay.am + (size_t)oy * 6;
        my[s] = 0;
        _Pragma("unroll") for (int t = 0; t < 6; ++t)
            my[s] |= (amy[t] != 0.0f ? 1u : 0u) << t;
        if (s < nrow) {
            rlo = min(rlo, ty[s]);
            rhi = max(rhi, ty[s] + 5);
        }
    }
    // source columns of the segment (chroma starts rise with the output column)
    const int c_lo = ax.start[sg * kSegCols];
    if constexpr (TILE) {
        const int c_hi = ax.start[min(sg * kSegCols + kSegCols - 1, ax.n - 1)] + 5;
        const int tw = c_hi - c_lo + 1, th = rhi - rlo + 1;  // at most kRadTileCols x kRadTileRows (host check)
        for (int e = threadIdx.x; e < tw * th; e += kBlock) {
            const int r = e / tw, c = e - r * tw;
            const size_t q = (size_t)clampi(rlo + r, 0, ch - 1) * cw + clampi(c_lo + c, 0, cw - 1);
            srct[r][c] = make_float2(u[q], v[q]);
        }
        __syncthreads();
    }
    for (int e = rd.segoff[sg] + wave; e < rd.segoff[sg + 1]; e += kBlock / 64) {
        const int4 en = rd.segw[e];
        const int cls = en.x, k = en.y + lane;
        if (k >= en.z)
            continue;
        const int ox = rd.xs[k];
        // tap activity is part of the phase: the class's first column has it
        const float *__restrict__ amx = ax.am + (size_t)rd.xs[rd.xoff[cls]] * 6;
        uint32_t mx = 0;
        _Pragma("unroll") for (int t = 0; t < 6; ++t)
            mx |= (amx[t] != 0.0f ? 1u : 0u) << t;
        const int tx0 = ax.start[ox];
        int col[6];
        _Pragma("unroll") for (int i = 0; i < 6; ++i)
            col[i] = clampi(tx0 + i, 0, cw - 1);
        double au[R], av[R];
        float mnu[R], mxu[R], mnv[R], mxv[R];
        _Pragma("unroll") for (int s = 0; s < R; ++s) {
            au[s] = av[s] = 0.0;
            mnu[s] = mnv[s] = 1e30f;
            mxu[s] = mxv[s] = -1e30f;
        }
        // source rows are loaded one ahead of their use; the row loop is
        // compiled once with every column tap active (the interior phases:
        // straight-line code) and once with the activity mask
        auto rows = [&](auto all6_tag) {
            constexpr bool ALL6 = decltype(all6_tag)::value;
            float na[6], nb[6];
            if constexpr (!TILE) {
                const size_t row = (size_t)clampi(rlo, 0, ch - 1) * cw;
                _Pragma("unroll") for (int i = 0; i < 6; ++i) {
                    na[i] = u[row + col[i]];
                    nb[i] = v[row + col[i]];
                }
            }
            const int co = tx0 - c_lo;
            for (int r = rlo; r <= rhi; ++r) {
                float a[6], b[6];
                if constexpr (TILE) {
                    const float2 *tr = &srct[r - rlo][co];
                    _Pragma("unroll") for (int i = 0; i < 6; ++i) {
                        const float2 t = tr[i];
                        a[i] = t.x;
                        b[i] = t.y;
                    }
                } else {
                    _Pragma("unroll") for (int i = 0; i < 6; ++i) {
                        a[i] = na[i];
                        b[i] = nb[i];
                    }
                    if (r < rhi) {
                        const size_t row = (size_t)clampi(r + 1, 0, ch - 1) * cw;
                        _Pragma("unroll") for (int i = 0; i < 6; ++i) {
                            na[i] = u[row + col[i]];
                            nb[i] = v[row + col[i]];
                        }
                    }
                }
                // the samples in double (exact) once for every output row
                double da[6], db[6];
                float rmnu = 1e30f, rmxu = -1e30f, rmnv = 1e30f, rmxv = -1e30f;
                _Pragma("unroll") for (int i = 0; i < 6; ++i) {
                    da[i] = (double)a[i];
                    db[i] = (double)b[i];
                    if (ALL6 || ((mx >> i) & 1u)) {
                        rmnu = fmn(rmnu, a[i]); rmxu = fmx(rmxu, a[i]);
                        rmnv = fmn(rmnv, b[i]); rmxv = fmx(rmxv, b[i]);
                    }
                }
                _Pragma("unroll") for (int s = 0; s < R; ++s) {
                    const int j = r - ty[s];
                    if (s >= nrow || j < 0 || j > 5 || !((my[s] >> j) & 1u))
                        continue;  // wave-uniform
                    // the pair's weights as doubles (the float weights,
                    // exactly), wave-uniform
                    const double *__restrict__ wr = rd.wd + (size_t)(ypair[s] + cls) * 36 + j * 6;
                    _Pragma("unroll") for (int i = 0; i < 6; ++i) {
                        if (!ALL6 && !((mx >> i) & 1u))
                            continue;
                        au[s] = __builtin_fma(wr[i], da[i], au[s]);
                        av[s] = __builtin_fma(wr[i], db[i], av[s]);
                    }
                    mnu[s] = fmn(mnu[s], rmnu); mxu[s] = fmx(mxu[s], rmxu);
                    mnv[s] = fmn(mnv[s], rmnv); mxv[s] = fmx(mxv[s], rmxv);
                }
            }
        };
        if (mx == 0x3fu)  // wave-uniform
            rows(std::true_type{});
        else
            rows(std::false_type{});
        const float scx = ax.pos[ox];
        _Pragma("unroll") for (int s = 0; s < R; ++s) {
            if (s >= nrow)
                break;
            const int pair = ypair[s] + cls;
            const double ws = rd.wsa[2 * pair], sa = rd.wsa[2 * pair + 1];
            double vu, vv;
            if (ws > 1e-6 && ws >= 0.05 * sa) {
                vu = au[s] / ws;
                vv = av[s] / ws;
            } else {
                const float scy = ay.pos[oy0 + s];
                vu = bilinear(u, cw, ch, scx, scy);
                vv = bilinear(v, cw, ch, scx, scy);
            }
            if (clamp_hull) {
                vu = dclamp(vu, (double)(mnu[s] - margin), (double)(mxu[s] + margin));
                vv = dclamp(vv, (double)(mnv[s] - margin), (double)(mxv[s] + margin));
            }
            tile[s][ox - sg * kSegCols] = make_float2((float)vu, (float)vv);
        }
    }
    __syncthreads();
    const int x = sg * kSegCols + (int)threadIdx.x;
    if (x < ax.n)
        for (int s = 0; s < nrow; ++s) {
            const size_t o = (size_t)(oy0 + s) * ax.n + x;
            const float2 t = tile[s][threadIdx.x];
            out_u[o] = t.x;
            out_v[o] = t.y;
        }
}

// ---------------------------------------------------------------------------
// guided reconstruction, algo 2 (reconstructChromaImpl, recon.cpp:14-819):
// one thread per active output pixel of the compacted workset.
// ---------------------------------------------------------------------------

struct GuidedArgs {
    const float *u, *v;
    int cw, ch;
    const float *y;
    int sw, sh;
    const float4 *yt;    // {Y, jxx, jxy, jyy} per luma pixel
    const float4 *luv;   // {lc, U, V, 0} per chroma sample
    const float *lc, *ms;
    AxisDev ax, ay;
    RadialDev rd;
    GuideParams gp;
    const uint32_t *row_off;
    const uint32_t *indices;
    float *out_u, *out_v;
    SelectorDev sel;
    TemporalDev tmp;
    const uint64_t *mbits;   // trust mask bits at mask resolution (tile kernel), nullptr = dense
    int mw, mh, mw64;
    // tile kernel with fused output: back-projection (bp_on) and packing
    // straight into the output planes
    RawOut po_u, po_v;
    float bp;
    int bp_on;
};

// One frame of a multi-frame tile launch: its arguments and tile grid.
struct BatchFrame {
    GuidedArgs A;
    int tiles_x, tiles;
};
// kernel variant of a fused multi-frame tile launch: temporal taps (TRecon,
// no plain base), plain hull, output sample type (1 = 8-bit, 2 = 16-bit, 3 =
// float)
constexpr int batch_variant(bool temporal, bool ph, int outt) { return (temporal ? 1 : 0) | (ph ? 2 : 0) | (outt << 2); }

// True when the float selector keeps compressed per-active arrays: a sparse
// workset that is not in dense fallback (plugin.cpp:455-456).
__device__ __forceinline__ bool sel_compressed(const uint32_t *row_off, const uint32_t *indices, int ow, int oh) {
    return indices && (uint64_t)row_off[oh] * 2 < (uint64_t)ow * (uint64_t)oh;
}

// Feature switches are template parameters so each launch carries only the
// code (and registers) of the features it uses.
// WIN: the pass-1 and tap loops read the chroma guide from an LDS window of
// kTWin x kTWin samples whose origin is (wx0, wy0), filled with luv at
// clamped coordinates (so it holds exactly what the global reads return).
constexpr int kTile = 8, kTWin = 16, kTStride = kTWin + 1;  // padded rows spread LDS banks

// Per-pixel inputs of the weighted tap sum.
struct TapIn {
    float L0, nxv, nyv, scx, scy, inv_s2, guide_fade, inv_g2, inv_g2p, rescue;
    int tx0, ty0;
};
struct TapAcc {
    double u = 0.0, v = 0.0, wsum = 0.0, sabs = 0.0;
};

// Guided tap sum (recon.cpp:581-627).  With fixed 6-tap supports the loops
// are straight-line code: inactive taps are computed and their weight is
// replaced by an exact zero, which leaves every accumulator unchanged.
// FAST: the caller has shown every reciprocal argument of the active taps
// lies in [1, 2^126), where rcp_nr is exact.  GUIDED: strength > 0.
// NORESCUE: the caller has shown rescue == 0 for the pixel.  Then
// rescue * bump is +0 (bump is finite), wx + 0 == wx and the weight is
// exactly sim * wx (a zero weight may change sign, which no accumulator
// sees), so the bump kernel and the projected offsets are not computed.
// Inactive taps of the 6-tap axis records carry an exact zero base weight.
// With finite samples every similarity and bump value is finite, so an
// inactive tap's weight sim * (wx + r * bump) is an exact (signed) zero
// when its rescue factor r is zero as well, which leaves every accumulator
// unchanged (no accumulator is ever -0.0).  NORESCUE (separable) thus needs
// no weight masks; ROWSON (separable, every vertical tap of the wave
// active) takes the rescue factor per column, zero on inactive columns, in
// place of a mask per tap.
template <bool RADIAL, int SUP, bool WIN, bool FAST, bool GUIDED, bool NORESCUE = false, bool ROWSON = false>
__device__ __forceinline__ void guided_taps(const GuidedArgs &A, const PixAxis<SUP> &px, const PixAxis<SUP> &py,
                                            int ox, int oy, const TapIn &t, const float4 *win, int wx0, int wy0,
                                            TapAcc &acc) {
    const GuideParams &gp = A.gp;
    const int cw = A.cw, ch = A.ch;
    const int SX = SUP ? SUP : A.ax.sup, SY = SUP ? SUP : A.ay.sup;
    const float *rbase = RADIAL ? radial_at(A.rd, ox, oy) : nullptr;
    // Window offset of the first tap.  The empty asm hides its value so the
    // compiler reloads the window here instead of keeping all 36 samples of
    // the first pass live in registers.
    int woff = (t.ty0 - wy0) * kTStride + (t.tx0 - wx0);
    int goff = 0;  // the same for the global chroma guide
    if (WIN)
        __asm__ volatile("" : "+v"(woff));
    else if (SUP)
        __asm__ volatile("" : "+s"(goff));
    PixAxis<SUP> qx = px, qy = py;
    if constexpr (SUP == 6) {
        // likewise the active-tap masks are rebuilt here, not kept from the
        // first pass
        __asm__ volatile("" : "+v"(qx.bits), "+v"(qy.bits));
    }
    constexpr int NH = SUP ? SUP : 1;
    float dxn[NH], dxt[NH];
    float resc[NH];
    if constexpr (ROWSON) {
        _Pragma("unroll") for (int i = 0; i < NH; ++i)
            resc[i] = qx.on(i) ? t.rescue : 0.0f;
    }
    if constexpr (SUP != 0 && !NORESCUE) {
        _Pragma("unroll") for (int i = 0; i < SUP; ++i) {
            const float dx = ((float)(t.tx0 + i) - t.scx) * gp.frw;
            dxn[i] = dx * t.nxv;
            dxt[i] = -dx * t.nyv;
        }
    }
    _Pragma("unroll") for (int j = 0; j < SY; ++j) {
        const bool onj = qy.on(j);
        if (!SUP && !onj)
            continue;
        const size_t row = (size_t)clampi(t.ty0 + j, 0, ch - 1) * cw;
        const float wy = qy.wt(j);
        float dyn = 0.0f, dyt = 0.0f;
        if constexpr (!NORESCUE) {
            const float dy = ((float)(t.ty0 + j) - t.scy) * gp.frh;
            dyn = dy * t.nyv;
            dyt = dy * t.nxv;
        }
        _Pragma("unroll") for (int i = 0; i < SX; ++i) {
            const bool on = onj && qx.on(i);
            if (!SUP && !on)
                continue;
            const float4 q = WIN ? win[woff + j * kTStride + i] : A.luv[row + goff + clampi(t.tx0 + i, 0, cw - 1)];
            const float wx = RADIAL ? radial_weight(A.rd, rbase, SX, i, j, t.tx0, t.ty0, t.scx, t.scy)
                                    : qx.wt(i) * wy;
            float w;
            if (GUIDED && NORESCUE) {
                const float dL = q.x - t.L0;
                const float ys = 1.0f + dL * dL * t.inv_s2;
                const float sim_raw = FAST ? rcp_nr(ys) : recip_ge1(ys);
                const float sim = 1.0f - t.guide_fade * (1.0f - sim_raw);
                w = sim * wx;
            } else if (GUIDED) {
                float dp, dt;
                if constexpr (SUP != 0) {
                    dp = dxn[i] + dyn;
                    dt = dxt[i] + dyt;
                } else {
                    const float dx = ((float)(t.tx0 + i) - t.scx) * gp.frw;
                    dp = dx * t.nxv + dyn;
                    dt = -dx * t.nyv + dyt;
                }
                const float dL = q.x - t.L0;
                const float ys = 1.0f + dL * dL * t.inv_s2;
                const float yb = 1.0f + dp * dp * t.inv_g2 + dt * dt * t.inv_g2p;
                const float sim_raw = FAST ? rcp_nr(ys) : recip_ge1(ys);
                const float sim = 1.0f - t.guide_fade * (1.0f - sim_raw);
                const float bump = FAST ? rcp_nr(yb) : recip_ge1(yb);
                w = sim * (wx + (ROWSON ? resc[i] : t.rescue) * bump);
            } else {
                w = wx;
            }
            constexpr bool kMaskFree = ROWSON || (NORESCUE && !RADIAL);
            if (SUP && !kMaskFree)
                w = on ? w : 0.0f;
            acc.u = dacc(acc.u, w, q.y);
            acc.v = dacc(acc.v, w, q.z);
            // |double(w)| == double(|w|): one conversion serves both sums
            const double wd = (double)w;
            acc.wsum += wd;
            acc.sabs += __builtin_fabs(wd);
        }
        // one row of taps in flight at a time keeps the register footprint
        // of the unrolled loop small
        if (SUP)
            __builtin_amdgcn_sched_barrier(0);
    }
}

// HULL_IN: the caller passes the chroma hull of the pixel's active taps
// (hull_in = {min U, max U, min V, max V}), which the plain base of the tile
// has formed over the same taps; pass 1 then only forms the luma range.
// With finite samples the minimum and maximum of a set do not depend on the
// order they are taken in, so the hull is the one pass 1 would form.
template <bool RADIAL, bool TEMPORAL, int SEL, bool CEDGE, int SUP, bool WIN = false, bool HULL_IN = false>
__device__ __attribute__((always_inline)) void guided_pixel(const GuidedArgs &A, int ox, int oy, uint32_t k,
                                                            const float4 *win = nullptr, int wx0 = 0,
                                                            int wy0 = 0, float2 *res = nullptr,
                                                            float4 hull_in = float4{}) {
    const GuideParams &gp = A.gp;
    const int cw = A.cw, ch = A.ch, sw = A.sw, sh = A.sh;
    const bool guided = gp.strength > 0.0f;
    const PixAxis<SUP> px = pix_axis<SUP>(A.ax, ox), py = pix_axis<SUP>(A.ay, oy);

    // --- guide metadata (recon.cpp:155-218)
    float L0 = 0.0f, nxv = 0.0f, nyv = 0.0f, coherence = 0.0f, min_diff = 0.0f, ridge = 0.0f;
    float phase = 0.0f, mfade = 1.0f;
    bool has_dir = false;
    const float scx = px.pos, scy = py.pos;
    if (guided) {
        const float4 t = bilinear_fast4(A.yt, sw, sh, px.lb_i0, px.lb_f, py.lb_i0, py.lb_f);
        L0 = t.x;
        const float jxx = t.y, jxy = t.z, jyy = t.w;
        const Dir d = principal_dir(jxx, jxy, jyy);
        coherence = d.coherence;
        nxv = d.nx;
        nyv = d.ny;
        has_dir = jxx + jyy > 1e-6f;
        const float lx = px.lpos, ly = py.lpos;
        if (has_dir) {
            const float sp = gp.sp_out;
            const float ya = bilinear_f(A.y, sw, sh, lx - sp * nxv, ly - sp * nyv);
            const float yb = bilinear_f(A.y, sw, sh, lx + sp * nxv, ly + sp * nyv);
            min_diff = fmn(fabsf(ya - L0), fabsf(yb - L0));
            if (gp.ridge) {
                const float dl = 1.5f * sp;
                const float dl2 = 2.0f * dl;
                const float ym2 = bilinear_f(A.y, sw, sh, lx - dl2 * nxv, ly - dl2 * nyv);
                const float ym1 = bilinear_f(A.y, sw, sh, lx - dl * nxv, ly - dl * nyv);
                const float yp1 = bilinear_f(A.y, sw, sh, lx + dl * nxv, ly + dl * nyv);
                const float yp2 = bilinear_f(A.y, sw, sh, lx + dl2 * nxv, ly + dl2 * nyv);
                const float df[4] = {ym1 - ym2, L0 - ym1, yp1 - L0, yp2 - yp1};
                for (int q = 0; q < 3; ++q) {
                    const float a = df[q], b = df[q + 1];
                    if (a > 0.0f && b < 0.0f)
                        ridge = fmx(ridge, fmn(a, -b));
                    if (a < 0.0f && b > 0.0f)
                        ridge = fmx(ridge, fmn(-a, b));
                }
            }
        }
        const float fx = scx - floorf(scx), fy = scy - floorf(scy);
        const float phx = fmn(fx, 1.0f - fx), phy = fmn(fy, 1.0f - fy);
        phase = (1.0f - 2.0f * phx) * (1.0f - 2.0f * phy);
        if (gp.ms_on) {
            const float gate = bilinear_fast(A.ms, cw, ch, px.cb_i0, px.cb_f, py.cb_i0, py.cb_f);
            mfade = 1.0f - gp.ms * (1.0f - gate);
        }
    }

    // --- pass 1: window luma range + chroma hull (recon.cpp:304-363)
    const int tx0 = px.start, ty0 = py.start;
    const int SX = SUP ? SUP : A.ax.sup, SY = SUP ? SUP : A.ay.sup;  // taps per axis
    float max_dl = 0.0f;
    float mnu = 1e30f, mxu = -1e30f, mnv = 1e30f, mxv = -1e30f;
    bool range_done = false;
    if constexpr (HULL_IN && SUP == 6 && WIN) {
        // every vertical tap of the wave active: the range over the active
        // taps is the maximum over the clamped column run (a repeated
        // active tap changes no maximum), read at fixed row offsets
        if (__all(py.bits == 0x3fu)) {
            int cx[6];
            tap_run(px.bits, cx, 1);
            const float4 *b0 = win + (ty0 - wy0) * kTStride + (tx0 - wx0);
            float m = 0.0f;
            _Pragma("unroll") for (int j = 0; j < 6; ++j) {
                float d[6];
                _Pragma("unroll") for (int i = 0; i < 6; ++i)
                    d[i] = b0[cx[i] + j * kTStride].x - L0;
                m = max3abs(m, d[0], d[1]);
                m = max3abs(m, d[2], d[3]);
                m = max3abs(m, d[4], d[5]);
            }
            max_dl = m;
            range_done = true;
        }
    }
    _Pragma("unroll") for (int j = 0; j < SY; ++j) {
        if (range_done)
            break;
        const bool onj = py.on(j);
        if (!SUP && !onj)
            continue;
        const size_t row = (size_t)clampi(ty0 + j, 0, ch - 1) * cw;
        _Pragma("unroll") for (int i = 0; i < SX; ++i) {
            const bool on = onj && px.on(i);
            if (!SUP && !on)
                continue;
            const float4 q = WIN ? win[(ty0 + j - wy0) * kTStride + (tx0 + i - wx0)]
                                 : A.luv[row + clampi(tx0 + i, 0, cw - 1)];
            if constexpr (HULL_IN) {
                // integer input: |q.x - L0| is finite and >= +0, max_dl
                // starts at +0
                max_dl = nn_max(max_dl, on ? fabsf(q.x - L0) : 0.0f);
            } else {
                const float m = fmx(max_dl, fabsf(q.x - L0));
                max_dl = on ? m : max_dl;
            }
            if constexpr (!HULL_IN) {
                const float a = q.y, b = q.z;
                const float nmu = fmn(mnu, a), nxu = fmx(mxu, a), nmv = fmn(mnv, b), nxv2 = fmx(mxv, b);
                mnu = on ? nmu : mnu;
                mxu = on ? nxu : mxu;
                mnv = on ? nmv : mnv;
                mxv = on ? nxv2 : mxv;
            }
        }
        if (SUP)
            __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (HULL_IN) {
        mnu = hull_in.x;
        mxu = hull_in.y;
        mnv = hull_in.z;
        mxv = hull_in.w;
    }
    if (!guided)
        max_dl = 0.0f;
    // The hull is read only at the end; the empty asm makes the compiler
    // finish it here rather than keep every window sample live until then.
    if (SUP)
        __asm__ volatile("" ::"v"(mnu), "v"(mxu), "v"(mnv), "v"(mxv), "v"(max_dl));

    // --- adaptive sigma and fades (recon.cpp:369-417)
    const float t2 = fclamp(min_diff / (0.1f * max_dl + 1e-6f), 0.0f, 1.0f);
    const float ss_ramp = t2 * t2 * (3.0f - 2.0f * t2);
    const float sig_mult = gp.sratio + (gp.sdb - gp.sratio) * ss_ramp;
    const float sig_loc = fmx(gp.sigma, sig_mult * max_dl);
    const float inv_s2 = 1.0f / (sig_loc * sig_loc);
    const float ridge_fade = 1.0f - fclamp(2.0f * ridge / (max_dl + 1e-6f), 0.0f, 1.0f);
    float cedge_fade = 1.0f;
    if (CEDGE && guided && has_dir) {
        const float rc = fmx(mxu - mnu, mxv - mnv);
        if (rc > 1e-6f) {
            float gmax = 0.0f;
            const float anx = fabsf(nxv), any = fabsf(nyv);
            _Pragma("unroll") for (int j = 0; j < SY; ++j) {
                const int ty = clampi(ty0 + j, 0, ch - 1), ty1 = min(ty + 1, ch - 1);
                _Pragma("unroll") for (int i = 0; i < SX; ++i) {
                    const int tx = clampi(tx0 + i, 0, cw - 1), tx1 = min(tx + 1, cw - 1);
                    const float c0u = A.u[(size_t)ty * cw + tx], c0v = A.v[(size_t)ty * cw + tx];
                    const float gu = fmx(fabsf(A.u[(size_t)ty * cw + tx1] - c0u) * anx,
                                         fabsf(A.u[(size_t)ty1 * cw + tx] - c0u) * any);
                    const float gv = fmx(fabsf(A.v[(size_t)ty * cw + tx1] - c0v) * anx,
                                         fabsf(A.v[(size_t)ty1 * cw + tx] - c0v) * any);
                    gmax = fmx(gmax, fmx(gu, gv));
                }
            }
            const float wc = rc / (gmax + 1e-6f);
            cedge_fade = fclamp((2.2f - wc) / 0.7f, 0.0f, 1.0f);
        }
    }
    const float guide_fade = ridge_fade * cedge_fade * mfade * gp.strength;

    // --- phase-0 rescue with kappa^2 (recon.cpp:425-475)
    float rescue = 0.0f;
    if (guided) {
        const int ncx = clampi((int)lroundf(scx), 0, cw - 1);
        const int ncy = clampi((int)lroundf(scy), 0, ch - 1);
        const float dl0 = A.lc[(size_t)ncy * cw + ncx] - L0;
        const float sim0 = 1.0f / (1.0f + dl0 * dl0 * inv_s2);
        rescue = gp.rescue * guide_fade * phase * phase * (1.0f - sim0);
        if (rescue > 0.0f) {
            float kappa;
            if (!RADIAL) {
                float mxX = 0, ssX = 0, mxY = 0, ssY = 0;
                _Pragma("unroll") for (int i = 0; i < SX; ++i) {
                    const bool on = px.on(i);
                    if (!SUP && !on)
                        continue;
                    const float w = px.wt(i);
                    const float m = fmx(mxX, fabsf(w)), s2 = ssX + w * w;
                    mxX = on ? m : mxX;
                    ssX = on ? s2 : ssX;
                }
                _Pragma("unroll") for (int j = 0; j < SY; ++j) {
                    const bool on = py.on(j);
                    if (!SUP && !on)
                        continue;
                    const float w = py.wt(j);
                    const float m = fmx(mxY, fabsf(w)), s2 = ssY + w * w;
                    mxY = on ? m : mxY;
                    ssY = on ? s2 : ssY;
                }
                kappa = (mxX * mxX * mxY * mxY) / (ssX * ssY + 1e-12f);
            } else {
                const float *base = radial_at(A.rd, ox, oy);
                float mx = 0, ss = 0;
                _Pragma("unroll") for (int j = 0; j < SY; ++j) {
                    if (!py.on(j))
                        continue;
                    _Pragma("unroll") for (int i = 0; i < SX; ++i) {
                        if (!px.on(i))
                            continue;
                        const float w = radial_weight(A.rd, base, SX, i, j, tx0, ty0, scx, scy);
                        mx = fmx(mx, fabsf(w));
                        ss += w * w;
                    }
                }
                kappa = (mx * mx) / (ss + 1e-12f);
            }
            rescue *= kappa * kappa;
        }
    }
    const float aniso = 1.0f + gp.stretch * coherence;
    const float inv_g2 = gp.inv_gsigma2;
    const float inv_g2p = inv_g2 / (aniso * aniso);

    // --- tap loop (recon.cpp:581-627)
    const TapIn ti{L0, nxv, nyv, scx, scy, inv_s2, guide_fade, inv_g2, inv_g2p, rescue, tx0, ty0};
    // Every active tap's reciprocal arguments are bounded by the window luma
    // range (max_dl) and the largest projected offset; when both bounds are
    // below 2^126 for the whole wave the branch-free reciprocal is exact.
    bool fast = true;
    if (guided) {
        const float ysm = 1.0f + max_dl * max_dl * inv_s2;
        const float dxm = fmx(fabsf((float)tx0 - scx), fabsf((float)(tx0 + SX - 1) - scx)) * gp.frw;
        const float dym = fmx(fabsf((float)ty0 - scy), fabsf((float)(ty0 + SY - 1) - scy)) * gp.frh;
        const float b = (dxm + dym) * (fabsf(nxv) + fabsf(nyv));
        const float ybm = 1.0f + 2.0f * b * b * (inv_g2 + inv_g2p);
        fast = ysm < 0x1p120f && ybm < 0x1p120f;
    }
    TapAcc acc;
    // guide_fade == 0 makes sim exactly 1 and rescue 0, so every weight is
    // exactly the base weight: the unguided sum (same order) gives the same
    // accumulators.  In the flat interior of the dilated trust mask the
    // mutual gate is 0 and whole waves take this path.  Both shortcuts need
    // finite samples (a NaN would otherwise reach the weights through the
    // similarity or the bump kernel).
    if (!guided || (gp.finite && __all(guide_fade == 0.0f && rescue == 0.0f)))
        guided_taps<RADIAL, SUP, WIN, true, false>(A, px, py, ox, oy, ti, win, wx0, wy0, acc);
    else if (gp.finite && __all(fast) && __all(rescue == 0.0f))
        guided_taps<RADIAL, SUP, WIN, true, true, true>(A, px, py, ox, oy, ti, win, wx0, wy0, acc);
    else if (__all(fast)) {
        // every vertical tap active for the whole wave (4:2:0 with centred
        // vertical siting, away from the borders): no weight masks
        if constexpr (SUP == 6 && !RADIAL && WIN) {
            if (gp.finite && __all(py.bits == 0x3fu))
                guided_taps<RADIAL, SUP, WIN, true, true, false, true>(A, px, py, ox, oy, ti, win, wx0, wy0, acc);
            else
                guided_taps<RADIAL, SUP, WIN, true, true>(A, px, py, ox, oy, ti, win, wx0, wy0, acc);
        } else {
            guided_taps<RADIAL, SUP, WIN, true, true>(A, px, py, ox, oy, ti, win, wx0, wy0, acc);
        }
    }
    else  // rare: exact reciprocals, weights from the tables, compact loop
        guided_taps<RADIAL, 0, WIN, false, true>(A, pix_axis<0>(A.ax, ox), pix_axis<0>(A.ay, oy), ox, oy, ti,
                                                 win, wx0, wy0, acc);
    double acc_u = acc.u, acc_v = acc.v, wsum = acc.wsum, sum_abs = acc.sabs;
    // --- temporal taps of TRecon (recon.cpp:642-745)
    if (TEMPORAL && guided && A.tmp.n > 0) {
        const TemporalDev &T = A.tmp;
        const int lxi = clampi((int)px.lpos, 0, sw - 1), lyi = clampi((int)py.lpos, 0, sh - 1);
        const float frw = gp.frw, frh = gp.frh;
        for (int q = 0; q < T.n; ++q) {
            const NbrDev nb = T.nbrs[q];
            const int bxx = min(nb.bw - 1, lxi / nb.block), byy = min(nb.bh - 1, lyi / nb.block);
            const size_t bi = (size_t)byy * nb.bw + bxx;
            const float tconf = gmem(nb.conf)[bi];
            if (tconf < 0.05f)
                continue;
            const int ivx = gmem(nb.mvx)[bi], ivy = gmem(nb.mvy)[bi];
            const float mx = (float)ivx, my = (float)ivy;
            // target sample, residual and kernel weight per (output column,
            // vector x) and (output row, vector y), tabulated per clip
            const int nvec = 2 * T.search + 1;
            const int jx = clampi(ivx + T.search, 0, nvec - 1), jy = clampi(ivy + T.search, 0, nvec - 1);
            const TempTap ex = T.tx[(size_t)ox * nvec + jx];
            const TempTap ey = T.ty[(size_t)oy * nvec + jy];
            const int tnx = ex.tn, tny = ey.tn;
            const float ddx = ex.dd, ddy = ey.dd;
            const bool stat = (fabsf(mx) + fabsf(my) <= 1.0f) && tconf > 0.5f;
            float dL;
            if (stat) {
                dL = gmem(nb.lc)[(size_t)tny * cw + tnx] - L0;
            } else {
                const double lx = (tnx + 0.5) * T.rw - 0.5 + T.sx - mx;
                const double ly = (tny + 0.5) * T.rh - 0.5 + T.sy - my;
                const int x0 = (int)ceil(lx - 0.5 * T.rw), x1 = (int)ceil(lx + 0.5 * T.rw);
                const int y0 = (int)ceil(ly - 0.5 * T.rh), y1 = (int)ceil(ly + 0.5 * T.rh);
                float md = 0.0f;
                for (int jj = y0; jj < y1; ++jj)
                    for (int ii = x0; ii < x1; ++ii)
                        md = fmx(md, fabsf(at(A.y, sw, sh, ii, jj) - L0));
                dL = md;
            }
            const float sig_t = stat ? sig_loc : 0.5f * sig_loc;
            const float sim_raw = 1.0f / (1.0f + dL * dL / (sig_t * sig_t));
            const float sim = stat ? 1.0f - guide_fade * (1.0f - sim_raw) : sim_raw;
            float kw;
            if (RADIAL) {
                const float dxc = ddx / frw * T.inv_wx, dyc = ddy / frh * T.inv_wy;
                kw = lut_lookup(A.rd.lut, A.rd.lut_n, sqrtf(dxc * dxc + dyc * dyc) * A.rd.lut_scale);
            } else {
                kw = (float)(ex.k * ey.k);
            }
            float nw = 1.0f;
            if (!stat) {
                nw = fmx(T.nsin[jx], T.nsin[nvec + jy]);
                nw *= 1.0f - ss_ramp;
            }
            const float w = tconf * sim * kw * nw;
            const size_t tk = (size_t)tny * cw + tnx;
            const float nu = gmem(nb.u)[tk], nv = gmem(nb.v)[tk];
            acc_u = dacc(acc_u, w, nu);
            acc_v = dacc(acc_v, w, nv);
            wsum += w;
            sum_abs += fabsf(w);
            mnu = fmn(mnu, nu); mxu = fmx(mxu, nu);
            mnv = fmn(mnv, nv); mxv = fmx(mxv, nv);
        }
    }
    double vu, vv;
    if (wsum > 1e-6 && wsum >= 0.05 * sum_abs) {
        vu = acc_u / wsum;
        vv = acc_v / wsum;
    } else {
        vu = bilinear(A.u, cw, ch, scx, scy);
        vv = bilinear(A.v, cw, ch, scx, scy);
    }
    if (gp.clamp_hull) {
        vu = dclamp(vu, (double)(mnu - gp.margin), (double)(mxu + gp.margin));
        vv = dclamp(vv, (double)(mnv - gp.margin), (double)(mxv + gp.margin));
    }
    const size_t o = (size_t)oy * A.ax.n + ox;
    const SelectorDev &S = A.sel;
    if (SEL != SEL_NONE) {
        // algo-4 selector weights (recon.cpp:762-811)
        const float diag = has_dir ? fmn(1.0f, 2.0f * fabsf(nxv * nyv)) : 0.0f;
        const float hedgy = (1.0f - ss_ramp) * guide_fade;
        const float w2 = hedgy * (1.0f - diag) * S.strength;
        const float w3 = hedgy * diag * S.strength;
        if (SEL == SEL_FLOAT) {
            if (sel_compressed(A.row_off, A.indices, A.ax.n, A.ay.n)) {
                const float bu = A.out_u[o], bv = A.out_v[o];
                if (w2 < 1e-4f && w3 < 1e-4f) {
                    S.du[k] = 0.0f;
                    S.dv[k] = 0.0f;
                    S.w3c[k] = 0.0f;
                } else {
                    S.du[k] = hedgy * (1.0f - diag) * ((float)vu - bu);
                    S.dv[k] = hedgy * (1.0f - diag) * ((float)vv - bv);
                    S.w3c[k] = hedgy * diag;
                }
            } else {
                S.w2p[o] = hedgy * (1.0f - diag);
                S.w3p[o] = hedgy * diag;
                S.gu[o] = (float)vu;
                S.gv[o] = (float)vv;
            }
            return;
        }
        // the plain base of the pixel: in `res` when the tile packs the
        // output itself, else in the output planes
        const float bu = res ? res->x : A.out_u[o], bv = res ? res->y : A.out_v[o];
        if (w2 >= 1e-4f || w3 >= 1e-4f) {
            const size_t nc = (size_t)cw * ch;
            const int cx = px.cb_i0, cy = py.cb_i0;
            const float cxf = px.cb_f, cyf = py.cb_f;
            const float au = bilinear_fast(S.lgf, cw, ch, cx, cxf, cy, cyf);
            const float bbu = bilinear_fast(S.lgf + nc, cw, ch, cx, cxf, cy, cyf);
            const float cu = bilinear_fast(S.lgf + 2 * nc, cw, ch, cx, cxf, cy, cyf);
            const float av = bilinear_fast(S.lgf + 3 * nc, cw, ch, cx, cxf, cy, cyf);
            const float bbv = bilinear_fast(S.lgf + 4 * nc, cw, ch, cx, cxf, cy, cyf);
            const float cv = bilinear_fast(S.lgf + 5 * nc, cw, ch, cx, cxf, cy, cyf);
            vu = bu + w2 * ((float)vu - bu) + w3 * cu * (au * L0 + bbu - bu);
            vv = bv + w2 * ((float)vv - bv) + w3 * cv * (av * L0 + bbv - bv);
        } else {
            vu = bu;
            vv = bv;
        }
    }
    if (res) {
        *res = make_float2((float)vu, (float)vv);
        return;
    }
    A.out_u[o] = (float)vu;
    A.out_v[o] = (float)vv;
}

template <bool RADIAL, bool TEMPORAL, int SEL, bool CEDGE, int SUP>
__global__ __launch_bounds__(kBlock, SUP == 6 ? (TEMPORAL ? 2 : 4) : 1) void k_guided(GuidedArgs A) {
    const int ow = A.ax.n, oh = A.ay.n;
    const uint32_t total = A.indices ? A.row_off[oh] : (uint32_t)ow * (uint32_t)oh;
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < total; k += gridDim.x * kBlock) {
        const uint32_t idx = A.indices ? A.indices[k] : k;
        guided_pixel<RADIAL, TEMPORAL, SEL, CEDGE, SUP>(A, (int)(idx % (uint32_t)ow), (int)(idx / (uint32_t)ow), k);
    }
}


// Hot path of the guided pass (6 taps per axis, separable, no cedge, no
// compressed selector): one wave per 8x8 output tile, lanes on the tile's
// pixels that are active in the trust mask, the tile's chroma-guide
// footprint staged in LDS.  The host enables it when every tile's footprint
// fits kTWin x kTWin.
//
// PLAIN: the wave also computes the plain separable base of its tile from
// the same window (every element as k_plain_fused computes it: horizontal
// pass of the 16 window rows into LDS, vertical pass per pixel) and stores
// it where the guided pass does not overwrite it; with a selector the
// guided pass reads it back as its base.
// One horizontal plain sum (plainChromaSeparable's row pass,
// recon.cpp:1151-1461) over six window samples: weighted U/V in double and
// the 1-D hulls of the active taps.  ALL: every tap is active, so the hull
// needs no masks (inactive taps carry zero weight in the axis records, so
// the sums need none either way).
template <bool ALL>
__device__ __forceinline__ void plain_h_row(const float4 *row, const PixAxis<6> &pc, float2 &val, float4 &hull) {
    double vu = 0.0, vv = 0.0;
    float mnu = 1e30f, mxu = -1e30f, mnv = 1e30f, mxv = -1e30f;
    _Pragma("unroll") for (int i = 0; i < 6; ++i) {
        const float4 q = row[i];
        const float a = q.y, b = q.z;
        vu = dacc(vu, pc.w[i], a);
        vv = dacc(vv, pc.w[i], b);
        const float nmu = fmn(mnu, a), nxu = fmx(mxu, a), nmv = fmn(mnv, b), nxv = fmx(mxv, b);
        const bool on = ALL || pc.on(i);
        mnu = on ? nmu : mnu;
        mxu = on ? nxu : mxu;
        mnv = on ? nmv : mnv;
        mxv = on ? nxv : mxv;
    }
    val = make_float2((float)vu, (float)vv);
    hull = make_float4(mnu, mxu, mnv, mxv);
}

// The vertical pass over six rows of horizontal results (column stride
// kTile in LDS).
template <bool ALL>
__device__ __forceinline__ void plain_v_col(const float2 *val, const float4 *hull, const PixAxis<6> &pr, double &vu,
                                            double &vv, float &mnu, float &mxu, float &mnv, float &mxv) {
    _Pragma("unroll") for (int j = 0; j < 6; ++j) {
        const float2 v = val[j * kTile];
        const float4 h = hull[j * kTile];
        vu = dacc(vu, pr.w[j], v.x);
        vv = dacc(vv, pr.w[j], v.y);
        const float nmu = fmn(mnu, h.x), nxu = fmx(mxu, h.y);
        const float nmv = fmn(mnv, h.z), nxv = fmx(mxv, h.w);
        const bool on = ALL || pr.on(j);
        mnu = on ? nmu : mnu;
        mxu = on ? nxu : mxu;
        mnv = on ? nmv : mnv;
        mxv = on ? nxv : mxv;
    }
}

// Samples of integer input (luma k * s, chroma k * s - 1/2: finite, never
// -0.0): the active taps of a 6-tap axis record form one run [lo, hi];
// reading tap clamp(i, lo, hi) in place of tap i repeats active samples,
// which changes no minimum or maximum, and the weight of an inactive tap is
// an exact zero, so the sums are unchanged too.  The hulls are then plain
// minimum / maximum chains over the six reads: for finite samples without
// -0.0 the minimum of a set is the same whichever way it is formed, so the
// three-operand VALU minimum / maximum give the ordered compares' result.
__device__ __forceinline__ float min3f(float a, float b, float c) {
    float r;
    __asm__("v_min3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ __forceinline__ float max3f(float a, float b, float c) {
    float r;
    __asm__("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ __forceinline__ void plain_h_row_fin(const float4 *row, const int off[6], const PixAxis<6> &pc,
                                                float2 &val, float4 &hull) {
    double vu = 0.0, vv = 0.0;
    float a[6], b[6];
    _Pragma("unroll") for (int i = 0; i < 6; ++i) {
        const float4 q = row[off[i]];
        a[i] = q.y;
        b[i] = q.z;
        vu = dacc(vu, pc.w[i], a[i]);
        vv = dacc(vv, pc.w[i], b[i]);
    }
    val = make_float2((float)vu, (float)vv);
    hull = make_float4(min3f(min3f(a[0], a[1], a[2]), a[3], min3f(a[4], a[5], a[5])),
                       max3f(max3f(a[0], a[1], a[2]), a[3], max3f(a[4], a[5], a[5])),
                       min3f(min3f(b[0], b[1], b[2]), b[3], min3f(b[4], b[5], b[5])),
                       max3f(max3f(b[0], b[1], b[2]), b[3], max3f(b[4], b[5], b[5])));
}

__device__ __forceinline__ void plain_v_col_fin(const float2 *val, const float4 *hull, const int off[6],
                                                const PixAxis<6> &pr, double &vu, double &vv, float &mnu,
                                                float &mxu, float &mnv, float &mxv) {
    float4 h[6];
    _Pragma("unroll") for (int j = 0; j < 6; ++j) {
        const float2 v = val[off[j]];
        h[j] = hull[off[j]];
        vu = dacc(vu, pr.w[j], v.x);
        vv = dacc(vv, pr.w[j], v.y);
    }
    mnu = min3f(min3f(h[0].x, h[1].x, h[2].x), h[3].x, min3f(h[4].x, h[5].x, h[5].x));
    mxu = max3f(max3f(h[0].y, h[1].y, h[2].y), h[3].y, max3f(h[4].y, h[5].y, h[5].y));
    mnv = min3f(min3f(h[0].z, h[1].z, h[2].z), h[3].z, min3f(h[4].z, h[5].z, h[5].z));
    mxv = max3f(max3f(h[0].w, h[1].w, h[2].w), h[3].w, max3f(h[4].w, h[5].w, h[5].w));
}

// OUTT: 0 = results to out_u/out_v as floats; otherwise the tile also
// applies back-projection (A.bp_on) and writes the output planes itself
// (1 = 8-bit, 2 = 16-bit, 3 = float samples).
// PH (with PLAIN, finite samples only): the guided pass takes the chroma
// hull of each pixel from the plain base's vertical pass.
// One wave computes tile t (wave-uniform) of the frame described by A, using
// its own LDS window wb and plain-base rows hv/hh.
template <bool TEMPORAL, int SEL, bool PLAIN, int OUTT, bool PH>
__device__ __attribute__((always_inline)) void guided_tile(const GuidedArgs &A, int t, int tiles_x, float4 *wb,
                                                           float2 (*hv)[kTile], float4 (*hh)[kTile], int lane) {
    static_assert(!PH || PLAIN, "the plain hull comes from the plain base in the tile");
    // fused output needs every inside pixel's result in the tile: the plain
    // base in the tiles (Recon), or a dense temporal pass (TRecon, no mask)
    static_assert(OUTT == 0 || (SEL != SEL_FLOAT && PLAIN != TEMPORAL), "fused output needs every pixel's result");
    const int ow = A.ax.n, oh = A.ay.n, cw = A.cw, ch = A.ch;
    const int tx = t % tiles_x, ty = t / tiles_x;
    const int ox = tx * kTile + (lane & (kTile - 1)), oy = ty * kTile + (lane / kTile);
    const bool inside = ox < ow && oy < oh;
    float2 res = make_float2(0.0f, 0.0f);  // this pixel's result (fused output)
    float4 phull = make_float4(1e30f, -1e30f, 1e30f, -1e30f);
    bool act = inside;
    if (act && A.mbits) {
        // nearest mask sample (buildSparseWorkset, maps.cpp:342-360); the
        // mask has the output size unless the clip resizes
        const bool same = A.mw == ow && A.mh == oh;
        const int mx = same ? ox : min(A.mw - 1, (int)((int64_t)ox * A.mw / ow));
        const int my = same ? oy : min(A.mh - 1, (int)((int64_t)oy * A.mh / oh));
        act = (A.mbits[(uint32_t)(my * A.mw64 + (mx >> 6))] >> (mx & 63)) & 1ull;
    }
    const bool any = __ballot(act) != 0;
    if (!PLAIN && !any)
        return;  // no active pixel in the tile (uniform)
    int x_lo = INT_MAX, y_lo = INT_MAX;
    for (int c = 0; c < kTile; ++c) {
        x_lo = min(x_lo, A.ax.start[min(tx * kTile + c, ow - 1)]);
        y_lo = min(y_lo, A.ay.start[min(ty * kTile + c, oh - 1)]);
    }
    // the whole kTWin x kTWin window from (x_lo, y_lo); the host checked that
    // every tile's taps lie inside it (guided_tiles_fit)
    _Pragma("unroll") for (int e = lane; e < kTWin * kTWin; e += 64) {
        const int r = e / kTWin, c = e % kTWin;
        wb[r * kTStride + c] = A.luv[(uint32_t)(clampi(y_lo + r, 0, ch - 1) * cw + clampi(x_lo + c, 0, cw - 1))];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    if constexpr (PLAIN) {
        // horizontal pass: lane -> column lane % 8, window rows lane / 8 and
        // lane / 8 + 8 (columns past the right edge repeat the last one)
        const int c8 = lane & (kTile - 1);
        const PixAxis<6> pc = pix_axis<6>(A.ax, min(ox, ow - 1));
        const int col0 = pc.start - x_lo;
        // inactive taps carry zero weight in the axis records; when every
        // lane's six taps are active (the interior) the hulls need no masks
        if constexpr (PH) {
            int hx[6];
            tap_run(pc.bits, hx, 1);
            _Pragma("unroll") for (int k = 0; k < 2; ++k)
                plain_h_row_fin(wb + ((lane >> 3) + 8 * k) * kTStride + col0, hx, pc,
                                hv[(lane >> 3) + 8 * k][c8], hh[(lane >> 3) + 8 * k][c8]);
        } else if (__all(pc.bits == 0x3fu)) {
            _Pragma("unroll") for (int k = 0; k < 2; ++k)
                plain_h_row<true>(wb + ((lane >> 3) + 8 * k) * kTStride + col0, pc, hv[(lane >> 3) + 8 * k][c8],
                                  hh[(lane >> 3) + 8 * k][c8]);
        } else {
            _Pragma("unroll") for (int k = 0; k < 2; ++k)
                plain_h_row<false>(wb + ((lane >> 3) + 8 * k) * kTStride + col0, pc,
                                   hv[(lane >> 3) + 8 * k][c8], hh[(lane >> 3) + 8 * k][c8]);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        // vertical pass, normalisation, fallback and hull clamp
        // (plainChromaSeparable, recon.cpp:1151-1461)
        const PixAxis<6> pr = pix_axis<6>(A.ay, min(oy, oh - 1));
        const bool vall = __all(pr.bits == 0x3fu);
        if (inside) {
            const int t0 = pr.start - y_lo;
            double vu = 0.0, vv = 0.0;
            float mnu = 1e30f, mxu = -1e30f, mnv = 1e30f, mxv = -1e30f;
            if constexpr (PH) {
                int hy[6];
                tap_run(pr.bits, hy, kTile);
                plain_v_col_fin(&hv[t0][c8], &hh[t0][c8], hy, pr, vu, vv, mnu, mxu, mnv, mxv);
            } else if (vall) {
                plain_v_col<true>(&hv[t0][c8], &hh[t0][c8], pr, vu, vv, mnu, mxu, mnv, mxv);
            } else {
                plain_v_col<false>(&hv[t0][c8], &hh[t0][c8], pr, vu, vv, mnu, mxu, mnv, mxv);
            }
            phull = make_float4(mnu, mxu, mnv, mxv);
            const double wsum = (double)pc.wsum * pr.wsum;
            const double asum = (double)pc.wabs * pr.wabs;
            if (wsum > 1e-6 && wsum >= 0.05 * asum) {
                vu /= wsum;
                vv /= wsum;
            } else {
                vu = bilinear_f(A.u, cw, ch, pc.pos, pr.pos);
                vv = bilinear_f(A.v, cw, ch, pc.pos, pr.pos);
            }
            if (A.gp.clamp_hull) {
                vu = dclamp(vu, (double)(mnu - A.gp.margin), (double)(mxu + A.gp.margin));
                vv = dclamp(vv, (double)(mnv - A.gp.margin), (double)(mxv + A.gp.margin));
            }
            if (OUTT != 0) {
                res = make_float2((float)vu, (float)vv);
            } else if (!act || SEL != SEL_NONE) {
                const uint32_t o = (uint32_t)(oy * ow + ox);
                A.out_u[o] = (float)vu;
                A.out_v[o] = (float)vv;
            }
        }
        if (OUTT == 0 && !any)
            return;
    }
    if (act)
        guided_pixel<false, TEMPORAL, SEL, false, 6, true, PH>(A, ox, oy, 0u, wb, x_lo, y_lo,
                                                               OUTT != 0 ? &res : nullptr, phull);
    if constexpr (OUTT != 0) {
        // back-projection (backProject, algos.cpp:663-677): same-size 4:2:0
        // with an even output, so the 2x2 block of chroma sample (ox/2,
        // oy/2) is lanes l, l^1, l^8, l^9 of this tile.  The block's top-left
        // lane forms p00 + p10 + p01 + p11 in that order and the other three
        // read its sum.
        float fu = res.x, fv = res.y;
        if (A.bp_on) {
            float su = ((fu + __shfl_xor(fu, 1)) + __shfl_xor(fu, 8)) + __shfl_xor(fu, 9);
            float sv = ((fv + __shfl_xor(fv, 1)) + __shfl_xor(fv, 8)) + __shfl_xor(fv, 9);
            su = __shfl(su, lane & ~9);
            sv = __shfl(sv, lane & ~9);
            if (inside) {
                const uint32_t ci = (uint32_t)((oy >> 1) * cw + (ox >> 1));
                fu += A.bp * (A.u[ci] - 0.25f * su);
                fv += A.bp * (A.v[ci] - 0.25f * sv);
            }
        }
        if (inside) {
            ptrdiff_t su, sv;
            char *bu = plane_out(A.po_u, su), *bv = plane_out(A.po_v, sv);
            if constexpr (OUTT == 1) {
                reinterpret_cast<uint8_t *>(bu + su * oy)[ox] = pack_sample<uint8_t>(fu, A.po_u);
                reinterpret_cast<uint8_t *>(bv + sv * oy)[ox] = pack_sample<uint8_t>(fv, A.po_v);
            } else if constexpr (OUTT == 2) {
                reinterpret_cast<uint16_t *>(bu + su * oy)[ox] = pack_sample<uint16_t>(fu, A.po_u);
                reinterpret_cast<uint16_t *>(bv + sv * oy)[ox] = pack_sample<uint16_t>(fv, A.po_v);
            } else {
                reinterpret_cast<float *>(bu + su * oy)[ox] = fu;
                reinterpret_cast<float *>(bv + sv * oy)[ox] = fv;
            }
        }
    }
}

template <bool TEMPORAL, int SEL, bool PLAIN, int OUTT = 0, bool PH = false>
__global__ __launch_bounds__(kBlock, TEMPORAL ? 4 : (PLAIN ? 5 : 6)) void k_guided_tiles(GuidedArgs A, int tiles_x,
                                                                                       int tiles) {
    __shared__ float4 wins[kBlock / 64][kTWin * kTStride];
    __shared__ float2 hval[PLAIN ? kBlock / 64 : 1][kTWin][kTile];
    __shared__ float4 hhull[PLAIN ? kBlock / 64 : 1][kTWin][kTile];
    const int lane = threadIdx.x & 63;
    // wave-uniform tile index, held in scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int t = blockIdx.x * (kBlock / 64) + wave;
    if (t >= tiles)
        return;  // whole wave
    const int hw = PLAIN ? wave : 0;
    guided_tile<TEMPORAL, SEL, PLAIN, OUTT, PH>(A, t, tiles_x, wins[wave], hval[hw], hhull[hw], lane);
}

// The fused tile pass over the tiles of many frames in one launch.
// stride == 0: one wave per tile, frame blockIdx.y, its tiles over
// blockIdx.x, the hardware handing out the workgroups as CUs free up.
// Persistent form (stride > 0): frame f's tiles are positions first[f] ..
// first[f + 1] - 1 (first[] ascending, first[0] = 0, at most 64 frames), the
// grid is what the device holds at once and wave g computes positions g,
// g + stride, g + 2 stride, ... (stride = the grid's wave count), so
// neighbouring positions run on the waves of one workgroup and the frames'
// tiles spread evenly over the CUs and XCDs; every wave leaves after its
// last position.
template <bool TEMPORAL, int SEL, bool PLAIN, int OUTT = 0, bool PH = false>
__global__ __launch_bounds__(kBlock, TEMPORAL ? 4 : (PLAIN ? 5 : 6)) void k_guided_tiles_batch(
    const BatchFrame *__restrict__ frames_in, const uint32_t *__restrict__ first, int nframes, uint32_t total,
    uint32_t stride) {
    // the frame table is read-only for the whole launch: reading it through
    // the constant address space keeps its loads scalar and the plane
    // pointers in it known to be global
    const __attribute__((address_space(4))) BatchFrame *frames =
        (const __attribute__((address_space(4))) BatchFrame *)frames_in;
    __shared__ float4 wins[kBlock / 64][kTWin * kTStride];
    __shared__ float2 hval[PLAIN ? kBlock / 64 : 1][kTWin][kTile];
    __shared__ float4 hhull[PLAIN ? kBlock / 64 : 1][kTWin][kTile];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int hw = PLAIN ? wave : 0;
    // lane f holds frame f's first position (lanes past the batch: never
    // below any position)
    if (stride == 0) {
        // one wave per tile: frame blockIdx.y, its tiles over blockIdx.x
        const int f = blockIdx.y;
        const int t = blockIdx.x * (kBlock / 64) + wave;
        if (t >= frames[f].tiles)
            return;
        const GuidedArgs &A = *(const GuidedArgs *)&frames[f].A;
        guided_tile<TEMPORAL, SEL, PLAIN, OUTT, PH>(A, t, frames[f].tiles_x, wins[wave], hval[hw], hhull[hw], lane);
        return;
    }
    const uint32_t lfirst = lane < nframes ? first[lane] : 0xffffffffu;
    for (uint32_t t = blockIdx.x * (kBlock / 64) + wave; t < total; t += stride) {
        const int f = __builtin_amdgcn_readfirstlane((int)__popcll(__ballot(lfirst <= t)) - 1);
        const uint32_t f0 = __builtin_amdgcn_readfirstlane(__shfl(lfirst, f));
        // the previous tile's LDS reads have completed before its window is
        // overwritten
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        const GuidedArgs &A = *(const GuidedArgs *)&frames[f].A;
        guided_tile<TEMPORAL, SEL, PLAIN, OUTT, PH>(A, (int)(t - f0), frames[f].tiles_x, wins[wave], hval[hw],
                                                    hhull[hw], lane);
    }
}

// ---------------------------------------------------------------------------
// algo 4: LGF regression maps and selector blend (algos.cpp:241-593)
// ---------------------------------------------------------------------------

// pointSampledLuma (algos.cpp:300-315)
__global__ __launch_bounds__(kBlock) void k_point_luma(const float *y, int w, int h, float *ls, int cw, int ch,
                                                       const int *xi, const float *xf, const int *yi,
                                                       const float *yf) {
    const int x = blockIdx.x * kBlock + threadIdx.x, j = blockIdx.y;
    if (x >= cw)
        return;
    ls[(size_t)j * cw + x] = bilinear_fast(y, w, h, xi[x], xf[x], yi[j], yf[j]);
}

// slidingMinMax, radius 2 (algos.cpp:70-139): in-bounds 5x5 window
__global__ __launch_bounds__(kBlock) void k_minmax5(const float *u, const float *v, int cw, int ch, float *mm) {
    const int x = blockIdx.x * kBlock + threadIdx.x, y = blockIdx.y;
    if (x >= cw)
        return;
    const size_t n = (size_t)cw * ch, i = (size_t)y * cw + x;
    for (int c = 0; c < 2; ++c) {
        const float *p = c ? v : u;
        float lo = p[(size_t)max(0, y - 2) * cw + max(0, x - 2)], hi = lo;
        for (int j = max(0, y - 2); j <= min(ch - 1, y + 2); ++j)
            for (int k = max(0, x - 2); k <= min(cw - 1, x + 2); ++k) {
                const float q = p[(size_t)j * cw + k];
                lo = fmn(lo, q);
                hi = fmx(hi, q);
            }
        mm[(2 * c) * n + i] = lo;
        mm[(2 * c + 1) * n + i] = hi;
    }
}

// buildLGFPair finalize (algos.cpp:381-414) over double 5x5 window sums,
// plus the mutual-structure confidence fade (plugin.cpp:418-425, 543-551).
__global__ __launch_bounds__(kBlock) void k_lgf(const float *__restrict__ ls, const float *__restrict__ u,
                                                const float *__restrict__ v, int cw, int ch, double eps, int cedge,
                                                const float *mm, const uint8_t *roi, const uint32_t *roi_count,
                                                const float *gate, float ms, float *lgf) {
    const int x = blockIdx.x * kBlock + threadIdx.x, y = blockIdx.y;
    if (x >= cw)
        return;
    const size_t n = (size_t)cw * ch, i = (size_t)y * cw + x;
    const bool dense = !roi || (uint64_t)roi_count[0] * 2 >= (uint64_t)n;
    if (!dense && !roi[i]) {
        for (int c = 0; c < 6; ++c)
            lgf[c * n + i] = 0.0f;
        return;
    }
    const int x0 = max(0, x - 2), x1 = min(cw - 1, x + 2), y0 = max(0, y - 2), y1 = min(ch - 1, y + 2);
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
    for (int j = y0; j <= y1; ++j)
        for (int k = x0; k <= x1; ++k) {
            const size_t q = (size_t)j * cw + k;
            const double l = ls[q], a = u[q], b = v[q];
            s0 += l; s1 += l * l; s2 += a; s3 += l * a; s4 += b; s5 += l * b;
        }
    const int cnt = (x1 - x0 + 1) * (y1 - y0 + 1);
    const double mean_y = s0 / cnt;
    const double var = s1 / cnt - mean_y * mean_y;
    const double mean_u = s2 / cnt, mean_v = s4 / cnt;
    const double su = (s3 / cnt - mean_y * mean_u) / (var + eps);
    const double sv = (s5 / cnt - mean_y * mean_v) / (var + eps);
    const float base = (float)(var / (var + eps));
    float fu = 1.0f, fv = 1.0f;
    if (cedge) {
        const float lgx = (at(ls, cw, ch, x + 1, y) - at(ls, cw, ch, x - 1, y)) * 0.5f;
        const float lgy = (at(ls, cw, ch, x, y + 1) - at(ls, cw, ch, x, y - 1)) * 0.5f;
        const float gl = hypotf(lgx, lgy);
        for (int c = 0; c < 2; ++c) {
            const float *P = c ? v : u;
            const float range = mm[(2 * c + 1) * n + i] - mm[(2 * c) * n + i];
            float f = 1.0f;
            if (!(gl <= 1e-9f || range <= 1e-6f)) {
                const float cgx = (at(P, cw, ch, x + 1, y) - at(P, cw, ch, x - 1, y)) * 0.5f;
                const float cgy = (at(P, cw, ch, x, y + 1) - at(P, cw, ch, x, y - 1)) * 0.5f;
                const float gn = fabsf(cgx * lgx / gl + cgy * lgy / gl);
                f = fclamp((2.2f - range / (gn + 1e-6f)) / 0.7f, 0.0f, 1.0f);
            }
            if (c) fv = f; else fu = f;
        }
    }
    float cu = base * fu, cv = base * fv;
    if (gate && (!roi || roi[i])) {
        const float f = 1.0f - ms * (1.0f - gate[i]);
        cu *= f;
        cv *= f;
    }
    lgf[i] = (float)su;
    lgf[n + i] = (float)(mean_u - su * mean_y);
    lgf[2 * n + i] = cu;
    lgf[3 * n + i] = (float)sv;
    lgf[4 * n + i] = (float)(mean_v - sv * mean_y);
    lgf[5 * n + i] = cv;
}

struct WorkList { const uint32_t *row_off; const uint32_t *indices; int ow, oh; };

__device__ __forceinline__ uint32_t work_total(const WorkList &L) {
    return L.indices ? L.row_off[L.oh] : (uint32_t)L.ow * (uint32_t)L.oh;
}

// ROI of the diagonal route (plugin.cpp:496-531)
__global__ __launch_bounds__(kBlock) void k_roi_mark(WorkList L, SelectorDev S, AxisDev ax, AxisDev ay,
                                                     uint8_t *roi, int cw, int ch) {
    const uint32_t total = work_total(L);
    const bool comp = sel_compressed(L.row_off, L.indices, L.ow, L.oh);
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < total; k += gridDim.x * kBlock) {
        const uint32_t idx = L.indices ? L.indices[k] : k;
        const float w3 = comp ? S.w3c[k] : S.w3p[idx];
        if (w3 < 1e-4f)
            continue;
        const int ox = (int)(idx % (uint32_t)L.ow), oy = (int)(idx / (uint32_t)L.ow);
        const int cx0 = ax.cb_i0[ox], cx1 = min(cx0 + 1, cw - 1);
        const int cy0 = ay.cb_i0[oy], cy1 = min(cy0 + 1, ch - 1);
        roi[(size_t)cy0 * cw + cx0] = 1;
        roi[(size_t)cy0 * cw + cx1] = 1;
        roi[(size_t)cy1 * cw + cx0] = 1;
        roi[(size_t)cy1 * cw + cx1] = 1;
    }
}

// blendSelectorCompressed / blendSelector (algos.cpp:453-593)
__global__ __launch_bounds__(kBlock) void k_blend4(WorkList L, const float *y, int sw, int sh, AxisDev ax,
                                                   AxisDev ay, int cw, int ch, SelectorDev S, float *out_u,
                                                   float *out_v) {
    const uint32_t total = work_total(L);
    const bool comp = sel_compressed(L.row_off, L.indices, L.ow, L.oh);
    const size_t n = (size_t)cw * ch;
    const float lam = S.strength;
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < total; k += gridDim.x * kBlock) {
        const uint32_t idx = L.indices ? L.indices[k] : k;
        const int ox = (int)(idx % (uint32_t)L.ow), oy = (int)(idx / (uint32_t)L.ow);
        const int cx = ax.cb_i0[ox], cy = ay.cb_i0[oy];
        const float cxf = ax.cb_f[ox], cyf = ay.cb_f[oy];
        if (comp) {
            const float pu = out_u[idx], pv = out_v[idx];
            const float w3 = lam * S.w3c[k];
            float vu = pu + lam * S.du[k], vv = pv + lam * S.dv[k];
            if (w3 != 0.0f) {
                const float luma = bilinear_fast(y, sw, sh, ax.lb_i0[ox], ax.lb_f[ox], ay.lb_i0[oy], ay.lb_f[oy]);
                const float au = bilinear_fast(S.lgf, cw, ch, cx, cxf, cy, cyf);
                const float bu = bilinear_fast(S.lgf + n, cw, ch, cx, cxf, cy, cyf);
                const float cu = bilinear_fast(S.lgf + 2 * n, cw, ch, cx, cxf, cy, cyf);
                const float av = bilinear_fast(S.lgf + 3 * n, cw, ch, cx, cxf, cy, cyf);
                const float bv = bilinear_fast(S.lgf + 4 * n, cw, ch, cx, cxf, cy, cyf);
                const float cv = bilinear_fast(S.lgf + 5 * n, cw, ch, cx, cxf, cy, cyf);
                vu += w3 * cu * (au * luma + bu - pu);
                vv += w3 * cv * (av * luma + bv - pv);
            }
            out_u[idx] = vu;
            out_v[idx] = vv;
        } else {
            const float w2v = S.w2p[idx] * lam, w3v = S.w3p[idx] * lam;
            if (w2v < 1e-4f && w3v < 1e-4f)
                continue;
            const float L0 = bilinear_fast(y, sw, sh, ax.lb_i0[ox], ax.lb_f[ox], ay.lb_i0[oy], ay.lb_f[oy]);
            const float au = bilinear_fast(S.lgf, cw, ch, cx, cxf, cy, cyf);
            const float bu = bilinear_fast(S.lgf + n, cw, ch, cx, cxf, cy, cyf);
            const float cu = bilinear_fast(S.lgf + 2 * n, cw, ch, cx, cxf, cy, cyf);
            const float av = bilinear_fast(S.lgf + 3 * n, cw, ch, cx, cxf, cy, cyf);
            const float bv = bilinear_fast(S.lgf + 4 * n, cw, ch, cx, cxf, cy, cyf);
            const float cv = bilinear_fast(S.lgf + 5 * n, cw, ch, cx, cxf, cy, cyf);
            const float ru = out_u[idx], rv = out_v[idx];
            out_u[idx] = ru + (w2v * (S.gu[idx] - ru) + w3v * cu * (au * L0 + bu - ru));
            out_v[idx] = rv + (w2v * (S.gv[idx] - rv) + w3v * cv * (av * L0 + bv - rv));
        }
    }
}


// ---------------------------------------------------------------------------
// algo 6: affine credibility maps (buildAffineMaps, algos.cpp:1382-1526),
// Nyquist detail (algos.cpp:1139-1380), detail transfer (algos.cpp:1587-1689)
// ---------------------------------------------------------------------------

__device__ __forceinline__ float median3(float a, float b, float c) {
    const float mn = fmn(fmn(a, b), c), mx = fmx(fmx(a, b), c);
    return a + b + c - mn - mx;
}

__device__ __forceinline__ double dmax0(double x) { return 0.0 < x ? x : 0.0; }

__global__ __launch_bounds__(kBlock) void k_affine(const float *__restrict__ u, const float *__restrict__ v, int cw,
                                                   int ch, const float *__restrict__ yc0,
                                                   const float *__restrict__ yc1, const float *__restrict__ yc2,
                                                   const uint8_t *cmask, const uint32_t *ccount, double eps,
                                                   double eps_sig, AffineDev m) {
    const int x = blockIdx.x * kBlock + threadIdx.x, y = blockIdx.y;
    if (x >= cw)
        return;
    const size_t n = (size_t)cw * ch, i = (size_t)y * cw + x;
    m.deg[i] = median3(yc0[i], yc1[i], yc2[i]);
    // sparse ROI: chroma mask present and not in dense fallback (algos.cpp:1414-1415)
    const bool sparse = cmask && (uint64_t)ccount[0] * 2 < (uint64_t)n;
    if (sparse && !cmask[i]) {
        m.au[i] = m.av[i] = m.g[i] = m.mnu[i] = m.mxu[i] = m.mnv[i] = m.mxv[i] = m.rng[i] = 0.0f;
        return;
    }
    const int top = max(0, y - 2), bot = min(ch - 1, y + 2), lef = max(0, x - 2), rig = min(cw - 1, x + 2);
    const float *ycs[3] = {yc0, yc1, yc2};
    double sum[16];
    for (int k = 0; k < 16; ++k)
        sum[k] = 0.0;
    float hmnu = u[(size_t)top * cw + lef], hmxu = hmnu, hmnv = v[(size_t)top * cw + lef], hmxv = hmnv;
    for (int py = top; py <= bot; ++py)
        for (int px = lef; px <= rig; ++px) {
            const size_t q = (size_t)py * cw + px;
            const double a = u[q], b = v[q];
            sum[0] += a; sum[1] += a * a; sum[2] += b; sum[3] += b * b;
            hmnu = fmn(hmnu, (float)a); hmxu = fmx(hmxu, (float)a);
            hmnv = fmn(hmnv, (float)b); hmxv = fmx(hmxv, (float)b);
            for (int k = 0; k < 3; ++k) {
                const double l = ycs[k][q];
                sum[4 + 4 * k] += l;
                sum[5 + 4 * k] += l * l;
                sum[6 + 4 * k] += l * a;
                sum[7 + 4 * k] += l * b;
            }
        }
    const int cnt = (rig - lef + 1) * (bot - top + 1);
    const double inv = 1.0 / cnt;
    // sparse path multiplies by 1/n (algos.cpp:1082-1102); dense path divides
    // (algos.cpp:1483-1497)
    auto mean = [&](double s) { return sparse ? s * inv : s / cnt; };
    const double meanU = mean(sum[0]), meanV = mean(sum[2]);
    const double varU = dmax0(mean(sum[1]) - meanU * meanU);
    const double varV = dmax0(mean(sum[3]) - meanV * meanV);
    double qk[3], aUk[3], aVk[3];
    for (int k = 0; k < 3; ++k) {
        const int b = 4 + 4 * k;
        const double meanY = mean(sum[b]);
        const double varY = dmax0(mean(sum[b + 1]) - meanY * meanY);
        const double covU = mean(sum[b + 2]) - meanY * meanU;
        const double covV = mean(sum[b + 3]) - meanY * meanV;
        aUk[k] = dclamp(covU / (varY + eps), -16.0, 16.0);
        aVk[k] = dclamp(covV / (varY + eps), -16.0, 16.0);
        qk[k] = dclamp((covU * covU + covV * covV) / ((varY + eps) * (varU + varV + eps)), 0.0, 1.0);
    }
    m.mnu[i] = hmnu; m.mxu[i] = hmxu; m.mnv[i] = hmnv; m.mxv[i] = hmxv;
    m.rng[i] = fmx(hmxu - hmnu, hmxv - hmnv);
    m.au[i] = median3((float)aUk[0], (float)aUk[1], (float)aUk[2]);
    m.av[i] = median3((float)aVk[0], (float)aVk[1], (float)aVk[2]);
    const double qmax = fmax(fmax(qk[0], qk[1]), qk[2]);
    const double qmin = fmin(fmin(qk[0], qk[1]), qk[2]);
    const double stab = qmax > 1e-9 ? qmin / qmax : 0.0;
    const double sig = (varU + varV) / (varU + varV + eps_sig);
    m.g[i] = (float)(((qk[0] + qk[1] + qk[2]) / 3.0) * stab * sig);
}

__global__ __launch_bounds__(kBlock) void k_detail_sub(const float *y, float *rec, size_t n) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        rec[i] = y[i] - rec[i];
}

// one dyadic [1/4, 1/2, 1/4] pass, edge clamped (algos.cpp:1195-1217, 1252-1278)
__global__ __launch_bounds__(kBlock) void k_nyq(const float *src, float *dst, int w, int h, int step, int vertical) {
    const int x = blockIdx.x * kBlock + threadIdx.x, y = blockIdx.y;
    if (x >= w)
        return;
    float a, b, c;
    if (vertical) {
        a = src[(size_t)max(0, y - step) * w + x];
        b = src[(size_t)y * w + x];
        c = src[(size_t)min(h - 1, y + step) * w + x];
    } else {
        const float *r = src + (size_t)y * w;
        a = r[max(0, x - step)];
        b = r[x];
        c = r[min(w - 1, x + step)];
    }
    dst[(size_t)y * w + x] = 0.25f * a + 0.5f * b + 0.25f * c;
}

__global__ __launch_bounds__(kBlock) void k_detail_transfer(WorkList L, const float4 *yt, int sw, int sh,
                                                            AxisDev ax, AxisDev ay,
                                                            int cw, int ch, AffineDev m, const float *ms,
                                                            const float *detail, float mss, float qs, float st,
                                                            float stretch, float ar, int clamp_hull, float *out_u,
                                                            float *out_v) {
    const uint32_t total = work_total(L);
    const int ow = L.ow, oh = L.oh;
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < total; k += gridDim.x * kBlock) {
        const uint32_t idx = L.indices ? L.indices[k] : k;
        const int ox = (int)(idx % (uint32_t)ow), oy = (int)(idx / (uint32_t)ow);
        const int lxi = ax.lb_i0[ox], lyi = ay.lb_i0[oy];
        const float lxf = ax.lb_f[ox], lyf = ay.lb_f[oy];
        const float4 t = bilinear_fast4(yt, sw, sh, lxi, lxf, lyi, lyf);
        const Dir d = principal_dir(t.y, t.z, t.w);
        const int cxi = ax.cb_i0[ox], cyi = ay.cb_i0[oy];
        const float cxf = ax.cb_f[ox], cyf = ay.cb_f[oy];
        const float q = bilinear_fast(m.g, cw, ch, cxi, cxf, cyi, cyf);
        float g = (1.0f - qs * (1.0f - q)) * st;
        if (ms)
            g *= 1.0f - mss * (1.0f - bilinear_fast(ms, cw, ch, cxi, cxf, cyi, cyf));
        const float au = bilinear_fast(m.au, cw, ch, cxi, cxf, cyi, cyf);
        const float av = bilinear_fast(m.av, cw, ch, cxi, cxf, cyi, cyf);
        const float coh = d.coherence;
        const float tx = -d.ny, ty = d.nx;
        const float dc = detail[idx];
        const float s = 1.0f + stretch * coh;
        const float dtp = bilinear(detail, ow, oh, (float)ox + tx * s, (float)oy + ty * s);
        const float dtm = bilinear(detail, ow, oh, (float)ox - tx * s, (float)oy - ty * s);
        const float dt = 0.25f * dtp + 0.5f * dc + 0.25f * dtm;
        const float det = dc + (dt - dc) * coh;
        const float cap = 0.5f * bilinear_fast(m.rng, cw, ch, cxi, cxf, cyi, cyf);
        const float cu = fclamp(g * au * det, -cap, cap);
        const float cv = fclamp(g * av * det, -cap, cap);
        const float bu = out_u[idx], bv = out_v[idx];
        float ou = bu + cu, ov = bv + cv;
        if (clamp_hull) {
            const float lou = fmn(bu, bilinear_fast(m.mnu, cw, ch, cxi, cxf, cyi, cyf) - ar);
            const float hiu = fmx(bu, bilinear_fast(m.mxu, cw, ch, cxi, cxf, cyi, cyf) + ar);
            const float lov = fmn(bv, bilinear_fast(m.mnv, cw, ch, cxi, cxf, cyi, cyf) - ar);
            const float hiv = fmx(bv, bilinear_fast(m.mxv, cw, ch, cxi, cxf, cyi, cyf) + ar);
            ou = fclamp(ou, lou, hiu);
            ov = fclamp(ov, lov, hiv);
        }
        out_u[idx] = ou;
        out_v[idx] = ov;
    }
}


// ---------------------------------------------------------------------------
// TRecon: integer-pel block matching (blockMatch, algos.cpp:683-755) and the
// chroma-consistency confidence (applyChromaConsistency, algos.cpp:757-807).
// One workgroup per block; each thread scores candidates sequentially in the
// reference's (j, i) order and the argmin reduces on (sad, |dx|+|dy|, scan
// index), which is exactly the first-wins scan of the reference.
// ---------------------------------------------------------------------------

constexpr int kMeThreads = 256;

template <bool GROUP>
__global__ __launch_bounds__(kMeThreads) void k_block_match(const float *__restrict__ cur, const float *__restrict__ nbr,
                                                            int w, int h, int block, int search, float tsad,
                                                            int16_t *mvx, int16_t *mvy, float *conf, GroupOff go) {
    if constexpr (GROUP) {
        // blockIdx.z: the neighbour; the centre frame is shared
        const long long gd = group_off(go);
        nbr = gshift(nbr, gd);
        mvx = gshift(mvx, gd);
        mvy = gshift(mvy, gd);
        conf = gshift(conf, gd);
    }
    __shared__ float sc[32 * 32];
    __shared__ float ksad[kMeThreads];
    __shared__ int kidx[kMeThreads];
    __shared__ float s_sad0;
    const int bw = (w + block - 1) / block;
    const int bx = blockIdx.x % bw, by = blockIdx.x / bw;
    const int x0 = bx * block, y0 = by * block;
    const int bw_ = min(block, w - x0), bh_ = min(block, h - y0);
    const int nx = (bw_ + 1) / 2, ny = (bh_ + 1) / 2;
    for (int t = threadIdx.x; t < nx * ny; t += kMeThreads) {
        const int i = 2 * (t % nx), j = 2 * (t / nx);
        sc[t] = at(cur, w, h, x0 + i, y0 + j);
    }
    if (threadIdx.x == 0)
        s_sad0 = 1e30f;
    __syncthreads();
    const int side = 2 * search + 1, ncand = side * side;
    float best = 1e30f;
    int bidx = 0x7fffffff;
    int bnorm = 0x7fffffff;
    for (int c = threadIdx.x; c < ncand; c += kMeThreads) {
        const int dy = c / side - search, dx = c % side - search;
        if (x0 + dx < 0 || y0 + dy < 0 || x0 + dx + bw_ > w || y0 + dy + bh_ > h)
            continue;
        float sad = 0;
        for (int t = 0; t < nx * ny; ++t) {
            const int i = 2 * (t % nx), j = 2 * (t / nx);
            sad += fabsf(sc[t] - nbr[(size_t)(y0 + dy + j) * w + (x0 + dx + i)]);
        }
        if (dx == 0 && dy == 0)
            s_sad0 = sad;
        const int norm = abs(dx) + abs(dy);
        if (sad < best || (sad == best && (norm < bnorm || (norm == bnorm && c < bidx)))) {
            best = sad;
            bidx = c;
            bnorm = norm;
        }
    }
    ksad[threadIdx.x] = best;
    kidx[threadIdx.x] = bidx;
    __syncthreads();
    for (int off = kMeThreads / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            const float s2 = ksad[threadIdx.x + off];
            const int i2 = kidx[threadIdx.x + off];
            const float s1 = ksad[threadIdx.x];
            const int i1 = kidx[threadIdx.x];
            bool take = false;
            if (i2 != 0x7fffffff) {
                if (i1 == 0x7fffffff || s2 < s1) {
                    take = true;
                } else if (s2 == s1) {
                    const int n1 = abs(i1 / side - search) + abs(i1 % side - search);
                    const int n2 = abs(i2 / side - search) + abs(i2 % side - search);
                    take = n2 < n1 || (n2 == n1 && i2 < i1);
                }
            }
            if (take) {
                ksad[threadIdx.x] = s2;
                kidx[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0)
        return;
    float bsad = ksad[0];
    int bdx = 0, bdy = 0;
    if (kidx[0] != 0x7fffffff) {
        bdy = kidx[0] / side - search;
        bdx = kidx[0] % side - search;
    }
    // a tie against the (0,0) candidate keeps zero motion: norm 0 is smallest
    double sum = 0.0, sum2 = 0.0;
    for (int t = 0; t < nx * ny; ++t) {
        const float v = sc[t];
        sum += v;
        sum2 = dacc(sum2, v, v);
    }
    const float npix = (float)(nx * ny);
    const float sad0 = s_sad0;
    if ((bdx != 0 || bdy != 0) && sad0 - bsad <= 0.10f * tsad * npix) {
        bdx = 0;
        bdy = 0;
        bsad = sad0;
    }
    const int idx = blockIdx.x;
    mvx[idx] = (int16_t)bdx;
    mvy[idx] = (int16_t)bdy;
    const float sad_px = bsad / npix;
    const double vv = sum2 / npix - (sum / npix) * (sum / npix);
    const float var = (float)(0.0 < vv ? vv : 0.0);
    const float inv_t2 = 1.0f / (tsad * tsad);
    const float match_q = 1.0f / (1.0f + sad_px * sad_px * inv_t2);
    const float uniq = var / (var + 0.25f * tsad * tsad);
    conf[idx] = match_q * uniq;
}

// The same search with the block's search window in LDS: one wave per block,
// WAVES blocks per workgroup.  The window holds every neighbour sample a
// candidate placement inside the image reads (rows y0-search ..
// y0+bh-1+search, columns likewise; samples outside the image are clamped
// copies no valid placement reads).  Each lane scores kMeCand candidates per
// pass, walking the block's samples once in the reference's (j, i) order
// with one float accumulator per candidate, so every SAD rounds exactly as
// the scalar loop; the argmin is a wave reduction on (sad, |dx|+|dy|, scan
// index).
constexpr int kMeCand = 3;
constexpr int kMeLdsMax = 64 * 1024;

__device__ __forceinline__ bool me_better(float s2, int n2, int i2, float s1, int n1, int i1) {
    if (i2 == 0x7fffffff)
        return false;
    if (i1 == 0x7fffffff || s2 < s1)
        return true;
    return s2 == s1 && (n2 < n1 || (n2 == n1 && i2 < i1));
}

// SAD of kMeCand candidates over an NX x NY block with compile-time loop
// bounds, so the window reads of a row use immediate LDS offsets.
template <int NX, int NY>
__device__ __forceinline__ void me_sad_fixed(const float *sc, const float *win, int ws, const int (&off)[kMeCand],
                                             float (&acc)[kMeCand]) {
#pragma unroll
    for (int j = 0; j < NY; ++j) {
        const float *wr = win + 2 * j * ws;
#pragma unroll
        for (int i = 0; i < NX; ++i) {
            const float v = sc[j * NX + i];
#pragma unroll
            for (int k = 0; k < kMeCand; ++k)
                acc[k] += fabsf(v - wr[off[k] + 2 * i]);
        }
    }
}

template <int WAVES, bool GROUP>
__global__ __launch_bounds__(64 * WAVES) void k_block_match_win(const float *__restrict__ cur,
                                                                const float *__restrict__ nbr, int w, int h, int block,
                                                                int search, float tsad, int16_t *mvx, int16_t *mvy,
                                                                float *conf, GroupOff go) {
    if constexpr (GROUP) {
        const long long gd = group_off(go);
        nbr = gshift(nbr, gd);
        mvx = gshift(mvx, gd);
        mvy = gshift(mvy, gd);
        conf = gshift(conf, gd);
    }
    extern __shared__ float me_lds[];
    __shared__ float s_sad0[WAVES], s_best[WAVES];
    __shared__ int s_bidx[WAVES];
    // wave-uniform: the block geometry below lives in scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int bw = (w + block - 1) / block, nblocks = bw * ((h + block - 1) / block);
    // a wave past the last block repeats the last one; its result is not
    // stored, and every wave reaches the barriers
    const int blk = min((int)blockIdx.x * WAVES + wave, nblocks - 1);
    const int bx = blk % bw, by = blk / bw;
    const int x0 = bx * block, y0 = by * block;
    const int bw_ = min(block, w - x0), bh_ = min(block, h - y0);
    const int nx = (bw_ + 1) / 2, ny = (bh_ + 1) / 2;
    const int half = (block + 1) / 2;
    const int ws = block + 2 * search;
    const int region = half * half + ws * ws;
    float *sc = me_lds + wave * region;  // nx*ny block samples
    float *win = sc + half * half;       // ws x ws window
    for (int t = lane; t < nx * ny; t += 64)
        sc[t] = at(cur, w, h, x0 + 2 * (t % nx), y0 + 2 * (t / nx));
    {
        // window element t = r * ws + q; (r, q) advance by 64 per step
        const int dq = 64 % ws, dr = 64 / ws, n = ws * ws;
        int q = lane % ws, r = lane / ws;
        for (int t0 = 0; t0 < n; t0 += 256) {
            float v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                v[k] = at(nbr, w, h, x0 - search + q, y0 - search + r);  // clamped: in bounds for any (q, r)
                q += dq;
                r += dr;
                if (q >= ws) {
                    q -= ws;
                    ++r;
                }
            }
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (t0 + lane + 64 * k < n)
                    win[t0 + lane + 64 * k] = v[k];
        }
    }
    if (lane == 0)
        s_sad0[wave] = 1e30f;
    __syncthreads();
    const int side = 2 * search + 1, ncand = side * side;
    float best = 1e30f;
    int bidx = 0x7fffffff, bnorm = 0x7fffffff;
    const bool full = nx == 8 && ny == 8;
    for (int base = 0; base < ncand; base += 64 * kMeCand) {
        float acc[kMeCand];
        int off[kMeCand], cdx[kMeCand], cdy[kMeCand];
        bool cok[kMeCand];
#pragma unroll
        for (int k = 0; k < kMeCand; ++k) {
            const int c = base + lane + 64 * k;
            cdy[k] = c / side - search;
            cdx[k] = c - (cdy[k] + search) * side - search;
            cok[k] = c < ncand && x0 + cdx[k] >= 0 && y0 + cdy[k] >= 0 && x0 + cdx[k] + bw_ <= w &&
                     y0 + cdy[k] + bh_ <= h;
            off[k] = cok[k] ? (cdy[k] + search) * ws + (cdx[k] + search) : 0;
            acc[k] = 0.0f;
        }
        if (full) {
            me_sad_fixed<8, 8>(sc, win, ws, off, acc);
        } else {
            for (int j = 0; j < ny; ++j) {
                const float *srow = sc + j * nx;
                const float *wr = win + 2 * j * ws;
                for (int i = 0; i < nx; ++i) {
                    const float v = srow[i];
#pragma unroll
                    for (int k = 0; k < kMeCand; ++k)
                        acc[k] += fabsf(v - wr[off[k] + 2 * i]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kMeCand; ++k) {
            if (!cok[k])
                continue;
            if (cdx[k] == 0 && cdy[k] == 0)
                s_sad0[wave] = acc[k];
            const int c = base + lane + 64 * k;
            const int norm = abs(cdx[k]) + abs(cdy[k]);
            if (me_better(acc[k], norm, c, best, bnorm, bidx)) {
                best = acc[k];
                bidx = c;
                bnorm = norm;
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float s2 = __shfl_xor(best, o);
        const int i2 = __shfl_xor(bidx, o);
        const int n2 = __shfl_xor(bnorm, o);
        if (me_better(s2, n2, i2, best, bnorm, bidx)) {
            best = s2;
            bidx = i2;
            bnorm = n2;
        }
    }
    if (lane == 0) {
        s_best[wave] = best;
        s_bidx[wave] = bidx;
    }
    __syncthreads();
    // lane k of wave 0 finishes block k of the workgroup: one sequential
    // pass over the block's samples (double sums in the reference's order)
    // per block instead of one per wave
    if (wave != 0 || lane >= WAVES || (int)blockIdx.x * WAVES + lane >= nblocks)
        return;
    const int fb = (int)blockIdx.x * WAVES + lane;
    const int fx0 = (fb % bw) * block, fy0 = (fb / bw) * block;
    const int fnx = (min(block, w - fx0) + 1) / 2, fny = (min(block, h - fy0) + 1) / 2;
    const float *fsc = me_lds + lane * region;
    float bsad = s_best[lane];
    const int fidx = s_bidx[lane];
    int bdx = 0, bdy = 0;
    if (fidx != 0x7fffffff) {
        bdy = fidx / side - search;
        bdx = fidx % side - search;
    }
    double sum = 0.0, sum2 = 0.0;
    for (int t = 0; t < fnx * fny; ++t) {
        const float v = fsc[t];
        sum += v;
        sum2 = dacc(sum2, v, v);
    }
    const float npix = (float)(fnx * fny);
    const float sad0 = s_sad0[lane];
    if ((bdx != 0 || bdy != 0) && sad0 - bsad <= 0.10f * tsad * npix) {
        bdx = 0;
        bdy = 0;
        bsad = sad0;
    }
    mvx[fb] = (int16_t)bdx;
    mvy[fb] = (int16_t)bdy;
    const float sad_px = bsad / npix;
    const double vv = sum2 / npix - (sum / npix) * (sum / npix);
    const float var = (float)(0.0 < vv ? vv : 0.0);
    const float inv_t2 = 1.0f / (tsad * tsad);
    const float match_q = 1.0f / (1.0f + sad_px * sad_px * inv_t2);
    const float uniq = var / (var + 0.25f * tsad * tsad);
    conf[fb] = match_q * uniq;
}

constexpr int kCcMax = 1024;

constexpr int kCcThreads = 64;  // one wave per block

template <bool GROUP>
__global__ __launch_bounds__(kCcThreads) void k_chroma_consistency(const float *cu, const float *cv, const float *nu,
                                                               const float *nv, int cw, int ch, int luma_w,
                                                               int luma_h, int block, const int16_t *mvx,
                                                               const int16_t *mvy, int bw, float *conf, GroupOff go) {
    if constexpr (GROUP) {
        // blockIdx.z: the neighbour; the centre frame is shared
        const long long gd = group_off(go);
        nu = gshift(nu, gd);
        nv = gshift(nv, gd);
        mvx = gshift(mvx, gd);
        mvy = gshift(mvy, gd);
        conf = gshift(conf, gd);
    }
    __shared__ float delta[kCcMax];
    __shared__ float lo_val, mid_val;
    const int bi = blockIdx.x;
    if (conf[bi] <= 0.0f)
        return;
    const int bx = bi % bw, by = bi / bw;
    const double rw = (double)luma_w / cw, rh = (double)luma_h / ch;
    const int lx0 = bx * block, ly0 = by * block;
    const int lx1 = min(luma_w, lx0 + block), ly1 = min(luma_h, ly0 + block);
    const int cx0 = clampi((int)floor(lx0 / rw), 0, cw - 1), cy0 = clampi((int)floor(ly0 / rh), 0, ch - 1);
    const int cx1 = clampi((int)ceil(lx1 / rw), cx0 + 1, cw), cy1 = clampi((int)ceil(ly1 / rh), cy0 + 1, ch);
    const double mx = mvx[bi] / rw, my = mvy[bi] / rh;
    const int nxc = cx1 - cx0, n = nxc * (cy1 - cy0);
    if (n > kCcMax)
        return;  // the host rejects block sizes that could overflow
    for (int t = threadIdx.x; t < n; t += kCcThreads) {
        const int x = cx0 + t % nxc, y = cy0 + t / nxc;
        const int qx = clampi((int)llround(x + mx), 0, cw - 1), qy = clampi((int)llround(y + my), 0, ch - 1);
        const size_t a = (size_t)y * cw + x, b = (size_t)qy * cw + qx;
        delta[t] = fmx(fabsf(cu[a] - nu[b]), fabsf(cv[a] - nv[b]));
    }
    __syncthreads();
    // order statistics by rank: element t has rank #{smaller} + #{equal, lower index}
    const int mid = n / 2;
    for (int t = threadIdx.x; t < n; t += kCcThreads) {
        const float d = delta[t];
        int rank = 0;
        for (int k = 0; k < n; ++k) {
            const float e = delta[k];
            rank += (e < d) || (e == d && k < t);
        }
        if (rank == mid)
            mid_val = d;
        if (rank == mid - 1)
            lo_val = d;
    }
    __syncthreads();
    if (threadIdx.x != 0)
        return;
    float median = mid_val;
    if ((n & 1) == 0)
        median = 0.5f * (lo_val + median);
    const float ratio = median / 0.04f;
    const float r2 = ratio * ratio;
    conf[bi] *= 1.0f / (1.0f + r2 * r2);
}

// ---------------------------------------------------------------------------
// back-projection (backProject, algos.cpp:663-677); the reference applies it
// only when the output is exactly 2x the chroma grid, so the 2x2 blocks are
// disjoint and every chroma sample can be processed in parallel.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_back_project(float *ou, float *ov, int ow, int oh, const float *u,
                                                         const float *v, int cw, int ch, float bp) {
    const int cx = blockIdx.x * kBlock + threadIdx.x, cy = blockIdx.y;
    if (cx >= cw)
        return;
    const int xa = min(2 * cx, ow - 1), xb = min(2 * cx + 1, ow - 1);
    const int ya = min(2 * cy, oh - 1), yb = min(2 * cy + 1, oh - 1);
    for (int c = 0; c < 2; ++c) {
        float *o = c ? ov : ou;
        const float *s = c ? v : u;
        const float p00 = o[(size_t)ya * ow + xa], p10 = o[(size_t)ya * ow + xb];
        const float p01 = o[(size_t)yb * ow + xa], p11 = o[(size_t)yb * ow + xb];
        const float r = s[(size_t)cy * cw + cx] - 0.25f * (p00 + p10 + p01 + p11);
        const float d = bp * r;
        o[(size_t)ya * ow + xa] = p00 + d;
        o[(size_t)ya * ow + xb] = p10 + d;
        o[(size_t)yb * ow + xa] = p01 + d;
        o[(size_t)yb * ow + xb] = p11 + d;
    }
}

// ---------------------------------------------------------------------------
// luma resampling for resize (resampleH/V, kernels.cpp:171-271; radial jinc,
// kernels.cpp:276-334)
// ---------------------------------------------------------------------------

// SUP > 0: that many taps, every load issued before the sums; 0: t.sup.
template <int SUP>
__global__ __launch_bounds__(kBlock) void k_resample_h(const float *src, int sw, int sh, float *dst, WeightsDev t) {
    const int x = blockIdx.x * kBlock + threadIdx.x, y = blockIdx.y;
    if (x >= t.n)
        return;
    const float *wp = t.w + (size_t)x * (SUP ? SUP : t.sup);
    const float *row = src + (size_t)y * sw;
    const int s = t.start[x];
    double acc = 0.0;
    if constexpr (SUP > 0) {
        float w[SUP], a[SUP];
        _Pragma("unroll") for (int j = 0; j < SUP; ++j) {
            w[j] = wp[j];
            a[j] = row[clampi(s + j, 0, sw - 1)];
        }
        _Pragma("unroll") for (int j = 0; j < SUP; ++j)
            acc = dacc(acc, w[j], a[j]);
    } else {
        for (int j = 0; j < t.sup; ++j)
            acc = dacc(acc, wp[j], row[clampi(s + j, 0, sw - 1)]);
    }
    dst[(size_t)y * t.n + x] = (float)acc;
}
template <int SUP>
__global__ __launch_bounds__(kBlock) void k_resample_v(const float *src, int sw, int sh, float *dst, WeightsDev t) {
    const int x = blockIdx.x * kBlock + threadIdx.x, y = blockIdx.y;
    if (x >= sw)
        return;
    const float *wp = t.w + (size_t)y * (SUP ? SUP : t.sup);
    const int s = t.start[y];
    double acc = 0.0;
    if constexpr (SUP > 0) {
        float a[SUP];
        _Pragma("unroll") for (int j = 0; j < SUP; ++j)
            a[j] = src[(size_t)clampi(s + j, 0, sh - 1) * sw + x];
        _Pragma("unroll") for (int j = 0; j < SUP; ++j)
            acc = dacc(acc, wp[j], a[j]);
    } else {
        for (int j = 0; j < t.sup; ++j)
            acc = dacc(acc, wp[j], src[(size_t)clampi(s + j, 0, sh - 1) * sw + x]);
    }
    dst[(size_t)y * sw + x] = (float)acc;
}

__global__ __launch_bounds__(kBlock) void k_resample_radial(const float *src, int sw, int sh, float *dst, int dw,
                                                            int dh, double sx, double sy, float inv_w,
                                                            int radius, const float *lut, int lut_n,
                                                            float lut_scale) {
    const int ox = blockIdx.x * kBlock + threadIdx.x, oy = blockIdx.y;
    if (ox >= dw)
        return;
    const int sup = 2 * radius;
    const double cy = (oy + 0.5) * sy - 0.5, cx = (ox + 0.5) * sx - 0.5;
    const int y0 = (int)ceil(cy - radius), x0 = (int)ceil(cx - radius);
    double acc = 0.0, wsum = 0.0;
    for (int j = 0; j < sup; ++j) {
        const int ty = clampi(y0 + j, 0, sh - 1);
        const float dy = (float)((y0 + j - cy) * inv_w);
        for (int i = 0; i < sup; ++i) {
            const int tx = clampi(x0 + i, 0, sw - 1);
            const float dx = (float)((x0 + i - cx) * inv_w);
            const float w = lut_lookup(lut, lut_n, sqrtf(dx * dx + dy * dy) * lut_scale);
            acc = dacc(acc, w, src[(size_t)ty * sw + tx]);
            wsum += w;
        }
    }
    dst[(size_t)oy * dw + ox] = (float)(wsum > 1e-6 ? acc / wsum : 0.0);
}

// resampleRadial with class tables: a block row holds output pixels of one
// x class in one output row, so the pixel's weight table is uniform across
// the wave and read through the scalar cache; each tap costs one source
// load, one conversion and one FP64 FMA.  The sums run in the reference's
// (j, i) order with the same weights, so results equal the per-tap path.
template <int SUP>
__global__ __launch_bounds__(kBlock) void k_resample_radial_tab(const float *__restrict__ src, int sw, int sh,
                                                                float *__restrict__ dst, int dw,
                                                                RadialResampleDev T) {
    const int cls = blockIdx.z, oy = blockIdx.y;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    const int beg = T.xoff[cls], cnt = T.xoff[cls + 1] - beg;
    if (k >= cnt)
        return;
    const int S = SUP ? SUP : T.sup;
    const int ox = T.xs[beg + k];
    const int pair = (int)T.ycls[oy] * T.ncx + cls;
    const double *__restrict__ wt = T.w + (size_t)pair * S * S;
    const int x0 = T.x0[ox], y0 = T.y0[oy];
    double acc = 0.0;
    if constexpr (SUP == 6) {
        // the pair's weights (wave-uniform) read up front
        double wr[36];
        _Pragma("unroll") for (int t = 0; t < 36; ++t)
            wr[t] = wt[t];
        _Pragma("unroll") for (int j = 0; j < 6; ++j) {
            const float *row = src + (size_t)clampi(y0 + j, 0, sh - 1) * sw;
            _Pragma("unroll") for (int i = 0; i < 6; ++i)
                acc = __builtin_fma(wr[j * 6 + i], (double)row[clampi(x0 + i, 0, sw - 1)], acc);
        }
    } else {
        for (int j = 0; j < S; ++j) {
            const float *row = src + (size_t)clampi(y0 + j, 0, sh - 1) * sw;
            _Pragma("unroll") for (int i = 0; i < S; ++i)
                acc = __builtin_fma(wt[j * S + i], (double)row[clampi(x0 + i, 0, sw - 1)], acc);
        }
    }
    const double wsum = T.wsum[pair];
    dst[(size_t)oy * dw + ox] = (float)(wsum > 1e-6 ? acc / wsum : 0.0);
}

// The class-table radial resample with 6x6 taps.  A block computes R
// consecutive output rows of one segment of kSegCols output columns, each
// wave the columns of one x class inside it (T.segw): the union of the R
// rows' source rows is walked once, each source row's six samples loaded
// once for every output row whose taps include it, and each output
// accumulates in the reference's (j, i) order.  The results go through LDS so
// the block writes whole rows of the segment.
template <int R>
__global__ __launch_bounds__(kBlock) void k_resample_radial_rows(const float *__restrict__ src, int sw, int sh,
                                                                 float *__restrict__ dst, int dw, int dh,
                                                                 RadialResampleDev T) {
    __shared__ float tile[R][kSegCols];
    const int sg = blockIdx.x, oy0 = blockIdx.y * R;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int nrow = min(R, dh - oy0);
    int y0[R], ypair[R];
    int rlo = INT_MAX, rhi = INT_MIN;
    _Pragma("unroll") for (int s = 0; s < R; ++s) {
        const int oy = min(oy0 + s, dh - 1);
        y0[s] = T.y0[oy];
        ypair[s] = (int)T.ycls[oy] * T.ncx;
        if (s < nrow) {
            rlo = min(rlo, y0[s]);
            rhi = max(rhi, y0[s] + 5);
        }
    }
    for (int e = T.segoff[sg] + wave; e < T.segoff[sg + 1]; e += kBlock / 64) {
        const int4 en = T.segw[e];
        const int cls = en.x, k = en.y + lane;
        if (k >= en.z)
            continue;
        const int ox = T.xs[k];
        const int x0 = T.x0[ox];
        int col[6];
        _Pragma("unroll") for (int i = 0; i < 6; ++i)
            col[i] = clampi(x0 + i, 0, sw - 1);
        double acc[R];
        _Pragma("unroll") for (int s = 0; s < R; ++s)
            acc[s] = 0.0;
        for (int r = rlo; r <= rhi; ++r) {
            const float *row = src + (size_t)clampi(r, 0, sh - 1) * sw;
            double a[6];
            _Pragma("unroll") for (int i = 0; i < 6; ++i)
                a[i] = (double)row[col[i]];
            _Pragma("unroll") for (int s = 0; s < R; ++s) {
                const int j = r - y0[s];
                if (s >= nrow || j < 0 || j > 5)
                    continue;  // wave-uniform
                const double *__restrict__ w = T.w + (size_t)(ypair[s] + cls) * 36 + j * 6;
                _Pragma("unroll") for (int i = 0; i < 6; ++i)
                    acc[s] = __builtin_fma(w[i], a[i], acc[s]);
            }
        }
        _Pragma("unroll") for (int s = 0; s < R; ++s) {
            if (s >= nrow)
                break;
            const double wsum = T.wsum[ypair[s] + cls];
            tile[s][ox - sg * kSegCols] = (float)(wsum > 1e-6 ? acc[s] / wsum : 0.0);
        }
    }
    __syncthreads();
    const int x = sg * kSegCols + (int)threadIdx.x;
    if (x < dw)
        for (int s = 0; s < nrow; ++s)
            dst[(size_t)(oy0 + s) * dw + x] = tile[s][threadIdx.x];
}

// ---------------------------------------------------------------------------
// collaborative chroma filter (collaborativeChromaFilter, bm.cpp:250-442).
// Pass 1: one 64-lane block per anchor patch matches up to 4 patches, runs
// the 2D + group Haar shrinkage in LDS and stores the clamped filtered
// patches.  Pass 2: every pixel gathers the contributions of the groups that
// cover it in the reference's anchor order (anchor row, anchor column, group
// member), so the float sums round exactly as the scalar scatter does.
// ---------------------------------------------------------------------------

constexpr int kBmP = 8, kBmG = 4;
constexpr float kBmInvSqrt2 = 0.7071067811865475f;

template <int N> __device__ __forceinline__ void haar_fwd(float *v) {
    float t[N];
#pragma unroll
    for (int len = N; len > 1; len /= 2) {
        const int half = len / 2;
#pragma unroll
        for (int i = 0; i < half; ++i) {
            const float a = v[2 * i], b = v[2 * i + 1];
            t[i] = (a + b) * kBmInvSqrt2;
            t[half + i] = (a - b) * kBmInvSqrt2;
        }
#pragma unroll
        for (int i = 0; i < len; ++i)
            v[i] = t[i];
    }
}

template <int N> __device__ __forceinline__ void haar_inv(float *v) {
    float t[N];
#pragma unroll
    for (int len = 2; len <= N; len *= 2) {
        const int half = len / 2;
#pragma unroll
        for (int i = 0; i < half; ++i) {
            const float a = v[i], b = v[half + i];
            t[2 * i] = (a + b) * kBmInvSqrt2;
            t[2 * i + 1] = (a - b) * kBmInvSqrt2;
        }
#pragma unroll
        for (int i = 0; i < len; ++i)
            v[i] = t[i];
    }
}

// group Haar across the G members of coefficient l with hard threshold,
// keeping the DC of the DC (bm.cpp:374-393); returns the retained count
template <int G> __device__ __forceinline__ int bm_shrink(float (*P)[kBmG][64], int l, float thr) {
    int kept = 0;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        float gv[G];
#pragma unroll
        for (int k = 0; k < G; ++k) gv[k] = P[c][k][l];
        haar_fwd<G>(gv);
#pragma unroll
        for (int k = 0; k < G; ++k) {
            if (!(l == 0 && k == 0) && fabsf(gv[k]) < thr)
                gv[k] = 0.0f;
            else
                ++kept;
        }
        haar_inv<G>(gv);
#pragma unroll
        for (int k = 0; k < G; ++k) P[c][k][l] = gv[k];
    }
    return kept;
}

// window weight of patch offset i: 0.25, 0.5, 0.75, 1, 1, 0.75, 0.5, 0.25 (bm.cpp:285-287)
__device__ __forceinline__ float bm_window(int i) { return 0.25f * (float)(i < 4 ? i + 1 : 8 - i); }

__device__ __forceinline__ int bm_start(int i, int n, int extent) { return i == n - 1 ? extent - kBmP : i * 8; }

struct BmArgs {
    const float *g;           // match guide, w x h
    float *u, *v;             // chroma, w x h (filtered in place by pass 2)
    int w, h, nx, ny;
    float limit, thr;
    int *gc;                  // per anchor: group size (0 = no group)
    int *pos;                 // per anchor: x0, y0, .. x3, y3
    float *gw;                // per anchor: 1 / (1 + retained)
    float *filt;              // per anchor: [plane][k][64]
};

__global__ __launch_bounds__(64) void k_bm_group(BmArgs A) {
    __shared__ float cost[25];
    __shared__ int cand[25];          // packed (by << 16) | bx, or -1 for duplicates
    __shared__ int sel[1 + kBmG];     // group size, then packed (y << 16) | x per member
    __shared__ float P[2][kBmG][64];
    __shared__ int retained;
    const int a = blockIdx.x, l = threadIdx.x;
    const int iy = a / A.nx, ix = a - iy * A.nx;
    const int ax = bm_start(ix, A.nx, A.w), ay = bm_start(iy, A.ny, A.h);
    const int w = A.w;
    if (l < 25) {
        const int bx = clampi(ax + (-8 + 4 * (l % 5)), 0, A.w - kBmP);
        const int by = clampi(ay + (-8 + 4 * (l / 5)), 0, A.h - kBmP);
        bool dup = false;
        for (int j = 0; j < l; ++j) {
            const int jx = clampi(ax + (-8 + 4 * (j % 5)), 0, A.w - kBmP);
            const int jy = clampi(ay + (-8 + 4 * (j / 5)), 0, A.h - kBmP);
            dup |= jx == bx && jy == by;
        }
        float c = 0.0f;
        if (!dup && !(bx == ax && by == ay)) {
            int samples = 0;
            for (int py = 0; py < kBmP; py += 2)
                for (int px = 0; px < kBmP; px += 2) {
                    const size_t qa = (size_t)(ay + py) * w + ax + px, qb = (size_t)(by + py) * w + bx + px;
                    const float dg = A.g[qa] - A.g[qb], du = A.u[qa] - A.u[qb], dv = A.v[qa] - A.v[qb];
                    c += dg * dg + 0.25f * (du * du + dv * dv);
                    ++samples;
                }
            c = c / samples;
        }
        cost[l] = c;
        cand[l] = dup ? -1 : (by << 16) | bx;
    }
    if (l == 0)
        retained = 0;
    __syncthreads();
    if (l == 0) {
        // insertMatch (bm.cpp:29-42) over the candidates in scan order
        float *mc = P[0][0];
        int *mp = sel + 1;
        int count = 0;
        for (int k = 0; k < kBmG; ++k)
            mp[k] = (ay << 16) | ax;
        for (int i = 0; i < 25; ++i) {
            if (cand[i] < 0)
                continue;
            const float c = cost[i];
            int position = count < kBmG - 1 ? count : kBmG - 1;
            if (count == kBmG && !(c < mc[kBmG - 1]))
                continue;
            if (count < kBmG)
                ++count;
            while (position > 0 && c < mc[position - 1]) {
                mc[position] = mc[position - 1];
                mp[position] = mp[position - 1];
                --position;
            }
            mc[position] = c;
            mp[position] = cand[i];
        }
        int eligible = 0;
        while (eligible < count && mc[eligible] <= A.limit)
            ++eligible;
        sel[0] = eligible >= 4 ? 4 : (eligible >= 2 ? 2 : 0);
    }
    __syncthreads();
    const int gc = sel[0];
    if (l == 0)
        A.gc[a] = gc;
    if (gc == 0)
        return;
    int mx[kBmG], my[kBmG];
#pragma unroll
    for (int k = 0; k < kBmG; ++k) {
        mx[k] = sel[1 + k] & 0xffff;
        my[k] = sel[1 + k] >> 16;
    }
    // gather (bm.cpp:353-371); lane l owns patch offset l
    const int py = l >> 3, px = l & 7;
    float mnu = INFINITY, mxu = -INFINITY, mnv = INFINITY, mxv = -INFINITY;
#pragma unroll
    for (int k = 0; k < kBmG; ++k) {
        if (k >= gc)
            break;
        const size_t q = (size_t)(my[k] + py) * w + mx[k] + px;
        const float cu = A.u[q], cv = A.v[q];
        P[0][k][l] = cu;
        P[1][k][l] = cv;
        mnu = fmn(mnu, cu); mxu = fmx(mxu, cu);
        mnv = fmn(mnv, cv); mxv = fmx(mxv, cv);
    }
    __syncthreads();
    // 2D forward: rows then columns of every patch; lane = (plane, k, line)
    const int tp = l >> 5, tk = (l >> 3) & 3, tl = l & 7;
    float r[kBmP];
    if (tk < gc) {
#pragma unroll
        for (int i = 0; i < kBmP; ++i) r[i] = P[tp][tk][tl * kBmP + i];
        haar_fwd<kBmP>(r);
#pragma unroll
        for (int i = 0; i < kBmP; ++i) P[tp][tk][tl * kBmP + i] = r[i];
    }
    __syncthreads();
    if (tk < gc) {
#pragma unroll
        for (int i = 0; i < kBmP; ++i) r[i] = P[tp][tk][i * kBmP + tl];
        haar_fwd<kBmP>(r);
#pragma unroll
        for (int i = 0; i < kBmP; ++i) P[tp][tk][i * kBmP + tl] = r[i];
    }
    __syncthreads();
    // group transform with hard threshold (bm.cpp:374-393); lane = coefficient
    const int kept = gc == 4 ? bm_shrink<4>(P, l, A.thr) : bm_shrink<2>(P, l, A.thr);
    atomicAdd(&retained, kept);
    __syncthreads();
    // 2D inverse: columns then rows
    if (tk < gc) {
#pragma unroll
        for (int i = 0; i < kBmP; ++i) r[i] = P[tp][tk][i * kBmP + tl];
        haar_inv<kBmP>(r);
#pragma unroll
        for (int i = 0; i < kBmP; ++i) P[tp][tk][i * kBmP + tl] = r[i];
    }
    __syncthreads();
    if (tk < gc) {
#pragma unroll
        for (int i = 0; i < kBmP; ++i) r[i] = P[tp][tk][tl * kBmP + i];
        haar_inv<kBmP>(r);
#pragma unroll
        for (int i = 0; i < kBmP; ++i) P[tp][tk][tl * kBmP + i] = r[i];
    }
    __syncthreads();
    float *f = A.filt + (size_t)a * (2 * kBmG * 64);
    for (int k = 0; k < gc; ++k) {
        f[k * 64 + l] = fclamp(P[0][k][l], mnu, mxu);
        f[(kBmG + k) * 64 + l] = fclamp(P[1][k][l], mnv, mxv);
    }
    if (l < kBmG) {
        A.pos[a * 2 * kBmG + 2 * l] = mx[l];
        A.pos[a * 2 * kBmG + 2 * l + 1] = my[l];
    }
    if (l == 0)
        A.gw[a] = 1.0f / (1.0f + retained);
}

// One wave per 8x8 pixel tile (a block covers 32x8 pixels).  A pixel's
// matches come from anchors starting within [p - 15, p + 8] on each axis
// (members lie within 8 of their anchor, clamped inward); the wave walks the
// anchors of its whole tile, [x0 - 15, x0 + 15] x [y0 - 15, y0 + 15], with
// wave-uniform bounds, and each lane keeps the members that cover its pixel,
// so every pixel sums its contributions in the reference's (anchor row,
// anchor column, member) order.
constexpr int kBmAggTiles = kBlock / 64;
__global__ __launch_bounds__(kBlock) void k_bm_aggregate(BmArgs A) {
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int x0 = ((int)blockIdx.x * kBmAggTiles + wave) * 8, y0 = (int)blockIdx.y * 8;
    if (x0 >= A.w)
        return;  // whole wave
    const int x = x0 + (lane & 7), y = y0 + (lane >> 3);
    const int ix0 = max(0, (x0 - 15) / 8 - 1), iy0 = max(0, (y0 - 15) / 8 - 1);
    float su = 0.0f, sv = 0.0f, sw = 0.0f;
    for (int iy = iy0; iy < A.ny; ++iy) {
        const int ay = bm_start(iy, A.ny, A.h);
        if (ay > y0 + 15)
            break;
        if (ay < y0 - 15)
            continue;
        for (int ix = ix0; ix < A.nx; ++ix) {
            const int ax = bm_start(ix, A.nx, A.w);
            if (ax > x0 + 15)
                break;
            if (ax < x0 - 15)
                continue;
            const int a = iy * A.nx + ix;
            const int gc = A.gc[a];
            if (gc == 0)
                continue;
            const int *ps = A.pos + a * 2 * kBmG;
            const float gwt = A.gw[a];
            const float *f = A.filt + (size_t)a * (2 * kBmG * 64);
            for (int k = 0; k < gc; ++k) {
                const int px = x - ps[2 * k], py = y - ps[2 * k + 1];
                if ((unsigned)px >= (unsigned)kBmP || (unsigned)py >= (unsigned)kBmP)
                    continue;
                const float wgt = gwt * bm_window(px) * bm_window(py);
                su += wgt * f[k * 64 + py * kBmP + px];
                sv += wgt * f[(kBmG + k) * 64 + py * kBmP + px];
                sw += wgt;
            }
        }
    }
    if (x >= A.w || y >= A.h)
        return;
    const size_t q = (size_t)y * A.w + x;
    if (sw > 0.0f) {
        const float s0 = A.u[q], s1 = A.v[q];
        const float fu = su / sw, fv = sv / sw;
        A.u[q] = s0 + 0.65f * (fu - s0);
        A.v[q] = s1 + 0.65f * (fv - s1);
    }
}

__global__ __launch_bounds__(kBlock) void k_bm_guide(const float *g, int gw, int gh, float *dst, int w, int h,
                                                     double sx, double sy) {
    const int x = blockIdx.x * kBlock + threadIdx.x, y = blockIdx.y;
    if (x >= w)
        return;
    dst[(size_t)y * w + x] = bilinear(g, gw, gh, (x + 0.5) * sx - 0.5, (y + 0.5) * sy - 0.5);
}

// ---------------------------------------------------------------------------
// Sharpen (sharpenPlane, algos.cpp:595-659): 5x5 luma-gated Laplacian with a
// slope-median sigma floor and hull clamp; lc == nullptr self-guides.
// ---------------------------------------------------------------------------

struct SharpenArgs {
    const float *src;
    float *dst;
    int w, h;
    const float *lc;        // chroma-grid footprint luma, or nullptr
    const float *gy;        // luma guide (gw x gh) for L0
    int gw, gh;
    double rw, rh, sx, sy;
    float inv_g2, sigma, sratio, alpha, margin;
    int clamp_hull;
};

// sharpenPlane (algos.cpp:595-659): a block computes a 64x4 tile; the
// tile's plane and guide samples with a 2-sample apron are staged in LDS at
// clamped coordinates, so every lookup equals the reference's clamped read,
// and the 24 spatial weights (the same for every pixel) are formed once per
// block.  Both reciprocals have arguments >= 1, where recip_ge1 equals the
// IEEE quotient.  Taps in the reference's (j, i) order.
constexpr int kShW = 64, kShH = 4, kShAW = kShW + 4, kShAH = kShH + 4;
__global__ __launch_bounds__(kBlock) void k_sharpen(SharpenArgs A) {
    __shared__ float ts[kShAH][kShAW + 1];
    __shared__ float tg[kShAH][kShAW + 1];
    __shared__ float spw[25];
    const int w = A.w, h = A.h;
    const int x0 = blockIdx.x * kShW, y0 = blockIdx.y * kShH;
    const float *src = A.src;
    const float *gd = A.lc ? A.lc : src;
    for (int e = threadIdx.x; e < kShAH * kShAW; e += kBlock) {
        const int r = e / kShAW, c = e % kShAW;
        const size_t q = (size_t)clampi(y0 - 2 + r, 0, h - 1) * w + clampi(x0 - 2 + c, 0, w - 1);
        ts[r][c] = src[q];
        tg[r][c] = gd[q];
    }
    if (threadIdx.x < 25) {
        const int j = (int)threadIdx.x / 5 - 2, i = (int)threadIdx.x % 5 - 2;
        spw[threadIdx.x] = recip_ge1(1.0f + (float)(i * i + j * j) * A.inv_g2);
    }
    __syncthreads();
    const int tx = threadIdx.x % kShW, ty = threadIdx.x / kShW;
    const int x = x0 + tx, y = y0 + ty;
    if (x >= w || y >= h)
        return;
    const float c0 = ts[ty + 2][tx + 2];
    const float l0 = A.lc ? bilinear(A.gy, A.gw, A.gh, (x + 0.5) * A.rw - 0.5 + A.sx, (y + 0.5) * A.rh - 0.5 + A.sy)
                          : c0;
    float max_dl = 0.0f, lo = 1e30f, hi = -1e30f;
    for (int j = 0; j < 5; ++j)
        for (int i = 0; i < 5; ++i) {
            const float sv = ts[ty + j][tx + i];
            max_dl = fmx(max_dl, fabsf(tg[ty + j][tx + i] - l0));
            lo = fmn(lo, sv);
            hi = fmx(hi, sv);
        }
    // median of the four neighbour slopes: mean of the middle two
    const float a = fabsf(ts[ty + 2][tx + 3] - c0), b = fabsf(ts[ty + 2][tx + 1] - c0);
    const float c = fabsf(ts[ty + 3][tx + 2] - c0), d = fabsf(ts[ty + 1][tx + 2] - c0);
    const float a1 = fmn(a, b), b1 = fmx(a, b), c1 = fmn(c, d), d1 = fmx(c, d);
    const float m1 = fmx(a1, c1), m2 = fmn(b1, d1);
    const float med = (fmn(m1, m2) + fmx(m1, m2)) * 0.5f;
    const float sig = fmx(fmx(A.sigma, A.sratio * max_dl), 3.0f * med);
    const float inv_s2 = 1.0f / (sig * sig);
    float hp = 0.0f, ws = 0.0f;
    for (int j = 0; j < 5; ++j)
        for (int i = 0; i < 5; ++i) {
            if (i == 2 && j == 2)
                continue;
            const float dl = tg[ty + j][tx + i] - l0;
            const float sim = recip_ge1(1.0f + dl * dl * inv_s2);
            const float n = spw[j * 5 + i] * sim;
            hp += n * (c0 - ts[ty + j][tx + i]);
            ws += n;
        }
    float out = c0;
    if (ws > 1e-6f)
        out = c0 + A.alpha * hp / ws;
    if (A.clamp_hull)
        out = fclamp(out, lo - A.margin, hi + A.margin);
    A.dst[(size_t)y * w + x] = out;
}

inline dim3 rows_grid(int w, int h, int z = 1) { return dim3((unsigned)((w + kBlock - 1) / kBlock), (unsigned)h, (unsigned)z); }

// ---------------------------------------------------------------------------
// Downsample 4:4:4 -> 4:2:0 (downsample.cpp).  Every value is the float
// expression of the scalar build in its order; samples are read raw and
// converted per read (SourcePlane::at, downsample.cpp:101-106).
// ---------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ float ds_at(const DsPlane &p, int w, int h, int x, int y) {
    x = clampi(x, 0, w - 1);
    y = clampi(y, 0, h - 1);
    const T *row = reinterpret_cast<const T *>(static_cast<const char *>(p.data) + p.stride * (ptrdiff_t)y);
    return (float)row[x] * p.scale + p.offset;
}

// Downsample baseline (BaselineRing, downsample.cpp:129-261: ensureRow's
// horizontal taps, then makeRow's vertical taps over clamped rows) in one
// kernel, one kDsBaseTileW x kDsBaseTileH block of outputs per block; the
// reference's row ring becomes the tile's row buffer in LDS.  The horizontal pass runs over the tile's source rows (clamped
// rows of its vertical taps) in chunks of eight: each chunk's source
// segment is converted into LDS, the next chunk's raw samples are requested
// (requesting every chunk up front measured slower), then every thread makes
// one horizontal sample of both planes into the row buffer; the vertical
// pass reads the row buffer.  SX / SY are the tap counts (12 for spline36
// and lanczos3 at 2x, 4 or 5 binomial), each sum runs over the taps in order
// from 0, and samples convert as SourcePlane::at (downsample.cpp:101-106).
constexpr int kDbChunk = kBlock / kDsBaseTileW;
constexpr int kDbLoads = (kDbChunk * kDsBaseSegW + kBlock - 1) / kBlock;
static_assert(kDbChunk * kDsBaseTileW == kBlock && kDsBaseTileH % kDbChunk == 0,
              "one horizontal sample per thread per chunk, whole output rows per thread");

template <typename T>
__device__ __forceinline__ T ds_raw(const DsPlane &p, int w, int x, int y) {
    x = clampi(x, 0, w - 1);
    return reinterpret_cast<const T *>(static_cast<const char *>(p.data) + p.stride * (ptrdiff_t)y)[x];
}

template <typename T, int SX, int SY> __global__ __launch_bounds__(kBlock) void k_ds_base(DsArgs a) {
    __shared__ float seg_u[kDbChunk][kDsBaseSegW], seg_v[kDbChunk][kDsBaseSegW];
    __shared__ float row_u[kDsBaseSegH][kDsBaseTileW], row_v[kDsBaseSegH][kDsBaseTileW];
    const int x0 = blockIdx.x * kDsBaseTileW, oy0 = blockIdx.y * kDsBaseTileH;
    const int xl = min(x0 + kDsBaseTileW, a.cw) - 1, oyl = min(oy0 + kDsBaseTileH, a.ch) - 1;
    const int sx0 = a.xw.start[x0], nsx = a.xw.start[xl] + SX - sx0;
    const int r0 = clampi(a.yw.start[oy0], 0, a.h - 1);
    const int nr = clampi(a.yw.start[oyl] + SY - 1, 0, a.h - 1) - r0 + 1;
    const int tx = threadIdx.x % kDsBaseTileW, tr = threadIdx.x / kDsBaseTileW;
    const int x = min(x0 + tx, xl);
    const float *wt = a.xw.w + (size_t)x * SX;
    const int o = a.xw.start[x] - sx0;
    // this thread's slots of a chunk: sample i = threadIdx.x + j * kBlock
    // is row i / nsx, column i % nsx (row kDbChunk: unused slot)
    int srow[kDbLoads], scol[kDbLoads];
#pragma unroll
    for (int j = 0; j < kDbLoads; ++j) {
        const int i = threadIdx.x + j * kBlock;
        srow[j] = i < kDbChunk * nsx ? i / nsx : kDbChunk;
        scol[j] = i - srow[j] * nsx;
    }
    T ru[kDbLoads], rv[kDbLoads];
    auto fetch = [&](int c0) {
#pragma unroll
        for (int j = 0; j < kDbLoads; ++j)
            if (srow[j] < kDbChunk && c0 + srow[j] < nr) {
                ru[j] = ds_raw<T>(a.u, a.w, sx0 + scol[j], r0 + c0 + srow[j]);
                rv[j] = ds_raw<T>(a.v, a.w, sx0 + scol[j], r0 + c0 + srow[j]);
            }
    };
    fetch(0);
    float wx[SX];
#pragma unroll
    for (int t = 0; t < SX; ++t)
        wx[t] = wt[t];
    for (int c0 = 0; c0 < nr; c0 += kDbChunk) {
        const int rows = min(kDbChunk, nr - c0);
#pragma unroll
        for (int j = 0; j < kDbLoads; ++j)
            if (srow[j] < rows) {
                seg_u[srow[j]][scol[j]] = (float)ru[j] * a.u.scale + a.u.offset;
                seg_v[srow[j]][scol[j]] = (float)rv[j] * a.v.scale + a.v.offset;
            }
        __syncthreads();
        if (c0 + kDbChunk < nr)
            fetch(c0 + kDbChunk);
        if (tr < rows) {
            float su = 0.0f, sv = 0.0f;
#pragma unroll
            for (int t = 0; t < SX; ++t) {
                su += wx[t] * seg_u[tr][o + t];
                sv += wx[t] * seg_v[tr][o + t];
            }
            row_u[c0 + tr][tx] = su;
            row_v[c0 + tr][tx] = sv;
        }
        __syncthreads();
    }
    const int ox = x0 + tx;
    if (ox >= a.cw)
        return;
    for (int k = 0; k < kDsBaseTileH / kDbChunk; ++k) {
        const int oy = oy0 + tr + k * kDbChunk;
        if (oy > oyl)
            break;
        const float *wy = a.yw.w + (size_t)oy * SY;
        const int sy = a.yw.start[oy] - r0;
        float bu = 0.0f, bv = 0.0f;
#pragma unroll
        for (int t = 0; t < SY; ++t) {
            const int r = clampi(sy + t, -r0, a.h - 1 - r0);
            bu += wy[t] * row_u[r][tx];
            bv += wy[t] * row_v[r][tx];
        }
        a.base_u[(size_t)oy * a.cw + ox] = bu;
        a.base_v[(size_t)oy * a.cw + ox] = bv;
    }
}

__device__ __forceinline__ float2 ds_dir(int d) {
    switch (d) {
    case 0: return make_float2(1.0f, 0.0f);
    case 1: return make_float2(0.9238795f, 0.3826834f);
    case 2: return make_float2(0.7071068f, 0.7071068f);
    case 3: return make_float2(0.3826834f, 0.9238795f);
    case 4: return make_float2(0.0f, 1.0f);
    case 5: return make_float2(-0.3826834f, 0.9238795f);
    case 6: return make_float2(-0.7071068f, 0.7071068f);
    default: return make_float2(-0.9238795f, 0.3826834f);
    }
}

struct DsStats {
    int primary = 0, secondary = 1;
    float mix = 0.0f, gate = 0.0f;
    bool isotropic = true;
};

// Source sample access for the guided passes: a block's LDS tile holds, for
// every coordinate its samples read, the value the planes give there
// (clamped, converted as SourcePlane::at does).
// Guided Downsample tiles: 32 x 8 chroma samples per block, their 64 x 16
// source samples and a halo of H on every side, converted once into LDS.
// The window of radius r around cx in {2x, 2x + 0.5}, cy in {2y, 2y + 0.5,
// 2y + 1} spans source columns 2x - r .. 2x + r and rows 2y - r .. 2y + 1 + r,
// and the structure tensor reads +-1 columns and +-2 rows around
// (2x or 2x + 1, 2y or 2y + 1), so H = max(r, 2) covers every read.
constexpr int kDsTW = 32, kDsTH = 8;
template <int H> struct DsTile {
    static constexpr int SW = 2 * kDsTW + 2 * H, SH = 2 * kDsTH + 2 * H, N = SW * SH;
    const float *ty, *tu, *tv;
    int ox, oy;
    __device__ float y(int x, int r) const { return ty[(r - oy) * SW + (x - ox)]; }
    __device__ float u(int x, int r) const { return tu[(r - oy) * SW + (x - ox)]; }
    __device__ float v(int x, int r) const { return tv[(r - oy) * SW + (x - ox)]; }
};

// classify (downsample.cpp:281-385)
template <typename S>
__device__ DsStats ds_classify(const S &src, float cx, float cy, int quality) {
    DsStats st;
    const int ix = (int)floorf(cx + 0.5f), iy = (int)floorf(cy + 0.5f);
    float jxx = 0.0f, jxy = 0.0f, jyy = 0.0f, cgx = 0.0f, cgy = 0.0f;
    if (quality == 0) {
        const float tl = src.y(ix - 1, iy - 1), tc = src.y(ix, iy - 1);
        const float tr = src.y(ix + 1, iy - 1);
        const float ml = src.y(ix - 1, iy), mr = src.y(ix + 1, iy);
        const float bl = src.y(ix - 1, iy + 1), bc = src.y(ix, iy + 1);
        const float br = src.y(ix + 1, iy + 1);
        cgx = (tr + 2.0f * mr + br - tl - 2.0f * ml - bl) * 0.125f;
        cgy = (bl + 2.0f * bc + br - tl - 2.0f * tc - tr) * 0.125f;
        jxx = cgx * cgx;
        jxy = cgx * cgy;
        jyy = cgy * cgy;
    } else {
        for (int k = -1; k <= 1; ++k) {
            const float gx = 0.5f * (src.y(ix + 1, iy + k) - src.y(ix - 1, iy + k));
            const float gy = 0.5f * (src.y(ix, iy + k + 1) - src.y(ix, iy + k - 1));
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
        const float2 n = ds_dir(d);
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
        const float ratio = fclamp(second / best, 0.0f, 1.0f);
        st.mix = 0.5f * fclamp((ratio - adj) / (1.0f - adj), 0.0f, 1.0f);
    }
    const float ugx = 0.5f * (src.u(ix + 1, iy) - src.u(ix - 1, iy));
    const float ugy = 0.5f * (src.u(ix, iy + 1) - src.u(ix, iy - 1));
    const float vgx = 0.5f * (src.v(ix + 1, iy) - src.v(ix - 1, iy));
    const float vgy = 0.5f * (src.v(ix, iy + 1) - src.v(ix, iy - 1));
    const float cen = ugx * ugx + ugy * ugy + vgx * vgx + vgy * vgy;
    const float2 nrm = ds_dir(st.primary);
    const float un = ugx * nrm.x + ugy * nrm.y, vn = vgx * nrm.x + vgy * nrm.y;
    const float aligned = un * un + vn * vn;
    const int nsx = (int)lroundf(nrm.x), nsy = (int)lroundf(nrm.y);
    const float cu = src.u(ix, iy), cv = src.v(ix, iy);
    const float curu = src.u(ix + nsx, iy + nsy) - 2.0f * cu + src.u(ix - nsx, iy - nsy);
    const float curv = src.v(ix + nsx, iy + nsy) - 2.0f * cv + src.v(ix - nsx, iy - nsy);
    const float curE = curu * curu + curv * curv;
    const float sharp = curE / (curE + cen + 1.0e-12f);
    const float luma_g = trace / (trace + 3.0e-5f);
    const float conf_g = coh2 / (coh2 + 0.08f);
    const float chroma_g = cen / (cen + 2.0e-5f);
    const float dir_g = aligned / (cen + 1.0e-12f);
    const float centerE = cgx * cgx + cgy * cgy;
    const float pos_g = centerE / (centerE + 1.0e-5f);
    st.gate = fclamp(luma_g * conf_g * chroma_g * dir_g * pos_g * sharp, 0.0f, 1.0f);
    if (quality == 0)
        st.gate *= 0.985f;
    else if (quality == 2)
        st.gate = fminf(1.0f, st.gate * 1.01f);
    if (sharp < (quality == 0 ? 0.20f : 0.12f))
        st.gate = 0.0f;
    return st;
}

struct DsEnv {
    float nx, ny;
    int radius;
    float aniso;
    bool iso;
};

// spatialWeight (downsample.cpp:402-411).  Both denominators are 1 plus
// non-negative terms and stay below 200 for windows of radius <= 4, where
// rcp_nr equals the division 1.0f / y exactly.
__device__ __forceinline__ float ds_spatial(const DsEnv &e, float dx, float dy) {
    const float r2 = dx * dx + dy * dy;
    if (e.iso)
        return rcp_nr(1.0f + 0.42f * r2);
    const float nd = dx * e.nx + dy * e.ny;
    const float td = -dx * e.ny + dy * e.nx;
    return rcp_nr(1.0f + (0.9f + e.aniso) * nd * nd + 0.16f * td * td);
}

// accumulateGuided (downsample.cpp:413-458) over N envelopes
template <typename S, int N>
__device__ void ds_accumulate(const S &src, float cx, float cy, const DsEnv (&env)[N], float (&aw)[N],
                              float (&au)[N], float (&av)[N]) {
    int maxr = 0;
    _Pragma("unroll") for (int i = 0; i < N; ++i) {
        aw[i] = au[i] = av[i] = 0.0f;
        maxr = max(maxr, env[i].radius);
    }
    const int x0 = (int)ceilf(cx - (float)maxr), x1 = (int)floorf(cx + (float)maxr);
    const int y0 = (int)ceilf(cy - (float)maxr), y1 = (int)floorf(cy + (float)maxr);
    const int cx0 = (int)floorf(cx), cy0 = (int)floorf(cy);
    const float fx = cx - (float)cx0, fy = cy - (float)cy0;
    const float ya = src.y(cx0, cy0), yb = src.y(cx0 + 1, cy0);
    const float yc = src.y(cx0, cy0 + 1), yd = src.y(cx0 + 1, cy0 + 1);
    const float cl = ya + (yb - ya) * fx + (yc - ya) * fy + (ya - yb - yc + yd) * fx * fy;
    const float rs2 = 2.25e-4f;
    for (int sy = y0; sy <= y1; ++sy)
        for (int sx = x0; sx <= x1; ++sx) {
            const float dx = (float)sx - cx, dy = (float)sy - cy;
            const float delta = src.y(sx, sy) - cl;
            const float rw = rs2 / (rs2 + delta * delta);
            const float su = src.u(sx, sy), sv = src.v(sx, sy);
            _Pragma("unroll") for (int i = 0; i < N; ++i) {
                if (fabsf(dx) > (float)env[i].radius || fabsf(dy) > (float)env[i].radius)
                    continue;
                const float wgt = rw * ds_spatial(env[i], dx, dy);
                aw[i] += wgt;
                au[i] += wgt * su;
                av[i] += wgt * sv;
            }
        }
}

__device__ __forceinline__ float ds_norm(float sum, float wgt, float fallback) {
    return wgt > 1e-12f ? sum / wgt : fallback;
}

// Stage the block's source tile (clamped, converted) in LDS.
template <typename T, int H>
__device__ DsTile<H> ds_load_tile(const DsArgs &a, float *ty, float *tu, float *tv) {
    constexpr int SW = DsTile<H>::SW;
    const int ox = (int)blockIdx.x * 2 * kDsTW - H, oy = (int)blockIdx.y * 2 * kDsTH - H;
    const int tid = threadIdx.x;
    for (int e = tid; e < DsTile<H>::N; e += kBlock) {
        const int r = e / SW, c = e - r * SW;
        ty[e] = ds_at<T>(a.y, a.w, a.h, ox + c, oy + r);
        tu[e] = ds_at<T>(a.u, a.w, a.h, ox + c, oy + r);
        tv[e] = ds_at<T>(a.v, a.w, a.h, ox + c, oy + r);
    }
    __syncthreads();
    return DsTile<H>{ty, tu, tv, ox, oy};
}

// writeRow (downsample.cpp:117-126)
template <typename T> __device__ __forceinline__ void ds_store(const RawOut &P, int x, int y, float v) {
    T *row = reinterpret_cast<T *>(static_cast<char *>(P.data) + P.stride * (ptrdiff_t)y);
    if constexpr (sizeof(T) == 4)
        row[x] = v;
    else
        row[x] = pack_sample<T>(v, P);
}

// Gated samples of a guided tile, compacted in LDS: a sample whose gate
// passes is appended (tile index, directions, isotropy, mix, gate) and the
// block's threads then run the window sums over the list, so only as many
// lanes as there are gated samples walk the windows.  Every sample's
// arithmetic is unchanged; only the thread that computes it differs.
// Envelopes are independent sums, so with few gated samples the list runs
// one lane per (sample, envelope): envelope k sums over its own window in
// the same row-major order as the all-envelope loop (a tap outside an
// envelope's radius is skipped there), and the sample's lane finishes from
// the per-envelope sums in LDS.
template <bool GATE> struct DsGated {
    int n;
    int key[kBlock];      // tile index | primary << 8 | secondary << 11 | isotropic << 14
    float mix[kBlock];
    float gate[GATE ? kBlock : 1];
};

__device__ __forceinline__ int ds_gated_key(int t, const DsStats &st) {
    return t | st.primary << 8 | st.secondary << 11 | (st.isotropic ? 1 << 14 : 0);
}

__device__ __forceinline__ void ds_gated_decode(int key, int &x, int &y, int &primary, int &secondary, bool &iso) {
    const int t = key & 0xff;
    x = blockIdx.x * kDsTW + t % kDsTW;
    y = blockIdx.y * kDsTH + t / kDsTW;
    primary = (key >> 8) & 7;
    secondary = (key >> 11) & 7;
    iso = (key >> 14) & 1;
}

// packed baseline (strength 0) or the guided result of quality 0/1
// (processFrame, downsample.cpp:686-712, with makeGuidedPixel,
// downsample.cpp:466-507).  Every thread classifies one chroma sample and
// packs the ungated ones; the gated ones go through the list.  H: 2 for
// quality 0 (radius 2), 3 for quality 1 (radius 3).
template <typename T, int H> __global__ __launch_bounds__(kBlock) void k_ds_output(DsArgs a) {
    __shared__ float ty[DsTile<H>::N], tu[DsTile<H>::N], tv[DsTile<H>::N];
    __shared__ DsGated<true> list;
    const int tid = threadIdx.x;
    const int x = blockIdx.x * kDsTW + (int)(tid % kDsTW), y = blockIdx.y * kDsTH + (int)(tid / kDsTW);
    const bool valid = x < a.cw && y < a.ch;
    if (a.strength == 0.0f) {
        if (valid) {
            const size_t i = (size_t)y * a.cw + x;
            ds_store<T>(a.out_u, x, y, a.base_u[i]);
            ds_store<T>(a.out_v, x, y, a.base_v[i]);
        }
        return;
    }
    if (tid == 0)
        list.n = 0;
    const DsTile<H> src = ds_load_tile<T, H>(a, ty, tu, tv);
    if (valid) {
        const size_t i = (size_t)y * a.cw + x;
        const float bu = a.base_u[i], bv = a.base_v[i];
        const float cx = 2.0f * (float)x + 0.5f + a.shift_x, cy = 2.0f * (float)y + 0.5f + a.shift_y;
        const DsStats st = ds_classify(src, cx, cy, a.quality);
        if (!(st.gate <= 1.0e-4f)) {
            const int e = atomicAdd(&list.n, 1);
            list.key[e] = ds_gated_key(tid, st);
            list.mix[e] = st.mix;
            list.gate[e] = st.gate;
        } else {
            // gate 0: the blend keeps the baseline (gu = bu)
            const float gate = 0.0f;
            ds_store<T>(a.out_u, x, y, bu + a.strength * gate * (bu - bu));
            ds_store<T>(a.out_v, x, y, bv + a.strength * gate * (bv - bv));
        }
    }
    __syncthreads();
    const int n = list.n;
    if constexpr (H == 2) {
        // quality 0: one envelope, one lane per gated sample
        for (int e = tid; e < n; e += kBlock) {
            int gx, gy, primary, secondary;
            bool iso;
            ds_gated_decode(list.key[e], gx, gy, primary, secondary, iso);
            const float gate = list.gate[e];
            const size_t i = (size_t)gy * a.cw + gx;
            const float bu = a.base_u[i], bv = a.base_v[i];
            const float cx = 2.0f * (float)gx + 0.5f + a.shift_x, cy = 2.0f * (float)gy + 0.5f + a.shift_y;
            const float2 n0 = ds_dir(primary);
            const DsEnv env[1] = {{n0.x, n0.y, 2, 1.8f, iso}};
            float aw[1], au[1], av[1];
            ds_accumulate<DsTile<H>, 1>(src, cx, cy, env, aw, au, av);
            const float gu = ds_norm(au[0], aw[0], bu), gv = ds_norm(av[0], aw[0], bv);
            ds_store<T>(a.out_u, gx, gy, bu + a.strength * gate * (gu - bu));
            ds_store<T>(a.out_v, gx, gy, bv + a.strength * gate * (gv - bv));
        }
    } else {
        // quality 1: primary and secondary envelopes, one lane each
        constexpr int NE = 2, PER = 64;
        __shared__ float sums[PER][NE][3];
        for (int e0 = 0; e0 < n; e0 += PER) {
            const int m = min(PER, n - e0);
            if (tid < m * NE) {
                const int e = e0 + tid / NE, k = tid % NE;
                int gx, gy, primary, secondary;
                bool iso;
                ds_gated_decode(list.key[e], gx, gy, primary, secondary, iso);
                const float cx = 2.0f * (float)gx + 0.5f + a.shift_x, cy = 2.0f * (float)gy + 0.5f + a.shift_y;
                const float2 d = ds_dir(k == 0 ? primary : secondary);
                const DsEnv env[1] = {{d.x, d.y, 3, 2.2f, iso}};
                float aw[1], au[1], av[1];
                ds_accumulate<DsTile<H>, 1>(src, cx, cy, env, aw, au, av);
                sums[tid / NE][k][0] = aw[0];
                sums[tid / NE][k][1] = au[0];
                sums[tid / NE][k][2] = av[0];
            }
            __syncthreads();
            if (tid < m) {
                const int e = e0 + tid;
                int gx, gy, primary, secondary;
                bool iso;
                ds_gated_decode(list.key[e], gx, gy, primary, secondary, iso);
                const float gate = list.gate[e], mix = list.mix[e];
                const size_t i = (size_t)gy * a.cw + gx;
                const float bu = a.base_u[i], bv = a.base_v[i];
                const float u0 = ds_norm(sums[tid][0][1], sums[tid][0][0], bu);
                const float v0 = ds_norm(sums[tid][0][2], sums[tid][0][0], bv);
                const float u1 = ds_norm(sums[tid][1][1], sums[tid][1][0], bu);
                const float v1 = ds_norm(sums[tid][1][2], sums[tid][1][0], bv);
                const float gu = u0 + mix * (u1 - u0), gv = v0 + mix * (v1 - v0);
                ds_store<T>(a.out_u, gx, gy, bu + a.strength * gate * (gu - bu));
                ds_store<T>(a.out_v, gx, gy, bv + a.strength * gate * (gv - bv));
            }
            __syncthreads();
        }
    }
}

// quality 2: gate and the four candidates of every chroma sample
// (generateCandidateRow, downsample.cpp:526-568); ungated samples copy the
// baseline, gated ones go through the list
template <typename T> __global__ __launch_bounds__(kBlock) void k_ds_candidates(DsArgs a) {
    constexpr int H = 5;  // radius 4; one spare column and row measured faster than 4
    __shared__ float ty[DsTile<H>::N], tu[DsTile<H>::N], tv[DsTile<H>::N];
    __shared__ DsGated<false> list;
    const int tid = threadIdx.x;
    const int x = blockIdx.x * kDsTW + (int)(tid % kDsTW), y = blockIdx.y * kDsTH + (int)(tid / kDsTW);
    if (tid == 0)
        list.n = 0;
    const DsTile<H> src = ds_load_tile<T, H>(a, ty, tu, tv);
    const size_t nc = (size_t)a.cw * a.ch;
    if (x < a.cw && y < a.ch) {
        const size_t i = (size_t)y * a.cw + x;
        const float cx = 2.0f * (float)x + 0.5f + a.shift_x, cy = 2.0f * (float)y + 0.5f + a.shift_y;
        const DsStats st = ds_classify(src, cx, cy, 2);
        const float gate = st.gate > 1.0e-4f ? st.gate : 0.0f;
        a.gate[i] = gate;
        if (gate == 0.0f) {
            const float bu = a.base_u[i], bv = a.base_v[i];
            _Pragma("unroll") for (int c = 0; c < 4; ++c) {
                a.cand_u[c * nc + i] = bu;
                a.cand_v[c * nc + i] = bv;
            }
        } else {
            const int e = atomicAdd(&list.n, 1);
            list.key[e] = ds_gated_key(tid, st);
            list.mix[e] = st.mix;
        }
    }
    __syncthreads();
    const int n = list.n;
    // envelopes: primary / secondary at radius 3 (aniso 2.2), at radius 4
    // (2.7), isotropic radius 4
    constexpr int NE = 5, PER = 32;
    __shared__ float sums[PER][NE][3];
    for (int e0 = 0; e0 < n; e0 += PER) {
        const int m = min(PER, n - e0);
        if (tid < m * NE) {
            const int e = e0 + tid / NE, k = tid % NE;
            int gx, gy, primary, secondary;
            bool iso;
            ds_gated_decode(list.key[e], gx, gy, primary, secondary, iso);
            const float cx = 2.0f * (float)gx + 0.5f + a.shift_x, cy = 2.0f * (float)gy + 0.5f + a.shift_y;
            const float2 d = ds_dir(k == 1 || k == 3 ? secondary : primary);
            const DsEnv env[1] = {{d.x, d.y, k < 2 ? 3 : 4, k < 2 ? 2.2f : (k < 4 ? 2.7f : 0.0f), k == 4 || iso}};
            float aw[1], au[1], av[1];
            ds_accumulate<DsTile<H>, 1>(src, cx, cy, env, aw, au, av);
            sums[tid / NE][k][0] = aw[0];
            sums[tid / NE][k][1] = au[0];
            sums[tid / NE][k][2] = av[0];
        }
        __syncthreads();
        if (tid < m) {
            int gx, gy, primary, secondary;
            bool iso;
            ds_gated_decode(list.key[e0 + tid], gx, gy, primary, secondary, iso);
            const size_t i = (size_t)gy * a.cw + gx;
            const float bu = a.base_u[i], bv = a.base_v[i];
            const float mix = list.mix[e0 + tid];
            const float u0 = ds_norm(sums[tid][0][1], sums[tid][0][0], bu);
            const float v0 = ds_norm(sums[tid][0][2], sums[tid][0][0], bv);
            const float u1 = ds_norm(sums[tid][1][1], sums[tid][1][0], bu);
            const float v1 = ds_norm(sums[tid][1][2], sums[tid][1][0], bv);
            a.cand_u[i] = u0 + mix * (u1 - u0);
            a.cand_v[i] = v0 + mix * (v1 - v0);
            _Pragma("unroll") for (int c = 1; c < 4; ++c) {
                a.cand_u[c * nc + i] = ds_norm(sums[tid][c + 1][1], sums[tid][c + 1][0], bu);
                a.cand_v[c * nc + i] = ds_norm(sums[tid][c + 1][2], sums[tid][c + 1][0], bv);
            }
        }
        __syncthreads();
    }
}

// quality 2: loopback score over the sample's 2x2 source block, best
// candidate, blend and packing (scoreCandidateRow, downsample.cpp:576-626;
// rows outside the frame clamp as the three-row ring's lookup does).  One
// block per 256 samples of a row: ungated samples pack the baseline, gated
// ones are listed and scored one lane per (sample, source pixel); the
// sample's lane adds the four pixels' terms in the scalar order.
template <typename T> __global__ __launch_bounds__(kBlock) void k_ds_score(DsArgs a) {
    constexpr int PER = kBlock / 4;
    __shared__ int n_list;
    __shared__ int list[kBlock];
    __shared__ float part[PER][4][4];  // [sample][source pixel][candidate]
    const int tid = threadIdx.x;
    const int x = blockIdx.x * kBlock + tid, y = blockIdx.y;
    const int cw = a.cw, ch = a.ch, w = a.w, h = a.h;
    const size_t nc = (size_t)cw * ch;
    if (tid == 0)
        n_list = 0;
    __syncthreads();
    if (x < cw) {
        const size_t i = (size_t)y * cw + x;
        if (a.gate[i] != 0.0f) {
            list[atomicAdd(&n_list, 1)] = x;
        } else {
            ds_store<T>(a.out_u, x, y, a.base_u[i]);
            ds_store<T>(a.out_v, x, y, a.base_v[i]);
        }
    }
    __syncthreads();
    const int n = n_list;
    for (int e0 = 0; e0 < n; e0 += PER) {
        const int m = min(PER, n - e0);
        if (tid < 4 * m) {
            const int sx = list[e0 + tid / 4], k = tid % 4;
            const int hy = 2 * y + k / 2, hx = 2 * sx + k % 2;
            const float gx = 0.5f * (ds_at<T>(a.y, w, h, hx + 1, hy) - ds_at<T>(a.y, w, h, hx - 1, hy));
            const float gy = 0.5f * (ds_at<T>(a.y, w, h, hx, hy + 1) - ds_at<T>(a.y, w, h, hx, hy - 1));
            const float e2 = gx * gx + gy * gy;
            const float ew = 0.125f + e2 / (e2 + 1.0e-5f);
            const float lx = ((float)hx - 0.5f - a.shift_x) * 0.5f;
            const float ly = ((float)hy - 0.5f - a.shift_y) * 0.5f;
            const float su = ds_at<T>(a.u, w, h, hx, hy), sv = ds_at<T>(a.v, w, h, hx, hy);
            const int rx0 = (int)floorf(lx), ry0 = (int)floorf(ly);
            const int x0 = clampi(rx0, 0, cw - 1), x1 = clampi(rx0 + 1, 0, cw - 1);
            const int r0 = clampi(ry0, 0, ch - 1), r1 = clampi(ry0 + 1, 0, ch - 1);
            const float fx = lx - (float)rx0, fy = ly - (float)ry0;
            const float wts[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy, fx * fy};
            const size_t idx[4] = {(size_t)r0 * cw + x0, (size_t)r0 * cw + x1, (size_t)r1 * cw + x0,
                                   (size_t)r1 * cw + x1};
            _Pragma("unroll") for (int c = 0; c < 4; ++c) {
                float recu = 0.0f, recv = 0.0f;
                _Pragma("unroll") for (int q = 0; q < 4; ++q) {
                    const size_t j = idx[q];
                    float vu = a.base_u[j], vv = a.base_v[j];
                    vu += a.gate[j] * (a.cand_u[c * nc + j] - vu);
                    vv += a.gate[j] * (a.cand_v[c * nc + j] - vv);
                    recu += wts[q] * vu;
                    recv += wts[q] * vv;
                }
                const float du = recu - su, dv = recv - sv;
                part[tid / 4][k][c] = ew * (du * du + dv * dv);
            }
        }
        __syncthreads();
        if (tid < m) {
            const int sx = list[e0 + tid];
            const size_t i = (size_t)y * cw + sx;
            float score[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            _Pragma("unroll") for (int k = 0; k < 4; ++k)
                _Pragma("unroll") for (int c = 0; c < 4; ++c)
                    score[c] += part[tid][k][c];
            int best = 0;
            _Pragma("unroll") for (int c = 1; c < 4; ++c)
                if (score[c] < score[best])
                    best = c;
            const float bu = a.base_u[i], bv = a.base_v[i], gate = a.gate[i];
            ds_store<T>(a.out_u, sx, y, bu + a.strength * gate * (a.cand_u[best * nc + i] - bu));
            ds_store<T>(a.out_v, sx, y, bv + a.strength * gate * (a.cand_v[best * nc + i] - bv));
        }
        __syncthreads();
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

hipError_t launch_to_float(const RawPlane *planes, int count, int type, hipStream_t s) {
    InBatch b{};
    int mw = 0, mh = 0;
    for (int i = 0; i < count; ++i) {
        b.p[i] = planes[i];
        mw = max(mw, planes[i].w);
        mh = max(mh, planes[i].h);
    }
    const dim3 g(1, (unsigned)mh, (unsigned)count);
    if (type == RAW_U8)
        hipLaunchKernelGGL(k_to_float<uint8_t>, g, dim3(kBlock), 0, s, b);
    else if (type == RAW_U16)
        hipLaunchKernelGGL(k_to_float<uint16_t>, g, dim3(kBlock), 0, s, b);
    else
        hipLaunchKernelGGL(k_copy_f32, g, dim3(kBlock), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_to_float_group(const RawPlane *planes, int count, int type, hipStream_t s) {
    if (count < 1 || count > 3 * kGroupMax)
        return hipErrorInvalidValue;
    InBatchG b{};
    int mh = 0;
    for (int i = 0; i < count; ++i) {
        b.p[i] = planes[i];
        mh = max(mh, planes[i].h);
    }
    const dim3 g(1, (unsigned)mh, (unsigned)count);
    if (type == RAW_U8)
        hipLaunchKernelGGL((k_to_float_group<uint8_t, false>), g, dim3(kBlock), 0, s, b);
    else if (type == RAW_U16)
        hipLaunchKernelGGL((k_to_float_group<uint16_t, false>), g, dim3(kBlock), 0, s, b);
    else
        hipLaunchKernelGGL((k_to_float_group<float, true>), g, dim3(kBlock), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_from_float(const RawOut *planes, int count, int type, hipStream_t s) {
    OutBatch b{};
    int mw = 0, mh = 0;
    for (int i = 0; i < count; ++i) {
        b.p[i] = planes[i];
        mw = max(mw, planes[i].w);
        mh = max(mh, planes[i].h);
    }
    const dim3 g(1, (unsigned)mh, (unsigned)count);
    if (type == RAW_U8)
        hipLaunchKernelGGL(k_from_float<uint8_t>, g, dim3(kBlock), 0, s, b);
    else if (type == RAW_U16)
        hipLaunchKernelGGL(k_from_float<uint16_t>, g, dim3(kBlock), 0, s, b);
    else
        hipLaunchKernelGGL(k_from_float_f32, g, dim3(kBlock), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_sobel_tensor(const float *y, int w, int h, float4 *yt, uint64_t *seed, float threshold,
                               hipStream_t s, int tile_rows, const GroupOff *go, int nf) {
    const GroupOff g0 = go ? *go : GroupOff{};
#define LGCR_SOBEL_G(TH_, G_)                                                                                  \
    hipLaunchKernelGGL((k_sobel_tensor<TH_, G_>),                                                              \
                       dim3((unsigned)((w + kSobelW - 1) / kSobelW), (unsigned)((h + TH_ - 1) / TH_), (unsigned)nf), \
                       dim3(kSobelW), 0, s, y, w, h, yt, seed, (w + 63) / 64, threshold, g0)
#define LGCR_SOBEL(TH_)                 \
    {                                   \
        if (go)                         \
            LGCR_SOBEL_G(TH_, true);    \
        else                            \
            LGCR_SOBEL_G(TH_, false);   \
    }
    if (tile_rows == 8)
        LGCR_SOBEL(8)
    else if (tile_rows == 32)
        LGCR_SOBEL(32)
    else
        LGCR_SOBEL(16)
#undef LGCR_SOBEL
#undef LGCR_SOBEL_G
    return hipGetLastError();
}

hipError_t launch_lc_map(const float *y, int w, int h, float *lc, int cw, int ch, const int *xb,
                         const int *xe, const int *yb, const int *ye, hipStream_t s, const float *u, const float *v,
                         float4 *luv, const GroupOff *go, int nf) {
    (void)h;
    if (go)
        hipLaunchKernelGGL(k_lc_map<true>, rows_grid(cw, ch, nf), dim3(kBlock), 0, s, y, w, lc, cw, ch, xb, xe, yb,
                           ye, u, v, luv, *go);
    else
        hipLaunchKernelGGL(k_lc_map<false>, rows_grid(cw, ch), dim3(kBlock), 0, s, y, w, lc, cw, ch, xb, xe, yb, ye,
                           u, v, luv, GroupOff{});
    return hipGetLastError();
}

hipError_t launch_trust_workset(const uint64_t *seed, uint64_t *tmp, uint64_t *mask, int w, int h, int radius,
                                uint32_t *row_counts, uint32_t *row_offsets, uint32_t *indices, uint8_t *chroma_mask,
                                int cw, int ch, const AxisDev &ax, const AxisDev &ay, uint32_t *counters,
                                hipStream_t s) {
    const int w64 = (w + 63) / 64;
    const size_t nc = (size_t)cw * ch;
    hipLaunchKernelGGL(k_dilate_hbits, dim3((unsigned)((w64 + kBlock - 1) / kBlock), (unsigned)h), dim3(kBlock), 0, s,
                       seed, tmp, w, w64, h, radius, chroma_mask, nc, counters);
    const bool same = ax.n == w && ay.n == h;
    hipLaunchKernelGGL(k_dilate_vbits, dim3((unsigned)h), dim3(64), 0, s, tmp, mask, w64, h, radius,
                       same ? row_counts : nullptr);
    const MaskBits m{mask, w, h, w64, ax.n, ay.n};
    if (!same)
        hipLaunchKernelGGL(k_row_count_bits, dim3((unsigned)ay.n), dim3(kBlock), 0, s, m, row_counts);
    hipLaunchKernelGGL(k_compact_bits, dim3((unsigned)ay.n), dim3(kBlock), 0, s, m, row_counts, row_offsets, indices,
                       ax.cb_i0, ay.cb_i0, chroma_mask, cw, ch);
    hipLaunchKernelGGL(k_count_u8x16, dim3((unsigned)std::min<size_t>(256, (nc / 16 + kBlock - 1) / kBlock + 1)),
                       dim3(kBlock), 0, s, chroma_mask, nc, counters);
    return hipGetLastError();
}

size_t trust_fused_lds(int w, int radius, int rows, int cw) {
    const size_t w64 = (size_t)(w + 63) / 64, srows = (size_t)rows + 2 * (size_t)radius;
    return (srows * (w64 + 2) + (size_t)rows * (w64 + 2) + (size_t)rows * w64) * 8 + (size_t)cw * 8;
}

hipError_t launch_trust_mask_tiles(const uint64_t *seed, uint64_t *mask, int w, int h, int radius, const int4 *band,
                                   int nbands, int cband, int max_rows, uint8_t *chroma_mask, int cw, int ch,
                                   const int2 *xr, const int2 *yr, uint32_t *counters, hipStream_t s,
                                   const GroupOff *go, int nf) {
    const int w64 = (w + 63) / 64;
    const size_t lds = trust_fused_lds(w, radius, max_rows, cw);
    if (go)
        hipLaunchKernelGGL(k_trust_fused<true>, dim3((unsigned)nbands, 1, (unsigned)nf), dim3(kBlock), lds, s, seed,
                           mask, w, w64, h, radius, band, cband, cw, ch, xr, yr, chroma_mask, counters, *go);
    else
        hipLaunchKernelGGL(k_trust_fused<false>, dim3((unsigned)nbands), dim3(kBlock), lds, s, seed, mask, w, w64, h,
                           radius, band, cband, cw, ch, xr, yr, chroma_mask, counters, GroupOff{});
    return hipGetLastError();
}

bool guided_tiled(int tiles_fit, int radial, int cedge, const AxisDev &ax, const AxisDev &ay, int sel_mode) {
    return tiles_fit && !radial && !cedge && ax.sup == 6 && ay.sup == 6 && ax.rec && ay.rec && sel_mode != SEL_FLOAT;
}

hipError_t launch_mutual_gate(const float *lc, const float *u, const float *v, int w, int h,
                              const uint8_t *chroma_mask, const uint32_t *count, int ncount, float *gate,
                              hipStream_t s, int tile_cols, const GroupOff *go, int nf) {
    const GroupOff g0 = go ? *go : GroupOff{};
#define LGCR_GATE_G(TW_, TH_, G_)                                                                              \
    hipLaunchKernelGGL((k_gate_fused<TW_, TH_, G_>),                                                           \
                       dim3((unsigned)((w + TW_ - 1) / TW_), (unsigned)((h + TH_ - 1) / TH_), (unsigned)nf),   \
                       dim3(kBlock), 0, s, lc, u, v, w, h, chroma_mask, count, ncount, gate, g0)
#define LGCR_GATE(TW_, TH_)               \
    {                                     \
        if (go)                           \
            LGCR_GATE_G(TW_, TH_, true);  \
        else                              \
            LGCR_GATE_G(TW_, TH_, false); \
    }
    if (tile_cols == 64)
        LGCR_GATE(64, 4)
    else if (tile_cols == 16)
        LGCR_GATE(16, 16)
    else
        LGCR_GATE(32, 8)
#undef LGCR_GATE
#undef LGCR_GATE_G
    return hipGetLastError();
}

hipError_t launch_plain_separable(const float *u, const float *v, int cw, int ch, const AxisDev &ax,
                                  const AxisDev &ay, float *hbuf, float *out_u, float *out_v, int clamp_hull,
                                  float margin, hipStream_t s) {
    if (ax.sup == 6 && ay.sup == 6 && ax.rec && ay.tile_rows16 > 0 && ay.tile_rows16 <= kPlainRows) {
        const dim3 g((unsigned)((ax.n + kPlainCols - 1) / kPlainCols), (unsigned)((ay.n + kPlainTileRows - 1) / kPlainTileRows));
        hipLaunchKernelGGL(k_plain_fused, g, dim3(kBlock), 0, s, u, v, cw, ch, ax, ay, out_u, out_v, clamp_hull,
                           margin);
        return hipGetLastError();
    }
    if (ax.sup == 6 && ax.rec)
        hipLaunchKernelGGL(k_plain_h<6>, rows_grid(ax.n, ch), dim3(kBlock), 0, s, u, v, cw, ch, ax, hbuf);
    else if (ax.sup == 8)
        hipLaunchKernelGGL(k_plain_h<8>, rows_grid(ax.n, ch), dim3(kBlock), 0, s, u, v, cw, ch, ax, hbuf);
    else
        hipLaunchKernelGGL(k_plain_h<0>, rows_grid(ax.n, ch), dim3(kBlock), 0, s, u, v, cw, ch, ax, hbuf);
    if (ay.sup == 6)
        hipLaunchKernelGGL(k_plain_v<6>, rows_grid(ax.n, ay.n), dim3(kBlock), 0, s, u, v, cw, ch, ax, ay, hbuf,
                           out_u, out_v, clamp_hull, margin);
    else if (ay.sup == 8)
        hipLaunchKernelGGL(k_plain_v<8>, rows_grid(ax.n, ay.n), dim3(kBlock), 0, s, u, v, cw, ch, ax, ay, hbuf,
                           out_u, out_v, clamp_hull, margin);
    else
        hipLaunchKernelGGL(k_plain_v<0>, rows_grid(ax.n, ay.n), dim3(kBlock), 0, s, u, v, cw, ch, ax, ay, hbuf,
                           out_u, out_v, clamp_hull, margin);
    return hipGetLastError();
}

hipError_t launch_plain_bilinear(const float *u, const float *v, int cw, int ch, const AxisDev &ax,
                                 const AxisDev &ay, float *out_u, float *out_v, hipStream_t s) {
    hipLaunchKernelGGL(k_plain_bilinear, rows_grid(ax.n, ay.n), dim3(kBlock), 0, s, u, v, cw, ch, ax, ay, out_u,
                       out_v);
    return hipGetLastError();
}

hipError_t launch_guided(const float *u, const float *v, int cw, int ch, const float *y, int sw, int sh,
                         const float4 *yt, const float4 *luv, const float *lc, const float *ms,
                         const AxisDev &ax, const AxisDev &ay, const RadialDev &rd, const GuideParams &gp,
                         const uint32_t *row_offsets, const uint32_t *indices, float *out_u, float *out_v,
                         const SelectorDev &sel, const TemporalDev &tmp, hipStream_t s, const uint64_t *mask_bits,
                         int mask_w, int mask_h, int tiles_fit, int plain_in_tiles, const RawOut *fused_out,
                         int out_type, float bp, int bp_on, GuidedDeferred *defer) {
    GuidedArgs A{u, v, cw, ch, y, sw, sh, yt, luv, lc, ms, ax, ay, rd, gp, row_offsets, indices, out_u, out_v,
                 sel, tmp, indices ? mask_bits : nullptr, mask_w, mask_h, (mask_w + 63) / 64};
    if (guided_tiled(tiles_fit, gp.radial, gp.cedge, ax, ay, sel.mode)) {
        const int tiles_x = (ax.n + kTile - 1) / kTile, tiles = tiles_x * ((ay.n + kTile - 1) / kTile);
        const dim3 g((unsigned)((tiles + kBlock / 64 - 1) / (kBlock / 64)));
#define LGCR_TILES(T, S_, P)                                                                                      \
    do {                                                                                                        \
        if (P && gp.finite)                                                                                     \
            hipLaunchKernelGGL((k_guided_tiles<T, S_, P, 0, P>), g, dim3(kBlock), 0, s, A, tiles_x, tiles);     \
        else                                                                                                    \
            hipLaunchKernelGGL((k_guided_tiles<T, S_, P, 0, false>), g, dim3(kBlock), 0, s, A, tiles_x, tiles); \
    } while (0)
        if (fused_out) {
            // the caller checked: no selector, and either the plain base in
            // the tiles without temporal taps (bp_on only for same-size 4:2:0
            // with even sizes) or a dense temporal pass without bp
            const bool recon = plain_in_tiles && tmp.n == 0 && !(bp_on && ((ax.n | ay.n) & 1));
            const bool trecon = !plain_in_tiles && tmp.n > 0 && !A.mbits && !bp_on && sel.mode == SEL_NONE;
            if (sel.mode == SEL_FLOAT || !(recon || trecon))
                return hipErrorInvalidValue;
            A.po_u = fused_out[0];
            A.po_v = fused_out[1];
            A.bp = bp;
            A.bp_on = bp_on;
            if (defer && sel.mode == SEL_NONE) {
                const bool ph = recon && gp.finite;
                BatchFrame bf{A, tiles_x, tiles};
                memcpy(defer->args, &bf, sizeof bf);
                defer->variant = batch_variant(!recon, ph, out_type == RAW_U8 ? 1 : out_type == RAW_U16 ? 2 : 3);
                defer->tiles = tiles;
                return hipSuccess;
            }
#define LGCR_FUSED(T, S_, P, H)                                                                                   \
    if (out_type == RAW_U8)                                                                                     \
        hipLaunchKernelGGL((k_guided_tiles<T, S_, P, 1, H>), g, dim3(kBlock), 0, s, A, tiles_x, tiles);         \
    else if (out_type == RAW_U16)                                                                               \
        hipLaunchKernelGGL((k_guided_tiles<T, S_, P, 2, H>), g, dim3(kBlock), 0, s, A, tiles_x, tiles);         \
    else                                                                                                        \
        hipLaunchKernelGGL((k_guided_tiles<T, S_, P, 3, H>), g, dim3(kBlock), 0, s, A, tiles_x, tiles);
            if (sel.mode == SEL_MAPS) {
                // algo 4 with integer output: the selector of every active
                // pixel is final in the tile (its LGF maps are built before)
                if (gp.finite) {
                    LGCR_FUSED(false, SEL_MAPS, true, true)
                } else {
                    LGCR_FUSED(false, SEL_MAPS, true, false)
                }
            } else if (recon && gp.finite) {
                LGCR_FUSED(false, SEL_NONE, true, true)
            } else if (recon) {
                LGCR_FUSED(false, SEL_NONE, true, false)
            } else {
                LGCR_FUSED(true, SEL_NONE, false, false)
            }
#undef LGCR_FUSED
            return hipGetLastError();
        }
        if (tmp.n > 0) {
            if (plain_in_tiles)
                LGCR_TILES(true, SEL_NONE, true);
            else
                LGCR_TILES(true, SEL_NONE, false);
        } else if (sel.mode == SEL_MAPS) {
            if (plain_in_tiles)
                LGCR_TILES(false, SEL_MAPS, true);
            else
                LGCR_TILES(false, SEL_MAPS, false);
        } else {
            if (plain_in_tiles)
                LGCR_TILES(false, SEL_NONE, true);
            else
                LGCR_TILES(false, SEL_NONE, false);
        }
#undef LGCR_TILES
        return hipGetLastError();
    }
    const size_t n = (size_t)ax.n * ay.n;
    const unsigned blocks = (unsigned)std::min<size_t>(8192, (n + kBlock - 1) / kBlock);
    const bool radial = gp.radial != 0, temporal = tmp.n > 0, cedge = gp.cedge != 0;
    // 6 taps per axis (Lanczos 3 / Spline36 at chroma upsampling) get fully
    // unrolled tap loops
    const int sup6 = (!radial && !cedge && ax.sup == 6 && ay.sup == 6 && ax.rec && ay.rec) ? 6 : 0;
#define LGCR_GUIDED(R, T, S_, C) \
    if (radial == R && temporal == T && sel.mode == S_ && cedge == C) { \
        if (!R && !C && sup6 == 6) \
            hipLaunchKernelGGL((k_guided<R, T, S_, C, (R || C) ? 0 : 6>), dim3(blocks), dim3(kBlock), 0, s, A); \
        else \
            hipLaunchKernelGGL((k_guided<R, T, S_, C, 0>), dim3(blocks), dim3(kBlock), 0, s, A); \
        return hipGetLastError(); \
    }
#define LGCR_GUIDED_SEL(R, T, C) LGCR_GUIDED(R, T, SEL_NONE, C) LGCR_GUIDED(R, T, SEL_MAPS, C) \
    LGCR_GUIDED(R, T, SEL_FLOAT, C)
    LGCR_GUIDED_SEL(false, false, false)
    LGCR_GUIDED_SEL(false, false, true)
    LGCR_GUIDED_SEL(true, false, false)
    LGCR_GUIDED_SEL(true, false, true)
    LGCR_GUIDED(false, true, SEL_NONE, false)
    LGCR_GUIDED(false, true, SEL_NONE, true)
    LGCR_GUIDED(true, true, SEL_NONE, false)
    LGCR_GUIDED(true, true, SEL_NONE, true)
#undef LGCR_GUIDED_SEL
#undef LGCR_GUIDED
    return hipErrorInvalidValue;
}

size_t guided_batch_table_bytes() {
    return sizeof(BatchFrame) * kBatchMaxFrames + sizeof(uint32_t) * kBatchMaxFrames;
}

template <bool T, bool P, int O, bool H>
static hipError_t batch_launch_one(const BatchFrame *frames, const uint32_t *first, int n, uint32_t total,
                                   int max_tiles, bool persistent, hipStream_t s) {
    auto kern = k_guided_tiles_batch<T, SEL_NONE, P, O, H>;
    // resident workgroups of the device, once per variant (one device type)
    struct Fill {
        hipError_t err;
        int blocks;
    };
    static const Fill fill = [&]() {
        int dev = 0, b = 0, c = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess)
            e = hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kern, kBlock, 0);
        return Fill{e, std::max(1, b) * std::max(1, c)};
    }();
    if (fill.err != hipSuccess)
        return fill.err;
    if (!persistent) {
        hipLaunchKernelGGL(kern, dim3((unsigned)((max_tiles + kBlock / 64 - 1) / (kBlock / 64)), (unsigned)n),
                           dim3(kBlock), 0, s, frames, first, n, total, 0u);
        return hipGetLastError();
    }
    const uint32_t need = (total + kBlock / 64 - 1) / (kBlock / 64);
    const uint32_t blocks = std::min<uint32_t>(need, (uint32_t)fill.blocks);
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), 0, s, frames, first, n, total, blocks * (kBlock / 64));
    return hipGetLastError();
}

hipError_t launch_guided_batch(const GuidedDeferred *items, int n, void *dev, void *host, bool persistent,
                               hipStream_t s) {
    static_assert(sizeof(BatchFrame) <= kGuidedDeferBytes, "deferred frame storage too small");
    if (n <= 0)
        return hipSuccess;
    if (n > kBatchMaxFrames)
        return hipErrorInvalidValue;
    const int variant = items[0].variant;
    // table layout: first[kBatchMaxFrames], then the frames (one upload)
    static_assert(sizeof(uint32_t) * kBatchMaxFrames % alignof(BatchFrame) == 0, "frames after first[] aligned");
    uint32_t *hfirst = static_cast<uint32_t *>(host);
    BatchFrame *hf = reinterpret_cast<BatchFrame *>(hfirst + kBatchMaxFrames);
    uint32_t total = 0;
    int max_tiles = 0;
    for (int k = 0; k < n; ++k) {
        if (items[k].variant != variant || items[k].tiles <= 0)
            return hipErrorInvalidValue;
        max_tiles = std::max(max_tiles, items[k].tiles);
        memcpy(&hf[k], items[k].args, sizeof(BatchFrame));
        hfirst[k] = total;
        total += (uint32_t)items[k].tiles;
    }
    uint32_t *dfirst = static_cast<uint32_t *>(dev);
    BatchFrame *df = reinterpret_cast<BatchFrame *>(dfirst + kBatchMaxFrames);
    const hipError_t e = hipMemcpyAsync(dfirst, hfirst, sizeof(uint32_t) * kBatchMaxFrames + sizeof(BatchFrame) * n,
                                        hipMemcpyHostToDevice, s);
    if (e != hipSuccess)
        return e;
#define LGCR_BATCH(T, P, O, H) \
    case batch_variant(T, H, O): return batch_launch_one<T, P, O, H>(df, dfirst, n, total, max_tiles, persistent, s);
    switch (variant) {
        LGCR_BATCH(false, true, 1, true)
        LGCR_BATCH(false, true, 2, true)
        LGCR_BATCH(false, true, 3, true)
        LGCR_BATCH(false, true, 1, false)
        LGCR_BATCH(false, true, 2, false)
        LGCR_BATCH(false, true, 3, false)
        LGCR_BATCH(true, false, 1, false)
        LGCR_BATCH(true, false, 2, false)
        LGCR_BATCH(true, false, 3, false)
    default: return hipErrorInvalidValue;
    }
#undef LGCR_BATCH
}

hipError_t launch_plain_radial(const float *u, const float *v, int cw, int ch, const AxisDev &ax,
                               const AxisDev &ay, const RadialDev &rd, float *out_u, float *out_v, int clamp_hull,
                               float margin, hipStream_t s) {
    if (rd.valid && rd.max_count > 0) {
        const dim3 g((unsigned)((rd.max_count + kBlock - 1) / kBlock), (unsigned)ay.n, (unsigned)rd.nx);
        if (ax.sup == 6 && ay.sup == 6) {
            constexpr int R = kRadialRows;
            static_assert(kSegCols == kBlock, "one output column per thread in the segment write");
            const dim3 gr((unsigned)rd.nseg, (unsigned)((ay.n + R - 1) / R));
            if (rd.tile_fit)
                hipLaunchKernelGGL((k_plain_radial_rows<R, true>), gr, dim3(kBlock), 0, s, u, v, cw, ch, ax, ay, rd,
                                   out_u, out_v, clamp_hull, margin);
            else
                hipLaunchKernelGGL((k_plain_radial_rows<R, false>), gr, dim3(kBlock), 0, s, u, v, cw, ch, ax, ay, rd,
                                   out_u, out_v, clamp_hull, margin);
        } else
            hipLaunchKernelGGL(k_plain_radial_tab<0>, g, dim3(kBlock), 0, s, u, v, cw, ch, ax, ay, rd, out_u, out_v,
                               clamp_hull, margin);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_plain_radial, rows_grid(ax.n, ay.n), dim3(kBlock), 0, s, u, v, cw, ch, ax, ay, rd, out_u,
                       out_v, clamp_hull, margin);
    return hipGetLastError();
}

hipError_t launch_point_luma(const float *y, int w, int h, float *ls, int cw, int ch, const int *xi, const float *xf,
                             const int *yi, const float *yf, hipStream_t s) {
    hipLaunchKernelGGL(k_point_luma, rows_grid(cw, ch), dim3(kBlock), 0, s, y, w, h, ls, cw, ch, xi, xf, yi, yf);
    return hipGetLastError();
}

hipError_t launch_minmax5(const float *u, const float *v, int cw, int ch, float *mm, hipStream_t s) {
    hipLaunchKernelGGL(k_minmax5, rows_grid(cw, ch), dim3(kBlock), 0, s, u, v, cw, ch, mm);
    return hipGetLastError();
}

hipError_t launch_lgf(const float *ls, const float *u, const float *v, int cw, int ch, double eps, int cedge,
                      const float *mm, const uint8_t *roi, const uint32_t *roi_count, const float *gate, float ms,
                      float *lgf, hipStream_t s) {
    hipLaunchKernelGGL(k_lgf, rows_grid(cw, ch), dim3(kBlock), 0, s, ls, u, v, cw, ch, eps, cedge, mm, roi,
                       roi_count, gate, ms, lgf);
    return hipGetLastError();
}

hipError_t launch_roi_mark(const uint32_t *row_offsets, const uint32_t *indices, int ow, int oh,
                           const SelectorDev &sel, const AxisDev &ax, const AxisDev &ay, uint8_t *roi, int cw, int ch,
                           uint32_t *roi_count, hipStream_t s) {
    const size_t nc = (size_t)cw * ch, n = (size_t)ow * oh;
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)std::min<size_t>(1024, (nc + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       s, roi, nc, (uint8_t)0);
    const WorkList L{row_offsets, indices, ow, oh};
    hipLaunchKernelGGL(k_roi_mark, dim3((unsigned)std::min<size_t>(8192, (n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       s, L, sel, ax, ay, roi, cw, ch);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, roi_count, 1);
    hipLaunchKernelGGL(k_count_u8, dim3((unsigned)std::min<size_t>(512, (nc + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       s, roi, nc, roi_count);
    return hipGetLastError();
}

hipError_t launch_blend4(const uint32_t *row_offsets, const uint32_t *indices, const float *y, int sw, int sh,
                         const AxisDev &ax, const AxisDev &ay, int cw, int ch, const SelectorDev &sel, float *out_u,
                         float *out_v, hipStream_t s) {
    const size_t n = (size_t)ax.n * ay.n;
    const WorkList L{row_offsets, indices, ax.n, ay.n};
    hipLaunchKernelGGL(k_blend4, dim3((unsigned)std::min<size_t>(8192, (n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       s, L, y, sw, sh, ax, ay, cw, ch, sel, out_u, out_v);
    return hipGetLastError();
}

hipError_t launch_affine(const float *u, const float *v, int cw, int ch, const float *yc0, const float *yc1,
                         const float *yc2, const uint8_t *chroma_mask, const uint32_t *chroma_count, double eps,
                         double eps_sig, const AffineDev &m, hipStream_t s) {
    hipLaunchKernelGGL(k_affine, rows_grid(cw, ch), dim3(kBlock), 0, s, u, v, cw, ch, yc0, yc1, yc2, chroma_mask,
                       chroma_count, eps, eps_sig, m);
    return hipGetLastError();
}

hipError_t launch_detail_sub(const float *y, float *rec, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_detail_sub, dim3((unsigned)std::min<size_t>(4096, (n + kBlock - 1) / kBlock)), dim3(kBlock),
                       0, s, y, rec, n);
    return hipGetLastError();
}

hipError_t launch_nyquist(float *a, float *b, int w, int h, int xlevels, int ylevels, float **result, hipStream_t s) {
    float *cur = a, *nxt = b;
    for (int pass = 0; pass < xlevels; ++pass) {
        hipLaunchKernelGGL(k_nyq, rows_grid(w, h), dim3(kBlock), 0, s, cur, nxt, w, h, 1 << pass, 0);
        std::swap(cur, nxt);
    }
    for (int pass = 0; pass < ylevels; ++pass) {
        hipLaunchKernelGGL(k_nyq, rows_grid(w, h), dim3(kBlock), 0, s, cur, nxt, w, h, 1 << pass, 1);
        std::swap(cur, nxt);
    }
    *result = cur;
    return hipGetLastError();
}

hipError_t launch_detail_transfer(const uint32_t *row_offsets, const uint32_t *indices, const float4 *yt,
                                  int sw, int sh, const AxisDev &ax,
                                  const AxisDev &ay, int cw, int ch, const AffineDev &m, const float *ms,
                                  const float *detail, float ms_strength, float qgate, float strength, float stretch,
                                  float ar, int clamp_hull, float *out_u, float *out_v, hipStream_t s) {
    const size_t n = (size_t)ax.n * ay.n;
    const WorkList L{row_offsets, indices, ax.n, ay.n};
    hipLaunchKernelGGL(k_detail_transfer, dim3((unsigned)std::min<size_t>(8192, (n + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, L, yt, sw, sh, ax, ay, cw, ch, m, ms, detail, ms_strength,
                       qgate, strength, stretch, ar, clamp_hull, out_u, out_v);
    return hipGetLastError();
}

hipError_t launch_block_match(const float *cur, const float *nbr, int w, int h, int block, int search, float tsad,
                              int16_t *mvx, int16_t *mvy, float *conf, hipStream_t s, const GroupOff *go, int nf) {
    if (block > 64)
        return hipErrorInvalidValue;  // LDS block cache holds 32x32 subsampled samples
    const int bw = (w + block - 1) / block, bh = (h + block - 1) / block;
    const int half = (block + 1) / 2, ws = block + 2 * search;
    const size_t lds = sizeof(float) * ((size_t)half * half + (size_t)ws * ws);
    const unsigned nb = (unsigned)(bw * bh);
    const GroupOff g0 = go ? *go : GroupOff{};
#define LGCR_ME(G_)                                                                                              \
    {                                                                                                            \
        if (4 * lds <= (size_t)kMeLdsMax)                                                                        \
            hipLaunchKernelGGL((k_block_match_win<4, G_>), dim3((nb + 3) / 4, 1, (unsigned)nf), dim3(256), 4 * lds, \
                               s, cur, nbr, w, h, block, search, tsad, mvx, mvy, conf, g0);                       \
        else if (lds <= (size_t)kMeLdsMax)                                                                       \
            hipLaunchKernelGGL((k_block_match_win<1, G_>), dim3(nb, 1, (unsigned)nf), dim3(64), lds, s, cur, nbr, w, \
                               h, block, search, tsad, mvx, mvy, conf, g0);                                      \
        else                                                                                                     \
            hipLaunchKernelGGL((k_block_match<G_>), dim3(nb, 1, (unsigned)nf), dim3(kMeThreads), 0, s, cur, nbr, w, \
                               h, block, search, tsad, mvx, mvy, conf, g0);                                      \
    }
    if (go)
        LGCR_ME(true)
    else
        LGCR_ME(false)
#undef LGCR_ME
    return hipGetLastError();
}

hipError_t launch_chroma_consistency(const float *cu, const float *cv, const float *nu, const float *nv, int cw,
                                     int ch, int luma_w, int luma_h, int block, const int16_t *mvx,
                                     const int16_t *mvy, int bw, int bh, float *conf, hipStream_t s, const GroupOff *go,
                                     int nf) {
    if (go)
        hipLaunchKernelGGL(k_chroma_consistency<true>, dim3((unsigned)(bw * bh), 1, (unsigned)nf), dim3(kCcThreads), 0,
                           s, cu, cv, nu, nv, cw, ch, luma_w, luma_h, block, mvx, mvy, bw, conf, *go);
    else
        hipLaunchKernelGGL(k_chroma_consistency<false>, dim3((unsigned)(bw * bh)), dim3(kCcThreads), 0, s, cu, cv, nu,
                           nv, cw, ch, luma_w, luma_h, block, mvx, mvy, bw, conf, GroupOff{});
    return hipGetLastError();
}

hipError_t launch_bp_pack(float *out_u, float *out_v, int ow, int oh, const float *u, const float *v, int cw,
                          int ch, float bp, const RawOut *planes, int type, int keep_float, hipStream_t s) {
    OutBatch b{};
    b.p[0] = planes[0];
    b.p[1] = planes[1];
    const dim3 g = rows_grid(cw, ch);
    if (type == RAW_U8)
        hipLaunchKernelGGL(k_bp_pack<uint8_t>, g, dim3(kBlock), 0, s, out_u, out_v, ow, oh, u, v, cw, ch, bp, b,
                           keep_float);
    else if (type == RAW_U16)
        hipLaunchKernelGGL(k_bp_pack<uint16_t>, g, dim3(kBlock), 0, s, out_u, out_v, ow, oh, u, v, cw, ch, bp, b,
                           keep_float);
    else
        hipLaunchKernelGGL(k_bp_pack<float>, g, dim3(kBlock), 0, s, out_u, out_v, ow, oh, u, v, cw, ch, bp, b,
                           keep_float);
    return hipGetLastError();
}

hipError_t launch_io_advance(IoTable *io, hipStream_t s) {
    hipLaunchKernelGGL(k_io_advance, dim3(1), dim3(64), 0, s, io);
    return hipGetLastError();
}

hipError_t launch_back_project(float *out_u, float *out_v, int ow, int oh, const float *u, const float *v,
                               int cw, int ch, float bp, hipStream_t s) {
    hipLaunchKernelGGL(k_back_project, rows_grid(cw, ch), dim3(kBlock), 0, s, out_u, out_v, ow, oh, u, v, cw,
                       ch, bp);
    return hipGetLastError();
}

hipError_t launch_resample_h(const float *src, int sw, int sh, float *dst, const WeightsDev &t, hipStream_t s) {
    const dim3 g = rows_grid(t.n, sh);
    if (t.sup == 4)
        hipLaunchKernelGGL(k_resample_h<4>, g, dim3(kBlock), 0, s, src, sw, sh, dst, t);
    else if (t.sup == 6)
        hipLaunchKernelGGL(k_resample_h<6>, g, dim3(kBlock), 0, s, src, sw, sh, dst, t);
    else if (t.sup == 8)
        hipLaunchKernelGGL(k_resample_h<8>, g, dim3(kBlock), 0, s, src, sw, sh, dst, t);
    else
        hipLaunchKernelGGL(k_resample_h<0>, g, dim3(kBlock), 0, s, src, sw, sh, dst, t);
    return hipGetLastError();
}

hipError_t launch_resample_v(const float *src, int sw, int sh, float *dst, const WeightsDev &t, hipStream_t s) {
    const dim3 g = rows_grid(sw, t.n);
    if (t.sup == 4)
        hipLaunchKernelGGL(k_resample_v<4>, g, dim3(kBlock), 0, s, src, sw, sh, dst, t);
    else if (t.sup == 6)
        hipLaunchKernelGGL(k_resample_v<6>, g, dim3(kBlock), 0, s, src, sw, sh, dst, t);
    else if (t.sup == 8)
        hipLaunchKernelGGL(k_resample_v<8>, g, dim3(kBlock), 0, s, src, sw, sh, dst, t);
    else
        hipLaunchKernelGGL(k_resample_v<0>, g, dim3(kBlock), 0, s, src, sw, sh, dst, t);
    return hipGetLastError();
}

hipError_t launch_resample_radial_tab(const float *src, int sw, int sh, float *dst, int dw, int dh,
                                      const RadialResampleDev &t, hipStream_t s) {
    const dim3 g((unsigned)((t.max_count + kBlock - 1) / kBlock), (unsigned)dh, (unsigned)t.ncx);
    if (t.sup == 6) {
        constexpr int R = kRadialRows;
        const dim3 gr((unsigned)t.nseg, (unsigned)((dh + R - 1) / R));
        hipLaunchKernelGGL(k_resample_radial_rows<R>, gr, dim3(kBlock), 0, s, src, sw, sh, dst, dw, dh, t);
    } else
        hipLaunchKernelGGL(k_resample_radial_tab<0>, g, dim3(kBlock), 0, s, src, sw, sh, dst, dw, t);
    return hipGetLastError();
}

hipError_t launch_resample_radial(const float *src, int sw, int sh, float *dst, int dw, int dh, double support,
                                  const float *lut, int lut_n, float lut_scale, hipStream_t s) {
    const double sx = (double)sw / dw, sy = (double)sh / dh;
    const double widen = fmax(1.0, fmax(sx, sy));
    const float inv_w = (float)(1.0 / widen);
    const int radius = (int)ceil(support * widen);
    hipLaunchKernelGGL(k_resample_radial, rows_grid(dw, dh), dim3(kBlock), 0, s, src, sw, sh, dst, dw, dh, sx, sy,
                       inv_w, radius, lut, lut_n, lut_scale);
    return hipGetLastError();
}

hipError_t launch_sharpen(const float *src, float *dst, int w, int h, const float *lc, const float *gy, int gw,
                          int gh, double rw, double rh, double sx, double sy, const SharpenParams &p,
                          hipStream_t s) {
    SharpenArgs A{src, dst, w, h, lc, gy, gw, gh, rw, rh, sx, sy, p.inv_g2, p.sigma, p.sratio, p.alpha, p.margin,
                  p.clamp_hull};
    static_assert(kShW * kShH == kBlock, "one pixel per thread");
    hipLaunchKernelGGL(k_sharpen, dim3((unsigned)((w + kShW - 1) / kShW), (unsigned)((h + kShH - 1) / kShH)),
                       dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

// Every 8x8 output tile's taps fit a kTWin x kTWin chroma window.
int guided_tiles_fit(const int *xs, int ow, const int *ys, int oh) {
    auto fit_rows = [](const int *st, int n, int rows) {
        for (int t0 = 0; t0 < n; t0 += rows) {
            int lo = INT_MAX, hi = INT_MIN;
            for (int c = 0; c < rows; ++c) {
                const int v = st[std::min(t0 + c, n - 1)];
                lo = std::min(lo, v);
                hi = std::max(hi, v);
            }
            if (hi - lo + 6 > kTWin)
                return false;
        }
        return true;
    };
    return fit_rows(xs, ow, kTile) && fit_rows(ys, oh, kTile) ? 1 : 0;
}

size_t bm_scratch_bytes(int w, int h) {
    if (w < kBmP || h < kBmP)
        return 0;
    const int nx = (w - kBmP) / 8 + 1 + ((w - kBmP) % 8 != 0), ny = (h - kBmP) / 8 + 1 + ((h - kBmP) % 8 != 0);
    const size_t na = (size_t)nx * ny;
    return na * (sizeof(int) * (1 + 2 * kBmG) + sizeof(float) * (1 + 2 * kBmG * 64)) + 4 * 256;
}

hipError_t launch_collaborative(const float *guide, int gw, int gh, float *u, float *v, int w, int h,
                                float guide_sigma, float *guide_tmp, void *scratch, hipStream_t s) {
    if (w < kBmP || h < kBmP)
        return hipSuccess;
    const int nx = (w - kBmP) / 8 + 1 + ((w - kBmP) % 8 != 0), ny = (h - kBmP) / 8 + 1 + ((h - kBmP) % 8 != 0);
    const size_t na = (size_t)nx * ny;
    auto carve = [](uint8_t *&p, size_t bytes) {
        uint8_t *r = p;
        p += (bytes + 255) & ~size_t(255);
        return r;
    };
    uint8_t *p = static_cast<uint8_t *>(scratch);
    BmArgs A;
    A.gc = reinterpret_cast<int *>(carve(p, na * sizeof(int)));
    A.pos = reinterpret_cast<int *>(carve(p, na * 2 * kBmG * sizeof(int)));
    A.gw = reinterpret_cast<float *>(carve(p, na * sizeof(float)));
    A.filt = reinterpret_cast<float *>(carve(p, na * 2 * kBmG * 64 * sizeof(float)));
    A.u = u;
    A.v = v;
    A.w = w;
    A.h = h;
    A.nx = nx;
    A.ny = ny;
    // bm.cpp:282-284
    const float sigma = fminf(fmaxf(0.6f * guide_sigma, 0.002f), 0.012f);
    A.thr = 2.7f * sigma;
    A.limit = std::max(4e-5f, 4.0f * guide_sigma * guide_sigma);
    A.g = guide;
    if (gw != w || gh != h) {
        hipLaunchKernelGGL(k_bm_guide, rows_grid(w, h), dim3(kBlock), 0, s, guide, gw, gh, guide_tmp, w, h,
                           (double)gw / w, (double)gh / h);
        A.g = guide_tmp;
    }
    hipLaunchKernelGGL(k_bm_group, dim3((unsigned)na), dim3(64), 0, s, A);
    hipLaunchKernelGGL(k_bm_aggregate, dim3((unsigned)((w + 8 * kBmAggTiles - 1) / (8 * kBmAggTiles)), (unsigned)((h + 7) / 8)),
                       dim3(kBlock), 0, s, A);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Downsample launchers
// ---------------------------------------------------------------------------

#define LGCR_DS_TYPED(K, grid)                                                                  \
    if (a.type == RAW_U8)                                                                       \
        hipLaunchKernelGGL(K<uint8_t>, grid, dim3(kBlock), 0, s, a);                            \
    else if (a.type == RAW_U16)                                                                 \
        hipLaunchKernelGGL(K<uint16_t>, grid, dim3(kBlock), 0, s, a);                           \
    else                                                                                        \
        hipLaunchKernelGGL(K<float>, grid, dim3(kBlock), 0, s, a);

template <int SX, int SY> void launch_ds_base_tiled(const DsArgs &a, dim3 g, hipStream_t s) {
    if (a.type == RAW_U8)
        hipLaunchKernelGGL((k_ds_base<uint8_t, SX, SY>), g, dim3(kBlock), 0, s, a);
    else if (a.type == RAW_U16)
        hipLaunchKernelGGL((k_ds_base<uint16_t, SX, SY>), g, dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL((k_ds_base<float, SX, SY>), g, dim3(kBlock), 0, s, a);
}

hipError_t launch_ds_base(const DsArgs &a, hipStream_t s) {
    const dim3 g((unsigned)((a.cw + kDsBaseTileW - 1) / kDsBaseTileW), (unsigned)((a.ch + kDsBaseTileH - 1) / kDsBaseTileH));
    const int sx = a.xw.sup, sy = a.yw.sup;
    if (sx == 12 && sy == 12)
        launch_ds_base_tiled<12, 12>(a, g, s);
    else if (sx == 4 && sy == 4)
        launch_ds_base_tiled<4, 4>(a, g, s);
    else if (sx == 4 && sy == 5)
        launch_ds_base_tiled<4, 5>(a, g, s);
    else if (sx == 5 && sy == 4)
        launch_ds_base_tiled<5, 4>(a, g, s);
    else if (sx == 5 && sy == 5)
        launch_ds_base_tiled<5, 5>(a, g, s);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <int H> void launch_ds_output_h(const DsArgs &a, dim3 g, hipStream_t s) {
    if (a.type == RAW_U8)
        hipLaunchKernelGGL((k_ds_output<uint8_t, H>), g, dim3(kBlock), 0, s, a);
    else if (a.type == RAW_U16)
        hipLaunchKernelGGL((k_ds_output<uint16_t, H>), g, dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL((k_ds_output<float, H>), g, dim3(kBlock), 0, s, a);
}

hipError_t launch_ds_output(const DsArgs &a, hipStream_t s) {
    static_assert(kDsTW * kDsTH == kBlock, "one thread per output of the tile");
    const dim3 g((unsigned)((a.cw + kDsTW - 1) / kDsTW), (unsigned)((a.ch + kDsTH - 1) / kDsTH));
    if (a.strength != 0.0f && a.quality == 1)
        launch_ds_output_h<3>(a, g, s);
    else
        launch_ds_output_h<2>(a, g, s);
    return hipGetLastError();
}

hipError_t launch_ds_candidates(const DsArgs &a, hipStream_t s) {
    const dim3 g((unsigned)((a.cw + kDsTW - 1) / kDsTW), (unsigned)((a.ch + kDsTH - 1) / kDsTH));
    LGCR_DS_TYPED(k_ds_candidates, g)
    return hipGetLastError();
}

hipError_t launch_ds_score(const DsArgs &a, hipStream_t s) {
    const dim3 g((unsigned)((a.cw + kBlock - 1) / kBlock), (unsigned)a.ch);
    LGCR_DS_TYPED(k_ds_score, g)
    return hipGetLastError();
}

#undef LGCR_DS_TYPED

namespace {
// Compares rcp_nr(y) with the IEEE quotient 1.0f / y, bit for bit, for every
// float bit pattern in [lo, hi).  out[0] counts mismatches; out[1] keeps the
// smallest mismatching bit pattern (0xffffffff if none).
__global__ void k_check_rcp(uint32_t lo, uint32_t hi, unsigned long long *out) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    unsigned long long bad = 0;
    uint32_t first = 0xffffffffu;
    for (uint64_t b = lo + uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; b < hi; b += stride) {
        const float y = __uint_as_float((uint32_t)b);
        const float q = 1.0f / y;
        const float r = rcp_nr(y);
        if (__float_as_uint(q) != __float_as_uint(r)) {
            ++bad;
            first = min(first, (uint32_t)b);
        }
    }
    if (bad) {
        atomicAdd(&out[0], bad);
        atomicMin(&out[1], (unsigned long long)first);
    }
}
}  // namespace

hipError_t launch_check_rcp(uint32_t lo, uint32_t hi, unsigned long long *out, hipStream_t s) {
    hipLaunchKernelGGL(k_check_rcp, dim3(2048), dim3(kBlock), 0, s, lo, hi, out);
    return hipGetLastError();
}

}  // namespace lgcr_hip
